// The scratch of vk_select_into, the selection entry that allocates nothing: one file-scope
// device array, `scratch`, of which every GPU holds its own copy until the process ends.
//
// Host code never uses the array's name as an address (on the host it is only a stand-in for
// the symbol): at context creation, once per device and with that device current,
// hipGetSymbolAddress resolves the device's copy, and that pointer is what launches receive.
//
// Layout (64-bit words): [0] the tile ticket (a u32 in the low half; zero before every launch,
// and the launch's last tile returns it to zero), [1] the kept count (written by the last tile
// of every launch, so never cleared), [2, 32) unused, [32, 32 + tiles) one look-back status
// word per 16384-row tile. Status words carry a 16-bit epoch that each call takes anew per
// device, so the words of earlier calls read as not published; when the epoch wraps they are
// cleared by a fill kernel (no memset, so the entry stays capturable). Calls that overlap on one
// GPU share the array and race.
#include <mutex>

#include "vk_device.h"
#include "vk_internal.h"

namespace vk {
namespace {

constexpr uint64_t kHeadWords = 32;
constexpr uint64_t kTileRows = 16384;
constexpr uint64_t kIntoTiles = VK_SELECT_INTO_MAX_ROWS / kTileRows;  // 131072 words, 1 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint64_t scratch[kHeadWords + kIntoTiles];

struct DevScratch {
  uint64_t *ptr = nullptr;  // this device's copy of `scratch`
  uint32_t epoch = 0;       // the last epoch taken on this device, 1..65535
};

std::mutex g_mu;
DevScratch g_dev[kMaxDevices];

}  // namespace

void select_into_setup(Ctx &c) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (c.device < 0 || c.device >= kMaxDevices) usage_error("device ordinal out of range");
  DevScratch &d = g_dev[c.device];
  if (d.ptr) return;
  void *p = nullptr;
  check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  if (!p) fail(VK_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned no address");
  fill_u64(c, static_cast<uint64_t *>(p), kHeadWords + kIntoTiles, 0);
  check_hip(hipStreamSynchronize(c.stream), "stream sync");
  d.ptr = static_cast<uint64_t *>(p);
}

IntoScratch select_into_scratch(Ctx &c, uint64_t rows) {
  if (rows > VK_SELECT_INTO_MAX_ROWS) usage_error("selection of " + std::to_string(rows) + " rows exceeds the limit");
  std::lock_guard<std::mutex> lock(g_mu);
  DevScratch &d = g_dev[c.device];
  if (!d.ptr) fail(VK_ERR_DEVICE, "the allocation-free scratch of this device is not set up");
  if (++d.epoch > 0xFFFF) {
    d.epoch = 1;
    fill_u64(c, d.ptr + kHeadWords, kIntoTiles, 0);
  }
  IntoScratch s;
  s.ticket = reinterpret_cast<uint32_t *>(d.ptr);
  s.count = d.ptr + 1;
  s.status = d.ptr + kHeadWords;
  s.epoch = d.epoch;
  return s;
}

}  // namespace vk
