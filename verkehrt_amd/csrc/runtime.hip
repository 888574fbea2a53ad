// The context, the caching device pool, device array views and per-context scratch.
#include <cstdio>
#include <cstring>

#include "vk_device.h"
#include "vk_internal.h"

namespace vk {

void fail(int code, std::string msg) { throw Error{code, std::move(msg)}; }
void runtime_error(std::string msg) { fail(VK_ERR_RUNTIME, std::move(msg)); }
void usage_error(std::string msg) { fail(VK_ERR_USAGE, std::move(msg)); }

void check_hip(hipError_t e, const char *what) {
  if (e != hipSuccess) fail(VK_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

const char *type_name(int32_t t) {
  switch (t) {
    case VK_F64: return "f64";
    case VK_F32: return "f32";
    case VK_I64: return "i64";
    case VK_U64: return "u64";
    case VK_I32: return "i32";
    case VK_U8: return "u8";
    case VK_BOOL: return "bool";
    default: return "?";
  }
}
bool valid_type(int32_t t) { return t >= VK_F64 && t <= VK_BOOL; }

// ------------------------------------------------------------------- pool

void *Pool::alloc(uint64_t bytes) {
  std::lock_guard<std::mutex> lock(mu_);
  if (bytes == 0) bytes = 256;
  bytes = (bytes + 255) & ~uint64_t(255);
  auto it = free_.lower_bound(bytes);
  if (it != free_.end() && it->first <= bytes + bytes / 4) {
    void *p = it->second;
    in_use_ += it->first;
    free_.erase(it);
    return p;
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    // Give cached blocks back and try once more before reporting.
    (void)hipGetLastError();
    trim_locked();
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  size_of_[p] = bytes;
  held_ += bytes;
  in_use_ += bytes;
  return p;
}

void Pool::free(void *p) {
  if (!p) return;
  std::lock_guard<std::mutex> lock(mu_);
  auto it = size_of_.find(p);
  if (it == size_of_.end()) return;
  in_use_ -= it->second;
  free_.emplace(it->second, p);
}

void Pool::trim() {
  std::lock_guard<std::mutex> lock(mu_);
  trim_locked();
}

void Pool::trim_locked() {
  for (auto &kv : free_) {
    (void)hipFree(kv.second);
    held_ -= kv.first;
    size_of_.erase(kv.second);
  }
  free_.clear();
}

Pool::~Pool() {
  trim_locked();
  for (auto &kv : size_of_) (void)hipFree(kv.first);
}

// ----------------------------------------------------------------- arrays

vk_array *new_array(Ctx &c, int32_t type, uint64_t len) {
  const uint64_t w = type_width(type);
  if (len > (uint64_t(1) << 62) / w)
    runtime_error("array of " + std::to_string(len) + " elements is too large");
  void *p = c.pool->alloc(len * w);
  if (!p) runtime_error("cannot allocate an array of " + std::to_string(len) + " elements");
  auto *b = new Buffer{&c, p, len * w, true, 1};
  ctx_retain(c);
  return new vk_array{1, b, 0, len, type};
}

void release(vk_array *a) {
  if (!a) return;
  if (--a->refs > 0) return;
  Buffer *b = a->buf;
  if (--b->refs == 0) {
    if (b->owned) {
      b->ctx->pool->free(b->ptr);
      ctx_release(b->ctx);
    }
    delete b;
  }
  delete a;
}

// ---------------------------------------------------------------- scratch

__global__ void k_fill_u64(uint64_t *p, uint64_t n, uint64_t v) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

void fill_u64(Ctx &c, uint64_t *p, uint64_t n, uint64_t value) {
  if (n == 0) return;
  k_fill_u64<<<stream_grid(n, kBlock), kBlock, 0, c.stream>>>(p, n, value);
  check_hip(hipGetLastError(), "fill");
}

uint64_t *status_words(Ctx &c, int family, uint64_t words) {
  if (words > c.scratch.status_words[family]) {
    uint64_t want = words < 4096 ? 4096 : words + words / 4;
    if (c.scratch.status[family]) c.pool->free(c.scratch.status[family]);
    c.scratch.status[family] = static_cast<uint64_t *>(c.pool->alloc(want * 8));
    if (!c.scratch.status[family]) runtime_error("cannot allocate look-back status words");
    c.scratch.status_words[family] = want;
    fill_u64(c, c.scratch.status[family], want, 0);
  }
  return c.scratch.status[family];
}

uint32_t next_epoch(Ctx &c) {
  if (++c.epoch > 0xFFFF) {
    c.epoch = 1;
    for (int f = 0; f < ST_FAMILIES; ++f)
      if (c.scratch.status[f]) fill_u64(c, c.scratch.status[f], c.scratch.status_words[f], 0);
  }
  return c.epoch;
}

void *partials(Ctx &c, uint64_t bytes) {
  if (bytes > c.scratch.partial_bytes) {
    if (c.scratch.partials) c.pool->free(c.scratch.partials);
    c.scratch.partials = c.pool->alloc(bytes);
    if (!c.scratch.partials) runtime_error("cannot allocate reduction partials");
    c.scratch.partial_bytes = bytes;
  }
  return c.scratch.partials;
}

void *tmp_alloc(Ctx &c, uint64_t bytes) {
  void *p = c.pool->alloc(bytes);
  if (!p) runtime_error("cannot allocate " + std::to_string(bytes) + " bytes of device scratch");
  return p;
}
void tmp_free(Ctx &c, void *p) { c.pool->free(p); }

void read_back(Ctx &c, const void *src, uint64_t bytes, void *host) {
  check_hip(hipMemcpyAsync(c.host_slots, src, bytes, hipMemcpyDeviceToHost, c.stream), "read back");
  check_hip(hipStreamSynchronize(c.stream), "stream sync");
  std::memcpy(host, c.host_slots, bytes);
}

void copy_d2d(Ctx &c, void *dst, const void *src, uint64_t bytes) {
  if (bytes == 0) return;
  check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c.stream), "copy");
}

BatchPlan batch_plan(Ctx &c, std::vector<BatchItem> &items, uint64_t rows_per_tile) {
  BatchPlan p;
  uint64_t tiles = 0;
  for (auto &it : items) {
    it.tile0 = tiles;
    tiles += (it.n + rows_per_tile - 1) / rows_per_tile;
  }
  if (tiles > 0xFFFFFFFFull) usage_error("batch has too many tiles");
  p.ntiles = static_cast<uint32_t>(tiles);
  if (items.empty()) return p;
  std::vector<uint32_t> map(tiles);
  for (size_t i = 0; i < items.size(); ++i) {
    const uint64_t t1 = i + 1 < items.size() ? items[i + 1].tile0 : tiles;
    for (uint64_t t = items[i].tile0; t < t1; ++t) map[t] = static_cast<uint32_t>(i);
  }
  p.items = static_cast<BatchItem *>(tmp_alloc(c, items.size() * sizeof(BatchItem)));
  p.tile_item = static_cast<uint32_t *>(tmp_alloc(c, (tiles ? tiles : 1) * 4));
  check_hip(hipMemcpyAsync(p.items, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice,
                           c.stream), "batch items");
  if (tiles)
    check_hip(hipMemcpyAsync(p.tile_item, map.data(), tiles * 4, hipMemcpyHostToDevice, c.stream),
              "batch tiles");
  // the host vectors die with this call: wait for the copies
  check_hip(hipStreamSynchronize(c.stream), "batch plan");
  return p;
}

namespace {

constexpr int kPlanThreads = 1024;

// Tile counts and first tiles of every item, by one block: thread t takes a contiguous range
// of items, the range totals are scanned in LDS, then each thread writes its items' tile0.
__global__ __launch_bounds__(kPlanThreads) void k_plan_items(BatchItem *items, uint32_t nitems,
                                                             const uint64_t *dev_n, uint32_t unit,
                                                             uint64_t per_tile, uint64_t min_rows,
                                                             uint32_t max_tiles, uint32_t *ntiles) {
  __shared__ uint64_t s_part[kPlanThreads];
  const uint32_t t = threadIdx.x;
  const uint32_t per = (nitems + kPlanThreads - 1) / kPlanThreads;
  const uint32_t i0 = t * per < nitems ? t * per : nitems;
  const uint32_t i1 = i0 + per < nitems ? i0 + per : nitems;
  // items below min_rows rows get no tiles (another kernel of the statement takes them)
  auto tiles_of = [&](uint64_t n) { return n < min_rows ? 0 : (n / unit + per_tile - 1) / per_tile; };
  uint64_t sum = 0;
  for (uint32_t i = i0; i < i1; ++i) {
    const uint64_t n = dev_n ? dev_n[i] : items[i].n;
    if (dev_n) items[i].n = n;
    sum += tiles_of(n);
  }
  s_part[t] = sum;
  __syncthreads();
  for (uint32_t d = 1; d < kPlanThreads; d <<= 1) {  // inclusive scan (Hillis-Steele)
    const uint64_t o = t >= d ? s_part[t - d] : 0;
    __syncthreads();
    s_part[t] += o;
    __syncthreads();
  }
  uint64_t run = s_part[t] - sum;
  for (uint32_t i = i0; i < i1; ++i) {
    items[i].tile0 = run;
    run += tiles_of(items[i].n);
  }
  // the grid of the launches is max_tiles, an upper bound the host proved from input sizes;
  // the clamp only keeps every access inside the map
  if (t == kPlanThreads - 1) *ntiles = static_cast<uint32_t>(s_part[t] < max_tiles ? s_part[t] : max_tiles);
}

// Tile t belongs to the last item whose first tile is at most t (items without tiles share
// their successor's first tile, so the search never stops on one).
__global__ void k_plan_map(const BatchItem *items, uint32_t nitems, uint32_t max_tiles, const uint32_t *ntiles,
                           uint32_t *tile_item) {
  const uint64_t t = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (t >= max_tiles || t >= *ntiles) return;
  uint32_t lo = 0, hi = nitems - 1;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo + 1) / 2;
    if (items[mid].tile0 <= t) lo = mid;
    else hi = mid - 1;
  }
  tile_item[t] = lo;
}

}  // namespace

void plan_tiles_dev(Ctx &c, const DevPlan &p, uint32_t nitems, const uint64_t *dev_n, uint32_t unit,
                    uint64_t per_tile, uint64_t min_rows) {
  if (nitems == 0) return;
  k_plan_items<<<1, kPlanThreads, 0, c.stream>>>(p.items, nitems, dev_n, unit, per_tile, min_rows, p.max_tiles,
                                                 p.ntiles);
  if (p.max_tiles)
    k_plan_map<<<(p.max_tiles + kBlock - 1) / kBlock, kBlock, 0, c.stream>>>(p.items, nitems, p.max_tiles, p.ntiles,
                                                                           p.tile_item);
  check_hip(hipGetLastError(), "device tile plan");
}

void batch_plan_free(Ctx &c, BatchPlan &p) {
  if (p.items) tmp_free(c, p.items);
  if (p.tile_item) tmp_free(c, p.tile_item);
  p = BatchPlan{};
}

Ctx *create_ctx(int device) {
  check_hip(hipSetDevice(device), "hipSetDevice");
  auto *c = new Ctx();
  c->device = device;
  c->cus = 0;
  check_hip(hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device), "CU count");
  if (c->cus < 1) c->cus = 1;
  check_hip(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream create");
  c->pool = new Pool(device);
  c->epoch = 0;
  c->trace = false;
  c->last_us = 0;
  check_hip(hipEventCreate(&c->ev0), "event");
  check_hip(hipEventCreate(&c->ev1), "event");
  check_hip(hipHostMalloc(reinterpret_cast<void **>(&c->host_slots), 64 * 8, hipHostMallocDefault),
            "pinned slots");
  void *counters = c->pool->alloc(256);
  void *slots = c->pool->alloc(64 * 8);
  if (!counters || !slots) runtime_error("cannot allocate context scratch");
  c->scratch.counters = static_cast<uint32_t *>(counters);
  c->scratch.slots = static_cast<uint64_t *>(slots);
  fill_u64(*c, static_cast<uint64_t *>(counters), 256 / 8, 0);
  check_hip(hipStreamSynchronize(c->stream), "stream sync");
  select_into_setup(*c);
  return c;
}

void ctx_retain(Ctx &c) { c.refs.fetch_add(1); }

void ctx_release(Ctx *c) {
  if (!c || c->refs.fetch_sub(1) != 1) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  delete c->pool;
  (void)hipHostFree(c->host_slots);
  (void)hipEventDestroy(c->ev0);
  (void)hipEventDestroy(c->ev1);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

void destroy_ctx(Ctx *c) { ctx_release(c); }

}  // namespace vk
