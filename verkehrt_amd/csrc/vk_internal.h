// Internal declarations of the hip backend: the context, device buffers and array views,
// the error type, and the host launchers of every kernel family. Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vk_hip.h"

namespace vk {

// An error that crosses the C ABI as a vk_status: `code` is VK_ERR_*, `msg` the text.
struct Error {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, std::string msg);
[[noreturn]] void runtime_error(std::string msg);  // VK_ERR_RUNTIME, reference text
[[noreturn]] void usage_error(std::string msg);    // VK_ERR_USAGE, reference text
void check_hip(hipError_t e, const char *what);

inline uint32_t type_width(int32_t t) {
  switch (t) {
    case VK_F64: case VK_I64: case VK_U64: return 8;
    case VK_F32: case VK_I32: return 4;
    default: return 1;
  }
}
const char *type_name(int32_t t);  // ScalarType::name, src/value.rs:57-71
bool valid_type(int32_t t);

// A caching device allocator for one context. Freed blocks are kept by size and reused
// for requests of at most 1.25x their size; all reuse is ordered by the context's single
// stream, so a block freed after its last use was enqueued is safe to hand out again.
class Pool {
 public:
  explicit Pool(int device) : device_(device) {}
  // Thread safe: an array's last reference may be dropped on any thread. Reuse is ordered by
  // the owning context's stream; memory used on another context's stream must not be
  // released before that stream has finished with it.
  void *alloc(uint64_t bytes);
  void free(void *p);
  void trim();
  uint64_t held() const { return held_; }
  uint64_t in_use() const { return in_use_; }
  ~Pool();

 private:
  void trim_locked();
  int device_;
  std::mutex mu_;
  std::multimap<uint64_t, void *> free_;
  std::map<void *, uint64_t> size_of_;
  uint64_t held_ = 0, in_use_ = 0;
};

}  // namespace vk

struct vk_ctx;

namespace vk {

// A device allocation shared by array views. `owned` buffers return to the pool.
struct Buffer {
  vk_ctx *ctx;
  void *ptr;
  uint64_t bytes;
  bool owned;
  std::atomic<int> refs;
};

}  // namespace vk

struct vk_array {
  std::atomic<int> refs;
  vk::Buffer *buf;
  uint64_t offset;  // bytes from buf->ptr to row 0
  uint64_t len;
  int32_t type;
};

namespace vk {

// Per-context device scratch reused by every launch of the context: look-back status
// words, tickets, done counters and small result slots. Grows on demand (between calls).
// Look-back status buffers, one per kernel family, each with one fixed word layout, so a
// word at a flag position was only ever written as a flag (older epochs read as not ready).
enum StatusFamily { ST_SELECT = 0, ST_SCAN, ST_SORT, ST_CANON, ST_FAMILIES };

struct Scratch {
  uint64_t *status[ST_FAMILIES] = {};  // look-back words
  uint64_t status_words[ST_FAMILIES] = {};
  uint32_t *counters = nullptr;  // [0] ticket, [1] done, [2..] spare; always returned to 0
  uint64_t *slots = nullptr;     // 64 x 8 bytes of small device results
  void *partials = nullptr;      // per-block partial results
  uint64_t partial_bytes = 0;
};

}  // namespace vk

struct vk_ctx {
  int device;
  int cus;  // compute units of the device (the grid of persistent kernels)
  hipStream_t stream;
  vk::Pool *pool;
  vk::Scratch scratch;
  uint32_t epoch;  // look-back epoch, 1..65535
  uint64_t *host_slots;  // pinned 64 x 8 bytes for small D2H results
  bool trace;
  std::string trace_buf;
  double last_us;
  hipEvent_t ev0, ev1;
  // Holders of the context: its creator (until vk_ctx_destroy), every owned buffer, reducer,
  // scanner, builder and plan. The stream, pool and scratch go when the last one lets go, so
  // an array released after vk_ctx_destroy still returns its memory to a live pool.
  std::atomic<int> refs{1};
};

namespace vk {

using Ctx = vk_ctx;

void ctx_retain(Ctx &c);
void ctx_release(Ctx *c);  // the last release tears the context down

// Arrays ------------------------------------------------------------------
vk_array *new_array(Ctx &c, int32_t type, uint64_t len);  // uninitialised
inline void *dptr(const vk_array *a) { return static_cast<char *>(a->buf->ptr) + a->offset; }
inline uint64_t abytes(const vk_array *a) { return a->len * type_width(a->type); }
void release(vk_array *a);

// Scratch -----------------------------------------------------------------
uint64_t *status_words(Ctx &c, int family, uint64_t words);  // zeroed at growth
uint32_t next_epoch(Ctx &c);  // 1..65535; clears every status buffer on wrap
void *partials(Ctx &c, uint64_t bytes);
void *tmp_alloc(Ctx &c, uint64_t bytes);  // pool memory for one call
void tmp_free(Ctx &c, void *p);
// Copies `bytes` of device memory at `src` to the pinned host slots and waits.
void read_back(Ctx &c, const void *src, uint64_t bytes, void *host);
void fill_u64(Ctx &c, uint64_t *p, uint64_t n, uint64_t value);

// Launch geometry -----------------------------------------------------------
constexpr int kBlock = 256;
inline uint32_t stream_grid(uint64_t work_items, uint32_t per_block, uint32_t cap = 2048) {
  uint64_t g = (work_items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<uint32_t>(g);
}

// Comparison codes in the order of CMPS (src/kernels/scalar/fused.rs:33-40).
enum Cmp { CMP_EQ = 0, CMP_NE, CMP_LT, CMP_LE, CMP_GT, CMP_GE };

// ---------------------------------------------------------------- launchers
// All launchers enqueue on c.stream. Device pointers are rows of the named type.

// Streaming reductions (k_reduce.hip). `state` is the reducer's device state; each
// launch merges the reduction of its rows into it.
struct ReduceSpec {
  enum Kind { SUM_INT, COUNT_CMP, COUNT_BOOL, EXTREME, ARG_EXTREME, BITS, PROD_INT, FSUM_INT } kind;
  int32_t type;       // element type of x
  int cmp;            // COUNT_CMP, FSUM_INT condsum
  uint64_t cbits;     // COUNT_CMP / condsum constant bits
  bool want_max;      // EXTREME / ARG_EXTREME
  int bitop;          // BITS: 0 and, 1 or, 2 xor
  uint64_t base_row;  // ARG_EXTREME: rows consumed before this update
  int fkind;          // FSUM_INT: 1 dot (two inputs, reduce_rows2), 2 sumsq, 3 condsum
  uint64_t ebits;     // FSUM_INT condsum: the value of rows that fail the comparison
};
// Reduce state layouts (device): SUM_INT/COUNT/BITS: u64; EXTREME: {u64 bits, u64 has};
// ARG_EXTREME: {u64 best_bits, i64 best_idx, i64 nan_idx}.
void reduce_rows(Ctx &c, const ReduceSpec &s, const void *x, uint64_t n, void *state);
// Integer dot (FSUM_INT, fkind 1): the wrapping sum of the widened wrapping products x[i] * y[i].
void reduce_rows2(Ctx &c, const ReduceSpec &s, const void *x, const void *y, uint64_t n, void *state);
void reduce_state_init(Ctx &c, const ReduceSpec &s, void *state);

// Canonical float reductions (k_canon.hip). kind: 0 sum, 1 dot, 2 sumsq, 3 condsum, 4 mul.
// `acc` is 16 doubles; blocks of 16 rows fold into them in the canonical order.
struct CanonSpec {
  int kind;
  int cmp;
  double c, e;
};
void canon_blocks(Ctx &c, const CanonSpec &s, const double *x, const double *y, uint64_t blocks,
                  double *acc);
// The same at f32 (k_canon32.hip): acc holds 16 f32 values as doubles; every step rounds to
// f32. canon32_finish writes the result's f32 bits (zero-extended) to out.
void canon32_blocks(Ctx &c, const CanonSpec &s, const float *x, const float *y, uint64_t blocks, double *acc);
void canon32_rest(Ctx &c, const CanonSpec &s, const float *x, const float *y, uint64_t rest, double *tail);
void canon32_finish(Ctx &c, bool mul, const double *acc, const double *tail, int tail_len, uint64_t *out);
// Writes the row values of rows [0, rest) (rest < 16) to `tail`.
void canon_rest(Ctx &c, const CanonSpec &s, const double *x, const double *y, uint64_t rest,
                double *tail);
void canon_finish(Ctx &c, bool mul, const double *acc, const double *tail, int tail_len,
                  double *out);
void canon_init(Ctx &c, bool mul, double *acc);
// Batched canonical reductions (declared after BatchItem below).

// Selections (k_select.hip): append the kept rows at `out`, the count to *d_count (device).
// kind 0 compress (mask), 1 filter (cmp with cbits), 2 where (base).
struct SelectSpec {
  int kind;
  int32_t type;  // row type (compress/filter); where writes i64
  int cmp;
  uint64_t cbits;
  uint64_t base;
};
void select_rows(Ctx &c, const SelectSpec &s, const void *x, const uint8_t *mask, uint64_t n,
                 void *out, uint64_t *d_count);
// The same launch with the caller's look-back scratch: `status` holds one word per tile of
// select_rows_per_tile() rows, written with `epoch` (words of other epochs read as not
// published); `ticket` is zero before the launch and zero again after it.
uint64_t select_rows_per_tile();
void select_rows_scratch(Ctx &c, const SelectSpec &s, const void *x, const uint8_t *mask, uint64_t n,
                         void *out, uint64_t *d_count, uint64_t *status, uint32_t epoch, uint32_t *ticket);

// Batched launches: many independent arrays in one launch per kernel phase. Tiles are
// numbered over all items; every tile belongs to one item, and look-backs stop at the
// item's first tile.
struct BatchItem {
  const void *a;   // first input rows
  const void *b;   // second input rows (mask, y), or null
  void *out;       // output rows
  uint64_t n;      // rows
  uint64_t tile0;  // first tile of the item in the launch
  uint64_t aux;    // kernel-specific (where: index base; sorts: first concatenated row)
  uint64_t aux2;   // kernel-specific (group_id: first table word)
  uint64_t aux3;   // kernel-specific (group_id: log2 of the table size)
};
struct BatchPlan {
  BatchItem *items = nullptr;     // device copy of the items
  uint32_t *tile_item = nullptr;  // device: the item of each tile
  uint32_t ntiles = 0;
};
// Sets tile0 of every item for `rows_per_tile` rows per tile and uploads the items and the
// tile -> item map.
BatchPlan batch_plan(Ctx &c, std::vector<BatchItem> &items, uint64_t rows_per_tile);
void batch_plan_free(Ctx &c, BatchPlan &p);
// Batched reductions: item i's final state (the reduce state layout, 8-byte words) at
// states + i * reduce_state_words(s); items are the rows a (n).
uint32_t reduce_state_words(const ReduceSpec &s);
void reduce_batch(Ctx &c, const ReduceSpec &s, std::vector<BatchItem> &items, uint64_t *states);
// Batched group_id_<t>: item i numbers its keys a (n rows) into the i64 ids at out; the
// number of groups of item i goes to d_groups[i] (device). Sets aux, aux2, aux3.
void group_id_batch(Ctx &c, int32_t type, std::vector<BatchItem> &items, uint64_t *d_groups);
// Batched key checks of reduce_into: item i (keys a, n rows, aux2 buckets) lowers d_bad[i] to
// its first row whose key is outside 0..buckets (all ones when none).
void check_keys_batch(Ctx &c, std::vector<BatchItem> &items, uint64_t *d_bad);
// Batched reduce_into_add_f64 (keys checked): item i folds values b by keys a (n rows) into
// the aux2 f64 buckets at out, in the order of reduce_into_add_f64. Sets aux3.
void reduce_into_add_f64_batch(Ctx &c, std::vector<BatchItem> &items);
// Batched launches whose tiles are planned on the device, so that no host read-back sits
// between statements and a statement sequence can be captured into a hipGraph: item row
// counts may come from an earlier launch's device results (a selection's kept counts), the
// launch grid is an upper bound of the tiles, and blocks past the planned count exit. The
// look-back kernels take their epoch from `epoch` (bumped on the device once per run) and their
// tickets from their own counter, so one plan's launches share no scratch with other calls.
struct DevPlan {
  BatchItem *items = nullptr;     // device items; n (from dev_n) and tile0 set on the device
  uint32_t *tile_item = nullptr;  // device tile -> item map, max_tiles entries
  uint32_t *ntiles = nullptr;     // device: the planned tile count
  uint32_t max_tiles = 0;         // the launch grid, an upper bound of the planned tiles
  uint32_t *epoch = nullptr;      // device epoch word of the look-back kernels
  uint64_t *status = nullptr;     // look-back status words (max_tiles x words per tile)
  uint32_t *ticket = nullptr;     // ticket counter, zero before a launch and after it
};
// Item i has n_i rows (dev_n[i] when dev_n is set, written to items[i].n; else items[i].n) and
// ceil(floor(n_i / unit) / per_tile) tiles, none when n_i < min_rows; writes items[i].tile0,
// the map and the tile count.
void plan_tiles_dev(Ctx &c, const DevPlan &p, uint32_t nitems, const uint64_t *dev_n, uint32_t unit,
                    uint64_t per_tile, uint64_t min_rows = 0);
// Rows per tile of the batched selection with 8 or 16 rows per thread (0: the persistent filter
// form, one block per compute unit over 16K-row tiles).
uint64_t select_tile_rows(int items_per_thread);
void select_batch_dev(Ctx &c, const SelectSpec &s, int items_per_thread, const DevPlan &p, uint32_t nitems,
                      uint64_t *d_counts);
// Batched canonical sums (plan unit 16, canon_tile_blocks() blocks per tile); acc and zero are
// nitems x 16 doubles of scratch.
uint64_t canon_tile_blocks();
// Batched canonical sums fold items of fewer rows than this one wave each, in order; larger
// items go by tiles (plan their tiles with min_rows = canon_tile_min_rows()).
uint64_t canon_tile_min_rows();
uint64_t canon_status_words_per_tile();
void canon_batch_dev(Ctx &c, const CanonSpec &s, const DevPlan &p, uint32_t nitems, double *acc, double *zero,
                     double *d_out);
// Batched reductions (plan unit 1, tile_rows rows per tile: 4096, 8192 or 16384); partials
// holds max_tiles accumulators of at most 24 bytes.
void reduce_batch_dev(Ctx &c, const ReduceSpec &s, uint64_t tile_rows, const DevPlan &p, uint32_t nitems,
                      void *partials, uint64_t *states);
// One selection launch over all items; the kept count of item i goes to d_counts[i].
void select_batch(Ctx &c, const SelectSpec &s, std::vector<BatchItem> &items, uint64_t *d_counts);
// One canonical-reduction launch over all items (a = x, b = y or null); item i's result to
// d_out[i].
void canon_batch(Ctx &c, const CanonSpec &s, std::vector<BatchItem> &items, double *d_out);

// Scans (k_scan.hip). `carry` is the scanner's device state {u64 acc, u64 has}.
// op 0 add (out width 8, input widened from `type`), 1 min, 2 max.
struct ScanSpec {
  int op;
  int32_t type;
  bool by;
};
void scan_rows(Ctx &c, const ScanSpec &s, const void *x, const uint8_t *starts, uint64_t n,
               void *out, uint64_t *carry);
void scan_add_f64(Ctx &c, const double *x, uint64_t n, double *out, double *state);
void canon_tile_starts(Ctx &c, const double *x, uint64_t blocks, double *acc, double *tstart);
// prev: row i is row i - 1, row 0 the carried last row or `fill_bits`.
void prev_rows(Ctx &c, uint32_t width, const void *x, uint64_t n, void *out, uint64_t *carry,
               uint64_t fill_bits);

// The allocation-free selection's scratch (noalloc.hip): select_into_setup resolves this
// device's copy of the fixed device array once (at context creation, with the device current);
// select_into_scratch hands a call its ticket, count slot and status words with a new epoch.
struct IntoScratch {
  uint32_t *ticket;
  uint64_t *count;
  uint64_t *status;
  uint32_t epoch;
};
void select_into_setup(Ctx &c);
IntoScratch select_into_scratch(Ctx &c, uint64_t rows);

// Gather (k_gather.hip): out[i] = xs[idx[i]]; the first bad row (or ~0) to *d_bad.
void gather_rows(Ctx &c, uint32_t width, const void *xs, uint64_t xlen, const int64_t *idx,
                 uint64_t n, void *out, uint64_t *d_bad);

// Buckets (k_bucket.hip): histogram / reduce_into; first bad key row to *d_bad.
// op: 0 count (histogram), 1 add i64, 2 min, 3 max (type f64 or i64).
void bucket_init(Ctx &c, int op, int32_t type, void *out, uint64_t buckets);
void bucket_rows(Ctx &c, int op, int32_t type, const int64_t *keys, const void *vals, uint64_t n,
                 uint64_t buckets, void *out, uint64_t *d_bad);
// reduce_into_add_f64 checks its keys itself (inside the sort's counting pass): it returns the
// first row whose key is outside 0..buckets, or ~0, and writes out only when there is none.
uint64_t reduce_into_add_f64(Ctx &c, const int64_t *keys, const double *vals, uint64_t n,
                             uint64_t buckets, double *out);
void check_keys(Ctx &c, const int64_t *keys, uint64_t n, uint64_t buckets, uint64_t *d_bad);

// Argsort (k_sort.hip): stable ascending positions as i64.
void argsort(Ctx &c, int32_t type, const void *keys, uint64_t n, int64_t *out);
// Batched argsort: item i sorts keys a (n rows) into the i64 positions at out; aux is set.
void argsort_batch(Ctx &c, int32_t type, std::vector<BatchItem> &items);
// Sorts inside batch plans (k_sort.hip): every launch reads its decisions from the device, so a
// plan captures them. Positions mode (pos) writes each item's i64 positions to item.out; (key,
// value) mode sorts the keys' images (type kChainKey: i64 bucket k of row r as k * 16 + r % 16,
// u64 images when wide, else u32) with the u64 values of item.b into sorted_keys / sorted_vals
// at each item's first row of the concatenated rows (item.aux, set here).
struct PlanSort {
  int32_t type = 0;
  bool pos = true, wide = true;
  int P = 8;
  uint64_t nitems = 0, total = 0;
  BatchPlan bp;
  uint32_t *hist = nullptr;     // per item and digit: counts, then bases; then the active digits
  int32_t *slots = nullptr;     // [0] active digits, [1 + k] the k-th
  void *kbuf[2] = {nullptr, nullptr}, *pbuf[2] = {nullptr, nullptr};
  void *sorted_keys = nullptr;
  uint64_t *sorted_vals = nullptr;
  uint64_t *status = nullptr;   // P slots x tiles x 256 look-back words
  uint64_t status_words = 0;
  uint32_t *ticket = nullptr;
};
int32_t chain_key_type();
// group_id inside batch plans (k_group.hip): item.a the keys, item.out the ids (n rows);
// groups[i] receives item i's group count. Status words of its selection: sel.status.
struct PlanGroup {
  int32_t type = 0;
  uint64_t nitems = 0, total = 0, small_words = 0;
  BatchPlan small_p, fill_p;
  BatchItem *items_full = nullptr, *items_final = nullptr;
  unsigned long long *small = nullptr, *full = nullptr;
  uint32_t *over = nullptr, *slot = nullptr;
  uint8_t *first = nullptr;
  int64_t *firsts = nullptr;
  uint64_t *groups = nullptr;
  DevPlan sel;
  uint64_t status_words = 0;
};
PlanGroup *plan_group_new(Ctx &c, int32_t type, std::vector<BatchItem> &items, uint32_t *epoch);
void plan_group_run(Ctx &c, PlanGroup &g);
void plan_group_free(Ctx &c, PlanGroup *g);
// reduce_into_add_f64 inside batch plans (k_bucket.hip): item.a the keys (group ids of a
// group_id statement), item.b the f64 values, item.n the rows, item.out the output (room for
// n rows); groups: the group_id statement's device group counts, the bucket counts.
struct PlanFold {
  uint64_t nitems = 0;
  const uint64_t *groups = nullptr;
  bool wide = false;
  void *sk = nullptr;
  double *sv = nullptr;
  PlanSort *sort = nullptr;
  DevPlan dp;
};
PlanFold *plan_fold_new(Ctx &c, std::vector<BatchItem> &items, const uint64_t *groups);
void plan_fold_run(Ctx &c, PlanFold &f, const uint32_t *epoch);
void plan_fold_free(Ctx &c, PlanFold *f);
PlanSort *plan_sort_new(Ctx &c, int32_t type, bool pos, bool wide, std::vector<BatchItem> &items, void *sorted_keys,
                        uint64_t *sorted_vals);
void plan_sort_run(Ctx &c, PlanSort &s, const uint32_t *epoch);
void plan_sort_free(Ctx &c, PlanSort *s);
// Stable sort of u64 keys: sorted keys and the positions they came from.
void sort_u64_pairs(Ctx &c, const uint64_t *keys, uint64_t n, uint64_t *sorted_keys, int64_t *perm);
// Stable sort of (key, 8-byte value) pairs by key: keys are u64 (wide) or u32.
void sort_key_values(Ctx &c, bool wide, const void *keys, const uint64_t *vals, uint64_t n,
                     void *sorted_keys, uint64_t *sorted_vals);
// The same for chain keys computed on the fly: row r of bucket keys[r] has chain key
// keys[r] * 16 + r % 16 (u32 unless wide). The counting pass also checks the bucket keys:
// *bad_row is set to the first row whose key is outside 0..buckets (or ~0), and nothing is
// sorted when there is one. With class_shift, the top non-constant digit of the chain keys is
// left unsorted when there are two or more (the rows are in stable order of the key bits below
// *class_shift, their class); otherwise *class_shift is 8 above the top non-constant digit.
void sort_chain_values(Ctx &c, bool wide, const int64_t *keys, const uint64_t *vals, uint64_t n,
                       void *sorted_keys, uint64_t *sorted_vals, uint64_t buckets, uint64_t *bad_row,
                       int *class_shift = nullptr);
// Batched: item i sorts its keys a with the values b (n rows) into sorted_keys / sorted_vals
// from row aux (set here) of the concatenated rows.
void sort_key_values_batch(Ctx &c, bool wide, std::vector<BatchItem> &items, void *sorted_keys,
                           uint64_t *sorted_vals);
// The batched form of sort_chain_values: each item's bucket keys at a, its values at b, rows
// relative to the item.
void sort_chain_values_batch(Ctx &c, bool wide, std::vector<BatchItem> &items, void *sorted_keys,
                             uint64_t *sorted_vals);

// group_id (k_group.hip): first-occurrence ids. `b` may be null (one key column).
void group_id(Ctx &c, int32_t ta, const void *a, int32_t tb, const void *b, uint64_t n,
              int64_t *out);

// Construction and elementwise (k_misc.hip).
void rand_f64(Ctx &c, uint64_t n, uint64_t seed, double *out);
void rand_i64(Ctx &c, uint64_t n, uint64_t seed, int64_t *out);
void iota(Ctx &c, uint64_t n, int64_t *out);
void fill_bits(Ctx &c, uint32_t width, uint64_t n, uint64_t bits, void *out);
// Elementwise ops over rows or splats (k_misc.hip). in_type is the operands' type (select: the
// type of its two choices; cast: the source), out_type the target of a cast. Operand a is the
// condition of select, b and the third operand e its choices; unary ops read a (and b's
// scalar: the divisor of div_const).
enum EwOp {
  EW_ADD = 0, EW_SUB, EW_MUL, EW_EQ, EW_NE, EW_LT, EW_LE, EW_GT, EW_GE,
  EW_AND, EW_OR, EW_XOR, EW_NOT, EW_CAST, EW_MIN, EW_MAX, EW_NEG, EW_ABS, EW_DIV, EW_SQRT,
  EW_DIVC, EW_BAND, EW_BOR, EW_BXOR, EW_BNOT, EW_SELECT
};
struct EwSpec {
  int op;
  int32_t in_type;
  int32_t out_type;
  bool a_splat, b_splat, c_splat;
  uint64_t a_bits, b_bits, c_bits;
};
void elementwise(Ctx &c, const EwSpec &s, const void *a, const void *b, const void *e, uint64_t n, void *out);

// A fused region's elementwise chain (one register, no intermediate arrays): v starts as row x,
// then each op combines v with its operand, a constant (src 0), row x (1), row y (2) or v
// itself (3); `swap` puts the operand on the left (k - v). Ops are EwOp arithmetic codes of
// the chain's element type; every op rounds (or wraps) exactly as its elementwise kernel does.
constexpr int kChainMax = 8;
struct ChainOp {
  int32_t op;
  int32_t src;
  int32_t swap;
  int32_t pad;
  uint64_t kbits;  // the constant's bits
};
struct ChainSpec {
  int32_t n;
  int32_t type;  // element type of x, y and every op
  ChainOp ops[kChainMax];
};
// Canonical sum (reduce_add_f64 order) of the chain's f64 row values: accumulators and tail
// exactly as canon_blocks / canon_rest update them; y may be null when no op reads it.
void canon_blocks_chain(Ctx &c, const ChainSpec &ch, const double *x, const double *y, uint64_t blocks,
                        double *acc);
void canon_rest_chain(Ctx &c, const ChainSpec &ch, const double *x, const double *y, uint64_t rest,
                      double *tail);
// A streaming reduction (sum, count, extreme, arg) of the chain's values over one input x
// (no op reads y), merged into the reducer state as reduce_rows does.
void reduce_rows_chain(Ctx &c, const ReduceSpec &s, const ChainSpec &ch, const void *x, uint64_t n, void *state);
void copy_d2d(Ctx &c, void *dst, const void *src, uint64_t bytes);

}  // namespace vk
