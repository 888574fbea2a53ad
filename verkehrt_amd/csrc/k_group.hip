// group_id_<t> and group_id2_<a>_<b> (GroupId / GroupId2, src/kernels/scalar/group.rs:
// 259-301): the dense i64 id of each row's key, numbered from 0 by first occurrence. The ids
// depend only on the order in which distinct keys first appear (group.rs:1-5), so the table
// layout is free.
//
// MI355X form, four passes:
//  1. insert: every row claims or finds its key's slot in an open-addressing table of
//     u64 (min row + 1) words, claiming with compare-and-swap and lowering the word with
//     atomicMin, so the word ends at the key's first row + 1;
//  2. probe: every row finds its slot again and flags itself if it is its key's first row;
//  3. the flagged rows, in row order (the selection kernel's `where`), are the keys in
//     first-occurrence order: the slot of the j-th first row gets id j;
//  4. every row reads the id of its slot.
// Keys compare by their zero-extended bit patterns (Column::bits, group.rs:84-116).
#include "vk_device.h"
#include "vk_internal.h"

namespace vk {
namespace {

struct KeyCol {
  int32_t type;
  const void *p;
  __device__ __forceinline__ uint64_t bits(uint64_t r) const {
    switch (type_width_d(type)) {
      case 8: return static_cast<const uint64_t *>(p)[r];
      case 4: return static_cast<const uint32_t *>(p)[r];
      default: return static_cast<const uint8_t *>(p)[r];
    }
  }
  __device__ static __forceinline__ int type_width_d(int32_t t) {
    return (t == VK_F64 || t == VK_I64 || t == VK_U64) ? 8 : (t == VK_F32 || t == VK_I32) ? 4 : 1;
  }
};

__device__ __forceinline__ uint64_t mix(uint64_t x) {
  return (x ^ (x >> 32)) * 0x9E3779B97F4A7C15ull;
}

template <bool PAIR>
__device__ __forceinline__ uint64_t slot_of(const KeyCol &a, const KeyCol &b, uint64_t ka,
                                            uint64_t kb, int shift) {
  const uint64_t h = PAIR ? mix(mix(ka) ^ kb) : mix(ka);
  return h >> shift;
}

template <bool PAIR>
__global__ __launch_bounds__(kBlock) void k_insert(KeyCol a, KeyCol b, uint64_t n,
                                                   unsigned long long *table, uint64_t mask,
                                                   int shift) {
  for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
       r += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t ka = a.bits(r), kb = PAIR ? b.bits(r) : 0;
    uint64_t h = slot_of<PAIR>(a, b, ka, kb, shift);
    while (true) {
      unsigned long long s = __hip_atomic_load(&table[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (s == 0) {
        const unsigned long long old = atomicCAS(&table[h], 0ull, r + 1);
        if (old == 0) break;
        s = old;
      }
      const uint64_t rep = s - 1;
      if (a.bits(rep) == ka && (!PAIR || b.bits(rep) == kb)) {
        if (r + 1 < s) atomicMin(&table[h], r + 1);
        break;
      }
      h = (h + 1) & mask;
    }
  }
}

template <bool PAIR>
__global__ __launch_bounds__(kBlock) void k_probe(KeyCol a, KeyCol b, uint64_t n,
                                                  const unsigned long long *table, uint64_t mask,
                                                  int shift, uint32_t *slot, uint8_t *first) {
  for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
       r += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t ka = a.bits(r), kb = PAIR ? b.bits(r) : 0;
    uint64_t h = slot_of<PAIR>(a, b, ka, kb, shift);
    while (true) {
      const uint64_t w = table[h];
      if (w == 0) {  // unreachable: every key was inserted on its own probe path
        slot[r] = 0;
        first[r] = 0;
        break;
      }
      const uint64_t rep = w - 1;
      if (a.bits(rep) == ka && (!PAIR || b.bits(rep) == kb)) {
        slot[r] = static_cast<uint32_t>(h);
        first[r] = rep == r;
        break;
      }
      h = (h + 1) & mask;
    }
  }
}

__global__ void k_number(const int64_t *__restrict__ firsts, uint64_t groups,
                         const uint32_t *__restrict__ slot, uint32_t *id_of_slot) {
  for (uint64_t j = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; j < groups;
       j += uint64_t(gridDim.x) * blockDim.x)
    id_of_slot[slot[firsts[j]]] = static_cast<uint32_t>(j);
}

__global__ void k_ids(const uint32_t *__restrict__ slot, const uint32_t *__restrict__ id_of_slot,
                      uint64_t n, int64_t *out) {
  for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
       r += uint64_t(gridDim.x) * blockDim.x)
    out[r] = id_of_slot[slot[r]];
}

// ------------------------------------------------------------------ batched form
// The same four passes for many arrays at once: each item has its own table (words aux2 ..
// aux2 + 2^aux3 of one allocation) and its own rows in the concatenated slot and first-flag
// buffers (from row aux); one launch per pass covers the tiles of every item.
constexpr uint64_t kGTile = 4096;

struct ItemRows {
  uint64_t t0, t1;
  __device__ ItemRows(const BatchItem &I, uint32_t tile) {
    t0 = (tile - I.tile0) * kGTile;
    t1 = t0 + kGTile < I.n ? t0 + kGTile : I.n;
  }
};

__global__ __launch_bounds__(kBlock) void k_insert_batch(int32_t type, const BatchItem *items,
                                                         const uint32_t *tile_item,
                                                         unsigned long long *tables) {
  const BatchItem &I = items[tile_item[blockIdx.x]];
  const ItemRows R(I, blockIdx.x);
  const KeyCol a{type, I.a};
  unsigned long long *table = tables + I.aux2;
  const int lg = static_cast<int>(I.aux3);
  const uint64_t mask = (uint64_t(1) << lg) - 1;
  for (uint64_t r = R.t0 + threadIdx.x; r < R.t1; r += kBlock) {
    const uint64_t ka = a.bits(r);
    uint64_t h = slot_of<false>(a, a, ka, 0, 64 - lg);
    while (true) {
      unsigned long long s = __hip_atomic_load(&table[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (s == 0) {
        const unsigned long long old = atomicCAS(&table[h], 0ull, r + 1);
        if (old == 0) break;
        s = old;
      }
      if (a.bits(s - 1) == ka) {
        if (r + 1 < s) atomicMin(&table[h], r + 1);
        break;
      }
      h = (h + 1) & mask;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_probe_batch(int32_t type, const BatchItem *items,
                                                        const uint32_t *tile_item,
                                                        const unsigned long long *tables,
                                                        uint32_t *slot, uint8_t *first) {
  const BatchItem &I = items[tile_item[blockIdx.x]];
  const ItemRows R(I, blockIdx.x);
  const KeyCol a{type, I.a};
  const unsigned long long *table = tables + I.aux2;
  const int lg = static_cast<int>(I.aux3);
  const uint64_t mask = (uint64_t(1) << lg) - 1;
  for (uint64_t r = R.t0 + threadIdx.x; r < R.t1; r += kBlock) {
    const uint64_t ka = a.bits(r);
    uint64_t h = slot_of<false>(a, a, ka, 0, 64 - lg);
    while (true) {
      const uint64_t w = table[h];
      if (w == 0) {  // unreachable: every key was inserted on its own probe path
        slot[I.aux + r] = 0;
        first[I.aux + r] = 0;
        break;
      }
      if (a.bits(w - 1) == ka) {
        slot[I.aux + r] = static_cast<uint32_t>(h);
        first[I.aux + r] = (w - 1) == r;
        break;
      }
      h = (h + 1) & mask;
    }
  }
}

// Row j < groups of an item is the item's j-th first row: its slot gets id j.
__global__ __launch_bounds__(kBlock) void k_number_batch(const BatchItem *items, const uint32_t *tile_item,
                                                         const int64_t *firsts, const uint64_t *groups,
                                                         const uint32_t *slot, unsigned long long *tables) {
  const uint32_t it = tile_item[blockIdx.x];
  const BatchItem &I = items[it];
  const ItemRows R(I, blockIdx.x);
  const uint64_t g = groups[it];
  uint32_t *id_of_slot = reinterpret_cast<uint32_t *>(tables + I.aux2);
  for (uint64_t j = R.t0 + threadIdx.x; j < R.t1 && j < g; j += kBlock)
    id_of_slot[slot[I.aux + firsts[I.aux + j]]] = static_cast<uint32_t>(j);
}

__global__ __launch_bounds__(kBlock) void k_ids_batch(const BatchItem *items, const uint32_t *tile_item,
                                                      const uint32_t *slot,
                                                      const unsigned long long *tables) {
  const BatchItem &I = items[tile_item[blockIdx.x]];
  const ItemRows R(I, blockIdx.x);
  const uint32_t *id_of_slot = reinterpret_cast<const uint32_t *>(tables + I.aux2);
  int64_t *out = static_cast<int64_t *>(I.out);
  for (uint64_t r = R.t0 + threadIdx.x; r < R.t1; r += kBlock) out[r] = id_of_slot[slot[I.aux + r]];
}

}  // namespace

void group_id_batch(Ctx &c, int32_t type, std::vector<BatchItem> &items, uint64_t *d_groups) {
  uint64_t total = 0, words = 0;
  for (auto &it : items) {
    if (it.n > 0x7FFFFFFFull) usage_error("group_id supports at most 2^31 - 1 rows");
    uint64_t slots = 2;
    int lg = 1;
    while (slots < 2 * it.n) {
      slots <<= 1;
      ++lg;
    }
    it.aux = total;
    it.aux2 = words;
    it.aux3 = static_cast<uint64_t>(lg);
    total += it.n;
    words += slots;
  }
  const uint64_t nitems = items.size();
  fill_u64(c, d_groups, nitems, 0);
  if (total == 0) return;
  auto *tables = static_cast<unsigned long long *>(tmp_alloc(c, words * 8));
  fill_u64(c, reinterpret_cast<uint64_t *>(tables), words, 0);
  auto *slot = static_cast<uint32_t *>(tmp_alloc(c, total * 4));
  auto *first = static_cast<uint8_t *>(tmp_alloc(c, total));
  auto *firsts = static_cast<int64_t *>(tmp_alloc(c, total * 8));
  BatchPlan p = batch_plan(c, items, kGTile);
  k_insert_batch<<<p.ntiles, kBlock, 0, c.stream>>>(type, p.items, p.tile_item, tables);
  k_probe_batch<<<p.ntiles, kBlock, 0, c.stream>>>(type, p.items, p.tile_item, tables, slot, first);
  check_hip(hipGetLastError(), "group_id batch table");
  // each item's first rows in row order
  std::vector<BatchItem> w(nitems);
  for (uint64_t i = 0; i < nitems; ++i) {
    w[i] = BatchItem{};
    w[i].a = first + items[i].aux;
    w[i].out = firsts + items[i].aux;
    w[i].n = items[i].n;
  }
  SelectSpec ss{2, VK_I64, 0, 0, 0};
  select_batch(c, ss, w, d_groups);
  k_number_batch<<<p.ntiles, kBlock, 0, c.stream>>>(p.items, p.tile_item, firsts, d_groups, slot, tables);
  k_ids_batch<<<p.ntiles, kBlock, 0, c.stream>>>(p.items, p.tile_item, slot, tables);
  check_hip(hipGetLastError(), "group_id batch ids");
  batch_plan_free(c, p);
  tmp_free(c, tables);
  tmp_free(c, slot);
  tmp_free(c, first);
  tmp_free(c, firsts);
}

void group_id(Ctx &c, int32_t ta, const void *a, int32_t tb, const void *b, uint64_t n,
              int64_t *out) {
  if (n == 0) return;
  if (n > 0x7FFFFFFFull) usage_error("group_id supports at most 2^31 - 1 rows");
  uint64_t slots = 2;
  int lg = 1;
  while (slots < 2 * n) {
    slots <<= 1;
    ++lg;
  }
  auto *table = static_cast<unsigned long long *>(tmp_alloc(c, slots * 8));
  fill_u64(c, reinterpret_cast<uint64_t *>(table), slots, 0);
  auto *slot = static_cast<uint32_t *>(tmp_alloc(c, n * 4));
  auto *first = static_cast<uint8_t *>(tmp_alloc(c, n));
  KeyCol ka{ta, a}, kb{tb, b};
  const uint32_t g = stream_grid(n, kBlock * 4, 8192);
  const int shift = 64 - lg;
  if (b) {
    k_insert<true><<<g, kBlock, 0, c.stream>>>(ka, kb, n, table, slots - 1, shift);
    k_probe<true><<<g, kBlock, 0, c.stream>>>(ka, kb, n, table, slots - 1, shift, slot, first);
  } else {
    k_insert<false><<<g, kBlock, 0, c.stream>>>(ka, kb, n, table, slots - 1, shift);
    k_probe<false><<<g, kBlock, 0, c.stream>>>(ka, kb, n, table, slots - 1, shift, slot, first);
  }
  check_hip(hipGetLastError(), "group_id table");
  // first rows in row order
  auto *firsts = static_cast<int64_t *>(tmp_alloc(c, n * 8));
  uint64_t *d_cnt = c.scratch.slots + 21;
  SelectSpec ss{2, VK_I64, 0, 0, 0};
  select_rows(c, ss, nullptr, first, n, firsts, d_cnt);
  uint64_t groups = 0;
  read_back(c, d_cnt, 8, &groups);
  // reuse the table words as the id of each slot (u32)
  auto *id_of_slot = reinterpret_cast<uint32_t *>(table);
  k_number<<<stream_grid(groups, kBlock, 8192), kBlock, 0, c.stream>>>(firsts, groups, slot, id_of_slot);
  k_ids<<<g, kBlock, 0, c.stream>>>(slot, id_of_slot, n, out);
  check_hip(hipGetLastError(), "group_id ids");
  tmp_free(c, table);
  tmp_free(c, slot);
  tmp_free(c, first);
  tmp_free(c, firsts);
}

}  // namespace vk
