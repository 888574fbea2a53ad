// group_id_<t> and group_id2_<a>_<b> (GroupId / GroupId2, src/kernels/scalar/group.rs:
// 259-301): the dense i64 id of each row's key, numbered from 0 by first occurrence. The ids
// depend only on the order in which distinct keys first appear (group.rs:1-5), so the table
// layout is free.
//
// The two-key form and the batched form, four passes over a global table (the one-key form
// puts a per-block LDS table in front of it; see the single-key section):
//  1. insert: every row claims or finds its key's slot in an open-addressing table of
//     u64 (min row + 1) words, claiming with compare-and-swap and lowering the word with
//     atomicMin, so the word ends at the key's first row + 1;
//  2. probe: every row finds its slot again and flags itself if it is its key's first row;
//  3. the flagged rows, in row order (the selection kernel's `where`), are the keys in
//     first-occurrence order: the slot of the j-th first row gets id j;
//  4. every row reads the id of its slot.
// Keys compare by their zero-extended bit patterns (Column::bits, group.rs:84-116).
#include <cmath>

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
  // Calls f(load) with the load of this column's width fixed, so that a loop inside f issues
  // all its loads together (a width switch per row would sit between them).
  template <typename F>
  __device__ __forceinline__ void with_bits(F &&f) const {
    switch (type_width_d(type)) {
      case 8: {
        const uint64_t *q = static_cast<const uint64_t *>(p);
        f([q](uint64_t r) -> uint64_t { return q[r]; });
        break;
      }
      case 4: {
        const uint32_t *q = static_cast<const uint32_t *>(p);
        f([q](uint64_t r) -> uint64_t { return q[r]; });
        break;
      }
      default: {
        const uint8_t *q = static_cast<const uint8_t *>(p);
        f([q](uint64_t r) -> uint64_t { return q[r]; });
        break;
      }
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

// ------------------------------------------------------------------ single-key form
// group_id_<t> over one key column, in five passes around a per-block LDS hash table that sits
// in front of the global table, so that the rows of a column with few distinct keys stay on
// the compute unit:
//  1. insert (k_insert1): rows fold into the block's LDS table (key -> least row); a key the
//     LDS table cannot take (its new-entry budget spent, a probe run too long, or the key equal
//     to the empty marker) goes straight to the global table, and at the end each LDS entry
//     is inserted into the global table once. The global table's slots hold the key itself
//     (claimed by compare-and-swap from the empty marker), the key's least row (atomicMin)
//     and, later, its id, in one 16-byte entry: a search step reads one line and the least-row
//     update hits the same line. The key equal to the empty marker has a slot of its own past
//     the table's end. The table starts at kSmallSlots slots; a probe run longer than kMaxRun
//     there marks overflow, and the pass is redone on 2n slots;
//  2-4. on the small table (at most kSmallSlots slots, so at most that many keys) the used
//     slots list themselves with their least rows (k_used_slots), the list is sorted by row,
//     and the slot at place j gets id j (k_number_slots). On the full table every used slot
//     flags its least row (k_first_flags), the flagged rows in row order (the selection
//     kernel) are the keys' first rows, and the slot of the j-th first row's key gets id j
//     (k_number1);
//  5. every row maps its key to its id through the LDS table again, each key new to the block
//     resolved once from the global table, and writes it (k_ids1); with more than kLdsGroups
//     keys, where LDS tables seldom hit, every row searches the global table (k_ids2).
constexpr int kGBlock = 1024;
constexpr int kGRows = 8;  // rows per thread per chunk
constexpr uint64_t kGChunk = uint64_t(kGBlock) * kGRows;
constexpr int kLBits = 12;
constexpr int kLSlots = 1 << kLBits;  // LDS entries: 32 KB of keys, 16 KB of values
constexpr int kLProbes = 32;
constexpr uint64_t kLEmpty = 0xFFF8DEADBEEF1357ull;  // LDS empty marker
constexpr int kSmallLg = 21;
constexpr uint64_t kSmallSlots = uint64_t(1) << kSmallLg;
constexpr int kMaxRun = 64;  // longest probe run allowed in the small table
constexpr int kNoLimit = 0x7FFFFFFF;
constexpr uint64_t kLdsGroups = uint64_t(1) << 16;  // more keys than this: k_ids2 instead of k_ids1

// The LDS entry of key k in a table of 2^BITS entries, or -1. An empty entry is claimed
// (*claimed set) while fewer than half the entries are taken; after that only existing entries
// are found.
template <int BITS = kLBits>
__device__ __forceinline__ int lds_find(unsigned long long *s_key, uint32_t *s_cnt, uint64_t k, bool *claimed) {
  constexpr int kSlots = 1 << BITS;
  *claimed = false;
  if (k == kLEmpty) return -1;
  int h = static_cast<int>(mix(k) >> (64 - BITS));
  for (int p = 0; p < kLProbes; ++p) {
    const unsigned long long cur = __hip_atomic_load(&s_key[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (cur == k) return h;
    if (cur == kLEmpty) {
      // the count is read only by rows about to claim: rows of known keys never touch it
      if (__hip_atomic_load(s_cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= uint32_t(kSlots / 2))
        return -1;
      const unsigned long long old = atomicCAS(&s_key[h], kLEmpty, static_cast<unsigned long long>(k));
      if (old == kLEmpty) {
        *claimed = true;
        atomicAdd(s_cnt, 1u);
        return h;
      }
      if (old == k) return h;
    }
    h = (h + 1) & (kSlots - 1);
  }
  return -1;
}

// A slot of the single-key global table: its key (kLEmpty when unused), the key's least row and
// its id. Slot `slots` (past the end) belongs to the key equal to kLEmpty.
struct alignas(16) Slot {
  unsigned long long key;
  uint32_t row;
  uint32_t id;
};

__global__ void k_slots_init(Slot *t, uint64_t n) {
  for (uint64_t h = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; h < n; h += uint64_t(gridDim.x) * blockDim.x)
    t[h] = Slot{kLEmpty, 0xFFFFFFFFu, 0u};
}

// Inserts (key k, row r) into the global table of mask + 1 slots: finds or claims the slot of
// k and lowers its least row to r. False when the probe run reaches max_run slots.
__device__ bool global_insert(Slot *t, uint64_t mask, int shift, uint64_t k, uint32_t r, int max_run) {
  if (k == kLEmpty) {
    atomicMin(&t[mask + 1].row, r);
    return true;
  }
  uint64_t h = mix(k) >> shift;
  for (int p = 0; p < max_run; ++p) {
    unsigned long long cur = __hip_atomic_load(&t[h].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == kLEmpty) cur = atomicCAS(&t[h].key, kLEmpty, static_cast<unsigned long long>(k));
    if (cur == kLEmpty || cur == k) {
      // the least row only decreases, so a value read at any time that is already at most r
      // needs no atomic (mostly so: rows arrive in increasing order per thread)
      if (__hip_atomic_load(&t[h].row, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > r) atomicMin(&t[h].row, r);
      return true;
    }
    h = (h + 1) & mask;
  }
  return false;
}

__global__ __launch_bounds__(kGBlock) void k_insert1(KeyCol a, uint64_t n, Slot *table, uint64_t mask, int shift,
                                                     int max_run, uint32_t *overflow) {
  __shared__ unsigned long long s_key[kLSlots];
  __shared__ uint32_t s_row[kLSlots];
  __shared__ uint32_t s_cnt;
  const int t = threadIdx.x;
  for (int i = t; i < kLSlots; i += kGBlock) {
    s_key[i] = kLEmpty;
    s_row[i] = 0xFFFFFFFFu;
  }
  if (t == 0) s_cnt = 0;
  __syncthreads();
  bool over = false;
  for (uint64_t c0 = blockIdx.x * kGChunk; c0 < n; c0 += uint64_t(gridDim.x) * kGChunk) {
    uint64_t k[kGRows];
    a.with_bits([&](auto bits) {
#pragma unroll
      for (int j = 0; j < kGRows; ++j) {
        const uint64_t r = c0 + uint64_t(j) * kGBlock + t;
        k[j] = r < n ? bits(r) : 0;
      }
    });
    if (max_run != kNoLimit) {
      // leave early once any block has overflowed (one read per wave, issued after the
      // chunk's loads so that its round trip overlaps theirs)
      uint32_t o = 0;
      if (lane_id() == 0) o = __hip_atomic_load(overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (__shfl(o, 0) != 0) {
        over = true;
        break;
      }
    }
#pragma unroll
    for (int j = 0; j < kGRows; ++j) {
      const uint64_t r = c0 + uint64_t(j) * kGBlock + t;
      if (r >= n) continue;
      bool claimed;
      const int h = lds_find(s_key, &s_cnt, k[j], &claimed);
      if (h >= 0) {
        // rows arrive in increasing order per thread, so the entry is mostly lower already
        if (__hip_atomic_load(&s_row[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) > r)
          atomicMin(&s_row[h], static_cast<uint32_t>(r));
      } else if (!global_insert(table, mask, shift, k[j], static_cast<uint32_t>(r), max_run)) {
        over = true;
      }
    }
    if (over) atomicOr(overflow, 1u);
    over = false;
  }
  __syncthreads();
  if (max_run != kNoLimit && __hip_atomic_load(overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
    return;  // the pass is redone on the full table
  for (int i = t; i < kLSlots; i += kGBlock) {
    const uint64_t k = s_key[i];
    if (k != kLEmpty && !global_insert(table, mask, shift, k, s_row[i], max_run)) atomicOr(overflow, 1u);
  }
}

// A HyperLogLog sketch of the distinct keys (4096 registers, standard error about 1.6%): each
// row's key is hashed (a 64-bit finalizer), its top 12 bits pick a register and the register
// keeps the most leading zeros, plus one, seen in the remaining bits; per-block LDS registers,
// merged into reg[] with one atomicMax per nonzero register.
constexpr int kHllBits = 12;
constexpr int kHllRegs = 1 << kHllBits;

__device__ __forceinline__ uint64_t hll_hash(uint64_t x) {
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ __launch_bounds__(kBlock) void k_hll(KeyCol a, uint64_t n, uint32_t *reg) {
  __shared__ uint32_t s_reg[kHllRegs];
  for (int i = threadIdx.x; i < kHllRegs; i += kBlock) s_reg[i] = 0;
  __syncthreads();
  a.with_bits([&](auto bits) {
    for (uint64_t r0 = blockIdx.x * uint64_t(kBlock) * 4; r0 < n; r0 += uint64_t(gridDim.x) * kBlock * 4) {
      uint64_t k[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint64_t r = r0 + uint64_t(u) * kBlock + threadIdx.x;
        k[u] = r < n ? bits(r) : 0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint64_t r = r0 + uint64_t(u) * kBlock + threadIdx.x;
        if (r >= n) continue;
        const uint64_t h = hll_hash(k[u]);
        const uint32_t idx = static_cast<uint32_t>(h >> (64 - kHllBits));
        const uint64_t rest = h << kHllBits;
        const uint32_t rho = rest ? static_cast<uint32_t>(__clzll(static_cast<long long>(rest))) + 1u
                                  : static_cast<uint32_t>(64 - kHllBits + 1);
        if (s_reg[idx] < rho) atomicMax(&s_reg[idx], rho);
      }
    }
  });
  __syncthreads();
  for (int i = threadIdx.x; i < kHllRegs; i += kBlock)
    if (s_reg[i]) atomicMax(&reg[i], s_reg[i]);
}

// Slots 0 .. slots (the last one the empty marker's own): a used slot flags its least row.
__global__ void k_first_flags(const Slot *__restrict__ table, uint64_t slots, uint8_t *first) {
  for (uint64_t h = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; h <= slots;
       h += uint64_t(gridDim.x) * blockDim.x) {
    const Slot e = table[h];
    if (h < slots ? e.key != kLEmpty : e.row != 0xFFFFFFFFu) first[e.row] = 1;
  }
}

// The slot of key k (every key is in the table).
__device__ __forceinline__ uint64_t slot_of_key(const Slot *table, uint64_t mask, int shift, uint64_t k) {
  if (k == kLEmpty) return mask + 1;
  uint64_t h = mix(k) >> shift;
  for (uint64_t p = 0; p <= mask; ++p) {
    if (table[h].key == k) return h;
    h = (h + 1) & mask;
  }
  return mask + 1;  // unreachable: every key has a slot on its probe path
}

// The slot of the j-th first row's key gets id j.
__global__ void k_number1(KeyCol a, const int64_t *__restrict__ firsts, uint64_t groups, Slot *table, uint64_t mask,
                          int shift) {
  for (uint64_t j = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; j < groups;
       j += uint64_t(gridDim.x) * blockDim.x)
    table[slot_of_key(table, mask, shift, a.bits(static_cast<uint64_t>(firsts[j])))].id = static_cast<uint32_t>(j);
}

// Every used slot of a table of slots + 1 entries lists its least row and its index, in no
// particular order; *count ends at the number of keys. One atomic per block places its list.
constexpr int kUsedPer = 16;
__global__ __launch_bounds__(kBlock) void k_used_slots(const Slot *__restrict__ table, uint64_t slots,
                                                       uint32_t *rows, uint64_t *where,
                                                       unsigned long long *count) {
  __shared__ uint32_t s_wave[kBlock / kWave];
  __shared__ unsigned long long s_base;
  const uint64_t base = blockIdx.x * uint64_t(kBlock * kUsedPer) + threadIdx.x;
  uint32_t row[kUsedPer];
  uint32_t used = 0, cnt = 0;
#pragma unroll
  for (int i = 0; i < kUsedPer; ++i) {
    const uint64_t h = base + uint64_t(i) * kBlock;
    row[i] = 0xFFFFFFFFu;
    if (h <= slots) {
      const Slot e = table[h];
      if (h < slots ? e.key != kLEmpty : e.row != 0xFFFFFFFFu) row[i] = e.row;
    }
  }
#pragma unroll
  for (int i = 0; i < kUsedPer; ++i) {
    const uint32_t u = row[i] != 0xFFFFFFFFu;
    used |= u << i;
    cnt += u;
  }
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  uint32_t inc = cnt;
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t o = __shfl_up(inc, d);
    if (lane >= d) inc += o;
  }
  if (lane == kWave - 1) s_wave[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < kBlock / kWave; ++w) {
      const uint32_t x = s_wave[w];
      s_wave[w] = t;
      t += x;
    }
    s_base = t ? atomicAdd(count, static_cast<unsigned long long>(t)) : 0;
  }
  __syncthreads();
  uint64_t pos = s_base + s_wave[wave] + inc - cnt;
#pragma unroll
  for (int i = 0; i < kUsedPer; ++i)
    if (used >> i & 1) {
      rows[pos] = row[i];
      where[pos] = base + uint64_t(i) * kBlock;
      ++pos;
    }
}

// The slot of the j-th least first row gets id j.
__global__ void k_number_slots(const uint64_t *__restrict__ where, uint64_t groups, Slot *table) {
  for (uint64_t j = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; j < groups;
       j += uint64_t(gridDim.x) * blockDim.x)
    table[where[j]].id = static_cast<uint32_t>(j);
}

__device__ __forceinline__ uint32_t global_id(const Slot *table, uint64_t mask, int shift, uint64_t k) {
  if (k == kLEmpty) return table[mask + 1].id;
  uint64_t h = mix(k) >> shift;
  for (uint64_t p = 0; p <= mask; ++p) {
    const Slot e = table[h];
    if (e.key == k) return e.id;
    h = (h + 1) & mask;
  }
  return 0;  // unreachable: every key has a slot on its probe path
}

// Every row's id from the global table, for more keys than the LDS tables hold.
__global__ __launch_bounds__(kBlock) void k_ids2(KeyCol a, uint64_t n, const Slot *table, uint64_t mask, int shift,
                                                 int64_t *out) {
  for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
       r += uint64_t(gridDim.x) * blockDim.x)
    out[r] = global_id(table, mask, shift, a.bits(r));
}

// Phase 1 of each chunk finds (or claims) the LDS entry of every row's key, the claiming lane
// resolving the key's id; after the barrier, phase 2 reads the ids. An entry's id is written
// once, before the barrier that follows its claim, and never changes.
__global__ __launch_bounds__(kGBlock) void k_ids1(KeyCol a, uint64_t n, const Slot *table, uint64_t mask,
                                                  int shift, int64_t *out) {
  __shared__ unsigned long long s_key[kLSlots];
  __shared__ uint32_t s_id[kLSlots];
  __shared__ uint32_t s_cnt;
  const int t = threadIdx.x;
  for (int i = t; i < kLSlots; i += kGBlock) s_key[i] = kLEmpty;
  if (t == 0) s_cnt = 0;
  __syncthreads();
  // the next chunk's keys load while this chunk is resolved and written
  const uint64_t stride = uint64_t(gridDim.x) * kGChunk;
  uint64_t kn[kGRows];
  auto load = [&](uint64_t c) {
    a.with_bits([&](auto bits) {
#pragma unroll
      for (int j = 0; j < kGRows; ++j) {
        const uint64_t r = c + uint64_t(j) * kGBlock + t;
        kn[j] = r < n ? bits(r) : 0;
      }
    });
  };
  if (blockIdx.x * kGChunk < n) load(blockIdx.x * kGChunk);
  for (uint64_t c0 = blockIdx.x * kGChunk; c0 < n; c0 += stride) {
    uint64_t k[kGRows];
    int h[kGRows];
#pragma unroll
    for (int j = 0; j < kGRows; ++j) k[j] = kn[j];
    if (c0 + stride < n) load(c0 + stride);
#pragma unroll
    for (int j = 0; j < kGRows; ++j) {
      const uint64_t r = c0 + uint64_t(j) * kGBlock + t;
      h[j] = -1;
      if (r >= n) continue;
      bool claimed;
      h[j] = lds_find(s_key, &s_cnt, k[j], &claimed);
      if (claimed) s_id[h[j]] = global_id(table, mask, shift, k[j]);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kGRows; ++j) {
      const uint64_t r = c0 + uint64_t(j) * kGBlock + t;
      if (r >= n) continue;
      out[r] = h[j] >= 0 ? s_id[h[j]] : global_id(table, mask, shift, k[j]);
    }
  }
}

// ------------------------------------------------------------------ batched form
// The same four passes for many arrays at once: each item has its own table (2^aux3 words at
// b) and its own rows in the concatenated slot and first-flag buffers (from row aux); one launch
// per pass covers the tiles of every item. Tables start small (at most 2^kBSmallLg words): an
// insert whose probe run there passes kMaxRun marks its item's overflow word, and only those
// items insert again into tables of 2n words before the other passes run.
// A block takes a tile of kGTile rows in sub-tiles of kGSub (16 rows per thread): the keys a
// tile repeats reach the item's table once per tile, through the block's LDS table.
constexpr uint64_t kGSub = 4096;
constexpr uint64_t kGTile = 4 * kGSub;

struct ItemRows {
  uint64_t t0, t1;
  __device__ ItemRows(const BatchItem &I, uint32_t tile) {
    t0 = (tile - I.tile0) * kGTile;
    t1 = t0 + kGTile < I.n ? t0 + kGTile : I.n;
  }
};

// The batched tables hold u64 words (first row + 1); a key is compared through the column.
// Inserts (key k, row r): claims an empty slot with r + 1 or lowers the word of k's slot.
// Returns false when the probe run passes max_run slots (the key is not inserted).
__device__ bool word_insert(const KeyCol &a, unsigned long long *table, uint64_t mask, int shift, uint64_t k,
                            uint64_t r, uint64_t max_run) {
  uint64_t h = mix(k) >> shift;
  for (uint64_t p = 0; p <= mask; ++p) {
    if (p > max_run) return false;
    unsigned long long s = __hip_atomic_load(&table[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (s == 0) {
      const unsigned long long old = atomicCAS(&table[h], 0ull, r + 1);
      if (old == 0) return true;
      s = old;
    }
    if (a.bits(s - 1) == k) {
      if (r + 1 < s) atomicMin(&table[h], r + 1);
      return true;
    }
    h = (h + 1) & mask;
  }
  return false;
}

// The slot of key k (every key is in the table).
__device__ uint64_t word_slot(const KeyCol &a, const unsigned long long *table, uint64_t mask, int shift,
                              uint64_t k) {
  uint64_t h = mix(k) >> shift;
  for (uint64_t p = 0; p <= mask; ++p) {
    const uint64_t w = table[h];
    if (w != 0 && a.bits(w - 1) == k) return h;
    h = (h + 1) & mask;
  }
  return 0;  // unreachable: every key was inserted on its own probe path
}

// A tile of an item (kGTile rows, 16 per thread) goes through an LDS table of 2^kBLBits keys
// first, as in the single-key form: rows of a key the tile has seen stay in LDS, and each of
// the tile's keys reaches the item's table once.
constexpr int kBLBits = 11;
constexpr int kBLSlots = 1 << kBLBits;
constexpr int kBRows = static_cast<int>(kGSub / kBlock);

// aux2: the item's longest probe run (kNoLimit in a table of 2n words); overflow[item] is set
// when a run passes it. With `only`, items whose word there is zero are skipped (a plan's
// second insert, into full tables, of the items whose small table overflowed).
__global__ __launch_bounds__(kBlock) void k_insert_batch(int32_t type, const BatchItem *items,
                                                         const uint32_t *tile_item, uint32_t *overflow,
                                                         const uint32_t *only) {
  __shared__ unsigned long long s_key[kBLSlots];
  __shared__ uint32_t s_row[kBLSlots];
  __shared__ uint32_t s_cnt;
  const uint32_t it = tile_item[blockIdx.x];
  if (only && only[it] == 0) return;  // uniform over the block
  const BatchItem &I = items[it];
  const ItemRows R(I, blockIdx.x);
  const KeyCol a{type, I.a};
  unsigned long long *table = static_cast<unsigned long long *>(const_cast<void *>(I.b));
  const int lg = static_cast<int>(I.aux3);
  const uint64_t mask = (uint64_t(1) << lg) - 1;
  const uint64_t max_run = I.aux2;
  const int t = threadIdx.x;
  bool over = false;
  for (int i = t; i < kBLSlots; i += kBlock) {
    s_key[i] = kLEmpty;
    s_row[i] = 0xFFFFFFFFu;
  }
  if (t == 0) s_cnt = 0;
  __syncthreads();
  for (uint64_t s0 = R.t0; s0 < R.t1; s0 += kGSub) {
    uint64_t k[kBRows];
    a.with_bits([&](auto bits) {
#pragma unroll
      for (int j = 0; j < kBRows; ++j) {
        const uint64_t r = s0 + uint64_t(j) * kBlock + t;
        k[j] = r < R.t1 ? bits(r) : 0;
      }
    });
#pragma unroll
    for (int j = 0; j < kBRows; ++j) {
      const uint64_t r = s0 + uint64_t(j) * kBlock + t;
      if (r >= R.t1) continue;
      bool claimed;
      const int h = lds_find<kBLBits>(s_key, &s_cnt, k[j], &claimed);
      if (h >= 0) {
        if (__hip_atomic_load(&s_row[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) > r)
          atomicMin(&s_row[h], static_cast<uint32_t>(r));
      } else {
        if (!word_insert(a, table, mask, 64 - lg, k[j], r, max_run)) over = true;
      }
    }
  }
  __syncthreads();
  for (int i = t; i < kBLSlots; i += kBlock) {
    const uint64_t kk = s_key[i];
    if (kk != kLEmpty && !word_insert(a, table, mask, 64 - lg, kk, s_row[i], max_run)) over = true;
  }
  if (over) atomicOr(&overflow[it], 1u);
}

// Phase 1 finds (or claims) every row's LDS entry, the claiming lane resolving the key's slot
// and first row; after the barrier, phase 2 writes each row's slot and first-row flag.
__global__ __launch_bounds__(kBlock) void k_probe_batch(int32_t type, const BatchItem *items,
                                                        const uint32_t *tile_item,
                                                        uint32_t *slot, uint8_t *first) {
  __shared__ unsigned long long s_key[kBLSlots];
  __shared__ uint32_t s_slot[kBLSlots];
  __shared__ uint32_t s_rep[kBLSlots];
  __shared__ uint32_t s_cnt;
  const BatchItem &I = items[tile_item[blockIdx.x]];
  const ItemRows R(I, blockIdx.x);
  const KeyCol a{type, I.a};
  const unsigned long long *table = static_cast<const unsigned long long *>(I.b);
  const int lg = static_cast<int>(I.aux3);
  const uint64_t mask = (uint64_t(1) << lg) - 1;
  const int t = threadIdx.x;
  for (int i = t; i < kBLSlots; i += kBlock) s_key[i] = kLEmpty;
  if (t == 0) s_cnt = 0;
  __syncthreads();
  // per sub-tile: phase 1, barrier, phase 2 (an entry claimed in an earlier sub-tile was
  // resolved before that sub-tile's barrier; phase 1 writes only entries it claims)
  for (uint64_t s0 = R.t0; s0 < R.t1; s0 += kGSub) {
    uint64_t k[kBRows];
    int h[kBRows];
    a.with_bits([&](auto bits) {
#pragma unroll
      for (int j = 0; j < kBRows; ++j) {
        const uint64_t r = s0 + uint64_t(j) * kBlock + t;
        k[j] = r < R.t1 ? bits(r) : 0;
      }
    });
#pragma unroll
    for (int j = 0; j < kBRows; ++j) {
      const uint64_t r = s0 + uint64_t(j) * kBlock + t;
      h[j] = -1;
      if (r >= R.t1) continue;
      bool claimed;
      h[j] = lds_find<kBLBits>(s_key, &s_cnt, k[j], &claimed);
      if (claimed) {
        const uint64_t g = word_slot(a, table, mask, 64 - lg, k[j]);
        s_slot[h[j]] = static_cast<uint32_t>(g);
        s_rep[h[j]] = static_cast<uint32_t>(table[g] - 1);
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kBRows; ++j) {
      const uint64_t r = s0 + uint64_t(j) * kBlock + t;
      if (r >= R.t1) continue;
      uint64_t g, rep;
      if (h[j] >= 0) {
        g = s_slot[h[j]];
        rep = s_rep[h[j]];
      } else {
        g = word_slot(a, table, mask, 64 - lg, k[j]);
        rep = table[g] - 1;
      }
      slot[I.aux + r] = static_cast<uint32_t>(g);
      first[I.aux + r] = rep == r;
    }
  }
}

// Row j < groups of an item is the item's j-th first row: its slot gets id j.
__global__ __launch_bounds__(kBlock) void k_number_batch(const BatchItem *items, const uint32_t *tile_item,
                                                         const int64_t *firsts, const uint64_t *groups,
                                                         const uint32_t *slot) {
  const uint32_t it = tile_item[blockIdx.x];
  const BatchItem &I = items[it];
  const ItemRows R(I, blockIdx.x);
  const uint64_t g = groups[it];
  uint32_t *id_of_slot = static_cast<uint32_t *>(const_cast<void *>(I.b));
  for (uint64_t j = R.t0 + threadIdx.x; j < R.t1 && j < g; j += kBlock)
    id_of_slot[slot[I.aux + firsts[I.aux + j]]] = static_cast<uint32_t>(j);
}

// A thread's kBRows rows: every slot load issued first, then every id lookup, then the stores
// (one load round trip per step instead of one per row). Lookups go through a per-block LDS
// cache of 2^kIdsLg entries in front of the item's id table: the entry of slot s sits at
// s mod 2^kIdsLg and holds ((s >> kIdsLg) + 1) << idbits | id in one 32-bit word, so a read
// racing a write from another wave sees either whole entry, and 0 is empty. An item whose
// group count does not fit in idbits bits reads the table directly.
constexpr int kIdsLg = 13;
template <bool CACHE>
__device__ __forceinline__ void ids_rows(const BatchItem &I, const ItemRows &R, const uint32_t *slot,
                                         const uint32_t *id_of_slot, uint32_t *cache, int idbits) {
  const uint32_t *sl = slot + I.aux;
  int64_t *out = static_cast<int64_t *>(I.out);
  const int t = threadIdx.x;
  const uint32_t idmask = (1u << idbits) - 1;
  for (uint64_t s0 = R.t0; s0 < R.t1; s0 += kGSub) {
    uint32_t s[kBRows], id[kBRows];
#pragma unroll
    for (int u = 0; u < kBRows; ++u) {
      const uint64_t r = s0 + uint64_t(u) * kBlock + t;
      s[u] = r < R.t1 ? sl[r] : 0;
    }
    if (CACHE) {
      uint32_t miss = 0;
#pragma unroll
      for (int u = 0; u < kBRows; ++u) {
        const uint32_t e = cache[s[u] & ((1u << kIdsLg) - 1)];
        const bool hit = (e >> idbits) == (s[u] >> kIdsLg) + 1;
        id[u] = e & idmask;
        const uint64_t r = s0 + uint64_t(u) * kBlock + t;
        if (!hit && r < R.t1) miss |= 1u << u;
      }
#pragma unroll
      for (int u = 0; u < kBRows; ++u)
        if (miss >> u & 1) id[u] = id_of_slot[s[u]];
#pragma unroll
      for (int u = 0; u < kBRows; ++u)
        if (miss >> u & 1) cache[s[u] & ((1u << kIdsLg) - 1)] = (((s[u] >> kIdsLg) + 1) << idbits) | id[u];
    } else {
#pragma unroll
      for (int u = 0; u < kBRows; ++u) {
        const uint64_t r = s0 + uint64_t(u) * kBlock + t;
        id[u] = r < R.t1 ? id_of_slot[s[u]] : 0;
      }
    }
#pragma unroll
    for (int u = 0; u < kBRows; ++u) {
      const uint64_t r = s0 + uint64_t(u) * kBlock + t;
      if (r < R.t1) out[r] = id[u];
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_ids_batch(const BatchItem *items, const uint32_t *tile_item,
                                                      const uint32_t *slot, const uint64_t *groups) {
  __shared__ uint32_t s_cache[1 << kIdsLg];
  const uint32_t it = tile_item[blockIdx.x];
  const BatchItem &I = items[it];
  const ItemRows R(I, blockIdx.x);
  const uint32_t *id_of_slot = static_cast<const uint32_t *>(I.b);
  const int lg = static_cast<int>(I.aux3);
  const int idbits = 31 - (lg > kIdsLg ? lg - kIdsLg : 0);
  if (idbits > 0 && groups[it] <= (uint64_t(1) << idbits)) {
    for (int i = threadIdx.x; i < (1 << kIdsLg); i += kBlock) s_cache[i] = 0;
    __syncthreads();
    ids_rows<true>(I, R, slot, id_of_slot, s_cache, idbits);
  } else {
    ids_rows<false>(I, R, slot, id_of_slot, nullptr, 31);
  }
}

}  // namespace

void group_id_batch(Ctx &c, int32_t type, std::vector<BatchItem> &items, uint64_t *d_groups) {
  constexpr int kBSmallLg = 13;  // first tables: at most 8192 words (64 KB) per item
  uint64_t total = 0, small_words = 0;
  std::vector<int> full_lg(items.size());
  for (size_t i = 0; i < items.size(); ++i) {
    BatchItem &it = items[i];
    if (it.n > 0x7FFFFFFFull) usage_error("group_id supports at most 2^31 - 1 rows");
    int lg = 1;
    while ((uint64_t(1) << lg) < 2 * it.n) ++lg;
    full_lg[i] = lg;
    it.aux = total;
    it.aux3 = static_cast<uint64_t>(lg < kBSmallLg ? lg : kBSmallLg);
    total += it.n;
    small_words += uint64_t(1) << it.aux3;
  }
  const uint64_t nitems = items.size();
  fill_u64(c, d_groups, nitems, 0);
  if (total == 0) return;
  auto *small = static_cast<unsigned long long *>(tmp_alloc(c, small_words * 8));
  fill_u64(c, reinterpret_cast<uint64_t *>(small), small_words, 0);
  auto *over = static_cast<uint32_t *>(tmp_alloc(c, nitems * 4 + 8));
  fill_u64(c, reinterpret_cast<uint64_t *>(over), (nitems * 4 + 8) / 8, 0);
  uint64_t w = 0;
  for (size_t i = 0; i < nitems; ++i) {
    items[i].b = small + w;
    items[i].aux2 = int(items[i].aux3) < full_lg[i] ? uint64_t(kMaxRun) : uint64_t(kNoLimit);
    w += uint64_t(1) << items[i].aux3;
  }
  {
    BatchPlan p1 = batch_plan(c, items, kGTile);
    k_insert_batch<<<p1.ntiles, kBlock, 0, c.stream>>>(type, p1.items, p1.tile_item, over, nullptr);
    check_hip(hipGetLastError(), "group_id batch insert");
    batch_plan_free(c, p1);
  }
  // items whose small table overflowed insert again into tables of 2n words
  std::vector<uint32_t> ov(nitems);
  check_hip(hipMemcpyAsync(ov.data(), over, nitems * 4, hipMemcpyDeviceToHost, c.stream), "overflow");
  check_hip(hipStreamSynchronize(c.stream), "sync");
  std::vector<BatchItem> again;
  std::vector<size_t> again_idx;
  uint64_t full_words = 0;
  for (size_t i = 0; i < nitems; ++i)
    if (ov[i]) {
      again_idx.push_back(i);
      full_words += uint64_t(1) << full_lg[i];
    }
  unsigned long long *full = nullptr;
  if (!again_idx.empty()) {
    full = static_cast<unsigned long long *>(tmp_alloc(c, full_words * 8));
    fill_u64(c, reinterpret_cast<uint64_t *>(full), full_words, 0);
    uint64_t f = 0;
    for (size_t i : again_idx) {
      items[i].b = full + f;
      items[i].aux3 = static_cast<uint64_t>(full_lg[i]);
      items[i].aux2 = uint64_t(kNoLimit);
      f += uint64_t(1) << full_lg[i];
      again.push_back(items[i]);
    }
    BatchPlan p2 = batch_plan(c, again, kGTile);
    k_insert_batch<<<p2.ntiles, kBlock, 0, c.stream>>>(type, p2.items, p2.tile_item, over, nullptr);
    check_hip(hipGetLastError(), "group_id batch insert (full tables)");
    batch_plan_free(c, p2);
  }
  auto *slot = static_cast<uint32_t *>(tmp_alloc(c, total * 4));
  auto *first = static_cast<uint8_t *>(tmp_alloc(c, total));
  auto *firsts = static_cast<int64_t *>(tmp_alloc(c, total * 8));
  BatchPlan p = batch_plan(c, items, kGTile);
  k_probe_batch<<<p.ntiles, kBlock, 0, c.stream>>>(type, p.items, p.tile_item, slot, first);
  check_hip(hipGetLastError(), "group_id batch table");
  // each item's first rows in row order
  std::vector<BatchItem> wl(nitems);
  for (uint64_t i = 0; i < nitems; ++i) {
    wl[i] = BatchItem{};
    wl[i].a = first + items[i].aux;
    wl[i].out = firsts + items[i].aux;
    wl[i].n = items[i].n;
  }
  SelectSpec ss{2, VK_I64, 0, 0, 0};
  select_batch(c, ss, wl, d_groups);
  k_number_batch<<<p.ntiles, kBlock, 0, c.stream>>>(p.items, p.tile_item, firsts, d_groups, slot);
  k_ids_batch<<<p.ntiles, kBlock, 0, c.stream>>>(p.items, p.tile_item, slot, d_groups);
  check_hip(hipGetLastError(), "group_id batch ids");
  batch_plan_free(c, p);
  tmp_free(c, small);
  if (full) tmp_free(c, full);
  tmp_free(c, over);
  tmp_free(c, slot);
  tmp_free(c, first);
  tmp_free(c, firsts);
}

namespace {

// Passes 2-4 of the file header over a filled two-key table: probe (per-row slot and first
// flag), the first rows in row order, number, ids.
void ids_from_slots(Ctx &c, KeyCol ka, KeyCol kb, uint64_t n, unsigned long long *table, uint64_t slots,
                    int lg, int64_t *out) {
  auto *slot = static_cast<uint32_t *>(tmp_alloc(c, n * 4));
  auto *first = static_cast<uint8_t *>(tmp_alloc(c, n));
  const uint32_t g = stream_grid(n, kBlock * 4, 8192);
  const int shift = 64 - lg;
  k_probe<true><<<g, kBlock, 0, c.stream>>>(ka, kb, n, table, slots - 1, shift, slot, first);
  check_hip(hipGetLastError(), "group_id probe");
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
  tmp_free(c, slot);
  tmp_free(c, first);
  tmp_free(c, firsts);
}

// The two-key form: the four passes of the file header over the global table alone.
void group_id_pair(Ctx &c, KeyCol ka, KeyCol kb, uint64_t n, int64_t *out) {
  uint64_t slots = 2;
  int lg = 1;
  while (slots < 2 * n) {
    slots <<= 1;
    ++lg;
  }
  auto *table = static_cast<unsigned long long *>(tmp_alloc(c, slots * 8));
  fill_u64(c, reinterpret_cast<uint64_t *>(table), slots, 0);
  const uint32_t g = stream_grid(n, kBlock * 4, 8192);
  k_insert<true><<<g, kBlock, 0, c.stream>>>(ka, kb, n, table, slots - 1, 64 - lg);
  check_hip(hipGetLastError(), "group_id table");
  ids_from_slots(c, ka, kb, n, table, slots, lg, out);
  tmp_free(c, table);
}

}  // namespace

namespace {

constexpr int kSizedRun = 1024;  // longest probe run allowed in a table sized by the sketch

// The sketched count of distinct keys of a (HyperLogLog: the registers' harmonic mean, with
// linear counting while the estimate is small and registers are still zero).
uint64_t distinct_estimate(Ctx &c, const KeyCol &a, uint64_t n) {
  auto *reg = static_cast<uint32_t *>(tmp_alloc(c, kHllRegs * 4));
  fill_u64(c, reinterpret_cast<uint64_t *>(reg), kHllRegs / 2, 0);
  k_hll<<<stream_grid(n, kBlock * 4, 2048), kBlock, 0, c.stream>>>(a, n, reg);
  check_hip(hipGetLastError(), "group_id sketch");
  std::vector<uint32_t> h(kHllRegs);
  check_hip(hipMemcpyAsync(h.data(), reg, kHllRegs * 4, hipMemcpyDeviceToHost, c.stream), "sketch");
  check_hip(hipStreamSynchronize(c.stream), "sync");
  tmp_free(c, reg);
  const double m = kHllRegs;
  double z = 0.0;
  int zeros = 0;
  for (uint32_t v : h) {
    z += std::ldexp(1.0, -static_cast<int>(v));
    zeros += v == 0;
  }
  double e = 0.7213 / (1.0 + 1.079 / m) * m * m / z;
  if (e <= 2.5 * m && zeros) e = m * std::log(m / zeros);
  return static_cast<uint64_t>(e) + 1;
}

}  // namespace

void group_id(Ctx &c, int32_t ta, const void *a, int32_t tb, const void *b, uint64_t n,
              int64_t *out) {
  if (n == 0) return;
  if (n > 0x7FFFFFFFull) usage_error("group_id supports at most 2^31 - 1 rows");
  KeyCol ka{ta, a}, kb{tb, b};
  if (b) {
    group_id_pair(c, ka, kb, n, out);
    return;
  }
  uint64_t full = 2;
  int lg_full = 1;
  while (full < 2 * n) {
    full <<= 1;
    ++lg_full;
  }
  const uint32_t g = static_cast<uint32_t>(std::min<uint64_t>((n + kGChunk - 1) / kGChunk, 512));
  auto *d_over = reinterpret_cast<uint32_t *>(c.scratch.slots + 22);
  auto new_table = [&](uint64_t slots) {  // slots + 1 entries: the last for the empty marker's key
    auto *t = static_cast<Slot *>(tmp_alloc(c, (slots + 1) * sizeof(Slot)));
    k_slots_init<<<stream_grid(slots + 1, kBlock * 4, 8192), kBlock, 0, c.stream>>>(t, slots + 1);
    return t;
  };
  Slot *table = nullptr;
  uint64_t slots = full;
  int lg = lg_full;
  if (full > kSmallSlots) {
    // few distinct keys fit the small table; a long probe run there means too many
    table = new_table(kSmallSlots);
    fill_u64(c, c.scratch.slots + 22, 1, 0);
    k_insert1<<<g, kGBlock, 0, c.stream>>>(ka, n, table, kSmallSlots - 1, 64 - kSmallLg, kMaxRun, d_over);
    check_hip(hipGetLastError(), "group_id insert");
    uint32_t over = 0;
    read_back(c, d_over, 4, &over);
    if (over) {
      tmp_free(c, table);
      table = nullptr;
    } else {
      slots = kSmallSlots;
      lg = kSmallLg;
    }
  }
  if (!table && full > kSmallSlots) {
    // more keys than the small table takes: a table of the next power of two above twice the
    // sketched count of distinct keys, when that is below 2n slots; a probe run past
    // kSizedRun there (an estimate far too low) sends the insert to 2n slots
    const uint64_t est = distinct_estimate(c, ka, n);
    uint64_t sized = kSmallSlots;
    int lg_sized = kSmallLg;
    while (sized < 2 * est) {
      sized <<= 1;
      ++lg_sized;
    }
    if (sized < full) {
      table = new_table(sized);
      fill_u64(c, c.scratch.slots + 22, 1, 0);
      k_insert1<<<g, kGBlock, 0, c.stream>>>(ka, n, table, sized - 1, 64 - lg_sized, kSizedRun, d_over);
      check_hip(hipGetLastError(), "group_id insert");
      uint32_t over = 0;
      read_back(c, d_over, 4, &over);
      if (over) {
        tmp_free(c, table);
        table = nullptr;
      } else {
        slots = sized;
        lg = lg_sized;
      }
    }
  }
  if (!table) {
    table = new_table(full);
    k_insert1<<<g, kGBlock, 0, c.stream>>>(ka, n, table, full - 1, 64 - lg_full, kNoLimit, d_over);
    check_hip(hipGetLastError(), "group_id insert");
  }
  const uint64_t mask = slots - 1;
  const int shift = 64 - lg;
  uint64_t *d_cnt = c.scratch.slots + 21;
  uint64_t groups = 0;
  if (slots <= kSmallSlots) {
    // the used slots listed with their least rows, then numbered in row order
    auto *rows = static_cast<uint32_t *>(tmp_alloc(c, (slots + 1) * 4));
    auto *where = static_cast<uint64_t *>(tmp_alloc(c, (slots + 1) * 8));
    fill_u64(c, d_cnt, 1, 0);
    constexpr uint64_t kPerBlock = uint64_t(kBlock) * kUsedPer;
    k_used_slots<<<static_cast<uint32_t>((slots + kPerBlock) / kPerBlock), kBlock, 0, c.stream>>>(
        table, slots, rows, where, reinterpret_cast<unsigned long long *>(d_cnt));
    check_hip(hipGetLastError(), "group_id slots");
    read_back(c, d_cnt, 8, &groups);
    auto *sorted_rows = tmp_alloc(c, groups * 4);
    auto *sorted_where = static_cast<uint64_t *>(tmp_alloc(c, groups * 8));
    sort_key_values(c, false, rows, where, groups, sorted_rows, sorted_where);
    k_number_slots<<<stream_grid(groups, kBlock, 8192), kBlock, 0, c.stream>>>(sorted_where, groups, table);
    tmp_free(c, rows);
    tmp_free(c, where);
    tmp_free(c, sorted_rows);
    tmp_free(c, sorted_where);
  } else {
    // first rows flagged from the table, then listed in row order
    auto *first = static_cast<uint8_t *>(tmp_alloc(c, (n + 7) / 8 * 8));
    fill_u64(c, reinterpret_cast<uint64_t *>(first), (n + 7) / 8, 0);
    k_first_flags<<<stream_grid(slots + 1, kBlock * 4, 8192), kBlock, 0, c.stream>>>(table, slots, first);
    auto *firsts = static_cast<int64_t *>(tmp_alloc(c, n * 8));
    SelectSpec ss{2, VK_I64, 0, 0, 0};
    select_rows(c, ss, nullptr, first, n, firsts, d_cnt);
    read_back(c, d_cnt, 8, &groups);
    k_number1<<<stream_grid(groups, kBlock, 8192), kBlock, 0, c.stream>>>(ka, firsts, groups, table, mask, shift);
    tmp_free(c, first);
    tmp_free(c, firsts);
  }
  if (groups <= kLdsGroups)
    k_ids1<<<g, kGBlock, 0, c.stream>>>(ka, n, table, mask, shift, out);
  else
    k_ids2<<<stream_grid(n, kBlock * 4, 8192), kBlock, 0, c.stream>>>(ka, n, table, mask, shift, out);
  check_hip(hipGetLastError(), "group_id ids");
  tmp_free(c, table);
}


// ------------------------------------------------------------ group_id inside batch plans
// group_id_batch with its one host read (which items overflowed their small table) moved to
// the device: every item with more than 4096 rows has a full table of 2n words allocated at
// compile time; a run zero-fills and inserts into it only for items whose small table
// overflowed (their blocks check the item's overflow word), then the table each item ended in
// is picked on the device, and probe, first rows (a device-planned `where` selection), numbering
// and ids run over it. The group count of item i is left in groups[i].

namespace {

constexpr uint64_t kFillTile = 65536;  // table words per block of the conditional fill

// Zero item it's full table (aux: the item's index) when its small table overflowed.
__global__ __launch_bounds__(kBlock) void k_fill_full_plan(const BatchItem *fitems, const uint32_t *tile_item,
                                                           const uint32_t *over) {
  const BatchItem &F = fitems[tile_item[blockIdx.x]];
  if (over[F.aux] == 0) return;
  unsigned long long *t = static_cast<unsigned long long *>(const_cast<void *>(F.a));
  const uint64_t w0 = (blockIdx.x - F.tile0) * kFillTile;
  const uint64_t w1 = w0 + kFillTile < F.n ? w0 + kFillTile : F.n;
  for (uint64_t w = w0 + threadIdx.x; w < w1; w += kBlock) t[w] = 0;
}

__global__ void k_pick_items(const BatchItem *small, const BatchItem *full, const uint32_t *over, uint64_t nitems,
                             BatchItem *fin) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < nitems) fin[i] = over[i] ? full[i] : small[i];
}

BatchItem *upload_items(Ctx &c, const std::vector<BatchItem> &v) {
  auto *d = static_cast<BatchItem *>(tmp_alloc(c, (v.size() ? v.size() : 1) * sizeof(BatchItem)));
  if (!v.empty())
    check_hip(hipMemcpyAsync(d, v.data(), v.size() * sizeof(BatchItem), hipMemcpyHostToDevice, c.stream), "items");
  check_hip(hipStreamSynchronize(c.stream), "items");
  return d;
}

}  // namespace

PlanGroup *plan_group_new(Ctx &c, int32_t type, std::vector<BatchItem> &items, uint32_t *epoch) {
  constexpr int kBSmallLg = 13;
  auto *g = new PlanGroup{};
  g->type = type;
  g->nitems = items.size();
  const uint64_t ni = g->nitems;
  uint64_t total = 0, small_words = 0, full_words = 0;
  std::vector<int> full_lg(ni);
  for (size_t i = 0; i < ni; ++i) {
    BatchItem &it = items[i];
    if (it.n > 0x7FFFFFFFull) usage_error("group_id supports at most 2^31 - 1 rows");
    int lg = 1;
    while ((uint64_t(1) << lg) < 2 * it.n) ++lg;
    full_lg[i] = lg;
    it.aux = total;
    it.aux3 = static_cast<uint64_t>(lg < kBSmallLg ? lg : kBSmallLg);
    total += it.n;
    small_words += uint64_t(1) << it.aux3;
    if (lg > kBSmallLg) full_words += uint64_t(1) << lg;
  }
  g->total = total;
  g->small_words = small_words;
  g->small = static_cast<unsigned long long *>(tmp_alloc(c, (small_words ? small_words : 1) * 8));
  g->full = static_cast<unsigned long long *>(tmp_alloc(c, (full_words ? full_words : 1) * 8));
  g->over = static_cast<uint32_t *>(tmp_alloc(c, ni * 4 + 8));
  std::vector<BatchItem> full_items(ni), fill;
  uint64_t w = 0, f = 0;
  for (size_t i = 0; i < ni; ++i) {
    items[i].b = g->small + w;
    items[i].aux2 = int(items[i].aux3) < full_lg[i] ? uint64_t(kMaxRun) : uint64_t(kNoLimit);
    w += uint64_t(1) << items[i].aux3;
    full_items[i] = items[i];
    if (full_lg[i] > kBSmallLg) {
      full_items[i].b = g->full + f;
      full_items[i].aux3 = static_cast<uint64_t>(full_lg[i]);
      full_items[i].aux2 = uint64_t(kNoLimit);
      BatchItem fi{};
      fi.a = g->full + f;
      fi.n = uint64_t(1) << full_lg[i];
      fi.aux = i;
      fill.push_back(fi);
      f += uint64_t(1) << full_lg[i];
    }
  }
  g->small_p = batch_plan(c, items, kGTile);  // sets tile0 (the same in full_items below)
  for (size_t i = 0; i < ni; ++i) full_items[i].tile0 = items[i].tile0;
  g->items_full = upload_items(c, full_items);
  g->items_final = static_cast<BatchItem *>(tmp_alloc(c, (ni ? ni : 1) * sizeof(BatchItem)));
  g->fill_p = batch_plan(c, fill, kFillTile);
  g->slot = static_cast<uint32_t *>(tmp_alloc(c, (total ? total : 1) * 4));
  g->first = static_cast<uint8_t *>(tmp_alloc(c, total + 8));
  g->firsts = static_cast<int64_t *>(tmp_alloc(c, (total ? total : 1) * 8));
  g->groups = static_cast<uint64_t *>(tmp_alloc(c, ni * 8 + 8));
  // the first-row selection: item i's flags at first + aux, its first rows to firsts + aux
  std::vector<BatchItem> wl(ni);
  uint64_t max_tiles = 0;
  const uint64_t per = select_tile_rows(16);
  for (size_t i = 0; i < ni; ++i) {
    wl[i] = BatchItem{};
    wl[i].a = g->first + items[i].aux;
    wl[i].out = g->firsts + items[i].aux;
    wl[i].n = items[i].n;
    max_tiles += (items[i].n + per - 1) / per;
  }
  if (max_tiles > 0x7FFFFFFFull) usage_error("plan statement too large");
  DevPlan &dp = g->sel;
  dp.items = upload_items(c, wl);
  dp.max_tiles = static_cast<uint32_t>(max_tiles);
  dp.tile_item = static_cast<uint32_t *>(tmp_alloc(c, (max_tiles ? max_tiles : 1) * 4));
  dp.ntiles = static_cast<uint32_t *>(tmp_alloc(c, 8));
  dp.ticket = static_cast<uint32_t *>(tmp_alloc(c, 8));
  fill_u64(c, reinterpret_cast<uint64_t *>(dp.ticket), 1, 0);
  dp.epoch = epoch;
  g->status_words = max_tiles;
  dp.status = static_cast<uint64_t *>(tmp_alloc(c, (max_tiles ? max_tiles : 1) * 8));
  fill_u64(c, dp.status, max_tiles ? max_tiles : 1, 0);
  return g;
}

void plan_group_run(Ctx &c, PlanGroup &g) {
  const uint64_t ni = g.nitems;
  fill_u64(c, g.groups, ni, 0);
  if (g.total == 0) return;
  fill_u64(c, reinterpret_cast<uint64_t *>(g.small), g.small_words, 0);
  fill_u64(c, reinterpret_cast<uint64_t *>(g.over), (ni * 4 + 8) / 8, 0);
  const BatchPlan &p = g.small_p;
  k_insert_batch<<<p.ntiles, kBlock, 0, c.stream>>>(g.type, p.items, p.tile_item, g.over, nullptr);
  if (g.fill_p.ntiles) {
    k_fill_full_plan<<<g.fill_p.ntiles, kBlock, 0, c.stream>>>(g.fill_p.items, g.fill_p.tile_item, g.over);
    k_insert_batch<<<p.ntiles, kBlock, 0, c.stream>>>(g.type, g.items_full, p.tile_item, g.over, g.over);
  }
  k_pick_items<<<stream_grid(ni, kBlock, 1u << 20), kBlock, 0, c.stream>>>(p.items, g.items_full, g.over, ni,
                                                                          g.items_final);
  k_probe_batch<<<p.ntiles, kBlock, 0, c.stream>>>(g.type, g.items_final, p.tile_item, g.slot, g.first);
  check_hip(hipGetLastError(), "plan group_id table");
  plan_tiles_dev(c, g.sel, static_cast<uint32_t>(ni), nullptr, 1, select_tile_rows(16));
  SelectSpec ss{2, VK_I64, 0, 0, 0};
  select_batch_dev(c, ss, 16, g.sel, static_cast<uint32_t>(ni), g.groups);
  k_number_batch<<<p.ntiles, kBlock, 0, c.stream>>>(g.items_final, p.tile_item, g.firsts, g.groups, g.slot);
  k_ids_batch<<<p.ntiles, kBlock, 0, c.stream>>>(g.items_final, p.tile_item, g.slot, g.groups);
  check_hip(hipGetLastError(), "plan group_id ids");
}

void plan_group_free(Ctx &c, PlanGroup *g) {
  if (!g) return;
  batch_plan_free(c, g->small_p);
  batch_plan_free(c, g->fill_p);
  tmp_free(c, g->small);
  tmp_free(c, g->full);
  tmp_free(c, g->over);
  tmp_free(c, g->items_full);
  tmp_free(c, g->items_final);
  tmp_free(c, g->slot);
  tmp_free(c, g->first);
  tmp_free(c, g->firsts);
  tmp_free(c, g->groups);
  tmp_free(c, g->sel.items);
  tmp_free(c, g->sel.tile_item);
  tmp_free(c, g->sel.ntiles);
  tmp_free(c, g->sel.ticket);
  tmp_free(c, g->sel.status);
  delete g;
}

}  // namespace vk
