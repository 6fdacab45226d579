// Device key table: groups a tensor's nonzeros by key.
//
// Role in the reference: buildHashTable (fksCuckoo.hpp:1044-1060) inserts
// nonzeros one by one into an FKS two-level table whose second level is a
// per-bucket cuckoo table of id lists; reorder_tensor (fksCuckooTC.hpp:31-77)
// then renumbers the nonzeros so every key's list ("slice") is contiguous.
//
// Here the same result is built in HBM in four passes:
//   1. insert + rank: each key is linearised to 64 bits; every lane inserts
//      its own key with atomicCAS into an open-addressing table of 16-byte
//      entries (8 to a 128-byte line, probing from line mix64(key)); the
//      wavefront then sorts its 64 slots (bitonic network over lanes) so a
//      run of equal keys takes its ranks with one atomicAdd on the entry's
//      length.  Small key spaces are addressed directly (slot = key) and
//      ranked by a stable counting sort with per-block histograms, with no
//      global atomics at all however hot a key is;
//   2. pack + scan: slot counts and occupancy are packed into one 64-bit
//      word ((occupied << 32) | count) and prefix-summed in one scan, which
//      yields each slice's offset and its compact slice index together;
//   3. finalize: entries receive (offset, length); a compact slice list is
//      written;
//   4. place: every nonzero goes to offset + rank, with its non-key tuple
//      (linearised) and its value, in the streaming layout the contraction
//      reads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>

#include "ptc_internal.h"

namespace ptc {

constexpr int KT_THREADS = 256;
// Automatic grouping sorts key spaces of at most this many bits (two 8-bit
// radix passes) and partitions wider ones, measured on MI355X: the 20-bit
// two-tensor B table (5 M nonzeros, 1 M keys) partitions in 0.31 ms, sorts
// in 0.51 ms; the 14-bit two-tensor A table (10000 keys) sorts in 0.37 ms,
// partitions in 0.55 ms, and its contraction runs 0.27 ms faster over the
// sorted table; 30-bit nips-shape B keys partition in 0.22 ms, sort in
// 0.63 ms.
constexpr int SORT_AUTO_BITS = 16;

__device__ __forceinline__ unsigned long long mix64(unsigned long long k) {
  // MurmurHash3 fmix64 finaliser
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

// Linearise the key of nonzero e (modes codec.mode[]); returns false if an
// index is outside the codec's radix.
__device__ __forceinline__ bool encode_key(const uint32_t* __restrict__ h, uint32_t d, uint64_t e,
                                           const KeyCodec& c, unsigned long long* key) {
  unsigned long long k = 0;
  for (uint32_t j = 0; j < c.n; ++j) {
    uint32_t x = h[e * d + c.mode[j]];
    if (x >= c.radix[j]) return false;
    k += (unsigned long long)x * c.stride[j];
  }
  *key = k;
  return true;
}

__device__ __forceinline__ unsigned long long load_key_l2(const KtEntry* e) {
  return __hip_atomic_load(&e->key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One nonzero's streaming payload (linearised non-key tuple, value), kept
// in input order so the grouped layout is gathered with one read per
// position.
template <class RestT, class ValT>
struct KtPay {
  RestT k;
  ValT v;
};
// 4-byte tuples with 8-byte values: the padding word carries the nonzero's
// id through the partitioned builds (no separate id array to write and read)
template <>
struct KtPay<uint32_t, double> {
  uint32_t k;
  uint32_t id;
  double v;
};
template <class RestT, class ValT>
constexpr bool kPayId = sizeof(RestT) == 4 && sizeof(ValT) == 8;

// Pass 1 (fused, the default for hashed tables): every lane inserts its own
// key with atomicCAS along the probe order above; then the wavefront sorts
// its 64 slots with a bitonic network over lanes, and each run of equal
// slots takes its ranks with one atomicAdd on the length field of the
// slot's own entry (the line the insert just touched), so a key repeated
// across the wave costs one atomic.  The nonzero's payload is written in
// input order on the way.
template <class RestT, class ValT>
__global__ void __launch_bounds__(KT_THREADS) k_kt_insert_rank(
    const uint32_t* __restrict__ h, uint32_t d, uint64_t N, KeyCodec codec, KtEntry* __restrict__ tab,
    uint64_t cap, uint32_t* __restrict__ slot_of, uint32_t* __restrict__ rank_of, KeyCodec rest,
    const ValT* __restrict__ val, KtPay<RestT, ValT>* __restrict__ pay) {
  __shared__ uint32_t xr[KT_THREADS];
  const int lane = threadIdx.x & 63, wbase = threadIdx.x - lane;
  for (uint64_t e0 = (uint64_t)blockIdx.x * KT_THREADS; e0 < N; e0 += (uint64_t)gridDim.x * KT_THREADS) {
    const uint64_t e = e0 + threadIdx.x;
    uint32_t slot = 0xFFFFFFFFu;
    if (e < N) {
      unsigned long long key = 0;
      encode_key(h, d, e, codec, &key);  // table's own tensor: always within radix
      uint64_t i = kt_home(mix64(key), cap);
      while (true) {
        const unsigned long long k = load_key_l2(&tab[i]);
        if (k == key) break;
        if (k == PTC_EMPTY_KEY) {
          const unsigned long long old = atomicCAS(&tab[i].key, PTC_EMPTY_KEY, key);
          if (old == PTC_EMPTY_KEY || old == key) break;
          continue;  // lost the race to another key: re-read this entry
        }
        i = (i + 1) & (cap - 1);
      }
      slot = (uint32_t)i;
      slot_of[e] = slot;
      unsigned long long rk = 0;
      for (uint32_t j = 0; j < rest.n; ++j) rk += (unsigned long long)h[e * d + rest.mode[j]] * rest.stride[j];
      KtPay<RestT, ValT> pl;
      pl.k = (RestT)rk;
      pl.v = val[e];
      pay[e] = pl;
    }
    // rank: runs of equal slots across the wave
    uint32_t key = slot, pl = (uint32_t)lane;
#pragma unroll
    for (int kk = 2; kk <= 64; kk <<= 1) {
#pragma unroll
      for (int j = kk >> 1; j > 0; j >>= 1) {
        const uint32_t ok = __shfl_xor(key, j, 64), op = __shfl_xor(pl, j, 64);
        const bool up = (lane & kk) == 0;
        const bool lower = (lane & j) == 0;
        const bool other_less = ok < key || (ok == key && op < pl);
        if (lower == up ? other_less : !other_less) {
          key = ok;
          pl = op;
        }
      }
    }
    const uint32_t prevk = __shfl_up(key, 1, 64);
    const bool head = lane == 0 || prevk != key;
    const unsigned long long hm = __ballot(head);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const int start = 63 - __clzll((long long)(hm & upto));
    const unsigned long long above = hm & ~upto;
    const int next = above ? (__ffsll((long long)above) - 1) : 64;
    uint32_t rbase = 0;
    if (head && key != 0xFFFFFFFFu) rbase = atomicAdd(&tab[key].len, (uint32_t)(next - start));
    rbase = __shfl(rbase, start, 64);
    xr[wbase + pl] = rbase + (uint32_t)(lane - start);
    __syncthreads();
    if (e < N) rank_of[e] = xr[threadIdx.x];
    __syncthreads();
  }
}

// Direct addressing (a perfect hash) when the key space is small: slot = key.
__global__ void k_kt_direct(const uint32_t* __restrict__ h, uint32_t d, uint64_t N, KeyCodec codec,
                            uint32_t* __restrict__ slot_of) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (uint64_t)gridDim.x * blockDim.x) {
    unsigned long long key = 0;
    encode_key(h, d, e, codec, &key);
    slot_of[e] = (uint32_t)key;
  }
}

// Pass 1b: rank of every nonzero within its slot.  Each wave sorts its 64
// slot ids with a bitonic network over lanes, so equal slots form runs; a
// run's head then finds the slot in the block's LDS table and takes the
// run's ranks with one LDS atomicAdd (a Zipf-hot key would otherwise hit one
// LDS counter from most lanes).  Finally each distinct slot of the block
// costs one global atomicAdd.
constexpr int RK_THREADS = 256;
constexpr int RK_ITEMS = 8;
constexpr int RK_CHUNK = RK_THREADS * RK_ITEMS;
constexpr int RK_SLOTS = 2 * RK_CHUNK;

__global__ void __launch_bounds__(RK_THREADS) k_kt_rank(const uint32_t* __restrict__ slot_of, uint64_t N,
                                                          uint32_t* __restrict__ cnt,
                                                          uint32_t* __restrict__ rank_of) {
  __shared__ uint32_t lkey[RK_SLOTS];
  __shared__ uint32_t lcnt[RK_SLOTS];
  __shared__ uint32_t xh[RK_THREADS];  // per original lane: LDS slot of its key
  __shared__ uint32_t xr[RK_THREADS];  // per original lane: rank within the block
  const int lane = threadIdx.x & 63, wbase = threadIdx.x - lane;
  for (int i = threadIdx.x; i < RK_SLOTS; i += RK_THREADS) {
    lkey[i] = 0xFFFFFFFFu;
    lcnt[i] = 0;
  }
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * RK_CHUNK;
  uint32_t ls[RK_ITEMS], lr[RK_ITEMS];
  for (int k = 0; k < RK_ITEMS; ++k) {
    const uint64_t e = base + (uint64_t)k * RK_THREADS + threadIdx.x;
    uint32_t key = e < N ? slot_of[e] : 0xFFFFFFFFu;
    uint32_t pl = (uint32_t)lane;
    // bitonic sort of (key, lane) across the wave, ascending
#pragma unroll
    for (int kk = 2; kk <= 64; kk <<= 1) {
#pragma unroll
      for (int j = kk >> 1; j > 0; j >>= 1) {
        const uint32_t ok = __shfl_xor(key, j, 64), op = __shfl_xor(pl, j, 64);
        const bool up = (lane & kk) == 0;
        const bool lower = (lane & j) == 0;
        const bool other_less = ok < key || (ok == key && op < pl);
        if (lower == up ? other_less : !other_less) {
          key = ok;
          pl = op;
        }
      }
    }
    const uint32_t prevk = __shfl_up(key, 1, 64);
    const bool head = lane == 0 || prevk != key;
    const unsigned long long hm = __ballot(head);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const int start = 63 - __clzll((long long)(hm & upto));
    const unsigned long long above = hm & ~upto;
    const int next = above ? (__ffsll((long long)above) - 1) : 64;
    uint32_t h = 0, rbase = 0;
    if (head && key != 0xFFFFFFFFu) {
      h = (key * 2654435761u) & (RK_SLOTS - 1);
      while (true) {  // at most RK_CHUNK keys in 2*RK_CHUNK slots
        const uint32_t cur = lkey[h];
        if (cur == key) break;
        if (cur == 0xFFFFFFFFu) {
          const uint32_t old = atomicCAS(&lkey[h], 0xFFFFFFFFu, key);
          if (old == 0xFFFFFFFFu || old == key) break;
        }
        h = (h + 1) & (RK_SLOTS - 1);
      }
      rbase = atomicAdd(&lcnt[h], (uint32_t)(next - start));
    }
    h = __shfl(h, start, 64);
    rbase = __shfl(rbase, start, 64);
    xh[wbase + pl] = h;
    xr[wbase + pl] = rbase + (uint32_t)(lane - start);
    __syncthreads();
    ls[k] = xh[threadIdx.x];
    lr[k] = xr[threadIdx.x];
    __syncthreads();
  }
  __syncthreads();
  // one global atomic per distinct slot of the chunk; lcnt becomes the base
  for (int i = threadIdx.x; i < RK_SLOTS; i += RK_THREADS) {
    const uint32_t s = lkey[i];
    if (s != 0xFFFFFFFFu) lcnt[i] = atomicAdd(&cnt[s], lcnt[i]);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < RK_ITEMS; ++k) {
    const uint64_t e = base + (uint64_t)k * RK_THREADS + threadIdx.x;
    if (e < N) rank_of[e] = lcnt[ls[k]] + lr[k];
  }
}

// Direct tables with a small key space: a stable-offset counting sort.  Each
// block counts its chunk per slot in LDS (wave-sorted runs, one LDS atomic
// per run) and writes its counts into a slot-major matrix hist[slot][block];
// one flat exclusive scan of that matrix then gives every (slot, block) its
// first output position -- no global atomics, however hot a key is.
constexpr uint32_t RD_MAX = 8192;

__global__ void __launch_bounds__(RK_THREADS) k_kt_rank_direct(const uint32_t* __restrict__ slot_of, uint64_t N,
                                                                 uint32_t space, uint64_t nblocks,
                                                                 uint32_t* __restrict__ hist,
                                                                 uint32_t* __restrict__ rank_of) {
  __shared__ uint32_t lcnt[RD_MAX];
  __shared__ uint32_t xr[RK_THREADS];
  const int lane = threadIdx.x & 63, wbase = threadIdx.x - lane;
  for (uint32_t i = threadIdx.x; i < space; i += RK_THREADS) lcnt[i] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * RK_CHUNK;
  uint32_t lr[RK_ITEMS];
  for (int k = 0; k < RK_ITEMS; ++k) {
    const uint64_t e = base + (uint64_t)k * RK_THREADS + threadIdx.x;
    uint32_t key = e < N ? slot_of[e] : 0xFFFFFFFFu;
    uint32_t pl = (uint32_t)lane;
#pragma unroll
    for (int kk = 2; kk <= 64; kk <<= 1) {
#pragma unroll
      for (int j = kk >> 1; j > 0; j >>= 1) {
        const uint32_t ok = __shfl_xor(key, j, 64), op = __shfl_xor(pl, j, 64);
        const bool up = (lane & kk) == 0;
        const bool lower = (lane & j) == 0;
        const bool other_less = ok < key || (ok == key && op < pl);
        if (lower == up ? other_less : !other_less) {
          key = ok;
          pl = op;
        }
      }
    }
    const uint32_t prevk = __shfl_up(key, 1, 64);
    const bool head = lane == 0 || prevk != key;
    const unsigned long long hm = __ballot(head);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const int start = 63 - __clzll((long long)(hm & upto));
    const unsigned long long above = hm & ~upto;
    const int next = above ? (__ffsll((long long)above) - 1) : 64;
    uint32_t rbase = 0;
    if (head && key != 0xFFFFFFFFu) rbase = atomicAdd(&lcnt[key], (uint32_t)(next - start));
    rbase = __shfl(rbase, start, 64);
    xr[wbase + pl] = rbase + (uint32_t)(lane - start);
    __syncthreads();
    lr[k] = xr[threadIdx.x];
    __syncthreads();
  }
  for (uint32_t i = threadIdx.x; i < space; i += RK_THREADS) hist[(uint64_t)i * nblocks + blockIdx.x] = lcnt[i];
#pragma unroll
  for (int k = 0; k < RK_ITEMS; ++k) {
    const uint64_t e = base + (uint64_t)k * RK_THREADS + threadIdx.x;
    if (e < N) rank_of[e] = lr[k];
  }
}

__global__ void k_kt_direct_cnt(const unsigned long long* __restrict__ hpre, uint32_t space, uint64_t nblocks,
                                uint32_t* __restrict__ cnt) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < space; i += (uint64_t)gridDim.x * blockDim.x)
    cnt[i] = (uint32_t)(hpre[(i + 1) * nblocks] - hpre[i * nblocks]);
}

// Pass 4a for direct tables: position = hist prefix of (slot, block) + rank.
__global__ void k_kt_scatter_direct(uint64_t N, const unsigned long long* __restrict__ slotinfo,
                                    const unsigned long long* __restrict__ hpre, uint64_t nblocks,
                                    const uint32_t* __restrict__ slot_of, const uint32_t* __restrict__ rank_of,
                                    uint32_t* __restrict__ perm, uint32_t* __restrict__ slice_of) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t s = slot_of[e];
    const uint64_t pos = hpre[(uint64_t)s * nblocks + e / RK_CHUNK] + rank_of[e];
    perm[pos] = (uint32_t)e;
    slice_of[pos] = (uint32_t)(slotinfo[s] >> 32);
  }
}

__global__ void k_kt_clear(KtEntry* __restrict__ tab, uint32_t* __restrict__ cnt, uint64_t cap) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * blockDim.x) {
    tab[i].key = PTC_EMPTY_KEY;
    tab[i].off = 0;
    tab[i].len = 0;
    if (cnt) cnt[i] = 0;
  }
}

// Pass 2a: pack (occupied << 32) | count.
// Counts come from cnt[] (direct tables) or from the entries' length fields
// (hashed tables, cnt == NULL).
__global__ void k_kt_pack(const uint32_t* __restrict__ cnt, const KtEntry* __restrict__ tab,
                          unsigned long long* __restrict__ packed, uint64_t cap) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t c = cnt ? cnt[i] : tab[i].len;
    packed[i] = ((unsigned long long)(c > 0) << 32) | c;
  }
}

// Pass 3: entries get (offset, length); compact slice list; pre[] is
// overwritten with (offset | slice index << 32) of every occupied slot, the
// one word the scatter reads per nonzero.
__global__ void k_kt_finalize(KtEntry* __restrict__ tab, const uint32_t* __restrict__ cnt,
                              unsigned long long* __restrict__ pre, uint64_t cap,
                              KtSlice* __restrict__ slices, int direct) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t c = cnt ? cnt[i] : tab[i].len;
    if (!c) continue;
    unsigned long long p = pre[i];
    uint32_t off = (uint32_t)(p & 0xFFFFFFFFull);
    uint32_t sid = (uint32_t)(p >> 32);
    if (direct) tab[i].key = i;
    tab[i].off = off;
    tab[i].len = c;
    slices[sid].key = tab[i].key;
    slices[sid].off = off;
    slices[sid].len = c;
  }
}

// Pass 4a for hashed tables: ids into grouped order (one 4-byte random
// store; the slice index of each position is derived later, only for a
// table that serves as A).
__global__ void k_kt_scatter_perm(uint64_t N, const unsigned long long* __restrict__ slotinfo,
                                  const uint32_t* __restrict__ slot_of, const uint32_t* __restrict__ rank_of,
                                  uint32_t* __restrict__ perm) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (uint64_t)gridDim.x * blockDim.x)
    perm[(slotinfo[slot_of[e]] & 0xFFFFFFFFull) + rank_of[e]] = (uint32_t)e;
}

// Pass 4b for hashed tables: the streaming layout in grouped order from the
// input-order payload (one random 8- or 16-byte read per position).
template <class RestT, class ValT>
__global__ void k_kt_gather_pay(uint64_t N, const uint32_t* __restrict__ perm, const KtPay<RestT, ValT>* __restrict__ pay,
                                RestT* __restrict__ rest_key, ValT* __restrict__ gval) {
  for (uint64_t pos = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; pos < N; pos += (uint64_t)gridDim.x * blockDim.x) {
    const KtPay<RestT, ValT> pl = pay[perm[pos]];
    rest_key[pos] = pl.k;
    gval[pos] = pl.v;
  }
}

// Slice index of every grouped position: binary search of the slice
// offsets (ascending with the slice index).
__global__ void k_kt_slice_of(uint64_t N, const KtSlice* __restrict__ slices,
                              const unsigned long long* __restrict__ total, uint32_t* __restrict__ slice_of) {
  const uint64_t nslices = *total >> 32;  // >= 1 when N > 0
  for (uint64_t pos = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; pos < N; pos += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t lo = 0, hi = nslices - 1;
    while (lo < hi) {
      const uint64_t mid = (lo + hi + 1) >> 1;
      if (slices[mid].off <= pos) lo = mid; else hi = mid - 1;
    }
    slice_of[pos] = (uint32_t)lo;
  }
}

// Pass 4a (unfused): scatter ids into grouped order (two 4-byte random stores).
__global__ void k_kt_scatter(uint64_t N, const unsigned long long* __restrict__ slotinfo,
                             const uint32_t* __restrict__ slot_of, const uint32_t* __restrict__ rank_of,
                             uint32_t* __restrict__ perm, uint32_t* __restrict__ slice_of) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long si = slotinfo[slot_of[e]];
    const uint64_t pos = (si & 0xFFFFFFFFull) + rank_of[e];
    perm[pos] = (uint32_t)e;
    slice_of[pos] = (uint32_t)(si >> 32);
  }
}

// Pass 4 (fused, the default): every nonzero e, read in input order (its
// coordinates and value are sequential reads), goes straight to its grouped
// position pos = slice offset + rank: id, slice id, linearised non-key tuple
// and value are stored there.  The stores are scattered, but the four
// destination arrays stay resident in the last-level cache while they fill,
// so no HBM line is read per nonzero -- unlike gathering in grouped order,
// which reads a random line of the coordinates and one of the values for
// every position.  COUNTING: direct tables ordered by the (slot, block)
// prefix of the counting sort.
template <class RestT, class ValT, bool COUNTING>
__global__ void k_kt_place(uint64_t N, const uint32_t* __restrict__ h, uint32_t d, KeyCodec rest,
                           const ValT* __restrict__ val, const unsigned long long* __restrict__ slotinfo,
                           const unsigned long long* __restrict__ hpre, uint64_t nblocks,
                           const uint32_t* __restrict__ slot_of, const uint32_t* __restrict__ rank_of,
                           uint32_t* __restrict__ perm, uint32_t* __restrict__ slice_of,
                           RestT* __restrict__ rest_key, ValT* __restrict__ gval) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t s = slot_of[e];
    const unsigned long long si = slotinfo[s];
    const uint64_t pos = (COUNTING ? hpre[(uint64_t)s * nblocks + e / RK_CHUNK] : (si & 0xFFFFFFFFull)) + rank_of[e];
    unsigned long long rk = 0;
    for (uint32_t j = 0; j < rest.n; ++j) rk += (unsigned long long)h[e * d + rest.mode[j]] * rest.stride[j];
    perm[pos] = (uint32_t)e;
    slice_of[pos] = (uint32_t)(si >> 32);
    rest_key[pos] = (RestT)rk;
    gval[pos] = val[e];
  }
}

// Pass 4b (unfused): the streaming layout in grouped order, written
// coalesced: the linearised non-key tuple and the value of every grouped
// position.
template <class RestT, class ValT>
__global__ void k_kt_gather(const uint32_t* __restrict__ h, const ValT* __restrict__ val, uint32_t d,
                            uint64_t N, KeyCodec rest, const uint32_t* __restrict__ perm,
                            RestT* __restrict__ rest_key, ValT* __restrict__ gval) {
  for (uint64_t pos = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; pos < N; pos += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t e = perm[pos];
    unsigned long long rk = 0;
    for (uint32_t j = 0; j < rest.n; ++j) rk += (unsigned long long)h[e * d + rest.mode[j]] * rest.stride[j];
    rest_key[pos] = (RestT)rk;
    gval[pos] = val[e];
  }
}


// ------------------------------------------------------------ sort grouping
//
// The same table built by sorting instead of hashing: a stable LSD radix sort
// of (32-bit key, id) pairs with 8-bit digits, then slices from the runs of
// equal keys, then the probe index over the distinct keys only.  No global
// atomics: digit counts are per tile (4096 nonzeros) in LDS, offsets come
// from one scan of the digit-major tile histograms, and ranks inside a tile
// come from matching equal digits across the lanes of a wavefront.

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_WAVES = RS_THREADS / 64;

// Lanes of the wavefront holding the same 8-bit digit (0 for invalid lanes).
__device__ __forceinline__ unsigned long long match_digit(uint32_t dg, bool valid) {
  unsigned long long m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = ((dg >> b) & 1u) != 0;
    const unsigned long long bb = __ballot(bit);
    m &= bit ? bb : ~bb;
  }
  return valid ? m : 0ull;
}

// keys, ids = 0..N-1, and the input-order payload (non-key tuple, value)
template <class RestT, class ValT>
__global__ void k_rs_keys(const uint32_t* __restrict__ h, uint32_t d, uint64_t N, KeyCodec codec, KeyCodec rest,
                          const ValT* __restrict__ val, uint32_t* __restrict__ keys, uint32_t* __restrict__ ids,
                          KtPay<RestT, ValT>* __restrict__ pay) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (uint64_t)gridDim.x * blockDim.x) {
    unsigned long long key = 0;
    encode_key(h, d, e, codec, &key);  // the table's own tensor: always within radix
    keys[e] = (uint32_t)key;
    ids[e] = (uint32_t)e;
    unsigned long long rk = 0;
    for (uint32_t j = 0; j < rest.n; ++j) rk += (unsigned long long)h[e * d + rest.mode[j]] * rest.stride[j];
    KtPay<RestT, ValT> pl;
    pl.k = (RestT)rk;
    pl.v = val[e];
    pay[e] = pl;
  }
}

// digit counts of one tile -> hist[digit * ntiles + tile]
__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const uint32_t* __restrict__ keys, uint64_t N, int shift,
                                                         uint64_t ntiles, uint32_t* __restrict__ hist) {
  __shared__ uint32_t lh[256];
  const int lane = threadIdx.x & 63;
  lh[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ITEMS; ++r) {
    const uint64_t i = base + (uint64_t)r * RS_THREADS + threadIdx.x;
    const bool valid = i < N;
    const uint32_t dg = valid ? (keys[i] >> shift) & 255u : 0u;
    const unsigned long long m = match_digit(dg, valid);
    if (valid && __ffsll((long long)m) - 1 == lane) atomicAdd(&lh[dg], (uint32_t)__popcll(m));
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = lh[threadIdx.x];
}

// Stable scatter of one tile: round r handles elements base + r*256 + t, in
// input order; an element's position is its (digit, tile) offset + the
// count of its digit in earlier rounds + in earlier waves of this round +
// its rank among the equal-digit lanes of its wave.
__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const uint32_t* __restrict__ kin,
                                                            const uint32_t* __restrict__ iin, uint64_t N, int shift,
                                                            uint64_t ntiles, const unsigned long long* __restrict__ hs,
                                                            uint32_t* __restrict__ kout, uint32_t* __restrict__ iout) {
  __shared__ unsigned long long next[256];
  __shared__ uint32_t wcnt[RS_WAVES][256];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  next[threadIdx.x] = hs[(uint64_t)threadIdx.x * ntiles + blockIdx.x];
  for (int ww = 0; ww < RS_WAVES; ++ww) wcnt[ww][threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int r = 0; r < RS_ITEMS; ++r) {
    const uint64_t i = base + (uint64_t)r * RS_THREADS + threadIdx.x;
    const bool valid = i < N;
    const uint32_t k = valid ? kin[i] : 0u, id = valid ? iin[i] : 0u;
    const uint32_t dg = (k >> shift) & 255u;
    const unsigned long long m = match_digit(dg, valid);
    if (valid && __ffsll((long long)m) - 1 == lane) wcnt[w][dg] = (uint32_t)__popcll(m);
    __syncthreads();
    if (valid) {
      unsigned long long pos = next[dg] + (unsigned long long)__popcll(m & below);
      for (int ww = 0; ww < w; ++ww) pos += wcnt[ww][dg];
      kout[pos] = k;
      iout[pos] = id;
    }
    __syncthreads();
    uint32_t tot = 0;
    for (int ww = 0; ww < RS_WAVES; ++ww) {
      tot += wcnt[ww][threadIdx.x];
      wcnt[ww][threadIdx.x] = 0;
    }
    next[threadIdx.x] += tot;
    __syncthreads();
  }
}

__global__ void k_rs_heads(const uint32_t* __restrict__ K, uint64_t N, uint32_t* __restrict__ head) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (uint64_t)gridDim.x * blockDim.x)
    head[i] = (i == 0 || K[i] != K[i - 1]) ? 1u : 0u;
}

// slice of every sorted position; heads write (key, off), tails write the
// end into len (fixed to a length by k_rs_index)
__global__ void k_rs_slices(const uint32_t* __restrict__ K, uint64_t N, const uint32_t* __restrict__ head,
                            const unsigned long long* __restrict__ sx, KtSlice* __restrict__ slices,
                            uint32_t* __restrict__ slice_of) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t sl = sx[i] + head[i] - 1;
    slice_of[i] = (uint32_t)sl;
    if (head[i]) {
      slices[sl].key = K[i];
      slices[sl].off = (uint32_t)i;
    }
    if (i + 1 == N || head[i + 1]) slices[sl].len = (uint32_t)(i + 1);
  }
}

// Slice lengths, then the probe index: direct tables write entry[key];
// hashed tables insert every distinct key once (no duplicates to resolve).
__global__ void k_rs_index(KtSlice* __restrict__ slices, const unsigned long long* __restrict__ nkeys, KtEntry* __restrict__ tab,
                           uint64_t cap, int direct, unsigned long long* __restrict__ total, uint64_t N) {
  const uint64_t nk = *nkeys;
  if (blockIdx.x == 0 && threadIdx.x == 0) *total = (nk << 32) | N;
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nk; s += (uint64_t)gridDim.x * blockDim.x) {
    KtSlice sl = slices[s];
    sl.len = sl.len - sl.off;
    slices[s].len = sl.len;
    uint64_t i;
    if (direct) {
      i = sl.key;
      tab[i].key = sl.key;
    } else {
      i = kt_home(mix64(sl.key), cap);
      while (true) {
        const unsigned long long old = atomicCAS(&tab[i].key, PTC_EMPTY_KEY, sl.key);
        if (old == PTC_EMPTY_KEY) break;  // keys are distinct: a claimed entry is never this key's
        i = (i + 1) & (cap - 1);
      }
    }
    tab[i].off = sl.off;
    tab[i].len = sl.len;
  }
}

static unsigned grid_for(uint64_t n, unsigned threads, unsigned cap = 8192) {
  uint64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  return (unsigned)std::min<uint64_t>(g, cap);
}

// Probe kernel for the C ABI: one thread per query key.
__global__ void k_kt_probe(const uint32_t* __restrict__ q, uint64_t n, KeyCodec codec,
                           const KtEntry* __restrict__ tab, uint64_t cap, int direct,
                           unsigned long long* __restrict__ off_len) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    unsigned long long key = 0;
    bool ok = true;
    for (uint32_t j = 0; j < codec.n; ++j) {
      uint32_t x = q[i * codec.n + j];
      if (x >= codec.radix[j]) ok = false;
      key += (unsigned long long)x * codec.stride[j];
    }
    unsigned long long res = 0;
    if (ok && direct) {
      if (key < cap) {
        KtEntry en = tab[key];
        if (en.len) res = ((unsigned long long)en.off << 32) | en.len;
      }
    } else if (ok) {
      uint64_t i = kt_home(mix64(key), cap);
      for (uint64_t probes = 0; probes < cap; ++probes) {
        KtEntry en = tab[i];
        if (en.key == key) {
          res = ((unsigned long long)en.off << 32) | en.len;
          break;
        }
        if (en.key == PTC_EMPTY_KEY) break;
        i = (i + 1) & (cap - 1);
      }
    }
    off_len[i] = res;
  }
}

// ------------------------------------------------------ partitioned hashing
//
// Wide keys group in two steps that keep every random access on chip:
//   1. partition: nonzeros are scattered into buckets by the top bits of
//      mix64(key).  Per-tile bucket counts and one scan give every (bucket,
//      tile) pair its output range; each tile is counting-sorted by bucket in
//      LDS first, so its nonzeros leave in runs of consecutive addresses,
//      carrying the key, the id and the payload;
//   2. group: one workgroup per bucket counts the bucket's keys in an LDS
//      open-addressing table, turns the counts into offsets with a block
//      scan, and places every nonzero at bucket start + its key's offset + a
//      running count.  A bucket larger than the LDS table runs the same
//      steps on its own region of a global table (bounded: distinct keys <=
//      the bucket's size = the region's slots).  The probe index is written
//      in the same kernel: a key's home line (kt_home) is numbered by the top
//      bits of its hash, so a bucket's keys start in the bucket's own region
//      of the index, which the bucket's workgroup clears and fills alone; a
//      key whose probe runs past its region's end is left to a last pass.
// A bucket's slices come out in offset order, so all slices are sorted by
// offset; a final pass moves them to their compact positions.

constexpr int PT_THREADS = 512;
constexpr int PT_GROUP_THREADS = 1024;     // bucket workgroups: two per CU fill it
// two 1024-thread workgroups per CU are 8 waves per SIMD, which leaves each
// lane 64 VGPRs; without the bound the compiler takes 77 and one workgroup
// per CU fits
constexpr int PT_GROUP_WAVES_PER_SIMD = 8;
constexpr uint32_t PT_TAB = 4096;          // LDS table slots per bucket workgroup
constexpr uint32_t PT_TILE = 4096;         // nonzeros per partition tile (16-bit indices in LDS)
constexpr uint32_t PT_MAX_BUCKETS = 4096;
constexpr uint64_t PT_BUCKET_ITEMS = 3072; // target mean bucket size
// Beyond this many nonzeros the mean bucket outgrows the LDS table: one
// global table (PTC_GROUP_HASH) is used instead.
constexpr uint64_t PT_MAX_N = (uint64_t)PT_MAX_BUCKETS * PT_BUCKET_ITEMS;

// Bucket of a key: the top bbits bits of mix64(key); with bbits < 0 the key
// itself (small direct key spaces, one key per bucket).
__device__ __forceinline__ uint32_t pt_bucket(unsigned long long key, int bbits) {
  return bbits < 0 ? (uint32_t)key : bbits ? (uint32_t)(mix64(key) >> (64 - bbits)) : 0u;
}

// Keys (and, with rk, linearised non-key tuples) of the nonzeros
// base + i0 + threadIdx.x + u * blockDim.x, u < PT_U, of a tile of n: mode by
// mode, so each mode's PT_U loads are in flight together (a loop over the
// nonzeros that finishes one key before the next load waits on memory PT_U
// times as often).  Entries past n are 0.
constexpr int PT_U = 8;  // nonzeros per thread in one sweep: PT_TILE / PT_THREADS
__device__ __forceinline__ void pt_load_keys(const uint32_t* __restrict__ h, uint32_t d, uint64_t base, uint32_t i0,
                                             uint32_t n, const KeyCodec& c, unsigned long long (&key)[PT_U]) {
#pragma unroll
  for (int u = 0; u < PT_U; ++u) key[u] = 0;
  for (uint32_t j = 0; j < c.n; ++j) {
    const uint32_t m = c.mode[j];
    const unsigned long long st = c.stride[j];
    uint32_t x[PT_U];
#pragma unroll
    for (int u = 0; u < PT_U; ++u) {
      const uint32_t i = i0 + threadIdx.x + u * blockDim.x;
      x[u] = i < n ? h[(base + i) * d + m] : 0u;
    }
#pragma unroll
    for (int u = 0; u < PT_U; ++u) key[u] += (unsigned long long)x[u] * st;
  }
}

// Keys and linearised non-key tuples of the same nonzeros together.  Four
// modes (the common 4-mode tensor, when the coordinates are 16-byte aligned)
// take one 16-byte load a nonzero instead of one strided 4-byte load per
// mode and codec.
__device__ __forceinline__ bool pt_vec4(const uint32_t* h, uint32_t d) {
  return d == 4 && ((uintptr_t)h & 15) == 0;
}
__device__ __forceinline__ unsigned long long pt_encode4(const uint4& q, const KeyCodec& c) {
  unsigned long long k = 0;
  for (uint32_t j = 0; j < c.n; ++j) {
    const uint32_t m = c.mode[j];
    const uint32_t x = m == 0 ? q.x : m == 1 ? q.y : m == 2 ? q.z : q.w;
    k += (unsigned long long)x * c.stride[j];
  }
  return k;
}
__device__ __forceinline__ void pt_load_kr(const uint32_t* __restrict__ h, uint32_t d, uint64_t base, uint32_t i0,
                                           uint32_t n, const KeyCodec& codec, const KeyCodec& rest,
                                           unsigned long long (&key)[PT_U], unsigned long long (&rk)[PT_U]) {
  if (pt_vec4(h, d)) {
    uint4 q[PT_U];
#pragma unroll
    for (int u = 0; u < PT_U; ++u) {
      const uint32_t i = i0 + threadIdx.x + u * blockDim.x;
      q[u] = i < n ? reinterpret_cast<const uint4*>(h)[base + i] : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < PT_U; ++u) {
      key[u] = pt_encode4(q[u], codec);
      rk[u] = pt_encode4(q[u], rest);
    }
    return;
  }
  pt_load_keys(h, d, base, i0, n, codec, key);
  pt_load_keys(h, d, base, i0, n, rest, rk);
}
__device__ __forceinline__ void pt_load_k(const uint32_t* __restrict__ h, uint32_t d, uint64_t base, uint32_t i0,
                                          uint32_t n, const KeyCodec& codec, unsigned long long (&key)[PT_U]) {
  if (pt_vec4(h, d)) {
    uint4 q[PT_U];
#pragma unroll
    for (int u = 0; u < PT_U; ++u) {
      const uint32_t i = i0 + threadIdx.x + u * blockDim.x;
      q[u] = i < n ? reinterpret_cast<const uint4*>(h)[base + i] : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < PT_U; ++u) key[u] = pt_encode4(q[u], codec);
    return;
  }
  pt_load_keys(h, d, base, i0, n, codec, key);
}

// Exclusive scan, in place, of a[0..n) in LDS (n <= 8 * blockDim.x).
__device__ void pt_block_scan(uint32_t* a, uint32_t n, uint32_t* s_wave) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint32_t per = (n + blockDim.x - 1) / blockDim.x, s0 = threadIdx.x * per;
  uint32_t v[8], sum = 0;
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j) {
    v[j] = j < per && s0 + j < n ? a[s0 + j] : 0u;
    sum += v[j];
  }
  uint32_t inc = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t x = __shfl_up(inc, o, 64);
    if (lane >= o) inc += x;
  }
  if (lane == 63) s_wave[wid] = inc;
  __syncthreads();
  uint32_t pre = inc - sum;
  for (int w = 0; w < wid; ++w) pre += s_wave[w];
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j)
    if (j < per && s0 + j < n) {
      a[s0 + j] = pre;
      pre += v[j];
    }
  __syncthreads();
}

// bucket counts of tile t -> hist[bucket * ntiles + t]
__device__ __forceinline__ void pt_hist_tile(const uint32_t* __restrict__ h, uint32_t d, uint64_t N,
                                             const KeyCodec& codec, int bbits, uint32_t nb, uint64_t ntiles,
                                             uint32_t* __restrict__ hist, uint64_t t) {
  __shared__ uint32_t lh[PT_MAX_BUCKETS];
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) lh[b] = 0;
  __syncthreads();
  const uint64_t base = t * PT_TILE;
  const uint32_t n = (uint32_t)min<uint64_t>(PT_TILE, N - base);
  for (uint32_t i0 = 0; i0 < n; i0 += PT_U * blockDim.x) {
    unsigned long long key[PT_U];
    pt_load_keys(h, d, base, i0, n, codec, key);  // the table's own tensor: always within radix
#pragma unroll
    for (int u = 0; u < PT_U; ++u)
      if (i0 + threadIdx.x + u * blockDim.x < n) atomicAdd(&lh[pt_bucket(key[u], bbits)], 1u);
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) hist[(uint64_t)b * ntiles + t] = lh[b];
}

__global__ void __launch_bounds__(PT_THREADS) k_pt_hist(const uint32_t* __restrict__ h, uint32_t d, uint64_t N,
                                                        KeyCodec codec, int bbits, uint32_t nb, uint64_t ntiles,
                                                        uint32_t* __restrict__ hist) {
  pt_hist_tile(h, d, N, codec, bbits, nb, ntiles, hist, blockIdx.x);
}

// One tile into bucket order.  A rank pass reads the tile once (coalesced),
// keeps every nonzero's key and non-key tuple in LDS and counting-sorts the
// tile by bucket; the write pass then emits the nonzeros in bucket order,
// so each bucket's run goes to consecutive addresses (key, id, payload).
// Only the values are gathered, from the tile's few KB in L2.  Tiles are
// dealt to XCDs in contiguous ranges (workgroup i runs on XCD i mod 8), so
// the runs that adjacent tiles write side by side meet in one L2.
// Dynamic LDS: keys (8 B), tuples (RestT) and slot order (2 B) per nonzero
// of the tile, then two words per bucket.
__device__ __forceinline__ uint64_t pt_tile_of(uint32_t wg, uint64_t ntiles) {
  const uint64_t per = (ntiles + 7) / 8;
  return (uint64_t)(wg % 8) * per + wg / 8;
}

template <class RestT, class ValT>
__device__ __forceinline__ void pt_scatter_tile(const uint32_t* __restrict__ h, uint32_t d, uint64_t N,
                                                const KeyCodec& codec, const KeyCodec& rest,
                                                const ValT* __restrict__ val, int bbits, uint32_t nb, uint64_t ntiles,
                                                const uint32_t* __restrict__ hist,
                                                const unsigned long long* __restrict__ off,
                                                unsigned long long* __restrict__ pkey, uint32_t* __restrict__ pid,
                                                KtPay<RestT, ValT>* __restrict__ ppay, uint64_t t) {
  extern __shared__ unsigned long long pt_smem[];
  unsigned long long* lkey = pt_smem;                       // PT_TILE
  RestT* lrk = reinterpret_cast<RestT*>(lkey + PT_TILE);    // PT_TILE
  uint16_t* ls = reinterpret_cast<uint16_t*>(lrk + PT_TILE); // PT_TILE: tile index of every bucket-ordered slot
  uint32_t* lcur = reinterpret_cast<uint32_t*>(ls + PT_TILE); // nb: bucket start in the tile, then its cursor
  uint32_t* gdst = lcur + nb;                               // nb: output position of slot 0 of a bucket's run
  __shared__ uint32_t s_wave[PT_THREADS / 64];
  const uint64_t base = t * PT_TILE;
  const uint32_t n = (uint32_t)min<uint64_t>(PT_TILE, N - base);
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
    lcur[b] = hist[(uint64_t)b * ntiles + t];
    gdst[b] = (uint32_t)off[(uint64_t)b * ntiles + t];  // positions < 2^32
  }
  __syncthreads();
  pt_block_scan(lcur, nb, s_wave);
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) gdst[b] -= lcur[b];  // modulo 2^32: slot j -> gdst + j
  __syncthreads();
  for (uint32_t i0 = 0; i0 < n; i0 += PT_U * blockDim.x) {
    unsigned long long key[PT_U], rk[PT_U];
    pt_load_keys(h, d, base, i0, n, codec, key);
    pt_load_keys(h, d, base, i0, n, rest, rk);
#pragma unroll
    for (int u = 0; u < PT_U; ++u) {
      const uint32_t i = i0 + threadIdx.x + u * blockDim.x;
      if (i >= n) continue;
      lkey[i] = key[u];
      lrk[i] = (RestT)rk[u];
      ls[atomicAdd(&lcur[pt_bucket(key[u], bbits)], 1u)] = (uint16_t)i;
    }
  }
  __syncthreads();
  for (uint32_t j0 = 0; j0 < n; j0 += PT_U * blockDim.x) {
    uint32_t ii[PT_U];
    ValT v[PT_U];
#pragma unroll
    for (int u = 0; u < PT_U; ++u) {  // the value gathers first, all in flight
      const uint32_t j = j0 + threadIdx.x + u * blockDim.x;
      ii[u] = j < n ? ls[j] : 0u;
      v[u] = j < n ? val[base + ii[u]] : ValT(0);
    }
#pragma unroll
    for (int u = 0; u < PT_U; ++u) {
      const uint32_t j = j0 + threadIdx.x + u * blockDim.x;
      if (j >= n) continue;
      const uint32_t i = ii[u];
      const unsigned long long key = lkey[i];
      const uint32_t p = gdst[pt_bucket(key, bbits)] + j;
      pkey[p] = key;
      KtPay<RestT, ValT> pl;
      if constexpr (kPayId<RestT, ValT>) pl.id = (uint32_t)(base + i);
      else pid[p] = (uint32_t)(base + i);
      pl.k = lrk[i];
      pl.v = v[u];
      ppay[p] = pl;
    }
  }
}

template <class RestT, class ValT>
__global__ void __launch_bounds__(PT_THREADS) k_pt_scatter(const uint32_t* __restrict__ h, uint32_t d, uint64_t N,
                                                           KeyCodec codec, KeyCodec rest, const ValT* __restrict__ val,
                                                           int bbits, uint32_t nb, uint64_t ntiles,
                                                           const uint32_t* __restrict__ hist,
                                                           const unsigned long long* __restrict__ off,
                                                           unsigned long long* __restrict__ pkey,
                                                           uint32_t* __restrict__ pid,
                                                           KtPay<RestT, ValT>* __restrict__ ppay) {
  const uint64_t t = pt_tile_of(blockIdx.x, ntiles);
  if (t >= ntiles) return;
  pt_scatter_tile<RestT, ValT>(h, d, N, codec, rest, val, bbits, nb, ntiles, hist, off, pkey, pid, ppay, t);
}

// ------------------------------------------------- small direct key spaces
//
// A key space of at most PT_MAX_BUCKETS keys is grouped in one counting
// pass: every key is its own bucket, so the bucket scan already gives every
// slice; the tiles are scattered straight into the final layout.

// The slices (nonempty keys in key order), the whole direct index and the
// key count, from the scanned count matrix; one workgroup.  Positions in
// off start at obase (the table's first position in a set of tables).
__device__ __forceinline__ void pd_slices_one(const unsigned long long* __restrict__ off, uint64_t ntiles,
                                              uint32_t nb, uint64_t N, uint64_t obase, KtSlice* __restrict__ slices,
                                              KtEntry* __restrict__ tab, uint32_t* __restrict__ cidx,
                                              unsigned long long* __restrict__ total) {
  __shared__ uint32_t lf[PT_MAX_BUCKETS];
  __shared__ uint32_t s_wave[1024 / 64];
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
    const uint64_t s0 = off[(uint64_t)b * ntiles] - obase;
    const uint64_t s1 = b + 1 < nb ? off[(uint64_t)(b + 1) * ntiles] - obase : N;
    lf[b] = s1 > s0 ? 1u : 0u;
  }
  __syncthreads();
  const uint32_t last = lf[nb - 1];
  __syncthreads();
  pt_block_scan(lf, nb, s_wave);
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
    const uint64_t s0 = off[(uint64_t)b * ntiles] - obase;
    const uint64_t s1 = b + 1 < nb ? off[(uint64_t)(b + 1) * ntiles] - obase : N;
    KtEntry en;
    en.key = PTC_EMPTY_KEY;
    en.off = 0;
    en.len = 0;
    if (s1 > s0) {
      KtSlice sl;
      sl.key = b;
      sl.off = (uint32_t)s0;
      sl.len = (uint32_t)(s1 - s0);
      slices[lf[b]] = sl;
      en.key = b;
      en.off = sl.off;
      en.len = sl.len;
    }
    cidx[b] = lf[b];
    tab[b] = en;
  }
  if (threadIdx.x == 0) *total = ((unsigned long long)(lf[nb - 1] + last) << 32) | N;
}

__global__ void __launch_bounds__(1024) k_pd_slices(const unsigned long long* __restrict__ off, uint64_t ntiles,
                                                    uint32_t nb, uint64_t N, KtSlice* __restrict__ slices,
                                                    KtEntry* __restrict__ tab, uint32_t* __restrict__ cidx,
                                                    unsigned long long* __restrict__ total) {
  pd_slices_one(off, ntiles, nb, N, 0, slices, tab, cidx, total);
}

// One tile in key order, written straight into the grouped layout (ids,
// non-key tuples, values, slice indices); as k_pt_scatter, with the key
// (16 bits) kept in LDS in place of the 64-bit key.
template <class RestT, class ValT>
__device__ __forceinline__ void pd_scatter_tile(const uint32_t* __restrict__ h, uint32_t d, uint64_t N,
                                                const KeyCodec& codec, const KeyCodec& rest,
                                                const ValT* __restrict__ val, uint32_t nb, uint64_t ntiles,
                                                const uint32_t* __restrict__ hist,
                                                const unsigned long long* __restrict__ off, uint64_t obase,
                                                const uint32_t* __restrict__ cidx, uint32_t* __restrict__ perm,
                                                RestT* __restrict__ rest_out, ValT* __restrict__ gval,
                                                uint32_t* __restrict__ slice_of, uint64_t t,
                                                uint16_t* __restrict__ rest16 = nullptr) {
  extern __shared__ unsigned long long pt_smem[];
  RestT* lrk = reinterpret_cast<RestT*>(pt_smem);           // PT_TILE
  uint16_t* lb = reinterpret_cast<uint16_t*>(lrk + PT_TILE); // PT_TILE: key of every tile item
  uint16_t* ls = lb + PT_TILE;                               // PT_TILE: tile index of every key-ordered slot
  uint32_t* lcur = reinterpret_cast<uint32_t*>(ls + PT_TILE); // nb
  uint32_t* gdst = lcur + nb;                                // nb
  __shared__ uint32_t s_wave[PT_THREADS / 64];
  const uint64_t base = t * PT_TILE;
  const uint32_t n = (uint32_t)min<uint64_t>(PT_TILE, N - base);
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
    lcur[b] = hist[(uint64_t)b * ntiles + t];
    gdst[b] = (uint32_t)(off[(uint64_t)b * ntiles + t] - obase);
  }
  __syncthreads();
  pt_block_scan(lcur, nb, s_wave);
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) gdst[b] -= lcur[b];
  __syncthreads();
  for (uint32_t i0 = 0; i0 < n; i0 += PT_U * blockDim.x) {
    unsigned long long key[PT_U], rk[PT_U];
    pt_load_keys(h, d, base, i0, n, codec, key);  // < nb
    pt_load_keys(h, d, base, i0, n, rest, rk);
#pragma unroll
    for (int u = 0; u < PT_U; ++u) {
      const uint32_t i = i0 + threadIdx.x + u * blockDim.x;
      if (i >= n) continue;
      lrk[i] = (RestT)rk[u];
      lb[i] = (uint16_t)key[u];
      ls[atomicAdd(&lcur[key[u]], 1u)] = (uint16_t)i;
    }
  }
  __syncthreads();
  for (uint32_t j0 = 0; j0 < n; j0 += PT_U * blockDim.x) {
    uint32_t ii[PT_U], sl[PT_U];
    ValT v[PT_U];
#pragma unroll
    for (int u = 0; u < PT_U; ++u) {  // the gathers first, all in flight
      const uint32_t j = j0 + threadIdx.x + u * blockDim.x;
      ii[u] = j < n ? ls[j] : 0u;
      v[u] = j < n ? val[base + ii[u]] : ValT(0);
      sl[u] = j < n && slice_of ? cidx[lb[ii[u]]] : 0u;
    }
#pragma unroll
    for (int u = 0; u < PT_U; ++u) {
      const uint32_t j = j0 + threadIdx.x + u * blockDim.x;
      if (j >= n) continue;
      const uint32_t i = ii[u];
      const uint32_t p = gdst[lb[i]] + j;
      perm[p] = (uint32_t)(base + i);
      rest_out[p] = lrk[i];
      if (rest16) rest16[p] = (uint16_t)lrk[i];
      gval[p] = v[u];
      if (slice_of) slice_of[p] = sl[u];
    }
  }
}

template <class RestT, class ValT>
__global__ void __launch_bounds__(PT_THREADS) k_pd_scatter(const uint32_t* __restrict__ h, uint32_t d, uint64_t N,
                                                           KeyCodec codec, KeyCodec rest, const ValT* __restrict__ val,
                                                           uint32_t nb, uint64_t ntiles,
                                                           const uint32_t* __restrict__ hist,
                                                           const unsigned long long* __restrict__ off,
                                                           const uint32_t* __restrict__ cidx,
                                                           uint32_t* __restrict__ perm, RestT* __restrict__ rest_out,
                                                           ValT* __restrict__ gval, uint32_t* __restrict__ slice_of) {
  const uint64_t t = pt_tile_of(blockIdx.x, ntiles);
  if (t >= ntiles) return;
  pd_scatter_tile<RestT, ValT>(h, d, N, codec, rest, val, nb, ntiles, hist, off, 0, cidx, perm, rest_out, gval,
                               slice_of, t);
}

// ----------------------------------------------- placement by column scan
//
// A single table places its tiles without global atomics and without a
// bucket-major count matrix: the counting pass writes every tile's count
// per bucket (or direct key) as one tile-major row (coalesced); a column
// scan gives every (tile, bucket) its offset among the bucket's earlier
// tiles -- first within segments of PC_SEG tiles (one wave per segment and
// 64 buckets, the segment's rows read coalesced), then across segments in
// one workgroup that also turns the bucket totals into bucket starts (and,
// for a direct key space, into the slices and the whole index); the placing
// pass reads its tile's rows and writes the tile in bucket order.  A
// bucket's runs stay in tile order, and adjacent tiles run on one XCD
// (pt_tile_of), so side-by-side runs merge in that XCD's L2.  Measured on
// MI355X: a count pass with one global fetch-add per (tile, bucket) took
// 32 us on the nips-shape A table against 10 us without it.

constexpr uint32_t PC_SEG = 32;  // tiles per column-scan segment

__global__ void __launch_bounds__(PT_THREADS) k_pc_count(const uint32_t* __restrict__ h, uint32_t d, uint64_t N,
                                                         KeyCodec codec, int bbits, uint32_t nb, uint64_t ntiles,
                                                         uint32_t* __restrict__ tcnt) {
  __shared__ uint32_t lh[PT_MAX_BUCKETS];
  const uint64_t t = pt_tile_of(blockIdx.x, ntiles);
  if (t >= ntiles) return;
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) lh[b] = 0;
  __syncthreads();
  const uint64_t base = t * PT_TILE;
  const uint32_t n = (uint32_t)min<uint64_t>(PT_TILE, N - base);
  for (uint32_t i0 = 0; i0 < n; i0 += PT_U * blockDim.x) {
    unsigned long long key[PT_U];
    pt_load_k(h, d, base, i0, n, codec, key);  // the table's own tensor: always within radix
#pragma unroll
    for (int u = 0; u < PT_U; ++u)
      if (i0 + threadIdx.x + u * blockDim.x < n) atomicAdd(&lh[pt_bucket(key[u], bbits)], 1u);
  }
  __syncthreads();
  uint32_t* rc = tcnt + t * nb;
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) rc[b] = lh[b];
}

// Offsets within segments: wave w of the grid takes buckets [64 g, 64 g +
// 64) of segment s; toff[t][b] = sum of tcnt[t'][b] over the segment's
// tiles t' < t, segtot[s][b] = the segment's total.
__global__ void __launch_bounds__(256) k_pc_colscan(const uint32_t* __restrict__ tcnt, uint32_t nb, uint64_t ntiles,
                                                    uint32_t* __restrict__ toff, uint32_t* __restrict__ segtot) {
  const uint32_t ng = (nb + 63) / 64;
  const uint64_t w = (uint64_t)blockIdx.x * 4 + threadIdx.x / 64;
  const uint64_t s = w / ng;
  const uint32_t b = (uint32_t)(w % ng) * 64 + (threadIdx.x & 63);
  const uint64_t t0 = s * PC_SEG;
  if (t0 >= ntiles || b >= nb) return;
  const uint32_t nts = (uint32_t)min<uint64_t>(PC_SEG, ntiles - t0);
  uint32_t c[PC_SEG];
#pragma unroll
  for (uint32_t k = 0; k < PC_SEG; ++k) c[k] = k < nts ? tcnt[(t0 + k) * nb + b] : 0u;  // all in flight
  uint32_t run = 0;
#pragma unroll
  for (uint32_t k = 0; k < PC_SEG; ++k) {
    if (k < nts) toff[(t0 + k) * nb + b] = run;
    run += c[k];
  }
  segtot[s * nb + b] = run;
}

// Segment prefixes, one thread per bucket: segpre[s][b] = the bucket's count
// in segments before s (in place of the segment totals), btot[b] = its total.
__global__ void __launch_bounds__(256) k_pc_segscan(uint32_t* __restrict__ segpre, uint32_t nb, uint32_t nseg,
                                                    uint32_t* __restrict__ btot) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  uint32_t c = 0;
  for (uint32_t s0 = 0; s0 < nseg; s0 += 8) {
    uint32_t x[8];
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) x[k] = s0 + k < nseg ? segpre[(uint64_t)(s0 + k) * nb + b] : 0u;  // in flight together
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
      if (s0 + k < nseg) segpre[(uint64_t)(s0 + k) * nb + b] = c;
      c += x[k];
    }
  }
  btot[b] = c;
}

// Bucket starts from the totals, one workgroup of 1024 threads (nb <=
// PT_MAX_BUCKETS, four consecutive buckets a thread): start[b], start[nb] =
// N.  A direct key space (every key its own bucket) also gets its slices
// (present keys in key order), its whole index, the compact slice index of
// every key and the (keys << 32 | nonzeros) total.  (present << 32 | count)
// is scanned as one word: counts sum to N < 2^32.
__global__ void __launch_bounds__(1024) k_pc_starts(const uint32_t* __restrict__ btot, uint32_t nb, uint64_t N,
                                                    unsigned long long* __restrict__ start, int direct,
                                                    KtSlice* __restrict__ slices, KtEntry* __restrict__ tab,
                                                    uint32_t* __restrict__ cidx,
                                                    unsigned long long* __restrict__ total,
                                                    uint32_t* __restrict__ zero_word) {
  __shared__ unsigned long long s_wave[1024 / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint32_t b0 = threadIdx.x * 4;
  if (zero_word && threadIdx.x == 0) *zero_word = 0;  // a counter of the next launches
  unsigned long long v[4], sum = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t c = b0 + j < nb ? btot[b0 + j] : 0u;
    v[j] = c ? (1ull << 32) | c : 0ull;
    sum += v[j];
  }
  unsigned long long inc = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned long long x = __shfl_up(inc, o, 64);
    if (lane >= o) inc += x;
  }
  if (lane == 63) s_wave[wid] = inc;
  __syncthreads();
  unsigned long long pre = inc - sum;
  for (int w = 0; w < wid; ++w) pre += s_wave[w];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t b = b0 + j;
    if (b >= nb) break;
    const uint32_t off = (uint32_t)pre, c = (uint32_t)v[j], si = (uint32_t)(pre >> 32);
    start[b] = off;
    if (direct) {
      KtEntry en;
      en.key = PTC_EMPTY_KEY;
      en.off = 0;
      en.len = 0;
      if (c) {
        KtSlice sl;
        sl.key = b;
        sl.off = off;
        sl.len = c;
        slices[si] = sl;
        en.key = b;
        en.off = off;
        en.len = c;
      }
      tab[b] = en;
      cidx[b] = si;
    }
    pre += v[j];
  }
  if (threadIdx.x == blockDim.x - 1) {
    start[nb] = N;
    if (direct) *total = (pre & 0xFFFFFFFF00000000ull) | N;
  }
}

// The tile's counts and offsets: lcur[b] = the bucket's start in the tile
// (scanned counts), gdst[b] + j = the output position of the tile's
// bucket-ordered slot j.
__device__ __forceinline__ void pc_tile_row(uint64_t t, uint32_t nb, const unsigned long long* __restrict__ start,
                                            const uint32_t* __restrict__ tcnt, const uint32_t* __restrict__ toff,
                                            const uint32_t* __restrict__ segpre, uint32_t* lcur, uint32_t* gdst,
                                            uint32_t* s_wave) {
  const uint32_t* rc = tcnt + t * nb;
  const uint32_t* ro = toff + t * nb;
  const uint32_t* rs = segpre + (t / PC_SEG) * nb;
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
    const uint32_t c = rc[b];
    lcur[b] = c;
    gdst[b] = c ? (uint32_t)start[b] + rs[b] + ro[b] : 0u;  // positions < 2^32
  }
  __syncthreads();
  pt_block_scan(lcur, nb, s_wave);
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) gdst[b] -= lcur[b];  // modulo 2^32: slot j -> gdst + j
  __syncthreads();
}

// Direct key spaces: one tile in key order, straight into the grouped layout
// (ids, non-key tuples, values; slice indices when slice_of is non-NULL).
// Dynamic LDS as k_pd_scatter.
template <class RestT, class ValT>
__global__ void __launch_bounds__(PT_THREADS) k_pd_place(const uint32_t* __restrict__ h, uint32_t d, uint64_t N,
                                                         KeyCodec codec, KeyCodec rest, const ValT* __restrict__ val,
                                                         uint32_t nb, uint64_t ntiles,
                                                         const unsigned long long* __restrict__ start,
                                                         const uint32_t* __restrict__ tcnt,
                                                         const uint32_t* __restrict__ toff,
                                                         const uint32_t* __restrict__ segpre,
                                                         const uint32_t* __restrict__ cidx,
                                                         uint32_t* __restrict__ perm, RestT* __restrict__ rest_out,
                                                         ValT* __restrict__ gval, uint32_t* __restrict__ slice_of,
                                                         uint16_t* __restrict__ rest16) {
  const uint64_t t = pt_tile_of(blockIdx.x, ntiles);
  if (t >= ntiles) return;
  extern __shared__ unsigned long long pt_smem[];
  RestT* lrk = reinterpret_cast<RestT*>(pt_smem);           // PT_TILE
  uint16_t* lb = reinterpret_cast<uint16_t*>(lrk + PT_TILE); // PT_TILE: key of every tile item
  uint16_t* ls = lb + PT_TILE;                               // PT_TILE: tile index of every key-ordered slot
  uint32_t* lcur = reinterpret_cast<uint32_t*>(ls + PT_TILE); // nb
  uint32_t* gdst = lcur + nb;                                // nb
  __shared__ uint32_t s_wave[PT_THREADS / 64];
  pc_tile_row(t, nb, start, tcnt, toff, segpre, lcur, gdst, s_wave);
  const uint64_t base = t * PT_TILE;
  const uint32_t n = (uint32_t)min<uint64_t>(PT_TILE, N - base);
  for (uint32_t i0 = 0; i0 < n; i0 += PT_U * blockDim.x) {
    unsigned long long key[PT_U], rk[PT_U];
    pt_load_kr(h, d, base, i0, n, codec, rest, key, rk);  // keys < nb
#pragma unroll
    for (int u = 0; u < PT_U; ++u) {
      const uint32_t i = i0 + threadIdx.x + u * blockDim.x;
      if (i >= n) continue;
      lrk[i] = (RestT)rk[u];
      lb[i] = (uint16_t)key[u];
      ls[atomicAdd(&lcur[key[u]], 1u)] = (uint16_t)i;
    }
  }
  __syncthreads();
  for (uint32_t j0 = 0; j0 < n; j0 += PT_U * blockDim.x) {
    uint32_t ii[PT_U], sl[PT_U];
    ValT v[PT_U];
#pragma unroll
    for (int u = 0; u < PT_U; ++u) {  // the gathers first, all in flight
      const uint32_t j = j0 + threadIdx.x + u * blockDim.x;
      ii[u] = j < n ? ls[j] : 0u;
      v[u] = j < n ? val[base + ii[u]] : ValT(0);
      sl[u] = j < n && slice_of ? cidx[lb[ii[u]]] : 0u;
    }
#pragma unroll
    for (int u = 0; u < PT_U; ++u) {
      const uint32_t j = j0 + threadIdx.x + u * blockDim.x;
      if (j >= n) continue;
      const uint32_t i = ii[u];
      const uint32_t p = gdst[lb[i]] + j;
      perm[p] = (uint32_t)(base + i);
      rest_out[p] = lrk[i];
      if (rest16) rest16[p] = (uint16_t)lrk[i];
      gval[p] = v[u];
      if (slice_of) slice_of[p] = sl[u];
    }
  }
}

// Wide keys: one tile into hash-bucket order (key, id, payload), as
// k_pt_scatter, its runs placed at the reserved offsets.  Dynamic LDS as
// k_pt_scatter.
template <class RestT, class ValT, class KeyT>
__global__ void __launch_bounds__(PT_THREADS) k_pt_place(const uint32_t* __restrict__ h, uint32_t d, uint64_t N,
                                                         KeyCodec codec, KeyCodec rest, const ValT* __restrict__ val,
                                                         int bbits, uint32_t nb, uint64_t ntiles,
                                                         const unsigned long long* __restrict__ start,
                                                         const uint32_t* __restrict__ tcnt,
                                                         const uint32_t* __restrict__ toff,
                                                         const uint32_t* __restrict__ segpre,
                                                         unsigned long long* __restrict__ pkey,
                                                         uint32_t* __restrict__ pid,
                                                         KtPay<RestT, ValT>* __restrict__ ppay) {
  const uint64_t t = pt_tile_of(blockIdx.x, ntiles);
  if (t >= ntiles) return;
  extern __shared__ unsigned long long pt_smem[];
  // keys of a space within 32 bits take 4 bytes of LDS: a tile then needs
  // 48 KB (three workgroups a CU, the whole nips-shape B tensor at once)
  KeyT* lkey = reinterpret_cast<KeyT*>(pt_smem);            // PT_TILE
  RestT* lrk = reinterpret_cast<RestT*>(lkey + PT_TILE);    // PT_TILE
  uint16_t* ls = reinterpret_cast<uint16_t*>(lrk + PT_TILE); // PT_TILE: tile index of every bucket-ordered slot
  uint32_t* lcur = reinterpret_cast<uint32_t*>(ls + PT_TILE); // nb
  uint32_t* gdst = lcur + nb;                               // nb
  __shared__ uint32_t s_wave[PT_THREADS / 64];
  pc_tile_row(t, nb, start, tcnt, toff, segpre, lcur, gdst, s_wave);
  const uint64_t base = t * PT_TILE;
  const uint32_t n = (uint32_t)min<uint64_t>(PT_TILE, N - base);
  for (uint32_t i0 = 0; i0 < n; i0 += PT_U * blockDim.x) {
    unsigned long long key[PT_U], rk[PT_U];
    pt_load_kr(h, d, base, i0, n, codec, rest, key, rk);
#pragma unroll
    for (int u = 0; u < PT_U; ++u) {
      const uint32_t i = i0 + threadIdx.x + u * blockDim.x;
      if (i >= n) continue;
      lkey[i] = (KeyT)key[u];
      lrk[i] = (RestT)rk[u];
      ls[atomicAdd(&lcur[pt_bucket(key[u], bbits)], 1u)] = (uint16_t)i;
    }
  }
  __syncthreads();
  for (uint32_t j0 = 0; j0 < n; j0 += PT_U * blockDim.x) {
    uint32_t ii[PT_U];
    ValT v[PT_U];
#pragma unroll
    for (int u = 0; u < PT_U; ++u) {  // the value gathers first, all in flight
      const uint32_t j = j0 + threadIdx.x + u * blockDim.x;
      ii[u] = j < n ? ls[j] : 0u;
      v[u] = j < n ? val[base + ii[u]] : ValT(0);
    }
#pragma unroll
    for (int u = 0; u < PT_U; ++u) {
      const uint32_t j = j0 + threadIdx.x + u * blockDim.x;
      if (j >= n) continue;
      const uint32_t i = ii[u];
      const unsigned long long key = lkey[i];
      const uint32_t p = gdst[pt_bucket(key, bbits)] + j;
      pkey[p] = key;
      KtPay<RestT, ValT> pl;
      if constexpr (kPayId<RestT, ValT>) pl.id = (uint32_t)(base + i);
      else pid[p] = (uint32_t)(base + i);
      pl.k = lrk[i];
      pl.v = v[u];
      ppay[p] = pl;
    }
  }
}

// Table state of one bucket: LDS arrays, or the bucket's region of a global
// table.  Either is touched by the bucket's workgroup alone, so every
// access is a workgroup-scope atomic: global ones are performed in the L2
// of the workgroup's own XCD, and __syncthreads orders them (no device-wide
// fence, which on gfx950 writes the L2 back).
__device__ __forceinline__ unsigned long long pt_ld64(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t pt_ld32(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ unsigned long long pt_cas64(unsigned long long* p, unsigned long long cmp,
                                                       unsigned long long v) {
  __hip_atomic_compare_exchange_strong(p, &cmp, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  return cmp;  // the value found
}
__device__ __forceinline__ uint32_t pt_add32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void pt_st64(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void pt_st32(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Slot of key in a table of cap slots (linear probing from mix64(key) mod
// cap); inserts the key when absent.  The global tables hold at most cap
// distinct keys (cap = the bucket's size).  The LDS table counts its keys in
// *ndist and refuses a new key (PT_FULL) once PT_LDS_KEYS are in: with at
// most one insert per thread in flight past that check, at most
// PT_LDS_KEYS + PT_GROUP_THREADS - 1 < PT_TAB keys ever enter, so every
// probe ends.
constexpr uint32_t PT_LDS_KEYS = PT_TAB - PT_GROUP_THREADS - 1;
constexpr uint32_t PT_FULL = 0xFFFFFFFFu;

template <bool GLOBAL>
__device__ __forceinline__ uint32_t pt_slot(unsigned long long* tk, uint32_t cap, unsigned long long key,
                                            uint32_t* ndist = nullptr) {
  uint32_t s = GLOBAL ? (uint32_t)(mix64(key) % cap) : (uint32_t)mix64(key) & (PT_TAB - 1);
  while (true) {
    const unsigned long long k = pt_ld64(&tk[s]);
    if (k == key) return s;
    if (k == PTC_EMPTY_KEY) {
      if (ndist && pt_ld32(ndist) >= PT_LDS_KEYS) return PT_FULL;
      const unsigned long long old = pt_cas64(&tk[s], PTC_EMPTY_KEY, key);
      if (old == PTC_EMPTY_KEY && ndist) pt_add32(ndist, 1u);
      if (old == PTC_EMPTY_KEY || old == key) return s;
      continue;  // another key took the slot: re-read it
    }
    s = GLOBAL ? (s + 1 == cap ? 0 : s + 1) : (s + 1) & (PT_TAB - 1);
  }
}

// The probe index of a hashed table, written per bucket region
// [r0, r1) (region mode).
struct PtIndex {
  KtEntry* tab;
  uint64_t cap;
  uint64_t r0, r1;
  int region;          // 1: this kernel writes the index
  uint32_t* map;       // LDS image of the region (slot -> local slice), or NULL
  uint32_t* ovf_cnt;   // keys whose probe ran past their region
  uint32_t* ovf_list;  // their slice records (indices into tmps)
  uint16_t* rest16;    // non-NULL: a 16-bit copy of the non-key tuples as well
};

__device__ __forceinline__ uint32_t pt_cas32(uint32_t* p, uint32_t cmp, uint32_t v) {
  __hip_atomic_compare_exchange_strong(p, &cmp, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  return cmp;
}

// Group the n nonzeros of one bucket (positions [bs, bs + n) of the
// partitioned arrays) with a table of cap >= n slots.
// Returns false (before writing anything outside LDS) when the LDS table
// fills up: the bucket is then grouped again in its global region.  The
// grouped layout and the slice offsets are positions less obase (the
// table's first position when several tables share the partitioned arrays).
template <bool GLOBAL, class RestT, class ValT>
__device__ bool pt_group_bucket(unsigned long long* tk, uint32_t* tc, uint32_t cap, uint64_t bs, uint32_t n,
                                uint64_t obase,
                                const unsigned long long* __restrict__ pkey, const uint32_t* __restrict__ pid,
                                const KtPay<RestT, ValT>* __restrict__ ppay, KtSlice* __restrict__ tmps,
                                uint32_t* __restrict__ nkeys_out, uint32_t* __restrict__ perm,
                                RestT* __restrict__ rest_out, ValT* __restrict__ gval, const PtIndex& X,
                                unsigned long long* s_carry, unsigned long long* s_wave, uint32_t* s_ndist,
                                unsigned long long* stage8 = nullptr, uint32_t* stage4 = nullptr) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (uint32_t s = threadIdx.x; s < cap; s += blockDim.x) {
    pt_st64(&tk[s], PTC_EMPTY_KEY);
    pt_st32(&tc[s], 0u);
  }
  if (X.region && X.map) {
    for (uint32_t s = threadIdx.x; s < (uint32_t)(X.r1 - X.r0); s += blockDim.x) X.map[s] = 0xFFFFFFFFu;
  } else if (X.region) {
    for (uint64_t s = X.r0 + threadIdx.x; s < X.r1; s += blockDim.x) {
      pt_st64(&X.tab[s].key, PTC_EMPTY_KEY);
      X.tab[s].off = 0;
      X.tab[s].len = 0;
    }
  }
  if (threadIdx.x == 0) {
    *s_carry = 0;
    *s_ndist = 0;
  }
  __syncthreads();
  // counts per key (four keys per thread loaded at a time); a bucket of at
  // most four nonzeros a thread keeps each one's slot in registers for the
  // placing pass (no second read of its key, no second probe)
  constexpr int U = 4;
  const bool kept = n <= U * blockDim.x;
  // a bucket of at most PT_LDS_KEYS nonzeros cannot fill the LDS table: its
  // inserts skip the distinct-key count (an LDS atomic on one word per key)
  const bool track = n > PT_LDS_KEYS;
  uint32_t slot[U];
  for (uint32_t i0 = threadIdx.x; i0 < n; i0 += U * blockDim.x) {
    unsigned long long k[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = i0 + u * blockDim.x;
      k[u] = i < n ? pkey[bs + i] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      slot[u] = PT_FULL;
      if (i0 + u * blockDim.x < n) {
        const uint32_t s = pt_slot<GLOBAL>(tk, cap, k[u], GLOBAL || !track ? nullptr : s_ndist);
        if (s != PT_FULL) pt_add32(&tc[s], 1u);
        slot[u] = s;
      }
    }
  }
  __syncthreads();
  if (!GLOBAL && track && pt_ld32(s_ndist) >= PT_LDS_KEYS) return false;  // the same value in every thread
  // counts -> offsets in slot order; (slices << 32 | nonzeros) scanned
  // together, SPT consecutive slots per thread, a chunk of SPT * blockDim
  // slots at a time (4 x 1024 covers the LDS table in one chunk)
  constexpr int SPT = 4;
  const uint32_t chunk = SPT * blockDim.x;
  for (uint32_t c0 = 0; c0 < cap; c0 += chunk) {
    const uint32_t s0 = c0 + SPT * threadIdx.x;
    unsigned long long v[SPT], sum = 0;
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
      const uint32_t s = s0 + j;
      v[j] = 0;
      if (s < cap) {
        const uint32_t c = pt_ld32(&tc[s]);
        if (c) v[j] = (1ull << 32) | c;
      }
      sum += v[j];
    }
    unsigned long long inc = sum;  // inclusive scan over the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const unsigned long long x = __shfl_up(inc, o, 64);
      if (lane >= o) inc += x;
    }
    if (lane == 63) s_wave[wid] = inc;
    __syncthreads();
    unsigned long long pre = *s_carry;
    for (int w = 0; w < wid; ++w) pre += s_wave[w];
    pre += inc - sum;
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
      const uint32_t s = s0 + j;
      if (v[j]) {
        KtSlice sl;
        sl.key = pt_ld64(&tk[s]);
        sl.off = (uint32_t)(bs - obase + (uint32_t)pre);
        sl.len = (uint32_t)v[j];
        const uint64_t ti = bs + (pre >> 32);
        tmps[ti] = sl;
        pt_st32(&tc[s], (uint32_t)pre);  // the key's offset: the placing cursor
        if (X.region && X.map) {
          // the region's image in LDS; written out whole below
          const uint32_t R = (uint32_t)(X.r1 - X.r0), lsi = (uint32_t)(pre >> 32);
          uint32_t e = (uint32_t)(kt_home(mix64(sl.key), X.cap) - X.r0);  // < R: its top bits are the bucket's
          while (e < R && pt_cas32(&X.map[e], 0xFFFFFFFFu, lsi) != 0xFFFFFFFFu) ++e;
          if (e >= R) X.ovf_list[atomicAdd(X.ovf_cnt, 1u)] = (uint32_t)ti;  // the last pass inserts it
        } else if (X.region) {
          uint64_t e = kt_home(mix64(sl.key), X.cap);  // inside [r0, r1): its top bits are the bucket's
          while (true) {
            if (e >= X.r1) {  // ran past the region: the last pass inserts it
              X.ovf_list[atomicAdd(X.ovf_cnt, 1u)] = (uint32_t)ti;
              break;
            }
            if (pt_cas64(&X.tab[e].key, PTC_EMPTY_KEY, sl.key) == PTC_EMPTY_KEY) {
              X.tab[e].off = sl.off;
              X.tab[e].len = sl.len;
              break;
            }
            ++e;
          }
        }
      }
      pre += v[j];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t = *s_carry;
      for (int w = 0; w < nw; ++w) t += s_wave[w];
      *s_carry = t;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *nkeys_out = (uint32_t)(*s_carry >> 32);
  if (X.region && X.map) {  // the region of the probe index, written in order
    // (R <= PT_TAB: at most four slots a thread, their slice reads in flight
    // together)
    const uint32_t R = (uint32_t)(X.r1 - X.r0);
    constexpr int W = PT_TAB / PT_GROUP_THREADS;
    for (uint32_t s0 = threadIdx.x; s0 < R; s0 += W * blockDim.x) {
      uint32_t j[W];
      KtSlice sl[W];
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const uint32_t s = s0 + w * blockDim.x;
        j[w] = s < R ? X.map[s] : 0xFFFFFFFFu;
      }
#pragma unroll
      for (int w = 0; w < W; ++w) {
        sl[w].key = PTC_EMPTY_KEY;
        sl[w].off = 0;
        sl[w].len = 0;
        if (j[w] != 0xFFFFFFFFu) sl[w] = tmps[bs + j[w]];
      }
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const uint32_t s = s0 + w * blockDim.x;
        if (s >= R) continue;
        KtEntry en;
        en.key = sl[w].key;
        en.off = sl[w].off;
        en.len = sl[w].len;
        X.tab[X.r0 + s] = en;
      }
    }
  }
  if (stage8 && kept && n <= PT_TAB) {
    // place through LDS: every nonzero's value and id go to its key's
    // cursor in LDS (the key table's words, free once the slices are out,
    // and the region map's), then out in order with coalesced stores; the
    // non-key tuples in the cursor words (4 bytes) or in a second round
    // (8 bytes).  Scattered stores of
    // 2-8 bytes to the bucket's range cost one cache-line access per lane.
    static_assert(PT_TAB * sizeof(unsigned long long) >= PT_TAB * sizeof(RestT) &&
                      PT_TAB * sizeof(unsigned long long) >= PT_TAB * sizeof(ValT),
                  "a bucket's values and tuples fit the staging words");
    ValT* sv = reinterpret_cast<ValT*>(stage8);
    RestT* sr = reinterpret_cast<RestT*>(stage8);
    uint32_t lp[U];
    RestT rk[U];
    {
      uint32_t id[U];
      KtPay<RestT, ValT> pl[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t i = threadIdx.x + u * blockDim.x;
        const uint64_t q = bs + (i < n ? i : 0u);
        pl[u] = ppay[q];
        if constexpr (kPayId<RestT, ValT>) id[u] = pl[u].id;
        else id[u] = pid[q];
      }
      __syncthreads();  // the region write-out has read the map
#pragma unroll
      for (int u = 0; u < U; ++u) {
        lp[u] = 0xFFFFFFFFu;
        if (threadIdx.x + u * blockDim.x >= n) continue;
        lp[u] = pt_add32(&tc[slot[u]], 1u);
        sv[lp[u]] = pl[u].v;
        stage4[lp[u]] = id[u];
        rk[u] = pl[u].k;
      }
    }
    __syncthreads();
    const uint64_t base = bs - obase;
    if constexpr (sizeof(RestT) == sizeof(uint32_t)) {
      // 4-byte tuples: staged in the cursor words, free once every cursor
      // has been taken; one round out for all three arrays
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (lp[u] != 0xFFFFFFFFu) tc[lp[u]] = (uint32_t)rk[u];
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const RestT k = (RestT)tc[i];
        perm[base + i] = stage4[i];
        gval[base + i] = sv[i];
        rest_out[base + i] = k;
        if (X.rest16) X.rest16[base + i] = (uint16_t)k;
      }
      return true;
    }
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
      perm[base + i] = stage4[i];
      gval[base + i] = sv[i];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (lp[u] != 0xFFFFFFFFu) sr[lp[u]] = rk[u];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
      const RestT k = sr[i];
      rest_out[base + i] = k;
      if (X.rest16) X.rest16[base + i] = (uint16_t)k;
    }
    return true;
  }
  // place: every nonzero at its key's cursor (loads of four at a time in
  // flight)
  constexpr int UP = U;
  for (uint32_t i0 = threadIdx.x; i0 < n; i0 += UP * blockDim.x) {
    unsigned long long k[UP];
    uint32_t id[UP];
    KtPay<RestT, ValT> pl[UP];
#pragma unroll
    for (int u = 0; u < UP; ++u) {
      const uint32_t i = i0 + u * blockDim.x;
      const uint64_t q = bs + (i < n ? i : 0u);
      k[u] = kept ? 0ull : pkey[q];
      pl[u] = ppay[q];
      if constexpr (kPayId<RestT, ValT>) id[u] = pl[u].id;
      else id[u] = pid[q];
    }
#pragma unroll
    for (int u = 0; u < UP; ++u) {
      if (i0 + u * blockDim.x >= n) continue;
      const uint32_t s = kept ? slot[u] : pt_slot<GLOBAL>(tk, cap, k[u]);  // present: a read-only probe
      const uint64_t p = bs - obase + pt_add32(&tc[s], 1u);
      perm[p] = id[u];
      rest_out[p] = pl[u].k;
      if (X.rest16) X.rest16[p] = (uint16_t)pl[u].k;
      gval[p] = pl[u].v;
    }
  }
  return true;
}

// LDS of one bucket workgroup (declared once per kernel, so a kernel that
// groups either key width still holds one copy)
struct PtGroupLds {
  unsigned long long lk[PT_TAB];
  uint32_t lc[PT_TAB];
  uint32_t lmap[PT_TAB];
  unsigned long long s_wave[PT_GROUP_THREADS / 64];
  unsigned long long s_carry;
  uint32_t s_ndist;
};

// Group bucket b of a table whose partitioned nonzeros start at obase and
// end at oend (positions in the shared partitioned arrays).
template <class RestT, class ValT>
__device__ __forceinline__ void pt_group_one(uint32_t b, const unsigned long long* __restrict__ off, uint64_t ntiles,
                                             uint32_t nb, uint64_t obase, uint64_t oend,
                                             const unsigned long long* __restrict__ pkey,
                                             const uint32_t* __restrict__ pid,
                                             const KtPay<RestT, ValT>* __restrict__ ppay,
                                             unsigned long long* __restrict__ gkey, uint32_t* __restrict__ gcnt,
                                             KtSlice* __restrict__ tmps, uint32_t* __restrict__ nkeys_out,
                                             uint32_t* __restrict__ perm, RestT* __restrict__ rest_out,
                                             ValT* __restrict__ gval, KtEntry* __restrict__ tab, uint64_t cap,
                                             int region, uint32_t* __restrict__ ovf_cnt,
                                             uint32_t* __restrict__ ovf_list, PtGroupLds& L,
                                             uint16_t* __restrict__ rest16 = nullptr) {
  const uint64_t bs = off[(uint64_t)b * ntiles];
  const uint64_t be = b + 1 < nb ? off[(uint64_t)(b + 1) * ntiles] : oend;
  const uint32_t n = (uint32_t)(be - bs);
  PtIndex X;
  X.tab = tab;
  X.cap = cap;
  X.region = region;
  X.r0 = region ? (cap / nb) * b : 0;
  X.r1 = region ? X.r0 + cap / nb : 0;
  X.map = region && cap / nb <= PT_TAB ? L.lmap : nullptr;
  X.ovf_cnt = ovf_cnt;
  X.ovf_list = ovf_list;
  X.rest16 = rest16;
  // the LDS table while the bucket's distinct keys fit, else the region
  if (!pt_group_bucket<false>(L.lk, L.lc, PT_TAB, bs, n, obase, pkey, pid, ppay, tmps, nkeys_out, perm, rest_out,
                              gval, X, &L.s_carry, L.s_wave, &L.s_ndist, L.lk, L.lmap)) {
    __syncthreads();
    pt_group_bucket<true>(gkey + bs, gcnt + bs, n, bs, n, obase, pkey, pid, ppay, tmps, nkeys_out, perm, rest_out,
                          gval, X, &L.s_carry, L.s_wave, &L.s_ndist);
  }
}

template <class RestT, class ValT>
__global__ void __launch_bounds__(PT_GROUP_THREADS, PT_GROUP_WAVES_PER_SIMD) k_pt_group(
    const unsigned long long* __restrict__ off, uint64_t ntiles, uint32_t nb, uint64_t N,
    const unsigned long long* __restrict__ pkey, const uint32_t* __restrict__ pid,
    const KtPay<RestT, ValT>* __restrict__ ppay, unsigned long long* __restrict__ gkey, uint32_t* __restrict__ gcnt,
    KtSlice* __restrict__ tmps, uint32_t* __restrict__ bkeys, uint32_t* __restrict__ perm,
    RestT* __restrict__ rest_out, ValT* __restrict__ gval, KtEntry* __restrict__ tab, uint64_t cap, int region,
    uint32_t* __restrict__ ovf_cnt, uint32_t* __restrict__ ovf_list, uint16_t* __restrict__ rest16) {
  __shared__ PtGroupLds L;
  pt_group_one<RestT, ValT>(blockIdx.x, off, ntiles, nb, 0, N, pkey, pid, ppay, gkey, gcnt, tmps, &bkeys[blockIdx.x],
                            perm, rest_out, gval, tab, cap, region, ovf_cnt, ovf_list, L, rest16);
}

// Slices of bucket b to their compact positions (sbase: the scanned key
// counts of the table's buckets, from its first); without region mode, the
// probe index as well (direct tables: entry[key]; hashed: one insert per
// distinct key).
__device__ __forceinline__ void pt_slices_one(uint32_t b, const unsigned long long* __restrict__ off,
                                              uint64_t ntiles, uint32_t nb, const uint32_t* __restrict__ bkeys,
                                              const unsigned long long* __restrict__ sbase,
                                              const KtSlice* __restrict__ tmps, KtSlice* __restrict__ slices,
                                              KtEntry* __restrict__ tab, uint64_t cap, int index, int direct,
                                              unsigned long long* __restrict__ total, uint64_t N) {
  if (b == 0 && threadIdx.x == 0) *total = ((sbase[nb] - sbase[0]) << 32) | N;
  const uint64_t bs = off[(uint64_t)b * ntiles], base = sbase[b] - sbase[0];
  const uint32_t nk = bkeys[b];
  for (uint32_t j = threadIdx.x; j < nk; j += blockDim.x) {
    const KtSlice sl = tmps[bs + j];
    slices[base + j] = sl;
    if (!index) continue;
    if (direct) {
      tab[sl.key].key = sl.key;
      tab[sl.key].off = sl.off;
      tab[sl.key].len = sl.len;
      continue;
    }
    uint64_t i = kt_home(mix64(sl.key), cap);
    while (true) {
      const unsigned long long old = atomicCAS(&tab[i].key, PTC_EMPTY_KEY, sl.key);
      if (old == PTC_EMPTY_KEY) break;  // keys are distinct: a claimed entry is never this key's
      i = (i + 1) & (cap - 1);
    }
    tab[i].off = sl.off;
    tab[i].len = sl.len;
  }
}

__global__ void k_pt_slices(const unsigned long long* __restrict__ off, uint64_t ntiles, uint32_t nb,
                            const uint32_t* __restrict__ bkeys, const unsigned long long* __restrict__ sbase,
                            const KtSlice* __restrict__ tmps, KtSlice* __restrict__ slices, KtEntry* __restrict__ tab,
                            uint64_t cap, int index, int direct, unsigned long long* __restrict__ total, uint64_t N) {
  pt_slices_one(blockIdx.x, off, ntiles, nb, bkeys, sbase, tmps, slices, tab, cap, index, direct, total, N);
}

// Keys whose probe ran past their region: inserted once every region is
// written, onward from their home line (the table keeps free slots: cap >
// distinct keys).
__global__ void k_pt_overflow(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ list,
                              const KtSlice* __restrict__ tmps, KtEntry* __restrict__ tab, uint64_t cap) {
  const uint32_t n = *cnt;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const KtSlice sl = tmps[list[k]];
    uint64_t i = kt_home(mix64(sl.key), cap);
    while (true) {
      const unsigned long long old = atomicCAS(&tab[i].key, PTC_EMPTY_KEY, sl.key);
      if (old == PTC_EMPTY_KEY) break;
      i = (i + 1) & (cap - 1);
    }
    tab[i].off = sl.off;
    tab[i].len = sl.len;
  }
}

// The last step of a single partitioned table, one workgroup per bucket:
// the bucket's slices to their compact positions (the bucket's place among
// the compact slices is the sum of the earlier buckets' key counts, which
// the workgroup adds up itself), the probe index without region mode, the
// key total (the last bucket); then, spread over all workgroups, the keys
// whose probe ran past their region (ovf_cnt non-NULL), inserted from their
// home line now that every region is written.  One launch for what a scan,
// pt_slices_one and k_pt_overflow do in three.
__global__ void __launch_bounds__(256) k_pt_finish(const unsigned long long* __restrict__ start, uint32_t nb,
                                                   const uint32_t* __restrict__ bkeys,
                                                   const KtSlice* __restrict__ tmps, KtSlice* __restrict__ slices,
                                                   KtEntry* __restrict__ tab, uint64_t cap, int index, int direct,
                                                   unsigned long long* __restrict__ total, uint64_t N,
                                                   const uint32_t* __restrict__ ovf_cnt,
                                                   const uint32_t* __restrict__ ovf_list) {
  __shared__ unsigned long long s_red[256 / 64];
  const uint32_t b = blockIdx.x;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned long long pre = 0;
  for (uint32_t j = threadIdx.x; j < b; j += blockDim.x) pre += bkeys[j];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) pre += __shfl_xor(pre, o, 64);
  if (lane == 0) s_red[wid] = pre;
  __syncthreads();
  pre = s_red[0] + s_red[1] + s_red[2] + s_red[3];
  const uint64_t bs = start[b];
  const uint32_t nk = bkeys[b];
  if (b + 1 == nb && threadIdx.x == 0) *total = ((pre + nk) << 32) | N;
  for (uint32_t j = threadIdx.x; j < nk; j += blockDim.x) {
    const KtSlice sl = tmps[bs + j];
    slices[pre + j] = sl;
    if (!index) continue;
    if (direct) {
      tab[sl.key].key = sl.key;
      tab[sl.key].off = sl.off;
      tab[sl.key].len = sl.len;
      continue;
    }
    uint64_t i = kt_home(mix64(sl.key), cap);
    while (true) {
      const unsigned long long old = atomicCAS(&tab[i].key, PTC_EMPTY_KEY, sl.key);
      if (old == PTC_EMPTY_KEY) break;  // keys are distinct: a claimed entry is never this key's
      i = (i + 1) & (cap - 1);
    }
    tab[i].off = sl.off;
    tab[i].len = sl.len;
  }
  if (!ovf_cnt) return;
  const uint32_t no = *ovf_cnt;
  for (uint32_t k = b * blockDim.x + threadIdx.x; k < no; k += nb * blockDim.x) {
    const KtSlice sl = tmps[ovf_list[k]];
    uint64_t i = kt_home(mix64(sl.key), cap);
    while (true) {
      const unsigned long long old = atomicCAS(&tab[i].key, PTC_EMPTY_KEY, sl.key);
      if (old == PTC_EMPTY_KEY) break;
      i = (i + 1) & (cap - 1);
    }
    tab[i].off = sl.off;
    tab[i].len = sl.len;
  }
}

// ---------------------------------------------------------- many tables
//
// A set of FP64 tables built together (a batch plan): the same counting
// and partitioned steps, with one launch per step for all tables.  Every
// table's count matrix, tiles, positions and buckets are laid end to end;
// a workgroup finds its table by a binary search over the bases.  The
// descriptors are read through global-address-space pointers so the
// kernels' loads and stores stay global (not flat) instructions.

#define PTC_MT_FIELDS(PTR)                                                                      \
  PTR(const uint32_t) hedges;                                                                   \
  PTR(const double) vals;                                                                       \
  PTR(uint32_t) perm;                                                                           \
  PTR(uint32_t) rest32;                                                                         \
  PTR(unsigned long long) rest64;                                                               \
  PTR(double) gval;                                                                             \
  PTR(KtEntry) entries;                                                                         \
  PTR(KtSlice) slices;                                                                          \
  PTR(uint32_t) slice_of;                                                                       \
  PTR(uint16_t) rest16;                                                                         \
  PTR(unsigned long long) total;                                                                \
  PTR(uint32_t) cidx;                                                                           \
  KeyCodec codec;                                                                               \
  KeyCodec rest;                                                                                \
  uint64_t N, cap, ntiles, tile_base, cnt_base, item_base, bucket_base;                         \
  uint32_t d, nb;                                                                               \
  int bbits, counting, direct, region, wide;

#define PTC_MT_PTR(T) T*
struct MtTable {
  PTC_MT_FIELDS(PTC_MT_PTR)
};
#ifdef __HIP_DEVICE_COMPILE__
#define PTC_MT_GPTR(T) __attribute__((address_space(1))) T*
#else
#define PTC_MT_GPTR(T) T*
#endif
struct MtTableG {
  PTC_MT_FIELDS(PTC_MT_GPTR)
};
static_assert(sizeof(MtTableG) == sizeof(MtTable), "device view of MtTable");

// the last table whose base (tile, bucket or position) is <= g
enum { MT_TILE = 0, MT_BUCKET = 1, MT_ITEM = 2 };
template <int WHICH>
__device__ __forceinline__ int mt_find(const MtTableG* __restrict__ T, int nt, uint64_t g) {
  int lo = 0, hi = nt - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const uint64_t b = WHICH == MT_TILE ? T[mid].tile_base : WHICH == MT_BUCKET ? T[mid].bucket_base : T[mid].item_base;
    if (b <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(PT_THREADS) k_mt_hist(const MtTableG* __restrict__ tabs, int nt,
                                                        uint32_t* __restrict__ hist) {
  const MtTableG& T = tabs[mt_find<MT_TILE>(tabs, nt, blockIdx.x)];
  const KeyCodec codec = T.codec;
  pt_hist_tile(T.hedges, T.d, T.N, codec, T.bbits, T.nb, T.ntiles, hist + T.cnt_base, blockIdx.x - T.tile_base);
}

// Direct tables grouped by partitioning write only their present keys:
// the rest of their index is cleared first.
__global__ void k_mt_clear(const MtTableG* __restrict__ tabs, int nt) {
  const MtTableG& T = tabs[blockIdx.y];
  if (T.counting || !T.direct) return;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < T.cap; i += (uint64_t)gridDim.x * blockDim.x) {
    KtEntry en;
    en.key = PTC_EMPTY_KEY;
    en.off = 0;
    en.len = 0;
    T.entries[i] = en;
  }
}

__global__ void __launch_bounds__(1024) k_mt_pd_slices(const MtTableG* __restrict__ tabs, int nt,
                                                       const unsigned long long* __restrict__ off) {
  const MtTableG& T = tabs[blockIdx.x];
  if (!T.counting) return;
  pd_slices_one(off + T.cnt_base, T.ntiles, T.nb, T.N, T.item_base, T.slices, T.entries, T.cidx, T.total);
}

__global__ void __launch_bounds__(PT_THREADS) k_mt_scatter(const MtTableG* __restrict__ tabs, int nt,
                                                           uint64_t gtiles, const uint32_t* __restrict__ hist,
                                                           const unsigned long long* __restrict__ off,
                                                           unsigned long long* __restrict__ pkey,
                                                           uint32_t* __restrict__ pid, void* __restrict__ ppay) {
  const uint64_t g = pt_tile_of(blockIdx.x, gtiles);
  if (g >= gtiles) return;
  const MtTableG& T = tabs[mt_find<MT_TILE>(tabs, nt, g)];
  const uint64_t lt = g - T.tile_base;
  const KeyCodec codec = T.codec, rest = T.rest;
  const uint32_t* hs = hist + T.cnt_base;
  const unsigned long long* os = off + T.cnt_base;
  if (T.counting) {
    if (T.wide)
      pd_scatter_tile<unsigned long long, double>(T.hedges, T.d, T.N, codec, rest, T.vals, T.nb, T.ntiles, hs, os,
                                                  T.item_base, T.cidx, T.perm, T.rest64, T.gval, T.slice_of, lt);
    else
      pd_scatter_tile<uint32_t, double>(T.hedges, T.d, T.N, codec, rest, T.vals, T.nb, T.ntiles, hs, os, T.item_base,
                                        T.cidx, T.perm, T.rest32, T.gval, T.slice_of, lt, T.rest16);
  } else {
    if (T.wide)
      pt_scatter_tile<unsigned long long, double>(T.hedges, T.d, T.N, codec, rest, T.vals, T.bbits, T.nb, T.ntiles, hs,
                                                  os, pkey, pid, (KtPay<unsigned long long, double>*)ppay, lt);
    else
      pt_scatter_tile<uint32_t, double>(T.hedges, T.d, T.N, codec, rest, T.vals, T.bbits, T.nb, T.ntiles, hs, os, pkey,
                                        pid, (KtPay<uint32_t, double>*)ppay, lt);
  }
}

__global__ void __launch_bounds__(PT_GROUP_THREADS, PT_GROUP_WAVES_PER_SIMD) k_mt_group(
    const MtTableG* __restrict__ tabs, int nt, const unsigned long long* __restrict__ off,
    const unsigned long long* __restrict__ pkey, const uint32_t* __restrict__ pid, const void* __restrict__ ppay,
    unsigned long long* __restrict__ gkey, uint32_t* __restrict__ gcnt, KtSlice* __restrict__ tmps,
    uint32_t* __restrict__ bkeys, uint32_t* __restrict__ ovf_cnt, uint32_t* __restrict__ ovf_list) {
  const MtTableG& T = tabs[mt_find<MT_BUCKET>(tabs, nt, blockIdx.x)];
  const uint32_t b = (uint32_t)(blockIdx.x - T.bucket_base);
  const unsigned long long* os = off + T.cnt_base;
  __shared__ PtGroupLds L;
  if (T.wide)
    pt_group_one<unsigned long long, double>(b, os, T.ntiles, T.nb, T.item_base, T.item_base + T.N, pkey, pid,
                                             (const KtPay<unsigned long long, double>*)ppay, gkey, gcnt, tmps,
                                             &bkeys[blockIdx.x], T.perm, T.rest64, T.gval, T.entries, T.cap, T.region,
                                             ovf_cnt, ovf_list, L);
  else
    pt_group_one<uint32_t, double>(b, os, T.ntiles, T.nb, T.item_base, T.item_base + T.N, pkey, pid,
                                   (const KtPay<uint32_t, double>*)ppay, gkey, gcnt, tmps, &bkeys[blockIdx.x], T.perm,
                                   T.rest32, T.gval, T.entries, T.cap, T.region, ovf_cnt, ovf_list, L, T.rest16);
}

__global__ void k_mt_slices(const MtTableG* __restrict__ tabs, int nt, const unsigned long long* __restrict__ off,
                            const uint32_t* __restrict__ bkeys, const unsigned long long* __restrict__ sbase,
                            const KtSlice* __restrict__ tmps) {
  const MtTableG& T = tabs[mt_find<MT_BUCKET>(tabs, nt, blockIdx.x)];
  pt_slices_one((uint32_t)(blockIdx.x - T.bucket_base), off + T.cnt_base, T.ntiles, T.nb, bkeys + T.bucket_base,
                sbase + T.bucket_base, tmps, T.slices, T.entries, T.cap, T.region ? 0 : 1, T.direct, T.total, T.N);
}

__global__ void k_mt_overflow(const MtTableG* __restrict__ tabs, int nt, const uint32_t* __restrict__ cnt,
                              const uint32_t* __restrict__ list, const KtSlice* __restrict__ tmps) {
  const uint32_t n = *cnt;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const MtTableG& T = tabs[mt_find<MT_ITEM>(tabs, nt, list[k])];
    const KtSlice sl = tmps[list[k]];
    const uint64_t cap = T.cap;
    uint64_t i = kt_home(mix64(sl.key), cap);
    while (true) {
      const unsigned long long old = atomicCAS(&T.entries[i].key, PTC_EMPTY_KEY, sl.key);
      if (old == PTC_EMPTY_KEY) break;
      i = (i + 1) & (cap - 1);
    }
    T.entries[i].off = sl.off;
    T.entries[i].len = sl.len;
  }
}

}  // namespace ptc

using namespace ptc;

namespace {

// Sort grouping of a table whose arrays (entries, perm, gval/gval32,
// rest_key32/64) are allocated; N > 0 and the key space fits 32 bits.
int build_sorted(ptc_ctx* ctx, ptc_hashtable* ht, const ptc_dtensor* t, bool f32, bool rest64) {
  const uint64_t N = t->N;
  hipStream_t st = ctx->stream;
  Workspace ws(ctx);
  const uint64_t ntiles = (N + RS_TILE - 1) / RS_TILE;
  // LSD passes over the significant bits of the key space; the ids of the
  // last pass land in perm
  int bits = 1;
  while (bits < 32 && (1ull << bits) < ht->codec.space) ++bits;
  const bool odd = ((bits + 7) / 8) % 2 == 1;
  uint32_t *k0 = ws.get<uint32_t>(N), *k1 = ws.get<uint32_t>(N);
  uint32_t *i0 = odd ? ws.get<uint32_t>(N) : ht->perm, *i1 = odd ? ht->perm : ws.get<uint32_t>(N);
  uint32_t* hist = ws.get<uint32_t>(256 * ntiles);
  unsigned long long* hs = ws.get<unsigned long long>(256 * ntiles + 1);
  uint32_t* head = ws.get<uint32_t>(N);
  unsigned long long* sx = ws.get<unsigned long long>(N + 1);
  const size_t pay_bytes = rest64 || !f32 ? 16 : 8;
  void* pay = ws.get<unsigned char>(N * pay_bytes);
  ht->slice_of = dalloc_n<uint32_t>(ctx, N);
  ht->d_total = dalloc_n<unsigned long long>(ctx, 1);
  ht->slices_cap = std::max<uint64_t>(1, std::min<uint64_t>(N, ht->cap));
  ht->slices = dalloc_n<KtSlice>(ctx, ht->slices_cap);
  if (!k0 || !i0 || !k1 || !i1 || !hist || !hs || !head || !sx || !pay || !ht->slice_of || !ht->d_total || !ht->slices)
    return PTC_ERR_NOMEM;
  const unsigned g = grid_for(N, 256, 16384);
#define PTC_RSKEYS(RT, VT, VALS)                                                                             \
  hipLaunchKernelGGL((k_rs_keys<RT, VT>), dim3(g), dim3(256), 0, st, t->hedges, t->d, N, ht->codec, ht->rest, \
                     (const VT*)(VALS), k0, i0, (KtPay<RT, VT>*)pay)
  if (rest64) {
    if (f32) PTC_RSKEYS(unsigned long long, float, t->vals32);
    else PTC_RSKEYS(unsigned long long, double, t->vals);
  } else {
    if (f32) PTC_RSKEYS(uint32_t, float, t->vals32);
    else PTC_RSKEYS(uint32_t, double, t->vals);
  }
#undef PTC_RSKEYS
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, st, (const uint32_t*)k0, N, shift,
                       ntiles, hist);
    int rc = scan_u32_to_u64(ctx, hist, hs, 256 * ntiles, hs + 256 * ntiles);
    if (rc) return rc;
    hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, st, (const uint32_t*)k0,
                       (const uint32_t*)i0, N, shift, ntiles, (const unsigned long long*)hs, k1, i1);
    std::swap(k0, k1);
    std::swap(i0, i1);
  }
  // slices from runs of equal keys
  hipLaunchKernelGGL(k_rs_heads, dim3(g), dim3(256), 0, st, (const uint32_t*)k0, N, head);
  int rc = scan_u32_to_u64(ctx, head, sx, N, sx + N);
  if (rc) return rc;
  hipLaunchKernelGGL(k_rs_slices, dim3(g), dim3(256), 0, st, (const uint32_t*)k0, N, (const uint32_t*)head,
                     (const unsigned long long*)sx, ht->slices, ht->slice_of);
  hipLaunchKernelGGL(k_kt_clear, dim3(grid_for(ht->cap, 256)), dim3(256), 0, st, ht->entries, (uint32_t*)nullptr,
                     ht->cap);
  hipLaunchKernelGGL(k_rs_index, dim3(grid_for(ht->slices_cap, 256, 16384)), dim3(256), 0, st, ht->slices,
                     (const unsigned long long*)(sx + N), ht->entries, ht->cap, ht->direct ? 1 : 0, ht->d_total, N);
  // the streaming layout from the payload, by the ids in grouped order (i0 == perm)
#define PTC_GATHER(RT, VT, RK, GV)                                                                         \
  hipLaunchKernelGGL((k_kt_gather_pay<RT, VT>), dim3(g), dim3(256), 0, st, N, (const uint32_t*)ht->perm, \
                     (const KtPay<RT, VT>*)pay, RK, GV)
  if (rest64) {
    if (f32) PTC_GATHER(unsigned long long, float, ht->rest_key64, ht->gval32);
    else PTC_GATHER(unsigned long long, double, ht->rest_key64, ht->gval);
  } else {
    if (f32) PTC_GATHER(uint32_t, float, ht->rest_key32, ht->gval32);
    else PTC_GATHER(uint32_t, double, ht->rest_key32, ht->gval);
  }
#undef PTC_GATHER
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

// The counting pass and the column scan of a single table's tiles: every
// (tile, bucket) gets its count (tcnt) and offset in its segment (toff), every
// (segment, bucket) its prefix (segpre), every bucket its start (start); a
// direct table also gets its slices, index, compact slice indices and total.
int place_offsets(ptc_ctx* ctx, const ptc_dtensor* t, const KeyCodec& codec, int bbits, uint32_t nb, uint64_t ntiles,
                  uint32_t nseg, uint32_t* tcnt, uint32_t* toff, uint32_t* segpre, uint32_t* btot,
                  unsigned long long* start, int direct, KtSlice* slices, KtEntry* tab, uint32_t* cidx,
                  unsigned long long* total, uint32_t* zero_word) {
  hipStream_t st = ctx->stream;
  hipLaunchKernelGGL(k_pc_count, dim3((unsigned)(8 * ((ntiles + 7) / 8))), dim3(PT_THREADS), 0, st, t->hedges, t->d,
                     t->N, codec, bbits, nb, ntiles, tcnt);
  const uint64_t waves = (uint64_t)nseg * ((nb + 63) / 64);
  hipLaunchKernelGGL(k_pc_colscan, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st, (const uint32_t*)tcnt, nb,
                     ntiles, toff, segpre);
  hipLaunchKernelGGL(k_pc_segscan, dim3((nb + 255) / 256), dim3(256), 0, st, segpre, nb, nseg, btot);
  hipLaunchKernelGGL(k_pc_starts, dim3(1), dim3(1024), 0, st, (const uint32_t*)btot, nb, t->N, start, direct, slices,
                     tab, cidx, total, zero_word);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

// Partitioned hashing (wide keys): partition by hash bucket, group each
// bucket in LDS, compact the slices and write the probe index.
int build_partitioned(ptc_ctx* ctx, ptc_hashtable* ht, const ptc_dtensor* t, bool f32, bool rest64) {
  const uint64_t N = t->N;
  hipStream_t st = ctx->stream;
  uint32_t nb = 1;
  int bbits = 0;
  while (nb < PT_MAX_BUCKETS && (uint64_t)nb * PT_BUCKET_ITEMS < N) {
    nb <<= 1;
    ++bbits;
  }
  const uint64_t ntiles = (N + PT_TILE - 1) / PT_TILE;
  // region mode: every bucket owns cap / nb slots of the probe index
  const bool region = !ht->direct && ht->cap / 8 >= nb;
  Workspace ws(ctx);
  const size_t pay_bytes = rest64 || !f32 ? 16 : 8;
  unsigned long long* off = ws.get<unsigned long long>(nb + 1);  // bucket starts
  uint32_t* tcnt = ws.get<uint32_t>((uint64_t)nb * ntiles);      // (tile, bucket) counts, tile-major
  uint32_t* toff = ws.get<uint32_t>((uint64_t)nb * ntiles);      // (tile, bucket) offsets in its segment
  const uint32_t nseg = (uint32_t)((ntiles + PC_SEG - 1) / PC_SEG);
  uint32_t* segpre = ws.get<uint32_t>((uint64_t)nb * nseg);       // (segment, bucket) totals, then prefixes
  uint32_t* btot = ws.get<uint32_t>(nb);                          // bucket totals
  unsigned long long* pkey = ws.get<unsigned long long>(N);
  uint32_t* pid = ws.get<uint32_t>(N);
  void* ppay = ws.get<unsigned char>(N * pay_bytes);
  unsigned long long* gkey = ws.get<unsigned long long>(N);
  uint32_t* gcnt = ws.get<uint32_t>(N);
  KtSlice* tmps = ws.get<KtSlice>(N);
  uint32_t* bkeys = ws.get<uint32_t>(nb);
  uint32_t* ovf = ws.get<uint32_t>(1);         // keys whose probe ran past their region: count,
  uint32_t* ovf_list = ws.get<uint32_t>(N);    // and list
  ht->d_total = dalloc_n<unsigned long long>(ctx, 1);
  ht->slices_cap = std::max<uint64_t>(1, std::min<uint64_t>(N, ht->cap));
  ht->slices = dalloc_n<KtSlice>(ctx, ht->slices_cap);
  // non-key tuples of at most 16 bits (the column keys of a dense
  // contraction) are also written as 16-bit words
  if (!rest64 && ht->rest.space <= 65536 && !(ht->rest_key16 = dalloc_n<uint16_t>(ctx, std::max<uint64_t>(N, 1))))
    return PTC_ERR_NOMEM;
  if (!tcnt || !toff || !segpre || !btot || !off || !pkey || !pid || !ppay || !gkey || !gcnt || !tmps || !bkeys || !ovf ||
      !ovf_list ||
      !ht->d_total || !ht->slices)
    return PTC_ERR_NOMEM;
  // (the overflow count is cleared by the offsets' last launch)
  int rc = place_offsets(ctx, t, ht->codec, bbits, nb, ntiles, nseg, tcnt, toff, segpre, btot, off, 0, nullptr, nullptr,
                         nullptr, nullptr, ovf);
  if (rc) return rc;
  if (!region)
    hipLaunchKernelGGL(k_kt_clear, dim3(grid_for(ht->cap, 256)), dim3(256), 0, st, ht->entries, (uint32_t*)nullptr,
                       ht->cap);
#define PTC_PLACE(RT, VT, KT, VALS)                                                                                    \
  do {                                                                                                               \
    const size_t lds = (size_t)PT_TILE * (sizeof(KT) + sizeof(RT) + 2) + (size_t)nb * 8;                          \
    static size_t lds_set = 0;                                                                                       \
    if (lds > 65536 && lds > lds_set) {                                                                              \
      PTC_HIP(ctx, hipFuncSetAttribute((const void*)k_pt_place<RT, VT, KT>,                                         \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));                      \
      lds_set = lds;                                                                                                 \
    }                                                                                                                \
    hipLaunchKernelGGL((k_pt_place<RT, VT, KT>), dim3((unsigned)(8 * ((ntiles + 7) / 8))), dim3(PT_THREADS), lds,    \
                       st, t->hedges, t->d, N,                                                                      \
                       ht->codec, ht->rest, (const VT*)(VALS), bbits, nb, ntiles, (const unsigned long long*)off,   \
                       (const uint32_t*)tcnt, (const uint32_t*)toff, (const uint32_t*)segpre, pkey, pid,            \
                       (KtPay<RT, VT>*)ppay);                                                                       \
  } while (0)
#define PTC_PT(RT, VT, VALS, RK, GV, R16)                                                                              \
  do {                                                                                                               \
    if (ht->codec.space <= (1ull << 32)) PTC_PLACE(RT, VT, uint32_t, VALS);                                         \
    else PTC_PLACE(RT, VT, unsigned long long, VALS);                                                                \
    hipLaunchKernelGGL((k_pt_group<RT, VT>), dim3(nb), dim3(PT_GROUP_THREADS), 0, st,                              \
                       (const unsigned long long*)off,                                                              \
                       (uint64_t)1, nb, N, (const unsigned long long*)pkey, (const uint32_t*)pid,                   \
                       (const KtPay<RT, VT>*)ppay, gkey, gcnt, tmps, bkeys, ht->perm, RK, GV, ht->entries, ht->cap, \
                       region ? 1 : 0, ovf, ovf_list, R16);                                                          \
  } while (0)
  if (rest64) {
    if (f32) PTC_PT(unsigned long long, float, t->vals32, ht->rest_key64, ht->gval32, (uint16_t*)nullptr);
    else PTC_PT(unsigned long long, double, t->vals, ht->rest_key64, ht->gval, (uint16_t*)nullptr);
  } else {
    if (f32) PTC_PT(uint32_t, float, t->vals32, ht->rest_key32, ht->gval32, ht->rest_key16);
    else PTC_PT(uint32_t, double, t->vals, ht->rest_key32, ht->gval, ht->rest_key16);
  }
#undef PTC_PT
#undef PTC_PLACE
  hipLaunchKernelGGL(k_pt_finish, dim3(nb), dim3(256), 0, st, (const unsigned long long*)off, nb,
                     (const uint32_t*)bkeys, (const KtSlice*)tmps, ht->slices, ht->entries, ht->cap,
                     region ? 0 : 1, ht->direct ? 1 : 0, ht->d_total, N, region ? (const uint32_t*)ovf : nullptr,
                     (const uint32_t*)ovf_list);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

// Small direct key spaces (<= PT_MAX_BUCKETS keys): one counting pass.
int build_direct_partitioned(ptc_ctx* ctx, ptc_hashtable* ht, const ptc_dtensor* t, bool f32, bool rest64) {
  const uint64_t N = t->N;
  hipStream_t st = ctx->stream;
  const uint32_t nb = (uint32_t)ht->codec.space;
  const uint64_t ntiles = (N + PT_TILE - 1) / PT_TILE;
  Workspace ws(ctx);
  uint32_t* tcnt = ws.get<uint32_t>((uint64_t)nb * ntiles);  // (tile, key) counts, tile-major
  uint32_t* toff = ws.get<uint32_t>((uint64_t)nb * ntiles);  // (tile, key) offsets in its segment
  const uint32_t nseg = (uint32_t)((ntiles + PC_SEG - 1) / PC_SEG);
  uint32_t* segpre = ws.get<uint32_t>((uint64_t)nb * nseg);   // (segment, key) totals, then prefixes
  uint32_t* btot = ws.get<uint32_t>(nb);                      // key totals
  unsigned long long* off = ws.get<unsigned long long>(nb + 1);
  uint32_t* cidx = ws.get<uint32_t>(nb);
  // (the slice index of every position is built on first use: only a
  // contraction that wants per-row features reads it)
  ht->d_total = dalloc_n<unsigned long long>(ctx, 1);
  ht->slices_cap = std::max<uint64_t>(1, std::min<uint64_t>(N, ht->cap));
  ht->slices = dalloc_n<KtSlice>(ctx, ht->slices_cap);
  if (!rest64 && ht->rest.space <= 65536 && !(ht->rest_key16 = dalloc_n<uint16_t>(ctx, std::max<uint64_t>(N, 1))))
    return PTC_ERR_NOMEM;
  if (!tcnt || !toff || !segpre || !btot || !off || !cidx || !ht->d_total || !ht->slices) return PTC_ERR_NOMEM;
  int rc = place_offsets(ctx, t, ht->codec, -1, nb, ntiles, nseg, tcnt, toff, segpre, btot, off, 1, ht->slices,
                         ht->entries, cidx, ht->d_total, nullptr);
  if (rc) return rc;
#define PTC_PD(RT, VT, VALS, RK, GV, R16)                                                                              \
  do {                                                                                                                \
    const size_t lds = (size_t)PT_TILE * (sizeof(RT) + 4) + (size_t)nb * 8;                                           \
    static size_t lds_set = 0;                                                                                        \
    if (lds > 65536 && lds > lds_set) {                                                                               \
      PTC_HIP(ctx, hipFuncSetAttribute((const void*)k_pd_place<RT, VT>,                                              \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));                       \
      lds_set = lds;                                                                                                  \
    }                                                                                                                 \
    hipLaunchKernelGGL((k_pd_place<RT, VT>), dim3((unsigned)(8 * ((ntiles + 7) / 8))), dim3(PT_THREADS), lds, st,     \
                       t->hedges, t->d, N, ht->codec, ht->rest, (const VT*)(VALS), nb, ntiles,                      \
                       (const unsigned long long*)off, (const uint32_t*)tcnt, (const uint32_t*)toff,              \
                       (const uint32_t*)segpre,                                                                     \
                       (const uint32_t*)cidx, ht->perm, RK, GV, (uint32_t*)nullptr, R16);                           \
  } while (0)
  if (rest64) {
    if (f32) PTC_PD(unsigned long long, float, t->vals32, ht->rest_key64, ht->gval32, (uint16_t*)nullptr);
    else PTC_PD(unsigned long long, double, t->vals, ht->rest_key64, ht->gval, (uint16_t*)nullptr);
  } else {
    if (f32) PTC_PD(uint32_t, float, t->vals32, ht->rest_key32, ht->gval32, ht->rest_key16);
    else PTC_PD(uint32_t, double, t->vals, ht->rest_key32, ht->gval, ht->rest_key16);
  }
#undef PTC_PD
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

}  // namespace

extern "C" int ptc_hashtable_build(ptc_ctx* ctx, const ptc_dtensor* t, ptc_hashtable** out) {
  return ptc_hashtable_build_ex(ctx, t, PTC_GROUP_AUTO, out);
}

extern "C" int ptc_hashtable_build_ex(ptc_ctx* ctx, const ptc_dtensor* t, int grouping, ptc_hashtable** out) {
  if (!ctx || !t || !out) return PTC_ERR_ARG;
  *out = nullptr;
  if (grouping < PTC_GROUP_AUTO || grouping > PTC_GROUP_PARTITION) return set_err(ctx, PTC_ERR_ARG, "grouping %d", grouping);
  // slices hold 32-bit nonzero ids, and the row passes address B by 32-bit
  // positions (a contraction's output can be larger than an upload)
  if (t->N >= 0xFFFFFFFFull) return set_err(ctx, PTC_ERR_RANGE, "N=%llu exceeds 2^32-2", (unsigned long long)t->N);
  PTC_HIP(ctx, hipSetDevice(ctx->device));
  ptc_hashtable* ht = new ptc_hashtable();
  ht->ctx = ctx;
  ctx_acquire(ctx);
  ht->t = t;
  if (ctx->timing) {
    if (hipEventCreate(&ht->ev_start) != hipSuccess || hipEventCreate(&ht->ev_end) != hipSuccess) {
      ptc_hashtable_free(ht);
      return hip_check(ctx, hipGetLastError(), "event");
    }
    (void)hipEventRecord(ht->ev_start, ctx->stream);
  }
  int rc = make_codec(ctx, t, t->dim_tc, t->nk, &ht->codec);
  if (!rc) rc = make_codec(ctx, t, t->dim_tc + t->nk, t->d - t->nk, &ht->rest);
  if (rc) {
    ptc_hashtable_free(ht);
    return rc;
  }
  const uint64_t N = t->N;
  // A small key space is addressed directly (slot = key); otherwise distinct
  // keys <= min(N, key space) and the table is sized for 80% worst-case load.
  ht->direct = ht->codec.space <= std::max<uint64_t>(2 * N, 1u << 16) && ht->codec.space < (1ull << 31);
  const uint64_t maxkeys = std::min<uint64_t>(N, ht->codec.space);
  uint64_t cap = 64;
  while (cap < maxkeys + maxkeys / 4) cap <<= 1;
  if (ht->direct) cap = std::max<uint64_t>(ht->codec.space, 1);
  ht->cap = cap;
  ht->entries = dalloc_n<KtEntry>(ctx, cap);
  ht->perm = dalloc_n<uint32_t>(ctx, N);
  const bool f32 = t->vals == nullptr && t->vals32 != nullptr;  // an FP32 tensor
  if (f32)
    ht->gval32 = dalloc_n<float>(ctx, N);
  else
    ht->gval = dalloc_n<double>(ctx, N);
  const bool rest64 = ht->rest.space > 0xFFFFFFFFull;
  if (rest64)
    ht->rest_key64 = dalloc_n<unsigned long long>(ctx, N);
  else
    ht->rest_key32 = dalloc_n<uint32_t>(ctx, N);
  const uint64_t nblk = (N + RK_CHUNK - 1) / RK_CHUNK;
  // sort grouping (32-bit key spaces): requested, or the default for hashed
  // key spaces
  const bool sortable = N > 0 && ht->codec.space <= 0xFFFFFFFFull;
  int key_bits = 1;
  while (key_bits < 64 && (1ull << key_bits) < ht->codec.space) ++key_bits;
  if (grouping == PTC_GROUP_SORT && !sortable) {
    ptc_hashtable_free(ht);
    return set_err(ctx, PTC_ERR_UNSUPPORTED, "sort grouping needs a nonempty tensor and a key space within 32 bits");
  }
  // one counting pass: small direct key spaces, automatically or when
  // partitioning is asked for
  if (N > 0 && N <= 0xFFFFFFFFull && ht->direct && ht->codec.space <= PT_MAX_BUCKETS &&
      (grouping == PTC_GROUP_AUTO || grouping == PTC_GROUP_PARTITION)) {
    if (!ht->entries || !ht->perm || !(ht->gval || ht->gval32) || !(ht->rest_key32 || ht->rest_key64)) {
      ptc_hashtable_free(ht);
      return PTC_ERR_NOMEM;
    }
    rc = build_direct_partitioned(ctx, ht, t, f32, rest64);
    if (rc) {
      ptc_hashtable_free(ht);
      return rc;
    }
    if (ht->ev_end) (void)hipEventRecord(ht->ev_end, ctx->stream);
    *out = ht;
    return PTC_OK;
  }
  if (sortable && (grouping == PTC_GROUP_SORT || (grouping == PTC_GROUP_AUTO && key_bits <= SORT_AUTO_BITS))) {
    if (!ht->entries || !ht->perm || !(ht->gval || ht->gval32) || !(ht->rest_key32 || ht->rest_key64)) {
      ptc_hashtable_free(ht);
      return PTC_ERR_NOMEM;
    }
    rc = build_sorted(ctx, ht, t, f32, rest64);
    if (rc) {
      ptc_hashtable_free(ht);
      return rc;
    }
    if (ht->ev_end) (void)hipEventRecord(ht->ev_end, ctx->stream);
    *out = ht;
    return PTC_OK;
  }
  // partitioned hashing: requested, or the default for key spaces too wide to
  // sort (direct ones too: the index is then written by key)
  const bool partitionable = N > 0 && N <= PT_MAX_N;
  if (grouping == PTC_GROUP_PARTITION && !partitionable) {
    ptc_hashtable_free(ht);
    return set_err(ctx, PTC_ERR_UNSUPPORTED, "partitioned grouping needs 1 to %llu nonzeros", (unsigned long long)PT_MAX_N);
  }
  if (partitionable && (grouping == PTC_GROUP_PARTITION || grouping == PTC_GROUP_AUTO)) {
    if (!ht->entries || !ht->perm || !(ht->gval || ht->gval32) || !(ht->rest_key32 || ht->rest_key64)) {
      ptc_hashtable_free(ht);
      return PTC_ERR_NOMEM;
    }
    rc = build_partitioned(ctx, ht, t, f32, rest64);
    if (rc) {
      ptc_hashtable_free(ht);
      return rc;
    }
    if (ht->ev_end) (void)hipEventRecord(ht->ev_end, ctx->stream);
    *out = ht;
    return PTC_OK;
  }
  // Three ways to group:
  //  hashed    insert + wave-aggregated rank (counts in the entries), ids
  //            scattered, streaming layout gathered from the input-order
  //            payload; slice indices only when needed (a table used as A);
  //  counting  direct table with a small key space: stable counting sort,
  //            every nonzero placed at once (k_kt_place);
  //  direct    other direct tables: block-aggregated ranks, ids and slice
  //            indices scattered, streaming layout gathered.
  const bool hashed = !ht->direct;
  const bool counting = ht->direct && cap <= RD_MAX && cap * nblk <= (64ull << 20);
  if (!hashed) ht->slice_of = dalloc_n<uint32_t>(ctx, N);
  Workspace ws(ctx);
  uint32_t* cnt = hashed ? nullptr : ws.get<uint32_t>(cap);
  unsigned long long* packed = ws.get<unsigned long long>(cap + 1);
  uint32_t* slot_of = ws.get<uint32_t>(N);
  uint32_t* rank_of = ws.get<uint32_t>(N);
  const size_t pay_bytes = rest64 || !f32 ? 16 : 8;
  void* pay = hashed ? (void*)ws.get<unsigned char>(N * pay_bytes) : nullptr;
  if (!ht->entries || !ht->perm || (!hashed && !ht->slice_of) || !(ht->gval || ht->gval32) ||
      !(ht->rest_key32 || ht->rest_key64) || (!hashed && !cnt) || !packed || !slot_of || !rank_of || (hashed && N && !pay)) {
    ptc_hashtable_free(ht);
    return PTC_ERR_NOMEM;
  }
  hipStream_t st = ctx->stream;
  hipLaunchKernelGGL(k_kt_clear, dim3(grid_for(cap, 256)), dim3(256), 0, st, ht->entries, cnt, cap);
  unsigned long long* hpre = nullptr;
  if (N && hashed) {
    const unsigned g = grid_for(N, KT_THREADS, 16384);
#define PTC_INSERT(RT, VT, VALS)                                                                                  \
  hipLaunchKernelGGL((k_kt_insert_rank<RT, VT>), dim3(g), dim3(KT_THREADS), 0, st, t->hedges, t->d, N, ht->codec, \
                     ht->entries, cap, slot_of, rank_of, ht->rest, (const VT*)(VALS), (KtPay<RT, VT>*)pay)
    if (rest64) {
      if (f32) PTC_INSERT(unsigned long long, float, t->vals32);
      else PTC_INSERT(unsigned long long, double, t->vals);
    } else {
      if (f32) PTC_INSERT(uint32_t, float, t->vals32);
      else PTC_INSERT(uint32_t, double, t->vals);
    }
#undef PTC_INSERT
  } else if (N) {
    hipLaunchKernelGGL(k_kt_direct, dim3(grid_for(N, 256)), dim3(256), 0, st, t->hedges, t->d, N, ht->codec,
                       slot_of);
    if (counting) {
      uint32_t* hist = ws.get<uint32_t>(cap * nblk);
      hpre = ws.get<unsigned long long>(cap * nblk + 1);
      if (!hist || !hpre) {
        ptc_hashtable_free(ht);
        return PTC_ERR_NOMEM;
      }
      hipLaunchKernelGGL(k_kt_rank_direct, dim3((unsigned)nblk), dim3(RK_THREADS), 0, st, (const uint32_t*)slot_of, N,
                         (uint32_t)cap, nblk, hist, rank_of);
      if ((rc = scan_u32_to_u64(ctx, hist, hpre, cap * nblk, hpre + cap * nblk))) {
        ptc_hashtable_free(ht);
        return rc;
      }
      hipLaunchKernelGGL(k_kt_direct_cnt, dim3(grid_for(cap, 256)), dim3(256), 0, st, (const unsigned long long*)hpre,
                         (uint32_t)cap, nblk, cnt);
    } else {
      hipLaunchKernelGGL(k_kt_rank, dim3((unsigned)nblk), dim3(RK_THREADS), 0, st, (const uint32_t*)slot_of, N, cnt,
                         rank_of);
    }
  }
  hipLaunchKernelGGL(k_kt_pack, dim3(grid_for(cap, 256)), dim3(256), 0, st, (const uint32_t*)cnt,
                     (const KtEntry*)ht->entries, packed, cap);
  // no read-back: slices are sized by the bound min(N, cap) on distinct keys
  ht->d_total = dalloc_n<unsigned long long>(ctx, 1);
  ht->slices_cap = std::max<uint64_t>(1, std::min<uint64_t>(N, cap));
  ht->slices = dalloc_n<KtSlice>(ctx, ht->slices_cap);
  if (!ht->d_total || !ht->slices) {
    ptc_hashtable_free(ht);
    return PTC_ERR_NOMEM;
  }
  rc = scan_u64(ctx, packed, packed, cap, ht->d_total);
  if (rc) {
    ptc_hashtable_free(ht);
    return rc;
  }
  hipLaunchKernelGGL(k_kt_finalize, dim3(grid_for(cap, 256)), dim3(256), 0, st, ht->entries, (const uint32_t*)cnt,
                     packed, cap, ht->slices, ht->direct ? 1 : 0);
  const unsigned g = grid_for(N, 256, 16384);
  const unsigned long long* sinfo = (const unsigned long long*)packed;
  const uint32_t *so = slot_of, *ro = rank_of;
  if (N && hashed) {
    hipLaunchKernelGGL(k_kt_scatter_perm, dim3(g), dim3(256), 0, st, N, sinfo, so, ro, ht->perm);
#define PTC_GATHER(RT, VT, RK, GV) \
  hipLaunchKernelGGL((k_kt_gather_pay<RT, VT>), dim3(g), dim3(256), 0, st, N, (const uint32_t*)ht->perm, \
                     (const KtPay<RT, VT>*)pay, RK, GV)
    if (rest64) {
      if (f32) PTC_GATHER(unsigned long long, float, ht->rest_key64, ht->gval32);
      else PTC_GATHER(unsigned long long, double, ht->rest_key64, ht->gval);
    } else {
      if (f32) PTC_GATHER(uint32_t, float, ht->rest_key32, ht->gval32);
      else PTC_GATHER(uint32_t, double, ht->rest_key32, ht->gval);
    }
#undef PTC_GATHER
  } else if (N && counting) {
    const unsigned long long* hp = (const unsigned long long*)hpre;
#define PTC_PLACE(RT, VT, RK, GV, VALS)                                                                        \
  hipLaunchKernelGGL((k_kt_place<RT, VT, true>), dim3(g), dim3(256), 0, st, N, (const uint32_t*)t->hedges, t->d, \
                     ht->rest, (const VT*)(VALS), sinfo, hp, nblk, so, ro, ht->perm, ht->slice_of, RK, GV)
    if (rest64) {
      if (f32) PTC_PLACE(unsigned long long, float, ht->rest_key64, ht->gval32, t->vals32);
      else PTC_PLACE(unsigned long long, double, ht->rest_key64, ht->gval, t->vals);
    } else {
      if (f32) PTC_PLACE(uint32_t, float, ht->rest_key32, ht->gval32, t->vals32);
      else PTC_PLACE(uint32_t, double, ht->rest_key32, ht->gval, t->vals);
    }
#undef PTC_PLACE
  } else if (N) {
    hipLaunchKernelGGL(k_kt_scatter, dim3(g), dim3(256), 0, st, N, sinfo, so, ro, ht->perm, ht->slice_of);
#define PTC_GATHER(RT, VT, RK, GV, VALS)                                                                       \
  hipLaunchKernelGGL((k_kt_gather<RT, VT>), dim3(g), dim3(256), 0, st, t->hedges, (const VT*)(VALS), t->d, N, \
                     ht->rest, (const uint32_t*)ht->perm, RK, GV)
    if (rest64) {
      if (f32) PTC_GATHER(unsigned long long, float, ht->rest_key64, ht->gval32, t->vals32);
      else PTC_GATHER(unsigned long long, double, ht->rest_key64, ht->gval, t->vals);
    } else {
      if (f32) PTC_GATHER(uint32_t, float, ht->rest_key32, ht->gval32, t->vals32);
      else PTC_GATHER(uint32_t, double, ht->rest_key32, ht->gval, t->vals);
    }
#undef PTC_GATHER
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ptc_hashtable_free(ht);
    return hip_check(ctx, e, "hashtable build");
  }
  if (ht->ev_end) (void)hipEventRecord(ht->ev_end, st);
  *out = ht;
  return PTC_OK;
}

namespace ptc {
// Key tables of n tensors at once (batch plans).  FP64 tensors with 1 to
// PT_MAX_N nonzeros are grouped together -- one counting pass for direct key
// spaces up to PT_MAX_BUCKETS keys, partitioned hashing for the rest -- with
// one launch per step; any other tensor is built on its own.  The tables
// hold the same slices as ptc_hashtable_build's.
int hashtable_build_many(ptc_ctx* ctx, int n, const ptc_dtensor* const* ts, ptc_hashtable** out) {
  for (int i = 0; i < n; ++i) out[i] = nullptr;
  PTC_HIP(ctx, hipSetDevice(ctx->device));
  std::vector<MtTable> T;
  uint64_t tiles = 0, cnts = 0, items = 0, buckets = 0, keyspace = 0;
  size_t lds = 0;
  int rc = PTC_OK;
  auto fail = [&](int code) {
    for (int i = 0; i < n; ++i) {
      ptc_hashtable_free(out[i]);
      out[i] = nullptr;
    }
    return code;
  };
  for (int i = 0; i < n; ++i) {
    const ptc_dtensor* t = ts[i];
    if (!t->vals || t->N == 0 || t->N > PT_MAX_N) {
      if ((rc = ptc_hashtable_build(ctx, t, &out[i]))) return fail(rc);
      continue;
    }
    ptc_hashtable* ht = new ptc_hashtable();
    out[i] = ht;
    ht->ctx = ctx;
    ctx_acquire(ctx);
    ht->t = t;
    if ((rc = make_codec(ctx, t, t->dim_tc, t->nk, &ht->codec)) ||
        (rc = make_codec(ctx, t, t->dim_tc + t->nk, t->d - t->nk, &ht->rest)))
      return fail(rc);
    const uint64_t N = t->N;
    ht->direct = ht->codec.space <= std::max<uint64_t>(2 * N, 1u << 16) && ht->codec.space < (1ull << 31);
    const uint64_t maxkeys = std::min<uint64_t>(N, ht->codec.space);
    uint64_t cap = 64;
    while (cap < maxkeys + maxkeys / 4) cap <<= 1;
    if (ht->direct) cap = std::max<uint64_t>(ht->codec.space, 1);
    ht->cap = cap;
    const bool wide = ht->rest.space > 0xFFFFFFFFull;
    const bool counting = ht->direct && ht->codec.space <= PT_MAX_BUCKETS;
    ht->entries = dalloc_n<KtEntry>(ctx, cap);
    ht->perm = dalloc_n<uint32_t>(ctx, N);
    ht->gval = dalloc_n<double>(ctx, N);
    if (wide) ht->rest_key64 = dalloc_n<unsigned long long>(ctx, N);
    else ht->rest_key32 = dalloc_n<uint32_t>(ctx, N);
    ht->d_total = dalloc_n<unsigned long long>(ctx, 1);
    ht->slices_cap = std::max<uint64_t>(1, std::min<uint64_t>(N, cap));
    ht->slices = dalloc_n<KtSlice>(ctx, ht->slices_cap);
    // (the slice index of every position is built on first use) and the
    // 16-bit copy of non-key tuples of at most 16 bits (the column keys that
    // dense contractions stream) is written while grouping
    if (!wide && ht->rest.space <= 65536 && !(ht->rest_key16 = dalloc_n<uint16_t>(ctx, std::max<uint64_t>(N, 1))))
      return fail(PTC_ERR_NOMEM);
    if (!ht->entries || !ht->perm || !ht->gval || !(ht->rest_key32 || ht->rest_key64) || !ht->d_total ||
        !ht->slices)
      return fail(PTC_ERR_NOMEM);
    MtTable m{};
    m.hedges = t->hedges;
    m.vals = t->vals;
    m.perm = ht->perm;
    m.rest32 = ht->rest_key32;
    m.rest64 = ht->rest_key64;
    m.gval = ht->gval;
    m.entries = ht->entries;
    m.slices = ht->slices;
    m.slice_of = nullptr;
    m.rest16 = ht->rest_key16;
    m.total = ht->d_total;
    m.codec = ht->codec;
    m.rest = ht->rest;
    m.N = N;
    m.cap = cap;
    m.d = t->d;
    m.counting = counting ? 1 : 0;
    m.direct = ht->direct ? 1 : 0;
    m.wide = wide ? 1 : 0;
    if (counting) {
      m.nb = (uint32_t)ht->codec.space;
      m.bbits = -1;
    } else {
      m.nb = 1;
      m.bbits = 0;
      while (m.nb < PT_MAX_BUCKETS && (uint64_t)m.nb * PT_BUCKET_ITEMS < N) {
        m.nb <<= 1;
        ++m.bbits;
      }
    }
    m.region = !counting && !ht->direct && cap / 8 >= m.nb ? 1 : 0;
    m.ntiles = (N + PT_TILE - 1) / PT_TILE;
    m.tile_base = tiles;
    m.cnt_base = cnts;
    m.item_base = items;
    m.bucket_base = buckets;
    m.cidx = reinterpret_cast<uint32_t*>(keyspace);  // an offset until the shared array exists
    tiles += m.ntiles;
    cnts += (uint64_t)m.nb * m.ntiles;
    items += N;
    if (counting) keyspace += m.nb;
    else buckets += m.nb;
    const size_t rs = wide ? 8 : 4;
    lds = std::max(lds, counting ? (size_t)PT_TILE * (rs + 4) + (size_t)m.nb * 8
                                 : (size_t)PT_TILE * (8 + rs + 2) + (size_t)m.nb * 8);
    T.push_back(m);
  }
  if (T.empty()) return PTC_OK;
  if (items > 0xFFFFFFFFull)
    return fail(set_err(ctx, PTC_ERR_RANGE, "%llu nonzeros in one table set", (unsigned long long)items));
  hipStream_t st = ctx->stream;
  Workspace ws(ctx);
  uint32_t* hist = ws.get<uint32_t>(cnts);
  unsigned long long* off = ws.get<unsigned long long>(cnts + 1);
  unsigned long long* pkey = ws.get<unsigned long long>(items);
  uint32_t* pid = ws.get<uint32_t>(items);
  void* ppay = ws.get<unsigned char>(items * 16);
  unsigned long long* gkey = ws.get<unsigned long long>(items);
  uint32_t* gcnt = ws.get<uint32_t>(items);
  KtSlice* tmps = ws.get<KtSlice>(items);
  uint32_t* bkeys = ws.get<uint32_t>(buckets + 1);
  unsigned long long* sbase = ws.get<unsigned long long>(buckets + 1);
  uint32_t* ovf = ws.get<uint32_t>(items + 1);
  uint32_t* cidx = ws.get<uint32_t>(keyspace + 1);
  MtTable* dT = ws.get<MtTable>(T.size());
  if (!hist || !off || !pkey || !pid || !ppay || !gkey || !gcnt || !tmps || !bkeys || !sbase || !ovf || !cidx || !dT)
    return fail(PTC_ERR_NOMEM);
  for (MtTable& m : T) m.cidx = cidx + reinterpret_cast<uintptr_t>(m.cidx);
  // a synchronous copy: the host vector is gone when the kernels run
  PTC_HIP(ctx, hipMemcpy(dT, T.data(), sizeof(MtTable) * T.size(), hipMemcpyHostToDevice));
  const int nt = (int)T.size();
  const MtTableG* dG = reinterpret_cast<const MtTableG*>(dT);
  if ((rc = fill_zero(ctx, ovf, 4))) return fail(rc);
  hipLaunchKernelGGL(k_mt_hist, dim3((unsigned)tiles), dim3(PT_THREADS), 0, st, dG, nt, hist);
  if ((rc = scan_u32_to_u64(ctx, hist, off, cnts, off + cnts))) return fail(rc);
  hipLaunchKernelGGL(k_mt_clear, dim3(16, (unsigned)nt), dim3(256), 0, st, dG, nt);
  hipLaunchKernelGGL(k_mt_pd_slices, dim3((unsigned)nt), dim3(1024), 0, st, dG, nt, (const unsigned long long*)off);
  static size_t lds_set = 0;
  if (lds > 65536 && lds > lds_set) {
    PTC_HIP(ctx, hipFuncSetAttribute((const void*)k_mt_scatter, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    lds_set = lds;
  }
  hipLaunchKernelGGL(k_mt_scatter, dim3((unsigned)(8 * ((tiles + 7) / 8))), dim3(PT_THREADS), lds, st, dG, nt, tiles,
                     (const uint32_t*)hist, (const unsigned long long*)off, pkey, pid, ppay);
  if (buckets) {
    hipLaunchKernelGGL(k_mt_group, dim3((unsigned)buckets), dim3(PT_GROUP_THREADS), 0, st, dG, nt,
                       (const unsigned long long*)off, (const unsigned long long*)pkey, (const uint32_t*)pid,
                       (const void*)ppay, gkey, gcnt, tmps, bkeys, ovf, ovf + 1);
    if ((rc = scan_u32_to_u64(ctx, bkeys, sbase, buckets, sbase + buckets))) return fail(rc);
    hipLaunchKernelGGL(k_mt_slices, dim3((unsigned)buckets), dim3(256), 0, st, dG, nt, (const unsigned long long*)off,
                       (const uint32_t*)bkeys, (const unsigned long long*)sbase, (const KtSlice*)tmps);
    hipLaunchKernelGGL(k_mt_overflow, dim3(64), dim3(256), 0, st, dG, nt, (const uint32_t*)ovf,
                       (const uint32_t*)(ovf + 1), (const KtSlice*)tmps);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(hip_check(ctx, e, "batched key tables"));
  return PTC_OK;
}

__global__ void k_narrow16(uint64_t N, const uint32_t* __restrict__ in, uint16_t* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (uint16_t)in[i];
}

int hashtable_col16(ptc_ctx* ctx, const ptc_hashtable* hc, const uint16_t** out) {
  ptc_hashtable* h = const_cast<ptc_hashtable*>(hc);
  if (!h->rest_key16) {
    if (!h->rest_key32 || h->rest.space > 65536)
      return set_err(ctx, PTC_ERR_INTERNAL, "16-bit keys of a %llu-key space", (unsigned long long)h->rest.space);
    const uint64_t N = h->t->N;
    uint16_t* k16 = dalloc_n<uint16_t>(ctx, std::max<uint64_t>(N, 1));
    if (!k16) return PTC_ERR_NOMEM;
    if (N) {
      hipLaunchKernelGGL(k_narrow16, dim3(grid_for(N, 256, 16384)), dim3(256), 0, ctx->stream, N,
                         (const uint32_t*)h->rest_key32, k16);
      PTC_HIP(ctx, hipGetLastError());
    }
    // a block of the allocation-free scratch lives for one call: not kept
    if (ctx->arena) {
      *out = k16;
      return PTC_OK;
    }
    h->rest_key16 = k16;
  }
  *out = h->rest_key16;
  return PTC_OK;
}

// Slice index of every grouped position, built on first use (hashed tables
// skip it at build time: only a table that serves as A needs it).
int hashtable_slice_of(ptc_ctx* ctx, const ptc_hashtable* hc, const uint32_t** out) {
  ptc_hashtable* h = const_cast<ptc_hashtable*>(hc);
  if (!h->slice_of) {
    const uint64_t N = h->t->N;
    uint32_t* so = dalloc_n<uint32_t>(ctx, std::max<uint64_t>(N, 1));
    if (!so) return PTC_ERR_NOMEM;
    if (N) {
      hipLaunchKernelGGL(k_kt_slice_of, dim3(grid_for(N, 256, 16384)), dim3(256), 0, ctx->stream, N,
                         (const KtSlice*)h->slices, (const unsigned long long*)h->d_total, so);
      PTC_HIP(ctx, hipGetLastError());
    }
    if (ctx->arena) {  // scratch block of one allocation-free call: not kept
      *out = so;
      return PTC_OK;
    }
    h->slice_of = so;
  }
  *out = h->slice_of;
  return PTC_OK;
}

int hashtable_num_keys(ptc_ctx* ctx, const ptc_hashtable* hc, uint64_t* out) {
  ptc_hashtable* h = const_cast<ptc_hashtable*>(hc);
  if (!h->keys_known) {
    unsigned long long total = 0;
    int rc = read_back(ctx, &total, h->d_total, sizeof(total));
    if (rc) return rc;
    if ((total & 0xFFFFFFFFull) != h->t->N)
      return set_err(ctx, PTC_ERR_INTERNAL, "key table grouped %llu of %llu nonzeros",
                     (unsigned long long)(total & 0xFFFFFFFFull), (unsigned long long)h->t->N);
    h->num_keys = total >> 32;
    h->keys_known = true;
  }
  *out = h->num_keys;
  return PTC_OK;
}

double hashtable_ms_build(const ptc_hashtable* hc) {
  ptc_hashtable* h = const_cast<ptc_hashtable*>(hc);
  if (h->ms_build < 0 && h->ev_start && h->ev_end) {
    float ms = 0.f;
    if (hipEventSynchronize(h->ev_end) == hipSuccess && hipEventElapsedTime(&ms, h->ev_start, h->ev_end) == hipSuccess)
      h->ms_build = ms;
    else
      (void)hipGetLastError();
  }
  return h->ms_build < 0 ? 0.0 : h->ms_build;
}

// ------------------------------------------------- tables from output rows
//
// A contraction writes its output grouped by row: row i (A's free key
// rows[i].key) occupies [off[i], off[i] + cnt[i]).  When the next
// contraction keys the output on exactly those modes, the row grouping is
// the reorder_tensor layout (fksCuckooTC.hpp:31-77) already: the slices are
// the nonempty rows, the grouped order is the output order (perm is the
// identity) and the values are the output's own.  Only the keys are
// re-encoded into the output's codec, the non-key tuples linearised, and the
// probe index written.

__global__ void k_row_flags(uint64_t nrows, const uint32_t* __restrict__ cnt, uint32_t* __restrict__ flag) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nrows; i += (uint64_t)gridDim.x * blockDim.x)
    flag[i] = cnt[i] ? 1u : 0u;
}

// Row i -> slice sx[i] (nonempty rows only).  Key position q of the table's
// codec is output mode keyc.mode[q], which is position keyc.mode[q] of the
// row codec.  len holds the end of the range, as k_rs_index expects.
__global__ void k_rows_to_slices(uint64_t nrows, const KtSlice* __restrict__ rows, const uint32_t* __restrict__ cnt,
                                 const unsigned long long* __restrict__ off, const unsigned long long* __restrict__ sx,
                                 KeyCodec rowc, KeyCodec keyc, KtSlice* __restrict__ slices) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nrows; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t n = cnt[i];
    if (!n) continue;
    const unsigned long long k = rows[i].key;
    unsigned long long nk = 0;
    for (uint32_t q = 0; q < keyc.n; ++q) {
      const uint32_t j = keyc.mode[q];
      const unsigned long long x = rowc.n == 1 ? k : (k / rowc.stride[j]) % rowc.radix[j];
      nk += x * keyc.stride[q];
    }
    KtSlice sl;
    sl.key = nk;
    sl.off = (uint32_t)off[i];
    sl.len = (uint32_t)(off[i] + n);
    slices[sx[i]] = sl;
  }
}

template <class RestT>
__global__ void k_rest_keys(const uint32_t* __restrict__ h, uint32_t d, uint64_t N, KeyCodec rest,
                            RestT* __restrict__ out) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N; p += (uint64_t)gridDim.x * blockDim.x) {
    unsigned long long k = 0;
    for (uint32_t j = 0; j < rest.n; ++j) k += (unsigned long long)h[p * d + rest.mode[j]] * rest.stride[j];
    out[p] = (RestT)k;
  }
}

__global__ void k_iota(uint64_t N, uint32_t* __restrict__ out) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N; p += (uint64_t)gridDim.x * blockDim.x)
    out[p] = (uint32_t)p;
}

int hashtable_from_rows(ptc_ctx* ctx, const ptc_dtensor* O, const KeyCodec& rowc, const KtSlice* rows,
                        uint64_t nrows, const uint32_t* cnt, const unsigned long long* off, ptc_hashtable** out) {
  *out = nullptr;
  const uint64_t N = O->N;
  if (N == 0 || N > 0xFFFFFFFFull || nrows == 0) return set_err(ctx, PTC_ERR_UNSUPPORTED, "output size for a row table");
  for (uint32_t q = 0; q < O->nk; ++q)
    if (O->dim_tc[q] >= rowc.n) return set_err(ctx, PTC_ERR_ARG, "key mode %u is not a row mode", O->dim_tc[q]);
  ptc_hashtable* ht = new ptc_hashtable();
  ht->ctx = ctx;
  ctx_acquire(ctx);
  ht->t = O;
  ht->shares_vals = true;
  hipStream_t st = ctx->stream;
  if (ctx->timing) {
    if (hipEventCreate(&ht->ev_start) != hipSuccess || hipEventCreate(&ht->ev_end) != hipSuccess) {
      ptc_hashtable_free(ht);
      return hip_check(ctx, hipGetLastError(), "event");
    }
    (void)hipEventRecord(ht->ev_start, st);
  }
  int rc = make_codec(ctx, O, O->dim_tc, O->nk, &ht->codec);
  if (!rc) rc = make_codec(ctx, O, O->dim_tc + O->nk, O->d - O->nk, &ht->rest);
  if (rc) {
    ptc_hashtable_free(ht);
    return rc;
  }
  // the index: the same direct/hashed rule as a build, sized for the rows
  ht->direct = ht->codec.space <= std::max<uint64_t>(2 * N, 1u << 16) && ht->codec.space < (1ull << 31);
  const uint64_t maxkeys = std::min<uint64_t>(nrows, ht->codec.space);
  uint64_t cap = 64;
  while (cap < maxkeys + maxkeys / 4) cap <<= 1;
  if (ht->direct) cap = std::max<uint64_t>(ht->codec.space, 1);
  ht->cap = cap;
  ht->entries = dalloc_n<KtEntry>(ctx, cap);
  ht->perm = dalloc_n<uint32_t>(ctx, N);
  ht->slices_cap = nrows;
  ht->slices = dalloc_n<KtSlice>(ctx, nrows);
  ht->d_total = dalloc_n<unsigned long long>(ctx, 1);
  const bool rest64 = ht->rest.space > 0xFFFFFFFFull;
  if (rest64) ht->rest_key64 = dalloc_n<unsigned long long>(ctx, N);
  else ht->rest_key32 = dalloc_n<uint32_t>(ctx, N);
  ht->gval = O->vals;
  ht->gval32 = O->vals32;
  Workspace ws(ctx);
  uint32_t* flag = ws.get<uint32_t>(nrows);
  unsigned long long* sx = ws.get<unsigned long long>(nrows + 1);
  if (!ht->entries || !ht->perm || !ht->slices || !ht->d_total || !(ht->rest_key32 || ht->rest_key64) || !flag || !sx) {
    ptc_hashtable_free(ht);
    return PTC_ERR_NOMEM;
  }
  const unsigned gr = grid_for(nrows, 256, 16384), gn = grid_for(N, 256, 16384);
  hipLaunchKernelGGL(k_row_flags, dim3(gr), dim3(256), 0, st, nrows, cnt, flag);
  if ((rc = scan_u32_to_u64(ctx, flag, sx, nrows, sx + nrows))) {
    ptc_hashtable_free(ht);
    return rc;
  }
  hipLaunchKernelGGL(k_rows_to_slices, dim3(gr), dim3(256), 0, st, nrows, rows, cnt, off,
                     (const unsigned long long*)sx, rowc, ht->codec, ht->slices);
  hipLaunchKernelGGL(k_kt_clear, dim3(grid_for(cap, 256)), dim3(256), 0, st, ht->entries, (uint32_t*)nullptr, cap);
  hipLaunchKernelGGL(k_rs_index, dim3(gr), dim3(256), 0, st, ht->slices, (const unsigned long long*)(sx + nrows),
                     ht->entries, cap, ht->direct ? 1 : 0, ht->d_total, N);
  hipLaunchKernelGGL(k_iota, dim3(gn), dim3(256), 0, st, N, ht->perm);
  if (rest64)
    hipLaunchKernelGGL((k_rest_keys<unsigned long long>), dim3(gn), dim3(256), 0, st, (const uint32_t*)O->hedges, O->d, N,
                       ht->rest, ht->rest_key64);
  else
    hipLaunchKernelGGL((k_rest_keys<uint32_t>), dim3(gn), dim3(256), 0, st, (const uint32_t*)O->hedges, O->d, N,
                       ht->rest, ht->rest_key32);
  if (ht->ev_end) (void)hipEventRecord(ht->ev_end, st);
  if (hipGetLastError() != hipSuccess) {
    ptc_hashtable_free(ht);
    return set_err(ctx, PTC_ERR_HIP, "row table kernels");
  }
  *out = ht;
  return PTC_OK;
}
}  // namespace ptc

extern "C" int ptc_hashtable_info(const ptc_hashtable* h, uint64_t* num_keys, uint64_t* num_items) {
  if (!h) return PTC_ERR_ARG;
  if (num_keys) {
    int rc = hashtable_num_keys(h->ctx, h, num_keys);
    if (rc) return rc;
  }
  if (num_items) *num_items = h->t ? h->t->N : 0;
  return PTC_OK;
}

extern "C" int ptc_hashtable_probe(ptc_ctx* ctx, const ptc_hashtable* h, uint64_t n,
                                   const uint32_t* keys, uint64_t* slice_len, uint64_t* slice_off,
                                   uint32_t* slice_ids, uint64_t cap) {
  if (!ctx || !h || (n && (!keys || !slice_len))) return PTC_ERR_ARG;
  if (n == 0) {
    if (slice_off) slice_off[0] = 0;
    return PTC_OK;
  }
  PTC_HIP(ctx, hipSetDevice(ctx->device));
  const uint32_t nk = h->codec.n;
  Workspace ws(ctx);
  uint32_t* dq = ws.get<uint32_t>(n * (nk ? nk : 1));
  unsigned long long* dres = ws.get<unsigned long long>(n);
  if (!dq || !dres) return PTC_ERR_NOMEM;
  if (nk) PTC_HIP(ctx, hipMemcpyAsync(dq, keys, sizeof(uint32_t) * n * nk, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_kt_probe, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, dq, n, h->codec,
                     h->entries, h->cap, h->direct ? 1 : 0, dres);
  PTC_HIP(ctx, hipGetLastError());
  std::vector<unsigned long long> res(n);
  PTC_HIP(ctx, hipMemcpyAsync(res.data(), dres, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost, ctx->stream));
  PTC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  uint64_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) {
    slice_len[i] = res[i] & 0xFFFFFFFFull;
    if (slice_off) slice_off[i] = acc;
    acc += slice_len[i];
  }
  if (slice_off) slice_off[n] = acc;
  if (slice_ids) {
    if (acc > cap) return set_err(ctx, PTC_ERR_RANGE, "slice_ids capacity %llu < %llu", (unsigned long long)cap, (unsigned long long)acc);
    uint64_t w = 0;
    for (uint64_t i = 0; i < n; ++i) {
      uint64_t len = res[i] & 0xFFFFFFFFull, off = res[i] >> 32;
      if (len)
        PTC_HIP(ctx, hipMemcpyAsync(slice_ids + w, h->perm + off, sizeof(uint32_t) * len, hipMemcpyDeviceToHost, ctx->stream));
      w += len;
    }
    PTC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return PTC_OK;
}

extern "C" void ptc_hashtable_free(ptc_hashtable* h) {
  if (!h) return;
  ptc_ctx* ctx = h->ctx;
  // after ptc_destroy the context released everything: no HIP call here
  if (ctx_open(ctx)) {
    dfree(ctx, h->entries);
    dfree(ctx, h->slices);
    dfree(ctx, h->perm);
    dfree(ctx, h->slice_of);
    dfree(ctx, h->rest_key32);
    dfree(ctx, h->rest_key64);
    dfree(ctx, h->rest_key16);
    if (!h->shares_vals) {
      dfree(ctx, h->gval);
      dfree(ctx, h->gval32);
    }
    dfree(ctx, h->d_total);
    if (h->ev_start) (void)hipEventDestroy(h->ev_start);
    if (h->ev_end) (void)hipEventDestroy(h->ev_end);
  }
  delete h;
  ctx_release(ctx);
}
