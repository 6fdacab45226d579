// The contraction O = A x B over the contracted modes (the role of
// tc_fksCuckoo, fksCuckooTC.hpp:80-371).
//
// Data on the device (all in HBM, grouped by the two key tables):
//   rows      HT_A slices: one per distinct A free tuple (an output row),
//             with the offset/length of its A positions in grouped order.
//   positions A nonzeros in row-grouped order: HT_A->rest_key (the
//             contracted tuple) and HT_A->gval (the value).
//   slices    HT_B: every B nonzero in key-grouped order with its free tuple
//             linearised (the column key) and its value.
//
// Kernels:
//   k_probe      one thread per A position: re-encode the contracted tuple
//                in B's codec, probe B's table -> (slice offset, length);
//                per-row hit count and longest slice by a wave-segmented
//                reduction (the reference's upper-bound preprocessing,
//                fksCuckooTC.hpp:117-154, plus the learned features).
//   scan         exclusive scan of slice lengths: every multiply-add of the
//                contraction gets a global index ("flop" index); a row's
//                flops are a contiguous range.
//   k_rows       per row: flops, work chunks, hash capacities.
//   k_item_p0    start position of each chunk of a heavy row.
//   k_pass       one workgroup per work item (a row, or a chunk of a heavy
//                row): stages its A positions in LDS, expands the flat flop
//                range (coalesced reads of B's column keys and values), and
//                accumulates per column in LDS -- a direct-mapped array when
//                the column space fits (a perfect hash), otherwise an
//                open-addressing LDS hash table -- or, for rows too large for
//                LDS, in an HBM hash table.  Heavy rows split over several
//                workgroups write partial slabs to HBM; a second launch of
//                the same kernel (MERGE) reduces a row's slabs in chunk
//                order and finishes the row.  SYMBOLIC counts distinct columns
//                per row; NUMERIC accumulates values and writes the row.
//   k_emit       learned/upper-bound methods: copy rows from their
//                capacity-sized staging area to the prefix-scanned output.
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "ptc_contract.h"
#include "ptc_internal.h"

namespace ptc {

// Groups of 64 flops a wave expands per step (2 loads each in flight).
// Measured on MI355X (numeric row pass): 2 -> 5.85 ms, 4 -> 5.37, 6 -> 5.19,
// 8 -> 5.07 ms on the two-tensor contraction, whose 64 KB accumulators
// leave 4 waves per SIMD, so each wave needs many loads in flight; the
// nips-shape pass (8 waves per SIMD) is flat from 4 to 8 (0.325 ms).  At 8
// the dense numeric pass takes 58 VGPRs, still 8 waves per SIMD; 12 and 16
// take 81 and 95, gain 1 % on the two-tensor pass and lose 22 % on nips.
// FP32 values accumulated in FP64 are the exception: 5.23 ms at 4, 7.00 ms
// at 8 on the two-tensor pass, so that value type keeps 4.
#ifndef PTC_PASS_UNROLL
#define PTC_PASS_UNROLL 8
#endif
constexpr int pass_unroll(int val) { return val == 3 ? 4 : PTC_PASS_UNROLL; }
// A positions staged in LDS at a time: 2 per thread of the launch
constexpr uint32_t NO_HEAVY = 0xFFFFFFFFu;

__device__ __forceinline__ unsigned long long mix64d(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

// ----------------------------------------------------------------- probe

template <class RestT>
__global__ void __launch_bounds__(256) k_probe(ProbeParams P) {
  const RestT* __restrict__ akey = (const RestT*)P.akey;
  if (P.zero4 && blockIdx.x == 0 && threadIdx.x < 4) P.zero4[threadIdx.x] = 0;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P.npos; p += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t len = 0, off = 0;
    {
      const unsigned long long ka = (unsigned long long)akey[p];
      // A's contracted tuple in B's key codec (the same key when both
      // codecs agree, as in a self-contraction)
      unsigned long long kb = ka;
      bool ok = true;
      if (!P.same_codec) {
        kb = 0;
        for (uint32_t j = 0; j < P.acodec.n; ++j) {
          const uint32_t x = (uint32_t)((ka / P.acodec.stride[j]) % P.acodec.radix[j]);
          if (x >= P.bcodec.radix[j]) ok = false;
          kb += (unsigned long long)x * P.bcodec.stride[j];
        }
      }
      if (ok && P.direct) {
        if (kb < P.cap) {
          const KtEntry en = P.tab[kb];
          len = en.len;
          off = en.off;
        }
      } else if (ok) {
        // same probe order as the insertion: the home line of mix64(k), then
        // onward; half a line (four entries, 64 bytes) is loaded at a time
        // -- the key is nearly always in the first half
        uint64_t i = kt_home(mix64d(kb), P.cap);
        bool done = false;
        while (!done) {
          KtEntry en[4];
#pragma unroll
          for (int s = 0; s < 4; ++s) en[s] = P.tab[(i + s) & (P.cap - 1)];
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            if (done) continue;
            if (en[s].key == kb) {
              len = en[s].len;
              off = en[s].off;
              done = true;
            } else if (en[s].key == PTC_EMPTY_KEY) {
              done = true;
            }
          }
          i = (i + 4) & (P.cap - 1);
        }
      }
      P.pos_len[p] = len;
      P.pos_off[p] = off;
    }
  }
}


// ------------------------------------------------------------------ rows

__global__ void k_rows(RowParams P) {
  const uint64_t live = P.live ? (*P.live >> 32) : P.nrows;
  unsigned long long chunk = P.chunk;
  if (P.total) chunk = max(chunk, (*P.total + P.max_items - 1) / P.max_items);
  chunk = min(chunk, MAX_ITEM_FLOPS);  // the growth rule passes 2^32 above ~1.3e13 flops
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < P.nrows; r += (uint64_t)gridDim.x * blockDim.x) {
    if (r >= live) {  // allocated beyond the table's rows: no work, zero sums
      P.row_nchunk[r] = 0;
      P.row_flops[r] = 0;
      P.packed[r] = 0;
      P.slabs[r] = 0;
      if (P.row_hits) {  // features of rows past the table (the scan wrote every row with positions)
        P.row_hits[r] = 0;
        P.row_maxlen[r] = 0;
      }
      continue;
    }
    KtSlice s = P.rows[r];
    unsigned long long f = P.apre[(uint64_t)s.off + s.len] - P.apre[s.off];
    uint32_t nch = f ? (uint32_t)((f + chunk - 1) / chunk) : 0u;
    P.row_nchunk[r] = nch;
    P.row_flops[r] = f;
    P.packed[r] = ((unsigned long long)(nch > 1) << 32) | nch;
    P.slabs[r] = nch > 1 ? nch : 0ull;
  }
}

// Rows, their two scans, their split-row list, (upper bound) their
// capacities and (zero non-NULL) the clear of nzero words in one workgroup, for at most ROWS_SMALL_THREADS * K rows: what
// k_rows, the in-place scans of packed and slabs (totals at [nrows]),
// k_items (but the items) and k_caps write, in one launch instead of four.  Thread t takes
// the K consecutive rows from t * K, all their loads in flight together.
constexpr int ROWS_SMALL_THREADS = 1024;
template <int K>
__global__ void __launch_bounds__(ROWS_SMALL_THREADS) k_rows_small(RowParams P, ItemParams Q, uint32_t* cap,
                                                                   unsigned long long ncols, uint32_t* zero,
                                                                   uint32_t nzero) {
  for (uint32_t i = threadIdx.x; i < nzero; i += ROWS_SMALL_THREADS) zero[i] = 0;
  __shared__ unsigned long long s_w[2][ROWS_SMALL_THREADS / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint64_t r0 = (uint64_t)threadIdx.x * K;
  // every load unconditional (a load under a condition is waited for
  // before the branch joins): the slices of rows clamped to the last one,
  // read before the row count arrives (rows < nrows are allocated); rows
  // past the count get the empty slice, whose flops apre[0] - apre[0] are 0
  KtSlice sl[K];
#pragma unroll
  for (int k = 0; k < K; ++k) sl[k] = P.rows[min<uint64_t>(r0 + k, P.nrows - 1)];
  const uint64_t live = P.live ? (*P.live >> 32) : P.nrows;
  unsigned long long chunk = P.chunk;
  if (P.total) chunk = max(chunk, (*P.total + P.max_items - 1) / P.max_items);
  chunk = min(chunk, MAX_ITEM_FLOPS);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint64_t r = r0 + k;
    if (r >= live || r >= P.nrows) {
      sl[k].off = 0;
      sl[k].len = 0;
    }
  }
  unsigned long long f[K];
#pragma unroll
  for (int k = 0; k < K; ++k) f[k] = P.apre[(uint64_t)sl[k].off + sl[k].len] - P.apre[sl[k].off];
  uint32_t nch[K];
  unsigned long long pk_sum = 0, sb_sum = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint64_t r = r0 + k;
    nch[k] = f[k] ? (uint32_t)((f[k] + chunk - 1) / chunk) : 0u;
    if (r < P.nrows) {
      P.row_nchunk[r] = nch[k];
      P.row_flops[r] = f[k];
      if (P.row_hits && r >= live) {
        P.row_hits[r] = 0;
        P.row_maxlen[r] = 0;
      }
      if (cap) cap[r] = (uint32_t)min(min(f[k], ncols), 0xFFFFFFFFull);
    }
    pk_sum += ((unsigned long long)(nch[k] > 1) << 32) | nch[k];
    sb_sum += nch[k] > 1 ? nch[k] : 0ull;
  }
  // exclusive scans of (heavy << 32 | chunks) and of the split rows' slabs
  unsigned long long x = pk_sum, y = sb_sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned long long a = __shfl_up(x, o, 64), b = __shfl_up(y, o, 64);
    if (lane >= o) {
      x += a;
      y += b;
    }
  }
  if (lane == 63) {
    s_w[0][wid] = x;
    s_w[1][wid] = y;
  }
  __syncthreads();
  unsigned long long pk = x - pk_sum, sb = y - sb_sum;
  for (int w = 0; w < wid; ++w) {
    pk += s_w[0][w];
    sb += s_w[1][w];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint64_t r = r0 + k;
    if (r < P.nrows) {
      P.packed[r] = pk;
      P.slabs[r] = sb;
      // the split-row list (k_items); the items themselves are written by
      // k_item_starts, which finds each item's row in the scanned counts
      const uint32_t hv = (uint32_t)(pk >> 32);
      Q.row_heavy[r] = nch[k] > 1 ? hv : NO_HEAVY;
      if (nch[k] > 1) Q.heavy_rows[hv] = (uint32_t)r;
    } else if (r == P.nrows) {  // the totals after the last row
      P.packed[r] = pk;
      P.slabs[r] = sb;
    }
    pk += ((unsigned long long)(nch[k] > 1) << 32) | nch[k];
    sb += nch[k] > 1 ? nch[k] : 0ull;
  }
}
int launch_rows_small(ptc_ctx* ctx, const RowParams& P, const ItemParams& Q, uint32_t* cap, unsigned long long ncols,
                      uint32_t* zero, uint32_t nzero) {
  // the kernel clamps its row loads to row nrows - 1
  if (P.nrows == 0) return set_err(ctx, PTC_ERR_INTERNAL, "no rows for the one-workgroup row pass");
  if (P.nrows + 1 <= (uint64_t)ROWS_SMALL_THREADS * 4)
    hipLaunchKernelGGL(k_rows_small<4>, dim3(1), dim3(ROWS_SMALL_THREADS), 0, ctx->stream, P, Q, cap, ncols, zero,
                       nzero);
  else if (P.nrows + 1 <= (uint64_t)ROWS_SMALL_THREADS * ROWS_SMALL_MAX_K)
    hipLaunchKernelGGL(k_rows_small<ROWS_SMALL_MAX_K>, dim3(1), dim3(ROWS_SMALL_THREADS), 0, ctx->stream, P, Q, cap,
                       ncols, zero, nzero);
  else
    return set_err(ctx, PTC_ERR_INTERNAL, "%llu rows for the one-workgroup row pass", (unsigned long long)P.nrows);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

// fills items: (row, chunk, first position, heavy index)
__global__ void k_items(ItemParams P) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < P.nrows; r += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t nch = P.row_nchunk[r];
    if (!nch) {
      P.row_heavy[r] = NO_HEAVY;
      continue;
    }
    unsigned long long pk = P.packed[r];
    uint64_t io = pk & 0xFFFFFFFFull;
    uint32_t hv = (uint32_t)(pk >> 32);
    P.row_heavy[r] = nch > 1 ? hv : NO_HEAVY;
    if (nch > 1) P.heavy_rows[hv] = (uint32_t)r;
    uint32_t p0 = P.rows[r].off;
    for (uint32_t k = 0; k < nch; ++k) P.items[io + k] = make_uint4((uint32_t)r, k, p0, 0u);
  }
}

// hash capacities: pow2 >= 2*bound (min 16); rows that must live in HBM
// (several chunks, or too big for the LDS table) get a global capacity
__global__ void k_hash_caps(HashCapParams P) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < P.nrows; r += (uint64_t)gridDim.x * blockDim.x) {
    unsigned long long bound = P.bound32 ? P.bound32[r] : P.bound64[r];
    if (P.cap_col && bound > P.cap_col) bound = P.cap_col;
    unsigned long long c = 16;
    while (c < 2 * bound) c <<= 1;
    uint32_t nch = P.row_nchunk[r];
    bool global = nch > 1 || c > P.lds_slots;
    P.row_hcap[r] = (uint32_t)std::min<unsigned long long>(c, 0xFFFFFFFFull);
    P.gcap[r] = (nch && global) ? c : 0ull;
  }
}

// -------------------------------------------------------------- row pass

template <class ColT>
struct ColEmpty {
  static constexpr ColT v = (ColT)~(ColT)0;
};

__device__ __forceinline__ void decode_tuple(unsigned long long key, const KeyCodec& c, uint32_t* out) {
  if (c.n == 1) {
    out[0] = (uint32_t)key;
    return;
  }
  for (uint32_t j = 0; j < c.n; ++j) out[j] = (uint32_t)((key / c.stride[j]) % c.radix[j]);
}

// Block-wide exclusive prefix of a 0/1 flag per thread; returns the rank and
// the block total.
__device__ __forceinline__ uint32_t block_rank(bool f, uint32_t* s_w, uint32_t* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned long long m = __ballot(f);
  uint32_t below = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) s_w[wid] = __popcll(m);
  __syncthreads();
  uint32_t base = 0, all = 0;
  const int nw = (int)(blockDim.x >> 6);
  for (int w = 0; w < nw; ++w) {
    uint32_t c = s_w[w];
    if (w < wid) base += c;
    all += c;
  }
  __syncthreads();
  *total = all;
  return base + below;
}

// write one output entry at position pos (row coordinates from LDS); an
// FP32 result is the pair (hi, lo) with hi = fl(sum), lo = sum - hi
template <class ColT, bool F32, class PP>
__device__ __forceinline__ void emit_entry(const PP& P, uint64_t pos, const uint32_t* s_rowc,
                                           ColT col, double v, float hi, float lo) {
  if (P.stg_col) {
    ((ColT*)P.stg_col)[pos] = col;
    if (F32) ((float2*)P.stg_val)[pos] = make_float2(hi, lo);
    else P.stg_val[pos] = v;
    return;
  }
  uint32_t* o = P.out_hedges + pos * P.dO;
  // row modes then column modes, or (column-wise products) the reverse
  uint32_t* orow = P.col_first ? o + P.colc.n : o;
  uint32_t* ocol = P.col_first ? o : o + P.nrowmodes;
  for (uint32_t j = 0; j < P.nrowmodes; ++j) orow[j] = s_rowc[j];
  if (P.colc.n == 1) {
    ocol[0] = (uint32_t)col;
  } else {
    // constant indices keep the codec in kernel-argument registers
#pragma unroll
    for (uint32_t j = 0; j < PTC_MAX_MODES; ++j)
      if (j < P.colc.n) ocol[j] = (uint32_t)(((unsigned long long)col / P.colc.stride[j]) % P.colc.radix[j]);
  }
  if (F32) {
    P.out_vals32[pos] = hi;
    P.out_err32[pos] = lo;
  } else {
    P.out_vals[pos] = v;
  }
}

// Insert-or-find col in an open-addressing LDS table of S slots (power of
// two).  Returns the slot, or -1 when all S slots were probed (table full).
// cnt (optional) counts keys this call inserted.
template <class ColT>
__device__ __forceinline__ int lds_insert(ColT* keys, uint32_t S, ColT col, uint32_t* cnt) {
  uint32_t h = (uint32_t)mix64d((unsigned long long)col) & (S - 1);
  for (uint32_t n = 0; n < S; ++n) {
    ColT kk = keys[h];
    if (kk == col) return (int)h;
    if (kk == ColEmpty<ColT>::v) {
      ColT old = atomicCAS(&keys[h], ColEmpty<ColT>::v, col);
      if (old == ColEmpty<ColT>::v) {
        if (cnt) atomicAdd(cnt, 1u);
        return (int)h;
      }
      if (old == col) return (int)h;
    }
    h = (h + 1) & (S - 1);
  }
  return -1;
}

// The same in an HBM table of mask+1 slots.
template <class ColT>
__device__ __forceinline__ long long hbm_insert(ColT* keys, uint32_t mask, ColT col, uint32_t* cnt) {
  uint32_t h = (uint32_t)mix64d((unsigned long long)col) & mask;
  for (uint64_t n = 0; n <= (uint64_t)mask; ++n) {
    ColT kk = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (kk == col) return (long long)h;
    if (kk == ColEmpty<ColT>::v) {
      ColT old = atomicCAS(&keys[h], ColEmpty<ColT>::v, col);
      if (old == ColEmpty<ColT>::v) {
        if (cnt) atomicAdd(cnt, 1u);
        return (long long)h;
      }
      if (old == col) return (long long)h;
    }
    h = (h + 1) & mask;
  }
  return -1;
}

// Bits of -0.0: the initial value of a dense accumulator entry.  When no
// product can be +-0 (no zero or tiny input values), an entry still holding
// exactly these bits after the pass was never touched, because x + (-0.0) = x
// for every x and a sum of nonzero terms is never -0.0.  Otherwise a flag
// byte per column records touches.
#define PTC_NEG_ZERO_BITS 0x8000000000000000ull

__device__ __forceinline__ bool touched(double v) {
  return (unsigned long long)__double_as_longlong(v) != PTC_NEG_ZERO_BITS;
}
__device__ __forceinline__ bool touched32(float v) { return __float_as_uint(v) != 0x80000000u; }

// FP32 compensated accumulation (error-free transformations):
//   TwoProduct  p = fl(a*b), ep = fma(a, b, -p)          (a*b = p + ep exactly)
//   TwoSum      s' = fl(o + p), e = (o - (s' - b')) + (p - b'), b' = s' - o
//                                                        (o + p = s' + e exactly)
// The sum lives in s and is updated by a hardware atomic add that returns
// the old value o; because the add rounds to nearest like the local o + p,
// the error e of that very update is recomputed locally and added, with ep,
// into the compensation word c.  The row's value is s + c: the products are
// summed as in twice the working precision, whatever the order of the
// atomics (Ogita, Rump and Oishi, "Accurate sum and dot product", SIAM J.
// Sci. Comput. 26(6), 2005, Algorithm Dot2, with the additions made atomic).
__device__ __forceinline__ float two_sum_err(float o, float p, float sn) {
  const float bp = sn - o;
  return (o - (sn - bp)) + (p - bp);
}
template <bool COMP>
__device__ __forceinline__ void acc_f32_lds(float* s, float* c, float a, float b) {
  const float p = a * b;
  if (COMP) {
    const float o = atomicAdd(s, p);
    const float e = two_sum_err(o, p, o + p) + fmaf(a, b, -p);
    if (e != 0.f) atomicAdd(c, e);
  } else {
    atomicAdd(s, p);
  }
}
template <bool COMP>
__device__ __forceinline__ void acc_f32_hbm(float* s, float* c, float a, float b) {
  const float p = a * b;
  if (COMP) {
    const float o = unsafeAtomicAdd(s, p);
    const float e = two_sum_err(o, p, o + p) + fmaf(a, b, -p);
    if (e != 0.f) unsafeAtomicAdd(c, e);
  } else {
    unsafeAtomicAdd(s, p);
  }
}
// A dense row's update: the sum word as above (FP32, a returning atomic, the
// exact error of that very update recomputed locally), the compensation --
// the TwoSum and TwoProduct errors -- collected in an FP64 word: gfx950's
// FP64 LDS atomic add runs several times faster than its FP32 one, so the
// second atomic of every update costs least there.
template <bool COMP>
__device__ __forceinline__ void acc_f32_lds_c64(float* s, double* c, float a, float b) {
  const float p = a * b;
  if (COMP) {
    const float o = atomicAdd(s, p);
    const float e = two_sum_err(o, p, o + p) + fmaf(a, b, -p);
    if (e != 0.f) atomicAdd(c, (double)e);
  } else {
    atomicAdd(s, p);
  }
}
// (s, c) -> (hi, lo): hi = fl32(s + c), lo = what that rounding left out
template <bool COMP>
__device__ __forceinline__ void f32_finish(float s, double c, float* hi, float* lo) {
  if (!COMP) {
    *hi = s;
    *lo = 0.f;
    return;
  }
  const double t = (double)s + c;
  const float h = (float)t;
  *hi = h;
  *lo = (float)(t - (double)h);
}

// Orders one wave's LDS accesses around an exchange between its lanes (no
// hardware wait: a wave's LDS accesses already complete in order).
__device__ __forceinline__ void wave_exchange_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS layout of one row pass (dynamic shared memory):
//   stage:  s_start[STAGE] u32, s_bias[STAGE] u32, s_aval[STAGE] f64,
//           s_bnd[64] u8 per wave (boundary flags of a group of 64 flops)
//   dense:  acc[ncols] f64 (+ flag[ncols] u8 when flags are needed)  (numeric)
//           bits[ceil(ncols/32)] u32                                 (symbolic)
//   hash:   keys[S] ColT, vals[S] f64             (S = lds_slots)
// MERGE = false: one workgroup per work item (a row or a chunk of a split
// row) expands its flops; a single-chunk row in LDS is finished at once, a
// chunk of a split row stores its partial (dense accumulator or bitmap) to
// its slab with plain stores, and a row accumulated in HBM is left there.
// MERGE = true (a separate launch, so the kernel boundary orders the slab
// stores before these loads -- no device-scope fence, which on gfx950 writes
// back and invalidates the whole L2, is ever issued): one workgroup per split
// or HBM row reduces the slabs in chunk order (deterministic) and finishes
// the row.
// VAL (ptc_value_type): 0 FP64 values; 1 FP32 values, FP32 accumulators
// with TwoProduct + TwoSum compensation; 2 FP32 values, plain FP32
// accumulators; 3 FP32 values, exact products accumulated in FP64.
// One work item of a row pass: widx indexes P.items (MERGE: P.merge_rows,
// or the row itself).  Every early return is uniform over the workgroup.
template <int PASS, bool DENSE, class ColT, bool MERGE, int VAL, class PP>
__device__ __forceinline__ void pass_item(const PP& P, uint32_t widx) {
  constexpr bool F32 = VAL == 1 || VAL == 2;  // FP32 (sum, compensation) accumulators
  constexpr bool F32IN = VAL != 0;            // FP32 values in, (hi, lo) FP32 pairs out
  constexpr bool COMP = VAL == 1;
  static_assert(!F32IN || PASS == PASS_NUMERIC, "FP32 values only matter to the numeric pass");
  const uint32_t NT = blockDim.x;  // a multiple of 64, at most PB_MAX
  const uint32_t STAGE = 2 * NT;
  const int NW = (int)(NT >> 6);
  extern __shared__ __align__(16) unsigned char smem[];
  uint32_t* s_start = (uint32_t*)smem;
  // B position of a staged A position's first flop, less the flop index,
  // modulo 2^32: adding any flop index of the position gives its B position
  uint32_t* s_bias = (uint32_t*)(smem + STAGE * 4);
  double* s_aval = (double*)(smem + STAGE * 8);
  float* s_aval32 = (float*)s_aval;  // FP32: the first half of the same region
  unsigned char* acc_base = smem + STAGE * 16 + NT;
  __shared__ uint32_t s_w[PB_MAX / 64];
  __shared__ uint32_t s_misc[4];
  __shared__ unsigned long long s_base;  // reservation mode: the row's first output entry
  __shared__ uint32_t s_rowc[PTC_MAX_MODES];

  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // dense rows stream 16-bit column keys (their column space is below 2^16)
  using BCol = typename std::conditional<DENSE, uint16_t, ColT>::type;
  const BCol* __restrict__ bcol = (const BCol*)P.bcol;
  uint4 it = make_uint4(0, 0, 0, 0);
  if (MERGE) {
    it.x = P.merge_rows ? P.merge_rows[widx] : widx;
  } else {
    it = P.items[widx];
  }
  const uint32_t r = it.x, k = it.y;
  const KtSlice row = P.rows[r];
  const uint32_t nch = P.row_nchunk[r];
  if (MERGE && DENSE && nch <= 1) return;                               // not a split row
  if (MERGE && !DENSE && (nch == 0 || P.row_gcap[r] == 0)) return;  // not an HBM row
  if (MERGE && !DENSE && PASS == PASS_SYMBOLIC) return;  // HBM rows counted with atomics
  const unsigned long long flo = P.apre[row.off];
  const unsigned long long fr = P.apre[(uint64_t)row.off + row.len] - flo;
  const unsigned long long per = (fr + nch - 1) / nch;
  const unsigned long long g0 = flo + (unsigned long long)k * per;
  const unsigned long long g1 = min(flo + fr, g0 + per);
  const uint64_t pend = (uint64_t)row.off + row.len;
  const uint64_t ncols = P.ncols;
  // dense rows: the column space (at most a few thousand columns) as 32 bits
  const uint32_t ncols32 = (uint32_t)min<uint64_t>(ncols, 0xFFFFFFFFull);
  // a B index or a column out of range (inconsistent tables): the lane drops
  // the multiply-add and the thread reports it once, after the item
  bool bad = false;
  const bool flags = P.use_flags != 0;

  // accumulator storage
  double* acc = (double*)acc_base;
  // FP32 accumulators: compensations in FP64 [ncols], then the FP32 sums [ncols]
  double* acc_c64 = (double*)acc_base;
  float* acc_s = (float*)(acc_base + ncols * 8);
  // touched columns (when values may give zero products): one bit per
  // column, set by an LDS atomic OR; the finish skips empty 64-column tiles
  uint32_t* fbits = (uint32_t*)(acc_base + ncols * (F32 ? 12 : 8));
  const uint32_t nfw = (uint32_t)((ncols + 31) / 32);
  uint32_t* bits = (uint32_t*)acc_base;
  ColT* hkeys = (ColT*)acc_base;
  uint32_t S = 0;  // LDS hash slots for this row (0: HBM table)
  ColT* gkeys = nullptr;
  double* gvals = nullptr;
  uint32_t gmask = 0;
  const uint32_t nbits = (uint32_t)((ncols + 31) / 32);

  if (DENSE && MERGE) {
    // reduce the row's slabs in chunk order
    const uint64_t slab0 = P.row_slab[r];
    // touched bits: a wave reduces 64 consecutive columns per step (the
    // columns of a step start at a multiple of 64), so each bitmap word is
    // assembled by a ballot and stored whole by one lane -- no LDS atomics
    if (PASS == PASS_SYMBOLIC) {
      for (uint32_t w = threadIdx.x; w < nbits; w += NT) {
        uint32_t o = 0;
        for (uint32_t kk = 0; kk < nch; ++kk) o |= P.hbits[(slab0 + kk) * nbits + w];
        bits[w] = o;
      }
    } else if (F32) {
      // slabs hold (sum, compensation) pairs; combine them in chunk order
      const float2* h2 = (const float2*)P.hacc;
      for (uint64_t c = threadIdx.x; c < ncols; c += NT) {
        float S = -0.0f;
        double Cc = 0.0;
        unsigned char fl = 0;
#pragma unroll 4
        for (uint32_t kk = 0; kk < nch; ++kk) {
          const float2 x = h2[(slab0 + kk) * ncols + c];
          const float sn = S + x.x;
          if (COMP) Cc += (double)x.y + (double)two_sum_err(S, x.x, sn);
          S = sn;
        }
        if (flags)
          for (uint32_t j = 0; j < nch; ++j) fl |= P.hflag[(slab0 + j) * ncols + c];
        acc_s[c] = S;
        acc_c64[c] = Cc;
        if (flags) {
          const unsigned long long m = __ballot(fl != 0);
          if ((lane & 31) == 0) fbits[c >> 5] = (uint32_t)(m >> (lane & 32));
        }
      }
    } else {
      // CQ columns a thread at a time, four chunks of each in flight
      // together (sixteen loads); every column's sum in chunk order
      constexpr int CQ = 4;
      for (uint64_t c0 = threadIdx.x; c0 < ncols; c0 += CQ * NT) {
        double sum[CQ];
        unsigned char fl[CQ];
#pragma unroll
        for (int q = 0; q < CQ; ++q) {
          sum[q] = __longlong_as_double((long long)PTC_NEG_ZERO_BITS);
          fl[q] = 0;
        }
        uint32_t kk = 0;
        for (; kk + 4 <= nch; kk += 4) {
          double x[CQ][4];
#pragma unroll
          for (int q = 0; q < CQ; ++q) {
            const uint64_t c = c0 + (uint64_t)q * NT;
            const double* a = P.hacc + (slab0 + kk) * ncols + min(c, ncols - 1);
#pragma unroll
            for (int j = 0; j < 4; ++j) x[q][j] = a[j * ncols];  // (a clamped column: read, not used)
          }
#pragma unroll
          for (int q = 0; q < CQ; ++q)
#pragma unroll
            for (int j = 0; j < 4; ++j) sum[q] += x[q][j];
        }
        for (; kk < nch; ++kk)
#pragma unroll
          for (int q = 0; q < CQ; ++q) sum[q] += P.hacc[(slab0 + kk) * ncols + min(c0 + (uint64_t)q * NT, ncols - 1)];
#pragma unroll
        for (int q = 0; q < CQ; ++q) {
          const uint64_t c = c0 + (uint64_t)q * NT;
          if (flags && c < ncols)
            for (uint32_t j = 0; j < nch; ++j) fl[q] |= P.hflag[(slab0 + j) * ncols + c];
          if (c < ncols) acc[c] = sum[q];
          if (flags) {  // lanes past the columns hold no flag
            const unsigned long long m = __ballot(fl[q] != 0);
            if ((lane & 31) == 0 && c < ncols) fbits[c >> 5] = (uint32_t)(m >> (lane & 32));
          }
        }
      }
    }
  } else if (DENSE) {
    if (PASS == PASS_SYMBOLIC) {
      for (uint32_t w = threadIdx.x; w < nbits; w += NT) bits[w] = 0;
    } else if (F32) {
      for (uint64_t c = threadIdx.x; c < ncols; c += NT) {
        acc_s[c] = -0.0f;
        acc_c64[c] = 0.0;
      }
      if (flags)
        for (uint32_t w = threadIdx.x; w < nfw; w += NT) fbits[w] = 0;
    } else {
      const double nz = __longlong_as_double((long long)PTC_NEG_ZERO_BITS);
      for (uint64_t c = threadIdx.x; c < ncols; c += NT) acc[c] = nz;
      if (flags)
        for (uint32_t w = threadIdx.x; w < nfw; w += NT) fbits[w] = 0;
    }
  } else {
    uint32_t hc = P.row_hcap[r];
    unsigned long long goff = P.row_goff[r];
    if (P.row_gcap[r] == 0) {
      S = hc;
      for (uint32_t i = threadIdx.x; i < S; i += NT) hkeys[i] = ColEmpty<ColT>::v;
      if (PASS == PASS_NUMERIC) {
        double* hv = (double*)(acc_base + (size_t)S * sizeof(ColT));
        for (uint32_t i = threadIdx.x; i < S; i += NT) hv[i] = 0.0;
      }
    } else {
      gkeys = (ColT*)P.gkeys + goff;
      gvals = P.gvals + goff;
      gmask = hc - 1;
    }
  }
  if (threadIdx.x < 4) s_misc[threadIdx.x] = 0;  // [0] LDS distinct count, [3] table full
  __syncthreads();
  double* hvals = (double*)(acc_base + (size_t)S * sizeof(ColT));
  float* hs32 = (float*)hvals;  // FP32: sums [S], then compensations [S]
  float* hc32 = hs32 + S;

  // ---- expand the flop range [g0, g1) ----
  uint64_t q0 = MERGE ? pend : it.z;
  while (q0 < pend) {
    const unsigned long long qstart = P.apre[q0];
    if (qstart >= g1) break;
    // stage STAGE raw positions, keeping (in order) those with flops in range
    const uint64_t qlim = min<uint64_t>(pend, q0 + STAGE);
    uint32_t keep[2];
    uint32_t nkeep = 0;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const uint64_t q = q0 + (uint64_t)threadIdx.x * 2 + j;
      keep[j] = 0;
      if (q < qlim) {
        const uint32_t len = P.pos_len[q];
        const unsigned long long a = P.apre[q];
        if (len && a + len > g0 && a < g1) {
          keep[j] = 1;
          ++nkeep;
        }
      }
    }
    // block-wide exclusive prefix of the per-thread kept counts
    uint32_t incl = nkeep;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t y = __shfl_up(incl, o, 64);
      if (lane >= o) incl += y;
    }
    if (lane == 63) s_w[wid] = incl;
    __syncthreads();
    uint32_t wbase = 0, ns = 0;
    for (int w = 0; w < NW; ++w) {
      if (w < wid) wbase += s_w[w];
      ns += s_w[w];
    }
    ns = (uint32_t)__builtin_amdgcn_readfirstlane((int)ns);  // the same in every lane: keep it scalar
    uint32_t dst = wbase + incl - nkeep;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      if (!keep[j]) continue;
      const uint64_t q = q0 + (uint64_t)threadIdx.x * 2 + j;
      const unsigned long long a = P.apre[q];
      s_start[dst] = (uint32_t)((a > g0 ? a : g0) - g0);
      s_bias[dst] = (uint32_t)((unsigned long long)P.pos_off[q] + g0 - a);
      if (PASS == PASS_NUMERIC) {
        if (F32IN) s_aval32[dst] = P.aval32[q];
        else s_aval[dst] = P.aval[q];
      }
      ++dst;
    }
    const unsigned long long qend = (qlim < pend) ? P.apre[qlim] : (flo + fr);
    const uint32_t fb = (uint32_t)((qstart > g0 ? qstart : g0) - g0);
    const uint32_t fe = (uint32_t)(min(qend, g1) - g0);
    __syncthreads();
    if (ns) {
      // each wave expands a contiguous, 64-aligned share of [fb, fe).  The
      // walk's state is the same in every lane of a wave and is kept in
      // scalar registers (readfirstlane / readlane), so its tests are
      // scalar branches rather than exec-masked regions.
      const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane(wid);
      const uint32_t nf = fe - fb;
      const uint32_t span = ((nf + NT - 1) / NT) * 64;
      const uint32_t wb = fb + wv * span;
      const uint32_t we = min(fe, wb + span);
      // position of the wave's first flop (searched once, by every lane)
      uint32_t qa = 0;
      if (wb < we) {
        uint32_t lo = 0, hi = ns - 1;
        while (lo < hi) {
          uint32_t mid = (lo + hi + 1) >> 1;
          if ((uint32_t)__builtin_amdgcn_readfirstlane((int)s_start[mid]) <= wb) lo = mid; else hi = mid - 1;
        }
        qa = lo;
      }
      // U groups of 64 flops per step: positions first (LDS only), then all
      // global loads, then the accumulations -- 2U loads in flight per wave
      constexpr int U = pass_unroll(VAL);
      unsigned char* s_bnd = smem + STAGE * 16 + wv * 64;  // this wave's boundary flags
      for (uint32_t f0 = wb; f0 < we; f0 += 64 * U) {
        uint32_t ii[U];
        uint32_t bb[U];
        bool act[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint32_t fu = f0 + 64 * u;
          act[u] = false;
          ii[u] = qa;
          bb[u] = 0;
          if (fu < we) {  // wave-uniform
            // every kept position has at least one flop, so the wave's 64
            // flops lie in positions [qa, qa + 63], and after this step
            // s_start[qa] <= fu < s_start[qa + 1] (= nx; past the last
            // position, no start)
            // both starts read at once (indices clamped into the staged
            // range, ns >= 1), one LDS round trip per group
            const uint32_t r1 = s_start[min(qa + 1, ns - 1)], r2 = s_start[min(qa + 2, ns - 1)];
            const uint32_t n1 = qa + 1 < ns ? (uint32_t)__builtin_amdgcn_readfirstlane((int)r1) : 0xFFFFFFFFu;
            const uint32_t n2 = qa + 2 < ns ? (uint32_t)__builtin_amdgcn_readfirstlane((int)r2) : 0xFFFFFFFFu;
            uint32_t nx = n1;
            if (n1 <= fu) {
              ++qa;
              nx = n2;
            }
            const uint32_t f = min(fu + lane, we - 1);
            uint32_t i = qa;
            if (nx <= min(fu + 63, we - 1)) {
              // the group crosses positions: lane k flags the flop where
              // position qa + 1 + k starts (a byte per flop of the group),
              // and a lane's position is qa + the flags at or before its flop
              const uint32_t bi = qa + 1 + lane;
              const uint32_t dk = bi < ns ? s_start[bi] - fu : 64u;
              // the lanes of the wave exchange the flags through LDS: the
              // wave's LDS accesses complete in order, and the wavefront-scope
              // fences keep the compiler from moving a flag access across the
              // exchange or answering a lane's load from its own store
              wave_exchange_fence();
              s_bnd[lane] = 0;
              if (dk < 64) s_bnd[dk] = 1;
              wave_exchange_fence();
              const unsigned long long m = __ballot(s_bnd[lane] != 0);
              i = qa + (uint32_t)__popcll(m & (~0ull >> (63 - lane)));
            }
            ii[u] = i;
            bb[u] = s_bias[i] + f;
            act[u] = fu + lane < we;
            const bool inb = (uint64_t)bb[u] < P.nb;  // cannot fail for consistent tables
            bad |= act[u] && !inb;
            act[u] = act[u] && inb;
            qa = (uint32_t)__builtin_amdgcn_readlane((int)i, 63);
          }
        }
        ColT col[U];
        double bv[U];
        float bv32[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          col[u] = 0;
          bv[u] = 0.0;
          bv32[u] = 0.f;
          if (act[u]) {
            col[u] = bcol[bb[u]];
            if (PASS == PASS_NUMERIC) {
              if (F32IN) bv32[u] = P.bval32[bb[u]];
              else bv[u] = P.bval[bb[u]];
            }
          }
        }
        if (DENSE && P.col_lo) {  // a column window: window-local columns (others wrap out of range)
#pragma unroll
          for (int u = 0; u < U; ++u) col[u] = (ColT)((uint32_t)col[u] - P.col_lo);
        }
        if constexpr (F32 && DENSE) {
          // all U sum updates first (a returning atomic each when
          // compensating), then the U compensation updates: one wait on LDS
          // returns covers the whole group
          float o[U], pr[U], ep[U];
          bool ok[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            ok[u] = act[u] && (uint32_t)col[u] < ncols32;
            bad |= act[u] && !ok[u] && !P.col_win;
            o[u] = 0.f;
            pr[u] = 0.f;
            ep[u] = 0.f;
            if (ok[u]) {
              const float a = s_aval32[ii[u]], b = bv32[u];
              pr[u] = a * b;
              if (COMP) {
                ep[u] = fmaf(a, b, -pr[u]);
                o[u] = atomicAdd(&acc_s[col[u]], pr[u]);
              } else {
                atomicAdd(&acc_s[col[u]], pr[u]);
              }
              if (flags) atomicOr(&fbits[(uint32_t)col[u] >> 5], 1u << ((uint32_t)col[u] & 31));
            }
          }
          if (COMP) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
              if (!ok[u]) continue;
              const float e = two_sum_err(o[u], pr[u], o[u] + pr[u]) + ep[u];
              if (e != 0.f) atomicAdd(&acc_c64[col[u]], (double)e);
            }
          }
          continue;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const ColT c = col[u];
          if (DENSE) {
            const bool ok = act[u] && (uint32_t)c < ncols32;
            bad |= act[u] && !ok && !P.col_win;  // a windowed row's other columns belong to other windows
            if (!ok) continue;
            if (PASS == PASS_SYMBOLIC) {
              atomicOr(&bits[c >> 5], 1u << (c & 31));
            } else {
              if (F32) acc_f32_lds_c64<COMP>(&acc_s[c], &acc_c64[c], s_aval32[ii[u]], bv32[u]);
              else if (F32IN) atomicAdd(&acc[c], (double)s_aval32[ii[u]] * (double)bv32[u]);
              else atomicAdd(&acc[c], s_aval[ii[u]] * bv[u]);
              if (flags) atomicOr(&fbits[(uint32_t)c >> 5], 1u << ((uint32_t)c & 31));
            }
          } else {
            if (!act[u]) continue;
            const double v = (PASS != PASS_NUMERIC || F32) ? 0.0
                           : F32IN ? (double)s_aval32[ii[u]] * (double)bv32[u] : s_aval[ii[u]] * bv[u];
            if (S) {
              const int h = lds_insert<ColT>(hkeys, S, c, PASS == PASS_SYMBOLIC ? &s_misc[0] : nullptr);
              if (h < 0) s_misc[3] = 1u;
              else if (PASS == PASS_NUMERIC) {
                if (F32) acc_f32_lds<COMP>(&hs32[h], &hc32[h], s_aval32[ii[u]], bv32[u]);
                else atomicAdd(&hvals[h], v);
              }
            } else {
              const long long h = hbm_insert<ColT>(gkeys, gmask, c, PASS == PASS_SYMBOLIC ? &P.row_nnz[r] : nullptr);
              if (h < 0) atomicOr(P.overflow, 0x80000000u);  // cannot happen: sized from exact or bound counts
              else if (PASS == PASS_NUMERIC) {
                if (F32) {
                  float* g2 = (float*)gvals;  // (sum, compensation) pairs
                  acc_f32_hbm<COMP>(&g2[2 * h], &g2[2 * h + 1], s_aval32[ii[u]], bv32[u]);
                } else {
                  unsafeAtomicAdd(&gvals[h], v);
                }
              }
            }
          }
        }
      }
    }
    __syncthreads();
    q0 = qlim;
  }

  if (bad) atomicOr(P.overflow, 0x40000000u);

  // ---- a split row (or a row accumulated in HBM) is finished by the merge launch ----
  if (!MERGE) {
    if (DENSE && nch > 1) {
      const uint64_t slab = P.row_slab[r] + k;
      if (PASS == PASS_SYMBOLIC) {
        uint32_t* gb = P.hbits + slab * nbits;
        for (uint32_t w = threadIdx.x; w < nbits; w += NT) gb[w] = bits[w];
      } else if (F32) {
        float2* ga = (float2*)P.hacc + slab * ncols;
        // the compensation rounded to FP32 (a second-order term of the chunk's sum)
        for (uint64_t c = threadIdx.x; c < ncols; c += NT) ga[c] = make_float2(acc_s[c], (float)acc_c64[c]);
        if (flags) {
          unsigned char* gf = P.hflag + slab * ncols;
          for (uint64_t c = threadIdx.x; c < ncols; c += NT) gf[c] = (unsigned char)((fbits[c >> 5] >> (c & 31)) & 1u);
        }
      } else {
        double* ga = P.hacc + slab * ncols;
        for (uint64_t c = threadIdx.x; c < ncols; c += NT) ga[c] = acc[c];
        if (flags) {
          unsigned char* gf = P.hflag + slab * ncols;
          for (uint64_t c = threadIdx.x; c < ncols; c += NT) gf[c] = (unsigned char)((fbits[c >> 5] >> (c & 31)) & 1u);
        }
      }
      return;
    }
    if (!DENSE && S == 0) return;
  }
  __syncthreads();

  // ---- finish the row ----
  if (PASS == PASS_SYMBOLIC) {
    if (DENSE) {
      // the row total, or (column windows) each window's count and the total
      const uint32_t nwin = P.win_cnt ? P.win_n : 1u;
      const uint32_t wpw = P.win_cnt ? P.win_width / 32 : nbits;  // bitmap words per window
      uint32_t tot = 0;
      for (uint32_t t = 0; t < nwin; ++t) {
        const uint32_t w0 = t * wpw, w1 = t + 1 == nwin ? nbits : min(nbits, w0 + wpw);
        uint32_t c = 0;
        for (uint32_t w = w0 + threadIdx.x; w < w1; w += NT) c += __popc(bits[w]);
        for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
        if (lane == 0) s_w[wid] = c;
        __syncthreads();
        if (threadIdx.x == 0) {
          uint32_t s = 0;
          for (int w = 0; w < NW; ++w) s += s_w[w];
          if (P.win_cnt) P.win_cnt[(uint64_t)t * P.win_stride + r] = s;
          tot += s;
        }
        __syncthreads();
      }
      if (threadIdx.x == 0) P.row_nnz[r] = tot;
    } else if (S) {
      if (threadIdx.x == 0) P.row_nnz[r] = s_misc[0];
    }
    // HBM hash rows counted themselves with atomics
    return;
  }

  // NUMERIC: count, check capacity, write
  if (threadIdx.x == 0) {
#pragma unroll
    for (uint32_t j = 0; j < PTC_MAX_MODES; ++j)
      if (j < P.nrowmodes)
        s_rowc[j] = P.rowc.n == 1 ? (uint32_t)row.key : (uint32_t)((row.key / P.rowc.stride[j]) % P.rowc.radix[j]);
  }
  __syncthreads();

  // The span (columns, LDS slots or HBM slots) is split into 64-entry tiles,
  // each wave owning a contiguous run of them: a wave counts its entries with
  // ballots, one block-wide exchange gives every wave its first output
  // position, and the wave writes its entries in span order (consecutive
  // lanes, consecutive outputs).  Tiles with nothing in them cost a ballot
  // (with touched bits, one broadcast LDS word), so a sparse row in a wide
  // column space is not charged per column.
  const uint64_t span = DENSE ? ncols : (S ? S : (uint64_t)gmask + 1);
  const uint64_t ntiles = (span + 63) / 64;
  const uint64_t tpw = (ntiles + NW - 1) / NW;
  const uint64_t t_lo = min<uint64_t>(ntiles, (uint64_t)wid * tpw), t_hi = min<uint64_t>(ntiles, t_lo + tpw);
  auto present = [&](uint64_t c) -> bool {
    if (c >= span) return false;
    if (DENSE) {
      if (flags) return (fbits[c >> 5] >> (c & 31)) & 1u;
      return F32 ? touched32(acc_s[c]) : touched(acc[c]);
    }
    if (S) return hkeys[c] != ColEmpty<ColT>::v;
    return __hip_atomic_load(&gkeys[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != ColEmpty<ColT>::v;
  };
  uint32_t wcnt = 0;
  for (uint64_t t = t_lo; t < t_hi; ++t) wcnt += (uint32_t)__popcll(__ballot(present(t * 64 + lane)));
  if (lane == 0) s_w[wid] = wcnt;
  __syncthreads();
  uint32_t total = 0, wbase = 0;
  for (int w = 0; w < NW; ++w) {
    const uint32_t x = s_w[w];
    if (w < wid) wbase += x;
    total += x;
  }
  uint32_t cap = P.out_cap && !P.out_soft ? P.out_cap[r] : 0xFFFFFFFFu;
  const bool full = s_misc[3] != 0;  // LDS table filled up: count unknown
  const bool soft = P.out_soft != 0;  // only the output's total limit applies
  if (threadIdx.x == 0) {
    if (soft) {
      P.row_nnz[r] = full ? 0xFFFFFFFFu : total;
      if (full) atomicAdd(P.overflow, 1u);
    } else if (P.out_cap) {
      // capacity mode: record the count; a row over its capacity (or whose
      // table filled) is redone by the exact two-pass path
      P.row_nnz[r] = full ? 0xFFFFFFFFu : total;
      if (full || total > cap) atomicAdd(P.overflow, 1u);
    } else if (full || P.row_nnz[r] != total) {
      atomicOr(P.overflow, 0x80000000u);  // exact mode: a mismatch is an internal error
    }
  }
  if (full || total > cap) return;
  // never write past the row's allocation (exact mode: the symbolic count)
  const uint32_t limit = soft ? total : min(total, P.out_cap ? cap : P.row_nnz[r]);
  uint64_t base;
  if (P.out_count) {
    // reservation: the row's entries go straight to the output, placed by
    // one fetch-add per row (the reference's append, fksCuckooTC.hpp:345-355;
    // rows land in completion order)
    if (threadIdx.x == 0) {
      unsigned long long b = atomicAdd(P.out_count, (unsigned long long)limit);
      if (b + limit > P.out_limit) {
        // soft: the output is smaller than the result (redone by the
        // caller); else it cannot happen: the output holds every capacity
        if (soft) atomicAdd(P.overflow, 1u);
        else atomicOr(P.overflow, 0x20000000u);
        b = ~0ull;
      } else if (P.out_rowoff) {
        P.out_rowoff[r] = b;
      }
      s_base = b;
    }
    __syncthreads();
    base = s_base;
    if (base == ~0ull) return;
  } else {
    base = P.out_base[r];
  }
  uint32_t run = wbase;
  for (uint64_t t = t_lo; t < t_hi; ++t) {
    const uint64_t c = t * 64 + lane;
    const bool f = present(c);
    const unsigned long long m = __ballot(f);
    if (!m) continue;  // wave-uniform
    const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (f && run + rank < limit) {
      ColT col = 0;
      double v = 0.0;
      float fs = 0.f;
      double fc = 0.0;
      if (DENSE) {
        col = (ColT)(c + P.col_lo);
        if (F32) {
          fs = acc_s[c];
          fc = acc_c64[c];
        } else {
          v = acc[c];
        }
      } else if (S) {
        col = hkeys[c];
        if (F32) {
          fs = hs32[c];
          fc = hc32[c];
        } else {
          v = hvals[c];
        }
      } else {
        col = __hip_atomic_load(&gkeys[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (F32) {
          float* g2 = (float*)gvals;
          fs = __hip_atomic_load(&g2[2 * c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          fc = __hip_atomic_load(&g2[2 * c + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
          v = __hip_atomic_load(&gvals[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      float hi = 0.f, lo = 0.f;
      if (F32) {
        f32_finish<COMP>(fs, fc, &hi, &lo);
      } else if (F32IN) {
        hi = (float)v;  // the FP64 sum as an FP32 pair
        lo = (float)(v - (double)hi);
      }
      emit_entry<ColT, F32IN>(P, base + run + rank, s_rowc, col, v, hi, lo);
    }
    run += (uint32_t)__popcll(m);
  }
}

template <int PASS, bool DENSE, class ColT, bool MERGE, int VAL>
__global__ void __launch_bounds__(PB_MAX) k_pass(PassParams P) {
  if (P.live_items) {  // a grid sized by a bound: work items (split rows) past the count exit
    const unsigned long long L = *P.live_items;
    if (blockIdx.x >= (MERGE ? (uint32_t)(L >> 32) : (uint32_t)L)) return;
  }
  pass_item<PASS, DENSE, ColT, MERGE, VAL>(P, blockIdx.x);
}

// Persistent row pass over the work of many contractions: each workgroup
// takes (job, work index) pairs from the queue until it is exhausted.  The
// queue lists the heaviest work first, so the tail of the launch is short
// items whatever the mix of job sizes.  The symbolic pass claims
// QUEUE_CLAIM consecutive entries per fetch-add: fetch-adds on one word
// serialise (tens of nanoseconds each on this GPU), and its light entries
// (column bitmaps only) came out bounded by the claims (256-job batch:
// 22.4 -> 16.5 ms of symbolic device time with the split-item counters
// aggregated per wave).  The numeric pass claims one entry at a time: its
// entries are long enough, and several heavy entries in one workgroup left
// a longer tail (15.7 -> 17.0 ms).
constexpr uint32_t QUEUE_CLAIM = 4;
template <int PASS, bool DENSE, class ColT, bool MERGE, int VAL>
__global__ void __launch_bounds__(PB_MAX) k_pass_batch(const PassParams* __restrict__ jobs, const uint2* __restrict__ queue,
                                                   uint32_t nq, uint32_t* __restrict__ head) {
  __shared__ uint32_t s_q;
  constexpr uint32_t claim = PASS == PASS_SYMBOLIC ? QUEUE_CLAIM : 1u;
  for (;;) {
    __syncthreads();  // the previous item is done with LDS
    if (threadIdx.x == 0) s_q = atomicAdd(head, claim);
    __syncthreads();
    const uint32_t q0 = s_q;
    if (q0 >= nq) return;
    const uint32_t q1 = min(nq, q0 + claim);
    for (uint32_t q = q0; q < q1; ++q) {
      if (q > q0) __syncthreads();  // the previous item is done with LDS
      const uint2 w = queue[q];
      // a private copy: the job's parameters are loaded once per item (into
      // scalar registers), not re-read after every atomic of the item
      const PassParamsG P = reinterpret_cast<const PassParamsG*>(jobs)[w.x];
      pass_item<PASS, DENSE, ColT, MERGE, VAL>(P, w.y);
    }
  }
}

// learned / upper-bound: staged rows -> final output
template <class ColT, bool F32>
__global__ void __launch_bounds__(PB) k_emit(EmitParams P) {
  const uint32_t r = blockIdx.x;
  if (r >= P.nrows) return;
  const uint32_t n = P.row_nnz[r];
  if (!n || n > P.cap[r] || (P.row_fb && P.row_fb[r])) return;  // fallback rows are redone exactly
  __shared__ uint32_t s_rowc[PTC_MAX_MODES];
  if (threadIdx.x == 0) {
    const unsigned long long key = P.rows[r].key;
#pragma unroll
    for (uint32_t j = 0; j < PTC_MAX_MODES; ++j)
      if (j < P.nrowmodes)
        s_rowc[j] = P.rowc.n == 1 ? (uint32_t)key : (uint32_t)((key / P.rowc.stride[j]) % P.rowc.radix[j]);
  }
  __syncthreads();
  const uint64_t src = P.cap_base[r], dst = P.out_base[r];
  const ColT* sc = (const ColT*)P.stg_col;
  for (uint32_t i = threadIdx.x; i < n; i += PB) {
    const ColT col = sc[src + i];
    uint32_t* o = P.out_hedges + (dst + i) * P.dO;
    for (uint32_t j = 0; j < P.nrowmodes; ++j) o[j] = s_rowc[j];
    o += P.nrowmodes;
    if (P.colc.n == 1) {
      o[0] = (uint32_t)col;
    } else {
#pragma unroll
      for (uint32_t j = 0; j < PTC_MAX_MODES; ++j)
        if (j < P.colc.n) o[j] = (uint32_t)(((unsigned long long)col / P.colc.stride[j]) % P.colc.radix[j]);
    }
    if (F32) {
      const float2 x = ((const float2*)P.stg_val)[src + i];
      P.out_vals32[dst + i] = x.x;
      P.out_err32[dst + i] = x.y;
    } else {
      P.out_vals[dst + i] = P.stg_val[src + i];
    }
  }
}

// ------------------------------------------------------------- launchers

static unsigned gridn(uint64_t n, unsigned t, unsigned cap = 16384) {
  uint64_t g = (n + t - 1) / t;
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

int launch_probe(ptc_ctx* ctx, const ProbeParams& P, bool rest64) {
  if (!P.npos) return PTC_OK;
  if (rest64)
    hipLaunchKernelGGL((k_probe<unsigned long long>), dim3(gridn(P.npos, 256)), dim3(256), 0, ctx->stream, P);
  else
    hipLaunchKernelGGL((k_probe<uint32_t>), dim3(gridn(P.npos, 256)), dim3(256), 0, ctx->stream, P);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

int launch_rows(ptc_ctx* ctx, const RowParams& P) {
  if (!P.nrows) return PTC_OK;
  hipLaunchKernelGGL(k_rows, dim3(gridn(P.nrows, 256)), dim3(256), 0, ctx->stream, P);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

// First position of every chunk k > 0 of a split row, one wavefront per
// work item: the last position p of the row with apre[p] <= the chunk's first
// flop (the position whose flop range holds it; zero-length positions share
// the apre of the position after them), found by a 64-way search -- a few
// rounds of 64 loads instead of a pass over every A position.
__global__ void __launch_bounds__(256) k_item_starts(ItemP0Params P) {
  const uint64_t n = P.live ? (*P.live & 0xFFFFFFFFull) : P.nitems;
  const int lane = threadIdx.x & 63;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
  for (uint64_t i = (uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; i < n; i += waves) {
    uint4 it;
    if (P.find_rows) {
      // the item's row: the last row whose first item (the low word of its
      // scanned count) is at or before i, found 64 rows at a time
      uint64_t lo = 0, hi = P.nrows;  // (packed[lo] & 0xFFFFFFFF) <= i throughout
      while (hi - lo > 64) {
        const uint64_t step = (hi - lo + 63) / 64;
        const uint64_t q = lo + (uint64_t)lane * step;
        const unsigned long long m = __ballot(q < hi && (P.packed[q] & 0xFFFFFFFFull) <= i);
        const uint64_t nlo = lo + (uint64_t)(63 - __clzll(m)) * step;
        hi = min(hi, nlo + step);
        lo = nlo;
      }
      const uint64_t q = lo + lane;
      const unsigned long long m = __ballot(q < hi && (P.packed[q] & 0xFFFFFFFFull) <= i);
      const uint32_t r = (uint32_t)(lo + (uint64_t)(63 - __clzll(m)));
      it = make_uint4(r, (uint32_t)(i - (P.packed[r] & 0xFFFFFFFFull)), P.rows[r].off, 0u);
      if (it.y == 0) {
        if (lane == 0) P.items[i] = it;
        continue;
      }
    } else {
      it = P.items[i];
      if (it.y == 0) continue;  // a row's first chunk starts at its first position
    }
    const uint32_t r = it.x, nch = P.row_nchunk[r];
    const KtSlice row = P.rows[r];
    const unsigned long long flo = P.apre[row.off], f = P.row_flops[r];
    const unsigned long long target = flo + (unsigned long long)it.y * ((f + nch - 1) / nch);
    uint64_t lo = row.off, hi = (uint64_t)row.off + row.len;  // apre[lo] <= target throughout
    while (hi - lo > 64) {
      const uint64_t step = (hi - lo + 63) / 64;
      const uint64_t q = lo + (uint64_t)lane * step;
      const unsigned long long m = __ballot(q < hi && P.apre[q] <= target);
      const uint64_t nlo = lo + (uint64_t)(63 - __clzll(m)) * step;  // lane 0 holds: m != 0
      hi = min(hi, nlo + step);
      lo = nlo;
    }
    const uint64_t q = lo + lane;
    const unsigned long long m = __ballot(q < hi && P.apre[q] <= target);
    it.z = (uint32_t)(lo + (uint64_t)(63 - __clzll(m)));
    if (lane == 0) {
      if (P.find_rows) P.items[i] = it;
      else P.items[i].z = it.z;
    }
  }
}

int launch_item_starts(ptc_ctx* ctx, const ItemP0Params& Q) {
  if (Q.npos && Q.nitems)
    hipLaunchKernelGGL(k_item_starts, dim3(gridn(Q.nitems, 4, 4096)), dim3(256), 0, ctx->stream, Q);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

int launch_items(ptc_ctx* ctx, const ItemParams& P, const ItemP0Params& Q) {
  if (P.nrows) hipLaunchKernelGGL(k_items, dim3(gridn(P.nrows, 256)), dim3(256), 0, ctx->stream, P);
  if (Q.npos && Q.nitems)
    hipLaunchKernelGGL(k_item_starts, dim3(gridn(Q.nitems, 4, 4096)), dim3(256), 0, ctx->stream, Q);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

int launch_hash_caps(ptc_ctx* ctx, const HashCapParams& P) {
  if (!P.nrows) return PTC_OK;
  hipLaunchKernelGGL(k_hash_caps, dim3(gridn(P.nrows, 256)), dim3(256), 0, ctx->stream, P);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

size_t pass_lds_bytes(int pass, bool dense, bool col64, uint64_t ncols, uint32_t lds_slots, bool flags,
                      uint32_t threads, int val) {
  size_t b = (size_t)2 * threads * 16 + threads;  // staging: 2 positions per thread, 64 flags per wave
  if (dense) {
    if (pass == PASS_SYMBOLIC) b += ((ncols + 31) / 32) * 4;
    else b += ncols * ((val == 1 || val == 2) ? 12 : 8) + (flags ? ((ncols + 31) / 32) * 4 : 0);  // accumulator (+ touched bits)
  } else {
    b += (size_t)lds_slots * ((col64 ? 8 : 4) + (pass == PASS_NUMERIC ? 8 : 0));
  }
  return (b + 15) & ~size_t(15);
}

template <int PASS, bool DENSE, class ColT, bool MERGE, int VAL = 0>
static int launch_pass_t(ptc_ctx* ctx, const PassParams& P, uint64_t nitems, size_t lds, uint32_t threads) {
  auto kern = k_pass<PASS, DENSE, ColT, MERGE, VAL>;
  static size_t lds_set = 0;  // largest dynamic LDS this kernel was opened to
  if (lds > lds_set) {
    if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
      return set_err(ctx, PTC_ERR_HIP, "opening %zu bytes of LDS for the row pass (%u threads): %s", lds, threads,
                     hipGetErrorString(hipGetLastError()));
    lds_set = lds;
  }
  hipLaunchKernelGGL(kern, dim3((unsigned)nitems), dim3(threads), lds, ctx->stream, P);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

template <bool MERGE, int VAL>
static int launch_numeric_v(ptc_ctx* ctx, bool dense, bool col64, const PassParams& P, uint64_t nitems, size_t lds,
                            uint32_t threads) {
  if (dense) return col64 ? launch_pass_t<PASS_NUMERIC, true, unsigned long long, MERGE, VAL>(ctx, P, nitems, lds, threads)
                          : launch_pass_t<PASS_NUMERIC, true, uint32_t, MERGE, VAL>(ctx, P, nitems, lds, threads);
  return col64 ? launch_pass_t<PASS_NUMERIC, false, unsigned long long, MERGE, VAL>(ctx, P, nitems, lds, threads)
               : launch_pass_t<PASS_NUMERIC, false, uint32_t, MERGE, VAL>(ctx, P, nitems, lds, threads);
}

template <bool MERGE>
static int launch_pass_m(ptc_ctx* ctx, int pass, bool dense, bool col64, const PassParams& P, uint64_t nitems,
                         size_t lds, uint32_t threads) {
  if (pass == PASS_SYMBOLIC) {
    if (dense) return col64 ? launch_pass_t<PASS_SYMBOLIC, true, unsigned long long, MERGE>(ctx, P, nitems, lds, threads)
                            : launch_pass_t<PASS_SYMBOLIC, true, uint32_t, MERGE>(ctx, P, nitems, lds, threads);
    return col64 ? launch_pass_t<PASS_SYMBOLIC, false, unsigned long long, MERGE>(ctx, P, nitems, lds, threads)
                 : launch_pass_t<PASS_SYMBOLIC, false, uint32_t, MERGE>(ctx, P, nitems, lds, threads);
  }
  switch (P.val) {
    case 1: return launch_numeric_v<MERGE, 1>(ctx, dense, col64, P, nitems, lds, threads);
    case 2: return launch_numeric_v<MERGE, 2>(ctx, dense, col64, P, nitems, lds, threads);
    case 3: return launch_numeric_v<MERGE, 3>(ctx, dense, col64, P, nitems, lds, threads);
    default: break;
  }
  if (dense) return col64 ? launch_pass_t<PASS_NUMERIC, true, unsigned long long, MERGE>(ctx, P, nitems, lds, threads)
                          : launch_pass_t<PASS_NUMERIC, true, uint32_t, MERGE>(ctx, P, nitems, lds, threads);
  return col64 ? launch_pass_t<PASS_NUMERIC, false, unsigned long long, MERGE>(ctx, P, nitems, lds, threads)
               : launch_pass_t<PASS_NUMERIC, false, uint32_t, MERGE>(ctx, P, nitems, lds, threads);
}

int launch_pass(ptc_ctx* ctx, int pass, bool dense, bool col64, const PassParams& P, uint64_t nitems,
                size_t lds, bool merge, uint32_t threads) {
  if (!nitems) return PTC_OK;
  if (nitems > 0x7FFFFFFFull) return set_err(ctx, PTC_ERR_RANGE, "too many work items");
  if (threads < 64 || threads > PB_MAX || threads % 64) return set_err(ctx, PTC_ERR_ARG, "row pass threads %u", threads);
  return merge ? launch_pass_m<true>(ctx, pass, dense, col64, P, nitems, lds, threads)
               : launch_pass_m<false>(ctx, pass, dense, col64, P, nitems, lds, threads);
}

// head == nullptr: only open the LDS attribute and report the grid.
template <int PASS, bool DENSE, class ColT, bool MERGE, int VAL>
static int launch_batch_t(ptc_ctx* ctx, const PassParams* jobs, const uint2* queue, uint32_t nq, uint32_t* head,
                          size_t lds, unsigned grid, unsigned* grid_out, uint32_t threads) {
  auto kern = k_pass_batch<PASS, DENSE, ColT, MERGE, VAL>;
  static size_t lds_set = 0;
  if (lds > lds_set) {
    if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
      return set_err(ctx, PTC_ERR_HIP, "opening %zu bytes of LDS for the batched row pass (%u threads): %s", lds, threads,
                     hipGetErrorString(hipGetLastError()));
    lds_set = lds;
  }
  if (!grid) {
    int per_cu = 0;
    PTC_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, lds));
    grid = (unsigned)std::max(1, per_cu) * (unsigned)std::max(1, ctx->prop.multiProcessorCount);
  }
  if (grid_out) *grid_out = grid;
  if (!head) return PTC_OK;
  grid = std::min<unsigned>(grid, nq);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), lds, ctx->stream, jobs, queue, nq, head);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

template <bool MERGE, int VAL>
static int launch_batch_v(ptc_ctx* ctx, int pass, bool dense, bool col64, const PassParams* jobs, const uint2* queue,
                          uint32_t nq, uint32_t* head, size_t lds, unsigned grid, unsigned* gout, uint32_t threads) {
#define PTC_LB(P_, D_, C_, V_) launch_batch_t<P_, D_, C_, MERGE, V_>(ctx, jobs, queue, nq, head, lds, grid, gout, threads)
  if (pass == PASS_SYMBOLIC) {
    if (dense) return col64 ? PTC_LB(PASS_SYMBOLIC, true, unsigned long long, 0) : PTC_LB(PASS_SYMBOLIC, true, uint32_t, 0);
    return col64 ? PTC_LB(PASS_SYMBOLIC, false, unsigned long long, 0) : PTC_LB(PASS_SYMBOLIC, false, uint32_t, 0);
  }
  if (dense) return col64 ? PTC_LB(PASS_NUMERIC, true, unsigned long long, VAL) : PTC_LB(PASS_NUMERIC, true, uint32_t, VAL);
  return col64 ? PTC_LB(PASS_NUMERIC, false, unsigned long long, VAL) : PTC_LB(PASS_NUMERIC, false, uint32_t, VAL);
#undef PTC_LB
}

int launch_pass_batch(ptc_ctx* ctx, int pass, bool dense, bool col64, int val, bool merge, const PassParams* jobs,
                      const uint2* queue, uint32_t nq, uint32_t* head, size_t lds, unsigned grid, uint32_t threads) {
  if (!nq) return PTC_OK;
  if (pass == PASS_SYMBOLIC) val = 0;
  if (val != 0) return set_err(ctx, PTC_ERR_UNSUPPORTED, "batched contractions compute in FP64");
  return merge ? launch_batch_v<true, 0>(ctx, pass, dense, col64, jobs, queue, nq, head, lds, grid, nullptr, threads)
               : launch_batch_v<false, 0>(ctx, pass, dense, col64, jobs, queue, nq, head, lds, grid, nullptr, threads);
}

int pass_batch_grid(ptc_ctx* ctx, int pass, bool dense, bool col64, int val, bool merge, size_t lds, unsigned* grid,
                    uint32_t threads) {
  if (pass == PASS_SYMBOLIC) val = 0;
  if (val != 0) return set_err(ctx, PTC_ERR_UNSUPPORTED, "batched contractions compute in FP64");
  return merge ? launch_batch_v<true, 0>(ctx, pass, dense, col64, nullptr, nullptr, 0, nullptr, lds, 0, grid, threads)
               : launch_batch_v<false, 0>(ctx, pass, dense, col64, nullptr, nullptr, 0, nullptr, lds, 0, grid, threads);
}

int launch_emit(ptc_ctx* ctx, const EmitParams& P, bool col64) {
  if (!P.nrows) return PTC_OK;
  if (P.f32) {
    if (col64)
      hipLaunchKernelGGL((k_emit<unsigned long long, true>), dim3((unsigned)P.nrows), dim3(PB), 0, ctx->stream, P);
    else
      hipLaunchKernelGGL((k_emit<uint32_t, true>), dim3((unsigned)P.nrows), dim3(PB), 0, ctx->stream, P);
  } else if (col64) {
    hipLaunchKernelGGL((k_emit<unsigned long long, false>), dim3((unsigned)P.nrows), dim3(PB), 0, ctx->stream, P);
  } else {
    hipLaunchKernelGGL((k_emit<uint32_t, false>), dim3((unsigned)P.nrows), dim3(PB), 0, ctx->stream, P);
  }
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

__global__ void k_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}
int fill_u64(ptc_ctx* ctx, unsigned long long* p, uint64_t n, unsigned long long v) {
  if (!n) return PTC_OK;
  hipLaunchKernelGGL(k_fill_u64, dim3(gridn(n, 256, 4096)), dim3(256), 0, ctx->stream, p, n, v);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}
int fill_u32(ptc_ctx* ctx, uint32_t* p, uint64_t n, uint32_t v) {
  if (!n) return PTC_OK;
  hipLaunchKernelGGL(k_fill_u32, dim3(gridn(n, 256, 4096)), dim3(256), 0, ctx->stream, p, n, v);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

// capacity of every row for the single-pass methods:
// upper bound min(flops, ncols); learned: clamp(scale * prediction, maxlen, bound)
__global__ void k_caps(CapParams P) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < P.nrows; r += (uint64_t)gridDim.x * blockDim.x) {
    unsigned long long f = P.row_flops[r];
    unsigned long long ub = f < P.ncols ? f : P.ncols;
    unsigned long long c = ub;
    if (P.pred) {
      double want = ceil((double)P.pred[r] * P.scale);
      unsigned long long lo = P.row_maxlen[r];
      unsigned long long w = want < 0 ? 0ull : (unsigned long long)min(want, 1.8e19);
      c = w < lo ? lo : w;
      if (c > ub) c = ub;
    }
    if (c > 0xFFFFFFFFull) c = 0xFFFFFFFFull;
    P.cap[r] = (uint32_t)c;
  }
}
int launch_caps(ptc_ctx* ctx, const CapParams& P) {
  if (!P.nrows) return PTC_OK;
  hipLaunchKernelGGL(k_caps, dim3(gridn(P.nrows, 256)), dim3(256), 0, ctx->stream, P);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

// rows that overflowed their predicted capacity -> a compact list and a
// per-row mark
__global__ void k_overflow_rows(const uint32_t* __restrict__ nnz, const uint32_t* __restrict__ cap,
                                uint64_t nrows, uint32_t* __restrict__ list, uint32_t* __restrict__ n,
                                uint32_t* __restrict__ mark) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += (uint64_t)gridDim.x * blockDim.x) {
    const bool o = nnz[r] > cap[r];
    mark[r] = o ? 1u : 0u;
    if (o) list[atomicAdd(n, 1u)] = (uint32_t)r;
  }
}
int launch_overflow_rows(ptc_ctx* ctx, const uint32_t* nnz, const uint32_t* cap, uint64_t nrows,
                         uint32_t* list, uint32_t* n, uint32_t* mark) {
  if (!nrows) return PTC_OK;
  hipLaunchKernelGGL(k_overflow_rows, dim3(gridn(nrows, 256)), dim3(256), 0, ctx->stream, nnz, cap, nrows, list, n, mark);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

// items of a subset of rows (fallback): items whose row is listed
__global__ void k_select_items(const uint4* __restrict__ items, uint64_t nitems, const uint32_t* __restrict__ nnz,
                               const uint32_t* __restrict__ cap, uint4* __restrict__ out, uint32_t* __restrict__ n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nitems; i += (uint64_t)gridDim.x * blockDim.x) {
    uint4 it = items[i];
    if (nnz[it.x] > cap[it.x]) out[atomicAdd(n, 1u)] = it;
  }
}
int launch_select_items(ptc_ctx* ctx, const uint4* items, uint64_t nitems, const uint32_t* nnz,
                        const uint32_t* cap, uint4* out, uint32_t* n) {
  if (!nitems) return PTC_OK;
  hipLaunchKernelGGL(k_select_items, dim3(gridn(nitems, 256)), dim3(256), 0, ctx->stream, items, nitems, nnz, cap, out, n);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

// reset per-row merge state of selected rows
__global__ void k_reset_rows(const uint32_t* __restrict__ list, uint32_t n, uint32_t* __restrict__ counts) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) counts[list[i]] = 0;
}
int launch_reset_rows(ptc_ctx* ctx, const uint32_t* list, uint32_t n, uint32_t* counts) {
  if (!n) return PTC_OK;
  hipLaunchKernelGGL(k_reset_rows, dim3(gridn(n, 256)), dim3(256), 0, ctx->stream, list, n, counts);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

// A small row is one work item (a row split over several items needs the
// dense share's slabs and merge).  One fetch-add per wave and list (the
// wave's items ranked by ballot): fetch-adds on one word serialise, tens of
// nanoseconds each on this GPU.
__global__ void k_split_items(const uint4* __restrict__ items, uint64_t nitems, const unsigned long long* __restrict__ row_flops,
                              const uint32_t* __restrict__ row_nchunk, uint4* __restrict__ dense, uint4* __restrict__ sparse,
                              uint32_t* __restrict__ cnt) {
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nitems; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint4 it = items[i];
    const bool sp = row_flops[it.x] <= SPARSE_FLOPS && row_nchunk[it.x] == 1;
    const unsigned long long ms = __ballot(sp), md = __ballot(!sp), act = ms | md;
    const int leader = __ffsll((long long)act) - 1;  // the lowest active lane
    uint32_t bs = 0, bd = 0;
    if (lane == leader) {
      if (ms) bs = atomicAdd(&cnt[1], (uint32_t)__popcll(ms));
      if (md) bd = atomicAdd(&cnt[0], (uint32_t)__popcll(md));
    }
    bs = __shfl(bs, leader, 64);
    bd = __shfl(bd, leader, 64);
    if (sp) sparse[bs + __popcll(ms & below)] = it;
    else dense[bd + __popcll(md & below)] = it;
  }
}
int launch_split_items(ptc_ctx* ctx, const uint4* items, uint64_t nitems, const unsigned long long* row_flops,
                       const uint32_t* row_nchunk, uint4* dense, uint4* sparse, uint32_t* cnt) {
  if (!nitems) return PTC_OK;
  hipLaunchKernelGGL(k_split_items, dim3(gridn(nitems, 256)), dim3(256), 0, ctx->stream, items, nitems, row_flops,
                     row_nchunk, dense, sparse, cnt);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

__global__ void k_sparse_caps(const unsigned long long* __restrict__ row_flops, uint64_t nrows, uint32_t* __restrict__ hcap) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long f = row_flops[r];
    uint32_t c = 16;
    while (c < 2 * f && c < SPARSE_SLOTS) c <<= 1;
    hcap[r] = c;
  }
}
int launch_sparse_caps(ptc_ctx* ctx, const unsigned long long* row_flops, uint64_t nrows, uint32_t* hcap) {
  if (!nrows) return PTC_OK;
  hipLaunchKernelGGL(k_sparse_caps, dim3(gridn(nrows, 256)), dim3(256), 0, ctx->stream, row_flops, nrows, hcap);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

__global__ void k_window_offsets(const unsigned long long* __restrict__ out_off, const uint32_t* __restrict__ cnt,
                                 uint64_t nrows, uint32_t nwin, unsigned long long* __restrict__ off) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += (uint64_t)gridDim.x * blockDim.x) {
    unsigned long long o = out_off[r];
    for (uint32_t t = 0; t < nwin; ++t) {
      off[(uint64_t)t * nrows + r] = o;
      o += cnt[(uint64_t)t * nrows + r];
    }
  }
}
int launch_window_offsets(ptc_ctx* ctx, const unsigned long long* out_off, const uint32_t* cnt, uint64_t nrows,
                          uint32_t nwin, unsigned long long* off) {
  if (!nrows || !nwin) return PTC_OK;
  hipLaunchKernelGGL(k_window_offsets, dim3(gridn(nrows, 256)), dim3(256), 0, ctx->stream, out_off, cnt, nrows, nwin, off);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

}  // namespace ptc
