// Runtime of the engine: contexts, the caching device allocator, device
// tensors, the .meta parser and the device prefix scan.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>

#include "ptc_internal.h"

namespace ptc {

int set_err(ptc_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return code;
}

int hip_check(ptc_ctx* ctx, hipError_t e, const char* what) {
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return set_err(ctx, PTC_ERR_NOMEM, "%s: %s", what, hipGetErrorString(e));
  return set_err(ctx, PTC_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

static size_t size_class(size_t bytes) {
  // 256-byte granularity below 1 MiB, then powers of two / 8 steps
  if (bytes <= 256) return 256;
  if (bytes <= (1u << 20)) return (bytes + 255) & ~size_t(255);
  size_t p = size_t(1) << (63 - __builtin_clzll(bytes));
  size_t step = p / 8;
  return (bytes + step - 1) / step * step;
}

void* dalloc(ptc_ctx* ctx, size_t bytes) {
  if (ctx->arena) {
    // allocation-free mode: a block of the call's scratch arena
    // blocks keep the allocator's size classes, so a kernel sees the same
    // padding behind an array in both modes
    const size_t off = (ctx->arena_used + 255) & ~size_t(255);
    const size_t need = size_class(bytes);
    if (off > ctx->arena_cap || need > ctx->arena_cap - off) {
      ctx->arena_overflow = true;
      set_err(ctx, PTC_ERR_RANGE, "scratch of %zu bytes exceeded", ctx->arena_cap);
      return nullptr;
    }
    ctx->arena_used = off + need;
    return ctx->arena + off;
  }
  size_t sz = size_class(bytes);
  auto it = ctx->free_blocks.lower_bound(sz);
  // reuse a cached block of the same class or at most 2x larger
  if (it != ctx->free_blocks.end() && it->first <= 2 * sz) {
    void* p = it->second;
    size_t have = it->first;
    ctx->free_blocks.erase(it);
    ctx->cached_bytes -= have;
    ctx->live[p] = have;
    ctx->live_bytes += have;
    ctx->peak_bytes = std::max(ctx->peak_bytes, ctx->live_bytes);
    return p;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, sz);
  if (e != hipSuccess) {
    // release the cache and retry once
    (void)hipGetLastError();
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->free_blocks) (void)hipFree(kv.second);
    ctx->free_blocks.clear();
    ctx->cached_bytes = 0;
    e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      set_err(ctx, PTC_ERR_NOMEM, "device allocation of %zu bytes failed", sz);
      return nullptr;
    }
  }
  ctx->live[p] = sz;
  ctx->live_bytes += sz;
  ctx->peak_bytes = std::max(ctx->peak_bytes, ctx->live_bytes);
  return p;
}

void ctx_acquire(ptc_ctx* ctx) {
  if (ctx) ++ctx->refs;
}

void ctx_release(ptc_ctx* ctx) {
  if (!ctx) return;
  if (ctx->refs) --ctx->refs;
  if (ctx->closed && ctx->refs == 0) delete ctx;
}

void dfree(ptc_ctx* ctx, void* p) {
  // a closed context has already released every block
  if (!p || !ctx_open(ctx)) return;
  auto it = ctx->live.find(p);
  if (it == ctx->live.end()) return;
  size_t sz = it->second;
  ctx->live.erase(it);
  ctx->live_bytes -= sz;
  // Blocks are only reused by work queued later on the same stream, so no
  // synchronisation is needed before caching them.
  ctx->free_blocks.emplace(sz, p);
  ctx->cached_bytes += sz;
}

int read_back(ptc_ctx* ctx, void* host, const void* dev, size_t bytes) {
  if (bytes > ctx->host_pinned_bytes) {
    if (ctx->arena) return set_err(ctx, PTC_ERR_INTERNAL, "read-back of %zu bytes in allocation-free mode", bytes);
    if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
    ctx->host_pinned = nullptr;
    size_t want = std::max<size_t>(bytes, 4096);
    PTC_HIP(ctx, hipHostMalloc(&ctx->host_pinned, want, hipHostMallocDefault));
    ctx->host_pinned_bytes = want;
  }
  PTC_HIP(ctx, hipMemcpyAsync(ctx->host_pinned, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
  PTC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  memcpy(host, ctx->host_pinned, bytes);
  return PTC_OK;
}

// ------------------------------------------------------------------ fill

__global__ void k_fill_zero(uint4* __restrict__ p, size_t n16, unsigned char* __restrict__ tail,
                            size_t ntail) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = i; k < n16; k += stride) p[k] = make_uint4(0, 0, 0, 0);
  if (i < ntail) tail[i] = 0;
}

// Clears device memory with a kernel (not hipMemsetAsync, so the work stays
// valid inside captured graphs).
int fill_zero(ptc_ctx* ctx, void* p, size_t bytes) {
  if (!bytes) return PTC_OK;
  uintptr_t a = (uintptr_t)p;
  if (a & 15) return set_err(ctx, PTC_ERR_INTERNAL, "fill_zero: unaligned pointer");
  size_t n16 = bytes / 16, ntail = bytes % 16;
  unsigned grid = (unsigned)std::min<size_t>((n16 + 255) / 256 + 1, 4096);
  hipLaunchKernelGGL(k_fill_zero, dim3(grid), dim3(256), 0, ctx->stream, (uint4*)p, n16,
                     (unsigned char*)p + n16 * 16, ntail);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

// ------------------------------------------------------------- collect
//
// Up to 8 device words (4 or 8 bytes each, zero-extended) copied into one
// array with a single launch, for a combined read-back.

__global__ void k_collect_words(WordSources W, unsigned long long* __restrict__ dst) {
  const uint32_t i = threadIdx.x;
  if (i < W.n)
    dst[i] = W.bytes[i] == 4 ? (unsigned long long)*(const uint32_t*)W.src[i] : *(const unsigned long long*)W.src[i];
}

int collect_words(ptc_ctx* ctx, unsigned long long* dst, const WordSources& W) {
  hipLaunchKernelGGL(k_collect_words, dim3(1), dim3(64), 0, ctx->stream, W, dst);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

// ------------------------------------------------------------------ scan
//
// Three-phase exclusive scan: per-block totals, a single-block scan of the
// totals (recursively for very long inputs), then per-block scans offset by
// their block's prefix.  Each block covers SCAN_TILE elements.

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

template <class In>
__device__ inline unsigned long long load_elem(const In* in, uint64_t i, uint64_t n) {
  return i < n ? (unsigned long long)in[i] : 0ull;
}

__device__ inline unsigned long long block_exclusive_scan(unsigned long long v,
                                                          unsigned long long* total) {
  __shared__ unsigned long long wsum[SCAN_THREADS / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned long long x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    unsigned long long y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  unsigned long long base = 0, all = 0;
#pragma unroll
  for (int w = 0; w < SCAN_THREADS / 64; ++w) {
    if (w < wid) base += wsum[w];
    all += wsum[w];
  }
  __syncthreads();
  *total = all;
  return base + x - v;
}

template <class In>
__global__ void k_scan_reduce(const In* __restrict__ in, uint64_t n,
                              unsigned long long* __restrict__ block_sums) {
  uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  unsigned long long s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k)
    s += load_elem(in, base + (uint64_t)k * SCAN_THREADS + threadIdx.x, n);
  unsigned long long tot;
  block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// The tile goes through LDS both ways, so global loads and stores are
// coalesced (element k * SCAN_THREADS + thread) while each thread scans
// SCAN_ITEMS consecutive elements; one pad word per SCAN_ITEMS keeps a
// thread's consecutive words in distinct banks across the wave.
__device__ __forceinline__ uint32_t scan_slot(uint32_t i) { return i + i / SCAN_ITEMS; }

template <class In>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(const In* __restrict__ in,
                                                             unsigned long long* __restrict__ out, uint64_t n,
                                                             const unsigned long long* __restrict__ block_prefix,
                                                             unsigned long long* __restrict__ total_out) {
  __shared__ unsigned long long s_tile[SCAN_TILE + SCAN_TILE / SCAN_ITEMS];
  const uint64_t b0 = (uint64_t)blockIdx.x * SCAN_TILE;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const uint32_t i = k * SCAN_THREADS + threadIdx.x;
    s_tile[scan_slot(i)] = load_elem(in, b0 + i, n);
  }
  __syncthreads();
  unsigned long long v[SCAN_ITEMS];
  unsigned long long s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = s_tile[scan_slot(threadIdx.x * SCAN_ITEMS + k)];
    s += v[k];
  }
  unsigned long long tot;
  unsigned long long run = block_exclusive_scan(s, &tot) + (block_prefix ? block_prefix[blockIdx.x] : 0ull);
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    s_tile[scan_slot(threadIdx.x * SCAN_ITEMS + k)] = run;
    run += v[k];
  }
  if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_THREADS - 1) *total_out = run;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const uint32_t i = k * SCAN_THREADS + threadIdx.x;
    if (b0 + i < n) out[b0 + i] = s_tile[scan_slot(i)];
  }
}

// Inputs of up to SCAN_SMALL elements: one workgroup, one launch.
constexpr int SCAN_SMALL_THREADS = 1024;
constexpr int SCAN_SMALL_ITEMS = 16;
constexpr uint64_t SCAN_SMALL = (uint64_t)SCAN_SMALL_THREADS * SCAN_SMALL_ITEMS;

template <class In>
__global__ void __launch_bounds__(SCAN_SMALL_THREADS) k_scan_small(const In* __restrict__ in,
                                                                   unsigned long long* __restrict__ out, uint64_t n,
                                                                   unsigned long long* __restrict__ total_out) {
  __shared__ unsigned long long wsum[SCAN_SMALL_THREADS / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint64_t base = (uint64_t)threadIdx.x * SCAN_SMALL_ITEMS;
  unsigned long long v[SCAN_SMALL_ITEMS], s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_SMALL_ITEMS; ++k) {
    v[k] = load_elem(in, base + k, n);
    s += v[k];
  }
  unsigned long long x = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned long long y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  unsigned long long run = x - s;
  for (int w = 0; w < wid; ++w) run += wsum[w];
#pragma unroll
  for (int k = 0; k < SCAN_SMALL_ITEMS; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (total_out && threadIdx.x == SCAN_SMALL_THREADS - 1) *total_out = run;
}

// Two-launch scan for long inputs: the per-tile totals (k_scan_reduce), then
// every tile sums the earlier tiles' totals itself (a few thousand words,
// read from L2) and scans its own elements.  No atomics and no waiting on
// other workgroups: on this GPU a chain of same-address device atomics (a
// ticket counter, a look-back flag) costs tens of nanoseconds per step.
//
// With features (the A positions' scan of the learned method) every row
// (a contiguous, non-empty range of elements) gets its count of elements
// with a nonzero value and its largest element, written once, by the tile
// in which the row ends, with no global atomics: the first launch also
// leaves each tile's FeatPart (its first and last rows and their partial
// sums in the tile); the second reduces its own rows with LDS atomics,
// writes the rows that end inside it, and for a row that entered from
// earlier tiles adds those tiles' partials, found by a parallel walk back
// over the parts (a row is the only row of every tile strictly inside it).

// per tile: the element total and the FeatPart
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce_feat(const uint32_t* __restrict__ in, uint64_t n,
                                                                   unsigned long long* __restrict__ block_sums,
                                                                   ScanFeatures F) {
  static_assert(SCAN_TILE == FEAT_TILE, "row features are kept per scan tile");
  __shared__ unsigned long long s_w[SCAN_THREADS / 64][5];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint64_t b0 = (uint64_t)blockIdx.x * SCAN_TILE;
  const uint64_t last = min<uint64_t>(n, b0 + SCAN_TILE) - 1;
  const uint32_t fr = F.row_of_pos[b0], lr = F.row_of_pos[last];
  unsigned long long sum = 0;
  uint32_t fh = 0, fm = 0, lh = 0, lm = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const uint64_t e = b0 + (uint64_t)k * SCAN_THREADS + threadIdx.x;
    if (e > last) continue;
    const uint32_t l = in[e], r = F.row_of_pos[e];
    sum += l;
    if (r == fr) {
      fh += l ? 1u : 0u;
      fm = max(fm, l);
    }
    if (r == lr) {
      lh += l ? 1u : 0u;
      lm = max(lm, l);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    sum += __shfl_xor(sum, o, 64);
    fh += __shfl_xor(fh, o, 64);
    fm = max(fm, (uint32_t)__shfl_xor(fm, o, 64));
    lh += __shfl_xor(lh, o, 64);
    lm = max(lm, (uint32_t)__shfl_xor(lm, o, 64));
  }
  if (lane == 0) {
    s_w[wid][0] = sum;
    s_w[wid][1] = fh;
    s_w[wid][2] = fm;
    s_w[wid][3] = lh;
    s_w[wid][4] = lm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    FeatPart P;
    P.fr = fr;
    P.lr = lr;
    P.fh = P.fm = P.lh = P.lm = 0;
    sum = 0;
    for (int w = 0; w < SCAN_THREADS / 64; ++w) {
      sum += s_w[w][0];
      P.fh += (uint32_t)s_w[w][1];
      P.fm = max(P.fm, (uint32_t)s_w[w][2]);
      P.lh += (uint32_t)s_w[w][3];
      P.lm = max(P.lm, (uint32_t)s_w[w][4]);
    }
    block_sums[blockIdx.x] = sum;
    F.parts[blockIdx.x] = P;
  }
}

template <class In, bool FEAT>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply_p(const In* __restrict__ in,
                                                               unsigned long long* __restrict__ out, uint64_t n,
                                                               const unsigned long long* __restrict__ block_sums,
                                                               unsigned long long* __restrict__ total_out,
                                                               ScanFeatures F) {
  __shared__ unsigned long long s_tile[SCAN_TILE + SCAN_TILE / SCAN_ITEMS];
  __shared__ uint32_t s_row[FEAT ? SCAN_TILE + SCAN_TILE / SCAN_ITEMS : 1];
  // once the tile's results are stored, its LDS holds the per-row sums
  // (at most one row per element): fewer bytes a workgroup, more of them
  // per CU
  static_assert(sizeof(s_tile) >= 2 * SCAN_TILE * sizeof(uint32_t), "row sums fit the tile's LDS");
  uint32_t* s_h = reinterpret_cast<uint32_t*>(s_tile);
  uint32_t* s_m = s_h + SCAN_TILE;
  __shared__ unsigned long long s_red[SCAN_THREADS / 64];
  __shared__ uint32_t s_stop, s_ch[SCAN_THREADS / 64], s_cm[SCAN_THREADS / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint64_t b0 = (uint64_t)blockIdx.x * SCAN_TILE;
  // this tile's prefix: the earlier tiles' totals
  unsigned long long pre = 0;
  for (uint32_t j = threadIdx.x; j < blockIdx.x; j += SCAN_THREADS) pre += block_sums[j];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) pre += __shfl_xor(pre, o, 64);
  if (lane == 0) s_red[wid] = pre;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const uint32_t i = k * SCAN_THREADS + threadIdx.x;
    s_tile[scan_slot(i)] = load_elem(in, b0 + i, n);
    if (FEAT) s_row[scan_slot(i)] = b0 + i < n ? F.row_of_pos[b0 + i] : 0xFFFFFFFFu;
  }
  __syncthreads();
  pre = 0;
#pragma unroll
  for (int w = 0; w < SCAN_THREADS / 64; ++w) pre += s_red[w];
  unsigned long long v[SCAN_ITEMS];
  unsigned long long sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = s_tile[scan_slot(threadIdx.x * SCAN_ITEMS + k)];
    sum += v[k];
  }
  unsigned long long tot;
  unsigned long long run = block_exclusive_scan(sum, &tot) + pre;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    s_tile[scan_slot(threadIdx.x * SCAN_ITEMS + k)] = run;
    run += v[k];
  }
  if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_THREADS - 1) *total_out = run;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const uint32_t i = k * SCAN_THREADS + threadIdx.x;
    if (b0 + i < n) out[b0 + i] = s_tile[scan_slot(i)];
  }
  if (!FEAT) return;
  uint32_t rows[SCAN_ITEMS];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) rows[k] = s_row[scan_slot(threadIdx.x * SCAN_ITEMS + k)];
  __syncthreads();  // every result is stored: the tile's LDS becomes the row sums
  for (uint32_t i = threadIdx.x; i < 2 * SCAN_TILE; i += SCAN_THREADS) s_h[i] = 0;
  __syncthreads();
  uint32_t fr = 0;
  {
    // each thread's run of consecutive elements, into the tile's per-row slots
    fr = s_row[scan_slot(0)];
    uint32_t row = 0xFFFFFFFFu, h = 0, m = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
      const uint32_t r = rows[k];
      if (r == 0xFFFFFFFFu) break;  // past n
      if (r != row) {
        if (row != 0xFFFFFFFFu) {
          if (h) atomicAdd(&s_h[row - fr], h);
          if (m) atomicMax(&s_m[row - fr], m);
        }
        row = r;
        h = 0;
        m = 0;
      }
      const uint32_t l = (uint32_t)v[k];
      h += l ? 1u : 0u;
      m = max(m, l);
    }
    if (row != 0xFFFFFFFFu) {
      if (h) atomicAdd(&s_h[row - fr], h);
      if (m) atomicMax(&s_m[row - fr], m);
    }
  }
  __syncthreads();
  {
    const uint32_t t = blockIdx.x, nt = gridDim.x;
    const uint32_t lr = s_row[scan_slot((uint32_t)min<uint64_t>(SCAN_TILE, n - b0) - 1)];
    for (uint32_t i = threadIdx.x + 1; i < lr - fr; i += SCAN_THREADS) {  // rows strictly inside the tile
      F.row_hits[fr + i] = s_h[i];
      F.row_maxlen[fr + i] = s_m[i];
    }
    // a row ends in this tile unless the next tile starts with it
    const bool next_same_f = t + 1 < nt && F.parts[t + 1].fr == fr;
    const bool next_same_l = t + 1 < nt && F.parts[t + 1].fr == lr;
    if (lr != fr && threadIdx.x == 0 && !next_same_l) {
      F.row_hits[lr] = s_h[lr - fr];
      F.row_maxlen[lr] = s_m[lr - fr];
    }
    const bool fr_ends = lr != fr || !next_same_f;
    if (fr_ends) {
      // the first row's sums from earlier tiles: every tile it fills whole,
      // then the tile where it starts (its last row there)
      uint32_t ch = 0, cm = 0;
      if (t > 0 && F.parts[t - 1].lr == fr) {
        for (int64_t base = (int64_t)t - 1;; base -= SCAN_THREADS) {
          if (threadIdx.x == 0) s_stop = SCAN_THREADS;
          __syncthreads();
          const int64_t tp = base - (int64_t)threadIdx.x;
          FeatPart P{};
          bool edge = tp < 0;
          if (!edge) {
            P = F.parts[tp];
            edge = P.fr != fr;  // the tile where the row starts
          }
          if (edge) atomicMin(&s_stop, (uint32_t)threadIdx.x);
          __syncthreads();
          const uint32_t stop = s_stop;
          uint32_t h = 0, m = 0;
          if (threadIdx.x < stop) {
            h = P.fh;  // a tile the row fills whole
            m = P.fm;
          } else if (threadIdx.x == stop && tp >= 0) {
            h = P.lh;
            m = P.lm;
          }
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) {
            h += __shfl_xor(h, o, 64);
            m = max(m, (uint32_t)__shfl_xor(m, o, 64));
          }
          if (lane == 0) {
            s_ch[wid] = h;
            s_cm[wid] = m;
          }
          __syncthreads();
#pragma unroll
          for (int w = 0; w < SCAN_THREADS / 64; ++w) {
            ch += s_ch[w];
            cm = max(cm, s_cm[w]);
          }
          if (stop < SCAN_THREADS) break;  // the same value in every thread
          __syncthreads();
        }
      }
      if (threadIdx.x == 0) {
        F.row_hits[fr] = ch + s_h[0];
        F.row_maxlen[fr] = max(cm, s_m[0]);
      }
    }
  }
}

template <class In, bool FEAT>
static int scan_two(ptc_ctx* ctx, const In* in, unsigned long long* out, uint64_t n, unsigned long long* total,
                    const ScanFeatures& F) {
  const uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  Workspace ws(ctx);
  unsigned long long* sums = ws.get<unsigned long long>(nb);
  if (!sums) return PTC_ERR_NOMEM;
  if constexpr (FEAT)
    hipLaunchKernelGGL(k_scan_reduce_feat, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, ctx->stream, in, n, sums, F);
  else
    hipLaunchKernelGGL((k_scan_reduce<In>), dim3((unsigned)nb), dim3(SCAN_THREADS), 0, ctx->stream, in, n, sums);
  hipLaunchKernelGGL((k_scan_apply_p<In, FEAT>), dim3((unsigned)nb), dim3(SCAN_THREADS), 0, ctx->stream, in, out, n,
                     (const unsigned long long*)sums, total, F);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}
// tiles whose prefix a workgroup sums itself (at most 64 words a thread)
constexpr uint64_t SCAN_TWO_MAX_TILES = 64ull * SCAN_THREADS;

// per tile: the element total; per half tile: the count of nonzero
// elements and the largest
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce_tiles(const uint32_t* __restrict__ in, uint64_t n,
                                                                    unsigned long long* __restrict__ block_sums,
                                                                    TilePart* __restrict__ tiles) {
  static_assert(SCAN_TILE == 2 * TILE_PART && SCAN_ITEMS % 2 == 0, "two tile parts a scan tile");
  __shared__ unsigned long long s_w[SCAN_THREADS / 64][5];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint64_t b0 = (uint64_t)blockIdx.x * SCAN_TILE;
  unsigned long long sum = 0;
  uint32_t h[2] = {0, 0}, m[2] = {0, 0};
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {  // element k * SCAN_THREADS + thread: the first half for k < SCAN_ITEMS / 2
    const uint64_t e = b0 + (uint64_t)k * SCAN_THREADS + threadIdx.x;
    const uint32_t l = e < n ? in[e] : 0u;
    const int half = k < SCAN_ITEMS / 2 ? 0 : 1;
    sum += l;
    h[half] += l ? 1u : 0u;
    m[half] = max(m[half], l);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    sum += __shfl_xor(sum, o, 64);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      h[j] += __shfl_xor(h[j], o, 64);
      m[j] = max(m[j], (uint32_t)__shfl_xor(m[j], o, 64));
    }
  }
  if (lane == 0) {
    s_w[wid][0] = sum;
    s_w[wid][1] = h[0];
    s_w[wid][2] = m[0];
    s_w[wid][3] = h[1];
    s_w[wid][4] = m[1];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    TilePart P0{0, 0}, P1{0, 0};
    sum = 0;
    for (int w = 0; w < SCAN_THREADS / 64; ++w) {
      sum += s_w[w][0];
      P0.h += (uint32_t)s_w[w][1];
      P0.m = max(P0.m, (uint32_t)s_w[w][2]);
      P1.h += (uint32_t)s_w[w][3];
      P1.m = max(P1.m, (uint32_t)s_w[w][4]);
    }
    block_sums[blockIdx.x] = sum;
    tiles[2 * blockIdx.x] = P0;
    tiles[2 * blockIdx.x + 1] = P1;
  }
}

template <class In>
static int scan_impl(ptc_ctx* ctx, const In* in, unsigned long long* out, uint64_t n,
                     unsigned long long* total);

int scan_u32_tiles(ptc_ctx* ctx, const uint32_t* in, unsigned long long* out, uint64_t n,
                   unsigned long long* total, TilePart* tiles) {
  if (n == 0) return total ? fill_zero(ctx, total, sizeof(unsigned long long)) : PTC_OK;
  const uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  Workspace ws(ctx);
  unsigned long long* sums = ws.get<unsigned long long>(nb);
  if (!sums) return PTC_ERR_NOMEM;
  hipLaunchKernelGGL(k_scan_reduce_tiles, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, ctx->stream, in, n, sums, tiles);
  if (nb <= SCAN_TWO_MAX_TILES) {
    hipLaunchKernelGGL((k_scan_apply_p<uint32_t, false>), dim3((unsigned)nb), dim3(SCAN_THREADS), 0, ctx->stream, in,
                       out, n, (const unsigned long long*)sums, total, ScanFeatures{});
  } else {
    PTC_HIP(ctx, hipGetLastError());
    const int rc = scan_impl<unsigned long long>(ctx, sums, sums, nb, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL((k_scan_apply<uint32_t>), dim3((unsigned)nb), dim3(SCAN_THREADS), 0, ctx->stream, in, out, n,
                       (const unsigned long long*)sums, total);
  }
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

int scan_u32_features(ptc_ctx* ctx, const uint32_t* in, unsigned long long* out, uint64_t n,
                      unsigned long long* total, const uint32_t* row_of_pos, uint32_t* row_hits,
                      uint32_t* row_maxlen, FeatPart* parts) {
  if (n == 0) return total ? fill_zero(ctx, total, sizeof(unsigned long long)) : PTC_OK;
  if ((n + SCAN_TILE - 1) / SCAN_TILE > SCAN_TWO_MAX_TILES)
    return set_err(ctx, PTC_ERR_RANGE, "feature scan of %llu elements", (unsigned long long)n);
  ScanFeatures F{row_of_pos, row_hits, row_maxlen, parts};
  return scan_two<uint32_t, true>(ctx, in, out, n, total, F);
}

// Two independent small scans in one launch: workgroup i scans array i.
__global__ void __launch_bounds__(SCAN_SMALL_THREADS) k_scan_small2(unsigned long long* __restrict__ a,
                                                                    unsigned long long* __restrict__ b, uint64_t n) {
  unsigned long long* x = blockIdx.x ? b : a;
  __shared__ unsigned long long wsum[SCAN_SMALL_THREADS / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint64_t base = (uint64_t)threadIdx.x * SCAN_SMALL_ITEMS;
  unsigned long long v[SCAN_SMALL_ITEMS], s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_SMALL_ITEMS; ++k) {
    v[k] = load_elem(x, base + k, n);
    s += v[k];
  }
  unsigned long long y = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned long long t = __shfl_up(y, o, 64);
    if (lane >= o) y += t;
  }
  if (lane == 63) wsum[wid] = y;
  __syncthreads();
  unsigned long long run = y - s;
  for (int w = 0; w < wid; ++w) run += wsum[w];
#pragma unroll
  for (int k = 0; k < SCAN_SMALL_ITEMS; ++k) {
    if (base + k < n) x[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == SCAN_SMALL_THREADS - 1) x[n] = run;  // the total after the last element
}

template <class In>
static int scan_impl(ptc_ctx* ctx, const In* in, unsigned long long* out, uint64_t n,
                     unsigned long long* total) {
  if (n == 0) {
    if (total) return fill_zero(ctx, total, sizeof(unsigned long long));
    return PTC_OK;
  }
  if (n <= SCAN_SMALL) {
    hipLaunchKernelGGL((k_scan_small<In>), dim3(1), dim3(SCAN_SMALL_THREADS), 0, ctx->stream, in, out, n, total);
    PTC_HIP(ctx, hipGetLastError());
    return PTC_OK;
  }
  if ((n + SCAN_TILE - 1) / SCAN_TILE <= SCAN_TWO_MAX_TILES) return scan_two<In, false>(ctx, in, out, n, total, ScanFeatures{});
  uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb == 1) {
    hipLaunchKernelGGL((k_scan_apply<In>), dim3(1), dim3(SCAN_THREADS), 0, ctx->stream, in, out, n,
                       (const unsigned long long*)nullptr, total);
    PTC_HIP(ctx, hipGetLastError());
    return PTC_OK;
  }
  Workspace ws(ctx);
  unsigned long long* sums = ws.get<unsigned long long>(nb);
  if (!sums) return PTC_ERR_NOMEM;
  hipLaunchKernelGGL((k_scan_reduce<In>), dim3((unsigned)nb), dim3(SCAN_THREADS), 0, ctx->stream,
                     in, n, sums);
  PTC_HIP(ctx, hipGetLastError());
  int rc = scan_impl<unsigned long long>(ctx, sums, sums, nb, nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL((k_scan_apply<In>), dim3((unsigned)nb), dim3(SCAN_THREADS), 0, ctx->stream,
                     in, out, n, (const unsigned long long*)sums, total);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

int scan_u32_to_u64(ptc_ctx* ctx, const uint32_t* in, unsigned long long* out, uint64_t n,
                    unsigned long long* total) {
  return scan_impl<uint32_t>(ctx, in, out, n, total);
}
int scan_u64(ptc_ctx* ctx, const unsigned long long* in, unsigned long long* out, uint64_t n,
             unsigned long long* total) {
  return scan_impl<unsigned long long>(ctx, in, out, n, total);
}
int scan_u64_two_inplace(ptc_ctx* ctx, unsigned long long* a, unsigned long long* b, uint64_t n) {
  if (n == 0 || n > SCAN_SMALL) {
    int rc = scan_u64(ctx, a, a, n, a + n);
    return rc ? rc : scan_u64(ctx, b, b, n, b + n);
  }
  hipLaunchKernelGGL(k_scan_small2, dim3(2), dim3(SCAN_SMALL_THREADS), 0, ctx->stream, a, b, n);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

// ------------------------------------------------------------- key codec

int make_codec(ptc_ctx* ctx, const ptc_dtensor* t, const uint32_t* modes, uint32_t n,
               KeyCodec* out) {
  KeyCodec c;
  c.n = n;
  unsigned __int128 space = 1;
  for (uint32_t j = 0; j < n; ++j) {
    uint32_t m = modes[j];
    if (m >= t->d) return set_err(ctx, PTC_ERR_ARG, "key mode %u out of range", m);
    c.mode[j] = m;
    c.radix[j] = (t->N ? t->mode_max[m] : 0u) + 1u;
    if (c.radix[j] == 0) c.radix[j] = 0xFFFFFFFFu;  // max index 2^32-1: radix saturates
    c.stride[j] = (uint64_t)space;
    space *= c.radix[j];
    if (space > (unsigned __int128)0xFFFFFFFFFFFFFFFEull)
      return set_err(ctx, PTC_ERR_UNSUPPORTED,
                     "key space over %u modes exceeds 64 bits (not supported)", n);
  }
  c.space = (uint64_t)space;
  *out = c;
  return PTC_OK;
}

// --------------------------------------------------------- mode maxima

__global__ void k_mode_max(const uint32_t* __restrict__ h, uint64_t N, uint32_t d,
                           unsigned int* __restrict__ mx) {
  __shared__ unsigned int smax[PTC_MAX_MODES];
  if (threadIdx.x < PTC_MAX_MODES) smax[threadIdx.x] = 0;
  __syncthreads();
  unsigned int local[PTC_MAX_MODES];
  for (uint32_t j = 0; j < d; ++j) local[j] = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (uint64_t)gridDim.x * blockDim.x)
    for (uint32_t j = 0; j < d; ++j) local[j] = max(local[j], h[i * d + j]);
  for (uint32_t j = 0; j < d; ++j) atomicMax(&smax[j], local[j]);
  __syncthreads();
  if (threadIdx.x < d) atomicMax(&mx[threadIdx.x], smax[threadIdx.x]);
}

}  // namespace ptc

using namespace ptc;

// ================================================================ C ABI

static int create_ctx(ptc_ctx** out, int device, hipStream_t caller_stream);

extern "C" int ptc_create(ptc_ctx** out, int device) { return create_ctx(out, device, nullptr); }

extern "C" int ptc_create_on_stream(ptc_ctx** out, int device, void* stream) {
  if (out) *out = nullptr;
  if (!stream) return PTC_ERR_ARG;
  return create_ctx(out, device, (hipStream_t)stream);
}

static int create_ctx(ptc_ctx** out, int device, hipStream_t caller_stream) {
  if (!out) return PTC_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return PTC_ERR_NODEV;
  }
  if (device < 0) {
    if (hipGetDevice(&device) != hipSuccess) device = 0;
  }
  if (device >= n) return PTC_ERR_ARG;
  ptc_ctx* ctx = new ptc_ctx();
  ctx->device = device;
  if (caller_stream) {
    ctx->stream = caller_stream;
    ctx->owns_stream = false;
    int sdev = -1;
    if (hipStreamGetDevice(caller_stream, &sdev) != hipSuccess || sdev != device) {
      (void)hipGetLastError();
      delete ctx;
      return PTC_ERR_ARG;
    }
  }
  if (hipSetDevice(device) != hipSuccess ||
      (!caller_stream && hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) ||
      hipGetDeviceProperties(&ctx->prop, device) != hipSuccess) {
    (void)hipGetLastError();
    if (ctx->owns_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PTC_ERR_HIP;
  }
  bool events_ok = true;
  for (auto& e : ctx->ev) events_ok = events_ok && hipEventCreate(&e) == hipSuccess;
  for (auto& e : ctx->kev) events_ok = events_ok && hipEventCreate(&e) == hipSuccess;
  if (!events_ok) {  // timing events are part of every call's stats
    (void)hipGetLastError();
    ptc_destroy(ctx);
    return PTC_ERR_HIP;
  }
  // pinned staging for read-backs, made now so that small reads never
  // allocate (the allocation-free entry point reads at most 32 bytes at once)
  ctx->host_pinned_bytes = 4096;
  if (hipHostMalloc(&ctx->host_pinned, ctx->host_pinned_bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    ctx->host_pinned = nullptr;
    ctx->host_pinned_bytes = 0;
    ptc_destroy(ctx);
    return PTC_ERR_NOMEM;
  }
  if (scratch_setup(ctx) != PTC_OK) {
    ptc_destroy(ctx);
    return PTC_ERR_HIP;
  }
  *out = ctx;
  return PTC_OK;
}

// Releases the context's device memory, events and stream at once.  Objects
// still holding a reference (tensors, tables, plans, models not freed yet)
// keep the host record alive; their later frees make no HIP call, and the
// last of them deletes the record (ctx_release).
extern "C" void ptc_destroy(ptc_ctx* ctx) {
  if (!ctx || ctx->closed) return;
  // teardown: nothing can be done about a failing release, so results are
  // not checked
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& kv : ctx->free_blocks) (void)hipFree(kv.second);
  for (auto& kv : ctx->live) (void)hipFree(kv.first);
  ctx->free_blocks.clear();
  ctx->live.clear();
  ctx->cached_bytes = 0;
  ctx->live_bytes = 0;
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->kev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
  ctx->host_pinned = nullptr;
  ctx->host_pinned_bytes = 0;
  if (ctx->stream && ctx->owns_stream) (void)hipStreamDestroy(ctx->stream);
  ctx->stream = nullptr;
  (void)hipGetLastError();
  ctx->closed = true;
  if (ctx->refs == 0) delete ctx;
}

extern "C" const char* ptc_last_error(const ptc_ctx* ctx) { return ctx ? ctx->err.c_str() : ""; }

extern "C" int ptc_synchronize(ptc_ctx* ctx) {
  if (!ctx) return PTC_ERR_ARG;
  PTC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PTC_OK;
}

extern "C" void* ptc_stream(ptc_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

extern "C" int ptc_set_timing(ptc_ctx* ctx, int on) {
  if (!ctx) return PTC_ERR_ARG;
  ctx->timing = on ? 1 : 0;
  return PTC_OK;
}

static int check_perm(ptc_ctx* ctx, uint32_t d, uint32_t nk, const uint32_t* dim_tc) {
  if (d == 0 || d > PTC_MAX_MODES) return set_err(ctx, PTC_ERR_ARG, "d=%u not in [1,%d]", d, PTC_MAX_MODES);
  if (nk > d) return set_err(ctx, PTC_ERR_ARG, "num_key_dimensions %u > d %u", nk, d);
  if (!dim_tc) return set_err(ctx, PTC_ERR_ARG, "dim_tc is NULL");
  uint32_t seen = 0;
  for (uint32_t j = 0; j < d; ++j) {
    if (dim_tc[j] >= d || (seen >> dim_tc[j]) & 1u)
      return set_err(ctx, PTC_ERR_ARG, "dim_tc is not a permutation of 0..%u", d - 1);
    seen |= 1u << dim_tc[j];
  }
  return PTC_OK;
}

// No product of two such values can be +-0, subnormal or infinite, so an
// untouched dense accumulator entry keeps its -0.0 initial value.
static bool values_safe(const double* v64, const float* v32, uint64_t n) {
  for (uint64_t i = 0; i < n; ++i) {
    const double a = v64 ? std::fabs(v64[i]) : std::fabs((double)v32[i]);
    const bool ok = v64 ? (a >= 0x1p-500 && a <= 0x1p500) : (a >= 0x1p-60 && a <= 0x1p60);
    if (!ok) return false;
  }
  return true;
}

// A coordinate of 2^32-1 in some mode: its key radix (max + 1) would not fit
// 32 bits.
static bool max_index_ok(const uint32_t* mode_max, uint32_t d) {
  for (uint32_t j = 0; j < d; ++j)
    if (mode_max[j] == 0xFFFFFFFFu) return false;
  return true;
}

// Upload of an FP64 (v64) or FP32 (v32) tensor.  vals_safe: every |value| is
// far enough from 0 and from overflow that no product is 0, subnormal or
// infinite (then untouched dense accumulator entries can be recognised by
// their -0.0 initial value).
static int upload(ptc_ctx* ctx, const ptc_tensor_view* v, const double* v64, const float* v32, ptc_dtensor** out) {
  if (!ctx || !v || !out) return PTC_ERR_ARG;
  *out = nullptr;
  int rc = check_perm(ctx, v->d, v->num_key_dimensions, v->dim_tc);
  if (rc) return rc;
  if (v->N >= 0xFFFFFFFFull) return set_err(ctx, PTC_ERR_RANGE, "N=%llu exceeds 2^32-2", (unsigned long long)v->N);
  if (v->N && (!v->hedges || !(v64 || v32))) return set_err(ctx, PTC_ERR_ARG, "NULL data arrays");
  PTC_HIP(ctx, hipSetDevice(ctx->device));
  ptc_dtensor* t = new ptc_dtensor();
  t->ctx = ctx;
  ctx_acquire(ctx);
  t->d = v->d;
  t->nk = v->num_key_dimensions;
  t->N = v->N;
  for (uint32_t j = 0; j < v->d; ++j) {
    t->dim_tc[j] = v->dim_tc[j];
    t->dims[j] = v->dims ? v->dims[j] : 0;
  }
  for (uint64_t i = 0; i < v->N; ++i)
    for (uint32_t j = 0; j < v->d; ++j) t->mode_max[j] = std::max(t->mode_max[j], v->hedges[i * v->d + j]);
  if (!max_index_ok(t->mode_max, t->d)) {
    ptc_dtensor_free(t);
    return set_err(ctx, PTC_ERR_RANGE, "coordinate 2^32-1 (the largest index is 2^32-2)");
  }
  t->vals_safe = values_safe(v64, v32, v->N);
  if (!v->dims)
    for (uint32_t j = 0; j < v->d; ++j) t->dims[j] = v->N ? t->mode_max[j] + 1 : 0;
  t->hedges = dalloc_n<uint32_t>(ctx, v->N * v->d);
  if (v64)
    t->vals = dalloc_n<double>(ctx, v->N);
  else
    t->vals32 = dalloc_n<float>(ctx, v->N);
  if (!t->hedges || !(t->vals || t->vals32)) {
    ptc_dtensor_free(t);
    return PTC_ERR_NOMEM;
  }
  if (v->N) {
    hipError_t e1 = hipMemcpyAsync(t->hedges, v->hedges, sizeof(uint32_t) * v->N * v->d, hipMemcpyHostToDevice, ctx->stream);
    hipError_t e2 = v64 ? hipMemcpyAsync(t->vals, v64, sizeof(double) * v->N, hipMemcpyHostToDevice, ctx->stream)
                        : hipMemcpyAsync(t->vals32, v32, sizeof(float) * v->N, hipMemcpyHostToDevice, ctx->stream);
    hipError_t e3 = hipStreamSynchronize(ctx->stream);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
      ptc_dtensor_free(t);
      return hip_check(ctx, e1 != hipSuccess ? e1 : (e2 != hipSuccess ? e2 : e3), "upload");
    }
  }
  *out = t;
  return PTC_OK;
}

extern "C" int ptc_dtensor_upload(ptc_ctx* ctx, const ptc_tensor_view* v, ptc_dtensor** out) {
  return upload(ctx, v, v ? v->vals : nullptr, nullptr, out);
}

extern "C" int ptc_dtensor_upload_f32(ptc_ctx* ctx, const ptc_tensor_view* v, const float* vals, ptc_dtensor** out) {
  if (v && v->N && !vals) return ctx ? set_err(ctx, PTC_ERR_ARG, "NULL FP32 values") : PTC_ERR_ARG;
  return upload(ctx, v, nullptr, vals, out);
}

extern "C" int ptc_dtensor_wrap(ptc_ctx* ctx, uint32_t d, uint32_t nk, uint64_t N,
                                const uint32_t* d_hedges, const double* d_vals,
                                const uint32_t* dim_tc, const uint32_t* dims,
                                const uint32_t* mode_max, ptc_dtensor** out) {
  if (!ctx || !out) return PTC_ERR_ARG;
  *out = nullptr;
  int rc = check_perm(ctx, d, nk, dim_tc);
  if (rc) return rc;
  if (N >= 0xFFFFFFFFull) return set_err(ctx, PTC_ERR_RANGE, "N exceeds 2^32-2");
  ptc_dtensor* t = new ptc_dtensor();
  t->ctx = ctx;
  ctx_acquire(ctx);
  t->d = d;
  t->nk = nk;
  t->N = N;
  t->owns = false;
  t->hedges = const_cast<uint32_t*>(d_hedges);
  t->vals = const_cast<double*>(d_vals);
  for (uint32_t j = 0; j < d; ++j) t->dim_tc[j] = dim_tc[j];
  if (mode_max) {
    for (uint32_t j = 0; j < d; ++j) t->mode_max[j] = mode_max[j];
  } else if (N) {
    Workspace ws(ctx);
    unsigned int* mx = ws.get<unsigned int>(PTC_MAX_MODES);
    if (!mx) {
      ptc_dtensor_free(t);
      return PTC_ERR_NOMEM;
    }
    rc = fill_zero(ctx, mx, sizeof(unsigned int) * PTC_MAX_MODES);
    if (!rc) {
      unsigned grid = (unsigned)std::min<uint64_t>((N + 255) / 256, 2048);
      hipLaunchKernelGGL(k_mode_max, dim3(grid), dim3(256), 0, ctx->stream, d_hedges, N, d, mx);
      rc = read_back(ctx, t->mode_max, mx, sizeof(uint32_t) * d);
    }
    if (rc) {
      ptc_dtensor_free(t);
      return rc;
    }
  }
  if (N && !max_index_ok(t->mode_max, d)) {
    ptc_dtensor_free(t);
    return set_err(ctx, PTC_ERR_RANGE, "coordinate 2^32-1 (the largest index is 2^32-2)");
  }
  for (uint32_t j = 0; j < d; ++j) t->dims[j] = dims ? dims[j] : (N ? t->mode_max[j] + 1 : 0);
  *out = t;
  return PTC_OK;
}

extern "C" int ptc_dtensor_set_keys(ptc_dtensor* t, uint32_t nk, const uint32_t* dim_tc) {
  if (!t) return PTC_ERR_ARG;
  int rc = check_perm(t->ctx, t->d, nk, dim_tc);
  if (rc) return rc;
  t->nk = nk;
  for (uint32_t j = 0; j < t->d; ++j) t->dim_tc[j] = dim_tc[j];
  return PTC_OK;
}

extern "C" int ptc_dtensor_info(const ptc_dtensor* t, uint32_t* d, uint64_t* N, uint32_t* nk) {
  if (!t) return PTC_ERR_ARG;
  if (d) *d = t->d;
  if (N) *N = t->N;
  if (nk) *nk = t->nk;
  return PTC_OK;
}

extern "C" int ptc_dtensor_device_ptrs(const ptc_dtensor* t, const uint32_t** h, const double** v,
                                       const float** v32) {
  if (!t) return PTC_ERR_ARG;
  if (h) *h = t->hedges;
  if (v) *v = t->vals;
  if (v32) *v32 = t->vals32;
  return PTC_OK;
}

__global__ void k_widen(const float* __restrict__ a, const float* __restrict__ err,
                        double* __restrict__ out, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (double)a[i] + (err ? (double)err[i] : 0.0);
}

extern "C" int ptc_dtensor_download(ptc_ctx* ctx, const ptc_dtensor* t, uint32_t* hedges,
                                    double* vals, uint32_t* dims) {
  if (!ctx || !t) return PTC_ERR_ARG;
  PTC_HIP(ctx, hipSetDevice(ctx->device));
  if (hedges && t->N)
    PTC_HIP(ctx, hipMemcpyAsync(hedges, t->hedges, sizeof(uint32_t) * t->N * t->d, hipMemcpyDeviceToHost, ctx->stream));
  if (vals && t->N) {
    if (t->vals) {
      PTC_HIP(ctx, hipMemcpyAsync(vals, t->vals, sizeof(double) * t->N, hipMemcpyDeviceToHost, ctx->stream));
    } else if (t->vals32) {
      // FP32 values, widened exactly (the lo part of an FP32 result comes
      // from ptc_dtensor_download_f32)
      Workspace ws(ctx);
      double* tmp = ws.get<double>(t->N);
      if (!tmp) return PTC_ERR_NOMEM;
      unsigned grid = (unsigned)std::min<uint64_t>((t->N + 255) / 256, 4096);
      hipLaunchKernelGGL(k_widen, dim3(grid), dim3(256), 0, ctx->stream, t->vals32, (const float*)nullptr, tmp, t->N);
      PTC_HIP(ctx, hipMemcpyAsync(vals, tmp, sizeof(double) * t->N, hipMemcpyDeviceToHost, ctx->stream));
      PTC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
  }
  if (dims)
    for (uint32_t j = 0; j < t->d; ++j) dims[j] = t->dims[j];
  PTC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PTC_OK;
}

extern "C" int ptc_dtensor_download_f32(ptc_ctx* ctx, const ptc_dtensor* t, float* vals, float* lo) {
  if (!ctx || !t || !vals) return PTC_ERR_ARG;
  if (!t->vals32) return set_err(ctx, PTC_ERR_ARG, "tensor has no FP32 values");
  if (lo && !t->err32) return set_err(ctx, PTC_ERR_ARG, "tensor is not an FP32 contraction result (no lo part)");
  if (t->N)
    PTC_HIP(ctx, hipMemcpyAsync(vals, t->vals32, sizeof(float) * t->N, hipMemcpyDeviceToHost, ctx->stream));
  if (t->N && lo)
    PTC_HIP(ctx, hipMemcpyAsync(lo, t->err32, sizeof(float) * t->N, hipMemcpyDeviceToHost, ctx->stream));
  PTC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PTC_OK;
}

extern "C" int ptc_dtensor_set_values(ptc_ctx* ctx, ptc_dtensor* t, const double* vals) {
  if (!ctx || !t || (t->N && !vals)) return PTC_ERR_ARG;
  if (!t->vals) return set_err(ctx, PTC_ERR_ARG, "tensor has no FP64 values");
  if (t->N) {
    PTC_HIP(ctx, hipMemcpyAsync(t->vals, vals, sizeof(double) * t->N, hipMemcpyHostToDevice, ctx->stream));
    PTC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  // zero, tiny or huge values make later contractions keep touch flags
  t->vals_safe = values_safe(vals, nullptr, t->N);
  return PTC_OK;
}

extern "C" void ptc_dtensor_free(ptc_dtensor* t) {
  if (!t) return;
  ptc_ctx* ctx = t->ctx;
  if (t->owns && ctx_open(ctx)) {
    dfree(ctx, t->hedges);
    dfree(ctx, t->vals);
    dfree(ctx, t->vals32);
    dfree(ctx, t->err32);
  }
  delete t;
  ctx_release(ctx);
}

// --------------------------------------------------------------- .meta

extern "C" int ptc_meta_parse(const char* path, uint32_t d, uint32_t* c_out, uint32_t* dim_tc_out) {
  if (!path || !c_out || !dim_tc_out || d == 0 || d > PTC_MAX_MODES) return PTC_ERR_ARG;
  std::ifstream in(path);
  if (!in.is_open()) return PTC_ERR_IO;
  std::string line;
  for (int i = 0; i < 2; ++i)
    if (!std::getline(in, line)) return PTC_ERR_IO;  // two free-text lines (tensor.hpp:159-160)
  long long c = -1;
  if (!(in >> c) || c < 0 || (uint64_t)c > d) return PTC_ERR_IO;
  uint32_t seen = 0;
  for (uint32_t j = 0; j < d; ++j) {
    long long m = -1;
    if (!(in >> m) || m < 0 || (uint64_t)m >= d || ((seen >> m) & 1u)) return PTC_ERR_IO;
    seen |= 1u << m;
    dim_tc_out[j] = (uint32_t)m;
  }
  *c_out = (uint32_t)c;
  return PTC_OK;
}
