// Learned row-size prediction.
//
// A 3-layer perceptron (6 -> 64 -> 64 -> 1, ReLU) maps cheap per-row
// features, known right after the key probes, to the row's number of output
// nonzeros.  Features of row r (all log1p, scaled by 1/8):
//   0 A-row nonzeros            1 flops = sum of probed B-slice lengths
//   2 contracted keys that hit  3 longest slice
//   4 column space              5 upper bound ub = min(flops, column space)
// The output z gives pred = ub * sigmoid(z); the loss is
// (log pred - log nnz)^2, a relative error, averaged over a batch.
//
// Kernels (all FP32, one thread per sample or per weight):
//   k_mlp_features  row arrays -> feature rows (on the device, per contraction)
//   k_mlp_predict   weights staged in LDS, one wavefront per row
//   k_mlp_forward   training forward: stores H1, H2 and dL/dz
//   k_mlp_backward  per-sample dH2, dH1
//   k_mlp_wgrad     per-weight reduction over the batch (dW = dH^T H)
//   k_mlp_adam      Adam update
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

#include "ptc_learned.h"

namespace ptc {

constexpr int MLP_IN = PTC_MLP_IN;
constexpr int MLP_H = PTC_MLP_HIDDEN;
// parameter layout: W1[H][IN], b1[H], W2[H][H], b2[H], w3[H], b3
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + MLP_H * MLP_IN;
constexpr int OFF_W2 = OFF_B1 + MLP_H;
constexpr int OFF_B2 = OFF_W2 + MLP_H * MLP_H;
constexpr int OFF_W3 = OFF_B2 + MLP_H;
constexpr int OFF_B3 = OFF_W3 + MLP_H;
constexpr int NPARAM = OFF_B3 + 1;
static_assert(NPARAM == PTC_MLP_NPARAM, "parameter count");

}  // namespace ptc

struct ptc_mlp {
  ptc_ctx* ctx = nullptr;
  float* w = nullptr;   // device, NPARAM
  float* m = nullptr;   // Adam moments
  float* v = nullptr;
  uint64_t step = 0;
};

namespace ptc {

__device__ __forceinline__ float softplus(float x) { return x > 20.f ? x : log1pf(expf(x)); }
__device__ __forceinline__ float sigmoidf_(float x) { return 1.f / (1.f + expf(-x)); }

// forward of one sample; h1, h2 are the post-ReLU activations
__device__ __forceinline__ float mlp_fwd(const float* __restrict__ w, const float* x, float* h1, float* h2) {
#pragma unroll 4
  for (int i = 0; i < MLP_H; ++i) {
    float s = w[OFF_B1 + i];
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) s += w[OFF_W1 + i * MLP_IN + k] * x[k];
    h1[i] = s > 0.f ? s : 0.f;
  }
  float z = w[OFF_B3];
  for (int i = 0; i < MLP_H; ++i) {
    float s = w[OFF_B2 + i];
#pragma unroll 8
    for (int k = 0; k < MLP_H; ++k) s += w[OFF_W2 + i * MLP_H + k] * h1[k];
    s = s > 0.f ? s : 0.f;
    h2[i] = s;
    z += w[OFF_W3 + i] * s;
  }
  return z;
}

__global__ void k_mlp_features(LearnedInputs in, float* __restrict__ X, float* __restrict__ ub) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < in.nrows; r += (uint64_t)gridDim.x * blockDim.x) {
    const double f = (double)in.row_flops[r];
    const double u = f < (double)in.ncols ? f : (double)in.ncols;
    float* x = X + r * MLP_IN;
    x[0] = (float)(log1p((double)in.rows[r].len) * 0.125);
    x[1] = (float)(log1p(f) * 0.125);
    x[2] = (float)(log1p((double)in.row_hits[r]) * 0.125);
    x[3] = (float)(log1p((double)in.row_maxlen[r]) * 0.125);
    x[4] = (float)(log1p((double)in.ncols) * 0.125);
    x[5] = (float)(log1p(u) * 0.125);
    ub[r] = (float)u;
  }
}

// Rows per 256-thread inference block: four wavefronts, two rows each, so a
// few thousand rows still spread over every CU.
constexpr int PREDICT_ROWS_PER_BLOCK = 8;

// Inference: one wavefront per row, lane i = hidden unit i.  The second
// layer's weights are staged transposed (W2T[k][i]) so the 64 lanes read
// consecutive LDS words; h1 is broadcast lane by lane with shuffles.
__global__ void __launch_bounds__(256) k_mlp_predict(const float* __restrict__ wg, const float* __restrict__ X,
                                                     const float* __restrict__ ub, uint64_t n,
                                                     float* __restrict__ pred) {
  static_assert(MLP_H == 64, "one lane per hidden unit");
  __shared__ float w2t[MLP_H * MLP_H];
  __shared__ float w1[MLP_H * MLP_IN], b1[MLP_H], b2[MLP_H], w3[MLP_H];
  for (int i = threadIdx.x; i < MLP_H * MLP_H; i += blockDim.x) {
    const int r = i / MLP_H, c = i % MLP_H;  // W2[r][c]
    w2t[c * MLP_H + r] = wg[OFF_W2 + i];
  }
  for (int i = threadIdx.x; i < MLP_H * MLP_IN; i += blockDim.x) w1[i] = wg[OFF_W1 + i];
  for (int i = threadIdx.x; i < MLP_H; i += blockDim.x) {
    b1[i] = wg[OFF_B1 + i];
    b2[i] = wg[OFF_B2 + i];
    w3[i] = wg[OFF_W3 + i];
  }
  __syncthreads();
  const float b3 = wg[OFF_B3];
  const int lane = threadIdx.x & 63;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
  for (uint64_t r = (uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; r < n; r += waves) {
    float a = b1[lane];
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) a += w1[lane * MLP_IN + k] * X[r * MLP_IN + k];
    const float h1 = a > 0.f ? a : 0.f;
    float s = b2[lane];
#pragma unroll 16
    for (int k = 0; k < MLP_H; ++k) s += w2t[k * MLP_H + lane] * __shfl(h1, k, 64);
    const float h2 = s > 0.f ? s : 0.f;
    float z = w3[lane] * h2;
    for (int o = 32; o > 0; o >>= 1) z += __shfl_down(z, o, 64);
    if (lane == 0) pred[r] = ub[r] * sigmoidf_(z + b3);
  }
}

// Features, inference and the row's capacity in one launch (the learned
// single-pass method): the features of k_mlp_features, the network of
// k_mlp_predict (the same operations in the same order), then the capacity
// rule of k_caps: clamp(ceil(scale * prediction), longest slice,
// min(flops, columns)).  Lane i computes hidden unit i from the weights,
// staged once per workgroup in LDS (W2 transposed, so the 64 lanes read 64
// consecutive words); each wavefront runs one row at a time, its features'
// loads all in flight together.
constexpr int CAPS_THREADS = 512;
constexpr int CAPS_WAVES_PER_SIMD = 4;  // two workgroups a CU: every workgroup of a few thousand rows resident at once
__device__ __forceinline__ float lane_value(float x, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), k));
}

// Row r's count of positions that hit B and its longest slice, over the
// wavefront (the same values in every lane), for a row of positions [a, b):
// the positions of the parts (TILE_PART positions each) the row only partly
// covers, at most two, in 16-byte words, all loads of a lane in flight
// together (pos_len holds three words past its last position, so the word
// of the last position is in bounds; positions outside the row are
// masked); the parts strictly inside the row by their TilePart.
constexpr int FEAT_HEAD_Q = (TILE_PART / 4 + 1 + 63) / 64;  // 16-byte words a lane, unaligned start
constexpr int FEAT_TAIL_Q = (TILE_PART / 4 + 63) / 64;      // 16-byte words a lane, part-aligned start
__device__ __forceinline__ void feat_word(const uint4& w, uint64_t q, uint64_t lo, uint64_t hi, uint32_t& h,
                                          uint32_t& m) {
  const uint32_t v[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const uint64_t e = 4 * q + c;
    const uint32_t l = e >= lo && e < hi ? v[c] : 0u;
    h += l ? 1u : 0u;
    m = max(m, l);
  }
}
__device__ __forceinline__ void row_features(const LearnedInputs& in, const KtSlice& sl, int lane, uint32_t& hits,
                                             uint32_t& maxlen) {
  const uint64_t a = sl.off, b = a + sl.len;  // len >= 1: the row has flops
  const uint64_t ta = a / TILE_PART, tb = (b - 1) / TILE_PART;
  const uint64_t e0 = ta == tb ? b : (ta + 1) * TILE_PART;  // the head: [a, e0)
  const uint64_t s1 = tb * TILE_PART;                         // the tail: [s1, b) when ta < tb
  const uint4* __restrict__ w4 = reinterpret_cast<const uint4*>(in.pos_len);
  const uint64_t hq0 = a >> 2, hq1 = (e0 + 3) >> 2, tq0 = s1 >> 2, tq1 = ta == tb ? tq0 : (b + 3) >> 2;
  uint4 hw[FEAT_HEAD_Q], tw[FEAT_TAIL_Q];
#pragma unroll
  for (int j = 0; j < FEAT_HEAD_Q; ++j) {
    const uint64_t q = hq0 + lane + 64 * j;
    hw[j] = q < hq1 ? w4[q] : make_uint4(0, 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < FEAT_TAIL_Q; ++j) {
    const uint64_t q = tq0 + lane + 64 * j;
    tw[j] = q < tq1 ? w4[q] : make_uint4(0, 0, 0, 0);
  }
  uint32_t h = 0, m = 0;
  for (uint64_t t0 = ta + 1 + lane; t0 < tb; t0 += 4 * 64) {  // four tile loads a lane in flight
    TilePart p[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) p[u] = t0 + u * 64 < tb ? in.tiles[t0 + u * 64] : TilePart{0, 0};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      h += p[u].h;
      m = max(m, p[u].m);
    }
  }
#pragma unroll
  for (int j = 0; j < FEAT_HEAD_Q; ++j) feat_word(hw[j], hq0 + lane + 64 * j, a, e0, h, m);
#pragma unroll
  for (int j = 0; j < FEAT_TAIL_Q; ++j) feat_word(tw[j], tq0 + lane + 64 * j, s1, ta == tb ? s1 : b, h, m);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    h += __shfl_xor(h, o, 64);
    m = max(m, (uint32_t)__shfl_xor(m, o, 64));
  }
  hits = h;
  maxlen = m;
}

__global__ void __launch_bounds__(CAPS_THREADS, CAPS_WAVES_PER_SIMD) k_mlp_caps(const float* __restrict__ wg, LearnedInputs in,
                                                           double scale, uint32_t* __restrict__ cap) {
  static_assert(MLP_H == 64, "one lane per hidden unit");
  __shared__ float w2t[MLP_H * MLP_H];  // W2 transposed: w2t[k][i] = W2[i][k]
  __shared__ float w1s[MLP_H * MLP_IN], b1s[MLP_H], b2s[MLP_H], w3s[MLP_H];
  const int lane = threadIdx.x & 63;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
  uint64_t r = (uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  // every staging load in flight before the first LDS write
  static_assert(OFF_W2 % 4 == 0, "W2 is 16-byte aligned");
  static_assert(MLP_H * MLP_H % (4 * CAPS_THREADS) == 0 && MLP_H * MLP_IN <= CAPS_THREADS, "staging shape");
  constexpr int W2Q = MLP_H * MLP_H / (4 * CAPS_THREADS);
  float4 q[W2Q];
#pragma unroll
  for (int j = 0; j < W2Q; ++j) q[j] = reinterpret_cast<const float4*>(wg + OFF_W2)[threadIdx.x + j * CAPS_THREADS];
  const float w1v = threadIdx.x < MLP_H * MLP_IN ? wg[OFF_W1 + threadIdx.x] : 0.f;
  const int hu = threadIdx.x & (MLP_H - 1);
  const float b1v = wg[OFF_B1 + hu], b2v = wg[OFF_B2 + hu], w3v = wg[OFF_W3 + hu];
  // the first row's flops and slice, in flight with the weights (the row's
  // slice is read with its flops: r < nrows <= the slices allocated; used
  // only when the row has flops)
  unsigned long long fl = 0;
  KtSlice sl{};
  if (r < in.nrows) {
    fl = in.row_flops[r];
    sl = in.rows[r];
  }
#pragma unroll
  for (int j = 0; j < W2Q; ++j) {
    const uint32_t t = threadIdx.x + j * CAPS_THREADS;
    const uint32_t i = (4 * t) / MLP_H, k = (4 * t) % MLP_H;
    w2t[(k + 0) * MLP_H + i] = q[j].x;
    w2t[(k + 1) * MLP_H + i] = q[j].y;
    w2t[(k + 2) * MLP_H + i] = q[j].z;
    w2t[(k + 3) * MLP_H + i] = q[j].w;
  }
  if (threadIdx.x < MLP_H * MLP_IN) w1s[threadIdx.x] = w1v;
  if (threadIdx.x < MLP_H) {
    b1s[threadIdx.x] = b1v;
    b2s[threadIdx.x] = b2v;
    w3s[threadIdx.x] = w3v;
  }
  const float b3 = wg[OFF_B3];
  __syncthreads();
  for (; r < in.nrows; r += waves) {
    uint32_t hits = 0, maxlen = 0;
    if (in.row_hits) {
      hits = in.row_hits[r];
      maxlen = in.row_maxlen[r];
    } else if (fl) {
      row_features(in, sl, lane, hits, maxlen);
    }
    if (fl) {
      const double f = (double)fl;
      const double u = f < (double)in.ncols ? f : (double)in.ncols;
      // lane k < MLP_IN computes feature k
      const double src = lane == 0 ? (double)sl.len : lane == 1 ? f : lane == 2 ? (double)hits
                       : lane == 3 ? (double)maxlen : lane == 4 ? (double)in.ncols : u;
      const float xl = lane < MLP_IN ? (float)(log1p(src) * 0.125) : 0.f;
      // lane k's value to every lane by readlane (a scalar register): the
      // 64-step sum below is one chain of adds, so each step must not wait
      // on a cross-lane LDS permute
      float a = b1s[lane];
#pragma unroll
      for (int k = 0; k < MLP_IN; ++k) a += w1s[lane * MLP_IN + k] * lane_value(xl, k);
      const float h1 = a > 0.f ? a : 0.f;
      float s = b2s[lane];
#pragma unroll
      for (int k = 0; k < MLP_H; ++k) s += w2t[k * MLP_H + lane] * lane_value(h1, k);
      const float h2 = s > 0.f ? s : 0.f;
      float z = w3s[lane] * h2;
      for (int o = 32; o > 0; o >>= 1) z += __shfl_down(z, o, 64);
      if (lane == 0) {
        const float pred = (float)u * sigmoidf_(z + b3);
        const unsigned long long ub = fl < in.ncols ? fl : in.ncols;
        const double want = ceil((double)pred * scale);
        const unsigned long long wv = want < 0 ? 0ull : (unsigned long long)fmin(want, 1.8e19);
        unsigned long long c = wv < maxlen ? maxlen : wv;
        if (c > ub) c = ub;
        cap[r] = c > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)c;
      }
    } else if (lane == 0) {
      cap[r] = 0;  // no flops (rows past the table too): capacity 0 whatever the network says
    }
    const uint64_t rn = r + waves;
    if (rn < in.nrows) {
      fl = in.row_flops[rn];
      sl = in.rows[rn];
    }
  }
}

// training forward: loss and dL/dz per sample (mean over the batch)
__global__ void __launch_bounds__(256) k_mlp_forward(const float* __restrict__ wg, const float* __restrict__ X,
                                                     const float* __restrict__ ub, const float* __restrict__ y,
                                                     uint64_t n, float* __restrict__ H1, float* __restrict__ H2,
                                                     float* __restrict__ dz, float* __restrict__ loss) {
  __shared__ float w[NPARAM];
  __shared__ float sl[256 / 64];
  for (int i = threadIdx.x; i < NPARAM; i += blockDim.x) w[i] = wg[i];
  __syncthreads();
  float l = 0.f;
  uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b < n) {
    float x[MLP_IN], h1[MLP_H], h2[MLP_H];
    for (int k = 0; k < MLP_IN; ++k) x[k] = X[b * MLP_IN + k];
    float z = mlp_fwd(w, x, h1, h2);
    for (int i = 0; i < MLP_H; ++i) {
      H1[b * MLP_H + i] = h1[i];
      H2[b * MLP_H + i] = h2[i];
    }
    // log pred = log ub + log sigmoid(z) = log ub - softplus(-z)
    float e = logf(fmaxf(ub[b], 1.f)) - softplus(-z) - logf(fmaxf(y[b], 1.f));
    l = e * e;
    // d/dz of e^2 / n: 2 e * sigmoid(-z) / n
    dz[b] = 2.f * e * sigmoidf_(-z) / (float)n;
  }
  for (int o = 32; o > 0; o >>= 1) l += __shfl_down(l, o, 64);
  if ((threadIdx.x & 63) == 0) sl[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int i = 0; i < 256 / 64; ++i) t += sl[i];
    atomicAdd(loss, t / (float)n);
  }
}

// per-sample backprop to the hidden layers
__global__ void __launch_bounds__(256) k_mlp_backward(const float* __restrict__ wg, uint64_t n,
                                                      const float* __restrict__ H1, const float* __restrict__ H2,
                                                      const float* __restrict__ dz, float* __restrict__ D1,
                                                      float* __restrict__ D2) {
  __shared__ float w[NPARAM];
  for (int i = threadIdx.x; i < NPARAM; i += blockDim.x) w[i] = wg[i];
  __syncthreads();
  uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  float d2[MLP_H];
  const float g = dz[b];
  for (int i = 0; i < MLP_H; ++i) {
    d2[i] = H2[b * MLP_H + i] > 0.f ? g * w[OFF_W3 + i] : 0.f;
    D2[b * MLP_H + i] = d2[i];
  }
  for (int k = 0; k < MLP_H; ++k) {
    float s = 0.f;
    for (int i = 0; i < MLP_H; ++i) s += w[OFF_W2 + i * MLP_H + k] * d2[i];
    D1[b * MLP_H + k] = H1[b * MLP_H + k] > 0.f ? s : 0.f;
  }
}

// one thread per parameter: sum over the batch
__global__ void __launch_bounds__(256) k_mlp_wgrad(const float* __restrict__ X, uint64_t n,
                                                   const float* __restrict__ H1, const float* __restrict__ H2,
                                                   const float* __restrict__ dz, const float* __restrict__ D1,
                                                   const float* __restrict__ D2, float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  double s = 0.0;
  if (p < OFF_B1) {
    const int i = (p - OFF_W1) / MLP_IN, k = (p - OFF_W1) % MLP_IN;
    for (uint64_t b = 0; b < n; ++b) s += (double)D1[b * MLP_H + i] * X[b * MLP_IN + k];
  } else if (p < OFF_W2) {
    const int i = p - OFF_B1;
    for (uint64_t b = 0; b < n; ++b) s += D1[b * MLP_H + i];
  } else if (p < OFF_B2) {
    const int i = (p - OFF_W2) / MLP_H, k = (p - OFF_W2) % MLP_H;
    for (uint64_t b = 0; b < n; ++b) s += (double)D2[b * MLP_H + i] * H1[b * MLP_H + k];
  } else if (p < OFF_W3) {
    const int i = p - OFF_B2;
    for (uint64_t b = 0; b < n; ++b) s += D2[b * MLP_H + i];
  } else if (p < OFF_B3) {
    const int i = p - OFF_W3;
    for (uint64_t b = 0; b < n; ++b) s += (double)dz[b] * H2[b * MLP_H + i];
  } else {
    for (uint64_t b = 0; b < n; ++b) s += dz[b];
  }
  grad[p] = (float)s;
}

__global__ void k_mlp_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ g, float lr, float b1t, float b2t) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float gm = b1 * m[p] + (1.f - b1) * g[p];
  float gv = b2 * v[p] + (1.f - b2) * g[p] * g[p];
  m[p] = gm;
  v[p] = gv;
  w[p] -= lr * (gm / (1.f - b1t)) / (sqrtf(gv / (1.f - b2t)) + eps);
}

static unsigned gridn(uint64_t n, unsigned t, unsigned cap = 8192) {
  uint64_t g = (n + t - 1) / t;
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

static const ptc_mlp* g_default = nullptr;

int learned_predict(ptc_ctx* ctx, const ptc_mlp* mlp, const LearnedInputs& in, float* pred) {
  if (!mlp) mlp = g_default;
  if (!mlp || !mlp->w) return set_err(ctx, PTC_ERR_ARG, "no row-size model");
  if (!in.nrows) return PTC_OK;
  Workspace ws(ctx);
  float* X = ws.get<float>(in.nrows * MLP_IN);
  float* ub = ws.get<float>(in.nrows);
  if (!X || !ub) return PTC_ERR_NOMEM;
  hipLaunchKernelGGL(k_mlp_features, dim3(gridn(in.nrows, 256)), dim3(256), 0, ctx->stream, in, X, ub);
  hipLaunchKernelGGL(k_mlp_predict, dim3(gridn(in.nrows, PREDICT_ROWS_PER_BLOCK, 2048)), dim3(256), 0, ctx->stream,
                     (const float*)mlp->w, (const float*)X, (const float*)ub, in.nrows, pred);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

int learned_caps(ptc_ctx* ctx, const ptc_mlp* mlp, const LearnedInputs& in, double scale, uint32_t* cap) {
  if (!mlp) mlp = g_default;
  if (!mlp || !mlp->w) return set_err(ctx, PTC_ERR_ARG, "no row-size model");
  if (!in.nrows) return PTC_OK;
  hipLaunchKernelGGL(k_mlp_caps, dim3(gridn(in.nrows, CAPS_THREADS / 64, 1024)), dim3(CAPS_THREADS), 0, ctx->stream,
                     (const float*)mlp->w, in, scale, cap);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

int learned_features(ptc_ctx* ctx, const LearnedInputs& in, float* X, float* ub) {
  if (!in.nrows) return PTC_OK;
  hipLaunchKernelGGL(k_mlp_features, dim3(gridn(in.nrows, 256)), dim3(256), 0, ctx->stream, in, X, ub);
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

// one optimisation step (or, with lr == 0, only the gradient) on a batch
static int train_batch(ptc_ctx* ctx, ptc_mlp* m, const float* X, const float* ub, const float* y, uint64_t n,
                       float lr, float* grad, float* loss, Workspace& ws) {
  float* H1 = ws.get<float>(n * MLP_H);
  float* H2 = ws.get<float>(n * MLP_H);
  float* D1 = ws.get<float>(n * MLP_H);
  float* D2 = ws.get<float>(n * MLP_H);
  float* dz = ws.get<float>(n);
  if (!H1 || !H2 || !D1 || !D2 || !dz) return PTC_ERR_NOMEM;
  hipStream_t st = ctx->stream;
  unsigned g = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_mlp_forward, dim3(g), dim3(256), 0, st, (const float*)m->w, X, ub, y, n, H1, H2, dz, loss);
  hipLaunchKernelGGL(k_mlp_backward, dim3(g), dim3(256), 0, st, (const float*)m->w, n, (const float*)H1,
                     (const float*)H2, (const float*)dz, D1, D2);
  hipLaunchKernelGGL(k_mlp_wgrad, dim3((NPARAM + 255) / 256), dim3(256), 0, st, X, n, (const float*)H1,
                     (const float*)H2, (const float*)dz, (const float*)D1, (const float*)D2, grad);
  if (lr > 0.f) {
    m->step++;
    float b1t = powf(0.9f, (float)m->step), b2t = powf(0.999f, (float)m->step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, st, m->w, m->m, m->v,
                       (const float*)grad, lr, b1t, b2t);
  }
  PTC_HIP(ctx, hipGetLastError());
  return PTC_OK;
}

}  // namespace ptc

using namespace ptc;

extern "C" const ptc_mlp* ptc_learned_default(void) { return g_default; }

extern "C" void ptc_learned_set_default(const ptc_mlp* m) { g_default = m; }

extern "C" int ptc_mlp_create(ptc_ctx* ctx, uint64_t seed, ptc_mlp** out) {
  if (!ctx || !out) return PTC_ERR_ARG;
  *out = nullptr;
  PTC_HIP(ctx, hipSetDevice(ctx->device));
  // He initialisation from the project's xoshiro stream
  std::vector<float> w(NPARAM, 0.f);
  Xoshiro rng(seed);
  auto normal = [&]() {
    double u1 = rng.uniform(), u2 = rng.uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  };
  for (int i = 0; i < MLP_H * MLP_IN; ++i) w[OFF_W1 + i] = (float)(normal() * std::sqrt(2.0 / MLP_IN));
  for (int i = 0; i < MLP_H * MLP_H; ++i) w[OFF_W2 + i] = (float)(normal() * std::sqrt(2.0 / MLP_H));
  for (int i = 0; i < MLP_H; ++i) w[OFF_W3 + i] = (float)(normal() * std::sqrt(1.0 / MLP_H));
  w[OFF_B3] = 2.0f;  // start near "dense": sigmoid(2) = 0.88
  ptc_mlp* m = new ptc_mlp();
  m->ctx = ctx;
  ctx_acquire(ctx);
  m->w = dalloc_n<float>(ctx, NPARAM);
  m->m = dalloc_n<float>(ctx, NPARAM);
  m->v = dalloc_n<float>(ctx, NPARAM);
  if (!m->w || !m->m || !m->v) {
    ptc_mlp_free(m);
    return PTC_ERR_NOMEM;
  }
  PTC_HIP(ctx, hipMemcpyAsync(m->w, w.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice, ctx->stream));
  int rc = fill_zero(ctx, m->m, sizeof(float) * NPARAM);
  if (!rc) rc = fill_zero(ctx, m->v, sizeof(float) * NPARAM);
  PTC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (rc) {
    ptc_mlp_free(m);
    return rc;
  }
  *out = m;
  return PTC_OK;
}

extern "C" void ptc_mlp_free(ptc_mlp* m) {
  if (!m) return;
  if (g_default == m) g_default = nullptr;
  ptc_ctx* ctx = m->ctx;
  if (ctx_open(ctx)) {
    dfree(ctx, m->w);
    dfree(ctx, m->m);
    dfree(ctx, m->v);
  }
  delete m;
  ctx_release(ctx);
}

extern "C" int ptc_mlp_num_params(void) { return NPARAM; }

extern "C" int ptc_mlp_get_params(const ptc_mlp* m, float* host) {
  if (!m || !host) return PTC_ERR_ARG;
  PTC_HIP(m->ctx, hipMemcpyAsync(host, m->w, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, m->ctx->stream));
  PTC_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return PTC_OK;
}

extern "C" int ptc_mlp_set_params(ptc_mlp* m, const float* host) {
  if (!m || !host) return PTC_ERR_ARG;
  PTC_HIP(m->ctx, hipMemcpyAsync(m->w, host, sizeof(float) * NPARAM, hipMemcpyHostToDevice, m->ctx->stream));
  PTC_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
  m->step = 0;
  int rc = fill_zero(m->ctx, m->m, sizeof(float) * NPARAM);
  if (!rc) rc = fill_zero(m->ctx, m->v, sizeof(float) * NPARAM);
  return rc;
}

// host arrays X (n x MLP_IN), ub (n), y (n)
static int upload_set(ptc_ctx* ctx, Workspace& ws, const float* X, const float* ub, const float* y, uint64_t n,
                      float** dX, float** dub, float** dy) {
  *dX = ws.get<float>(n * MLP_IN);
  *dub = ws.get<float>(n);
  *dy = y ? ws.get<float>(n) : nullptr;
  if (!*dX || !*dub || (y && !*dy)) return PTC_ERR_NOMEM;
  PTC_HIP(ctx, hipMemcpyAsync(*dX, X, sizeof(float) * n * MLP_IN, hipMemcpyHostToDevice, ctx->stream));
  PTC_HIP(ctx, hipMemcpyAsync(*dub, ub, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
  if (y) PTC_HIP(ctx, hipMemcpyAsync(*dy, y, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
  return PTC_OK;
}

extern "C" int ptc_mlp_grad(ptc_ctx* ctx, ptc_mlp* m, const float* X, const float* ub, const float* y, uint64_t n,
                            float* grad_out, float* loss_out) {
  if (!ctx || !m || !X || !ub || !y || !grad_out || n == 0) return PTC_ERR_ARG;
  Workspace ws(ctx);
  float *dX, *dub, *dy;
  int rc = upload_set(ctx, ws, X, ub, y, n, &dX, &dub, &dy);
  if (rc) return rc;
  float* g = ws.get<float>(NPARAM);
  float* loss = ws.get<float>(1);
  if (!g || !loss) return PTC_ERR_NOMEM;
  if ((rc = fill_zero(ctx, loss, 4))) return rc;
  if ((rc = train_batch(ctx, m, dX, dub, dy, n, 0.f, g, loss, ws))) return rc;
  PTC_HIP(ctx, hipMemcpyAsync(grad_out, g, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, ctx->stream));
  float l = 0.f;
  PTC_HIP(ctx, hipMemcpyAsync(&l, loss, 4, hipMemcpyDeviceToHost, ctx->stream));
  PTC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (loss_out) *loss_out = l;
  return PTC_OK;
}

extern "C" int ptc_mlp_train(ptc_ctx* ctx, ptc_mlp* m, const float* X, const float* ub, const float* y, uint64_t n,
                             uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float* loss_out) {
  if (!ctx || !m || !X || !ub || !y || n == 0 || batch == 0) return PTC_ERR_ARG;
  // shuffle on the host, then upload once
  std::vector<uint64_t> order(n);
  for (uint64_t i = 0; i < n; ++i) order[i] = i;
  Xoshiro rng(seed);
  std::vector<float> Xs(n * MLP_IN), us(n), ys(n);
  Workspace ws(ctx);
  float* dX = ws.get<float>(n * MLP_IN);
  float* du = ws.get<float>(n);
  float* dy = ws.get<float>(n);
  float* g = ws.get<float>(NPARAM);
  float* loss = ws.get<float>(epochs ? epochs : 1);
  if (!dX || !du || !dy || !g || !loss) return PTC_ERR_NOMEM;
  int rc = fill_zero(ctx, loss, sizeof(float) * (epochs ? epochs : 1));
  if (rc) return rc;
  for (uint32_t ep = 0; ep < epochs; ++ep) {
    for (uint64_t i = n - 1; i > 0; --i) {
      uint64_t j = rng.next() % (i + 1);
      std::swap(order[i], order[j]);
    }
    for (uint64_t i = 0; i < n; ++i) {
      memcpy(&Xs[i * MLP_IN], X + order[i] * MLP_IN, sizeof(float) * MLP_IN);
      us[i] = ub[order[i]];
      ys[i] = y[order[i]];
    }
    PTC_HIP(ctx, hipMemcpyAsync(dX, Xs.data(), sizeof(float) * n * MLP_IN, hipMemcpyHostToDevice, ctx->stream));
    PTC_HIP(ctx, hipMemcpyAsync(du, us.data(), sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    PTC_HIP(ctx, hipMemcpyAsync(dy, ys.data(), sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    for (uint64_t b0 = 0; b0 < n; b0 += batch) {
      uint64_t nb = std::min<uint64_t>(batch, n - b0);
      Workspace step_ws(ctx);
      if ((rc = train_batch(ctx, m, dX + b0 * MLP_IN, du + b0, dy + b0, nb, lr, g, loss + ep, step_ws))) return rc;
    }
    PTC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  if (loss_out && epochs) {
    std::vector<float> l(epochs);
    PTC_HIP(ctx, hipMemcpyAsync(l.data(), loss, sizeof(float) * epochs, hipMemcpyDeviceToHost, ctx->stream));
    PTC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // mean batch loss per epoch
    uint64_t nbatches = (n + batch - 1) / batch;
    for (uint32_t e = 0; e < epochs; ++e) loss_out[e] = l[e] / (float)nbatches;
  }
  return PTC_OK;
}

extern "C" int ptc_mlp_predict(ptc_ctx* ctx, const ptc_mlp* m, const float* X, const float* ub, uint64_t n,
                               float* pred_out) {
  if (!ctx || !m || !X || !ub || !pred_out) return PTC_ERR_ARG;
  if (!n) return PTC_OK;
  Workspace ws(ctx);
  float *dX, *du, *dy;
  int rc = upload_set(ctx, ws, X, ub, nullptr, n, &dX, &du, &dy);
  if (rc) return rc;
  float* dp = ws.get<float>(n);
  if (!dp) return PTC_ERR_NOMEM;
  hipLaunchKernelGGL(k_mlp_predict, dim3(gridn(n, PREDICT_ROWS_PER_BLOCK, 2048)), dim3(256), 0, ctx->stream, (const float*)m->w,
                     (const float*)dX, (const float*)du, n, dp);
  PTC_HIP(ctx, hipGetLastError());
  PTC_HIP(ctx, hipMemcpyAsync(pred_out, dp, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
  PTC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PTC_OK;
}

// file: "PTCMLP1\0", uint32 nparam, float[nparam]
extern "C" int ptc_mlp_save(const ptc_mlp* m, const char* path) {
  if (!m || !path) return PTC_ERR_ARG;
  std::vector<float> w(NPARAM);
  int rc = ptc_mlp_get_params(m, w.data());
  if (rc) return rc;
  std::ofstream f(path, std::ios::binary);
  if (!f) return PTC_ERR_IO;
  const char magic[8] = {'P', 'T', 'C', 'M', 'L', 'P', '1', 0};
  uint32_t np = NPARAM;
  f.write(magic, 8);
  f.write((const char*)&np, 4);
  f.write((const char*)w.data(), sizeof(float) * NPARAM);
  return f ? PTC_OK : PTC_ERR_IO;
}

extern "C" int ptc_mlp_load(ptc_ctx* ctx, const char* path, ptc_mlp** out) {
  if (!ctx || !path || !out) return PTC_ERR_ARG;
  *out = nullptr;
  std::ifstream f(path, std::ios::binary);
  if (!f) return PTC_ERR_IO;
  char magic[8];
  uint32_t np = 0;
  f.read(magic, 8);
  f.read((char*)&np, 4);
  if (!f || memcmp(magic, "PTCMLP1", 7) != 0 || np != NPARAM) return PTC_ERR_IO;
  std::vector<float> w(NPARAM);
  f.read((char*)w.data(), sizeof(float) * NPARAM);
  if (!f) return PTC_ERR_IO;
  ptc_mlp* m = nullptr;
  int rc = ptc_mlp_create(ctx, 0, &m);
  if (rc) return rc;
  if ((rc = ptc_mlp_set_params(m, w.data()))) {
    ptc_mlp_free(m);
    return rc;
  }
  *out = m;
  return PTC_OK;
}
