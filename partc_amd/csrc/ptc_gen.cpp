// Deterministic synthetic COO tensors: the inputs every workload of this
// engine is measured and checked on.
//
// The reference reads real tensors from FROSTT files (main.cpp:59-65,
// mtxTnsIO.cpp:30-65) and contracts a tensor with itself
// (sparse_tensor_generator.hpp:4-88 returns B := A).  The synthetic tensors
// here have the same layout as the reference's Tensor (tensor.hpp:19-26):
// nnz x d row-major uint32 coordinates (0-based) and float64 values.
//
//  * RNG: SplitMix64 seeds a xoshiro256** stream (Blackman & Vigna 2018).
//  * Coordinates per mode: uniform over [0, dim) or truncated Zipf(alpha)
//    over ranks 1..dim mapped to index rank-1, sampled by inverse CDF.
//  * Coordinates are unique: a candidate tuple that was already drawn is
//    rejected; first occurrences are kept in draw order.  The reference
//    requires unique coordinates (fksCuckooTC.hpp:221-252 copies a slice
//    without de-duplication).
//  * Values: U(0,1] as (u53 + 1) * 2^-53.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "partc.h"
#include "ptc_internal.h"

namespace ptc {

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

Xoshiro::Xoshiro(uint64_t seed) {
  uint64_t sm = seed;
  for (int i = 0; i < 4; ++i) s[i] = splitmix64(sm);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t Xoshiro::next() {
  const uint64_t result = rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}

// uniform in [0, 1) with 53 random bits
double Xoshiro::uniform() { return (double)(next() >> 11) * 0x1.0p-53; }

// uniform integer in [0, n) by Lemire's multiply-shift (n < 2^32)
uint32_t Xoshiro::below(uint32_t n) {
  return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32);
}

}  // namespace ptc

namespace {

struct ModeSampler {
  uint32_t dim = 0;
  bool zipf = false;
  std::vector<double> cdf;  // cdf[r] = P(rank <= r+1)

  void init(uint32_t d, int dist, double alpha) {
    dim = d;
    zipf = (dist == PTC_DIST_ZIPF);
    if (!zipf) return;
    cdf.resize(d);
    double acc = 0.0;
    for (uint32_t r = 0; r < d; ++r) {
      acc += std::pow((double)(r + 1), -alpha);
      cdf[r] = acc;
    }
    for (uint32_t r = 0; r < d; ++r) cdf[r] /= acc;
    cdf[d - 1] = 1.0;
  }

  uint32_t draw(ptc::Xoshiro& rng) const {
    if (!zipf) return rng.below(dim);
    double u = rng.uniform();
    // first rank whose cdf exceeds u
    uint32_t lo = 0, hi = dim - 1;
    while (lo < hi) {
      uint32_t mid = lo + (hi - lo) / 2;
      if (cdf[mid] > u) hi = mid; else lo = mid + 1;
    }
    return lo;
  }
};

// Open-addressing set of coordinate tuples, keyed by index into the output
// coordinate array (so tuples of any width are compared exactly).
struct TupleSet {
  std::vector<uint32_t> slot;  // nonzero id + 1, 0 = empty
  uint64_t mask = 0;
  uint32_t d = 0;
  const uint32_t* coords = nullptr;

  void init(uint64_t cap_items, uint32_t d_, const uint32_t* c) {
    uint64_t cap = 64;
    while (cap < 2 * cap_items) cap <<= 1;
    slot.assign(cap, 0u);
    mask = cap - 1;
    d = d_;
    coords = c;
  }
  static uint64_t hash(const uint32_t* t, uint32_t d) {
    uint64_t h = 0x243F6A8885A308D3ull;
    for (uint32_t j = 0; j < d; ++j) {
      h ^= t[j];
      h *= 0xff51afd7ed558ccdull;
      h ^= h >> 32;
    }
    return h;
  }
  // returns true if inserted (tuple new); `id` is where the tuple was written
  bool insert(uint32_t id) {
    const uint32_t* t = coords + (uint64_t)id * d;
    uint64_t i = hash(t, d) & mask;
    while (true) {
      uint32_t s = slot[i];
      if (s == 0) {
        slot[i] = id + 1;
        return true;
      }
      const uint32_t* u = coords + (uint64_t)(s - 1) * d;
      if (std::memcmp(t, u, sizeof(uint32_t) * d) == 0) return false;
      i = (i + 1) & mask;
    }
  }
};

}  // namespace

extern "C" int ptc_gen_coo(uint64_t seed, uint32_t d, const uint32_t* dims, uint64_t nnz,
                           int dist, double alpha, uint32_t* idx_out, double* val_out) {
  if (d == 0 || d > PTC_MAX_MODES || !dims || !idx_out || !val_out) return PTC_ERR_ARG;
  if (dist != PTC_DIST_UNIFORM && dist != PTC_DIST_ZIPF) return PTC_ERR_ARG;
  if (nnz >= 0xFFFFFFFFull) return PTC_ERR_RANGE;
  long double space = 1.0L;
  for (uint32_t j = 0; j < d; ++j) {
    if (dims[j] == 0) return PTC_ERR_ARG;
    space *= (long double)dims[j];
  }
  if ((long double)nnz > space) return PTC_ERR_RANGE;
  if (nnz == 0) return PTC_OK;

  std::vector<ModeSampler> modes(d);
  for (uint32_t j = 0; j < d; ++j) modes[j].init(dims[j], dist, alpha);

  ptc::Xoshiro rng(seed);
  TupleSet set;
  set.init(nnz, d, idx_out);
  // Guard against a request the distribution cannot fill (all draws landing on
  // already-seen tuples): give up after a generous number of rejections.
  const uint64_t max_draws = 64 * nnz + 1000000;
  uint64_t draws = 0;
  uint64_t n = 0;
  while (n < nnz) {
    if (++draws > max_draws) return PTC_ERR_RANGE;
    uint32_t* t = idx_out + n * d;
    for (uint32_t j = 0; j < d; ++j) t[j] = modes[j].draw(rng);
    if (set.insert((uint32_t)n)) {
      val_out[n] = (double)((rng.next() >> 11) + 1) * 0x1.0p-53;
      ++n;
    }
  }
  return PTC_OK;
}

// Shape of one job of the batched workload: 3 or 4 modes with equal
// probability, dims log-uniform in [dim_lo, dim_hi], nnz log-uniform in
// [nnz_lo, nnz_hi] capped at half the index space.
extern "C" int ptc_gen_job_shape(uint64_t seed, uint32_t dim_lo, uint32_t dim_hi, uint64_t nnz_lo,
                                 uint64_t nnz_hi, uint32_t* d_out, uint32_t* dims_out,
                                 uint64_t* nnz_out) {
  if (!d_out || !dims_out || !nnz_out || dim_lo == 0 || dim_hi < dim_lo || nnz_lo == 0 ||
      nnz_hi < nnz_lo)
    return PTC_ERR_ARG;
  ptc::Xoshiro rng(seed ^ 0x5DEECE66Dull);
  uint32_t d = (rng.next() >> 63) ? 4u : 3u;
  long double space = 1.0L;
  for (uint32_t j = 0; j < d; ++j) {
    double u = rng.uniform();
    double v = std::exp(std::log((double)dim_lo) + u * (std::log((double)dim_hi) - std::log((double)dim_lo)));
    uint32_t dim = (uint32_t)std::floor(v);
    if (dim < dim_lo) dim = dim_lo;
    if (dim > dim_hi) dim = dim_hi;
    dims_out[j] = dim;
    space *= dim;
  }
  double u = rng.uniform();
  double v = std::exp(std::log((double)nnz_lo) + u * (std::log((double)nnz_hi) - std::log((double)nnz_lo)));
  uint64_t nnz = (uint64_t)std::floor(v);
  if (nnz < nnz_lo) nnz = nnz_lo;
  long double cap = space * 0.5L;
  if ((long double)nnz > cap) nnz = (uint64_t)cap;
  if (nnz == 0) nnz = 1;
  *d_out = d;
  *nnz_out = nnz;
  return PTC_OK;
}
