// The allocation-free variant of ptc_tc_fksCuckoo, for callers that cannot
// allocate device memory (tight serving loops, harnesses that forbid
// hipMalloc).  Every scratch array of the contraction (key probes, flop
// prefix, per-row counts, work items, split-row slabs, HBM accumulators, the
// lazily made column/row index copies of the tables) is carved from one
// file-scope device array, `scratch`, with a bump pointer; the output goes
// into arrays the caller owns.
//
// The array is one per device, for the whole process, and shared by every
// call on that device: calls that overlap on one device (two host threads,
// or two contexts on one device, with work in flight at once) race on it.
// It keeps the last call's contents; every kernel of the path clears what it
// needs cleared (counters, flags, accumulators), as it does on memory reused
// from the caching allocator.
#include <hip/hip_runtime.h>

#include <mutex>

#include "ptc_internal.h"
#include "ptc_run.h"

// 64 MiB: enough for contractions of up to about a million A nonzeros (16
// bytes of probe state each, plus per-row arrays, work items and split-row
// slabs); a larger one is PTC_ERR_RANGE.
#define PTC_NOALLOC_SCRATCH_BYTES (64ull << 20)

static __device__ __align__(256) unsigned char scratch[PTC_NOALLOC_SCRATCH_BYTES];

namespace ptc {

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch[MAX_DEVICES] = {};
}  // namespace

// At context creation, with the context's device current: the device
// address of `scratch` on that device, looked up once per device.  (In host
// code `scratch` names a host-side stand-in; only this address may be passed
// to kernels or copies.)
int scratch_setup(ptc_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= MAX_DEVICES) return PTC_OK;  // the entry point then reports it
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[ctx->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return set_err(ctx, PTC_ERR_HIP, "scratch address: %s", hipGetErrorString(e));
    }
    g_scratch[ctx->device] = p;
  }
  ctx->scratch = g_scratch[ctx->device];
  return PTC_OK;
}

}  // namespace ptc

using namespace ptc;

extern "C" int ptc_tc_fksCuckoo_noalloc(ptc_ctx* ctx, const ptc_hashtable* HT_A, const ptc_hashtable* HT_B,
                                        const ptc_dtensor* A, const ptc_dtensor* B, const ptc_options* opt,
                                        uint32_t* out_hedges, double* out_vals, uint64_t out_cap,
                                        uint64_t* nnz_out, ptc_stats* stats) {
  if (!ctx || !nnz_out) return PTC_ERR_ARG;
  *nnz_out = 0;
  if (!ctx_open(ctx)) return PTC_ERR_ARG;
  if (!ctx->scratch) return set_err(ctx, PTC_ERR_UNSUPPORTED, "no scratch array on device %d", ctx->device);
  if (ctx->arena) return set_err(ctx, PTC_ERR_ARG, "allocation-free call inside another on this context");
  ctx->arena = static_cast<unsigned char*>(ctx->scratch);
  ctx->arena_cap = PTC_NOALLOC_SCRATCH_BYTES;
  ctx->arena_used = 0;
  ctx->arena_overflow = false;
  OutInto into{out_hedges, out_vals, out_cap, nnz_out};
  int rc = tc_into(ctx, HT_A, HT_B, A, B, opt, into, stats);
  const bool over = ctx->arena_overflow;
  ctx->arena = nullptr;
  ctx->arena_cap = ctx->arena_used = 0;
  ctx->arena_overflow = false;
  if (over) {
    // work queued before the block that did not fit is harmless: it only
    // touched scratch; wait for it so a retry starts on an idle array
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipGetLastError();
    rc = set_err(ctx, PTC_ERR_RANGE, "the contraction needs more than the %llu MiB scratch of the allocation-free "
                 "entry point (use ptc_tc_fksCuckoo)", (unsigned long long)(PTC_NOALLOC_SCRATCH_BYTES >> 20));
  }
  return rc;
}
