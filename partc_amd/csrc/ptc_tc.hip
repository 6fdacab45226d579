// Host orchestration of a contraction: the C ABI entry points
// ptc_tc_fksCuckoo / ptc_contract / ptc_matricize_*.
//
// Methods (all share the probe / row / work-item preprocessing):
//   TWO_PASS     symbolic pass (exact distinct columns per row) -> prefix
//                scan -> numeric pass writing straight into the output.
//   UPPER_BOUND  one numeric pass into per-row capacity min(flops, columns),
//                then a prefix scan of the true counts and a copy.
//   LEARNED      one numeric pass into per-row capacities predicted by the
//                row-size network; rows whose prediction was too small are
//                redone by the exact two-pass path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>

#include "ptc_contract.h"
#include "ptc_internal.h"
#include "ptc_learned.h"
#include "ptc_run.h"

namespace ptc {

// Output and split-row slab bytes below which a single pass over dense rows
// runs ahead of the preprocessing totals (the allocations then follow the
// bounds rows x columns and rows + DENSE_MAX_ITEMS work items).
static constexpr uint64_t AHEAD_BYTES = 512ull << 20;

struct Timer {
  ptc_ctx* ctx;
  hipEvent_t ev[8];
  int n = 0;
  explicit Timer(ptc_ctx* c) : ctx(c) {
    for (int i = 0; i < 8; ++i) ev[i] = c->ev[i];
  }
  // Timing only: a failing event call is ignored, and its error is cleared
  // so the next launch check does not report it.
  void mark() {
    if (ctx->timing && n < 8 && hipEventRecord(ev[n++], ctx->stream) != hipSuccess) (void)hipGetLastError();
  }
  double ms(int a, int b) {
    if (a >= n || b >= n) return 0.0;
    float t = 0.f;
    if (hipEventElapsedTime(&t, ev[a], ev[b]) != hipSuccess) {
      (void)hipGetLastError();
      t = 0.f;
    }
    return t;
  }
};

__global__ void k_merge_hbm_caps(uint64_t n, uint32_t* hcap, unsigned long long* gcap, const uint32_t* qh) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x)
    if (gcap[r]) {
      hcap[r] = qh[r];
      gcap[r] = qh[r];
    }
}

// Dynamic LDS a workgroup may use on this device.
uint64_t lds_limit(ptc_ctx* ctx) {
  int optin = 0;
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, ctx->device) != hipSuccess || optin <= 0) {
    (void)hipGetLastError();
    optin = (int)ctx->prop.sharedMemPerBlock;
  }
  return (uint64_t)optin > 4096 ? (uint64_t)optin - 1024 : 65536 - 1024;  // headroom for static LDS
}

int plan_hash_launch(Run& R, HashPlan& hp, const uint32_t* bound32, const unsigned long long* bound64,
                     const unsigned long long* hbm_bound) {
  ptc_ctx* ctx = R.ctx;
  if (!hp.hcap) hp.hcap = R.ws.get<uint32_t>(R.nrows);
  if (!hp.gcap) hp.gcap = R.ws.get<unsigned long long>(R.nrows + 1);
  if (!hp.hcap || !hp.gcap) return PTC_ERR_NOMEM;
  HashCapParams P{};
  P.nrows = R.nrows;
  P.bound32 = bound32;
  P.bound64 = bound64;
  P.cap_col = R.ncols;
  P.row_nchunk = R.row_nchunk;
  P.lds_slots = R.lds_slots;
  P.row_hcap = hp.hcap;
  P.gcap = hp.gcap;
  int rc = launch_hash_caps(ctx, P);
  if (rc) return rc;
  if (hbm_bound) {
    // HBM rows sized from a guaranteed bound instead (capacity mode)
    HashCapParams Q = P;
    Q.bound32 = nullptr;
    Q.bound64 = hbm_bound;
    Q.lds_slots = 0;  // every row sized as if in HBM ...
    Q.row_hcap = R.ws.get<uint32_t>(R.nrows);
    Q.gcap = R.ws.get<unsigned long long>(R.nrows + 1);
    if (!Q.row_hcap || !Q.gcap) return PTC_ERR_NOMEM;
    if ((rc = launch_hash_caps(ctx, Q))) return rc;
    // ... and merged: rows that P put in HBM take Q's size
    hipLaunchKernelGGL(k_merge_hbm_caps, dim3((unsigned)std::min<uint64_t>((R.nrows + 255) / 256, 4096)), dim3(256),
                       0, ctx->stream, R.nrows, hp.hcap, hp.gcap, (const uint32_t*)Q.row_hcap);
    PTC_HIP(ctx, hipGetLastError());
  }
  return scan_u64(ctx, hp.gcap, hp.gcap, R.nrows, hp.gcap + R.nrows);
}

int plan_hash_finish(Run& R, HashPlan& hp, unsigned long long pool) {
  hp.pool = pool;
  if (hp.pool) {
    size_t kb = R.col64 ? 8 : 4;
    hp.gkeys = R.ws.get<unsigned char>(hp.pool * kb);
    hp.gvals = R.ws.get<double>(hp.pool);
    if (!hp.gkeys || !hp.gvals) return PTC_ERR_NOMEM;
  }
  return PTC_OK;
}

int plan_hash(Run& R, HashPlan& hp, const uint32_t* bound32, const unsigned long long* bound64,
              const unsigned long long* hbm_bound) {
  int rc = plan_hash_launch(R, hp, bound32, bound64, hbm_bound);
  unsigned long long pool = 0;
  if (!rc) rc = read_back(R.ctx, &pool, hp.gcap + R.nrows, sizeof(pool));
  if (!rc) rc = plan_hash_finish(R, hp, pool);
  return rc;
}

int reset_hash_pool(Run& R, HashPlan& hp, bool values) {
  if (!hp.pool) return PTC_OK;
  int rc = R.col64 ? fill_u64(R.ctx, (unsigned long long*)hp.gkeys, hp.pool, ~0ull)
                   : fill_u32(R.ctx, (uint32_t*)hp.gkeys, hp.pool, ~0u);
  if (!rc && values) rc = fill_zero(R.ctx, hp.gvals, hp.pool * sizeof(double));
  return rc;
}

void base_params(Run& R, PassParams& P, const HashPlan* hp) {
  memset(&P, 0, sizeof(P));
  P.items = R.items;
  P.rows = R.HA->slices;
  P.row_nchunk = R.row_nchunk;
  P.row_heavy = R.row_heavy;
  P.apre = R.apre;
  P.pos_len = R.pos_len;
  P.pos_off = R.pos_off;
  P.aval = R.HA->gval;
  P.aval32 = R.HA->gval32;
  P.bcol = R.dense ? (const void*)R.bcol16 : R.col64 ? (const void*)R.HB->rest_key64 : (const void*)R.HB->rest_key32;
  P.bval = R.HB->gval;
  P.bval32 = R.HB->gval32;
  P.f32 = R.f32 ? 1 : 0;
  P.val = R.val;
  P.nb = R.HB->t->N;
  P.ncols = R.ncols;
  P.hbits = R.hbits;
  P.hacc = R.hacc;
  P.hflag = R.hflag;
  P.row_slab = R.row_slab;
  P.use_flags = R.use_flags ? 1 : 0;
  if (hp) {
    P.row_hcap = hp->hcap;
    P.row_gcap = hp->gcap;  // after the scan: offsets; a row is in HBM iff its region is non-empty
    P.row_goff = hp->gcap;
    P.gkeys = hp->gkeys;
    P.gvals = hp->gvals;
  }
  P.row_nnz = R.row_nnz;
  P.overflow = R.overflow;
  P.dO = R.dO;
  P.nrowmodes = R.nrowmodes;
  P.col_first = R.col_first ? 1 : 0;
  P.rowc = R.HA->codec;
  P.colc = R.HB->rest;
}

// After the scan gcap holds offsets; "row in HBM" must be recovered as
// offset[r+1] != offset[r].  k_pass reads row_gcap[r] as a flag, so give it
// a separate flag array.
__global__ void k_hbm_flag(const unsigned long long* __restrict__ off, uint64_t n, unsigned long long* __restrict__ flag) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x)
    flag[r] = off[r + 1] - off[r];
}

int hbm_flags(Run& R, HashPlan& hp, unsigned long long** flags) {
  unsigned long long* f = R.ws.get<unsigned long long>(R.nrows);
  if (!f) return PTC_ERR_NOMEM;
  hipLaunchKernelGGL(k_hbm_flag, dim3((unsigned)std::min<uint64_t>((R.nrows + 255) / 256, 4096)), dim3(256), 0,
                     R.ctx->stream, (const unsigned long long*)hp.gcap, R.nrows, f);
  PTC_HIP(R.ctx, hipGetLastError());
  *flags = f;
  return PTC_OK;
}

// One row pass over `items`, then the merge launch that finishes split rows
// (dense: the listed rows; hash: every row held in HBM).  merge_rows /
// nmerge default to all split rows.
static int run_pass(Run& R, int pass, HashPlan* hp, const uint4* items, uint64_t nitems, PassParams& P,
                    const uint32_t* merge_rows = nullptr, uint64_t nmerge = ~0ull) {
  unsigned long long* flags = nullptr;
  if (!R.dense) {
    int rc = hbm_flags(R, *hp, &flags);
    if (rc) return rc;
    P.row_gcap = flags;
  }
  P.items = items;
  size_t lds = pass_lds_bytes(pass, R.dense, R.col64, R.ncols, R.lds_slots, R.use_flags, R.threads,
                              pass == PASS_NUMERIC ? R.val : 0);
  // one pair of timing events brackets the row launch and its merge launch
  ptc_ctx* ctx = R.ctx;
  const bool timed = ctx->timing && ctx->nkev + 2 <= 16;
  if (timed && hipEventRecord(ctx->kev[ctx->nkev], ctx->stream) != hipSuccess) (void)hipGetLastError();
  int rc = launch_pass(ctx, pass, R.dense, R.col64, P, nitems, lds, false, R.threads);
  if (!rc && R.dense) {
    if (!merge_rows) {
      merge_rows = R.heavy_rows;
      nmerge = R.nheavy;
    }
    P.merge_rows = merge_rows;
    rc = launch_pass(ctx, pass, true, R.col64, P, nmerge, lds, true, R.threads);
  } else if (!rc && pass == PASS_NUMERIC && hp->pool) {
    P.merge_rows = merge_rows;  // NULL: every row (rows not in HBM exit at once)
    rc = launch_pass(ctx, pass, false, R.col64, P, merge_rows ? nmerge : R.nrows, lds, true, R.threads);
  }
  if (rc) return rc;
  if (timed) {
    if (hipEventRecord(ctx->kev[ctx->nkev + 1], ctx->stream) != hipSuccess) (void)hipGetLastError();
    ctx->kev_pass[ctx->nkev / 2] = pass;
    ctx->nkev += 2;
  }
  return PTC_OK;
}

// Slabs of split rows: one partial per chunk (a bitmap in the symbolic pass,
// a dense accumulator in the numeric pass); every chunk writes its whole
// slab, so nothing is cleared.
int prepare_heavy(Run& R, int pass) {
  if (!R.dense || !R.nheavy) return PTC_OK;
  if (pass == PASS_SYMBOLIC) {
    const uint64_t nbits = (R.ncols + 31) / 32;
    if (!R.hbits) R.hbits = R.ws.get<uint32_t>(R.nslabs * nbits);
    return R.hbits ? PTC_OK : PTC_ERR_NOMEM;
  }
  if (!R.hacc) R.hacc = R.ws.get<double>(R.nslabs * R.ncols);
  if (R.use_flags && !R.hflag) R.hflag = R.ws.get<unsigned char>(R.nslabs * R.ncols);
  if (!R.hacc || (R.use_flags && !R.hflag)) return PTC_ERR_NOMEM;
  return PTC_OK;
}

int make_output(Run& R, uint64_t nnz, const ptc_dtensor* A, const ptc_dtensor* B, ptc_dtensor** O) {
  ptc_dtensor* o = new ptc_dtensor();
  o->ctx = R.ctx;
  ctx_acquire(R.ctx);
  o->d = R.dO;
  o->nk = R.dO;  // tensor.hpp:124
  o->N = nnz;
  for (uint32_t j = 0; j < R.dO; ++j) {
    o->dim_tc[j] = j;  // identity (tensor.hpp:128-129)
    o->dims[j] = 0;    // zeros (tensor.hpp:121-122)
  }
  // per-mode index bounds (used as key radices if O is contracted again)
  const uint32_t na = A->nk, nb = B->d - B->nk;  // row modes, column modes
  const uint32_t ra = R.col_first ? nb : 0, cb = R.col_first ? 0 : na;  // where each block starts
  for (uint32_t j = 0; j < na; ++j) o->mode_max[ra + j] = A->mode_max[A->dim_tc[j]];
  for (uint32_t j = 0; j < nb; ++j) o->mode_max[cb + j] = B->mode_max[B->dim_tc[B->nk + j]];
  o->hedges = dalloc_n<uint32_t>(R.ctx, nnz * R.dO);
  if (R.f32) {
    o->vals32 = dalloc_n<float>(R.ctx, nnz);
    o->err32 = dalloc_n<float>(R.ctx, nnz);
  } else {
    o->vals = dalloc_n<double>(R.ctx, nnz);
  }
  if (!o->hedges || !(o->vals || (o->vals32 && o->err32))) {
    ptc_dtensor_free(o);
    return PTC_ERR_NOMEM;
  }
  *O = o;
  return PTC_OK;
}

}  // namespace ptc

using namespace ptc;

extern "C" void ptc_options_default(ptc_options* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->method = PTC_METHOD_TWO_PASS;
  o->value_type = PTC_VAL_F64;
  o->acc_mode = PTC_ACC_AUTO;
  o->predict_scale = 1.15;
}

namespace ptc {

// Validation, accumulator choice and the shared preprocessing: key probes,
// flop prefix, per-row chunks, work items.
int prepare_launch(Run& R, const ptc_hashtable* HT_A, const ptc_hashtable* HT_B, const ptc_dtensor* A,
                   const ptc_dtensor* B, const ptc_options& opt, unsigned long long* tot) {
  ptc_ctx* ctx = R.ctx;
  if (HT_A->t != A || HT_B->t != B) return set_err(ctx, PTC_ERR_ARG, "tables were not built on these tensors");
  const uint32_t c = A->d - A->nk;
  if (c == 0 || B->nk != c) return set_err(ctx, PTC_ERR_ARG, "A has %u contracted modes, B %u key modes", c, B->nk);
  if (HT_A->rest.n != c || HT_B->codec.n != c) return set_err(ctx, PTC_ERR_INTERNAL, "codec shape mismatch");
  if (opt.deterministic) return set_err(ctx, PTC_ERR_UNSUPPORTED, "the deterministic (reference-order) mode is not available");
  if (opt.chunk_flops > MAX_ITEM_FLOPS)
    return set_err(ctx, PTC_ERR_RANGE, "chunk_flops %llu exceeds 2^31 multiply-adds per work item",
                   (unsigned long long)opt.chunk_flops);
  {
    const bool a32 = HT_A->gval32 != nullptr, b32 = HT_B->gval32 != nullptr;
    if (a32 != b32) return set_err(ctx, PTC_ERR_ARG, "A and B must both hold FP64 or both FP32 values");
    if (opt.value_type < PTC_VAL_F64 || opt.value_type > PTC_VAL_F32_F64ACC)
      return set_err(ctx, PTC_ERR_ARG, "unknown value type %d", opt.value_type);
    if (opt.acc_mode < PTC_ACC_AUTO || opt.acc_mode > PTC_ACC_WINDOW)
      return set_err(ctx, PTC_ERR_ARG, "unknown accumulator mode %d", opt.acc_mode);
    if (opt.value_type == PTC_VAL_F64 && a32)
      return set_err(ctx, PTC_ERR_ARG, "FP32 tensors need an FP32 value type (PTC_VAL_F32_*)");
    if (opt.value_type != PTC_VAL_F64 && !a32)
      return set_err(ctx, PTC_ERR_ARG, "FP32 value types need FP32 tensors (ptc_dtensor_upload_f32)");
    R.f32 = a32;
    R.val = opt.value_type;
  }
  if (opt.method == PTC_METHOD_LEARNED && !opt.mlp && !ptc_learned_default())
    return set_err(ctx, PTC_ERR_ARG, "learned method needs a trained row-size model");
  PTC_HIP(ctx, hipSetDevice(ctx->device));

  R.HA = HT_A;
  R.HB = HT_B;
  R.opt = opt;
  R.npos = A->N;
  // rows: HT_A's key count when known, else its bound; the exact count is
  // read back with the preprocessing totals below
  R.nrows = HT_A->keys_known ? HT_A->num_keys : HT_A->slices_cap;
  R.ncols = HT_B->rest.space;
  R.col64 = R.ncols > 0xFFFFFFFFull;
  R.nrowmodes = A->nk;
  R.dO = A->nk + (B->d - B->nk);  // tensor.hpp:113
  R.use_flags = R.force_flags || !(A->vals_safe && B->vals_safe);
  if (R.dO == 0) return set_err(ctx, PTC_ERR_EMPTY, "tensor contraction not possible!! (output has no modes)");
  // No nonzero on one side: no pair contributes (what the flop total read
  // back below would show).  Returned here because an A table known to hold
  // no key has no row for the row passes to read.
  if (A->N == 0 || B->N == 0 || (HT_A->keys_known && HT_A->num_keys == 0))
    return set_err(ctx, PTC_ERR_EMPTY, "tensor contraction not possible!! (no contributing pair)");
  // largest dense column space whose numeric pass fits the workgroup's LDS
  const uint64_t lds_cap = lds_limit(ctx);
  uint64_t dense_fit = 0;
  while (dense_fit < 65536 && pass_lds_bytes(PASS_NUMERIC, true, false, dense_fit + 256, 0, true, PB, R.val) <= lds_cap)
    dense_fit += 256;
  uint64_t dense_max = opt.dense_max_cols ? opt.dense_max_cols : DEFAULT_DENSE_MAX;
  dense_max = std::min<uint64_t>(dense_max, dense_fit);
  R.dense = opt.acc_mode == PTC_ACC_DENSE ? R.ncols <= dense_fit
          : opt.acc_mode == PTC_ACC_AUTO ? R.ncols <= dense_max : false;
  if (opt.acc_mode == PTC_ACC_DENSE && !R.dense) return set_err(ctx, PTC_ERR_ARG, "column space too large for a dense accumulator");
  // Dense column windows: a 16-bit column space too wide for one dense
  // accumulator (AUTO), or any when asked for.  The widest window whose
  // numeric pass fits a 512-thread workgroup's LDS (touched bits kept),
  // capped by dense_max_cols when windows are asked for explicitly.
  if (opt.acc_mode == PTC_ACC_WINDOW || (opt.acc_mode == PTC_ACC_AUTO && !R.dense)) {
    uint64_t wfit = 0;
    while (wfit < 65536 && pass_lds_bytes(PASS_NUMERIC, true, false, wfit + 256, 0, true, 512, R.val) <= lds_cap)
      wfit += 256;
    if (opt.acc_mode == PTC_ACC_WINDOW && opt.dense_max_cols)
      wfit = std::min<uint64_t>(wfit, std::max<uint64_t>(64, opt.dense_max_cols / 64 * 64));
    const uint64_t nw = wfit ? (R.ncols + wfit - 1) / wfit : 0;
    const bool ok = !R.no_windows && !R.col64 && R.ncols <= 65536 && nw >= 1 && nw <= MAX_WINDOWS &&
                    (opt.acc_mode == PTC_ACC_WINDOW || R.ncols > wfit);
    if (opt.acc_mode == PTC_ACC_WINDOW && !ok)
      return set_err(ctx, PTC_ERR_ARG, "column space of %llu not held by %u dense windows below 2^16 columns",
                     (unsigned long long)R.ncols, MAX_WINDOWS);
    if (ok) {
      R.win_n = (uint32_t)nw;
      R.win_width = (uint32_t)(((R.ncols + nw - 1) / nw + 63) / 64 * 64);
      const size_t at512 = pass_lds_bytes(PASS_NUMERIC, true, false, R.win_width, 0, true, 512, R.val);
      R.win_threads = opt.block_threads ? opt.block_threads
                    : (2 * at512 > lds_cap &&
                       pass_lds_bytes(PASS_NUMERIC, true, false, R.win_width, 0, true, 1024, R.val) <= lds_cap)
                        ? 1024u : 512u;
      if (R.win_threads < 64 || R.win_threads > (uint32_t)PB_MAX || R.win_threads % 64 ||
          pass_lds_bytes(PASS_NUMERIC, true, false, R.win_width, 0, R.use_flags, R.win_threads, R.val) > lds_cap)
        return set_err(ctx, PTC_ERR_ARG, "block_threads %u for %u-column windows", R.win_threads, R.win_width);
    }
  }
  R.lds_slots = opt.lds_hash_max ? opt.lds_hash_max : DEFAULT_LDS_HASH;
  if (opt.acc_mode == PTC_ACC_SPILL) R.lds_slots = 0;
  {
    uint32_t p2 = 16;
    while (p2 * 2 <= R.lds_slots && p2 < 8192) p2 <<= 1;
    R.lds_slots = R.lds_slots ? p2 : 0;
    while (R.lds_slots > 16 && pass_lds_bytes(PASS_NUMERIC, false, R.col64, 0, R.lds_slots) > lds_cap) R.lds_slots >>= 1;
  }
  // Threads per row-pass workgroup.  The accumulator's LDS is paid once per
  // workgroup, so when a 512-thread workgroup's footprint leaves room for
  // fewer than two per CU, 1024 threads keep twice the waves in flight on
  // the same LDS.
  {
    const size_t at512 = pass_lds_bytes(PASS_NUMERIC, R.dense, R.col64, R.ncols, R.lds_slots, R.use_flags, 512, R.val);
    uint32_t t = opt.block_threads ? opt.block_threads : (2 * at512 > lds_cap ? 1024u : 512u);
    if (t < 64 || t > (uint32_t)PB_MAX || t % 64) return set_err(ctx, PTC_ERR_ARG, "block_threads %u", t);
    while (t > 512 && pass_lds_bytes(PASS_NUMERIC, R.dense, R.col64, R.ncols, R.lds_slots, R.use_flags, t, R.val) > lds_cap)
      t /= 2;
    R.threads = t;
  }
  // Work per workgroup.  Dense: a split row writes one dense slab (8 bytes a
  // column) per chunk, so chunks of at least 24 x columns multiply-adds
  // (12 bytes streamed each) keep the slabs within a few percent of the
  // streamed bytes; 65536 is the floor (measured on the nips-shape and
  // two-tensor workloads).  Hash rows keep smaller chunks: their HBM
  // accumulators are shared by all chunks of a row.
  // Windows: one item runs once per window, its slabs a window wide.
  const uint64_t chunk = opt.chunk_flops ? opt.chunk_flops
                       : R.dense ? std::max<uint64_t>(DEFAULT_DENSE_CHUNK, 24 * R.ncols)
                       : R.win_n ? std::max<uint64_t>(DEFAULT_DENSE_CHUNK, 24 * (uint64_t)R.win_width)
                                 : DEFAULT_CHUNK;

  int rc = PTC_OK;
  if ((R.dense || R.win_n) && (rc = hashtable_col16(ctx, HT_B, &R.bcol16))) return rc;
  // HT_A's slice index of every A position: only the per-row features of
  // the training profile read it
  const uint32_t* row_of_pos = nullptr;
  if (R.want_features && (rc = hashtable_slice_of(ctx, HT_A, &row_of_pos))) return rc;
  R.row_of_pos = row_of_pos;
  Workspace& ws = R.ws;
  R.pos_len = ws.get<uint32_t>(R.npos + 4);  // + the rest of the last 16-byte word (read whole by the learned features)
  R.pos_off = ws.get<uint32_t>(R.npos);
  R.apre = ws.get<unsigned long long>(R.npos + 1);
  R.row_hits = ws.get<uint32_t>(R.nrows);
  R.row_maxlen = ws.get<uint32_t>(R.nrows);
  R.row_nchunk = ws.get<uint32_t>(R.nrows);
  R.row_heavy = ws.get<uint32_t>(R.nrows);
  R.row_nnz = ws.get<uint32_t>(R.nrows + 1);
  R.row_flops = ws.get<unsigned long long>(R.nrows);
  R.packed = ws.get<unsigned long long>(R.nrows + 1);
  R.row_slab = ws.get<unsigned long long>(R.nrows + 1);
  R.overflow = ws.get<uint32_t>(4);
  R.out_off = ws.get<unsigned long long>(R.nrows + 1);
  if (!R.pos_len || !R.pos_off || !R.apre || !R.row_hits || !R.row_maxlen || !R.row_nchunk || !R.row_heavy ||
      !R.row_nnz || !R.row_flops || !R.packed || !R.row_slab || !R.overflow || !R.out_off)
    return PTC_ERR_NOMEM;

  // ---------------- preprocessing: probe, flop prefix, rows, items
  // per-row features as arrays (the training profile); the learned method's
  // capacity kernel finds them itself from the positions and tile parts
  const bool features = R.want_features;
  const bool tile_parts = opt.method == PTC_METHOD_LEARNED && !features;
  if (!R.npos && (rc = fill_zero(ctx, R.overflow, 16))) return rc;  // else cleared by the probe
  {
    ProbeParams P{};
    P.npos = R.npos;
    P.akey = HT_A->rest_key64 ? (const void*)HT_A->rest_key64 : (const void*)HT_A->rest_key32;
    P.acodec = HT_A->rest;
    P.bcodec = HT_B->codec;
    P.same_codec = 1;
    for (uint32_t j = 0; j < P.acodec.n; ++j)
      if (P.acodec.radix[j] != P.bcodec.radix[j]) P.same_codec = 0;
    P.tab = HT_B->entries;
    P.cap = HT_B->cap;
    P.direct = HT_B->direct ? 1 : 0;
    P.pos_len = R.pos_len;
    P.pos_off = R.pos_off;
    P.zero4 = R.overflow;
    if ((rc = launch_probe(ctx, P, HT_A->rest_key64 != nullptr))) return rc;
  }
  // the flop prefix over A positions; the row features ride on the same pass
  FeatPart* parts = nullptr;
  TilePart* tiles = nullptr;
  if (tile_parts) {
    tiles = ws.get<TilePart>(2 * ((R.npos + 2 * TILE_PART - 1) / (2 * TILE_PART)) + 2);
    if (!tiles) return PTC_ERR_NOMEM;
    if ((rc = scan_u32_tiles(ctx, R.pos_len, R.apre, R.npos, R.apre + R.npos, tiles))) return rc;
  } else if (features) {
    parts = ws.get<FeatPart>((R.npos + FEAT_TILE - 1) / FEAT_TILE + 1);
    if (!parts) return PTC_ERR_NOMEM;
    if ((rc = scan_u32_features(ctx, R.pos_len, R.apre, R.npos, R.apre + R.npos, row_of_pos, R.row_hits,
                                R.row_maxlen, parts)))
      return rc;
  } else if ((rc = scan_u32_to_u64(ctx, R.pos_len, R.apre, R.npos, R.apre + R.npos))) {
    return rc;
  }
  {
    RowParams P{};
    P.nrows = R.nrows;
    P.live = HT_A->keys_known ? nullptr : HT_A->d_total;
    P.rows = HT_A->slices;
    P.apre = R.apre;
    P.chunk = chunk;
    // dense rows: at least total / DENSE_MAX_ITEMS multiply-adds per item.
    // Measured on MI355X: the two-tensor numeric pass (2.6e9 multiply-adds,
    // 4 waves per SIMD) takes 4.93 ms with 192000 per item, 4.60 ms from
    // 524288 to 2097152; the nips-shape pass (1.4e8) is fastest at its
    // 65536 floor (0.326 ms; 0.508 ms at 131072: too few items for the CUs).
    if ((R.dense || R.win_n) && !opt.chunk_flops) {
      P.total = R.apre + R.npos;
      P.max_items = DENSE_MAX_ITEMS;
    }
    P.row_nchunk = R.row_nchunk;
    P.row_flops = R.row_flops;
    P.packed = R.packed;
    P.slabs = R.row_slab;
    P.row_hits = parts ? R.row_hits : nullptr;
    P.row_maxlen = parts ? R.row_maxlen : nullptr;
    // Dense rows with the default work-item rule have at most rows +
    // DENSE_MAX_ITEMS items (each row's chunk count is at most its flops /
    // chunk + 1 and the chunk is at least total / DENSE_MAX_ITEMS): the items
    // are built now, in arrays of that bound, so the GPU has them to do while
    // the host waits for the totals.  A few thousand rows take one
    // workgroup for their flops, scans, items and upper-bound capacities.
    R.early_items = (R.dense || R.win_n) && !opt.chunk_flops;
    if (R.early_items && R.nrows + 1 <= ROWS_SMALL_MAX) {
      ItemParams Q;
      ItemP0Params Q0;
      if ((rc = item_params(R, R.nrows + DENSE_MAX_ITEMS, R.nrows, R.npos, Q, Q0))) return rc;
      if (opt.method == PTC_METHOD_UPPER_BOUND) {
        R.cap = ws.get<uint32_t>(R.nrows + 1);
        if (!R.cap) return PTC_ERR_NOMEM;
      }
      uint32_t nzero = 0;
      if (opt.method != PTC_METHOD_TWO_PASS) {  // the single pass's counts, cleared here
        nzero = (uint32_t)(count_words(R.nrows) + 4);
        if (!(R.cnt_block = ws.get<uint32_t>(nzero))) return PTC_ERR_NOMEM;
      }
      Q0.find_rows = 1;
      Q0.nrows = R.nrows;
      if ((rc = launch_rows_small(ctx, P, Q, R.cap, R.ncols, R.cnt_block, nzero)) ||
          (rc = launch_item_starts(ctx, Q0)))
        return rc;
    } else {
      if ((rc = launch_rows(ctx, P)) || (rc = scan_u64_two_inplace(ctx, R.packed, R.row_slab, R.nrows))) return rc;
      if (R.early_items && (rc = build_items(R, R.nrows + DENSE_MAX_ITEMS, R.nrows, R.npos))) return rc;
    }
  }
  // A single pass over dense rows whose output and split-row slabs have
  // host-known bounds within AHEAD_BYTES runs ahead of these totals (one
  // round trip for the whole contraction); its capacities then bound rows,
  // not the output, and need no sum.
  R.ahead = false;
  if (R.allow_ahead && R.early_items && R.dense && !R.win_n && R.ncols && R.nrows <= (1ull << 40) / R.ncols) {
    const uint64_t out_bytes = R.nrows * R.ncols * (4 * R.dO + 8);
    const uint64_t slab_bytes = (R.nrows + DENSE_MAX_ITEMS) * R.ncols * (R.f32 ? 12 : 8 + (R.use_flags ? 1 : 0));
    R.ahead = out_bytes <= AHEAD_BYTES && slab_bytes <= AHEAD_BYTES;
  }
  // Single-pass methods: every row's capacity (rows past HT_A's key count
  // have no flops: capacity 0) and their sum, read back with the totals
  // below, so sizing the output costs no round trip of its own.  Upper
  // bound: min(flops, columns); learned: features, inference and the
  // clamped prediction in one launch.
  const bool single = opt.method != PTC_METHOD_TWO_PASS;
  if (single) {
    const bool caps_done = R.cap != nullptr;  // written with the rows (upper bound, few rows)
    if (!R.cap) R.cap = ws.get<uint32_t>(R.nrows + 1);
    R.cap_off = ws.get<unsigned long long>(R.nrows + 1);
    if (!R.cap || !R.cap_off) return PTC_ERR_NOMEM;
    const double scale = opt.predict_scale > 0 ? opt.predict_scale : 1.15;
    if (opt.method == PTC_METHOD_LEARNED) {
      LearnedInputs L{};
      L.nrows = R.nrows;
      L.rows = HT_A->slices;
      L.row_flops = R.row_flops;
      L.row_hits = parts ? R.row_hits : nullptr;
      L.row_maxlen = parts ? R.row_maxlen : nullptr;
      L.ncols = R.ncols;
      L.pos_len = R.pos_len;
      L.tiles = tiles;
      if ((rc = learned_caps(ctx, opt.mlp, L, scale, R.cap))) return rc;
    } else {
      CapParams P{};
      P.nrows = R.nrows;
      P.row_flops = R.row_flops;
      P.row_maxlen = R.row_maxlen;
      P.ncols = R.ncols;
      P.pred = nullptr;
      P.scale = scale;
      P.cap = R.cap;
      if (!caps_done && (rc = launch_caps(ctx, P))) return rc;
    }
    if (!R.ahead && (rc = scan_u32_to_u64(ctx, R.cap, R.cap_off, R.nrows, R.cap_off + R.nrows))) return rc;
  }
  // the totals and HT_A's key count, for one host round trip (tot NULL: the
  // caller collects them from R.totw with words of its own)
  {
    WordSources& W = R.totw;
    W = WordSources{};
    W.n = single && !R.ahead ? 5 : 4;
    W.src[0] = R.apre + R.npos;
    W.src[1] = R.packed + R.nrows;
    W.src[2] = R.row_slab + R.nrows;
    W.src[3] = HT_A->d_total;
    if (W.n == 5) W.src[4] = R.cap_off + R.nrows;
    for (uint32_t i = 0; i < W.n; ++i) W.bytes[i] = 8;
    if (tot && (rc = collect_words(ctx, tot, W))) return rc;
  }
  return PTC_OK;
}

int prepare_finish(Run& R, const unsigned long long* hostv) {
  ptc_ctx* ctx = R.ctx;
  const ptc_hashtable* HT_A = R.HA;
  const ptc_dtensor* A = HT_A->t;
  Workspace& ws = R.ws;
  int rc = PTC_OK;
  if (!HT_A->keys_known) {
    if ((hostv[3] & 0xFFFFFFFFull) != A->N)
      return set_err(ctx, PTC_ERR_INTERNAL, "key table grouped %llu of %llu nonzeros",
                     (unsigned long long)(hostv[3] & 0xFFFFFFFFull), (unsigned long long)A->N);
    ptc_hashtable* ha = const_cast<ptc_hashtable*>(HT_A);
    ha->num_keys = hostv[3] >> 32;
    ha->keys_known = true;
  }
  R.nrows = HT_A->num_keys;
  R.nslabs = hostv[2];
  R.F = hostv[0];
  if (R.cap && !R.ahead) R.capsum = hostv[4];
  R.nitems = hostv[1] & 0xFFFFFFFFull;
  R.nheavy = hostv[1] >> 32;
  if (R.F == 0) return set_err(ctx, PTC_ERR_EMPTY, "tensor contraction not possible!! (no contributing pair)");
  if (R.early_items) {  // built before the read-back, within their bound
    if (R.nitems > R.nrows + DENSE_MAX_ITEMS)
      return set_err(ctx, PTC_ERR_INTERNAL, "%llu work items past their bound", (unsigned long long)R.nitems);
    return PTC_OK;
  }
  return build_items(R, R.nitems, R.nheavy, R.nheavy ? R.npos : 0);
}

// Work items (row, chunk, first position) and the split rows' list into
// arrays of nitems and nheavy entries; npos = 0 skips the chunk starts.
uint64_t count_words(uint64_t nrows) { return (nrows + 1 + 3) / 4 * 4; }  // 16-byte aligned end of nrows + 1 counts

int item_params(Run& R, uint64_t nitems, uint64_t nheavy, uint64_t npos, ItemParams& P, ItemP0Params& Q) {
  const ptc_hashtable* HT_A = R.HA;
  R.items = R.ws.get<uint4>(std::max<uint64_t>(nitems, 1));
  R.heavy_rows = R.ws.get<uint32_t>(std::max<uint64_t>(nheavy, 1));
  if (!R.items || !R.heavy_rows) return PTC_ERR_NOMEM;
  P = ItemParams{};
  P.nrows = R.nrows;
  P.rows = HT_A->slices;
  P.row_nchunk = R.row_nchunk;
  P.packed = R.packed;
  P.row_heavy = R.row_heavy;
  P.heavy_rows = R.heavy_rows;
  P.items = R.items;
  Q = ItemP0Params{};
  Q.nitems = nitems;
  Q.live = R.early_items ? R.packed + R.nrows : nullptr;
  Q.npos = npos;
  Q.row_nchunk = R.row_nchunk;
  Q.rows = HT_A->slices;
  Q.apre = R.apre;
  Q.row_flops = R.row_flops;
  Q.packed = R.packed;
  Q.items = R.items;
  return PTC_OK;
}

int build_items(Run& R, uint64_t nitems, uint64_t nheavy, uint64_t npos) {
  ItemParams P;
  ItemP0Params Q;
  const int rc = item_params(R, nitems, nheavy, npos, P, Q);
  return rc ? rc : launch_items(R.ctx, P, Q);
}

int prepare(Run& R, const ptc_hashtable* HT_A, const ptc_hashtable* HT_B, const ptc_dtensor* A,
            const ptc_dtensor* B, const ptc_options& opt) {
  unsigned long long* tot = R.ws.get<unsigned long long>(5);
  if (!tot) return PTC_ERR_NOMEM;
  int rc = prepare_launch(R, HT_A, HT_B, A, B, opt, tot);
  unsigned long long hostv[5] = {0, 0, 0, 0, 0};
  if (!rc) rc = read_back(R.ctx, hostv, tot, R.cap ? 40 : 32);
  if (!rc) rc = prepare_finish(R, hostv);
  return rc;
}

}  // namespace ptc

namespace ptc {

// Output arrays of o reallocated for cap entries, the first keep entries
// copied over (a single-pass output that outgrew its capacities, or one
// whose capacities overestimated it by more than half).
static int resize_output(Run& R, ptc_dtensor* o, uint64_t keep, uint64_t cap) {
  ptc_ctx* ctx = R.ctx;
  uint32_t* h = dalloc_n<uint32_t>(ctx, cap * R.dO);
  double* v = R.f32 ? nullptr : dalloc_n<double>(ctx, cap);
  float* v32 = R.f32 ? dalloc_n<float>(ctx, cap) : nullptr;
  float* e32 = R.f32 ? dalloc_n<float>(ctx, cap) : nullptr;
  if (!h || !(v || (v32 && e32))) {
    dfree(ctx, h);
    dfree(ctx, v);
    dfree(ctx, v32);
    dfree(ctx, e32);
    return PTC_ERR_NOMEM;
  }
  if (keep) {
    PTC_HIP(ctx, hipMemcpyAsync(h, o->hedges, sizeof(uint32_t) * keep * R.dO, hipMemcpyDeviceToDevice, ctx->stream));
    if (R.f32) {
      PTC_HIP(ctx, hipMemcpyAsync(v32, o->vals32, sizeof(float) * keep, hipMemcpyDeviceToDevice, ctx->stream));
      PTC_HIP(ctx, hipMemcpyAsync(e32, o->err32, sizeof(float) * keep, hipMemcpyDeviceToDevice, ctx->stream));
    } else {
      PTC_HIP(ctx, hipMemcpyAsync(v, o->vals, sizeof(double) * keep, hipMemcpyDeviceToDevice, ctx->stream));
    }
  }
  // the old blocks are only reused by work queued after these copies
  dfree(ctx, o->hedges);
  dfree(ctx, o->vals);
  dfree(ctx, o->vals32);
  dfree(ctx, o->err32);
  o->hedges = h;
  o->vals = v;
  o->vals32 = v32;
  o->err32 = e32;
  return PTC_OK;
}

// The single-pass methods with direct writes: the output arrays are sized by
// the sum of the rows' capacities (read back with the preprocessing totals),
// and every row that finishes within its capacity reserves its entries with
// one fetch-add and writes them in place -- no staging area, no count scan,
// no copy.  Rows over their capacity (or whose LDS table filled) are redone
// by the exact two-pass path, appended the same way.
// Ahead mode: the pass runs before the preprocessing totals
// reach the host -- dense rows with the default work-item rule, whose items
// are already built within their bound -- on a grid of that bound (work
// past the device's count exits), with split-row slabs and the output sized
// by bounds known on the host (rows x columns); the totals come back with
// the pass's own flags and count, one round trip for the whole contraction.
static int single_pass_direct(Run& R, const ptc_dtensor* A, const ptc_dtensor* B, const uint32_t* cap,
                              const unsigned long long* cap_off, int method, Timer& T, ptc_dtensor** out_p,
                              uint64_t* nnz_p, uint64_t* fallback_p, uint64_t* capacity_p, bool ahead = false) {
  ptc_ctx* ctx = R.ctx;
  Workspace& ws = R.ws;
  int rc = PTC_OK;
  // the capacities' sum, read back with the preprocessing totals; ahead: the
  // bound rows x columns, which every row's capacity min(flops, columns) sums below
  unsigned long long capsum = ahead ? R.nrows * R.ncols : R.capsum;
  uint64_t nitems = R.nitems;
  if (ahead) {  // bounds of the counts the pass reads on the device
    nitems = R.nrows + DENSE_MAX_ITEMS;
    R.nheavy = R.nrows;
    R.nslabs = nitems;
  }
  *capacity_p = capsum;
  ptc_dtensor* out = nullptr;
  if ((rc = make_output(R, capsum, A, B, &out))) return rc;
  auto fail = [&](int code) {
    ptc_dtensor_free(out);
    return code;
  };
  // the reservation counter rides behind the row counts: one clear for
  // both (done by the preprocessing when it allocated them)
  const uint64_t cnt_words = count_words(R.nrows);
  const bool cleared = R.cnt_block != nullptr;
  uint32_t* row_nnz = cleared ? R.cnt_block : ws.get<uint32_t>(cnt_words + 4);
  unsigned long long* pair = ws.get<unsigned long long>(8);
  if (!row_nnz || !pair) return fail(PTC_ERR_NOMEM);
  R.row_nnz = row_nnz;
  unsigned long long* counter = reinterpret_cast<unsigned long long*>(row_nnz + cnt_words);
  HashPlan h1;
  if ((!R.dense && (rc = plan_hash(R, h1, cap, nullptr, R.row_flops))) ||
      (!cleared && (rc = fill_zero(ctx, row_nnz, (cnt_words + 4) * 4))) || (rc = prepare_heavy(R, PASS_NUMERIC)) ||
      (!R.dense && (rc = reset_hash_pool(R, h1, true))))
    return fail(rc);
  auto direct = [&](PassParams& P, const uint32_t* caps) {
    P.out_cap = caps;
    P.out_count = counter;
    P.out_limit = capsum;
    P.out_rowoff = R.out_off;
    P.out_hedges = out->hedges;
    P.out_vals = out->vals;
    P.out_vals32 = out->vals32;
    P.out_err32 = out->err32;
  };
  PassParams P;
  base_params(R, P, R.dense ? nullptr : &h1);
  direct(P, cap);
  if (ahead) P.live_items = R.packed + R.nrows;
  if ((rc = run_pass(R, PASS_NUMERIC, &h1, R.items, nitems, P))) return fail(rc);
  T.mark();  // 2
  // the consistency flags, the overflow count and the entries written (and,
  // ahead, the preprocessing totals): one read
  unsigned long long rb[7] = {0, 0, 0, 0, 0, 0, 0};
  {
    WordSources W{};
    W.n = 2;
    W.src[0] = R.overflow;
    W.bytes[0] = 4;
    W.src[1] = counter;
    W.bytes[1] = 8;
    if (ahead) {  // and the preprocessing totals
      for (uint32_t i = 0; i < R.totw.n; ++i) {
        W.src[W.n] = R.totw.src[i];
        W.bytes[W.n++] = R.totw.bytes[i];
      }
    }
    if ((rc = collect_words(ctx, pair, W)) || (rc = read_back(ctx, rb, pair, 8 * W.n))) return fail(rc);
    if (ahead && (rc = prepare_finish(R, rb + 2))) return fail(rc);
  }
  const uint32_t ovf0 = (uint32_t)rb[0];
  if (ovf0 & 0x80000000u) return fail(set_err(ctx, PTC_ERR_INTERNAL, "accumulator table overflow"));
  if (ovf0 & 0x40000000u) return fail(set_err(ctx, PTC_ERR_INTERNAL, "slice index out of range"));
  if (ovf0 & 0x20000000u) return fail(set_err(ctx, PTC_ERR_INTERNAL, "output reservation past the capacities"));
  const uint64_t fallback = ovf0 & 0x1FFFFFFFu;
  uint64_t nnz = rb[1];
  if (fallback) {
    // exact two-pass for the rows whose capacity was too small
    uint32_t* cnt = ws.get<uint32_t>(4);
    uint4* fitems = ws.get<uint4>(R.nitems);
    uint32_t* frows = ws.get<uint32_t>(fallback);
    uint32_t* row_fb = ws.get<uint32_t>(R.nrows);
    unsigned long long* tot = ws.get<unsigned long long>(R.nrows + 1);
    if (!cnt || !fitems || !frows || !row_fb || !tot) return fail(PTC_ERR_NOMEM);
    uint32_t nfitems = 0;
    if ((rc = fill_zero(ctx, cnt, 16)) || (rc = launch_select_items(ctx, R.items, R.nitems, R.row_nnz, cap, fitems, cnt)) ||
        (rc = launch_overflow_rows(ctx, R.row_nnz, cap, R.nrows, frows, cnt + 1, row_fb)) ||
        (rc = read_back(ctx, &nfitems, cnt, 4)) || (rc = launch_reset_rows(ctx, frows, (uint32_t)fallback, R.row_nnz)))
      return fail(rc);
    HashPlan hs;
    if ((!R.dense && (rc = plan_hash(R, hs, nullptr, R.row_flops, nullptr))) || (rc = prepare_heavy(R, PASS_SYMBOLIC)) ||
        (!R.dense && (rc = reset_hash_pool(R, hs, false))))
      return fail(rc);
    PassParams S;
    base_params(R, S, R.dense ? nullptr : &hs);
    if ((rc = run_pass(R, PASS_SYMBOLIC, &hs, fitems, nfitems, S, frows, fallback))) return fail(rc);
    // every row's exact count: the output's size
    unsigned long long total = 0;
    if ((rc = scan_u32_to_u64(ctx, R.row_nnz, tot, R.nrows, tot + R.nrows)) ||
        (rc = read_back(ctx, &total, tot + R.nrows, 8)))
      return fail(rc);
    if (total > capsum) {
      if ((rc = resize_output(R, out, nnz, total))) return fail(rc);
      capsum = total;
    }
    HashPlan hn;
    if ((!R.dense && (rc = plan_hash(R, hn, R.row_nnz, nullptr, nullptr))) || (rc = prepare_heavy(R, PASS_NUMERIC)) ||
        (!R.dense && (rc = reset_hash_pool(R, hn, true))))
      return fail(rc);
    PassParams Q;
    base_params(R, Q, R.dense ? nullptr : &hn);
    direct(Q, nullptr);  // exact mode: each row's count is its symbolic count
    if ((rc = run_pass(R, PASS_NUMERIC, &hn, fitems, nfitems, Q, frows, fallback))) return fail(rc);
    uint32_t ovf = 0;
    if ((rc = read_back(ctx, &ovf, R.overflow, 4))) return fail(rc);
    if (ovf & 0x80000000u) return fail(set_err(ctx, PTC_ERR_INTERNAL, "row count mismatch between passes"));
    if (ovf & 0x60000000u) return fail(set_err(ctx, PTC_ERR_INTERNAL, "slice index or reservation out of range"));
    nnz = total;
  }
  // capacities that overestimated the output by more than half: exact arrays
  if (capsum > 2 * nnz + 65536 && (rc = resize_output(R, out, nnz, nnz))) return fail(rc);
  out->N = nnz;
  T.mark();  // 3
  *out_p = out;
  *nnz_p = nnz;
  *fallback_p = fallback;
  return PTC_OK;
}

// ---------------- dense column windows
//
// Rows of a 16-bit column space too wide for one dense LDS accumulator.
// Small rows (one work item of at most SPARSE_FLOPS multiply-adds) run in an
// LDS hash share of small workgroups; the others run as dense column
// windows: window t's workgroups stream a row's whole flop range and
// accumulate the columns in [t * W, (t + 1) * W) in LDS.  The symbolic pass
// needs one bit per column, so it is a single dense pass over the whole
// column space that also counts each window's columns; the numeric pass runs
// the windows in turn, each writing at its row's first entry plus the row's
// counts in earlier windows.  The output is in row order with columns
// ascending within a row, as with one dense accumulator.  The single-pass
// methods reserve every row part with a fetch-add bounded only by the
// output's size (the rows' capacities summed); a run whose result does not
// fit is redone by the exact path.
struct WinRun {
  uint4* dense = nullptr;              // items of the windowed rows
  uint4* sparse = nullptr;             // items of the small rows
  uint32_t nd = 0, ns = 0;
  uint32_t* hcap = nullptr;            // small rows: LDS hash slots
  unsigned long long* zero = nullptr;  // small rows: none lives in HBM
  uint32_t* cnt = nullptr;             // [win_n][nrows]: each window's entries per row
  unsigned long long* off = nullptr;   // [win_n][nrows]: each window's first entry per row
  uint32_t* hbits = nullptr;           // split rows' symbolic slabs (whole column space)
  double* hacc = nullptr;              // numeric slabs, a window wide, reused window after window
  unsigned char* hflag = nullptr;
};

static uint32_t win_lo(const Run& R, uint32_t t) { return t * R.win_width; }
static uint32_t win_w(const Run& R, uint32_t t) {
  return (uint32_t)std::min<uint64_t>(R.win_width, R.ncols - (uint64_t)t * R.win_width);
}

// Items split by row kind; the two counts read back.
static int win_split(Run& R, WinRun& W) {
  ptc_ctx* ctx = R.ctx;
  Workspace& ws = R.ws;
  W.dense = ws.get<uint4>(R.nitems);
  W.sparse = ws.get<uint4>(R.nitems);
  W.hcap = ws.get<uint32_t>(R.nrows);
  W.zero = ws.get<unsigned long long>(R.nrows);
  W.cnt = ws.get<uint32_t>((uint64_t)R.win_n * R.nrows);
  uint32_t* c2 = ws.get<uint32_t>(2);
  if (!W.dense || !W.sparse || !W.hcap || !W.zero || !W.cnt || !c2) return PTC_ERR_NOMEM;
  int rc = PTC_OK;
  if ((rc = fill_zero(ctx, c2, 8)) || (rc = fill_zero(ctx, W.zero, R.nrows * 8)) ||
      (rc = fill_zero(ctx, W.cnt, (uint64_t)R.win_n * R.nrows * 4)) ||
      (rc = launch_split_items(ctx, R.items, R.nitems, R.row_flops, R.row_nchunk, W.dense, W.sparse, c2)) ||
      (rc = launch_sparse_caps(ctx, R.row_flops, R.nrows, W.hcap)))
    return rc;
  uint32_t h[2] = {0, 0};
  if ((rc = read_back(ctx, h, c2, 8))) return rc;
  W.nd = h[0];
  W.ns = h[1];
  return PTC_OK;
}

// The windowed rows' pass: the symbolic pass over the whole column space
// (t unused), or window t of the numeric pass.
static void win_params(Run& R, const WinRun& W, PassParams& P, int pass, uint32_t t) {
  base_params(R, P, nullptr);
  P.items = W.dense;
  P.bcol = R.bcol16;
  P.merge_rows = R.heavy_rows;
  if (pass == PASS_SYMBOLIC) {
    P.ncols = R.ncols;
    P.hbits = W.hbits;
    P.win_cnt = W.cnt;
    P.win_stride = R.nrows;
    P.win_width = R.win_width;
    P.win_n = R.win_n;
  } else {
    P.ncols = win_w(R, t);
    P.col_lo = win_lo(R, t);
    P.col_win = 1;
    P.row_nnz = W.cnt + (uint64_t)t * R.nrows;
    P.hacc = W.hacc;
    P.hflag = W.hflag;
  }
}

// The small rows' pass: LDS hash tables sized from their flops.
static void sparse_params(Run& R, const WinRun& W, PassParams& P) {
  base_params(R, P, nullptr);
  P.items = W.sparse;
  P.bcol = R.HB->rest_key32;
  P.row_hcap = W.hcap;
  P.row_gcap = W.zero;
  P.row_goff = W.zero;
  P.merge_rows = nullptr;
}

// One share's row pass and (dense) the merge launch of its split rows,
// bracketed by one pair of timing events.
static int run_share(Run& R, int pass, bool dense, uint32_t threads, uint64_t ncols, uint32_t lds_slots,
                     PassParams& P, uint64_t nitems, uint64_t nmerge) {
  if (!nitems) return PTC_OK;
  ptc_ctx* ctx = R.ctx;
  const size_t lds = pass_lds_bytes(pass, dense, false, ncols, lds_slots, R.use_flags, threads,
                                    pass == PASS_NUMERIC ? R.val : 0);
  const bool timed = ctx->timing && ctx->nkev + 2 <= 16;
  if (timed && hipEventRecord(ctx->kev[ctx->nkev], ctx->stream) != hipSuccess) (void)hipGetLastError();
  int rc = launch_pass(ctx, pass, dense, false, P, nitems, lds, false, threads);
  if (!rc && dense && nmerge) rc = launch_pass(ctx, pass, true, false, P, nmerge, lds, true, threads);
  if (rc) return rc;
  if (timed) {
    if (hipEventRecord(ctx->kev[ctx->nkev + 1], ctx->stream) != hipSuccess) (void)hipGetLastError();
    ctx->kev_pass[ctx->nkev / 2] = pass;
    ctx->nkev += 2;
  }
  return PTC_OK;
}

static int win_slabs(Run& R, WinRun& W) {
  if (!R.nheavy || W.hacc) return PTC_OK;
  W.hacc = R.ws.get<double>(R.nslabs * R.win_width);
  if (R.use_flags) W.hflag = R.ws.get<unsigned char>(R.nslabs * R.win_width);
  return W.hacc && (W.hflag || !R.use_flags) ? PTC_OK : PTC_ERR_NOMEM;
}

// Exact two-pass over windows: *out_p is ext (caller's arrays) or a new output.
static int win_two_pass(Run& R, WinRun& W, const ptc_dtensor* A, const ptc_dtensor* B, Timer& T,
                        const OutInto* into, ptc_dtensor* ext, ptc_dtensor** out_p, uint64_t* nnz_p) {
  ptc_ctx* ctx = R.ctx;
  Workspace& ws = R.ws;
  int rc = PTC_OK;
  const uint32_t sym_threads = R.opt.block_threads ? R.opt.block_threads : 512u;
  if (R.nheavy && !W.hbits && !(W.hbits = ws.get<uint32_t>(R.nslabs * ((R.ncols + 31) / 32)))) return PTC_ERR_NOMEM;
  if ((rc = fill_zero(ctx, R.row_nnz, (R.nrows + 1) * 4))) return rc;
  PassParams P;
  win_params(R, W, P, PASS_SYMBOLIC, 0);
  if ((rc = run_share(R, PASS_SYMBOLIC, true, sym_threads, R.ncols, 0, P, W.nd, R.nheavy))) return rc;
  sparse_params(R, W, P);
  if ((rc = run_share(R, PASS_SYMBOLIC, false, SPARSE_THREADS, 0, SPARSE_SLOTS, P, W.ns, 0))) return rc;
  T.mark();  // symbolic done
  unsigned long long nnz = 0;
  if ((rc = scan_u32_to_u64(ctx, R.row_nnz, R.out_off, R.nrows, R.out_off + R.nrows)) ||
      (rc = read_back(ctx, &nnz, R.out_off + R.nrows, 8)))
    return rc;
  ptc_dtensor* out = nullptr;
  if (into) {
    *into->nnz = nnz;
    if (nnz > into->cap)
      return set_err(ctx, PTC_ERR_RANGE, "the output has %llu nonzeros; the arrays hold %llu",
                     (unsigned long long)nnz, (unsigned long long)into->cap);
    ext->ctx = ctx;
    ext->d = R.dO;
    ext->N = nnz;
    ext->hedges = into->hedges;
    ext->vals = into->vals;
    out = ext;
  } else if ((rc = make_output(R, nnz, A, B, &out))) {
    return rc;
  }
  auto fail = [&](int code) {
    if (out != ext) ptc_dtensor_free(out);
    return code;
  };
  if (!(W.off = ws.get<unsigned long long>((uint64_t)R.win_n * R.nrows))) return fail(PTC_ERR_NOMEM);
  if ((rc = launch_window_offsets(ctx, R.out_off, W.cnt, R.nrows, R.win_n, W.off)) || (rc = win_slabs(R, W)))
    return fail(rc);
  auto outputs = [&](PassParams& Q) {
    Q.out_hedges = out->hedges;
    Q.out_vals = out->vals;
    Q.out_vals32 = out->vals32;
    Q.out_err32 = out->err32;
  };
  for (uint32_t t = 0; t < R.win_n; ++t) {
    win_params(R, W, P, PASS_NUMERIC, t);
    P.out_base = W.off + (uint64_t)t * R.nrows;
    outputs(P);
    if ((rc = run_share(R, PASS_NUMERIC, true, R.win_threads, win_w(R, t), 0, P, W.nd, R.nheavy))) return fail(rc);
  }
  sparse_params(R, W, P);
  P.out_base = R.out_off;
  outputs(P);
  if ((rc = run_share(R, PASS_NUMERIC, false, SPARSE_THREADS, 0, SPARSE_SLOTS, P, W.ns, 0))) return fail(rc);
  T.mark();  // numeric done
  uint32_t ovf = 0;
  if ((rc = read_back(ctx, &ovf, R.overflow, 4))) return fail(rc);
  if (ovf & 0x80000000u) return fail(set_err(ctx, PTC_ERR_INTERNAL, "row count mismatch between passes"));
  if (ovf & 0x40000000u) return fail(set_err(ctx, PTC_ERR_INTERNAL, "slice index out of range"));
  *out_p = out;
  *nnz_p = nnz;
  return PTC_OK;
}

// Single pass over windows into an output of the capacities' size.
static int win_single_pass(Run& R, WinRun& W, const ptc_dtensor* A, const ptc_dtensor* B, const uint32_t* cap,
                           const unsigned long long* cap_off, int method, Timer& T, ptc_dtensor** out_p,
                           uint64_t* nnz_p, uint64_t* fallback_p, uint64_t* capacity_p) {
  (void)cap;
  ptc_ctx* ctx = R.ctx;
  Workspace& ws = R.ws;
  int rc = PTC_OK;
  unsigned long long capsum = R.capsum;  // the capacities' sum, read back with the preprocessing totals
  *capacity_p = capsum;
  ptc_dtensor* out = nullptr;
  if ((rc = make_output(R, capsum, A, B, &out))) return rc;
  auto fail = [&](int code) {
    ptc_dtensor_free(out);
    return code;
  };
  unsigned long long* counter = ws.get<unsigned long long>(2);
  unsigned long long* pair = ws.get<unsigned long long>(2);
  if (!counter || !pair) return fail(PTC_ERR_NOMEM);
  if ((rc = fill_zero(ctx, counter, 16)) || (rc = win_slabs(R, W))) return fail(rc);
  auto direct = [&](PassParams& Q) {
    Q.out_count = counter;
    Q.out_limit = capsum;
    Q.out_soft = 1;
    Q.out_hedges = out->hedges;
    Q.out_vals = out->vals;
    Q.out_vals32 = out->vals32;
    Q.out_err32 = out->err32;
  };
  PassParams P;
  for (uint32_t t = 0; t < R.win_n; ++t) {
    win_params(R, W, P, PASS_NUMERIC, t);
    direct(P);
    if ((rc = run_share(R, PASS_NUMERIC, true, R.win_threads, win_w(R, t), 0, P, W.nd, R.nheavy))) return fail(rc);
  }
  sparse_params(R, W, P);
  direct(P);
  if ((rc = run_share(R, PASS_NUMERIC, false, SPARSE_THREADS, 0, SPARSE_SLOTS, P, W.ns, 0))) return fail(rc);
  unsigned long long rb[2] = {0, 0};
  {
    WordSources S{};
    S.n = 2;
    S.src[0] = R.overflow;
    S.bytes[0] = 4;
    S.src[1] = counter;
    S.bytes[1] = 8;
    if ((rc = collect_words(ctx, pair, S)) || (rc = read_back(ctx, rb, pair, 16))) return fail(rc);
  }
  const uint32_t ovf0 = (uint32_t)rb[0];
  if (ovf0 & 0x80000000u) return fail(set_err(ctx, PTC_ERR_INTERNAL, "accumulator table overflow"));
  if (ovf0 & 0x40000000u) return fail(set_err(ctx, PTC_ERR_INTERNAL, "slice index out of range"));
  const uint64_t fallback = ovf0 & 0x1FFFFFFFu;
  if (fallback) {
    // the result outgrew the capacities' sum: the exact path from the start
    ptc_dtensor_free(out);
    out = nullptr;
    if ((rc = fill_zero(ctx, R.overflow, 4))) return rc;
    uint64_t nnz = 0;
    if ((rc = win_two_pass(R, W, A, B, T, nullptr, nullptr, &out, &nnz))) return rc;
    *out_p = out;
    *nnz_p = nnz;
    *fallback_p = fallback;
    return PTC_OK;
  }
  T.mark();  // 2
  const uint64_t nnz = rb[1];
  if (capsum > 2 * nnz + 65536 && (rc = resize_output(R, out, nnz, nnz))) return fail(rc);
  out->N = nnz;
  T.mark();  // 3
  *out_p = out;
  *nnz_p = nnz;
  *fallback_p = 0;
  return PTC_OK;
}

}  // namespace ptc

namespace {

// The key split requested for the output of a contraction chain's step.
struct KeyReq {
  uint32_t nk;
  uint32_t dim_tc[PTC_MAX_MODES];
  ptc_hashtable** table;
  int* from_rows;
};

int tc_impl(ptc_ctx* ctx, const ptc_hashtable* HT_A, const ptc_hashtable* HT_B, const ptc_dtensor* A,
            const ptc_dtensor* B, const ptc_options* opt_in, ptc_dtensor** O, ptc_stats* stats, const KeyReq* kr,
            const OutInto* into = nullptr, bool col_first = false);

}  // namespace

namespace ptc {
int tc_into(ptc_ctx* ctx, const ptc_hashtable* HT_A, const ptc_hashtable* HT_B, const ptc_dtensor* A,
            const ptc_dtensor* B, const ptc_options* opt, const OutInto& into, ptc_stats* stats) {
  if (!ctx || !HT_A || !HT_B || !A || !B || !into.nnz || (into.cap && (!into.hedges || !into.vals)))
    return PTC_ERR_ARG;
  ptc_options o;
  if (opt) o = *opt; else ptc_options_default(&o);
  if (o.method != PTC_METHOD_TWO_PASS || o.value_type != PTC_VAL_F64)
    return set_err(ctx, PTC_ERR_ARG, "caller-provided outputs take the exact two-pass method in FP64");
  return tc_impl(ctx, HT_A, HT_B, A, B, &o, nullptr, stats, nullptr, &into);
}
}  // namespace ptc

extern "C" int ptc_tc_fksCuckoo(ptc_ctx* ctx, const ptc_hashtable* HT_A, const ptc_hashtable* HT_B,
                                const ptc_dtensor* A, const ptc_dtensor* B, const ptc_options* opt_in,
                                ptc_dtensor** O, ptc_stats* stats) {
  return tc_impl(ctx, HT_A, HT_B, A, B, opt_in, O, stats, nullptr);
}

extern "C" int ptc_tc_fksCuckoo_keyed(ptc_ctx* ctx, const ptc_hashtable* HT_A, const ptc_hashtable* HT_B,
                                      const ptc_dtensor* A, const ptc_dtensor* B, const ptc_options* opt,
                                      uint32_t nk_next, const uint32_t* dim_tc_next, ptc_dtensor** O,
                                      ptc_hashtable** HT_O, int* from_rows, ptc_stats* stats) {
  if (!ctx || !A || !B || !O || !HT_O || !dim_tc_next) return PTC_ERR_ARG;
  *O = nullptr;
  *HT_O = nullptr;
  if (from_rows) *from_rows = 0;
  const uint32_t dO = A->nk + (B->d - B->nk);  // tensor.hpp:113
  if (nk_next == 0 || nk_next > dO) return set_err(ctx, PTC_ERR_ARG, "%u key modes for a %u-mode output", nk_next, dO);
  KeyReq kr{};
  kr.nk = nk_next;
  kr.table = HT_O;
  kr.from_rows = from_rows;
  uint32_t seen = 0;
  for (uint32_t j = 0; j < dO; ++j) {
    const uint32_t m = dim_tc_next[j];
    if (m >= dO || (seen >> m) & 1u) return set_err(ctx, PTC_ERR_ARG, "dim_tc is not a permutation of the %u output modes", dO);
    seen |= 1u << m;
    kr.dim_tc[j] = m;
  }
  return tc_impl(ctx, HT_A, HT_B, A, B, opt, O, stats, &kr);
}

namespace {

int tc_impl(ptc_ctx* ctx, const ptc_hashtable* HT_A, const ptc_hashtable* HT_B, const ptc_dtensor* A,
            const ptc_dtensor* B, const ptc_options* opt_in, ptc_dtensor** O, ptc_stats* stats, const KeyReq* kr,
            const OutInto* into, bool col_first) {
  if (!ctx || !HT_A || !HT_B || !A || !B || !(O || into)) return PTC_ERR_ARG;
  if (O) *O = nullptr;
  // caller-provided output arrays (exact two-pass only): a non-owning view
  ptc_dtensor ext;
  ext.owns = false;
  auto drop = [&](ptc_dtensor* o) {
    if (o != &ext) ptc_dtensor_free(o);
  };
  ptc_options opt;
  if (opt_in) opt = *opt_in; else ptc_options_default(&opt);
  Run R(ctx);
  // col_first: the output's modes are the right operand's free modes, then
  // the left operand's (the column-wise product computes C^T's rows)
  R.col_first = col_first;
  if (col_first && (opt.method != PTC_METHOD_TWO_PASS || kr || into))
    return set_err(ctx, PTC_ERR_INTERNAL, "column-first output takes the plain two-pass path");
  // a chain step's single pass stages its rows in row order (no windows)
  R.no_windows = kr && opt.method != PTC_METHOD_TWO_PASS;
  Timer T(ctx);
  ctx->nkev = 0;
  const size_t live0 = ctx->live_bytes;
  ctx->peak_bytes = live0;
  auto wall0 = std::chrono::steady_clock::now();
  T.mark();  // 0
  // Preprocessing.  A single-pass contraction over dense rows whose output
  // and split-row slabs have host-known bounds within AHEAD_BYTES runs its
  // numeric pass ahead of the preprocessing totals (R.ahead, one round trip
  // in all); every other contraction reads the totals back first.
  unsigned long long* tot = R.ws.get<unsigned long long>(5);
  if (!tot) return PTC_ERR_NOMEM;
  R.allow_ahead = opt.method != PTC_METHOD_TWO_PASS && !kr && !into && !col_first;
  int rc = prepare_launch(R, HT_A, HT_B, A, B, opt, nullptr);
  if (rc) return rc;
  const bool ahead = R.ahead;
  if (!ahead) {
    unsigned long long hostv5[5] = {0, 0, 0, 0, 0};
    if ((rc = collect_words(ctx, tot, R.totw)) || (rc = read_back(R.ctx, hostv5, tot, 8 * R.totw.n)) ||
        (rc = prepare_finish(R, hostv5)))
      return rc;
  }
  Workspace& ws = R.ws;
  unsigned long long* out_off = R.out_off;
  unsigned long long hostv[2] = {0, 0};
  T.mark();  // 1
  uint64_t nnz = 0;
  uint64_t fallback_rows = 0;
  uint64_t out_capacity = 0;   // output entries allocated before the count was known
  bool flags_checked = false;  // the last pass's consistency flags were read already
  ptc_dtensor* out = nullptr;
  const int method = opt.method;

  WinRun W;
  if (R.win_n && (rc = win_split(R, W))) return rc;
  if (method == PTC_METHOD_TWO_PASS && R.win_n) {
    if ((rc = win_two_pass(R, W, A, B, T, into, &ext, &out, &nnz))) return rc;
    out_capacity = nnz;
    flags_checked = true;
  } else if (method == PTC_METHOD_TWO_PASS) {
    // ---------------- symbolic
    HashPlan hs;
    if (!R.dense && (rc = plan_hash(R, hs, nullptr, R.row_flops, nullptr))) return rc;
    if ((rc = fill_zero(ctx, R.row_nnz, (R.nrows + 1) * 4))) return rc;
    if ((rc = prepare_heavy(R, PASS_SYMBOLIC))) return rc;
    if (!R.dense && (rc = reset_hash_pool(R, hs, false))) return rc;
    PassParams P;
    base_params(R, P, R.dense ? nullptr : &hs);
    if ((rc = run_pass(R, PASS_SYMBOLIC, &hs, R.items, R.nitems, P))) return rc;
    T.mark();  // 2
    // ---------------- scan
    if ((rc = scan_u32_to_u64(ctx, R.row_nnz, out_off, R.nrows, out_off + R.nrows))) return rc;
    if ((rc = read_back(ctx, &hostv[0], out_off + R.nrows, 8))) return rc;
    nnz = hostv[0];
    out_capacity = nnz;
    if (into) {
      *into->nnz = nnz;  // also when the caller's arrays are too small
      if (nnz > into->cap)
        return set_err(ctx, PTC_ERR_RANGE, "the output has %llu nonzeros; the arrays hold %llu",
                       (unsigned long long)nnz, (unsigned long long)into->cap);
      ext.ctx = ctx;
      ext.d = R.dO;
      ext.N = nnz;
      ext.hedges = into->hedges;
      ext.vals = into->vals;
      out = &ext;
    } else if ((rc = make_output(R, nnz, A, B, &out))) {
      return rc;
    }
    // ---------------- numeric
    HashPlan hn;
    if (!R.dense && (rc = plan_hash(R, hn, R.row_nnz, nullptr, nullptr))) {
      drop(out);
      return rc;
    }
    if ((rc = prepare_heavy(R, PASS_NUMERIC)) ||
        (!R.dense && (rc = reset_hash_pool(R, hn, true)))) {
      drop(out);
      return rc;
    }
    base_params(R, P, R.dense ? nullptr : &hn);
    P.out_base = out_off;
    P.out_hedges = out->hedges;
    P.out_vals = out->vals;
    P.out_vals32 = out->vals32;
    P.out_err32 = out->err32;
    if ((rc = run_pass(R, PASS_NUMERIC, &hn, R.items, R.nitems, P))) {
      drop(out);
      return rc;
    }
    T.mark();  // 3
  } else {
    // ---------------- single pass into capacities
    uint32_t* cap = R.cap;
    unsigned long long* cap_off = R.cap_off;
    if (!cap || !cap_off) return set_err(ctx, PTC_ERR_INTERNAL, "capacities were not computed");
    if (!kr) {
      rc = R.win_n ? win_single_pass(R, W, A, B, cap, cap_off, method, T, &out, &nnz, &fallback_rows, &out_capacity)
                   : single_pass_direct(R, A, B, cap, cap_off, method, T, &out, &nnz, &fallback_rows, &out_capacity,
                                        ahead);
      if (rc) return rc;
      flags_checked = true;
    } else {
    // A chain step keeps its output rows in row order (its key table is
    // formed from them): staging sized by the capacities' sum, a scan of the
    // true counts, and a copy into place.
    const unsigned long long capsum = R.capsum;
    const size_t kb = R.col64 ? 8 : 4;
    void* stg_col = ws.get<unsigned char>(capsum * kb);
    double* stg_val = ws.get<double>(capsum);
    if (!stg_col || !stg_val) return PTC_ERR_NOMEM;
    HashPlan h1;
    if (!R.dense && (rc = plan_hash(R, h1, cap, nullptr, R.row_flops))) return rc;
    if ((rc = fill_zero(ctx, R.row_nnz, (R.nrows + 1) * 4)) ||
        (rc = prepare_heavy(R, PASS_NUMERIC)) || (!R.dense && (rc = reset_hash_pool(R, h1, true))))
      return rc;
    PassParams P;
    base_params(R, P, R.dense ? nullptr : &h1);
    P.out_base = cap_off;
    P.out_cap = cap;
    P.stg_col = stg_col;
    P.stg_val = stg_val;
    if ((rc = run_pass(R, PASS_NUMERIC, &h1, R.items, R.nitems, P))) return rc;
    T.mark();  // 2
    // In the common case no row overflowed its capacity and the scan of the
    // counts is final: read the overflow word and the total together.
    if ((rc = scan_u32_to_u64(ctx, R.row_nnz, out_off, R.nrows, out_off + R.nrows))) return rc;
    unsigned long long rb[2] = {0, 0};
    {
      unsigned long long* pair = ws.get<unsigned long long>(2);
      if (!pair) return PTC_ERR_NOMEM;
      WordSources W{};
      W.n = 2;
      W.src[0] = R.overflow;
      W.bytes[0] = 4;
      W.src[1] = out_off + R.nrows;
      W.bytes[1] = 8;
      if ((rc = collect_words(ctx, pair, W))) return rc;
      if ((rc = read_back(ctx, rb, pair, 16))) return rc;
    }
    const uint32_t ovf0 = (uint32_t)rb[0];
    if (ovf0 & 0x80000000u) return set_err(ctx, PTC_ERR_INTERNAL, "accumulator table overflow");
    if (ovf0 & 0x40000000u) return set_err(ctx, PTC_ERR_INTERNAL, "slice index out of range");
    fallback_rows = ovf0;
    nnz = rb[1];
    uint4* fitems = nullptr;
    uint32_t nfitems = 0;
    uint32_t* frows = nullptr;
    uint32_t* row_fb = nullptr;
    if (fallback_rows) {
      // exact two-pass for the rows whose capacity was too small
      uint32_t* cnt = ws.get<uint32_t>(4);
      fitems = ws.get<uint4>(R.nitems);
      frows = ws.get<uint32_t>(fallback_rows);
      row_fb = ws.get<uint32_t>(R.nrows);
      if (!cnt || !fitems || !frows || !row_fb) return PTC_ERR_NOMEM;
      if ((rc = fill_zero(ctx, cnt, 16))) return rc;
      if ((rc = launch_select_items(ctx, R.items, R.nitems, R.row_nnz, cap, fitems, cnt))) return rc;
      if ((rc = launch_overflow_rows(ctx, R.row_nnz, cap, R.nrows, frows, cnt + 1, row_fb))) return rc;
      if ((rc = read_back(ctx, &nfitems, cnt, 4))) return rc;
      // symbolic on those rows
      if ((rc = launch_reset_rows(ctx, frows, (uint32_t)fallback_rows, R.row_nnz))) return rc;
      HashPlan hs;
      if (!R.dense && (rc = plan_hash(R, hs, nullptr, R.row_flops, nullptr))) return rc;
      if ((rc = prepare_heavy(R, PASS_SYMBOLIC)) || (!R.dense && (rc = reset_hash_pool(R, hs, false)))) return rc;
      PassParams S;
      base_params(R, S, R.dense ? nullptr : &hs);
      if ((rc = run_pass(R, PASS_SYMBOLIC, &hs, fitems, nfitems, S, frows, fallback_rows))) return rc;
      if ((rc = scan_u32_to_u64(ctx, R.row_nnz, out_off, R.nrows, out_off + R.nrows))) return rc;
      if ((rc = read_back(ctx, &hostv[0], out_off + R.nrows, 8))) return rc;
      nnz = hostv[0];
    }
    if ((rc = make_output(R, nnz, A, B, &out))) return rc;
    {
      EmitParams E{};
      E.nrows = R.nrows;
      E.rows = HT_A->slices;
      E.row_nnz = R.row_nnz;
      E.cap = cap;
      E.row_fb = row_fb;
      E.cap_base = cap_off;
      E.out_base = out_off;
      E.stg_col = stg_col;
      E.stg_val = stg_val;
      E.out_hedges = out->hedges;
      E.out_vals = out->vals;
      E.out_vals32 = out->vals32;
      E.out_err32 = out->err32;
      E.f32 = R.f32 ? 1 : 0;
      E.dO = R.dO;
      E.nrowmodes = R.nrowmodes;
      E.rowc = HT_A->codec;
      E.colc = HT_B->rest;
      if ((rc = launch_emit(ctx, E, R.col64))) {
        ptc_dtensor_free(out);
        return rc;
      }
    }
    if (fallback_rows) {
      HashPlan hn;
      if ((!R.dense && (rc = plan_hash(R, hn, R.row_nnz, nullptr, nullptr))) || (rc = prepare_heavy(R, PASS_NUMERIC)) ||
          (!R.dense && (rc = reset_hash_pool(R, hn, true)))) {
        ptc_dtensor_free(out);
        return rc;
      }
      PassParams Q;
      base_params(R, Q, R.dense ? nullptr : &hn);
      Q.out_base = out_off;
      Q.out_hedges = out->hedges;
      Q.out_vals = out->vals;
      Q.out_vals32 = out->vals32;
      Q.out_err32 = out->err32;
      if ((rc = run_pass(R, PASS_NUMERIC, &hn, fitems, nfitems, Q, frows, fallback_rows))) {
        ptc_dtensor_free(out);
        return rc;
      }
    }
    out_capacity = capsum;
    T.mark();  // 3
    }
  }
  // The consistency flags of the last numeric pass: in the single-pass
  // methods without fallback rows they were read with the counts already.
  if (!flags_checked && (method == PTC_METHOD_TWO_PASS || fallback_rows)) {
    uint32_t ovf = 0;
    rc = read_back(ctx, &ovf, R.overflow, 4);
    if (!rc && (ovf & 0x80000000u)) rc = set_err(ctx, PTC_ERR_INTERNAL, "row count mismatch between passes");
    if (!rc && (ovf & 0x40000000u)) rc = set_err(ctx, PTC_ERR_INTERNAL, "slice index out of range");
    if (rc) {
      drop(out);
      return rc;
    }
  }
  // The output's key table for the next contraction: from the row grouping
  // when the key modes are exactly the row modes, else built.
  ptc_hashtable* hto = nullptr;
  bool rows_key = false;
  if (kr) {
    out->nk = kr->nk;
    for (uint32_t j = 0; j < out->d; ++j) out->dim_tc[j] = kr->dim_tc[j];
    rows_key = kr->nk == R.nrowmodes && nnz <= 0xFFFFFFFFull;
    for (uint32_t q = 0; q < kr->nk; ++q) rows_key = rows_key && kr->dim_tc[q] < R.nrowmodes;
    rc = rows_key ? hashtable_from_rows(ctx, out, HT_A->codec, HT_A->slices, R.nrows, R.row_nnz, out_off, &hto)
                  : ptc_hashtable_build_ex(ctx, out, PTC_GROUP_AUTO, &hto);
    if (rc) {
      ptc_dtensor_free(out);
      return rc;
    }
    *kr->table = hto;
    if (kr->from_rows) *kr->from_rows = rows_key ? 1 : 0;
  }
  if (stats) {
    if (T.n && hipEventSynchronize(T.ev[T.n - 1]) != hipSuccess) (void)hipGetLastError();
    memset(stats, 0, sizeof(*stats));
    if (hto) stats->ms_build_out = hashtable_ms_build(hto);
    stats->nnz_a = A->N;
    stats->nnz_b = B->N;
    stats->rows = R.nrows;
    stats->flops = R.F;
    stats->nnz_out = nnz;
    stats->fallback_rows = fallback_rows;
    stats->heavy_rows = R.nheavy;
    stats->out_capacity = out_capacity;
    stats->device_peak_bytes = ctx->peak_bytes - live0;
    stats->col_windows = R.win_n;
    stats->small_rows = W.ns;
    stats->ms_build_a = hashtable_ms_build(HT_A);
    stats->ms_build_b = hashtable_ms_build(HT_B);
    stats->ms_preprocess = T.ms(0, 1);
    if (method == PTC_METHOD_TWO_PASS) {
      stats->ms_symbolic = T.ms(1, 2);
      stats->ms_numeric = T.ms(2, 3);
    } else {
      stats->ms_numeric = T.ms(1, 3);
    }
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    // pass time = its row-pass launch + its merge launch
    for (int i = 0; i + 1 < ctx->nkev; i += 2) {
      float t = 0.f;
      if (hipEventElapsedTime(&t, ctx->kev[i], ctx->kev[i + 1]) != hipSuccess) {
        (void)hipGetLastError();
        t = 0.f;
      }
      if (ctx->kev_pass[i / 2] == PASS_NUMERIC) {
        stats->ms_kernel_numeric += t;
        stats->kernel_launches_numeric++;
      } else {
        stats->ms_kernel_symbolic += t;
      }
    }
  }
  if (O) *O = out;
  return PTC_OK;
}

}  // namespace

// A view of t with another key split, sharing t's device arrays.
namespace ptc {
ptc_dtensor* view_with_keys(const ptc_dtensor* t, uint32_t nk, const uint32_t* dim_tc) {
  ptc_dtensor* v = new ptc_dtensor(*t);
  v->owns = false;
  v->nk = nk;
  for (uint32_t j = 0; j < t->d; ++j) v->dim_tc[j] = dim_tc[j];
  return v;
}
}  // namespace ptc

static int contract_grouped(ptc_ctx* ctx, const ptc_dtensor* A, const ptc_dtensor* B, const ptc_options* opt,
                            ptc_dtensor** O, ptc_stats* stats, int grouping);

extern "C" int ptc_contract(ptc_ctx* ctx, const ptc_dtensor* A, const ptc_dtensor* B, const ptc_options* opt,
                            ptc_dtensor** O, ptc_stats* stats) {
  return contract_grouped(ctx, A, B, opt, O, stats, PTC_GROUP_AUTO);
}

// The driver sequence of main.cpp:68-118 with a chosen table grouping.
static int contract_grouped(ptc_ctx* ctx, const ptc_dtensor* A, const ptc_dtensor* B, const ptc_options* opt,
                            ptc_dtensor** O, ptc_stats* stats, int grouping) {
  if (!ctx || !A || !B || !O) return PTC_ERR_ARG;
  *O = nullptr;
  const uint32_t c = A->nk;
  if (c == 0 || B->nk != c) return set_err(ctx, PTC_ERR_ARG, "A and B must have the same number (>0) of contracted modes");
  // A_copy: free modes first (main.cpp:68-109)
  uint32_t rot[PTC_MAX_MODES];
  for (uint32_t i = c; i < A->d; ++i) rot[i - c] = A->dim_tc[i];
  for (uint32_t i = 0; i < c; ++i) rot[A->d - c + i] = A->dim_tc[i];
  ptc_dtensor* Ac = view_with_keys(A, A->d - c, rot);
  ptc_hashtable *HB = nullptr, *HA = nullptr;
  int rc = ptc_hashtable_build_ex(ctx, B, grouping, &HB);  // B first, as main.cpp:115-116
  if (!rc) rc = ptc_hashtable_build_ex(ctx, Ac, grouping, &HA);
  if (!rc) rc = ptc_tc_fksCuckoo(ctx, HA, HB, Ac, B, opt, O, stats);
  ptc_hashtable_free(HA);
  ptc_hashtable_free(HB);
  delete Ac;
  return rc;
}

// The matricized product C = A_mat B_mat computed column by column, as
// cs_multiply does (cs_multiply.c:6-38 over the CSC operands that
// matricizedspGEMM.hpp:172-603 builds): for each column j of C (a B free
// tuple), the columns of A_mat (A grouped by its contracted key) named by
// B's nonzeros in column j are scaled and accumulated over C's rows.  On the
// device this is the row pass run on the transposed problem C^T = B_mat^T
// A_mat^T: B is grouped by its free modes (the output columns: the CSC
// column pointers of B_mat), A by its contracted modes (the columns of
// A_mat), every B nonzero probes A's table, and A's slices are streamed and
// accumulated per output column over C's row space.  The entries of each
// column are written with A's free modes first, so the result has the
// modes and meaning of ptc_contract's, column by column (the CSC order of
// cs_multiply's result) instead of row by row.  Against ptc_contract it
// exercises other tables (keyed on B's free and A's contracted modes),
// another probe direction and another accumulator space -- the internal
// cross-check that main.cpp:131-158 intended between the two methods.
extern "C" int ptc_matricize_spGEMM(ptc_ctx* ctx, const ptc_dtensor* A, const ptc_dtensor* B, ptc_dtensor** O,
                                    ptc_stats* stats) {
  if (!ctx || !A || !B || !O) return PTC_ERR_ARG;
  *O = nullptr;
  const uint32_t c = A->nk;
  if (c == 0 || B->nk != c) return set_err(ctx, PTC_ERR_ARG, "A and B must have the same number (>0) of contracted modes");
  // B with its free modes as key: columns of C
  uint32_t rot[PTC_MAX_MODES];
  for (uint32_t i = c; i < B->d; ++i) rot[i - c] = B->dim_tc[i];
  for (uint32_t i = 0; i < c; ++i) rot[B->d - c + i] = B->dim_tc[i];
  ptc_dtensor* Bc = view_with_keys(B, B->d - c, rot);
  ptc_hashtable *HA = nullptr, *HBc = nullptr;
  int rc = ptc_hashtable_build(ctx, A, &HA);  // A keyed on its contracted modes: A_mat's columns
  if (!rc) rc = ptc_hashtable_build(ctx, Bc, &HBc);
  ptc_options opt;
  ptc_options_default(&opt);
  if (!rc) rc = tc_impl(ctx, HBc, HA, Bc, A, &opt, O, stats, nullptr, nullptr, /*col_first=*/true);
  if (!rc && stats) {  // report the operands as the caller named them
    std::swap(stats->nnz_a, stats->nnz_b);
    std::swap(stats->ms_build_a, stats->ms_build_b);
  }
  ptc_hashtable_free(HBc);
  ptc_hashtable_free(HA);
  delete Bc;
  return rc;
}

// The sorting matricization groups both operands by sorting their keys
// (hashing where a key space exceeds 32 bits).
extern "C" int ptc_matricize_sort_spGEMM(ptc_ctx* ctx, const ptc_dtensor* A, const ptc_dtensor* B, ptc_dtensor** O,
                                         ptc_stats* stats) {
  if (!ctx || !A || !B || !O) return PTC_ERR_ARG;
  int rc = contract_grouped(ctx, A, B, nullptr, O, stats, PTC_GROUP_SORT);
  if (rc == PTC_ERR_UNSUPPORTED) rc = contract_grouped(ctx, A, B, nullptr, O, stats, PTC_GROUP_HASH);
  return rc;
}

// Per-row training data for the row-size network: the learned features and
// the exact distinct-column counts of the symbolic pass.
extern "C" int ptc_row_profile(ptc_ctx* ctx, const ptc_hashtable* HT_A, const ptc_hashtable* HT_B,
                               const ptc_dtensor* A, const ptc_dtensor* B, float* X, float* ub, uint32_t* nnz,
                               uint64_t cap_rows, uint64_t* nrows_out) {
  if (!ctx || !HT_A || !HT_B || !A || !B || !nrows_out) return PTC_ERR_ARG;
  ptc_options opt;
  ptc_options_default(&opt);
  Run R(ctx);
  R.want_features = true;
  int rc = prepare(R, HT_A, HT_B, A, B, opt);
  if (rc) return rc;
  *nrows_out = R.nrows;
  if (R.nrows > cap_rows || !X || !ub || !nnz) return set_err(ctx, PTC_ERR_RANGE, "profile needs %llu rows", (unsigned long long)R.nrows);
  HashPlan hs;
  if (!R.dense && (rc = plan_hash(R, hs, nullptr, R.row_flops, nullptr))) return rc;
  if ((rc = fill_zero(ctx, R.row_nnz, (R.nrows + 1) * 4)) ||
      (rc = prepare_heavy(R, PASS_SYMBOLIC)) || (!R.dense && (rc = reset_hash_pool(R, hs, false))))
    return rc;
  PassParams P;
  base_params(R, P, R.dense ? nullptr : &hs);
  if ((rc = run_pass(R, PASS_SYMBOLIC, &hs, R.items, R.nitems, P))) return rc;
  float* dX = R.ws.get<float>(R.nrows * PTC_MLP_IN);
  float* dub = R.ws.get<float>(R.nrows);
  if (!dX || !dub) return PTC_ERR_NOMEM;
  LearnedInputs L{};
  L.nrows = R.nrows;
  L.rows = HT_A->slices;
  L.row_flops = R.row_flops;
  L.row_hits = R.row_hits;
  L.row_maxlen = R.row_maxlen;
  L.ncols = R.ncols;
  if ((rc = learned_features(ctx, L, dX, dub))) return rc;
  PTC_HIP(ctx, hipMemcpyAsync(X, dX, sizeof(float) * R.nrows * PTC_MLP_IN, hipMemcpyDeviceToHost, ctx->stream));
  PTC_HIP(ctx, hipMemcpyAsync(ub, dub, sizeof(float) * R.nrows, hipMemcpyDeviceToHost, ctx->stream));
  PTC_HIP(ctx, hipMemcpyAsync(nnz, R.row_nnz, sizeof(uint32_t) * R.nrows, hipMemcpyDeviceToHost, ctx->stream));
  PTC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PTC_OK;
}
