// Batched contractions: many independent contractions of uneven sizes.
//
// Each job is a ptc_contract (.meta-style tensors, A_copy made free-first,
// HT_B then HT_A, exact two-pass method), but the batch is organised so that
// the device is not driven job by job:
//
//   plan   all key tables are enqueued without waiting; every step that ends
//          in a device -> host read (preprocessing totals, HBM accumulator
//          pools, output sizes) is launched for every job first and read back
//          for the whole batch in one round trip; the symbolic passes of all
//          jobs run as persistent launches over one work queue per bin.
//   run    the numeric phase -- refresh of the grouped values from the input
//          tensors, accumulator resets, one persistent numeric launch per bin
//          and its merge launch -- is captured once in a hipGraph and
//          launched as a whole.  Replaying it recomputes every output for new
//          values on the same sparsity pattern.
//
// Bins: jobs are grouped by accumulator kind (dense LDS columns / LDS or HBM
// hash), column key width and LDS footprint class, so one launch's dynamic
// LDS fits all its jobs without sizing small jobs for the largest.  Within a
// bin the queue lists the jobs with the heaviest work items first.
//
// No reference counterpart: main.cpp:59-118 runs one contraction per
// process.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "ptc_contract.h"
#include "ptc_internal.h"
#include "ptc_run.h"

namespace ptc {

// One table's value refresh: gval[pos] = vals[perm[pos]].
struct RefreshSeg {
  const uint32_t* perm;
  const double* vals;
  double* gval;
  uint64_t n;
};

// One fill: n 32-bit words (or 64-bit when wide) set to v.
struct FillSeg {
  void* p;
  uint64_t n;
  unsigned long long v;
  int wide;
};

__global__ void k_refresh_batch(const RefreshSeg* __restrict__ segs, const unsigned long long* __restrict__ start,
                                int nseg, unsigned long long total) {
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (unsigned long long)gridDim.x * blockDim.x) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (start[mid] <= i) lo = mid; else hi = mid - 1;
    }
    const RefreshSeg s = segs[lo];
    const uint64_t pos = i - start[lo];
    s.gval[pos] = s.vals[s.perm[pos]];
  }
}

// all segments as one flat index space (start[k] = first index of segment k)
__global__ void k_fill_batch(const FillSeg* __restrict__ segs, const unsigned long long* __restrict__ start, int nseg,
                             unsigned long long total) {
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (unsigned long long)gridDim.x * blockDim.x) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (start[mid] <= i) lo = mid; else hi = mid - 1;
    }
    const FillSeg s = segs[lo];
    const uint64_t k = i - start[lo];
    if (s.wide) ((unsigned long long*)s.p)[k] = s.v;
    else ((uint32_t*)s.p)[k] = (uint32_t)s.v;
  }
}

// Device fill of a list of segments: the launch (capturable) and its arrays.
struct FillPlan {
  FillSeg* segs = nullptr;
  unsigned long long* start = nullptr;
  int nseg = 0;
  unsigned long long total = 0;
};

__global__ void k_gather_words(const uint32_t* const* __restrict__ src, int n, uint32_t* __restrict__ dst) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = *src[i];
}

// A launch group: one kernel variant and LDS size over several jobs.
struct Bin {
  bool dense = true, col64 = false;
  size_t lds = 0;
  uint32_t threads = PB;
  std::vector<int> jobs;             // batch job indices
  PassParams* d_params = nullptr;    // device, one per job of the bin
  uint2* d_queue = nullptr;          // main work
  uint32_t nq = 0;
  uint2* d_mqueue = nullptr;         // merge work
  uint32_t nmq = 0;
  uint32_t* d_head = nullptr;        // two counters: main, merge
  unsigned grid = 0, mgrid = 0;      // device-filling grids (main, merge)
};

// Rows of a dense job that are small against its column space (or of a hash
// job, whose LDS tables are sized for its largest rows), in a column space of
// at least SPLIT_MIN_COLS, run in an LDS hash bin of small workgroups (the
// row-kind split of ptc_contract.h: SPARSE_FLOPS, SPARSE_SLOTS,
// SPARSE_THREADS).
constexpr uint64_t SPLIT_MIN_COLS = 4096;

// One kernel kind's share of a job: its parameters (items pointing at the
// share's work list), kernel shape and work counts.
struct Work {
  int job = 0;
  bool dense = true, col64 = false;
  uint32_t threads = PB, lds_slots = 0;
  uint64_t ncols = 0;   // accumulator columns (dense work)
  PassParams params{};
  uint64_t nitems = 0, nmerge = 0;
  double weight = 0.0;  // multiply-adds per item (queue order)
};

// A dense job whose small rows run in a hash bin (device arrays in b->ws).
struct Split {
  uint32_t nd = 0, ns = 0;     // dense and sparse items
  uint4* dense = nullptr;
  uint4* sparse = nullptr;
  uint32_t* hcap = nullptr;    // LDS hash slots of every row
  unsigned long long* zero = nullptr;  // no row of the share lives in HBM
};

// A hash job whose column space is at most MAX_WINDOWS dense accumulators
// wide (a 16-bit column space): its large rows run as dense column windows
// -- each window's workgroups stream the row's whole flop range and
// accumulate the columns in [lo, lo + w) in LDS -- instead of LDS or HBM hash
// tables, and its small rows in the sparse share.  Every window (and the
// sparse share) counts its own entries per row and writes its own part of
// the job's output.
struct Win {
  uint32_t T = 0;  // windows (0: not windowed)
  uint32_t threads = PB;
  uint32_t lo[MAX_WINDOWS] = {}, w[MAX_WINDOWS] = {};
  const uint16_t* col16 = nullptr;
  uint32_t* row_nnz[MAX_WINDOWS + 1] = {};  // [T]: the sparse share
  unsigned long long* out_off[MAX_WINDOWS + 1] = {};
  uint32_t* hbits[MAX_WINDOWS] = {};
  double* hacc[MAX_WINDOWS] = {};
  unsigned char* hflag[MAX_WINDOWS] = {};
  unsigned long long base[MAX_WINDOWS + 1] = {};  // first output entry of each part
};

// LDS footprint classes: a bin's launch opens the class size, so jobs of a
// bin differ in footprint by at most the class step.
static size_t lds_class(size_t b) {
  static const size_t cls[] = {16384, 24576, 32768, 49152, 65536, 98304, 131072, 163840};
  for (size_t c : cls)
    if (b <= c) return c;
  return b;
}

}  // namespace ptc

using namespace ptc;

struct ptc_batch {
  ptc_ctx* ctx = nullptr;
  int n = 0;
  ptc_options opt{};
  std::vector<const ptc_dtensor*> A, B;
  std::vector<ptc_dtensor*> Ac;
  std::vector<ptc_hashtable*> HA, HB;
  std::vector<std::unique_ptr<Run>> runs;
  std::vector<HashPlan> hn;
  std::vector<ptc_dtensor*> out;
  std::vector<Bin> bins;
  std::vector<Split> split;          // per job (ns = 0: not split)
  std::vector<Win> win;              // per job (T = 0: not windowed)
  Workspace* ws = nullptr;           // batch-level device arrays
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev[8] = {};
  uint32_t** d_ovf_src = nullptr;
  uint32_t* d_ovf = nullptr;
  std::vector<std::vector<unsigned char>> host_keep;  // sources of asynchronous uploads
  ptc_batch_stats st{};
};

namespace {

// Device copy of v; the host bytes are kept by the batch until it is freed,
// so the asynchronous copy never reads freed memory.
template <class T>
T* upload_vec(ptc_batch* b, const std::vector<T>& v) {
  if (v.empty()) return nullptr;
  T* d = b->ws->get<T>(v.size());
  if (!d) return nullptr;
  b->host_keep.emplace_back((const unsigned char*)v.data(), (const unsigned char*)v.data() + sizeof(T) * v.size());
  if (hipMemcpyAsync(d, b->host_keep.back().data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, b->ctx->stream) !=
      hipSuccess)
    return nullptr;
  return d;
}

// Host round trip of one 64-bit word per job (dst[j] = *src[j] on the device).
int read_words64(ptc_batch* b, const std::vector<const unsigned long long*>& src, std::vector<unsigned long long>& out) {
  ptc_ctx* ctx = b->ctx;
  const int n = (int)src.size();
  out.assign(n, 0);
  if (!n) return PTC_OK;
  unsigned long long* dst = b->ws->get<unsigned long long>(n);
  if (!dst) return PTC_ERR_NOMEM;
  for (int j = 0; j < n; ++j)
    if (src[j]) PTC_HIP(ctx, hipMemcpyAsync(dst + j, src[j], 8, hipMemcpyDeviceToDevice, ctx->stream));
  int rc = read_back(ctx, out.data(), dst, sizeof(unsigned long long) * n);
  for (int j = 0; j < n; ++j)
    if (!src[j]) out[j] = 0;
  return rc;
}

void stats_mark(ptc_batch* b, int i) { (void)hipEventRecord(b->ev[i], b->ctx->stream); }
double stats_ms(ptc_batch* b, int i, int j) {
  float t = 0.f;
  if (hipEventElapsedTime(&t, b->ev[i], b->ev[j]) != hipSuccess) {
    (void)hipGetLastError();
    return 0.0;
  }
  return t;
}

// Bins of the works for one pass and their queues.
int make_bins(ptc_batch* b, int pass, const std::vector<Work>& works, std::vector<Bin>& bins) {
  std::map<std::tuple<bool, bool, size_t, uint32_t>, int> key;
  bins.clear();
  for (int w = 0; w < (int)works.size(); ++w) {
    const Work& W = works[w];
    const Run& R = *b->runs[W.job];
    const size_t need = pass_lds_bytes(pass, W.dense, W.col64, W.ncols ? W.ncols : R.ncols, W.lds_slots, R.use_flags,
                                       W.threads);
    const size_t lds = std::max(need, std::min<size_t>(lds_class(need), lds_limit(b->ctx)));
    auto k = std::make_tuple(W.dense, W.col64, lds, W.threads);
    auto it = key.find(k);
    if (it == key.end()) {
      it = key.emplace(k, (int)bins.size()).first;
      Bin nb;
      nb.dense = W.dense;
      nb.col64 = W.col64;
      nb.lds = lds;
      nb.threads = W.threads;
      bins.push_back(nb);
    }
    bins[it->second].jobs.push_back(w);
  }
  for (Bin& bn : bins) {
    // heaviest work items first
    std::stable_sort(bn.jobs.begin(), bn.jobs.end(),
                     [&](int x, int y) { return works[x].weight > works[y].weight; });
    std::vector<PassParams> bp;
    std::vector<uint2> q, mq;
    for (size_t s = 0; s < bn.jobs.size(); ++s) {
      const Work& W = works[bn.jobs[s]];
      bp.push_back(W.params);
      for (uint64_t i = 0; i < W.nitems; ++i) q.push_back(make_uint2((uint32_t)s, (uint32_t)i));
      for (uint64_t h = 0; h < W.nmerge; ++h) mq.push_back(make_uint2((uint32_t)s, (uint32_t)h));
    }
    if (q.size() > 0xFFFFFFFFull || mq.size() > 0xFFFFFFFFull) return set_err(b->ctx, PTC_ERR_RANGE, "batch queue too long");
    bn.nq = (uint32_t)q.size();
    bn.nmq = (uint32_t)mq.size();
    bn.d_params = upload_vec(b, bp);
    bn.d_queue = upload_vec(b, q);
    bn.d_mqueue = upload_vec(b, mq);
    bn.d_head = b->ws->get<uint32_t>(2);
    if (!bn.d_params || (bn.nq && !bn.d_queue) || (bn.nmq && !bn.d_mqueue) || !bn.d_head) return PTC_ERR_NOMEM;
    int rc = pass_batch_grid(b->ctx, pass, bn.dense, bn.col64, 0, false, bn.lds, &bn.grid, bn.threads);
    if (!rc) rc = pass_batch_grid(b->ctx, pass, bn.dense, bn.col64, 0, true, bn.lds, &bn.mgrid, bn.threads);
    if (rc) return rc;
  }
  return PTC_OK;
}

// The works of every job for one pass: the job as one work, or -- for a
// split dense job -- its dense share (heavy rows, with their merges) and its
// small rows in an LDS hash share.
std::vector<Work> make_works(ptc_batch* b, int pass, const std::vector<PassParams>& params,
                             const std::vector<HashPlan*>& hp) {
  std::vector<Work> works;
  for (int j = 0; j < b->n; ++j) {
    const Run& R = *b->runs[j];
    const Split& sp = b->split[j];
    Work W;
    W.job = j;
    W.dense = R.dense;
    W.col64 = R.col64;
    W.threads = R.threads;
    W.lds_slots = R.lds_slots;
    W.params = params[j];
    W.weight = (double)R.F / std::max<uint64_t>(R.nitems, 1);
    const Win& wn = b->win[j];
    if (wn.T) {
      // dense column windows over the large rows
      for (uint32_t t = 0; t < wn.T && (sp.nd || R.nheavy); ++t) {
        Work D = W;
        D.dense = true;
        D.col64 = false;
        D.threads = wn.threads;
        D.lds_slots = 0;
        D.ncols = wn.w[t];
        PassParams& P = D.params;
        P.items = sp.dense;
        P.bcol = wn.col16;
        P.ncols = wn.w[t];
        P.col_lo = wn.lo[t];
        P.col_win = 1;
        P.row_nnz = wn.row_nnz[t];
        P.merge_rows = R.heavy_rows;
        P.row_hcap = nullptr;
        P.row_gcap = nullptr;
        P.row_goff = nullptr;
        P.gkeys = nullptr;
        P.gvals = nullptr;
        if (pass == PASS_SYMBOLIC) {
          P.hbits = wn.hbits[t];
        } else {
          P.hacc = wn.hacc[t];
          P.hflag = wn.hflag[t];
          P.out_base = wn.out_off[t];
          P.out_hedges = b->out[j]->hedges + wn.base[t] * R.dO;
          P.out_vals = b->out[j]->vals + wn.base[t];
        }
        D.nitems = sp.nd;
        D.nmerge = R.nheavy;
        works.push_back(D);
      }
    }
    if (!sp.ns && !wn.T) {
      W.nitems = R.nitems;
      W.nmerge = R.dense ? R.nheavy : ((pass == PASS_NUMERIC && hp[j] && hp[j]->pool) ? R.nrows : 0);
      works.push_back(W);
      continue;
    }
    if (!wn.T && (sp.nd || R.nheavy)) {
      W.params.items = sp.dense;
      W.nitems = sp.nd;
      W.nmerge = R.dense ? R.nheavy : ((pass == PASS_NUMERIC && hp[j] && hp[j]->pool) ? R.nrows : 0);
      works.push_back(W);
    }
    Work H = W;
    H.dense = false;
    H.col64 = false;
    H.threads = SPARSE_THREADS;
    H.lds_slots = SPARSE_SLOTS;
    H.params.items = sp.sparse;
    H.params.bcol = R.HB->rest_key32;  // 32-bit column keys (the dense share streams 16-bit ones)
    H.params.row_hcap = sp.hcap;
    H.params.row_gcap = sp.zero;
    H.params.row_goff = sp.zero;
    H.params.gkeys = nullptr;
    H.params.gvals = nullptr;
    H.params.merge_rows = nullptr;
    H.nitems = sp.ns;
    H.nmerge = 0;
    H.weight = 0.0;
    if (wn.T) {  // the sparse share writes its own part of a windowed job's output
      H.params.row_nnz = wn.row_nnz[wn.T];
      if (pass == PASS_NUMERIC) {
        H.params.out_base = wn.out_off[wn.T];
        H.params.out_hedges = b->out[j]->hedges + wn.base[wn.T] * R.dO;
        H.params.out_vals = b->out[j]->vals + wn.base[wn.T];
      }
    }
    if (sp.ns) works.push_back(H);
  }
  return works;
}

int launch_bins(ptc_batch* b, int pass, std::vector<Bin>& bins, std::vector<FillSeg>* fills) {
  ptc_ctx* ctx = b->ctx;
  int rc = PTC_OK;
  for (Bin& bn : bins) {
    if ((rc = launch_pass_batch(ctx, pass, bn.dense, bn.col64, 0, false, bn.d_params, bn.d_queue, bn.nq, bn.d_head,
                                bn.lds, bn.grid, bn.threads)))
      return rc;
    if ((rc = launch_pass_batch(ctx, pass, bn.dense, bn.col64, 0, true, bn.d_params, bn.d_mqueue, bn.nmq,
                                bn.d_head + 1, bn.lds, bn.mgrid, bn.threads)))
      return rc;
  }
  (void)fills;
  return PTC_OK;
}

int make_fill(ptc_batch* b, const std::vector<FillSeg>& f, FillPlan* fp) {
  std::vector<unsigned long long> st;
  fp->total = 0;
  for (const FillSeg& s : f) {
    st.push_back(fp->total);
    fp->total += s.n;
  }
  fp->nseg = (int)f.size();
  if (f.empty()) return PTC_OK;
  fp->segs = upload_vec(b, f);
  fp->start = upload_vec(b, st);
  return fp->segs && fp->start ? PTC_OK : PTC_ERR_NOMEM;
}

int launch_fill(ptc_batch* b, const FillPlan& fp) {
  if (!fp.nseg || !fp.total) return PTC_OK;
  const unsigned grid = (unsigned)std::min<unsigned long long>((fp.total + 255) / 256, 16384);
  hipLaunchKernelGGL(k_fill_batch, dim3(grid), dim3(256), 0, b->ctx->stream, (const FillSeg*)fp.segs,
                     (const unsigned long long*)fp.start, fp.nseg, fp.total);
  PTC_HIP(b->ctx, hipGetLastError());
  return PTC_OK;
}

// Zeroes the bins' queue heads (before a pass).
int reset_heads(ptc_batch* b, std::vector<Bin>& bins) {
  std::vector<FillSeg> f;
  for (Bin& bn : bins) f.push_back(FillSeg{bn.d_head, 2, 0ull, 0});
  FillPlan fp;
  int rc = make_fill(b, f, &fp);
  return rc ? rc : launch_fill(b, fp);
}

// Key tables of every job (independent of the tuning knobs).
int plan_tables(ptc_batch* b) {
  ptc_ctx* ctx = b->ctx;
  const int n = b->n;
  int rc = PTC_OK;
  stats_mark(b, 0);
  // ---- key tables: B and A_copy of every job (main.cpp:115-116), all
  // built together, no waits
  std::vector<const ptc_dtensor*> ts(2 * (size_t)n);
  for (int j = 0; j < n; ++j) {
    const ptc_dtensor* A = b->A[j];
    const uint32_t c = A->nk;
    if (c == 0 || b->B[j]->nk != c)
      return set_err(ctx, PTC_ERR_ARG, "job %d: A and B must have the same number (>0) of contracted modes", j);
    uint32_t rot[PTC_MAX_MODES];
    for (uint32_t i = c; i < A->d; ++i) rot[i - c] = A->dim_tc[i];
    for (uint32_t i = 0; i < c; ++i) rot[A->d - c + i] = A->dim_tc[i];
    b->Ac[j] = view_with_keys(A, A->d - c, rot);
    ts[2 * j] = b->B[j];
    ts[2 * j + 1] = b->Ac[j];
  }
  std::vector<ptc_hashtable*> hs(2 * (size_t)n, nullptr);
  if ((rc = hashtable_build_many(ctx, 2 * n, ts.data(), hs.data()))) return rc;
  for (int j = 0; j < n; ++j) {
    b->HB[j] = hs[2 * j];
    b->HA[j] = hs[2 * j + 1];
  }
  stats_mark(b, 1);
  return PTC_OK;
}

// Everything after the tables: preprocessing, symbolic passes, outputs and
// the captured numeric graph, under b->opt.
int plan_rest(ptc_batch* b) {
  ptc_ctx* ctx = b->ctx;
  const int n = b->n;
  int rc = PTC_OK;
  // ---- preprocessing: one read-back for all jobs
  unsigned long long* tot = b->ws->get<unsigned long long>(4 * (size_t)n);
  if (!tot) return PTC_ERR_NOMEM;
  for (int j = 0; j < n; ++j) {
    b->runs[j].reset(new Run(ctx));
    b->runs[j]->force_flags = true;  // replays may bring any values
    b->runs[j]->no_windows = true;   // the batch plans its own windows
    if ((rc = prepare_launch(*b->runs[j], b->HA[j], b->HB[j], b->Ac[j], b->B[j], b->opt, tot + 4 * j)))
      return set_err(ctx, rc, "job %d: %s", j, ctx->err.c_str());
  }
  std::vector<unsigned long long> hv(4 * (size_t)n);
  if ((rc = read_back(ctx, hv.data(), tot, sizeof(unsigned long long) * 4 * n))) return rc;
  for (int j = 0; j < n; ++j) {
    Run& R = *b->runs[j];
    if ((rc = prepare_finish(R, &hv[4 * (size_t)j]))) return set_err(ctx, rc, "job %d: %s", j, ctx->err.c_str());
    b->st.flops += R.F;
    b->st.queue_items += (uint32_t)R.nitems;
  }
  // ---- small rows of wide dense jobs to LDS hash shares: item lists split
  // on the device, their sizes read back once
  b->split.assign(n, Split());
  {
    std::vector<const unsigned long long*> csrc(n, nullptr);
    for (int j = 0; j < n; ++j) {
      Run& R = *b->runs[j];
      // dense jobs in wide column spaces, and every hash job (whose LDS
      // tables are sized for its largest rows)
      if ((R.dense && R.ncols < SPLIT_MIN_COLS) || R.col64 || !R.nitems) continue;
      Split& sp = b->split[j];
      sp.dense = b->ws->get<uint4>(R.nitems);
      sp.sparse = b->ws->get<uint4>(R.nitems);
      sp.hcap = b->ws->get<uint32_t>(R.nrows);
      sp.zero = b->ws->get<unsigned long long>(R.nrows);
      uint32_t* cnt = b->ws->get<uint32_t>(2);
      if (!sp.dense || !sp.sparse || !sp.hcap || !sp.zero || !cnt) return PTC_ERR_NOMEM;
      if ((rc = fill_zero(ctx, cnt, 8)) || (rc = fill_zero(ctx, sp.zero, R.nrows * 8))) return rc;
      if ((rc = launch_split_items(ctx, R.items, R.nitems, R.row_flops, R.row_nchunk, sp.dense, sp.sparse, cnt)) ||
          (rc = launch_sparse_caps(ctx, R.row_flops, R.nrows, sp.hcap)))
        return rc;
      csrc[j] = (const unsigned long long*)cnt;  // (dense, sparse) as one 64-bit word
    }
    std::vector<unsigned long long> cnts;
    if ((rc = read_words64(b, csrc, cnts))) return rc;
    for (int j = 0; j < n; ++j) {
      if (!csrc[j]) continue;
      b->split[j].nd = (uint32_t)cnts[j];
      b->split[j].ns = (uint32_t)(cnts[j] >> 32);
    }
  }
  // ---- dense column windows for hash jobs in 16-bit column spaces
  b->win.assign(n, Win());
  {
    const uint64_t lds_cap = lds_limit(ctx);
    uint64_t wfit = 0;  // widest dense window (touched bits kept) at 512 threads
    while (wfit < 65536 && pass_lds_bytes(PASS_NUMERIC, true, false, wfit + 256, 0, true, 512) <= lds_cap) wfit += 256;
    for (int j = 0; j < n && wfit; ++j) {
      Run& R = *b->runs[j];
      const Split& sp = b->split[j];
      if (R.dense || R.col64 || R.ncols > 65536 || !sp.dense) continue;
      const uint64_t T = (R.ncols + wfit - 1) / wfit;
      if (T > MAX_WINDOWS) continue;
      Win& wn = b->win[j];
      if ((rc = hashtable_col16(ctx, R.HB, &wn.col16))) return rc;
      wn.T = (uint32_t)T;
      const uint64_t width = ((R.ncols + T - 1) / T + 63) / 64 * 64;
      uint64_t wmax = 0;
      for (uint32_t t = 0; t < wn.T; ++t) {
        wn.lo[t] = (uint32_t)(t * width);
        wn.w[t] = (uint32_t)std::min<uint64_t>(width, R.ncols - t * width);
        wmax = std::max<uint64_t>(wmax, wn.w[t]);
      }
      const size_t at512 = pass_lds_bytes(PASS_NUMERIC, true, false, wmax, 0, true, 512);
      wn.threads = (2 * at512 > lds_cap && pass_lds_bytes(PASS_NUMERIC, true, false, wmax, 0, true, 1024) <= lds_cap)
                       ? 1024u : 512u;
      for (uint32_t x = 0; x <= wn.T; ++x) {
        wn.row_nnz[x] = b->ws->get<uint32_t>(R.nrows + 1);
        wn.out_off[x] = b->ws->get<unsigned long long>(R.nrows + 1);
        if (!wn.row_nnz[x] || !wn.out_off[x]) return PTC_ERR_NOMEM;
      }
    }
  }
  // ---- symbolic: HBM pools of hash jobs in one read-back, then one
  // persistent launch per bin
  std::vector<HashPlan> hs(n);
  std::vector<const unsigned long long*> src(n, nullptr);
  for (int j = 0; j < n; ++j) {
    Run& R = *b->runs[j];
    if (!R.dense && !b->win[j].T) {
      if ((rc = plan_hash_launch(R, hs[j], nullptr, R.row_flops, nullptr))) return rc;
      src[j] = hs[j].gcap + R.nrows;
    }
  }
  std::vector<unsigned long long> pools;
  if ((rc = read_words64(b, src, pools))) return rc;
  std::vector<PassParams> params(n);
  std::vector<HashPlan*> hptr(n, nullptr);
  for (int j = 0; j < n; ++j) {
    Run& R = *b->runs[j];
    Win& wn = b->win[j];
    if (!R.dense && !wn.T) {
      if ((rc = plan_hash_finish(R, hs[j], pools[j])) || (rc = reset_hash_pool(R, hs[j], false))) return rc;
      hptr[j] = &hs[j];
    }
    if ((rc = fill_zero(ctx, R.row_nnz, (R.nrows + 1) * 4)) || (rc = prepare_heavy(R, PASS_SYMBOLIC))) return rc;
    for (uint32_t x = 0; wn.T && x <= wn.T; ++x) {
      if ((rc = fill_zero(ctx, wn.row_nnz[x], (R.nrows + 1) * 4))) return rc;
      if (x < wn.T && R.nheavy) {  // symbolic slabs of split rows: one bitmap of the window per chunk
        wn.hbits[x] = b->ws->get<uint32_t>(R.nslabs * ((wn.w[x] + 31) / 32));
        if (!wn.hbits[x]) return PTC_ERR_NOMEM;
      }
    }
    base_params(R, params[j], R.dense ? nullptr : &hs[j]);
    if (!R.dense && !wn.T) {
      unsigned long long* flags = nullptr;
      if ((rc = hbm_flags(R, hs[j], &flags))) return rc;
      params[j].row_gcap = flags;
    }
    params[j].merge_rows = R.dense ? R.heavy_rows : nullptr;
  }
  std::vector<Bin> sbins;
  if ((rc = make_bins(b, PASS_SYMBOLIC, make_works(b, PASS_SYMBOLIC, params, hptr), sbins)) ||
      (rc = reset_heads(b, sbins)))
    return rc;
  stats_mark(b, 6);
  if ((rc = launch_bins(b, PASS_SYMBOLIC, sbins, nullptr))) return rc;
  stats_mark(b, 7);
  stats_mark(b, 2);
  // ---- output sizes and the symbolic pass's consistency flags: one read-back
  std::vector<const unsigned long long*> fsrc(n);
  for (int j = 0; j < n; ++j) {
    Run& R = *b->runs[j];
    if ((rc = scan_u32_to_u64(ctx, R.row_nnz, R.out_off, R.nrows, R.out_off + R.nrows))) return rc;
    src[j] = R.out_off + R.nrows;
    fsrc[j] = (const unsigned long long*)R.overflow;  // the first two flag words
  }
  std::vector<unsigned long long> nnz, sflags;
  if ((rc = read_words64(b, src, nnz)) || (rc = read_words64(b, fsrc, sflags))) return rc;
  {  // windowed jobs: each part's size, read back together; parts laid end to end
    std::vector<const unsigned long long*> wsrc;
    for (int j = 0; j < n; ++j) {
      Run& R = *b->runs[j];
      const Win& wn = b->win[j];
      for (uint32_t x = 0; wn.T && x <= wn.T; ++x) {
        if ((rc = scan_u32_to_u64(ctx, wn.row_nnz[x], wn.out_off[x], R.nrows, wn.out_off[x] + R.nrows))) return rc;
        wsrc.push_back(wn.out_off[x] + R.nrows);
      }
    }
    std::vector<unsigned long long> wtot;
    if ((rc = read_words64(b, wsrc, wtot))) return rc;
    size_t k = 0;
    for (int j = 0; j < n; ++j) {
      Win& wn = b->win[j];
      if (!wn.T) continue;
      unsigned long long run = 0;
      for (uint32_t x = 0; x <= wn.T; ++x) {
        wn.base[x] = run;
        run += wtot[k++];
      }
      nnz[j] = run;
    }
  }
  for (int j = 0; j < n; ++j)
    if ((uint32_t)sflags[j] & 0xC0000000u)
      return set_err(ctx, PTC_ERR_INTERNAL, "job %d: symbolic pass consistency flags %08x", j, (uint32_t)sflags[j]);
  for (int j = 0; j < n; ++j) {
    if ((rc = make_output(*b->runs[j], nnz[j], b->Ac[j], b->B[j], &b->out[j]))) return rc;
    b->st.nnz_out += nnz[j];
  }
  // ---- numeric plan: HBM pools from the exact counts, one read-back
  for (int j = 0; j < n; ++j) {
    Run& R = *b->runs[j];
    src[j] = nullptr;
    if (!R.dense && !b->win[j].T) {
      if ((rc = plan_hash_launch(R, b->hn[j], R.row_nnz, nullptr, nullptr))) return rc;
      src[j] = b->hn[j].gcap + R.nrows;
    }
  }
  if ((rc = read_words64(b, src, pools))) return rc;
  std::vector<RefreshSeg> rseg;
  std::vector<unsigned long long> rstart;
  unsigned long long rtotal = 0;
  std::vector<FillSeg> fills;
  for (int j = 0; j < n; ++j) {
    Run& R = *b->runs[j];
    Win& wn = b->win[j];
    hptr[j] = nullptr;
    for (uint32_t x = 0; x < wn.T && R.nheavy; ++x) {  // numeric slabs of split rows per window
      wn.hacc[x] = b->ws->get<double>(R.nslabs * wn.w[x]);
      wn.hflag[x] = b->ws->get<unsigned char>(R.nslabs * wn.w[x]);
      if (!wn.hacc[x] || !wn.hflag[x]) return PTC_ERR_NOMEM;
    }
    if (!R.dense && !wn.T) {
      if ((rc = plan_hash_finish(R, b->hn[j], pools[j]))) return rc;
      hptr[j] = &b->hn[j];
      if (b->hn[j].pool) {
        fills.push_back(FillSeg{b->hn[j].gkeys, b->hn[j].pool, R.col64 ? ~0ull : 0xFFFFFFFFull, R.col64 ? 1 : 0});
        fills.push_back(FillSeg{b->hn[j].gvals, b->hn[j].pool, 0ull, 1});
      }
    }
    if ((rc = prepare_heavy(R, PASS_NUMERIC))) return rc;
    base_params(R, params[j], R.dense ? nullptr : &b->hn[j]);
    if (!R.dense && !wn.T) {
      unsigned long long* flags = nullptr;
      if ((rc = hbm_flags(R, b->hn[j], &flags))) return rc;
      params[j].row_gcap = flags;
    }
    params[j].merge_rows = R.dense ? R.heavy_rows : nullptr;
    params[j].out_base = R.out_off;
    params[j].out_hedges = b->out[j]->hedges;
    params[j].out_vals = b->out[j]->vals;
    fills.push_back(FillSeg{R.overflow, 1, 0ull, 0});
    // grouped values of both tables, re-read from the input tensors
    const ptc_hashtable* tabs[2] = {b->HA[j], b->HB[j]};
    for (const ptc_hashtable* t : tabs) {
      rseg.push_back(RefreshSeg{t->perm, t->t->vals, t->gval, t->t->N});
      rstart.push_back(rtotal);
      rtotal += t->t->N;
    }
  }
  if ((rc = make_bins(b, PASS_NUMERIC, make_works(b, PASS_NUMERIC, params, hptr), b->bins))) return rc;
  for (Bin& bn : b->bins) fills.push_back(FillSeg{bn.d_head, 2, 0ull, 0});
  RefreshSeg* d_rseg = upload_vec(b, rseg);
  unsigned long long* d_rstart = upload_vec(b, rstart);
  FillPlan fplan;
  if ((rc = make_fill(b, fills, &fplan))) return rc;
  std::vector<const uint32_t*> ovf_src(n);
  for (int j = 0; j < n; ++j) ovf_src[j] = b->runs[j]->overflow;
  b->d_ovf_src = (uint32_t**)upload_vec(b, ovf_src);
  b->d_ovf = b->ws->get<uint32_t>(n);
  if (!d_rseg || !d_rstart || !b->d_ovf_src || !b->d_ovf) return PTC_ERR_NOMEM;
  b->st.bins = (uint32_t)b->bins.size();
  // ---- capture the numeric phase (no allocation, read-back or memset inside)
  PTC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  stats_mark(b, 3);
  PTC_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const unsigned rgrid = (unsigned)std::min<unsigned long long>((rtotal + 255) / 256, 16384);
  hipLaunchKernelGGL(k_refresh_batch, dim3(std::max(1u, rgrid)), dim3(256), 0, ctx->stream, (const RefreshSeg*)d_rseg,
                     (const unsigned long long*)d_rstart, (int)rseg.size(), rtotal);
  rc = launch_fill(b, fplan);
  if (!rc) rc = launch_bins(b, PASS_NUMERIC, b->bins, nullptr);
  hipError_t ce = hipStreamEndCapture(ctx->stream, &b->graph);
  if (rc) return rc;
  if (ce != hipSuccess) return hip_check(ctx, ce, "graph capture");
  PTC_HIP(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  size_t nodes = 0;
  if (hipGraphGetNodes(b->graph, nullptr, &nodes) == hipSuccess) b->st.graph_nodes = nodes;
  b->st.ms_build = stats_ms(b, 0, 1);
  b->st.ms_symbolic = stats_ms(b, 1, 2);
  b->st.ms_symbolic_pass = stats_ms(b, 6, 7);
  return PTC_OK;
}

// Drops what plan_rest made, keeping the key tables.
void clear_rest(ptc_batch* b) {
  (void)hipStreamSynchronize(b->ctx->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  for (auto*& o : b->out) {
    ptc_dtensor_free(o);
    o = nullptr;
  }
  for (auto& r : b->runs) r.reset();
  b->hn.assign(b->n, HashPlan());
  b->bins.clear();
  delete b->ws;
  b->ws = new Workspace(b->ctx);
  b->host_keep.clear();
  b->d_ovf_src = nullptr;
  b->d_ovf = nullptr;
  const double ms_build = b->st.ms_build;
  b->st = ptc_batch_stats{};
  b->st.jobs = (uint32_t)b->n;
  b->st.ms_build = ms_build;
}

}  // namespace

extern "C" int ptc_batch_plan(ptc_ctx* ctx, int n, const ptc_dtensor* const* A, const ptc_dtensor* const* B,
                              const ptc_options* opt, ptc_batch** out, ptc_batch_stats* st) {
  if (!ctx || n <= 0 || !A || !B || !out) return PTC_ERR_ARG;
  *out = nullptr;
  for (int j = 0; j < n; ++j)
    if (!A[j] || !B[j]) return set_err(ctx, PTC_ERR_ARG, "job %d: NULL tensor", j);
  PTC_HIP(ctx, hipSetDevice(ctx->device));
  auto w0 = std::chrono::steady_clock::now();
  ptc_batch* b = new ptc_batch();
  b->ctx = ctx;
  ctx_acquire(ctx);
  b->n = n;
  if (opt) b->opt = *opt; else ptc_options_default(&b->opt);
  b->opt.method = PTC_METHOD_TWO_PASS;
  if (b->opt.value_type != PTC_VAL_F64) {
    ptc_batch_free(b);
    return set_err(ctx, PTC_ERR_UNSUPPORTED, "batched contractions compute in FP64");
  }
  b->A.assign(A, A + n);
  b->B.assign(B, B + n);
  b->Ac.assign(n, nullptr);
  b->HA.assign(n, nullptr);
  b->HB.assign(n, nullptr);
  b->runs.resize(n);
  b->hn.resize(n);
  b->out.assign(n, nullptr);
  b->ws = new Workspace(ctx);
  b->st.jobs = (uint32_t)n;
  int rc = PTC_OK;
  for (auto& e : b->ev)
    if (!rc && hipEventCreate(&e) != hipSuccess) rc = set_err(ctx, PTC_ERR_HIP, "batch timing events");
  if (!rc) rc = plan_tables(b);
  if (!rc) rc = plan_rest(b);
  if (rc) {
    (void)hipGetLastError();
    ptc_batch_free(b);
    return rc;
  }
  b->st.ms_plan = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
  if (st) *st = b->st;
  *out = b;
  return PTC_OK;
}

extern "C" int ptc_batch_run(ptc_batch* b, ptc_batch_stats* st) {
  if (!b || !b->exec) return PTC_ERR_ARG;
  ptc_ctx* ctx = b->ctx;
  PTC_HIP(ctx, hipSetDevice(ctx->device));
  stats_mark(b, 4);
  PTC_HIP(ctx, hipGraphLaunch(b->exec, ctx->stream));
  stats_mark(b, 5);
  // consistency flags of every job, one read-back
  hipLaunchKernelGGL(k_gather_words, dim3((unsigned)std::min(1024, (b->n + 255) / 256)), dim3(256), 0, ctx->stream,
                     (const uint32_t* const*)b->d_ovf_src, b->n, b->d_ovf);
  std::vector<uint32_t> ovf(b->n);
  int rc = read_back(ctx, ovf.data(), b->d_ovf, sizeof(uint32_t) * b->n);
  if (rc) return rc;
  for (int j = 0; j < b->n; ++j) {
    if (ovf[j] & 0x80000000u) return set_err(ctx, PTC_ERR_INTERNAL, "job %d: row count mismatch between passes", j);
    if (ovf[j] & 0x40000000u) return set_err(ctx, PTC_ERR_INTERNAL, "job %d: slice index out of range", j);
  }
  b->st.ms_numeric = stats_ms(b, 4, 5);
  b->st.runs++;
  if (st) *st = b->st;
  return PTC_OK;
}

extern "C" int ptc_batch_output(const ptc_batch* b, int j, const ptc_dtensor** O) {
  if (!b || !O || j < 0 || j >= b->n) return PTC_ERR_ARG;
  *O = b->out[j];
  return PTC_OK;
}

extern "C" void ptc_batch_free(ptc_batch* b) {
  if (!b) return;
  ptc_ctx* ctx = b->ctx;
  // after ptc_destroy: no HIP call (the context released its memory); the
  // tables and outputs below release their own references the same way
  const bool open = ctx_open(ctx);
  if (open) {
    (void)hipStreamSynchronize(ctx->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
  }
  for (auto* o : b->out) ptc_dtensor_free(o);
  b->runs.clear();
  for (auto* h : b->HA) ptc_hashtable_free(h);
  for (auto* h : b->HB) ptc_hashtable_free(h);
  for (auto* v : b->Ac) delete v;
  delete b->ws;
  if (open)
    for (auto& e : b->ev)
      if (e) (void)hipEventDestroy(e);
  delete b;
  ctx_release(ctx);
}

// Coordinate descent over the knobs that shape a batch's row passes:
// dense LDS column limit (LDS vs hash accumulators), LDS hash slots (LDS vs
// HBM accumulators) and work per item (chunk).  Each trial re-plans the
// batch on the same key tables and times its symbolic phase and one numeric
// graph launch on the device.
extern "C" int ptc_batch_autotune(ptc_ctx* ctx, int n, const ptc_dtensor* const* A, const ptc_dtensor* const* B,
                                  const ptc_options* base, ptc_options* best, ptc_tune_report* report) {
  if (!ctx || n <= 0 || !A || !B || !best) return PTC_ERR_ARG;
  auto w0 = std::chrono::steady_clock::now();
  ptc_options cur;
  if (base) cur = *base; else ptc_options_default(&cur);
  if (!cur.dense_max_cols) cur.dense_max_cols = DEFAULT_DENSE_MAX;
  if (!cur.lds_hash_max) cur.lds_hash_max = DEFAULT_LDS_HASH;
  // chunk_flops 0 stays "automatic" (a rule of the accumulator kind and the
  // column space) and is one of the candidates
  ptc_batch* b = nullptr;
  // the tables, once, through a plan under the starting options
  int rc = ptc_batch_plan(ctx, n, A, B, &cur, &b, nullptr);
  if (rc) return rc;
  ptc_tune_report rep{};
  auto trial = [&](const ptc_options& o, double* ms) -> int {
    clear_rest(b);
    b->opt = o;
    b->opt.method = PTC_METHOD_TWO_PASS;
    int r = plan_rest(b);
    if (!r) r = ptc_batch_run(b, nullptr);  // warm
    if (!r) r = ptc_batch_run(b, nullptr);
    if (r) return r;
    *ms = b->st.ms_symbolic_pass + b->st.ms_numeric;
    if (rep.trials < PTC_TUNE_MAX_TRIALS) {
      rep.trial_dense[rep.trials] = o.dense_max_cols;
      rep.trial_hash[rep.trials] = o.lds_hash_max;
      rep.trial_chunk[rep.trials] = o.chunk_flops;
      rep.trial_ms[rep.trials] = *ms;
    }
    rep.trials++;
    return PTC_OK;
  };
  double best_ms = 0;
  if ((rc = trial(cur, &best_ms))) {
    ptc_batch_free(b);
    return rc;
  }
  rep.default_ms = best_ms;
  const uint32_t dense_c[] = {2048, 4096, 8192, 16384};
  const uint32_t hash_c[] = {512, 1024, 2048, 4096, 8192};
  const uint64_t chunk_c[] = {0, 16384, 32768, 65536, 131072};
  for (int knob = 0; knob < 3 && !rc; ++knob) {
    const int nc = knob == 0 ? 4 : 5;
    for (int k = 0; k < nc && !rc; ++k) {
      ptc_options o = cur;
      if (knob == 0) {
        if (dense_c[k] == cur.dense_max_cols) continue;
        o.dense_max_cols = dense_c[k];
      } else if (knob == 1) {
        if (hash_c[k] == cur.lds_hash_max) continue;
        o.lds_hash_max = hash_c[k];
      } else {
        if (chunk_c[k] == cur.chunk_flops) continue;
        o.chunk_flops = chunk_c[k];
      }
      double ms = 0;
      rc = trial(o, &ms);
      if (!rc && ms < best_ms) {
        best_ms = ms;
        cur = o;
      }
    }
  }
  ptc_batch_free(b);
  if (rc) return rc;
  *best = cur;
  rep.best_ms = best_ms;
  rep.dense_max_cols = cur.dense_max_cols;
  rep.lds_hash_max = cur.lds_hash_max;
  rep.chunk_flops = cur.chunk_flops;
  rep.ms_wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
  if (report) *report = rep;
  return PTC_OK;
}

extern "C" int ptc_contract_batched(ptc_ctx* ctx, int n, const ptc_dtensor* const* A, const ptc_dtensor* const* B,
                                    const ptc_options* opt, ptc_dtensor** O, ptc_batch_stats* st) {
  if (!O) return PTC_ERR_ARG;
  auto w0 = std::chrono::steady_clock::now();
  ptc_batch* b = nullptr;
  int rc = ptc_batch_plan(ctx, n, A, B, opt, &b, nullptr);
  if (rc) return rc;
  rc = ptc_batch_run(b, nullptr);
  if (!rc) {
    for (int j = 0; j < n; ++j) {
      O[j] = b->out[j];  // ownership passes to the caller
      b->out[j] = nullptr;
    }
    b->st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    if (st) *st = b->st;
  }
  ptc_batch_free(b);
  return rc;
}
