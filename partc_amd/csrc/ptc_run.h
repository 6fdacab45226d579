// Host-side state of one contraction and the steps it is made of, shared by
// the single-contraction entry points (ptc_tc.hip) and the batched driver
// (ptc_batch.hip).  Steps that end in a device -> host read are split into a
// launch half and a finish half, so a batch reads all its jobs' values back
// in one round trip.
#ifndef PTC_RUN_H
#define PTC_RUN_H

#include "ptc_contract.h"
#include "ptc_internal.h"

namespace ptc {

constexpr uint32_t DEFAULT_DENSE_MAX = 8192;
constexpr uint32_t DEFAULT_LDS_HASH = 2048;
constexpr uint64_t DEFAULT_CHUNK = 32768;         // multiply-adds per work item, hash accumulators
constexpr uint64_t DEFAULT_DENSE_CHUNK = 65536;   // floor for dense accumulators
constexpr uint64_t DENSE_MAX_ITEMS = 3072;         // dense chunks grow to total flops / this
constexpr uint32_t MAX_WINDOWS = 4;                // dense column windows of one contraction

// Hash accumulator plan of every row: LDS slots or an HBM region.
struct HashPlan {
  uint32_t* hcap = nullptr;
  unsigned long long* gcap = nullptr;  // scanned in place into offsets
  unsigned long long pool = 0;         // total HBM slots
  void* gkeys = nullptr;
  double* gvals = nullptr;
};

struct Run {
  ptc_ctx* ctx;
  Workspace ws;
  const ptc_hashtable* HA;
  const ptc_hashtable* HB;
  ptc_options opt;
  bool dense = true, col64 = false;
  uint64_t npos = 0, nrows = 0, ncols = 0;
  uint32_t dO = 0, nrowmodes = 0, lds_slots = DEFAULT_LDS_HASH;
  // preprocessing
  uint32_t *pos_len = nullptr, *pos_off = nullptr;
  unsigned long long* apre = nullptr;
  uint32_t *row_hits = nullptr, *row_maxlen = nullptr, *row_nchunk = nullptr, *row_heavy = nullptr;
  uint32_t* row_nnz = nullptr;
  unsigned long long *row_flops = nullptr, *packed = nullptr;
  uint4* items = nullptr;
  uint64_t F = 0, nitems = 0, nheavy = 0;
  // heavy dense merge buffers
  uint32_t* hbits = nullptr;
  double* hacc = nullptr;
  unsigned char* hflag = nullptr;
  unsigned long long* row_slab = nullptr;  // first partial-row slab of each split row
  uint32_t* heavy_rows = nullptr;          // split rows (heavy index -> row)
  uint64_t nslabs = 0;
  bool use_flags = true;
  uint32_t* overflow = nullptr;
  unsigned long long* out_off = nullptr;
  bool want_features = false;  // per-row hit counts and longest slices (learned features)
  bool early_items = false;    // work items built before the preprocessing read-back (bounded count)
  bool allow_ahead = false;    // the caller lets a single pass run ahead of the preprocessing totals
  bool ahead = false;          // ... and this one does
  // the single pass's row counts and reservation counter (cnt_words + 4
  // words), already cleared on the device by the preprocessing (or NULL)
  uint32_t* cnt_block = nullptr;
  const uint32_t* row_of_pos = nullptr;  // HT_A's slice index of every A position
  // single-pass methods: every row's output capacity, their offsets and sum
  // (computed with the preprocessing, the sum read back with its totals)
  uint32_t* cap = nullptr;
  unsigned long long* cap_off = nullptr;
  unsigned long long capsum = 0;
  WordSources totw{};  // the preprocessing totals' device words
  uint32_t threads = PB;       // threads per row-pass workgroup
  bool force_flags = false;    // touch flags whatever the values (set before prepare)
  const uint16_t* bcol16 = nullptr;  // dense rows: B's column keys as 16 bits
  bool f32 = false;            // FP32 values (both tables carry gval32)
  int val = PTC_VAL_F64;       // ptc_value_type of the numeric pass
  bool col_first = false;      // output modes: column modes, then row modes
  // Dense column windows (a 16-bit column space too wide for one dense LDS
  // accumulator): window t holds columns [t * win_width, (t + 1) * win_width).
  bool no_windows = false;     // keep hash accumulators (set before prepare)
  uint32_t win_n = 0, win_width = 0, win_threads = PB;
  explicit Run(ptc_ctx* c) : ctx(c), ws(c) {}
};

uint64_t lds_limit(ptc_ctx* ctx);

// Preprocessing (key probes, flop prefix, rows, work items and, for the
// single-pass methods, the rows' output capacities).  The launch half copies
// four totals to tot4 (device): flops, packed item/heavy counts, slabs,
// HT_A's key total -- and for a single-pass method a fifth, the capacities'
// sum; the finish half takes them on the host.
int prepare_launch(Run& R, const ptc_hashtable* HT_A, const ptc_hashtable* HT_B, const ptc_dtensor* A,
                   const ptc_dtensor* B, const ptc_options& opt, unsigned long long* tot4);
int prepare_finish(Run& R, const unsigned long long* hostv);
// Work items and split-row list into arrays of the given sizes (npos = 0:
// no chunk starts).
int build_items(Run& R, uint64_t nitems, uint64_t nheavy, uint64_t npos);
// Words of a single pass's row counts before its reservation counter.
uint64_t count_words(uint64_t nrows);
// The arrays of build_items and the parameters of its two launches.
int item_params(Run& R, uint64_t nitems, uint64_t nheavy, uint64_t npos, ItemParams& P, ItemP0Params& Q);
int prepare(Run& R, const ptc_hashtable* HT_A, const ptc_hashtable* HT_B, const ptc_dtensor* A,
            const ptc_dtensor* B, const ptc_options& opt);

// Hash accumulator plan: per-row LDS slots or HBM regions.  The launch half
// leaves the HBM pool size at hp.gcap[R.nrows] (device).
int plan_hash_launch(Run& R, HashPlan& hp, const uint32_t* bound32, const unsigned long long* bound64,
                     const unsigned long long* hbm_bound);
int plan_hash_finish(Run& R, HashPlan& hp, unsigned long long pool);
int plan_hash(Run& R, HashPlan& hp, const uint32_t* bound32, const unsigned long long* bound64,
              const unsigned long long* hbm_bound);
int reset_hash_pool(Run& R, HashPlan& hp, bool values);
int hbm_flags(Run& R, HashPlan& hp, unsigned long long** flags);
void base_params(Run& R, PassParams& P, const HashPlan* hp);
int prepare_heavy(Run& R, int pass);
int make_output(Run& R, uint64_t nnz, const ptc_dtensor* A, const ptc_dtensor* B, ptc_dtensor** O);
ptc_dtensor* view_with_keys(const ptc_dtensor* t, uint32_t nk, const uint32_t* dim_tc);

// Output arrays the caller owns (cap entries of d_O coordinates and a value):
// the exact two-pass contraction writes into them instead of allocating an
// output.  *nnz receives the output's size, also when cap is too small
// (PTC_ERR_RANGE, nothing written).
struct OutInto {
  uint32_t* hedges;
  double* vals;
  uint64_t cap;
  uint64_t* nnz;
};
int tc_into(ptc_ctx* ctx, const ptc_hashtable* HT_A, const ptc_hashtable* HT_B, const ptc_dtensor* A,
            const ptc_dtensor* B, const ptc_options* opt, const OutInto& into, ptc_stats* stats);

}  // namespace ptc

#endif
