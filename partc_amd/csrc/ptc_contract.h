// Kernel parameter blocks and launchers of the contraction passes.
#ifndef PTC_CONTRACT_H
#define PTC_CONTRACT_H

#include "ptc_internal.h"

#ifndef PTC_PASS_THREADS
#define PTC_PASS_THREADS 512
#endif

namespace ptc {

enum { PASS_SYMBOLIC = 0, PASS_NUMERIC = 1 };
// Largest work item, in multiply-adds: a row pass keeps flop offsets inside
// an item as 32-bit values (staged position starts, spans), so an item's
// range stays below 2^32 with room for the per-wave rounding.
constexpr unsigned long long MAX_ITEM_FLOPS = 1ull << 31;
constexpr int PB = PTC_PASS_THREADS;  // default threads per workgroup of the row passes
constexpr int PB_MAX = 1024;          // largest row-pass workgroup (its launch bound)

struct ProbeParams {
  uint64_t npos;
  const void* akey;               // HT_A rest keys (A's contracted tuple)
  KeyCodec acodec, bcodec;        // A's contracted codec, B's key codec
  int same_codec;                 // the two codecs linearise identically
  const KtEntry* tab;             // HT_B entries
  uint64_t cap;                   // HT_B slots
  int direct;                     // HT_B addressed by key
  uint32_t* pos_len;
  uint32_t* pos_off;
  uint32_t* zero4;                // non-NULL: four words cleared by the launch (the passes' flags)
};

struct RowParams {
  uint64_t nrows;                 // rows allocated (a bound when live is set)
  const unsigned long long* live; // HT_A's device total: rows = *live >> 32 (NULL: nrows)
  const KtSlice* rows;
  const unsigned long long* apre;
  unsigned long long chunk;
  // non-NULL: the contraction's flop total on the device; chunks then grow
  // to total / max_items when that is larger (few, large work items for
  // very large contractions)
  const unsigned long long* total;
  uint64_t max_items;
  uint32_t* row_nchunk;
  unsigned long long* row_flops;
  unsigned long long* packed;     // (heavy << 32) | nchunk, scanned later
  unsigned long long* slabs;      // nchunk of split rows, 0 otherwise (scanned later)
  // row features (non-NULL): zeroed for rows past the table's key count;
  // the positions' scan wrote every other row's
  uint32_t* row_hits;
  uint32_t* row_maxlen;
};

struct ItemParams {
  uint64_t nrows;
  const KtSlice* rows;
  const uint32_t* row_nchunk;
  const unsigned long long* packed;  // scanned
  uint32_t* row_heavy;
  uint32_t* heavy_rows;              // heavy index -> row
  uint4* items;
};

struct ItemP0Params {
  uint64_t nitems;                   // work items (a bound when live is set)
  const unsigned long long* live;    // non-NULL: the item count on the device (low 32 bits)
  int find_rows;                     // write whole items, each row found in packed (k_items did not run)
  uint64_t nrows;                    // rows of packed (find_rows)
  uint64_t npos;                     // A positions (none: no launch)
  const uint32_t* row_nchunk;
  const KtSlice* rows;
  const unsigned long long* apre;
  const unsigned long long* row_flops;
  const unsigned long long* packed;
  uint4* items;
};

struct HashCapParams {
  uint64_t nrows;
  const uint32_t* bound32;            // one of the two
  const unsigned long long* bound64;
  unsigned long long cap_col;         // clamp bound to the column space (0: none)
  const uint32_t* row_nchunk;
  uint32_t lds_slots;
  uint32_t* row_hcap;
  unsigned long long* gcap;           // scanned into offsets later
};

struct CapParams {
  uint64_t nrows;
  const unsigned long long* row_flops;
  const uint32_t* row_maxlen;
  unsigned long long ncols;
  const float* pred;                  // NULL: upper bound
  double scale;
  uint32_t* cap;
};

// Parameters of a row pass.  The field list is written once and expanded
// twice: PassParams, filled by host code, and PassParamsG, which in device
// code is the same layout with every pointer typed as a global-memory
// pointer.  A kernel that reads its parameters from memory (the persistent
// batch pass) uses PassParamsG, so its loads and atomics are global
// instructions rather than flat ones, which would also count against the LDS
// wait counter and serialise global loads with LDS traffic.
#define PTC_PASS_FIELDS(PTR)                                                                  \
  /* work */                                                                                  \
  PTR(const uint4) items;                                                                     \
  PTR(const KtSlice) rows;                                                                    \
  PTR(const uint32_t) row_nchunk;                                                             \
  PTR(const uint32_t) row_heavy;                                                              \
  /* positions (A, row-grouped) */                                                            \
  PTR(const unsigned long long) apre;                                                         \
  PTR(const uint32_t) pos_len;                                                                \
  PTR(const uint32_t) pos_off;                                                                \
  PTR(const double) aval;                                                                     \
  PTR(const float) aval32; /* FP32 path: A values in grouped order */                         \
  /* B slices */                                                                              \
  PTR(const void) bcol;                                                                       \
  PTR(const double) bval;                                                                     \
  PTR(const float) bval32; /* FP32 path: B values in grouped order */                         \
  uint64_t nb;             /* number of B nonzeros */                                         \
  uint64_t ncols;          /* column space (dense accumulators) */                            \
  /* heavy dense rows: symbolic bitmaps; numeric slabs (one partial row per chunk, */         \
  /* first slab of row r at row_slab[r]) */                                                   \
  PTR(uint32_t) hbits;                                                                        \
  PTR(double) hacc;                                                                           \
  PTR(unsigned char) hflag;                                                                   \
  PTR(const uint32_t) merge_rows; /* merge launch: rows to finish (NULL: all rows) */         \
  /* a grid sized by a bound: (split rows << 32 | items) on the device; work past it exits */ \
  PTR(const unsigned long long) live_items;                                                   \
  PTR(const unsigned long long) row_slab;                                                     \
  int use_flags; /* 1: a flag byte per column marks touches */                                \
  /* hash accumulators */                                                                     \
  PTR(const uint32_t) row_hcap;                                                               \
  PTR(const unsigned long long) row_gcap;                                                     \
  PTR(const unsigned long long) row_goff;                                                     \
  PTR(void) gkeys;                                                                            \
  PTR(double) gvals;                                                                          \
  /* results */                                                                               \
  PTR(uint32_t) row_nnz;                                                                      \
  PTR(const unsigned long long) out_base;                                                     \
  PTR(const uint32_t) out_cap; /* capacity mode when set */                                   \
  /* reservation mode (out_count set): a finished row takes its entries with one fetch-add */ \
  /* on out_count, never past out_limit, and records its first entry in out_rowoff */         \
  PTR(unsigned long long) out_count;                                                          \
  uint64_t out_limit;                                                                         \
  PTR(unsigned long long) out_rowoff;                                                         \
  /* out_soft = 1 (with out_count): no per-row capacity; a row part whose reservation would */ \
  /* pass out_limit writes nothing and is counted in overflow (the caller redoes the run) */  \
  uint32_t out_soft, out_pad;                                                                 \
  /* dense symbolic pass: the touched columns of each window [t * win_width, ...) counted */  \
  /* into win_cnt[t * win_stride + row] (win_n windows; win_cnt NULL: the row total only) */  \
  PTR(uint32_t) win_cnt;                                                                      \
  uint64_t win_stride;                                                                        \
  uint32_t win_width, win_n;                                                                  \
  PTR(uint32_t) overflow;                                                                     \
  PTR(void) stg_col; /* capacity mode staging (col key) */                                    \
  PTR(double) stg_val;                                                                        \
  PTR(uint32_t) out_hedges; /* exact mode: final output */                                    \
  PTR(double) out_vals;                                                                       \
  PTR(float) out_vals32; /* FP32 path: hi = fl(sum) and lo = sum - hi */                      \
  PTR(float) out_err32;                                                                       \
  int f32; /* FP32 values in, (hi, lo) FP32 pairs out */                                      \
  int val; /* ptc_value_type: accumulation mode of the numeric pass */                        \
  int col_first; /* 1: an output entry lists its column modes, then its row modes */          \
  /* dense column window: the accumulator holds columns [col_lo, col_lo + ncols) and the */  \
  /* row's other columns are left to other windows (col_win = 1) */                          \
  uint32_t col_lo, col_win;                                                                   \
  uint32_t dO, nrowmodes;                                                                     \
  KeyCodec rowc, colc;

#define PTC_PTR(T) T*
struct PassParams {
  PTC_PASS_FIELDS(PTC_PTR)
};
#if defined(__HIP_DEVICE_COMPILE__)
#define PTC_GPTR(T) __attribute__((address_space(1))) T*
#else
#define PTC_GPTR(T) T*  // host code never fills the device view
#endif
struct PassParamsG {
  PTC_PASS_FIELDS(PTC_GPTR)
};
static_assert(sizeof(PassParamsG) == sizeof(PassParams), "device view of PassParams");

struct EmitParams {
  uint64_t nrows;
  const KtSlice* rows;
  const uint32_t* row_nnz;
  const uint32_t* cap;
  const uint32_t* row_fb;             // rows redone by the exact fallback (skipped)
  const unsigned long long* cap_base;
  const unsigned long long* out_base;
  const void* stg_col;
  const double* stg_val;
  uint32_t* out_hedges;
  double* out_vals;
  float* out_vals32;
  float* out_err32;
  int f32;
  uint32_t dO, nrowmodes;
  KeyCodec rowc, colc;
};

int launch_probe(ptc_ctx* ctx, const ProbeParams& P, bool rest64);
int launch_rows(ptc_ctx* ctx, const RowParams& P);
// k_rows, the scans of packed and slabs, the split-row list of k_items,
// (cap non-NULL) the upper-bound capacities and the clear of nzero words at
// zero, in one workgroup: for nrows + 1 <= ROWS_SMALL_MAX
constexpr int ROWS_SMALL_MAX_K = 8;
constexpr uint64_t ROWS_SMALL_MAX = 1024ull * ROWS_SMALL_MAX_K;
int launch_rows_small(ptc_ctx* ctx, const RowParams& P, const ItemParams& Q, uint32_t* cap, unsigned long long ncols,
                      uint32_t* zero, uint32_t nzero);
// k_item_starts alone (the items written by launch_rows_small)
int launch_item_starts(ptc_ctx* ctx, const ItemP0Params& Q);
int launch_items(ptc_ctx* ctx, const ItemParams& P, const ItemP0Params& Q);
int launch_hash_caps(ptc_ctx* ctx, const HashCapParams& P);
int launch_caps(ptc_ctx* ctx, const CapParams& P);
// val: the numeric pass's ptc_value_type (FP32 accumulators take 12 bytes a
// dense column: an FP32 sum and an FP64 compensation)
size_t pass_lds_bytes(int pass, bool dense, bool col64, uint64_t ncols, uint32_t lds_slots, bool flags = true,
                      uint32_t threads = PB, int val = 0);
int launch_pass(ptc_ctx* ctx, int pass, bool dense, bool col64, const PassParams& P, uint64_t nitems,
                size_t lds, bool merge = false, uint32_t threads = PB);
int launch_emit(ptc_ctx* ctx, const EmitParams& P, bool col64);
// Persistent row pass over the queued work of several contractions (jobs:
// device array of their PassParams; queue: (job, work index) pairs; head: a
// zeroed device counter).  grid 0 = as many workgroups as fit on the device.
int launch_pass_batch(ptc_ctx* ctx, int pass, bool dense, bool col64, int val, bool merge, const PassParams* jobs,
                      const uint2* queue, uint32_t nq, uint32_t* head, size_t lds, unsigned grid = 0,
                      uint32_t threads = PB);
// The same kernel's device-filling grid at this LDS size (opens the LDS
// attribute too: call before a graph capture, then launch with the grid).
int pass_batch_grid(ptc_ctx* ctx, int pass, bool dense, bool col64, int val, bool merge, size_t lds, unsigned* grid,
                    uint32_t threads = PB);
int launch_overflow_rows(ptc_ctx* ctx, const uint32_t* nnz, const uint32_t* cap, uint64_t nrows,
                         uint32_t* list, uint32_t* n, uint32_t* mark);
int launch_select_items(ptc_ctx* ctx, const uint4* items, uint64_t nitems, const uint32_t* nnz,
                        const uint32_t* cap, uint4* out, uint32_t* n);
int launch_reset_rows(ptc_ctx* ctx, const uint32_t* list, uint32_t n, uint32_t* counts);

// Small rows in wide column spaces: at most SPARSE_FLOPS multiply-adds (so
// at most that many distinct columns) and one work item.  They run in an LDS
// hash pass of SPARSE_THREADS-thread workgroups with tables of at most
// SPARSE_SLOTS slots (about 29 KB of LDS: five workgroups per CU) instead of
// a dense accumulator spanning every column (up to ~144 KB: one workgroup
// per CU), so many small rows are in flight per CU at once and none pays for
// the whole column space.
constexpr unsigned long long SPARSE_FLOPS = 1024;
constexpr uint32_t SPARSE_SLOTS = 2048;
constexpr uint32_t SPARSE_THREADS = 128;
// items split by row kind into dense[] and sparse[] (cnt[0], cnt[1]: their
// counts, zeroed by the caller; order within each list is free)
int launch_split_items(ptc_ctx* ctx, const uint4* items, uint64_t nitems, const unsigned long long* row_flops,
                       const uint32_t* row_nchunk, uint4* dense, uint4* sparse, uint32_t* cnt);
// LDS hash slots of the small rows: a power of two >= 2 x their flops (>= 16)
int launch_sparse_caps(ptc_ctx* ctx, const unsigned long long* row_flops, uint64_t nrows, uint32_t* hcap);
// Column windows: off[t * nrows + r] = out_off[r] + sum of cnt[t' * nrows + r]
// over t' < t (each window's first output entry in its row's range).
int launch_window_offsets(ptc_ctx* ctx, const unsigned long long* out_off, const uint32_t* cnt, uint64_t nrows,
                          uint32_t nwin, unsigned long long* off);
int fill_u64(ptc_ctx* ctx, unsigned long long* p, uint64_t n, unsigned long long v);
int fill_u32(ptc_ctx* ctx, uint32_t* p, uint64_t n, uint32_t v);

}  // namespace ptc

#endif
