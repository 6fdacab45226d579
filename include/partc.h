/*
 * partc.h -- C ABI of the MI355X (gfx950) sparse tensor contraction engine.
 *
 * The reference (ssomesh/parTC) has no FFI: its boundary is the C++ header
 * API that main.cpp calls.  Each entry point below replaces one reference
 * interface; the citation on each names it.  All functions return PTC_OK (0)
 * or a negative ptc_status; nothing calls exit() (the reference exits on
 * every error, e.g. tensor.hpp:116-119, fksCuckoo.hpp:439-441).
 *
 * Conventions
 *  - Coordinates are uint32, 0-based, nnz x d row-major ("hedges", as in the
 *    reference Tensor, tensor.hpp:19-26); values are float64.
 *  - dim_tc is the reference's mode permutation: its first
 *    num_key_dimensions entries are the key modes of a table built on the
 *    tensor (fksCuckoo.hpp:33-41, :664).
 *  - A ptc_ctx owns one HIP device and one HIP stream; call it from one host
 *    thread at a time.
 *  - Counts are 64-bit (the reference's 32-bit counts overflow above 2^32
 *    multiply-adds, fksCuckooTC.hpp:156).
 */
#ifndef PARTC_H
#define PARTC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PTC_MAX_MODES 8

typedef enum {
  PTC_OK = 0,
  PTC_ERR_ARG = -1,         /* invalid argument / shape / permutation       */
  PTC_ERR_HIP = -2,         /* HIP runtime error                            */
  PTC_ERR_NOMEM = -3,       /* device or host allocation failed             */
  PTC_ERR_EMPTY = -4,       /* no contributing pair: the reference prints
                               "tensor contraction not possible!!" and exits
                               (tensor.hpp:116-119)                          */
  PTC_ERR_RANGE = -5,       /* index/count range exceeded                   */
  PTC_ERR_UNSUPPORTED = -6, /* a key space wider than 64 bits               */
  PTC_ERR_IO = -7,          /* file could not be opened/parsed              */
  PTC_ERR_NODEV = -8,       /* no HIP device                                */
  PTC_ERR_INTERNAL = -9     /* a device-side consistency check failed       */
} ptc_status;

/* ---------------------------------------------------------------- tensors */

/* Host view of a COO tensor, field for field the reference struct Tensor
 * (tensor.hpp:19-26).  N is 64-bit here. */
typedef struct {
  uint32_t d;
  uint32_t num_key_dimensions;
  uint64_t N;
  const uint32_t* hedges;   /* N*d, row-major                */
  const uint32_t* dims;     /* d (may be NULL: max index + 1) */
  const double* vals;       /* N                             */
  const uint32_t* dim_tc;   /* d, a permutation of 0..d-1    */
} ptc_tensor_view;

typedef struct ptc_ctx ptc_ctx;
typedef struct ptc_dtensor ptc_dtensor;     /* device-resident COO tensor   */
typedef struct ptc_hashtable ptc_hashtable; /* device key table (FKS role)  */
typedef struct ptc_mlp ptc_mlp;             /* learned row-size predictor   */

/* Context: one device, one stream.  device = -1 picks the current device.
 *
 * Lifetime rule: free every object made on a context (device tensors, key
 * tables, batch plans, row-size models) before ptc_destroy.  ptc_destroy
 * releases all of the context's device memory, events and its stream at
 * once, whatever is still live; an object freed after it only releases its
 * host record (no HIP call is made), and the context's own host record is
 * kept until the last such object is freed.  Any other use of an object
 * after its context was destroyed is an error.  Call ptc_destroy before the
 * process exits (HIP's own exit handlers may already have run when static
 * destructors or atexit handlers of the caller run). */
int ptc_create(ptc_ctx** ctx, int device);
/* The same with a caller stream: the context orders all its work on
 * `stream` (a hipStream_t the caller created on that device; not NULL, since
 * batch plans capture graphs on it) instead of creating its own, so a
 * caller's kernels before and after a call need no extra synchronisation.
 * The stream stays the caller's: ptc_destroy does not destroy it, and it
 * must outlive the context.  No reference counterpart (the reference has no
 * device work to order). */
int ptc_create_on_stream(ptc_ctx** ctx, int device, void* stream);
void ptc_destroy(ptc_ctx* ctx);
const char* ptc_last_error(const ptc_ctx* ctx);
int ptc_synchronize(ptc_ctx* ctx);
/* Raw hipStream_t of the context, for callers that order their own work. */
void* ptc_stream(ptc_ctx* ctx);
/* Phase timing on (1) or off (0, the default).  On, key-table builds and
 * contractions record HIP events between their phases and fill the ms_*
 * fields of ptc_stats and ms_build of key tables; off, those read 0.  Each
 * event leaves the GPU idle for several microseconds between two kernels
 * (on MI355X a nips-shape contraction with its two key tables runs 0.72 ms
 * untimed against 0.77 ms timed), so timing is for profiling runs.  Batch
 * plans always time their phases (the autotuner needs them).  No reference
 * counterpart (main.cpp times its phases with host clocks). */
int ptc_set_timing(ptc_ctx* ctx, int on);

/* Copy a host tensor to the device (the reference keeps every tensor in host
 * malloc memory; main.cpp:59-116).  Per-mode maxima are computed here.  A
 * coordinate of 2^32-1 is PTC_ERR_RANGE (key radices are max index + 1 in 32
 * bits); the largest index is 2^32-2. */
int ptc_dtensor_upload(ptc_ctx* ctx, const ptc_tensor_view* t, ptc_dtensor** out);
/* The same for a tensor with FP32 values (t->vals is ignored; vals holds N
 * floats).  Key tables built on it stream 4-byte values, and contracting two
 * FP32 tensors needs an FP32 value_type (PTC_VAL_F32_*).  No
 * reference counterpart: the reference's Tensor holds doubles only
 * (tensor.hpp:25). */
int ptc_dtensor_upload_f32(ptc_ctx* ctx, const ptc_tensor_view* t, const float* vals, ptc_dtensor** out);
/* Wrap device arrays that the caller owns (not freed by ptc_dtensor_free).
 * mode_max may be NULL: the maxima are then reduced on the device. */
int ptc_dtensor_wrap(ptc_ctx* ctx, uint32_t d, uint32_t num_key_dimensions, uint64_t N,
                     const uint32_t* d_hedges, const double* d_vals, const uint32_t* dim_tc,
                     const uint32_t* dims, const uint32_t* mode_max, ptc_dtensor** out);
/* Re-label the key modes of a device tensor without copying data (what
 * main.cpp:68-109 does with a full copy to build A_copy). */
int ptc_dtensor_set_keys(ptc_dtensor* t, uint32_t num_key_dimensions, const uint32_t* dim_tc);
int ptc_dtensor_info(const ptc_dtensor* t, uint32_t* d, uint64_t* N, uint32_t* num_key_dimensions);
/* Device pointers of a tensor (hedges; vals for FP64, vals32 for FP32). */
int ptc_dtensor_device_ptrs(const ptc_dtensor* t, const uint32_t** hedges, const double** vals,
                            const float** vals32);
/* Copy a device tensor back.  hedges: N*d uint32, vals: N float64 (FP32
 * values are widened exactly).  dims may be NULL. */
int ptc_dtensor_download(ptc_ctx* ctx, const ptc_dtensor* t, uint32_t* hedges, double* vals,
                         uint32_t* dims);
/* FP32 values of a tensor.  For an FP32 contraction result, lo (may be NULL)
 * receives the compensation: hi = vals is the sum rounded to FP32 and
 * hi + lo carries it to about twice FP32 precision. */
int ptc_dtensor_download_f32(ptc_ctx* ctx, const ptc_dtensor* t, float* vals, float* lo);
/* Overwrite the FP64 values of a device tensor in place (same coordinates):
 * the input of a batch plan's replay (ptc_batch_run), which re-reads them.
 * Key tables built before the call keep the values they grouped (build new
 * ones for ptc_tc_fksCuckoo).  The values are rescanned as at upload, so
 * zero, tiny or huge values make later contractions track touched output
 * entries with flags (an exactly-zero sum stays in the output, as the
 * reference keeps it, fksCuckooTC.hpp:318-330). */
int ptc_dtensor_set_values(ptc_ctx* ctx, ptc_dtensor* t, const double* vals);
void ptc_dtensor_free(ptc_dtensor* t);

/* ------------------------------------------------------------ file readers */

/* A COO tensor in host memory allocated by the library (free with
 * ptc_host_tensor_free).  Coordinates 0-based, nnz x d row-major. */
typedef struct {
  uint32_t d;
  uint64_t N;
  uint32_t* hedges;
  double* vals;
  uint32_t dims[PTC_MAX_MODES];
} ptc_host_tensor;

/* FROSTT .tns: d 1-based coordinates and a value per line, d fixed by the
 * first data line; dims = largest label per mode.  Replaces readTensor
 * (mtxTnsIO.cpp:31-63 over pigo::Tensor, pigo.hpp:4363-4456). */
int ptc_read_tns(const char* path, ptc_host_tensor* out);
/* Matrix Market coordinate file: d = 2, dims = the header's rows and
 * columns, entries must number the header's nnz.  Replaces readMatrix
 * (mtxTnsIO.cpp:5-29 over pigo::COO, pigo.hpp:3149-3203). */
int ptc_read_mtx(const char* path, ptc_host_tensor* out);
void ptc_host_tensor_free(ptc_host_tensor* t);
/* Write N nonzeros as .tns (1-based, values to 17 significant digits). */
int ptc_write_tns(const char* path, uint32_t d, uint64_t N, const uint32_t* hedges, const double* vals);

/* ------------------------------------------------------------ .meta spec */

/* Parse a .meta file: two free-text lines, then c, then d mode ids whose
 * first c are the contracted modes.  Replaces Tensor::populate_meta_info
 * (tensor.hpp:151-173); fails instead of exiting. */
int ptc_meta_parse(const char* path, uint32_t d, uint32_t* c_out, uint32_t* dim_tc_out);

/* ------------------------------------------------------------- key table */

/* Group the nonzeros of t by their key (the first num_key_dimensions modes of
 * dim_tc): a device hash over the keys, built by wavefront-cooperative
 * atomicCAS insertion, then a count -> scan -> scatter that lays out every
 * key's nonzeros contiguously ("slices").  Replaces buildHashTable
 * (fksCuckoo.hpp:1044-1060) together with reorder_tensor (fksCuckooTC.hpp:
 * 31-77), which the reference runs inside tc_fksCuckoo.  The table keeps a
 * reference to t; t must outlive it. */
int ptc_hashtable_build(ptc_ctx* ctx, const ptc_dtensor* t, ptc_hashtable** out);
/* Number of distinct keys and nonzeros grouped. */
int ptc_hashtable_info(const ptc_hashtable* h, uint64_t* num_keys, uint64_t* num_items);
/* Probe n keys (n x nk uint32 host array, key modes in dim_tc order).
 * slice_len[i] = number of nonzeros with that key (0 if absent);
 * slice_ids (optional, capacity cap) receives the concatenated original
 * nonzero ids of every found slice, slice_off[i] (n+1) their offsets.
 * Replaces Hash_Table::find_element_new (fksCuckoo.hpp:867-978). */
int ptc_hashtable_probe(ptc_ctx* ctx, const ptc_hashtable* h, uint64_t n, const uint32_t* keys,
                        uint64_t* slice_len, uint64_t* slice_off, uint32_t* slice_ids,
                        uint64_t cap);
void ptc_hashtable_free(ptc_hashtable* h);
/* The same table with a chosen grouping: PTC_GROUP_HASH (atomicCAS insert
 * into one open-addressing table, ranks by wave-aggregated atomics),
 * PTC_GROUP_SORT (stable LSD radix sort of the keys, slices from runs of
 * equal keys, the probe index over distinct keys only; key spaces within 32
 * bits), PTC_GROUP_PARTITION (nonzeros partitioned into buckets by key hash,
 * each bucket grouped in an LDS table, the probe index over distinct keys
 * only; fewer than 2^32 nonzeros) or PTC_GROUP_AUTO (one counting pass up
 * to 4096 direct keys, sorting up to 16-bit key spaces, partitioning
 * above).  Sorting is the sort-based matricization
 * of the reference (matricizedspGEMM_sorting.hpp:194-754, which exits before
 * finishing) applied to the grouping step. */
enum { PTC_GROUP_AUTO = 0, PTC_GROUP_HASH = 1, PTC_GROUP_SORT = 2, PTC_GROUP_PARTITION = 3 };
int ptc_hashtable_build_ex(ptc_ctx* ctx, const ptc_dtensor* t, int grouping, ptc_hashtable** out);

/* ------------------------------------------------------------ contraction */

typedef enum {
  PTC_METHOD_TWO_PASS = 0, /* exact symbolic count -> scan -> numeric        */
  PTC_METHOD_LEARNED = 1,  /* MLP-sized single pass, exact fallback rows     */
  PTC_METHOD_UPPER_BOUND = 2 /* single pass sized by min(flops, columns)     */
} ptc_method;

typedef enum {
  PTC_VAL_F64 = 0,        /* default and oracle path                          */
  PTC_VAL_F32_TWOSUM = 1, /* FP32 values; products and sums compensated by
                             TwoProduct (fma) + TwoSum, error ~ FP32 rounding
                             of the exact result whatever the row length    */
  PTC_VAL_F32_PLAIN = 2,  /* FP32 values, plain FP32 atomic accumulation
                             (the uncompensated baseline)                    */
  PTC_VAL_F32_F64ACC = 3  /* FP32 values, exact FP32 x FP32 products summed in
                             FP64 accumulators (gfx950's FP64 LDS atomic adds
                             run several times faster than its FP32 ones);
                             the same (hi, lo) FP32 pair out                 */
} ptc_value_type;

typedef enum {
  PTC_ACC_AUTO = 0,
  PTC_ACC_DENSE = 1,      /* direct-mapped (perfect-hash) LDS accumulator     */
  PTC_ACC_HASH = 2,       /* open-addressing LDS hash accumulator            */
  PTC_ACC_SPILL = 3,      /* HBM hash accumulator (heavy rows)               */
  PTC_ACC_WINDOW = 4      /* dense LDS column windows: each window's workgroups
                             stream a row's multiply-adds and keep the columns
                             in their range; rows of at most 1024 multiply-adds
                             in small LDS hash tables.  Needs a column space
                             below 2^16 and at most 4 windows; windows are at
                             most dense_max_cols wide when that is set, else
                             as wide as the LDS holds.  AUTO takes it for
                             column spaces too wide for one dense accumulator. */
} ptc_acc_mode;

typedef struct {
  int method;             /* ptc_method                                      */
  int value_type;         /* ptc_value_type                                  */
  int acc_mode;           /* ptc_acc_mode; AUTO lets the tuner decide        */
  uint32_t dense_max_cols;/* largest column space held densely in LDS (0 = default) */
  uint32_t lds_hash_max;  /* largest row nnz held in an LDS hash (0 = default) */
  uint64_t chunk_flops;   /* target multiply-adds per workgroup (0 = default; at most
                             2^31, else PTC_ERR_RANGE) */
  double predict_scale;   /* learned path: capacity = scale * prediction (0 = default 1.15) */
  const ptc_mlp* mlp;     /* learned path: predictor (NULL = built-in)       */
  int deterministic;      /* reserved, must be 0 (PTC_ERR_UNSUPPORTED otherwise) */
  uint32_t block_threads; /* threads per row-pass workgroup, a multiple of 64 up to
                             1024 (0 = 1024 when the accumulator's LDS allows
                             fewer than two 512-thread workgroups per CU)     */
} ptc_options;

typedef struct {
  uint64_t nnz_a, nnz_b;
  uint64_t rows;          /* nonempty rows of A (distinct free keys)         */
  uint64_t flops;         /* useful multiply-adds F = sum_k |A_k| |B_k|      */
  uint64_t nnz_out;
  uint64_t fallback_rows; /* learned path: rows whose prediction overflowed  */
  uint64_t heavy_rows;    /* rows split over several workgroups              */
  uint64_t spill_rows;    /* rows accumulated in HBM                          */
  /* phase times (ms): filled with phase timing on (ptc_set_timing), else 0;
   * ms_total is the call's wall time either way */
  double ms_build_b, ms_build_a, ms_preprocess, ms_symbolic, ms_numeric, ms_total;
  /* HIP-event durations of the row-pass kernels alone (k_pass launches on
   * the context stream), summed over their launches in this call; with
   * phase timing on, else 0 */
  double ms_kernel_symbolic, ms_kernel_numeric;
  uint64_t kernel_launches_numeric;
  double ms_build_out;    /* ptc_tc_fksCuckoo_keyed: device time of the output's key table */
  /* output entries allocated before the output's size was known: the
   * single-pass methods size the output by the rows' capacities (learned:
   * their sum; upper bound: min(F, rows x columns)); two-pass: nnz_out */
  uint64_t out_capacity;
  /* peak device memory the call held beyond what was live when it started
   * (its scratch and the output) */
  uint64_t device_peak_bytes;
  /* dense column windows the numeric pass ran (PTC_ACC_WINDOW; 0: one dense
   * or hash accumulator per row) and the rows kept in small LDS hash tables
   * beside them */
  uint64_t col_windows, small_rows;
} ptc_stats;

void ptc_options_default(ptc_options* opt);

/* O = A x B over the contracted modes.  Replaces tc_fksCuckoo
 * (fksCuckooTC.hpp:80-371) with the same preconditions: A is free-first
 * (A.num_key_dimensions = d_A - c, A.dim_tc = [free..., contracted...]),
 * B is contracted-first (B.num_key_dimensions = c); HT_A / HT_B were built
 * on A / B; A's contracted modes pair position-wise with B's.  The output
 * has d = (d_A - c) + (d_B - c) modes: A's free modes then B's free modes;
 * dims are zero and dim_tc is the identity (tensor.hpp:106-149).  A
 * contraction with no contributing pair returns PTC_ERR_EMPTY. */
int ptc_tc_fksCuckoo(ptc_ctx* ctx, const ptc_hashtable* HT_A, const ptc_hashtable* HT_B,
                     const ptc_dtensor* A, const ptc_dtensor* B, const ptc_options* opt,
                     ptc_dtensor** O, ptc_stats* stats);

/* One step of a contraction chain: ptc_tc_fksCuckoo, with the output handed
 * over ready for the next contraction.  O's key split is set to
 * (nk_next, dim_tc_next) -- dim_tc_next a permutation of O's modes, the
 * first nk_next the next contraction's key -- and *HT_O is O's key table,
 * the reorder_tensor layout (fksCuckooTC.hpp:31-77) that ptc_hashtable_build
 * would give.  When the key modes are exactly O's row modes (A's free modes,
 * in any order) the table is formed from the output's own row grouping with
 * no build (*from_rows = 1): the slices are the output rows, in output
 * order, and the table shares O's values (free the table before O).  Other
 * key splits build the table (*from_rows = 0).  stats->ms_build_out is the
 * table's device time.  No counterpart in the reference, which rebuilds a
 * table per contraction (main.cpp:115-116). */
int ptc_tc_fksCuckoo_keyed(ptc_ctx* ctx, const ptc_hashtable* HT_A, const ptc_hashtable* HT_B,
                           const ptc_dtensor* A, const ptc_dtensor* B, const ptc_options* opt,
                           uint32_t nk_next, const uint32_t* dim_tc_next, ptc_dtensor** O,
                           ptc_hashtable** HT_O, int* from_rows, ptc_stats* stats);

/* ptc_tc_fksCuckoo that allocates nothing: no counterpart in the reference;
 * for callers that cannot allocate device memory (serving loops and
 * harnesses that forbid hipMalloc).  The exact two-pass method in FP64 only
 * (opt may be NULL; any other method or value type is PTC_ERR_ARG).  The
 * output goes into caller-owned device arrays: out_hedges (out_cap x d_O
 * uint32) and out_vals (out_cap doubles); *nnz_out receives the output's
 * nonzeros, also when out_cap is too small (then PTC_ERR_RANGE and nothing
 * is written: call again with larger arrays).  Scratch comes from one
 * 64 MiB device array per device, held for the life of the process; a
 * contraction that needs more is PTC_ERR_RANGE (use ptc_tc_fksCuckoo).
 * Every call on a device shares that array: calls that overlap on one
 * device (several host threads or contexts with work in flight at once)
 * race.  The key tables are inputs, built beforehand. */
int ptc_tc_fksCuckoo_noalloc(ptc_ctx* ctx, const ptc_hashtable* HT_A, const ptc_hashtable* HT_B,
                             const ptc_dtensor* A, const ptc_dtensor* B, const ptc_options* opt,
                             uint32_t* out_hedges, double* out_vals, uint64_t out_cap, uint64_t* nnz_out,
                             ptc_stats* stats);

/* Convenience: both tensors carry .meta-style dim_tc (contracted modes
 * first, num_key_dimensions = c, as main.cpp:65 leaves A).  Builds both
 * tables and contracts, as main.cpp:68-118 does.  Replaces the driver path
 * and matricize_spGEMM (matricizedspGEMM.hpp:172-603), whose result it
 * equals. */
int ptc_contract(ptc_ctx* ctx, const ptc_dtensor* A, const ptc_dtensor* B,
                 const ptc_options* opt, ptc_dtensor** O, ptc_stats* stats);

/* matricize_spGEMM (matricizedspGEMM.hpp:172) and matricize_sort_spGEMM
 * (matricizedspGEMM_sorting.hpp:194; the reference exits before computing
 * it, :754): both tensors contracted-first; same result as ptc_contract.
 * ptc_matricize_spGEMM is a second method: the column-wise product of
 * cs_multiply (cs_multiply.c:6-38) -- B grouped by its free modes (the
 * output columns), A by its contracted modes, each output column
 * accumulated over the output's rows -- so the output comes column by
 * column; stats->rows counts nonempty output columns.  The sorting variant
 * is the row-wise method with both operands grouped by radix sort
 * (PTC_GROUP_SORT). */
int ptc_matricize_spGEMM(ptc_ctx* ctx, const ptc_dtensor* A, const ptc_dtensor* B,
                         ptc_dtensor** O, ptc_stats* stats);
int ptc_matricize_sort_spGEMM(ptc_ctx* ctx, const ptc_dtensor* A, const ptc_dtensor* B,
                              ptc_dtensor** O, ptc_stats* stats);

/* --------------------------------------------------- batched contractions */

/* Many independent contractions of uneven sizes, each with ptc_contract's
 * meaning (A and B contracted-first, the exact two-pass method, FP64).  No
 * reference counterpart: main.cpp:59-118 runs one contraction per process.
 *
 * ptc_batch_plan builds every job's key tables, runs the preprocessing and
 * symbolic passes of all jobs (one host round trip per step for the whole
 * batch; each pass one persistent launch per work bin) and allocates the
 * outputs.  The numeric phase is captured once as a hipGraph: re-reading the
 * values of A and B from their device arrays, then one persistent numeric
 * launch per bin.  ptc_batch_run launches it; calling it again after the
 * values of A or B changed (same coordinates) recomputes every output.  A
 * plan keeps its tables, work arrays and outputs until ptc_batch_free; the
 * tensors must outlive it. */
typedef struct ptc_batch ptc_batch;
typedef struct {
  uint32_t jobs;
  uint32_t bins;          /* numeric launch bins (kernel variant x LDS class) */
  uint32_t queue_items;   /* work items over all jobs                         */
  uint32_t runs;          /* numeric graph launches so far                    */
  uint64_t flops, nnz_out, graph_nodes;
  double ms_build;        /* all key tables (device time)                     */
  double ms_symbolic;     /* preprocessing + symbolic phase (device timeline) */
  double ms_symbolic_pass;/* the symbolic persistent launches alone           */
  double ms_plan;         /* ptc_batch_plan, wall                             */
  double ms_numeric;      /* last numeric graph launch (device time)          */
  double ms_total;        /* ptc_contract_batched, wall                       */
} ptc_batch_stats;

int ptc_batch_plan(ptc_ctx* ctx, int n, const ptc_dtensor* const* A, const ptc_dtensor* const* B,
                   const ptc_options* opt, ptc_batch** out, ptc_batch_stats* stats);
int ptc_batch_run(ptc_batch* b, ptc_batch_stats* stats);
/* Job j's output, owned by the plan. */
int ptc_batch_output(const ptc_batch* b, int j, const ptc_dtensor** O);
void ptc_batch_free(ptc_batch* b);
/* Autotuner of a batch's row passes: coordinate descent over the dense LDS
 * column limit (LDS vs hash accumulators), the LDS hash size (LDS vs HBM
 * accumulators) and the multiply-adds per work item, each trial timed on the
 * device (symbolic phase + numeric graph) on the batch's own key tables.
 * best receives base with the chosen knobs. */
#define PTC_TUNE_MAX_TRIALS 24
typedef struct {
  uint32_t trials;
  uint32_t dense_max_cols, lds_hash_max;
  uint64_t chunk_flops;
  double default_ms, best_ms;  /* symbolic launches + numeric graph, device  */
  double ms_wall;              /* the whole search                          */
  /* every trial: its knobs and its time */
  uint32_t trial_dense[PTC_TUNE_MAX_TRIALS], trial_hash[PTC_TUNE_MAX_TRIALS];
  uint64_t trial_chunk[PTC_TUNE_MAX_TRIALS];
  double trial_ms[PTC_TUNE_MAX_TRIALS];
} ptc_tune_report;
int ptc_batch_autotune(ptc_ctx* ctx, int n, const ptc_dtensor* const* A, const ptc_dtensor* const* B,
                       const ptc_options* base, ptc_options* best, ptc_tune_report* report);
/* Plan, run once and hand the outputs over (free each with ptc_dtensor_free). */
int ptc_contract_batched(ptc_ctx* ctx, int n, const ptc_dtensor* const* A, const ptc_dtensor* const* B,
                         const ptc_options* opt, ptc_dtensor** O, ptc_batch_stats* stats);

/* ------------------------------------------------ learned row-size model */

/* A 3-layer perceptron (PTC_MLP_IN -> 64 -> 64 -> 1) that predicts each
 * output row's nonzeros from features known after the key probes, so the
 * LEARNED method can size accumulators and output rows without the symbolic
 * pass.  No reference counterpart: the reference over-allocates every row to
 * its flop bound (fksCuckooTC.hpp:156, :200-205).  Training and inference run
 * on the GPU (hand-written kernels). */
#define PTC_MLP_IN 6
#define PTC_MLP_HIDDEN 64
#define PTC_MLP_NPARAM (PTC_MLP_HIDDEN * PTC_MLP_IN + PTC_MLP_HIDDEN + PTC_MLP_HIDDEN * PTC_MLP_HIDDEN + \
                        PTC_MLP_HIDDEN + PTC_MLP_HIDDEN + 1)
int ptc_mlp_create(ptc_ctx* ctx, uint64_t seed, ptc_mlp** out);
void ptc_mlp_free(ptc_mlp* m);
int ptc_mlp_num_params(void);
int ptc_mlp_get_params(const ptc_mlp* m, float* host);
int ptc_mlp_set_params(ptc_mlp* m, const float* host);
/* Loss (log pred - log nnz)^2 averaged over n samples and its gradient
 * (host arrays: X n x PTC_MLP_IN, ub and y n). */
int ptc_mlp_grad(ptc_ctx* ctx, ptc_mlp* m, const float* X, const float* ub, const float* y, uint64_t n,
                 float* grad_out, float* loss_out);
/* Adam over shuffled mini-batches; loss_out (optional) gets each epoch's mean loss. */
int ptc_mlp_train(ptc_ctx* ctx, ptc_mlp* m, const float* X, const float* ub, const float* y, uint64_t n,
                  uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float* loss_out);
int ptc_mlp_predict(ptc_ctx* ctx, const ptc_mlp* m, const float* X, const float* ub, uint64_t n,
                    float* pred_out);
int ptc_mlp_save(const ptc_mlp* m, const char* path);
int ptc_mlp_load(ptc_ctx* ctx, const char* path, ptc_mlp** out);
/* The model the LEARNED method uses when ptc_options.mlp is NULL. */
void ptc_learned_set_default(const ptc_mlp* m);
const ptc_mlp* ptc_learned_default(void);
/* Training data of one contraction: per-row features X (rows x
 * PTC_MLP_IN), upper bounds ub and exact nonzeros from the symbolic pass.
 * *nrows_out is always set; PTC_ERR_RANGE if cap_rows is too small. */
int ptc_row_profile(ptc_ctx* ctx, const ptc_hashtable* HT_A, const ptc_hashtable* HT_B,
                    const ptc_dtensor* A, const ptc_dtensor* B, float* X, float* ub, uint32_t* nnz,
                    uint64_t cap_rows, uint64_t* nrows_out);

/* ------------------------------------------------------ synthetic tensors */

enum { PTC_DIST_UNIFORM = 0, PTC_DIST_ZIPF = 1 };

/* Seeded unique-coordinate COO tensor (host arrays sized nnz*d and nnz). */
int ptc_gen_coo(uint64_t seed, uint32_t d, const uint32_t* dims, uint64_t nnz, int dist,
                double alpha, uint32_t* idx_out, double* val_out);
/* Shape of one job of a batch of uneven contractions. */
int ptc_gen_job_shape(uint64_t seed, uint32_t dim_lo, uint32_t dim_hi, uint64_t nnz_lo,
                      uint64_t nnz_hi, uint32_t* d_out, uint32_t* dims_out, uint64_t* nnz_out);

#ifdef __cplusplus
}
#endif

#endif /* PARTC_H */
