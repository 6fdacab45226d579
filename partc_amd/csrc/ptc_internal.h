// Internal declarations of the engine: context, device tensors, key tables,
// the device allocator and kernel launchers shared by the .hip files.
#ifndef PTC_INTERNAL_H
#define PTC_INTERNAL_H

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "partc.h"

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#endif

namespace ptc {

uint64_t splitmix64(uint64_t& s);

struct Xoshiro {
  uint64_t s[4];
  explicit Xoshiro(uint64_t seed);
  uint64_t next();
  double uniform();
  uint32_t below(uint32_t n);
};

}  // namespace ptc

#ifdef __HIPCC__

#define PTC_EMPTY_KEY 0xFFFFFFFFFFFFFFFFull

// Mixed-radix linearisation of a key tuple: key = sum_j x_j * stride[j] with
// stride[j] = prod_{i<j} radix[i].  A tuple with x_j >= radix[j] has no key
// (it cannot match anything the table holds).
struct KeyCodec {
  uint32_t n = 0;                    // number of modes in the key
  uint32_t mode[PTC_MAX_MODES] = {}; // tensor mode of key position j
  uint32_t radix[PTC_MAX_MODES] = {};
  uint64_t stride[PTC_MAX_MODES] = {};
  uint64_t space = 1;                // prod radix (number of possible keys)
};

// Table entry: 16 bytes, eight to a 128-byte L2 line.
struct KtEntry {
  unsigned long long key;  // PTC_EMPTY_KEY when free
  uint32_t off;            // first position of the slice in the grouped order
  uint32_t len;            // slice length
};

// One slice (a distinct key) in compact order.
struct KtSlice {
  unsigned long long key;
  uint32_t off;
  uint32_t len;
};

// First slot of a key's probe sequence in a hashed table of cap slots (a
// power of two >= 64, eight slots to a 128-byte line): the line numbered by
// the top bits of the key's hash.  Keys whose hashes share their top bits
// start in one contiguous region of the table, so a build that partitions
// keys by those bits writes each partition's region on its own.
__device__ __forceinline__ uint64_t kt_home(unsigned long long hash, uint64_t cap) {
  return (hash >> (64 - __builtin_ctzll(cap >> 3))) << 3;
}

struct ptc_block {
  void* ptr = nullptr;
  size_t bytes = 0;
};

#endif  // __HIPCC__

#ifdef __HIPCC__

struct ptc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool owns_stream = true;   // false: the caller's stream (ptc_create_on_stream)
  std::string err;
  hipDeviceProp_t prop{};
  // caching allocator: size class -> free blocks
  std::multimap<size_t, void*> free_blocks;
  std::map<void*, size_t> live;
  size_t cached_bytes = 0;
  size_t live_bytes = 0;     // bytes held by live blocks (cache excluded)
  size_t peak_bytes = 0;     // high-water mark of live_bytes since a call reset it
  hipEvent_t ev[8] = {};
  // phase timing (ptc_set_timing): HIP events between the phases of key
  // table builds and contractions, off by default
  int timing = 0;
  hipEvent_t kev[16] = {};   // kernel-level timing events
  int nkev = 0;
  int kev_pass[8] = {};
  // pinned host staging for small device->host reads
  void* host_pinned = nullptr;
  size_t host_pinned_bytes = 0;
  // Lifetime: every device tensor, key table, batch plan and row-size model
  // made on the context holds a reference.  ptc_destroy releases all device
  // memory, events and the stream and marks the context closed; the host
  // record itself stays until the last object is freed, so a free after
  // ptc_destroy reads a live record, sees `closed` and makes no HIP call.
  uint64_t refs = 0;
  bool closed = false;
  // Device address of the allocation-free entry point's scratch array on
  // this device (set at creation), and the arena that entry point opens over
  // it for one call: while `arena` is set, dalloc carves 256-byte aligned
  // blocks from it with a bump pointer (nothing is freed, nothing is
  // allocated) and fails past arena_cap.
  void* scratch = nullptr;
  unsigned char* arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  bool arena_overflow = false;
};

struct ptc_dtensor {
  ptc_ctx* ctx = nullptr;
  uint32_t d = 0;
  uint32_t nk = 0;
  uint64_t N = 0;
  uint32_t dim_tc[PTC_MAX_MODES] = {};
  uint32_t dims[PTC_MAX_MODES] = {};
  uint32_t mode_max[PTC_MAX_MODES] = {};  // max index per mode (0 if N == 0)
  uint32_t* hedges = nullptr;             // device, N*d
  double* vals = nullptr;                 // device, N (FP64 tensors)
  float* vals32 = nullptr;                // device, N: FP32 values (vals is then NULL)
  float* err32 = nullptr;                 // device, N: FP32 results only, lo = sum - vals32
  bool owns = true;
  // every value finite with magnitude >= 2^-500, so no product of two values
  // is +-0 (lets dense accumulators detect touched columns without flags)
  bool vals_safe = false;
};

struct ptc_hashtable {
  ptc_ctx* ctx = nullptr;
  const ptc_dtensor* t = nullptr;
  KeyCodec codec;
  uint64_t cap = 0;            // table slots (power of two; the key space when direct)
  bool direct = false;         // slot = key (small key spaces)
  KtEntry* entries = nullptr;  // device, cap
  // Distinct keys.  A build does not wait for the device: the count stays in
  // d_total ((keys << 32) | grouped nonzeros) until hashtable_num_keys reads
  // it back once.
  uint64_t num_keys = 0;
  bool keys_known = false;
  unsigned long long* d_total = nullptr;
  uint64_t slices_cap = 0;     // slices allocated: min(N, cap) >= num_keys
  KtSlice* slices = nullptr;   // device, num_keys used (slot order)
  uint32_t* perm = nullptr;    // device, N: original ids grouped by key
  uint32_t* slice_of = nullptr;// device, N: compact slice index of each grouped position
                               // (hashed tables: built on first use, hashtable_slice_of)
  // Streaming layout of the non-key modes, in grouped order: the linearised
  // non-key tuple (its codec below) and the value.  Built for every table so
  // the same table serves as B (columns) or as A (rows).
  KeyCodec rest;
  uint32_t* rest_key32 = nullptr;        // when rest.space <= 2^32
  unsigned long long* rest_key64 = nullptr;
  uint16_t* rest_key16 = nullptr;        // 16-bit copy (rest.space <= 2^16): written by the partitioned builds,
                                         // else made on first use
  double* gval = nullptr;      // device, N: values in grouped order
  float* gval32 = nullptr;     // device, N: values in grouped order, FP32 tensors (gval is then NULL)
  bool shares_vals = false;    // gval/gval32 are the tensor's own values (grouped order = tensor order)
  hipEvent_t ev_start = nullptr, ev_end = nullptr;  // bracket the build on the stream
  double ms_build = -1.0;      // from the events, on first request (hashtable_ms_build; 0 without timing)
};

#endif  // __HIPCC__

#ifdef __HIPCC__
namespace ptc {

int set_err(ptc_ctx* ctx, int code, const char* fmt, ...);
int hashtable_slice_of(ptc_ctx* ctx, const ptc_hashtable* h, const uint32_t** out);
// The non-key keys of a table as 16 bits (rest.space <= 2^16), made on
// first use: what dense row passes stream.
int hashtable_col16(ptc_ctx* ctx, const ptc_hashtable* h, const uint16_t** out);
// Key tables of n tensors built together (one launch per step for all of
// them); the same slices as ptc_hashtable_build gives each.
int hashtable_build_many(ptc_ctx* ctx, int n, const ptc_dtensor* const* ts, ptc_hashtable** out);
int hashtable_num_keys(ptc_ctx* ctx, const ptc_hashtable* h, uint64_t* out);
double hashtable_ms_build(const ptc_hashtable* h);
// Key table of a contraction's output O keyed on output row modes only
// (O->dim_tc[0..nk) a permutation of 0..rowc.n-1), formed from the output's
// row grouping: row i of the contraction (key rows[i].key in rowc) holds
// O's positions [off[i], off[i] + cnt[i]).  The table shares O's values.
int hashtable_from_rows(ptc_ctx* ctx, const ptc_dtensor* O, const KeyCodec& rowc, const KtSlice* rows,
                        uint64_t nrows, const uint32_t* cnt, const unsigned long long* off, ptc_hashtable** out);
int hip_check(ptc_ctx* ctx, hipError_t e, const char* what);
#define PTC_HIP(ctx, call)                                 \
  do {                                                     \
    hipError_t _e = (call);                                \
    if (_e != hipSuccess) return ptc::hip_check(ctx, _e, #call); \
  } while (0)

// Reference of an object on its context (ctx_acquire at creation,
// ctx_release when the object is freed; the release that drops the last
// reference of a closed context deletes it).
void ctx_acquire(ptc_ctx* ctx);
void ctx_release(ptc_ctx* ctx);
// True when device work and HIP calls on the context are still allowed.
inline bool ctx_open(const ptc_ctx* ctx) { return ctx && !ctx->closed; }

void* dalloc(ptc_ctx* ctx, size_t bytes);
void dfree(ptc_ctx* ctx, void* p);
template <class T>
T* dalloc_n(ptc_ctx* ctx, size_t n) {
  return static_cast<T*>(dalloc(ctx, sizeof(T) * (n ? n : 1)));
}

// Scoped device workspace: returns every block to the context on exit.
struct Workspace {
  ptc_ctx* ctx;
  std::vector<void*> blocks;
  explicit Workspace(ptc_ctx* c) : ctx(c) {}
  ~Workspace() {
    for (void* p : blocks) dfree(ctx, p);
  }
  template <class T>
  T* get(size_t n) {
    T* p = dalloc_n<T>(ctx, n);
    if (p && !ctx->arena) blocks.push_back(p);  // arena blocks are not freed
    return p;
  }
};

// Scratch of the allocation-free entry point: its device address on the
// context's device (ptc_create calls this with the device current).
int scratch_setup(ptc_ctx* ctx);

// Copy a small device value to the host through pinned memory, synchronously
// on the context stream.
int read_back(ptc_ctx* ctx, void* host, const void* dev, size_t bytes);

// Exclusive prefix sums on the context stream.  out may alias in.  total
// (device, optional) receives the sum of all elements.
int scan_u32_to_u64(ptc_ctx* ctx, const uint32_t* in, unsigned long long* out, uint64_t n,
                    unsigned long long* total);
int scan_u64(ptc_ctx* ctx, const unsigned long long* in, unsigned long long* out, uint64_t n,
             unsigned long long* total);
// Two arrays of n + 1 words scanned in place, each total written at [n]
// (one launch while n is small).
int scan_u64_two_inplace(ptc_ctx* ctx, unsigned long long* a, unsigned long long* b, uint64_t n);

int fill_zero(ptc_ctx* ctx, void* p, size_t bytes);

// Per scan tile of FEAT_TILE positions: the rows of its first and last
// positions and their partial counts of positions that hit B (h) and
// longest slices (m) within the tile.
constexpr uint32_t FEAT_TILE = 2048;
struct FeatPart {
  uint32_t fr, lr, fh, fm, lh, lm;
};
struct ScanFeatures {
  const uint32_t* row_of_pos;  // row of every element (non-decreasing; every row non-empty)
  uint32_t* row_hits;          // per row: elements with a nonzero value
  uint32_t* row_maxlen;        // per row: the largest element
  FeatPart* parts;             // scratch: one per tile
};

// Per TILE_PART elements (half a scan tile): the count of nonzero elements
// and the largest element.
constexpr uint32_t TILE_PART = 1024;
struct TilePart {
  uint32_t h, m;
};
// Exclusive scan of n u32 elements that also leaves the TilePart of every
// TILE_PART elements (tiles: 2 * ceil(n / (2 * TILE_PART)) entries).
int scan_u32_tiles(ptc_ctx* ctx, const uint32_t* in, unsigned long long* out, uint64_t n,
                   unsigned long long* total, TilePart* tiles);

// Exclusive scan of the A positions' slice lengths that also writes the row
// features (every row that has positions; rows past them are left alone).
int scan_u32_features(ptc_ctx* ctx, const uint32_t* in, unsigned long long* out, uint64_t n,
                      unsigned long long* total, const uint32_t* row_of_pos, uint32_t* row_hits,
                      uint32_t* row_maxlen, FeatPart* parts);

// Device words (4 or 8 bytes, zero-extended to 8) gathered into dst[0..n)
// by one launch on the context stream.
struct WordSources {
  const void* src[8];
  uint32_t bytes[8];
  uint32_t n;
};
int collect_words(ptc_ctx* ctx, unsigned long long* dst, const WordSources& W);

// Key codec of a tensor over the given modes, from the tensor's maxima.
int make_codec(ptc_ctx* ctx, const ptc_dtensor* t, const uint32_t* modes, uint32_t n,
               KeyCodec* out);

}  // namespace ptc
#endif  // __HIPCC__

#endif  // PTC_INTERNAL_H
