/*
 * oracle.c -- CPU restatement of the reference's hash-based sparse tensor
 * contraction (ssomesh/parTC), used ONLY as a checker and as the CPU
 * baseline.  This is test infrastructure: the product engine never links,
 * loads or calls it.
 *
 * What is restated (each function cites the reference lines it follows):
 *   - fastmod (extern/fastmod-master/include/fastmod.h:45-84)
 *   - the FKS two-level table with per-bucket cuckoo hashing
 *     (fksCuckoo.hpp:21-94, 152-1060), its probe find_element_new and
 *     buildHashTable
 *   - reorder_tensor and tc_fksCuckoo (fksCuckooTC.hpp:31-371), including
 *     the single-nonzero-row branch and the append by fetch-add
 *   - the per-row sparse accumulator Hash_Table_SPA (fksCuckooSPA.hpp:
 *     159-1069): <= 4 keys in a linear list, then cuckoo with 16 slots
 *   - matricize_spGEMM's relabel -> CSC -> column Gustavson product
 *     (matricizedspGEMM.hpp:172-603, extern/cxsparse/cs_multiply.c:6-38,
 *     extern/cxsparse/cs_scatter.c:44-63)
 *
 * Deliberate, documented deviations (none changes the result):
 *   - kDict is drawn from this file's SplitMix64 stream instead of a
 *     default-seeded std::mt19937 + uniform_int_distribution
 *     (fksCuckoo.hpp:122-129), so bucket/slot placement differs.  Output
 *     coordinates and values do not depend on placement.
 *   - per-row accumulator storage is freed after each row (the reference
 *     leaks it, fksCuckooTC.hpp:340); the output is over-allocated to the
 *     total flop count exactly as the reference does (fksCuckooTC.hpp:156).
 *   - errors return codes instead of exit().
 *   - counts are 64-bit where the reference's 32-bit counts would wrap.
 *   - the numeric loop runs on one thread, as the reference really does
 *     (fksCuckooTC.hpp:171-173 leaves a team of one).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"

/* ------------------------------------------------------------------ misc */

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* per thread, so independent contractions may run on several host threads
 * at once (orc_seed seeds the calling thread's stream) */
static _Thread_local uint64_t orc_rng_state = 0x1234ABCDull;
static uint64_t splitmix64(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
/* uniform in [0, p) -- role of randomnumberintExt (fksCuckoo.hpp:122-129) */
static uint32_t rand_below(uint32_t p) {
  return (uint32_t)(((splitmix64(&orc_rng_state) >> 32) * (uint64_t)p) >> 32);
}
void orc_seed(uint64_t s) { orc_rng_state = s; }

/* Planted mistakes, for tests that show the checks catch them (0 = none):
 *   1  drop the first nonzero of every multi-nonzero row (the SPA seed)
 *   2  pair A's contracted modes with B's in reverse order
 *   3  add A.val + B.val instead of multiplying
 *   4  take the output column from B's first key mode instead of its free
 *      modes */
static int orc_plant = 0;
void orc_set_plant(int p) { orc_plant = p; }

/* fastmod.h:76-78: M = ceil(2^64 / d) */
uint64_t orc_fastmod_M(uint32_t d) { return UINT64_C(0xFFFFFFFFFFFFFFFF) / d + 1; }
/* fastmod.h:45-47 (mul128_u32) and :81-84 (fastmod_u32): a mod d */
uint32_t orc_fastmod_u32(uint32_t a, uint64_t M, uint32_t d) {
  uint64_t lowbits = M * a;
  return (uint32_t)(((__uint128_t)lowbits * d) >> 64);
}

/* fksCuckoo.hpp:21-25: one Mersenne fold of x by p = 2^q - 1 */
static inline uint64_t fold_mersenne(uint64_t x, uint64_t p, unsigned q) {
  uint64_t y = (x & p) + (x >> q);
  return y >= p ? y - p : y;
}

/* ---------------------------------------------------------------- keys */

/* A keyed view of item ids: the key of item e is
 * hedges[e*d + dim[j]] for j < nk (fksCuckoo.hpp:33-41; for the SPA the
 * key is the whole SPA row, fksCuckooSPA.hpp:13-46). */
typedef struct {
  const uint32_t* hedges;
  uint32_t d;
  uint32_t nk;
  const uint32_t* dim; /* NULL = identity */
} keyview;

static inline uint32_t key_at(const keyview* kv, uint64_t e, uint32_t j) {
  return kv->hedges[e * kv->d + (kv->dim ? kv->dim[j] : j)];
}

static int key_equal_ids(const keyview* kv, uint64_t a, uint64_t b) {
  for (uint32_t j = 0; j < kv->nk; ++j)
    if (key_at(kv, a, j) != key_at(kv, b, j)) return 0;
  return 1;
}
/* isEqual(arr, ele, A) (fksCuckoo.hpp:48-61) */
static int key_equal_arr(const keyview* kv, const uint32_t* arr, uint64_t e) {
  for (uint32_t j = 0; j < kv->nk; ++j)
    if (arr[j] != key_at(kv, e, j)) return 0;
  return 1;
}

/* ------------------------------------------------------------ the table */

/* Aux_data (fksCuckoo.hpp:158-172): growable id list.  The SPA's
 * Aux_data_SPA (fksCuckooSPA.hpp:159-173) holds exactly one id. */
typedef struct {
  uint32_t* ids;
  uint32_t size, cap;
} item;

static item* item_new(uint32_t id) {
  item* it = (item*)malloc(sizeof(item));
  it->ids = (uint32_t*)malloc(sizeof(uint32_t));
  it->ids[0] = id;
  it->size = 1;
  it->cap = 1;
  return it;
}
/* Aux_data::insert (fksCuckoo.hpp:233-255): append with doubling */
static int item_append(item* it, uint32_t id) {
  if (it->size + 1 > it->cap) {
    uint32_t nc = it->cap * 2;
    uint32_t* t = (uint32_t*)realloc(it->ids, sizeof(uint32_t) * nc);
    if (!t) return ORC_ERR_NOMEM;
    it->ids = t;
    it->cap = nc;
  }
  it->ids[it->size++] = id;
  return ORC_OK;
}
static void item_free(item* it) {
  if (!it) return;
  free(it->ids);
  free(it);
}

/* Bucket (fksCuckoo.hpp:175-195) / BucketSPA (fksCuckooSPA.hpp:175-195) */
typedef struct {
  uint32_t ndist;  /* num_distinct_elements */
  uint32_t nslots;
  uint32_t k[2];
  item** slots;    /* aux_ptrs */
} bucket;

/* Hash_Table (fksCuckoo.hpp:197-224) / Hash_Table_SPA (fksCuckooSPA.hpp:197-225) */
typedef struct {
  int spa;            /* 0: main table, 1: SPA rules */
  uint32_t num_buckets;
  uint32_t p, q, max_keys;
  uint64_t M;
  bucket** buckets;
  uint64_t* kdict;    /* max_keys x nk */
  keyview kv;
  uint64_t size;
} table;

/* get_index (fksCuckoo.hpp:66-94) and get_index_SPA: slot of item e in a
 * bucket with n slots under key row `keyid`.  The reference sums in two
 * interleaved partial sums; addition mod 2^64 is associative, so one sum is
 * the same value.  (1 + x) is 32-bit unsigned arithmetic as in the
 * reference. */
static inline uint64_t hash_sum_ids(const table* t, uint64_t e, uint32_t keyid) {
  uint64_t s = 0;
  for (uint32_t j = 0; j < t->kv.nk; ++j)
    s += (uint64_t)(uint32_t)(1u + key_at(&t->kv, e, j)) * t->kdict[keyid * t->kv.nk + j];
  return s;
}
static inline uint64_t hash_sum_arr(const table* t, const uint32_t* arr, uint32_t keyid) {
  uint64_t s = 0;
  for (uint32_t j = 0; j < t->kv.nk; ++j)
    s += (uint64_t)(uint32_t)(1u + arr[j]) * t->kdict[keyid * t->kv.nk + j];
  return s;
}
static inline uint32_t slot_of(const table* t, uint64_t e, uint32_t keyid, uint32_t n) {
  return (uint32_t)(fold_mersenne(hash_sum_ids(t, e, keyid), t->p, t->q) % n);
}
static inline uint32_t slot_of_arr(const table* t, const uint32_t* arr, uint32_t keyid, uint32_t n) {
  return (uint32_t)(fold_mersenne(hash_sum_arr(t, arr, keyid), t->p, t->q) % n);
}

/* Hash_Table::Hash_Table (fksCuckoo.hpp:604-656); Hash_Table_SPA ctor
 * (fksCuckooSPA.hpp:637-695). */
static int table_init(table* t, int spa, uint32_t num_buckets, keyview kv, uint32_t largest_dim) {
  memset(t, 0, sizeof(*t));
  t->spa = spa;
  t->num_buckets = num_buckets;
  t->kv = kv;
  t->buckets = (bucket**)calloc(num_buckets ? num_buckets : 1, sizeof(bucket*));
  if (!t->buckets) return ORC_ERR_NOMEM;
  t->M = orc_fastmod_M(num_buckets ? num_buckets : 1);
  t->max_keys = 32;
  t->p = 2147483647u;
  t->q = 31;
  uint32_t larger = num_buckets > largest_dim ? num_buckets : largest_dim;
  if (t->p < larger) return ORC_ERR_RANGE; /* reference: exit(12), :642-648 */
  t->kdict = (uint64_t*)malloc(sizeof(uint64_t) * t->max_keys * (kv.nk ? kv.nk : 1));
  if (!t->kdict) return ORC_ERR_NOMEM;
  for (uint32_t i = 0; i < t->max_keys; ++i)
    for (uint32_t j = 0; j < kv.nk; ++j) t->kdict[i * kv.nk + j] = rand_below(t->p);
  return ORC_OK;
}

static void bucket_free(bucket* b, int free_items) {
  if (!b) return;
  if (free_items)
    for (uint32_t i = 0; i < b->nslots; ++i) item_free(b->slots[i]);
  free(b->slots);
  free(b);
}
static void table_free(table* t) {
  if (t->buckets)
    for (uint32_t i = 0; i < t->num_buckets; ++i) bucket_free(t->buckets[i], 1);
  free(t->buckets);
  free(t->kdict);
  t->buckets = NULL;
  t->kdict = NULL;
}

/* get_bucket_id (fksCuckoo.hpp:658-672): fastmod of the folded key-0 sum */
static inline uint32_t bucket_of(const table* t, uint64_t e) {
  return orc_fastmod_u32((uint32_t)fold_mersenne(hash_sum_ids(t, e, 0), t->p, t->q), t->M,
                         t->num_buckets);
}
static inline uint32_t bucket_of_arr(const table* t, const uint32_t* arr) {
  return orc_fastmod_u32((uint32_t)fold_mersenne(hash_sum_arr(t, arr, 0), t->p, t->q), t->M,
                         t->num_buckets);
}

/* manageBucketSize (fksCuckoo.hpp:267-295); the SPA variant jumps 4 -> 16
 * (fksCuckooSPA.hpp:247-289). */
static int bucket_grow(const table* t, bucket* b) {
  if (b->ndist + b->ndist + 2 > b->nslots) {
    uint32_t old = b->nslots;
    uint32_t ns = (t->spa && b->nslots == 4) ? 16 : b->nslots << 1;
    item** s = (item**)realloc(b->slots, sizeof(item*) * ns);
    if (!s) return ORC_ERR_NOMEM;
    for (uint32_t i = old; i < ns; ++i) s[i] = NULL;
    b->slots = s;
    b->nslots = ns;
  }
  return ORC_OK;
}

/* augment (fksCuckoo.hpp:444-499): walk the BFS predecessor chain from the
 * free slot back to the new item, moving every item on it to its other
 * slot. */
static void bucket_augment(const table* t, bucket* b, uint32_t slot, item** prev, uint32_t k1,
                           uint32_t k2, uint8_t* marked) {
  memset(marked, 0, b->nslots);
  uint32_t r = slot;
  while (prev[r] != NULL && !marked[r]) {
    marked[r] = 1;
    item* it = prev[r];
    uint32_t s1 = slot_of(t, it->ids[0], k1, b->nslots);
    uint32_t s2 = slot_of(t, it->ids[0], k2, b->nslots);
    uint32_t nxt = (s1 == r) ? s2 : s1;
    b->slots[r] = it;
    r = nxt;
  }
}

/* The BFS over the cuckoo graph of one bucket that cuckooHash
 * (fksCuckoo.hpp:504-597) and rehash (:355-420) both run: find a free slot
 * reachable from `start` by displacements and augment along the path.
 * Returns 1 on success. */
static int bucket_bfs_insert(const table* t, bucket* b, item* start, uint32_t k1, uint32_t k2,
                             item** prev, item** queue, uint8_t* visited, uint8_t* marked) {
  for (uint32_t i = 0; i < b->nslots; ++i) prev[i] = NULL;
  memset(visited, 0, b->nslots);
  queue[0] = start;
  uint32_t qh = 0, qs = 1;
  while (qs > qh) {
    item* qi = queue[qh++];
    uint32_t e = qi->ids[0];
    uint32_t s1 = slot_of(t, e, k1, b->nslots);
    if (b->slots[s1] == NULL) {
      prev[s1] = qi;
      bucket_augment(t, b, s1, prev, k1, k2, marked);
      return 1;
    }
    uint32_t s2 = slot_of(t, e, k2, b->nslots);
    if (b->slots[s2] == NULL) {
      prev[s2] = qi;
      bucket_augment(t, b, s2, prev, k1, k2, marked);
      return 1;
    }
    if (!visited[s1]) {
      prev[s1] = qi;
      visited[s1] = 1;
      queue[qs++] = b->slots[s1];
    } else if (!visited[s2]) {
      prev[s2] = qi;
      visited[s2] = 1;
      queue[qs++] = b->slots[s2];
    }
  }
  return 0;
}

typedef struct {
  item** prev;
  item** queue;
  uint8_t* visited;
  uint8_t* marked;
  uint32_t cap;
} scratch;

static int scratch_fit(scratch* s, uint32_t n) {
  if (n <= s->cap) return ORC_OK;
  uint32_t c = s->cap ? s->cap : 16;
  while (c < n) c <<= 1;
  item** p = (item**)realloc(s->prev, sizeof(item*) * c);
  if (!p) return ORC_ERR_NOMEM;
  s->prev = p;
  p = (item**)realloc(s->queue, sizeof(item*) * (c + 1));
  if (!p) return ORC_ERR_NOMEM;
  s->queue = p;
  uint8_t* v = (uint8_t*)realloc(s->visited, c);
  if (!v) return ORC_ERR_NOMEM;
  s->visited = v;
  v = (uint8_t*)realloc(s->marked, c);
  if (!v) return ORC_ERR_NOMEM;
  s->marked = v;
  s->cap = c;
  return ORC_OK;
}
static void scratch_free(scratch* s) {
  free(s->prev);
  free(s->queue);
  free(s->visited);
  free(s->marked);
  memset(s, 0, sizeof(*s));
}

/* rehash (fksCuckoo.hpp:297-442; SPA :291-452): collect the bucket's items
 * plus the new one, grow if needed, then try key pairs (k1, k2) after the
 * current one until every item is placed: direct placement first, BFS
 * augmentation for the rest. */
static int bucket_rehash(const table* t, bucket* b, uint32_t new_id, scratch* sc) {
  uint32_t n = b->ndist + 1;
  item** all = (item**)malloc(sizeof(item*) * n);
  item** left = (item**)malloc(sizeof(item*) * n);
  if (!all || !left) {
    free(all);
    free(left);
    return ORC_ERR_NOMEM;
  }
  uint32_t c = 0;
  for (uint32_t i = 0; i < b->nslots; ++i)
    if (b->slots[i]) all[c++] = b->slots[i];
  all[b->ndist] = item_new(new_id);
  int rc = bucket_grow(t, b);
  if (rc) {
    free(all);
    free(left);
    return rc;
  }
  if ((rc = scratch_fit(sc, b->nslots + 1))) {
    free(all);
    free(left);
    return rc;
  }
  for (uint32_t k1 = b->k[0]; k1 < t->max_keys; ++k1) {
    for (uint32_t k2 = (k1 == b->k[0]) ? (b->k[1] + 1) : 0; k2 < t->max_keys; ++k2) {
      if (k1 == k2) continue;
      for (uint32_t i = 0; i < b->nslots; ++i) b->slots[i] = NULL;
      int collide = 0;
      uint32_t nleft = 0;
      for (uint32_t i = 0; i < n; ++i) {
        uint32_t e = all[i]->ids[0];
        uint32_t s1 = slot_of(t, e, k1, b->nslots);
        uint32_t s2 = slot_of(t, e, k2, b->nslots);
        if (s1 == s2) {
          collide = 1;
          break;
        }
        if (!b->slots[s1]) {
          b->slots[s1] = all[i];
          continue;
        }
        if (!b->slots[s2]) {
          b->slots[s2] = all[i];
          continue;
        }
        left[nleft++] = all[i];
      }
      if (collide) continue;
      int ok = 1;
      for (uint32_t i = 0; i < nleft; ++i) {
        if (!bucket_bfs_insert(t, b, left[i], k1, k2, sc->prev, sc->queue, sc->visited,
                               sc->marked)) {
          ok = 0;
          break;
        }
      }
      if (ok) {
        b->k[0] = k1;
        b->k[1] = k2;
        b->ndist += 1;
        free(all);
        free(left);
        return ORC_OK;
      }
      /* the reference moves on to the next k1 after a failed augmentation
       * (fksCuckoo.hpp:418-420 breaks out of the k2 loop only for the
       * item; the next k2 is tried) */
    }
  }
  free(all);
  free(left);
  return ORC_ERR_KEYS; /* "None of the keys in kDict worked", :439-441 */
}

/* cuckooHash (fksCuckoo.hpp:504-597): BFS augmentation under the bucket's
 * current keys; on failure restore and rehash. */
static int bucket_cuckoo(const table* t, bucket* b, uint32_t new_id, scratch* sc) {
  int rc = scratch_fit(sc, b->nslots + 1);
  if (rc) return rc;
  item** saved = (item**)malloc(sizeof(item*) * b->nslots);
  if (!saved) return ORC_ERR_NOMEM;
  memcpy(saved, b->slots, sizeof(item*) * b->nslots);
  item* it = item_new(new_id);
  if (bucket_bfs_insert(t, b, it, b->k[0], b->k[1], sc->prev, sc->queue, sc->visited,
                        sc->marked)) {
    b->ndist += 1;
    free(saved);
    return ORC_OK;
  }
  memcpy(b->slots, saved, sizeof(item*) * b->nslots);
  free(saved);
  item_free(it);
  return bucket_rehash(t, b, new_id, sc);
}

static bucket* bucket_new(uint32_t id) {
  bucket* b = (bucket*)calloc(1, sizeof(bucket));
  b->ndist = 1;
  b->nslots = 1;
  b->slots = (item**)malloc(sizeof(item*));
  b->slots[0] = item_new(id);
  return b;
}

/* perform_second_level_hashing, main table (fksCuckoo.hpp:674-839) */
static int main_insert_at(table* t, uint32_t e, uint32_t bid, scratch* sc) {
  bucket* b = t->buckets[bid];
  if (!b) {
    t->buckets[bid] = bucket_new(e);
    return ORC_OK;
  }
  if (b->ndist == 1) {
    uint32_t e1 = b->slots[0]->ids[0];
    if (key_equal_ids(&t->kv, e, e1)) return item_append(b->slots[0], e);
    /* second distinct key: 8 slots, search a key pair (:700-775) */
    item** s = (item**)calloc(8, sizeof(item*));
    s[0] = b->slots[0];
    free(b->slots);
    b->slots = s;
    b->nslots = 8;
    for (uint32_t k1 = 0; k1 < t->max_keys; ++k1) {
      for (uint32_t k2 = 0; k2 < t->max_keys; ++k2) {
        if (k1 == k2) continue;
        uint32_t a1 = slot_of(t, e1, k1, 8), a2 = slot_of(t, e1, k2, 8);
        if (a1 == a2) continue;
        uint32_t n1 = slot_of(t, e, k1, 8), n2 = slot_of(t, e, k2, 8);
        if (n1 == n2) continue;
        if ((a1 == n1 && a2 == n2) || (a1 == n2 && a2 == n1)) continue;
        if (a1 > 0) {
          b->slots[a1] = b->slots[0];
          b->slots[0] = NULL;
        }
        uint32_t ns = (a1 == n1) ? n2 : n1;
        b->slots[ns] = item_new(e);
        b->k[0] = k1;
        b->k[1] = k2;
        b->ndist += 1;
        return ORC_OK;
      }
    }
    return ORC_ERR_KEYS;
  }
  /* >= 2 distinct keys (:776-837) */
  uint32_t s1 = slot_of(t, e, b->k[0], b->nslots);
  uint32_t s2 = slot_of(t, e, b->k[1], b->nslots);
  if (s1 == s2) return bucket_rehash(t, b, e, sc);
  if (b->slots[s1] && key_equal_ids(&t->kv, e, b->slots[s1]->ids[0]))
    return item_append(b->slots[s1], e);
  if (b->slots[s2] && key_equal_ids(&t->kv, e, b->slots[s2]->ids[0]))
    return item_append(b->slots[s2], e);
  if (!b->slots[s1]) {
    b->slots[s1] = item_new(e);
    b->ndist += 1;
    return ORC_OK;
  }
  if (!b->slots[s2]) {
    b->slots[s2] = item_new(e);
    b->ndist += 1;
    return ORC_OK;
  }
  return bucket_cuckoo(t, b, e, sc);
}

/* perform_second_level_hashing, SPA rules (fksCuckooSPA.hpp:709-822).  The
 * caller has already established that the key is new (it probes first,
 * fksCuckooTC.hpp:316-330); the seeding loop inserts without a probe
 * (:283), which is correct because one B slice has distinct free tuples. */
static int spa_insert_at(table* t, uint32_t e, uint32_t bid, scratch* sc) {
  bucket* b = t->buckets[bid];
  if (!b) {
    t->buckets[bid] = bucket_new(e);
    return ORC_OK;
  }
  if (b->ndist == 1) { /* :724-736 */
    item** s = (item**)realloc(b->slots, 2 * sizeof(item*));
    if (!s) return ORC_ERR_NOMEM;
    b->slots = s;
    b->nslots = 2;
    b->slots[1] = item_new(e);
    b->ndist += 1;
    return ORC_OK;
  }
  if (b->ndist <= 3) { /* :739-755 */
    if (b->ndist == 2) {
      item** s = (item**)realloc(b->slots, 4 * sizeof(item*));
      if (!s) return ORC_ERR_NOMEM;
      b->slots = s;
      b->nslots = 4;
      b->slots[2] = NULL;
      b->slots[3] = NULL;
    }
    b->slots[b->ndist] = item_new(e);
    b->ndist += 1;
    return ORC_OK;
  }
  if (b->ndist == 4) { /* :758-767; k[1] is left unset there -- taken as 0 */
    int rc = bucket_grow(t, b);
    if (rc) return rc;
    b->k[0] = 1;
    b->k[1] = 0;
    return bucket_rehash(t, b, e, sc);
  }
  uint32_t s1 = slot_of(t, e, b->k[0], b->nslots);
  uint32_t s2 = slot_of(t, e, b->k[1], b->nslots);
  if (s1 == s2) return bucket_rehash(t, b, e, sc);
  if (!b->slots[s1]) {
    b->slots[s1] = item_new(e);
    b->ndist += 1;
    return ORC_OK;
  }
  if (!b->slots[s2]) {
    b->slots[s2] = item_new(e);
    b->ndist += 1;
    return ORC_OK;
  }
  return bucket_cuckoo(t, b, e, sc);
}

static int table_insert(table* t, uint32_t e, scratch* sc) {
  uint32_t bid = bucket_of(t, e);
  t->size++;
  return t->spa ? spa_insert_at(t, e, bid, sc) : main_insert_at(t, e, bid, sc);
}

/* find_element_new (fksCuckoo.hpp:867-978; SPA fksCuckooSPA.hpp:851-991):
 * returns the item holding the key, or NULL; *bid_out gets the bucket. */
static item* table_find(const table* t, const uint32_t* key, uint32_t* bid_out) {
  uint32_t bid = bucket_of_arr(t, key);
  if (bid_out) *bid_out = bid;
  bucket* b = t->buckets[bid];
  if (!b) return NULL;
  if (b->ndist == 1) {
    item* it = b->slots[0];
    return key_equal_arr(&t->kv, key, it->ids[0]) ? it : NULL;
  }
  if (t->spa && b->ndist <= 4) {
    for (uint32_t i = 0; i < b->ndist; ++i)
      if (key_equal_arr(&t->kv, key, b->slots[i]->ids[0])) return b->slots[i];
    return NULL;
  }
  uint32_t s1 = slot_of_arr(t, key, b->k[0], b->nslots);
  if (b->slots[s1] && key_equal_arr(&t->kv, key, b->slots[s1]->ids[0])) return b->slots[s1];
  uint32_t s2 = slot_of_arr(t, key, b->k[1], b->nslots);
  if (b->slots[s2] && key_equal_arr(&t->kv, key, b->slots[s2]->ids[0])) return b->slots[s2];
  return NULL;
}

/* buildHashTable (fksCuckoo.hpp:1044-1060): sequential insert of 0..N-1 */
static int build_table(table* t, const orc_tensor* T, scratch* sc) {
  keyview kv = {T->hedges, T->d, T->nk, T->dim_tc};
  uint32_t largest = 0;
  for (uint32_t j = 0; j < T->nk; ++j)
    if (T->dims[T->dim_tc[j]] > largest) largest = T->dims[T->dim_tc[j]];
  if (T->N > 0xFFFFFFFFull) return ORC_ERR_RANGE;
  int rc = table_init(t, 0, (uint32_t)(T->N ? T->N : 1), kv, largest);
  if (rc) return rc;
  for (uint64_t i = 0; i < T->N; ++i)
    if ((rc = table_insert(t, (uint32_t)i, sc))) return rc;
  return ORC_OK;
}

/* ---------------------------------------------------------- contraction */

static inline void plant_key(uint32_t* k, uint32_t c) {
  if (orc_plant != 2) return;
  for (uint32_t i = 0; i < c / 2; ++i) {
    uint32_t t = k[i];
    k[i] = k[c - 1 - i];
    k[c - 1 - i] = t;
  }
}
static inline uint32_t plant_col(uint32_t pos, uint32_t nk) { return orc_plant == 4 ? 0 : pos; }

void orc_result_free(orc_result* r) {
  if (!r) return;
  free(r->hedges);
  free(r->vals);
  r->hedges = NULL;
  r->vals = NULL;
}

/* reorder_tensor (fksCuckooTC.hpp:31-77): walk the table's buckets and
 * slots, renumber ids so each key's slice is contiguous, rewrite the id
 * lists in place and scatter B into the new order. */
static int reorder_tensor(table* HT_B, const orc_tensor* B, uint32_t** hedges_out,
                          double** vals_out) {
  uint32_t* o2n = (uint32_t*)malloc(sizeof(uint32_t) * (B->N ? B->N : 1));
  uint32_t* h = (uint32_t*)malloc(sizeof(uint32_t) * B->N * B->d + 4);
  double* v = (double*)malloc(sizeof(double) * B->N + 8);
  if (!o2n || !h || !v) {
    free(o2n);
    free(h);
    free(v);
    return ORC_ERR_NOMEM;
  }
  uint64_t nid = 0;
  for (uint32_t i = 0; i < HT_B->num_buckets && nid < B->N; ++i) {
    bucket* b = HT_B->buckets[i];
    if (!b) continue;
    for (uint32_t j = 0; j < b->nslots; ++j) {
      item* it = b->slots[j];
      if (!it) continue;
      for (uint32_t x = 0; x < it->size; ++x) {
        o2n[it->ids[x]] = (uint32_t)nid;
        it->ids[x] = (uint32_t)nid;
        nid++;
      }
    }
  }
  for (uint64_t i = 0; i < B->N; ++i) {
    uint64_t n = o2n[i];
    memcpy(h + n * B->d, B->hedges + i * B->d, sizeof(uint32_t) * B->d);
    v[n] = B->vals[i];
  }
  free(o2n);
  *hedges_out = h;
  *vals_out = v;
  return ORC_OK;
}

/* tc_fksCuckoo (fksCuckooTC.hpp:80-371).  A is free-first (A.nk = d_A - c),
 * B contracted-first (B.nk = c).  HT_A / HT_B are built here, in the order
 * main.cpp:115-116 uses (B first).  max_rows > 0 stops after that many rows
 * of the numeric loop (a bounded sample for timing). */
int orc_tc_fksCuckoo(const orc_tensor* A, const orc_tensor* B, uint64_t max_rows,
                     orc_result* out) {
  memset(out, 0, sizeof(*out));
  uint32_t c = A->d - A->nk;
  if (B->nk != c || c == 0 || A->d > ORC_MAX_MODES || B->d > ORC_MAX_MODES) return ORC_ERR_ARG;
  scratch sc = {0};
  table HT_B, HT_A;
  memset(&HT_B, 0, sizeof(HT_B));
  memset(&HT_A, 0, sizeof(HT_A));
  uint32_t* Bh = NULL;
  double* Bv = NULL;
  uint32_t* row_nnz_bound = NULL;
  uint32_t* row_first = NULL;
  item** row_items = NULL;
  int rc;

  double t0 = now_s();
  if ((rc = build_table(&HT_B, B, &sc))) goto done;
  double t1 = now_s();
  if ((rc = build_table(&HT_A, A, &sc))) goto done;
  double t2 = now_s();
  out->t_build_b = t1 - t0;
  out->t_build_a = t2 - t1;

  /* --- preprocessing (:91-161) --- */
  if ((rc = reorder_tensor(&HT_B, B, &Bh, &Bv))) goto done;
  HT_B.kv.hedges = Bh; /* the table now describes B_reordered */
  orc_tensor Br = *B;
  Br.hedges = Bh;
  Br.vals = Bv;

  uint64_t nrows = 0;
  for (uint32_t i = 0; i < HT_A.num_buckets; ++i) {
    bucket* b = HT_A.buckets[i];
    if (!b) continue;
    for (uint32_t j = 0; j < b->nslots; ++j)
      if (b->slots[j]) nrows++;
  }
  row_nnz_bound = (uint32_t*)calloc(nrows + 1, sizeof(uint32_t));
  row_first = (uint32_t*)malloc(sizeof(uint32_t) * (nrows + 1));
  row_items = (item**)malloc(sizeof(item*) * (nrows + 1));
  if (!row_nnz_bound || !row_first || !row_items) {
    rc = ORC_ERR_NOMEM;
    goto done;
  }
  uint64_t flops = 0;
  uint32_t ckey[ORC_MAX_MODES];
  uint64_t r = 0;
  for (uint32_t i = 0; i < HT_A.num_buckets && r < nrows; ++i) {
    bucket* b = HT_A.buckets[i];
    if (!b) continue;
    for (uint32_t j = 0; j < b->nslots; ++j) {
      item* it = b->slots[j];
      if (!it) continue;
      row_first[r] = it->ids[0];
      row_items[r] = it;
      uint64_t bound = 0;
      for (uint32_t x = 0; x < it->size; ++x) {
        uint64_t e = it->ids[x];
        for (uint32_t cc = A->nk; cc < A->d; ++cc) ckey[cc - A->nk] = A->hedges[e * A->d + A->dim_tc[cc]];
        plant_key(ckey, c);
        item* sl = table_find(&HT_B, ckey, NULL);
        if (!sl) continue;
        bound += sl->size;
      }
      flops += bound;
      if (bound > 0xFFFFFFFFull) {
        rc = ORC_ERR_RANGE;
        goto done;
      }
      row_nnz_bound[r] = (uint32_t)bound;
      r++;
    }
  }
  out->flops = flops;
  out->rows = nrows;
  const uint32_t dO = A->nk + (B->d - B->nk); /* tensor.hpp:113 */
  out->d = dO;
  if (dO == 0 || flops == 0) { /* tensor.hpp:116-119 */
    rc = ORC_ERR_EMPTY;
    goto done;
  }
  out->hedges = (uint32_t*)malloc(sizeof(uint32_t) * flops * dO);
  out->vals = (double*)malloc(sizeof(double) * flops);
  if (!out->hedges || !out->vals) {
    rc = ORC_ERR_NOMEM;
    goto done;
  }
  double t3 = now_s();
  out->t_preprocess = t3 - t2;

  /* --- numeric row loop (:166-357), one thread --- */
  const uint32_t dS = B->d - B->nk; /* SPA.d, :199 */
  uint64_t nout = 0;
  uint64_t rows_done = 0;
  uint32_t* spa_h = NULL;
  double* spa_v = NULL;
  uint64_t spa_cap = 0;
  uint32_t bkey[ORC_MAX_MODES];
  uint32_t rowkey[ORC_MAX_MODES];
  for (uint64_t i = 0; i < nrows; ++i) {
    if (max_rows && rows_done >= max_rows) break;
    if (row_nnz_bound[i] == 0) continue; /* :183 */
    rows_done++;
    uint64_t cap = row_nnz_bound[i];
    if (cap > spa_cap) {
      free(spa_h);
      free(spa_v);
      spa_h = (uint32_t*)malloc(sizeof(uint32_t) * cap * (dS ? dS : 1));
      spa_v = (double*)malloc(sizeof(double) * cap);
      if (!spa_h || !spa_v) {
        rc = ORC_ERR_NOMEM;
        goto done_row;
      }
      spa_cap = cap;
    }
    uint64_t cnt = 0;
    uint32_t first = row_first[i];
    for (uint32_t cc = 0; cc < A->nk; ++cc) rowkey[cc] = A->hedges[(uint64_t)first * A->d + A->dim_tc[cc]];
    item* row = table_find(&HT_A, rowkey, NULL); /* :220 */
    if (row != row_items[i]) {
      rc = ORC_ERR_INTERNAL;
      goto done_row;
    }
    if (row->size == 1) {
      /* single-nonzero row: copy the scaled slice (:221-252) */
      uint64_t e0 = row->ids[0];
      double av = A->vals[e0];
      for (uint32_t cc = A->nk; cc < A->d; ++cc) ckey[cc - A->nk] = A->hedges[e0 * A->d + A->dim_tc[cc]];
      plant_key(ckey, c);
      item* sl = table_find(&HT_B, ckey, NULL);
      if (sl) {
        for (uint32_t j = 0; j < sl->size; ++j) {
          uint64_t b2 = sl->ids[j];
          double prod = orc_plant == 3 ? av + Br.vals[b2] : av * Br.vals[b2];
          for (uint32_t cc = 0; cc < dS; ++cc) spa_h[cnt * dS + cc] = Br.hedges[b2 * B->d + B->dim_tc[plant_col(cc + B->nk, B->nk)]];
          spa_v[cnt] = prod;
          cnt++;
        }
      }
    } else {
      /* per-row SPA (:255-341): table with as many buckets as the bound */
      table spa;
      keyview skv = {spa_h, dS, dS, NULL};
      if ((rc = table_init(&spa, 1, (uint32_t)cap, skv, 0))) {
        table_free(&spa);
        goto done_row;
      }
      uint64_t e0 = row->ids[0];
      double av = A->vals[e0];
      for (uint32_t cc = A->nk; cc < A->d; ++cc) ckey[cc - A->nk] = A->hedges[e0 * A->d + A->dim_tc[cc]];
      plant_key(ckey, c);
      item* sl = orc_plant == 1 ? NULL : table_find(&HT_B, ckey, NULL);
      if (sl) { /* seed with the first nonzero's slice (:269-290) */
        for (uint32_t j = 0; j < sl->size; ++j) {
          uint64_t b2 = sl->ids[j];
          double prod = orc_plant == 3 ? av + Br.vals[b2] : av * Br.vals[b2];
          for (uint32_t cc = 0; cc < dS; ++cc) spa_h[cnt * dS + cc] = Br.hedges[b2 * B->d + B->dim_tc[plant_col(cc + B->nk, B->nk)]];
          spa_v[cnt] = prod;
          if ((rc = table_insert(&spa, (uint32_t)cnt, &sc))) {
            table_free(&spa);
            goto done_row;
          }
          cnt++;
        }
      }
      for (uint32_t k = 1; k < row->size; ++k) { /* :294-337 */
        uint64_t e1 = row->ids[k];
        double av1 = A->vals[e1];
        for (uint32_t cc = A->nk; cc < A->d; ++cc) ckey[cc - A->nk] = A->hedges[e1 * A->d + A->dim_tc[cc]];
        plant_key(ckey, c);
        item* s2 = table_find(&HT_B, ckey, NULL);
        if (!s2) continue;
        for (uint32_t j = 0; j < s2->size; ++j) {
          uint64_t b2 = s2->ids[j];
          for (uint32_t cc = B->nk; cc < B->d; ++cc) bkey[cc - B->nk] = Br.hedges[b2 * B->d + B->dim_tc[plant_col(cc, B->nk)]];
          double prod = orc_plant == 3 ? av1 + Br.vals[b2] : av1 * Br.vals[b2]; /* product, then add (:313, :318) */
          uint32_t bid;
          item* hit = table_find(&spa, bkey, &bid);
          if (hit) {
            spa_v[hit->ids[0]] += prod;
          } else {
            for (uint32_t cc = 0; cc < dS; ++cc) spa_h[cnt * dS + cc] = bkey[cc];
            spa_v[cnt] = prod;
            spa.size++;
            if ((rc = spa_insert_at(&spa, (uint32_t)cnt, bid, &sc))) {
              table_free(&spa);
              goto done_row;
            }
            cnt++;
          }
        }
      }
      table_free(&spa);
    }
    /* append (:345-355) */
    uint64_t start = nout;
    nout += cnt;
    for (uint64_t x = 0; x < cnt; ++x) {
      uint32_t* o = out->hedges + (start + x) * dO;
      for (uint32_t cc = 0; cc < A->nk; ++cc) o[cc] = rowkey[cc];
      for (uint32_t cc = 0; cc < dS; ++cc) o[A->nk + cc] = spa_h[x * dS + cc];
      out->vals[start + x] = spa_v[x];
    }
  }
  rc = ORC_OK;
done_row:
  free(spa_h);
  free(spa_v);
  out->N = nout; /* :360 */
  out->rows_done = rows_done;
  out->t_contract = now_s() - t3;
done:
  table_free(&HT_B);
  table_free(&HT_A);
  scratch_free(&sc);
  free(Bh);
  free(Bv);
  free(row_nnz_bound);
  free(row_first);
  free(row_items);
  if (rc) {
    uint32_t d = out->d;
    orc_result_free(out);
    out->d = d;
    out->N = 0;
  }
  return rc;
}

/* The main.cpp driver (main.cpp:59-118): A and B carry .meta dim_tc with
 * the c contracted modes first; A_copy rotates A's permutation to
 * free-first (:68-109) and B := A's layout for B (sparse_tensor_generator.hpp
 * copies dim_tc, :27-30). */
int orc_contract_meta(const orc_tensor* A, const orc_tensor* B, uint64_t max_rows,
                      orc_result* out) {
  orc_tensor Ac = *A;
  uint32_t rot[ORC_MAX_MODES];
  uint32_t c = A->nk;
  if (A->d > ORC_MAX_MODES || c > A->d) return ORC_ERR_ARG;
  for (uint32_t i = c; i < A->d; ++i) rot[i - c] = A->dim_tc[i];
  for (uint32_t i = 0; i < c; ++i) rot[A->d - c + i] = A->dim_tc[i];
  Ac.dim_tc = rot;
  Ac.nk = A->d - c;
  return orc_tc_fksCuckoo(&Ac, B, max_rows, out);
}

/* ------------------------------------------------------ matricized path */

/* Tuple -> dense id map (the role of google::sparse_hash_map with
 * hashalaFKS, matricizedspGEMM.hpp:144-225): open addressing over tuple
 * references. */
typedef struct {
  const uint32_t** keys;
  uint32_t* vals;
  uint64_t mask;
  uint32_t w;
  uint32_t count;
} tuplemap;

static int tm_init(tuplemap* m, uint64_t n, uint32_t w) {
  uint64_t cap = 16;
  while (cap < 2 * n + 2) cap <<= 1;
  m->keys = (const uint32_t**)calloc(cap, sizeof(uint32_t*));
  m->vals = (uint32_t*)malloc(sizeof(uint32_t) * cap);
  m->mask = cap - 1;
  m->w = w;
  m->count = 0;
  return (m->keys && m->vals) ? ORC_OK : ORC_ERR_NOMEM;
}
static void tm_free(tuplemap* m) {
  free(m->keys);
  free(m->vals);
}
static uint64_t tm_hash(const uint32_t* t, uint32_t w) {
  uint64_t h = 1469598103934665603ull;
  for (uint32_t j = 0; j < w; ++j) {
    h ^= t[j];
    h *= 1099511628211ull;
    h ^= h >> 29;
  }
  return h;
}
/* returns the id of tuple t, inserting it with the next id when new */
static uint32_t tm_get(tuplemap* m, const uint32_t* t) {
  uint64_t i = tm_hash(t, m->w) & m->mask;
  while (m->keys[i]) {
    if (memcmp(m->keys[i], t, sizeof(uint32_t) * m->w) == 0) return m->vals[i];
    i = (i + 1) & m->mask;
  }
  m->keys[i] = t;
  m->vals[i] = m->count++;
  return m->vals[i];
}
static int tm_find(const tuplemap* m, const uint32_t* t, uint32_t* id) {
  uint64_t i = tm_hash(t, m->w) & m->mask;
  while (m->keys[i]) {
    if (memcmp(m->keys[i], t, sizeof(uint32_t) * m->w) == 0) {
      *id = m->vals[i];
      return 1;
    }
    i = (i + 1) & m->mask;
  }
  return 0;
}

/* CSC matrix as cs_di (extern/cxsparse/cs.h) */
typedef struct {
  uint32_t m, n;
  uint64_t* p;
  uint32_t* i;
  double* x;
} csc;

/* COO triples -> CSC by counting sort on the column, stable in input order
 * (matricizedspGEMM.hpp:344-489 sorts the triples then converts). */
static int csc_from_triples(uint32_t m, uint32_t n, uint64_t nz, const uint32_t* ri,
                            const uint32_t* cj, const double* x, csc* C) {
  C->m = m;
  C->n = n;
  C->p = (uint64_t*)calloc((uint64_t)n + 1, sizeof(uint64_t));
  C->i = (uint32_t*)malloc(sizeof(uint32_t) * (nz ? nz : 1));
  C->x = (double*)malloc(sizeof(double) * (nz ? nz : 1));
  if (!C->p || !C->i || !C->x) return ORC_ERR_NOMEM;
  for (uint64_t k = 0; k < nz; ++k) C->p[cj[k] + 1]++;
  for (uint32_t j = 0; j < n; ++j) C->p[j + 1] += C->p[j];
  uint64_t* w = (uint64_t*)malloc(sizeof(uint64_t) * ((uint64_t)n + 1));
  if (!w) return ORC_ERR_NOMEM;
  memcpy(w, C->p, sizeof(uint64_t) * ((uint64_t)n + 1));
  for (uint64_t k = 0; k < nz; ++k) {
    uint64_t q = w[cj[k]]++;
    C->i[q] = ri[k];
    C->x[q] = x[k];
  }
  free(w);
  return ORC_OK;
}
static void csc_free(csc* C) {
  free(C->p);
  free(C->i);
  free(C->x);
}

/* matricize_spGEMM (matricizedspGEMM.hpp:172-603): both tensors are
 * contracted-first (nk = c).  Contraction tuples of A and B get shared dense
 * ids (:184-225), A free tuples become rows and B free tuples columns
 * (:237-339); C = A_mat * B_mat by cs_multiply's column formulation
 * (cs_multiply.c:6-38: for each column j of B, scatter A(:,k)*B(k,j) for
 * every k in B(:,j), cs_scatter.c:44-63); CSC -> COO (:564-591). */
int orc_matricize_spGEMM(const orc_tensor* A, const orc_tensor* B, orc_result* out) {
  memset(out, 0, sizeof(*out));
  uint32_t c = A->nk;
  if (B->nk != c || A->d > ORC_MAX_MODES || B->d > ORC_MAX_MODES) return ORC_ERR_ARG;
  uint32_t fa = A->d - c, fb = B->d - c;
  out->d = fa + fb;
  int rc = ORC_OK;
  uint64_t NA = A->N, NB = B->N;
  /* gather tuples into flat arrays so maps can reference them */
  uint32_t* kA = (uint32_t*)malloc(sizeof(uint32_t) * (NA * c + 1));
  uint32_t* kB = (uint32_t*)malloc(sizeof(uint32_t) * (NB * c + 1));
  uint32_t* rA = (uint32_t*)malloc(sizeof(uint32_t) * (NA * fa + 1));
  uint32_t* cB = (uint32_t*)malloc(sizeof(uint32_t) * (NB * fb + 1));
  uint32_t *rowA = (uint32_t*)malloc(sizeof(uint32_t) * (NA + 1)), *colA = (uint32_t*)malloc(sizeof(uint32_t) * (NA + 1));
  uint32_t *rowB = (uint32_t*)malloc(sizeof(uint32_t) * (NB + 1)), *colB = (uint32_t*)malloc(sizeof(uint32_t) * (NB + 1));
  tuplemap mk, mr, mc;
  memset(&mk, 0, sizeof(mk));
  memset(&mr, 0, sizeof(mr));
  memset(&mc, 0, sizeof(mc));
  csc Am = {0}, Bm = {0};
  uint32_t** rev_r = NULL;
  uint32_t** rev_c = NULL;
  uint32_t* w = NULL;
  double* xw = NULL;
  if (!kA || !kB || !rA || !cB || !rowA || !colA || !rowB || !colB) {
    rc = ORC_ERR_NOMEM;
    goto done;
  }
  for (uint64_t e = 0; e < NA; ++e) {
    for (uint32_t j = 0; j < c; ++j) kA[e * c + j] = A->hedges[e * A->d + A->dim_tc[j]];
    for (uint32_t j = 0; j < fa; ++j) rA[e * fa + j] = A->hedges[e * A->d + A->dim_tc[c + j]];
  }
  for (uint64_t e = 0; e < NB; ++e) {
    for (uint32_t j = 0; j < c; ++j) kB[e * c + j] = B->hedges[e * B->d + B->dim_tc[j]];
    for (uint32_t j = 0; j < fb; ++j) cB[e * fb + j] = B->hedges[e * B->d + B->dim_tc[c + j]];
  }
  if ((rc = tm_init(&mk, NA + NB, c)) || (rc = tm_init(&mr, NA, fa)) || (rc = tm_init(&mc, NB, fb)))
    goto done;
  for (uint64_t e = 0; e < NA; ++e) colA[e] = tm_get(&mk, kA + e * c);
  for (uint64_t e = 0; e < NB; ++e) rowB[e] = tm_get(&mk, kB + e * c);
  for (uint64_t e = 0; e < NA; ++e) rowA[e] = tm_get(&mr, rA + e * fa);
  for (uint64_t e = 0; e < NB; ++e) colB[e] = tm_get(&mc, cB + e * fb);
  uint32_t K = mk.count, R = mr.count, S = mc.count;
  if ((rc = csc_from_triples(R, K, NA, rowA, colA, A->vals, &Am))) goto done;
  if ((rc = csc_from_triples(K, S, NB, rowB, colB, B->vals, &Bm))) goto done;
  /* reverse maps id -> tuple */
  rev_r = (uint32_t**)malloc(sizeof(uint32_t*) * (R + 1));
  rev_c = (uint32_t**)malloc(sizeof(uint32_t*) * (S + 1));
  if (!rev_r || !rev_c) {
    rc = ORC_ERR_NOMEM;
    goto done;
  }
  for (uint64_t e = 0; e < NA; ++e) rev_r[rowA[e]] = rA + e * fa;
  for (uint64_t e = 0; e < NB; ++e) rev_c[colB[e]] = cB + e * fb;
  /* Gustavson, column by column; C grows as needed (cs_multiply.c:20-35) */
  w = (uint32_t*)malloc(sizeof(uint32_t) * (R + 1));
  xw = (double*)malloc(sizeof(double) * (R + 1));
  if (!w || !xw) {
    rc = ORC_ERR_NOMEM;
    goto done;
  }
  for (uint32_t i = 0; i < R; ++i) w[i] = 0;
  uint64_t cap = 1024, nz = 0;
  uint32_t* Ci = (uint32_t*)malloc(sizeof(uint32_t) * cap);
  double* Cx = (double*)malloc(sizeof(double) * cap);
  uint32_t* Cj = (uint32_t*)malloc(sizeof(uint32_t) * cap);
  uint64_t flops = 0;
  for (uint32_t j = 0; j < S && Ci && Cx && Cj; ++j) {
    if (nz + R > cap) {
      while (nz + R > cap) cap *= 2;
      Ci = (uint32_t*)realloc(Ci, sizeof(uint32_t) * cap);
      Cx = (double*)realloc(Cx, sizeof(double) * cap);
      Cj = (uint32_t*)realloc(Cj, sizeof(uint32_t) * cap);
      if (!Ci || !Cx || !Cj) break;
    }
    uint64_t colstart = nz;
    for (uint64_t pb = Bm.p[j]; pb < Bm.p[j + 1]; ++pb) {
      uint32_t k = Bm.i[pb];
      double beta = Bm.x[pb];
      /* cs_scatter (cs_scatter.c:44-63): x(i) += A(i,k)*beta, marking w */
      for (uint64_t pa = Am.p[k]; pa < Am.p[k + 1]; ++pa) {
        uint32_t i = Am.i[pa];
        flops++;
        if (w[i] < j + 1) {
          w[i] = j + 1;
          Ci[nz] = i;
          Cj[nz] = j;
          nz++;
          xw[i] = beta * Am.x[pa];
        } else {
          xw[i] += beta * Am.x[pa];
        }
      }
    }
    for (uint64_t q = colstart; q < nz; ++q) Cx[q] = xw[Ci[q]];
  }
  if (!Ci || !Cx || !Cj) {
    free(Ci);
    free(Cx);
    free(Cj);
    rc = ORC_ERR_NOMEM;
    goto done;
  }
  out->flops = flops;
  out->rows = R;
  out->N = nz;
  out->hedges = (uint32_t*)malloc(sizeof(uint32_t) * (nz * out->d + 1));
  out->vals = (double*)malloc(sizeof(double) * (nz + 1));
  if (!out->hedges || !out->vals) {
    rc = ORC_ERR_NOMEM;
  } else {
    for (uint64_t q = 0; q < nz; ++q) {
      uint32_t* o = out->hedges + q * out->d;
      memcpy(o, rev_r[Ci[q]], sizeof(uint32_t) * fa);
      memcpy(o + fa, rev_c[Cj[q]], sizeof(uint32_t) * fb);
      out->vals[q] = Cx[q];
    }
  }
  free(Ci);
  free(Cx);
  free(Cj);
  if (nz == 0 && rc == ORC_OK) rc = ORC_ERR_EMPTY;
done:
  free(kA);
  free(kB);
  free(rA);
  free(cB);
  free(rowA);
  free(colA);
  free(rowB);
  free(colB);
  tm_free(&mk);
  tm_free(&mr);
  tm_free(&mc);
  csc_free(&Am);
  csc_free(&Bm);
  free(rev_r);
  free(rev_c);
  free(w);
  free(xw);
  if (rc && rc != ORC_ERR_EMPTY) orc_result_free(out);
  return rc;
}

/* Number of useful multiply-adds F = sum_k |A_k| |B_k| and the number of
 * nonempty output rows, counted by key maps (a closed form the contraction
 * must reproduce). */
int orc_flops(const orc_tensor* A, const orc_tensor* B, uint64_t* flops_out) {
  uint32_t c = A->nk;
  if (B->nk != c) return ORC_ERR_ARG;
  tuplemap m;
  uint32_t* kB = (uint32_t*)malloc(sizeof(uint32_t) * (B->N * c + 1));
  uint32_t* kA = (uint32_t*)malloc(sizeof(uint32_t) * (c + 1));
  uint64_t* cnt = (uint64_t*)calloc(B->N + 1, sizeof(uint64_t));
  if (!kB || !kA || !cnt || tm_init(&m, B->N, c)) {
    free(kB);
    free(kA);
    free(cnt);
    return ORC_ERR_NOMEM;
  }
  for (uint64_t e = 0; e < B->N; ++e) {
    for (uint32_t j = 0; j < c; ++j) kB[e * c + j] = B->hedges[e * B->d + B->dim_tc[j]];
    cnt[tm_get(&m, kB + e * c)]++;
  }
  uint64_t f = 0;
  for (uint64_t e = 0; e < A->N; ++e) {
    for (uint32_t j = 0; j < c; ++j) kA[j] = A->hedges[e * A->d + A->dim_tc[j]];
    uint32_t id;
    if (tm_find(&m, kA, &id)) f += cnt[id];
  }
  *flops_out = f;
  tm_free(&m);
  free(kB);
  free(kA);
  free(cnt);
  return ORC_OK;
}
