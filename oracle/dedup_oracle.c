/*
 * dedup_oracle.c -- serial CPU oracle (test infrastructure, see dedup_oracle.h).
 *
 * Everything here is single-threaded and written for clarity over speed: it is
 * the checker for the HIP path, never the thing measured or shipped.
 */
#include "dedup_oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* Murmur3 x64-128, include/kokkos_murmur3.hpp:213-292, seed 0 (:431-437).    */
/* ------------------------------------------------------------------------- */

/* The reference declares rotl64 as returning uint32_t
 * (include/kokkos_murmur3.hpp:30-31), so every 64-bit rotate keeps only its
 * low 32 bits, zero-extended back to 64 bits.  `quirk` = 0 gives textbook
 * MurmurHash3_x64_128 and exists only so the tests can check the structure
 * of this function against the published SMHasher verification value. */
static uint64_t rot64(uint64_t x, int r, int quirk) {
  uint64_t v = (x << r) | (x >> (64 - r));
  return quirk ? (uint64_t)(uint32_t)v : v;
}

static uint64_t fmix64(uint64_t k) { /* kokkos_murmur3.hpp:44-53 */
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

static uint64_t load_le64(const uint8_t* p) {
  uint64_t v = 0;
  for (int i = 7; i >= 0; i--) v = (v << 8) | p[i];
  return v;
}

void or_murmur3_x64_128(const void* key, uint64_t len, uint32_t seed, int quirk, uint8_t out[16]) {
  const uint8_t* data = (const uint8_t*)key;
  const uint64_t nblocks = len / 16;
  uint64_t h1 = seed, h2 = seed;
  const uint64_t c1 = 0x87c37b91114253d5ULL;
  const uint64_t c2 = 0x4cf5ad432745937fULL;

  for (uint64_t i = 0; i < nblocks; i++) { /* body, :229-241 */
    uint64_t k1 = load_le64(data + 16 * i);
    uint64_t k2 = load_le64(data + 16 * i + 8);
    k1 *= c1; k1 = rot64(k1, 31, quirk); k1 *= c2; h1 ^= k1;
    h1 = rot64(h1, 27, quirk); h1 += h2; h1 = h1 * 5 + 0x52dce729;
    k2 *= c2; k2 = rot64(k2, 33, quirk); k2 *= c1; h2 ^= k2;
    h2 = rot64(h2, 31, quirk); h2 += h1; h2 = h2 * 5 + 0x38495ab5;
  }

  const uint8_t* tail = data + nblocks * 16; /* tail, :245-271 */
  uint64_t k1 = 0, k2 = 0;
  unsigned rem = (unsigned)(len & 15);
  if (rem > 8) {
    for (unsigned i = rem; i > 8; i--) k2 ^= (uint64_t)tail[i - 1] << (8 * (i - 9));
    k2 *= c2; k2 = rot64(k2, 33, quirk); k2 *= c1; h2 ^= k2;
  }
  if (rem > 0) {
    unsigned top = rem > 8 ? 8 : rem;
    for (unsigned i = top; i > 0; i--) k1 ^= (uint64_t)tail[i - 1] << (8 * (i - 1));
    k1 *= c1; k1 = rot64(k1, 31, quirk); k1 *= c2; h1 ^= k1;
  }

  h1 ^= len; h2 ^= len; /* finalization, :277-289 */
  h1 += h2; h2 += h1;
  h1 = fmix64(h1); h2 = fmix64(h2);
  h1 += h2; h2 += h1;
  for (int i = 0; i < 8; i++) {
    out[i] = (uint8_t)(h1 >> (8 * i));
    out[8 + i] = (uint8_t)(h2 >> (8 * i));
  }
}

/* ------------------------------------------------------------------------- */
/* MD5, RFC 1321; the reference's PolarSSL-derived version                    */
/* (include/kokkos_md5.hpp:48-323) equals it for lengths <= INT_MAX.           */
/* ------------------------------------------------------------------------- */

static const uint32_t MD5_T[64] = {
  0xd76aa478, 0xe8c7b756, 0x242070db, 0xc1bdceee, 0xf57c0faf, 0x4787c62a, 0xa8304613, 0xfd469501,
  0x698098d8, 0x8b44f7af, 0xffff5bb1, 0x895cd7be, 0x6b901122, 0xfd987193, 0xa679438e, 0x49b40821,
  0xf61e2562, 0xc040b340, 0x265e5a51, 0xe9b6c7aa, 0xd62f105d, 0x02441453, 0xd8a1e681, 0xe7d3fbc8,
  0x21e1cde6, 0xc33707d6, 0xf4d50d87, 0x455a14ed, 0xa9e3e905, 0xfcefa3f8, 0x676f02d9, 0x8d2a4c8a,
  0xfffa3942, 0x8771f681, 0x6d9d6122, 0xfde5380c, 0xa4beea44, 0x4bdecfa9, 0xf6bb4b60, 0xbebfbc70,
  0x289b7ec6, 0xeaa127fa, 0xd4ef3085, 0x04881d05, 0xd9d4d039, 0xe6db99e5, 0x1fa27cf8, 0xc4ac5665,
  0xf4292244, 0x432aff97, 0xab9423a7, 0xfc93a039, 0x655b59c3, 0x8f0ccc92, 0xffeff47d, 0x85845dd1,
  0x6fa87e4f, 0xfe2ce6e0, 0xa3014314, 0x4e0811a1, 0xf7537e82, 0xbd3af235, 0x2ad7d2bb, 0xeb86d391};
static const int MD5_S[4][4] = {{7, 12, 17, 22}, {5, 9, 14, 20}, {4, 11, 16, 23}, {6, 10, 15, 21}};

static uint32_t rol32(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }

static void md5_block(uint32_t st[4], const uint8_t blk[64]) {
  uint32_t X[16];
  for (int i = 0; i < 16; i++)
    X[i] = (uint32_t)blk[4 * i] | ((uint32_t)blk[4 * i + 1] << 8) | ((uint32_t)blk[4 * i + 2] << 16) |
           ((uint32_t)blk[4 * i + 3] << 24);
  uint32_t a = st[0], b = st[1], c = st[2], d = st[3];
  for (int i = 0; i < 64; i++) {
    uint32_t f;
    int g, round = i / 16;
    if (round == 0) { f = (b & c) | (~b & d); g = i; }
    else if (round == 1) { f = (d & b) | (~d & c); g = (5 * i + 1) & 15; }
    else if (round == 2) { f = b ^ c ^ d; g = (3 * i + 5) & 15; }
    else { f = c ^ (b | ~d); g = (7 * i) & 15; }
    uint32_t tmp = d;
    d = c;
    c = b;
    b = b + rol32(a + f + MD5_T[i] + X[g], MD5_S[round][i & 3]);
    a = tmp;
  }
  st[0] += a; st[1] += b; st[2] += c; st[3] += d;
}

void or_md5(const void* data, uint64_t len, uint8_t out[16]) {
  const uint8_t* p = (const uint8_t*)data;
  uint32_t st[4] = {0x67452301, 0xefcdab89, 0x98badcfe, 0x10325476};
  uint64_t full = len / 64;
  for (uint64_t i = 0; i < full; i++) md5_block(st, p + 64 * i);
  uint8_t buf[128];
  uint64_t rem = len - 64 * full;
  memset(buf, 0, sizeof buf);
  memcpy(buf, p + 64 * full, (size_t)rem);
  buf[rem] = 0x80;
  uint64_t padded = (rem < 56) ? 64 : 128;
  uint64_t bits = len * 8;
  for (int i = 0; i < 8; i++) buf[padded - 8 + i] = (uint8_t)(bits >> (8 * i));
  md5_block(st, buf);
  if (padded == 128) md5_block(st, buf + 64);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) out[4 * i + j] = (uint8_t)(st[i] >> (8 * j));
}

/* Dispatch, include/hash_functions.hpp:16-20 (compile-time choice there). */
void or_hash(int kind, const void* data, uint64_t len, uint8_t out[16]) {
  if (kind == OR_HASH_MD5) or_md5(data, len, out);
  else or_murmur3_x64_128(data, len, 0, 1, out);
}

/* ------------------------------------------------------------------------- */
/* Digest -> NodeID map with insert-never-overwrites semantics                */
/* (include/map_helpers.hpp:38-58,105-107).  Layout is irrelevant to results. */
/* ------------------------------------------------------------------------- */

typedef struct { uint32_t node, tree; } node_id;
typedef struct {
  uint8_t* keys;   /* cap * 16 */
  node_id* vals;   /* cap */
  uint8_t* used;   /* cap */
  uint64_t cap, size;
} dmap;

static uint64_t key_hash(const uint8_t* k) {
  uint64_t a = load_le64(k), b = load_le64(k + 8);
  return fmix64(a ^ (b * 0x9e3779b97f4a7c15ULL));
}

static int dmap_init(dmap* m, uint64_t cap) {
  uint64_t c = 16;
  while (c < cap) c <<= 1;
  m->keys = (uint8_t*)malloc(c * 16);
  m->vals = (node_id*)malloc(c * sizeof(node_id));
  m->used = (uint8_t*)calloc(c, 1);
  m->cap = c;
  m->size = 0;
  return (m->keys && m->vals && m->used) ? OR_OK : OR_ENOMEM;
}

static void dmap_free(dmap* m) {
  free(m->keys); free(m->vals); free(m->used);
  memset(m, 0, sizeof *m);
}

/* Returns slot index of `k`, or ~0 if absent. */
static uint64_t dmap_find(const dmap* m, const uint8_t* k) {
  uint64_t s = key_hash(k) & (m->cap - 1);
  while (m->used[s]) {
    if (memcmp(m->keys + 16 * s, k, 16) == 0) return s;
    s = (s + 1) & (m->cap - 1);
  }
  return ~(uint64_t)0;
}

static int dmap_grow(dmap* m) {
  dmap n;
  if (dmap_init(&n, m->cap * 2) != OR_OK) { dmap_free(&n); return OR_ENOMEM; }
  for (uint64_t s = 0; s < m->cap; s++) {
    if (!m->used[s]) continue;
    uint64_t t = key_hash(m->keys + 16 * s) & (n.cap - 1);
    while (n.used[t]) t = (t + 1) & (n.cap - 1);
    n.used[t] = 1;
    memcpy(n.keys + 16 * t, m->keys + 16 * s, 16);
    n.vals[t] = m->vals[s];
    n.size++;
  }
  dmap_free(m);
  *m = n;
  return OR_OK;
}

/* Kokkos::UnorderedMap::insert: returns 1 on a new entry, 0 if the key exists;
 * *slot receives the entry's index either way. */
static int dmap_insert(dmap* m, const uint8_t* k, node_id v, uint64_t* slot) {
  if (2 * (m->size + 1) > m->cap) dmap_grow(m);
  uint64_t s = key_hash(k) & (m->cap - 1);
  while (m->used[s]) {
    if (memcmp(m->keys + 16 * s, k, 16) == 0) { *slot = s; return 0; }
    s = (s + 1) & (m->cap - 1);
  }
  m->used[s] = 1;
  memcpy(m->keys + 16 * s, k, 16);
  m->vals[s] = v;
  m->size++;
  *slot = s;
  return 1;
}

/* ------------------------------------------------------------------------- */
/* Heap-ordered tree geometry                                                  */
/* ------------------------------------------------------------------------- */

static unsigned ilog2_u64(uint64_t x) { unsigned r = 0; while (x >>= 1) r++; return r; }

/* In-order leaves of `node`: the deepest-level run then the run one level up
 * (see header comment).  Equals leftmost_leaf..rightmost_leaf of
 * include/kokkos_merkle_tree.hpp:116-142 when C is a power of two. */
uint64_t or_subtree_leaves(uint32_t node, uint32_t C, uint32_t* out) {
  uint64_t N = 2 * (uint64_t)C - 1;
  if (node >= N) return 0;
  if ((uint64_t)node >= (uint64_t)C - 1) {
    if (out) out[0] = node;
    return 1;
  }
  unsigned D = ilog2_u64(N);
  unsigned d = ilog2_u64((uint64_t)node + 1);
  uint64_t cnt = 0;
  uint64_t lo = (((uint64_t)node + 1) << (D - d)) - 1;
  uint64_t hi = (((uint64_t)node + 2) << (D - d)) - 2;
  if (hi > N - 1) hi = N - 1;
  for (uint64_t x = lo; x <= hi && lo <= N - 1; x++) { if (out) out[cnt] = (uint32_t)x; cnt++; }
  if (d + 1 <= D) {
    uint64_t lo1 = (((uint64_t)node + 1) << (D - 1 - d)) - 1;
    uint64_t hi1 = (((uint64_t)node + 2) << (D - 1 - d)) - 2;
    if (lo1 < (uint64_t)C - 1) lo1 = (uint64_t)C - 1;
    for (uint64_t x = lo1; x <= hi1; x++) { if (out) out[cnt] = (uint32_t)x; cnt++; }
  }
  return cnt;
}

/* ------------------------------------------------------------------------- */
/* Deduplicator state                                                          */
/* ------------------------------------------------------------------------- */

struct or_dedup {
  uint32_t chunk_size;
  int hash_kind;
  uint32_t current_id, baseline_id;
  int have_tree;
  uint64_t data_len;
  uint32_t C;
  uint8_t* tree;     /* (2C-1) * 16, the stored tree (interior zero, :27,121) */
  uint8_t* labels;   /* last checkpoint's labels */
  dmap map;
  uint64_t nfirst, nshift;
  uint32_t options;  /* OR_OPT_* */
};

void or_set_options(or_dedup* d, uint32_t options) { d->options = options; }

or_dedup* or_create(uint32_t chunk_size, int hash_kind) {
  if (chunk_size == 0) return NULL;
  or_dedup* d = (or_dedup*)calloc(1, sizeof *d);
  if (!d) return NULL;
  d->chunk_size = chunk_size;
  d->hash_kind = hash_kind;
  return d;
}

void or_destroy(or_dedup* d) {
  if (!d) return;
  free(d->tree);
  free(d->labels);
  dmap_free(&d->map);
  free(d);
}

uint32_t or_current_id(const or_dedup* d) { return d->current_id; }
void or_counts(const or_dedup* d, uint64_t* nf, uint64_t* ns) { *nf = d->nfirst; *ns = d->nshift; }
uint64_t or_map_size(const or_dedup* d) { return d->map.size; }
void or_free(void* p) { free(p); }

int or_last_labels(const or_dedup* d, uint8_t* out, uint64_t n) {
  if (!d->labels || n != 2 * (uint64_t)d->C - 1) return OR_EINVAL;
  memcpy(out, d->labels, n);
  return OR_OK;
}

int or_last_leaf_digests(const or_dedup* d, uint8_t* out, uint64_t n) {
  if (!d->tree || n != 16 * (uint64_t)d->C) return OR_EINVAL;
  memcpy(out, d->tree + 16 * ((uint64_t)d->C - 1), n);
  return OR_OK;
}

typedef struct { uint32_t node, prev_node, prev_tree; } shift_rec;

static int cmp_shift(const void* a, const void* b) {
  const shift_rec* x = (const shift_rec*)a;
  const shift_rec* y = (const shift_rec*)b;
  if (x->prev_tree != y->prev_tree) return x->prev_tree < y->prev_tree ? -1 : 1;
  if (x->node != y->node) return x->node < y->node ? -1 : 1;
  return 0;
}

static void put32(uint8_t* p, uint32_t v) { for (int i = 0; i < 4; i++) p[i] = (uint8_t)(v >> (8 * i)); }
static void put64(uint8_t* p, uint64_t v) { for (int i = 0; i < 8; i++) p[i] = (uint8_t)(v >> (8 * i)); }
static uint32_t get32(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }
static uint64_t get64(const uint8_t* p) { return (uint64_t)get32(p) | ((uint64_t)get32(p + 4) << 32); }

/* Low-root first occurrences: the TreeLowRootDeduplicator labeling
 * (include/reference_impl.hpp:265-472, dedup_low_root_ref), restated as a
 * deterministic rule.  A leaf is new when it changed and its digest has no
 * entry from an earlier checkpoint (the reference's FIRST_DUPL, :307-323); an
 * interior node is new when both children are (:327-344).  Among the new
 * leaves with one digest, the first occurrence is the leaf whose highest new
 * ancestor (the reference's possible_root, :374-399: walk up while the parent
 * is new) has the smallest node id, ties to the lowest leaf (the reference
 * keeps the first of its list with a strictly smaller root, and lists leaves
 * in ascending order); the others are shifted duplicates of it.  Not kept:
 * the reference's pass that erases roots holding duplicate leaves
 * (:347-372) while iterating a std::unordered_map, whose result depends on
 * that map's iteration order.  Runs after the low-offset leaf pass, whose
 * labels it corrects: the map entry of each new digest (this checkpoint's)
 * is moved to the selected leaf. */
static int low_root_select(or_dedup* d, uint8_t* lab, const uint8_t* curr, uint64_t C, uint32_t K) {
  const uint64_t N = 2 * C - 1;
  uint8_t* nw = (uint8_t*)calloc(N, 1);
  uint64_t* slot = (uint64_t*)malloc(N * sizeof(uint64_t));
  uint64_t* best = (uint64_t*)malloc(d->map.cap * sizeof(uint64_t));
  if (!nw || !slot || !best) { free(nw); free(slot); free(best); return OR_ENOMEM; }
  for (uint64_t leaf = C - 1; leaf < N; leaf++) {
    if (lab[leaf] != OR_FIRST_OCUR && lab[leaf] != OR_SHIFT_DUPL) continue;
    const uint64_t s = dmap_find(&d->map, curr + 16 * leaf);
    if (s == ~(uint64_t)0 || d->map.vals[s].tree != K) continue; /* an earlier checkpoint's digest */
    nw[leaf] = 1;
    slot[leaf] = s;
    best[s] = ~(uint64_t)0;
  }
  if (C >= 2)
    for (uint64_t node = C - 2;; node--) {
      nw[node] = nw[2 * node + 1] && nw[2 * node + 2];
      if (node == 0) break;
    }
  for (uint64_t leaf = C - 1; leaf < N; leaf++) {
    if (!nw[leaf]) continue;
    uint64_t t = leaf;
    while (t > 0 && nw[(t - 1) / 2]) t = (t - 1) / 2;
    const uint64_t key = (t << 32) | leaf;
    if (key < best[slot[leaf]]) best[slot[leaf]] = key;
  }
  for (uint64_t leaf = C - 1; leaf < N; leaf++) {
    if (!nw[leaf]) continue;
    const uint64_t s = slot[leaf];
    if ((uint32_t)best[s] == (uint32_t)leaf) {
      lab[leaf] = OR_FIRST_OCUR;
      d->map.vals[s].node = (uint32_t)leaf;
    } else {
      lab[leaf] = OR_SHIFT_DUPL;
    }
  }
  free(nw); free(slot); free(best);
  return OR_OK;
}

int or_checkpoint(or_dedup* d, const uint8_t* data, uint64_t len, int make_baseline,
                  uint8_t** out, uint64_t* out_len) {
  *out = NULL;
  *out_len = 0;
  uint64_t cs = d->chunk_size;
  if (len == 0) return OR_EINVAL;
  uint64_t C64 = (len + cs - 1) / cs;
  if (C64 > 0x7fffffffULL) return OR_EINVAL; /* node ids are uint32_t */
  uint32_t C = (uint32_t)C64;
  uint64_t N = 2 * C64 - 1;
  const uint32_t K = d->current_id;

  /* Allocation policy, src/tree_approach.cpp:1600-1623,1644-1653: a baseline
   * (requested, the first call, or checkpoint id 0) starts a fresh tree
   * (all-zero digests) and a fresh map.  An incremental checkpoint must keep
   * the buffer length of the chain. */
  int baseline = make_baseline || !d->have_tree || K == 0;
  if (!baseline && len != d->data_len) return OR_EINVAL;
  if (baseline) {
    free(d->tree);
    d->tree = (uint8_t*)calloc(N, 16);
    dmap_free(&d->map);
    if (!d->tree || dmap_init(&d->map, 2 * N) != OR_OK) return OR_ENOMEM;
    d->have_tree = 1;
    d->data_len = len;
    d->C = C;
  }
  if (baseline) d->baseline_id = K;

  /* --- dedup_low_offset_ref, include/reference_impl.hpp:15-132 --- */
  uint8_t* prev = d->tree;
  uint8_t* curr = (uint8_t*)calloc(N, 16);
  uint8_t* lab = (uint8_t*)malloc(N);
  uint8_t* is_root = (uint8_t*)calloc(N, 1);
  if (!curr || !lab || !is_root) { free(curr); free(lab); free(is_root); return OR_ENOMEM; }
  memset(lab, OR_DONE, N);

  for (uint64_t leaf = C64 - 1; leaf < N; leaf++) { /* leaves, :43-71 */
    uint64_t off = (leaf - (C64 - 1)) * cs;
    uint64_t nb = (leaf == N - 1) ? len - off : cs;
    uint8_t* dg = curr + 16 * leaf;
    or_hash(d->hash_kind, data + off, nb, dg);
    uint64_t slot;
    node_id nv = {(uint32_t)leaf, K};
    int inserted = dmap_insert(&d->map, dg, nv, &slot);
    if (memcmp(prev + 16 * leaf, dg, 16) == 0) {
      lab[leaf] = OR_FIXED_DUPL;
    } else if (inserted) {
      lab[leaf] = OR_FIRST_OCUR;
    } else {
      node_id* info = &d->map.vals[slot];
      if (info->tree == K && info->node > leaf) { /* unreachable in serial order */
        lab[leaf] = OR_FIRST_OCUR;
        lab[info->node] = OR_SHIFT_DUPL;
        info->node = (uint32_t)leaf;
      } else {
        lab[leaf] = OR_SHIFT_DUPL;
      }
    }
  }
  if (d->options & OR_OPT_LOW_ROOT) {
    int lr = low_root_select(d, lab, curr, C64, K);
    if (lr != OR_OK) { free(curr); free(lab); free(is_root); return lr; }
  }
  if (C64 >= 2) {
    for (uint64_t node = C64 - 2;; node--) { /* first-occurrence forest, :73-82 */
      uint64_t l = 2 * node + 1, r = 2 * node + 2;
      if (lab[l] == OR_FIRST_OCUR && lab[r] == OR_FIRST_OCUR) {
        lab[node] = OR_FIRST_OCUR;
        or_hash(d->hash_kind, curr + 16 * l, 32, curr + 16 * node);
        uint64_t slot;
        node_id nv = {(uint32_t)node, K};
        dmap_insert(&d->map, curr + 16 * node, nv, &slot);
      }
      if (node == 0) break;
    }
    for (uint64_t node = C64 - 2;; node--) { /* duplicate forest, :83-104 */
      uint64_t l = 2 * node + 1, r = 2 * node + 2;
      if (lab[l] != lab[r]) {
        lab[node] = OR_DONE;
        if (lab[l] != OR_FIXED_DUPL && lab[l] != OR_DONE) is_root[l] = 1;
        if (lab[r] != OR_FIXED_DUPL && lab[r] != OR_DONE) is_root[r] = 1;
      } else if (lab[l] == OR_FIXED_DUPL) {
        lab[node] = OR_FIXED_DUPL;
      } else if (lab[l] == OR_SHIFT_DUPL) {
        or_hash(d->hash_kind, curr + 16 * l, 32, curr + 16 * node);
        if (dmap_find(&d->map, curr + 16 * node) != ~(uint64_t)0) {
          lab[node] = OR_SHIFT_DUPL;
        } else {
          lab[node] = OR_DONE;
          is_root[l] = 1;
          is_root[r] = 1;
        }
      }
      if (node == 0) break;
    }
  }
  if (lab[0] == OR_FIRST_OCUR) is_root[0] = 1; /* :105-106 */
  /* Option of this build: a shifted node 0 (the whole buffer equals an
   * earlier checkpoint) is emitted too, so the checkpoint restarts to its own
   * bytes. */
  if ((d->options & OR_OPT_EMIT_ROOT_SHIFT) && lab[0] == OR_SHIFT_DUPL) is_root[0] = 1;

  /* Root lists ascending by node id (std::set order, :108-119). */
  uint64_t nf = 0, ns = 0;
  for (uint64_t v = 0; v < N; v++) {
    if (!is_root[v] || lab[v] == OR_DONE) continue;
    if (lab[v] == OR_FIRST_OCUR) nf++;
    else if (lab[v] == OR_SHIFT_DUPL) ns++;
  }
  uint32_t* flist = (uint32_t*)malloc((nf ? nf : 1) * sizeof(uint32_t));
  shift_rec* slist = (shift_rec*)malloc((ns ? ns : 1) * sizeof(shift_rec));
  uint64_t* prior = (uint64_t*)calloc((uint64_t)K + 1, sizeof(uint64_t));
  if (!flist || !slist || !prior) {
    free(flist); free(slist); free(prior); free(curr); free(lab); free(is_root);
    return OR_ENOMEM;
  }
  uint64_t fi = 0, si = 0, dchunks = 0;
  for (uint64_t v = 0; v < N; v++) {
    if (!is_root[v] || lab[v] == OR_DONE) continue;
    if (lab[v] == OR_FIRST_OCUR) {
      flist[fi++] = (uint32_t)v;
      dchunks += or_subtree_leaves((uint32_t)v, C, NULL);
    } else if (lab[v] == OR_SHIFT_DUPL) {
      /* collect_diff :572-585,690-696: the source is map.find(tree(node)). */
      uint64_t s = dmap_find(&d->map, curr + 16 * v);
      node_id pv = d->map.vals[s];
      slist[si].node = (uint32_t)v;
      slist[si].prev_node = pv.node;
      slist[si].prev_tree = pv.tree;
      prior[pv.tree] += 1;
      si++;
    }
  }
  /* Canonical order of the shifted list: by source checkpoint, then node
   * (restart needs the grouping, src/tree_approach.cpp:880-898). */
  qsort(slist, ns, sizeof(shift_rec), cmp_shift);
  uint64_t np = 0;
  for (uint64_t t = 0; t <= K; t++) np += prior[t] ? 1 : 0;

  /* --- collect_diff byte layout, src/tree_approach.cpp:593,622-696 --- */
  uint64_t meta = 4 * nf + 8 * np + 8 * ns;
  uint64_t blen = 32 + meta + dchunks * cs;
  uint8_t* buf = (uint8_t*)calloc(blen, 1);
  if (!buf) {
    free(flist); free(slist); free(prior); free(curr); free(lab); free(is_root);
    return OR_ENOMEM;
  }
  /* header_t, include/utils.hpp:29-37 */
  put32(buf + 0, d->baseline_id);
  put32(buf + 4, K);
  put64(buf + 8, len);
  put32(buf + 16, d->chunk_size);
  put32(buf + 20, (uint32_t)nf);
  put32(buf + 24, (uint32_t)np);
  put32(buf + 28, (uint32_t)ns);
  uint8_t* p = buf + 32;
  for (uint64_t i = 0; i < nf; i++, p += 4) put32(p, flist[i]);
  for (uint64_t t = 0; t <= K; t++) {
    if (!prior[t]) continue;
    put32(p, (uint32_t)t);
    put32(p + 4, (uint32_t)prior[t]);
    p += 8;
  }
  for (uint64_t i = 0; i < ns; i++, p += 8) {
    put32(p, slist[i].node);
    put32(p + 4, slist[i].prev_node);
  }
  uint32_t* leaves = (uint32_t*)malloc(C64 * sizeof(uint32_t));
  if (!leaves) {
    free(buf); free(flist); free(slist); free(prior); free(curr); free(lab); free(is_root);
    return OR_ENOMEM;
  }
  for (uint64_t i = 0; i < nf; i++) { /* Gather: Copy data, :635-650 */
    uint64_t n = or_subtree_leaves(flist[i], C, leaves);
    for (uint64_t j = 0; j < n; j++, p += cs) {
      uint64_t chunk = leaves[j] - (C64 - 1);
      uint64_t src = chunk * cs;
      uint64_t w = (src + cs > len) ? len - src : cs;
      memcpy(p, data + src, w);
    }
  }
  free(leaves);

  free(d->tree);
  d->tree = curr;        /* :121, deep_copy(tree.tree_d, curr_tree) */
  free(d->labels);
  d->labels = lab;
  d->nfirst = nf;
  d->nshift = ns;
  d->current_id = K + 1; /* :1748 */
  free(flist); free(slist); free(prior); free(is_root);
  *out = buf;
  *out_len = blen;
  return OR_OK;
}

/* ------------------------------------------------------------------------- */
/* Restart, src/tree_approach.cpp:712-1128, serial per-chunk restatement.      */
/* ------------------------------------------------------------------------- */

typedef struct {
  uint32_t ref_id, chkpt_id;
  uint64_t datalen;
  uint32_t chunk_size, nfirst, nprior, nshift;
  const uint8_t* first;  /* u32[nfirst] */
  const uint8_t* table;  /* {u32 ckpt, u32 count}[nprior] */
  const uint8_t* shift;  /* {u32 node, u32 prev}[nshift] */
  const uint8_t* data;
  uint64_t data_bytes;
} parsed_diff;

static int parse_diff(const uint8_t* b, uint64_t n, parsed_diff* h) {
  if (!b || n < 32) return OR_ECORRUPT;
  h->ref_id = get32(b);
  h->chkpt_id = get32(b + 4);
  h->datalen = get64(b + 8);
  h->chunk_size = get32(b + 16);
  h->nfirst = get32(b + 20);
  h->nprior = get32(b + 24);
  h->nshift = get32(b + 28);
  uint64_t meta = 4ULL * h->nfirst + 8ULL * h->nprior + 8ULL * h->nshift;
  if (h->chunk_size == 0 || h->datalen == 0 || 32 + meta > n) return OR_ECORRUPT;
  h->first = b + 32;
  h->table = h->first + 4ULL * h->nfirst;
  h->shift = h->table + 8ULL * h->nprior;
  h->data = h->shift + 8ULL * h->nshift;
  h->data_bytes = n - 32 - meta;
  return OR_OK;
}

int or_restart(const uint8_t* const* chain, const uint64_t* lens, uint32_t nchain,
               uint32_t id, uint8_t* out, uint64_t out_len) {
  if (id >= nchain) return OR_EINVAL;
  parsed_diff h;
  int rc = parse_diff(chain[id], lens[id], &h);
  if (rc) return rc;
  if (h.datalen != out_len) return OR_EINVAL;
  uint64_t cs = h.chunk_size;
  uint64_t C64 = (h.datalen + cs - 1) / cs;
  if (C64 > 0x7fffffffULL) return OR_ECORRUPT;
  uint32_t C = (uint32_t)C64;
  uint64_t N = 2 * C64 - 1;

  node_id* node_list = (node_id*)malloc(C64 * sizeof(node_id));
  int64_t* dist = (int64_t*)malloc(N * sizeof(int64_t));     /* leaf -> data offset */
  node_id* rep = (node_id*)malloc(N * sizeof(node_id));      /* leaf -> source leaf */
  uint32_t* la = (uint32_t*)malloc(C64 * sizeof(uint32_t));
  uint32_t* lb = (uint32_t*)malloc(C64 * sizeof(uint32_t));
  if (!node_list || !dist || !rep || !la || !lb) { rc = OR_ENOMEM; goto done; }
  for (uint64_t c = 0; c < C64; c++) node_list[c].node = node_list[c].tree = 0xffffffffu;

  int main_ckpt = 1;
  int64_t idx = id;
  uint32_t ref_id = h.ref_id;
  while (idx >= (int64_t)ref_id) {
    if (!main_ckpt) {
      rc = parse_diff(chain[idx], lens[idx], &h);
      if (rc) goto done;
      if (h.datalen != out_len || h.chunk_size != cs) { rc = OR_ECORRUPT; goto done; }
      ref_id = h.ref_id;
    }
    const uint32_t cur = h.chkpt_id;
    for (uint64_t v = 0; v < N; v++) { dist[v] = -1; rep[v].node = rep[v].tree = 0xffffffffu; }

    /* First-occurrence regions, :803-865 and :999-1033. */
    uint64_t off = 0;
    for (uint32_t i = 0; i < h.nfirst; i++) {
      uint32_t node = get32(h.first + 4ULL * i);
      if (node >= N) { rc = OR_ECORRUPT; goto done; }
      uint64_t n = or_subtree_leaves(node, C, la);
      if ((off + n) * cs > h.data_bytes) { rc = OR_ECORRUPT; goto done; }
      for (uint64_t j = 0; j < n; j++) dist[la[j]] = (int64_t)((off + j) * cs);
      off += n;
    }
    /* Prior-checkpoint table -> cumulative counts, :868-883. */
    uint64_t* cum = (uint64_t*)calloc((uint64_t)cur + 2, sizeof(uint64_t));
    if (!cum) { rc = OR_ENOMEM; goto done; }
    for (uint32_t i = 0; i < h.nprior; i++) {
      uint32_t ck = get32(h.table + 8ULL * i);
      uint32_t cnt = get32(h.table + 8ULL * i + 4);
      if (ck > cur) { free(cum); rc = OR_ECORRUPT; goto done; }
      cum[ck] = cnt;
    }
    for (uint64_t t = 1; t <= cur; t++) cum[t] += cum[t - 1];
    /* Shifted duplicates: leaf-by-leaf source mapping, :887-925, :1043-1063. */
    for (uint32_t i = 0; i < h.nshift; i++) {
      uint32_t node = get32(h.shift + 8ULL * i);
      uint32_t prev = get32(h.shift + 8ULL * i + 4);
      uint32_t tree = 0;
      for (uint64_t t = (uint64_t)cur + 1; t-- > 0;) if (i < cum[t]) tree = (uint32_t)t;
      if (node >= N || prev >= N) { free(cum); rc = OR_ECORRUPT; goto done; }
      uint64_t n = or_subtree_leaves(node, C, la);
      uint64_t m = or_subtree_leaves(prev, C, lb);
      if (n != m) { free(cum); rc = OR_ECORRUPT; goto done; }
      for (uint64_t j = 0; j < n; j++) { rep[la[j]].node = lb[j]; rep[la[j]].tree = tree; }
    }
    free(cum);

    if (main_ckpt) {
      /* Main checkpoint: write its own data, record everything else. */
      for (uint64_t c = 0; c < C64; c++) {
        uint64_t leaf = C64 - 1 + c;
        uint64_t dst = c * cs;
        uint64_t w = dst + cs > h.datalen ? h.datalen - dst : cs;
        if (dist[leaf] >= 0) {
          memcpy(out + dst, h.data + dist[leaf], w);
          node_list[c].node = (uint32_t)leaf;
          node_list[c].tree = cur;
        } else if (rep[leaf].node != 0xffffffffu) {
          node_list[c] = rep[leaf];
          if (rep[leaf].tree == cur) {
            int64_t so = dist[rep[leaf].node];
            if (so < 0) { rc = OR_ECORRUPT; goto done; }
            memcpy(out + dst, h.data + so, w);
          }
        } else {
          node_list[c].node = (uint32_t)leaf; /* "Fill same entries", :932-937 */
          node_list[c].tree = cur - 1;
        }
      }
      main_ckpt = 0;
    } else {
      /* "Fill data middle chkpts", :1083-1118 */
      for (uint64_t c = 0; c < C64; c++) {
        if (node_list[c].tree != cur) continue;
        uint32_t x = node_list[c].node;
        uint64_t dst = c * cs;
        uint64_t w = dst + cs > h.datalen ? h.datalen - dst : cs;
        if (dist[x] >= 0) {
          memcpy(out + dst, h.data + dist[x], w);
        } else if (rep[x].node != 0xffffffffu) {
          node_id pv = rep[x];
          if (pv.tree == cur) {
            int64_t so = dist[pv.node];
            if (so < 0) { rc = OR_ECORRUPT; goto done; }
            memcpy(out + dst, h.data + so, w);
          } else {
            node_list[c] = pv;
          }
        } else {
          node_list[c].tree = cur - 1;
        }
      }
    }
    idx--;
  }
  rc = OR_OK;
done:
  free(node_list); free(dist); free(rep); free(la); free(lb);
  return rc;
}

/* ------------------------------------------------------------------------- */
/* List approach, src/list_approach.cpp.                                       */
/* ------------------------------------------------------------------------- */

struct or_list {
  uint32_t chunk_size;
  int hash_kind;
  uint32_t current_id, baseline_id;
  int have_list;
  uint64_t data_len;
  uint64_t C;
  uint8_t* list; /* C * 16: HashList, include/kokkos_hash_list.hpp:11-40 */
  dmap map;
  uint64_t nfirst, nshift;
};

or_list* or_list_create(uint32_t chunk_size, int hash_kind) {
  if (chunk_size == 0) return NULL;
  or_list* d = (or_list*)calloc(1, sizeof *d);
  if (!d) return NULL;
  d->chunk_size = chunk_size;
  d->hash_kind = hash_kind;
  return d;
}

void or_list_destroy(or_list* d) {
  if (!d) return;
  free(d->list);
  dmap_free(&d->map);
  free(d);
}

void or_list_counts(const or_list* d, uint64_t* nf, uint64_t* ns) { *nf = d->nfirst; *ns = d->nshift; }
uint32_t or_list_current_id(const or_list* d) { return d->current_id; }

int or_list_checkpoint(or_list* d, const uint8_t* data, uint64_t len, int make_baseline,
                       uint8_t** out, uint64_t* out_len) {
  *out = NULL;
  *out_len = 0;
  const uint64_t cs = d->chunk_size;
  if (len == 0) return OR_EINVAL;
  const uint64_t C = (len + cs - 1) / cs;
  if (C > 0x7fffffffULL) return OR_EINVAL;
  const uint32_t K = d->current_id;
  /* checkpoint(), :860-893: a baseline starts a zero list and an empty map. */
  int baseline = make_baseline || !d->have_list || K == 0;
  if (!baseline && len != d->data_len) return OR_EINVAL;
  if (baseline) {
    free(d->list);
    d->list = (uint8_t*)calloc(C, 16);
    dmap_free(&d->map);
    if (!d->list || dmap_init(&d->map, 2 * C) != OR_OK) return OR_ENOMEM;
    d->have_list = 1;
    d->data_len = len;
    d->C = C;
    d->baseline_id = K;
  }

  /* dedup_data, :23-62, in chunk order. */
  uint32_t* flist = (uint32_t*)malloc((C ? C : 1) * sizeof(uint32_t));
  shift_rec* slist = (shift_rec*)malloc((C ? C : 1) * sizeof(shift_rec));
  uint64_t* prior = (uint64_t*)calloc((uint64_t)K + 1, sizeof(uint64_t));
  if (!flist || !slist || !prior) { free(flist); free(slist); free(prior); return OR_ENOMEM; }
  uint64_t nf = 0, ns = 0;
  for (uint64_t i = 0; i < C; i++) {
    const uint64_t off = i * cs;
    const uint64_t nb = (i == C - 1) ? len - off : cs;
    uint8_t dg[16];
    or_hash(d->hash_kind, data + off, nb, dg);
    if (memcmp(d->list + 16 * i, dg, 16) == 0) continue; /* unchanged chunk */
    uint64_t slot;
    node_id nv = {(uint32_t)i, K};
    if (dmap_insert(&d->map, dg, nv, &slot)) {
      flist[nf++] = (uint32_t)i;
    } else {
      node_id pv = d->map.vals[slot];
      slist[ns].node = (uint32_t)i;
      slist[ns].prev_node = pv.node;
      slist[ns].prev_tree = pv.tree;
      prior[pv.tree] += 1;
      ns++;
    }
    memcpy(d->list + 16 * i, dg, 16);
  }
  /* collect_diff, :64-256: shifted entries sorted by source checkpoint
   * (BinSort, :165-183), chunk order within a checkpoint. */
  qsort(slist, ns, sizeof(shift_rec), cmp_shift);
  uint64_t np = 0;
  for (uint64_t t = 0; t <= K; t++) np += prior[t] ? 1 : 0;
  const uint64_t blen = 32 + 4 * nf + 8 * np + 8 * ns + nf * cs;
  uint8_t* buf = (uint8_t*)calloc(blen, 1);
  if (!buf) { free(flist); free(slist); free(prior); return OR_ENOMEM; }
  put32(buf + 0, d->baseline_id); /* header, :228-234 */
  put32(buf + 4, K);
  put64(buf + 8, len);
  put32(buf + 16, d->chunk_size);
  put32(buf + 20, (uint32_t)nf);
  put32(buf + 24, (uint32_t)np);
  put32(buf + 28, (uint32_t)ns);
  uint8_t* p = buf + 32;
  for (uint64_t i = 0; i < nf; i++, p += 4) put32(p, flist[i]);
  for (uint64_t t = 0; t <= K; t++) {
    if (!prior[t]) continue;
    put32(p, (uint32_t)t);
    put32(p + 4, (uint32_t)prior[t]);
    p += 8;
  }
  for (uint64_t i = 0; i < ns; i++, p += 8) {
    put32(p, slist[i].node);
    put32(p + 4, slist[i].prev_node);
  }
  for (uint64_t i = 0; i < nf; i++, p += cs) { /* "Copy data", :207-222 */
    const uint64_t src = (uint64_t)flist[i] * cs;
    const uint64_t w = (flist[i] == C - 1) ? len - src : cs;
    memcpy(p, data + src, w);
  }
  d->nfirst = nf;
  d->nshift = ns;
  d->current_id = K + 1;
  free(flist); free(slist); free(prior);
  *out = buf;
  *out_len = blen;
  return OR_OK;
}

/* Per checkpoint of the walk: data offset of each first-occurrence chunk and
 * the (prev, source ckpt) of each shifted chunk, indexed by chunk. */
static int list_tables(const parsed_diff* h, uint64_t C, int64_t* dist, node_id* rep) {
  const uint64_t cs = h->chunk_size;
  for (uint64_t c = 0; c < C; c++) { dist[c] = -1; rep[c].node = rep[c].tree = 0xffffffffu; }
  if ((uint64_t)h->nfirst * cs > h->data_bytes) return OR_ECORRUPT;
  for (uint32_t i = 0; i < h->nfirst; i++) {
    const uint32_t node = get32(h->first + 4ULL * i);
    if (node >= C) return OR_ECORRUPT;
    dist[node] = (int64_t)((uint64_t)i * cs);
  }
  /* Source checkpoint of shifted entry i: the first t with i < cum[t]
   * (:334-354). */
  uint64_t* cum = (uint64_t*)calloc((uint64_t)h->chkpt_id + 2, sizeof(uint64_t));
  if (!cum) return OR_ENOMEM;
  for (uint32_t i = 0; i < h->nprior; i++) {
    const uint32_t ck = get32(h->table + 8ULL * i), cnt = get32(h->table + 8ULL * i + 4);
    if (ck > h->chkpt_id) { free(cum); return OR_ECORRUPT; }
    cum[ck] = cnt;
  }
  for (uint64_t t = 1; t <= h->chkpt_id; t++) cum[t] += cum[t - 1];
  for (uint32_t i = 0; i < h->nshift; i++) {
    const uint32_t node = get32(h->shift + 8ULL * i), prev = get32(h->shift + 8ULL * i + 4);
    if (node >= C || prev >= C) { free(cum); return OR_ECORRUPT; }
    uint32_t tree = 0;
    for (uint64_t t = (uint64_t)h->chkpt_id + 1; t-- > 0;) if (i < cum[t]) tree = (uint32_t)t;
    rep[node].node = prev;
    rep[node].tree = tree;
  }
  free(cum);
  return OR_OK;
}

int or_list_restart(const uint8_t* const* chain, const uint64_t* lens, uint32_t nchain,
                    uint32_t id, uint8_t* out, uint64_t out_len) {
  if (id >= nchain) return OR_EINVAL;
  parsed_diff h;
  int rc = parse_diff(chain[id], lens[id], &h);
  if (rc) return rc;
  if (h.datalen != out_len) return OR_EINVAL;
  const uint64_t cs = h.chunk_size;
  const uint64_t C = (h.datalen + cs - 1) / cs;
  node_id* node_list = (node_id*)malloc(C * sizeof(node_id));
  int64_t* dist = (int64_t*)malloc(C * sizeof(int64_t));
  node_id* rep = (node_id*)malloc(C * sizeof(node_id));
  uint8_t* written = (uint8_t*)calloc(C ? C : 1, 1);
  if (!node_list || !dist || !rep || !written) { rc = OR_ENOMEM; goto done; }
  for (uint64_t c = 0; c < C; c++) node_list[c].node = node_list[c].tree = 0xffffffffu;

  int main_ckpt = 1;
  int64_t idx = id;
  uint32_t ref_id = h.ref_id;
  while (idx >= (int64_t)ref_id) {
    if (!main_ckpt) {
      rc = parse_diff(chain[idx], lens[idx], &h);
      if (rc) goto done;
      if (h.datalen != out_len || h.chunk_size != cs) { rc = OR_ECORRUPT; goto done; }
    }
    const uint32_t cur = h.chkpt_id;
    rc = list_tables(&h, C, dist, rep);
    if (rc) goto done;
    for (uint64_t c = 0; c < C; c++) {
      const uint64_t dst = c * cs;
      const uint64_t w = dst + cs > h.datalen ? h.datalen - dst : cs;
      if (main_ckpt) {
        /* Newest checkpoint, :303-363: its own chunks, shifted chunks of its
         * own data, and the rest default to (chunk, cur-1). */
        if (dist[c] >= 0) {
          memcpy(out + dst, h.data + dist[c], w);
          written[c] = 1;
          node_list[c].node = (uint32_t)c;
          node_list[c].tree = cur;
        } else if (rep[c].node != 0xffffffffu) {
          node_list[c] = rep[c];
          if (rep[c].tree == cur) {
            if (dist[rep[c].node] < 0) { rc = OR_ECORRUPT; goto done; }
            memcpy(out + dst, h.data + dist[rep[c].node], w);
            written[c] = 1;
          }
        } else {
          node_list[c].node = (uint32_t)c;
          node_list[c].tree = cur - 1;
        }
      } else {
        /* Older checkpoints, :366-468. */
        if (node_list[c].tree != cur) continue;
        const uint32_t x = node_list[c].node;
        if (dist[x] >= 0) {
          memcpy(out + dst, h.data + dist[x], w);
          written[c] = 1;
        } else if (rep[x].node != 0xffffffffu) {
          const node_id pv = rep[x];
          if (pv.tree == cur) {
            if (dist[pv.node] < 0) { rc = OR_ECORRUPT; goto done; }
            memcpy(out + dst, h.data + dist[pv.node], w);
            written[c] = 1;
          } else {
            node_list[c] = pv;
          }
        } else {
          node_list[c].tree = cur - 1;
        }
      }
    }
    ref_id = h.ref_id;
    main_ckpt = 0;
    idx--;
  }
  /* Every chunk must have been written by some checkpoint of the chain. */
  for (uint64_t c = 0; c < C; c++) {
    if (!written[c]) { rc = OR_ECORRUPT; goto done; }
  }
  rc = OR_OK;
done:
  free(node_list); free(dist); free(rep); free(written);
  return rc;
}
