/*
 * dedup_omp.c -- parallel (OpenMP) restatement of the tree-approach checkpoint.
 *
 * TEST INFRASTRUCTURE / CPU BASELINE ONLY.  Only tests/ and the cpu_baseline
 * leg of bench.py load it; the product never links it.
 *
 * It restates the reference's parallel checkpoint path -- dedup_data and
 * collect_diff (src/tree_approach.cpp:261-710), which the reference runs as
 * Kokkos parallel loops, e.g. on its OpenMP backend -- on host threads, with
 * the results of the serial oracle (include/reference_impl.hpp:15-132 plus the
 * canonical diff order of dedup_oracle.c), byte for byte:
 *   1. leaves hashed in parallel and compared with the previous tree
 *      (:305-351, reference_impl.hpp:43-71);
 *   2. changed leaves inserted into a concurrent digest map: a slot is claimed
 *      with a 64-bit compare-and-swap, and entries of one checkpoint with one
 *      digest settle on the lowest leaf with an atomic minimum -- the leaf the
 *      serial loop visits first; then each changed leaf is FIRST if the entry
 *      holds it, else SHIFT;
 *   3. the first-occurrence forest level by level (:371-404,
 *      reference_impl.hpp:73-82), then the map inserts of its nodes (atomic
 *      maximum: the serial loop visits interior nodes in descending order);
 *   4. the duplicate forest level by level (:406-477, reference_impl.hpp:83-104);
 *   5. roots (FIRST or SHIFT under a DONE parent; node 0 only when FIRST) in
 *      node order, a stable counting sort of the shifted roots by source
 *      checkpoint, and the diff layout of collect_diff (:593,622-706).
 * The serial oracle's chunk hash (or_hash) and subtree leaf order
 * (or_subtree_leaves) are reused.
 */
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "dedup_oracle.h"

#define SLOT_EMPTY (~(uint64_t)0)
#define LAB_PENDING 0x80

typedef struct {
  uint64_t k0, k1, val, pad;
} om_slot;

typedef struct {
  om_slot* s;
  uint64_t cap, count;
} om_map;

struct om_dedup {
  uint32_t chunk_size;
  int hash_kind;
  uint32_t current_id, baseline_id;
  int have_tree;
  uint64_t data_len, C;
  uint64_t* prev;    /* 2 x u64 per node: previous tree (leaf digests) */
  uint64_t* curr;    /* this checkpoint's tree */
  uint8_t* lab;
  uint32_t* slot_of; /* map slot of changed leaves and shifted interior nodes */
  om_map map;
  uint64_t nfirst, nshift;
};

static uint64_t mix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

static uint64_t home_of(const om_map* m, uint64_t k0, uint64_t k1) {
  return mix64(k0 ^ (k1 * 0x9e3779b97f4a7c15ULL)) & (m->cap - 1);
}

static int map_alloc(om_map* m, uint64_t cap) {
  uint64_t c = 1024;
  while (c < cap) c <<= 1;
  om_slot* s = (om_slot*)malloc(c * sizeof(om_slot));
  if (!s) return OR_ENOMEM;
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < c; i++) {
    s[i].k0 = s[i].k1 = 0;
    s[i].val = SLOT_EMPTY;
    s[i].pad = 0;
  }
  m->s = s;
  m->cap = c;
  m->count = 0;
  return OR_OK;
}

/* Rehash into a table of at least `cap` slots.  Every entry's key is in its
 * slot (entries of earlier checkpoints). */
static int map_grow(om_map* m, uint64_t cap) {
  om_map n;
  if (map_alloc(&n, cap) != OR_OK) return OR_ENOMEM;
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < m->cap; i++) {
    const om_slot e = m->s[i];
    if (e.val == SLOT_EMPTY) continue;
    uint64_t t = home_of(&n, e.k0, e.k1);
    for (;;) {
      uint64_t exp = SLOT_EMPTY;
      if (__atomic_compare_exchange_n(&n.s[t].val, &exp, e.val, 0, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE)) {
        n.s[t].k0 = e.k0;
        n.s[t].k1 = e.k1;
        break;
      }
      t = (t + 1) & (n.cap - 1);
    }
  }
  n.count = m->count;
  free(m->s);
  *m = n;
  return OR_OK;
}

/* Key of an entry: entries claimed during checkpoint K may not have their key
 * stored yet, so theirs is read from the current tree (written before any
 * insert of the phase). */
static void entry_key(const om_map* m, uint64_t s, uint64_t v, uint32_t K, const uint64_t* curr, uint64_t* k0,
                      uint64_t* k1) {
  if ((uint32_t)(v >> 32) == K) {
    const uint32_t n = (uint32_t)v;
    *k0 = curr[2 * (uint64_t)n];
    *k1 = curr[2 * (uint64_t)n + 1];
  } else {
    *k0 = m->s[s].k0;
    *k1 = m->s[s].k1;
  }
}

/* Insert (k0,k1) -> mine; returns the slot, *claimed = 1 for a new entry.
 * Same-checkpoint entries settle on the minimum (use_max = 0) or maximum node. */
static uint64_t map_insert(om_map* m, uint64_t k0, uint64_t k1, uint64_t mine, int use_max, uint32_t K,
                           const uint64_t* curr, int* claimed) {
  uint64_t s = home_of(m, k0, k1);
  for (;;) {
    uint64_t v = __atomic_load_n(&m->s[s].val, __ATOMIC_ACQUIRE);
    if (v == SLOT_EMPTY) {
      uint64_t exp = SLOT_EMPTY;
      if (__atomic_compare_exchange_n(&m->s[s].val, &exp, mine, 0, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE)) {
        m->s[s].k0 = k0;
        m->s[s].k1 = k1;
        *claimed = 1;
        return s;
      }
      v = exp;
    }
    uint64_t e0, e1;
    entry_key(m, s, v, K, curr, &e0, &e1);
    if (e0 == k0 && e1 == k1) {
      if ((uint32_t)(v >> 32) == K) {
        uint64_t cur = v;
        while (use_max ? mine > cur : mine < cur) {
          if (__atomic_compare_exchange_n(&m->s[s].val, &cur, mine, 0, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE)) break;
        }
      }
      *claimed = 0;
      return s;
    }
    s = (s + 1) & (m->cap - 1);
  }
}

static uint64_t map_find(const om_map* m, uint64_t k0, uint64_t k1, uint32_t K, const uint64_t* curr) {
  uint64_t s = home_of(m, k0, k1);
  for (;;) {
    const uint64_t v = m->s[s].val;
    if (v == SLOT_EMPTY) return SLOT_EMPTY;
    uint64_t e0, e1;
    entry_key(m, s, v, K, curr, &e0, &e1);
    if (e0 == k0 && e1 == k1) return s;
    s = (s + 1) & (m->cap - 1);
  }
}

/* Leaf `idx` of the subtree of `node` in or_subtree_leaves' order: the run on
 * the deepest level, then the run one level up. */
static uint64_t subtree_leaf(uint64_t node, uint64_t C, uint64_t idx) {
  const uint64_t N = 2 * C - 1;
  if (node >= C - 1) return node;
  unsigned D = 0, d = 0;
  while ((N >> (D + 1)) > 0) D++;
  while (((node + 1) >> (d + 1)) > 0) d++;
  const uint64_t lo = ((node + 1) << (D - d)) - 1;
  uint64_t hi = ((node + 2) << (D - d)) - 2;
  if (hi > N - 1) hi = N - 1;
  const uint64_t n0 = lo <= N - 1 ? hi - lo + 1 : 0;
  if (idx < n0) return lo + idx;
  uint64_t lo1 = ((node + 1) << (D - 1 - d)) - 1;
  if (lo1 < C - 1) lo1 = C - 1;
  return lo1 + (idx - n0);
}

static void put32(uint8_t* p, uint32_t v) { memcpy(p, &v, 4); }
static void put64(uint8_t* p, uint64_t v) { memcpy(p, &v, 8); }

om_dedup* om_create(uint32_t chunk_size, int hash_kind) {
  if (chunk_size == 0) return NULL;
  om_dedup* d = (om_dedup*)calloc(1, sizeof *d);
  if (!d) return NULL;
  d->chunk_size = chunk_size;
  d->hash_kind = hash_kind;
  return d;
}

static void release_state(om_dedup* d) {
  free(d->prev);
  free(d->curr);
  free(d->lab);
  free(d->slot_of);
  free(d->map.s);
  d->prev = d->curr = NULL;
  d->lab = NULL;
  d->slot_of = NULL;
  memset(&d->map, 0, sizeof d->map);
}

void om_destroy(om_dedup* d) {
  if (!d) return;
  release_state(d);
  free(d);
}

void om_counts(const om_dedup* d, uint64_t* nf, uint64_t* ns) {
  *nf = d->nfirst;
  *ns = d->nshift;
}

uint32_t om_current_id(const om_dedup* d) { return d->current_id; }

int om_max_threads(void) { return omp_get_max_threads(); }

/* Interior nodes of level l: [2^l - 1, min(2^(l+1) - 1, C - 1)). */
static void level_range(uint64_t C, unsigned l, uint64_t* lo, uint64_t* hi) {
  *lo = (1ULL << l) - 1;
  uint64_t h = (1ULL << (l + 1)) - 1;
  *hi = h < C - 1 ? h : C - 1;
}

static int is_root(const uint8_t* lab, uint64_t v) {
  const uint8_t L = lab[v];
  if (L != OR_FIRST_OCUR && L != OR_SHIFT_DUPL) return 0;
  if (v == 0) return L == OR_FIRST_OCUR;
  return lab[(v - 1) >> 1] == OR_DONE;
}

int om_checkpoint(om_dedup* d, const uint8_t* data, uint64_t len, int make_baseline, uint8_t** out,
                  uint64_t* out_len) {
  *out = NULL;
  *out_len = 0;
  const uint64_t cs = d->chunk_size;
  if (len == 0) return OR_EINVAL;
  const uint64_t C = (len + cs - 1) / cs;
  if (C > 0x7fffffffULL) return OR_EINVAL;
  const uint64_t N = 2 * C - 1;
  const uint32_t K = d->current_id;
  const int baseline = make_baseline || !d->have_tree || K == 0;
  if (!baseline && len != d->data_len) return OR_EINVAL;
  if (baseline) {
    release_state(d);
    d->prev = (uint64_t*)calloc(2 * N, 8);
    d->curr = (uint64_t*)malloc(2 * N * 8);
    d->lab = (uint8_t*)malloc(N);
    d->slot_of = (uint32_t*)malloc(N * 4);
    if (!d->prev || !d->curr || !d->lab || !d->slot_of || map_alloc(&d->map, 4 * N) != OR_OK) {
      release_state(d);
      d->have_tree = 0;
      return OR_ENOMEM;
    }
    d->have_tree = 1;
    d->data_len = len;
    d->C = C;
    d->baseline_id = K;
  }
  uint64_t* prev = d->prev;
  uint64_t* curr = d->curr;
  uint8_t* lab = d->lab;
  uint32_t* slot_of = d->slot_of;
  const uint64_t L0 = C - 1; /* node of chunk 0 */

  /* 1. leaves */
  uint64_t npend = 0;
#pragma omp parallel for schedule(static) reduction(+ : npend)
  for (uint64_t k = 0; k < C; k++) {
    const uint64_t leaf = L0 + k, off = k * cs;
    uint8_t dg[16];
    or_hash(d->hash_kind, data + off, k == C - 1 ? len - off : cs, dg);
    uint64_t a, b;
    memcpy(&a, dg, 8);
    memcpy(&b, dg + 8, 8);
    curr[2 * leaf] = a;
    curr[2 * leaf + 1] = b;
    const int same = a == prev[2 * leaf] && b == prev[2 * leaf + 1];
    /* A fixed leaf of an incremental checkpoint: its digest is in the map
     * already, so its insert is a no-op.  A baseline inserts every leaf. */
    lab[leaf] = same ? (uint8_t)(OR_FIXED_DUPL | (baseline ? LAB_PENDING : 0)) : (uint8_t)LAB_PENDING;
    npend += (lab[leaf] & LAB_PENDING) ? 1 : 0;
  }
  /* map capacity: at most one new entry per changed leaf and per interior node */
  if (2 * (d->map.count + 2 * npend) > d->map.cap && map_grow(&d->map, 4 * (d->map.count + 2 * npend)) != OR_OK)
    return OR_ENOMEM;

  /* 2. leaf inserts, then labels */
  om_map* m = &d->map;
  uint64_t claims = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : claims)
  for (uint64_t leaf = L0; leaf < N; leaf++) {
    if (!(lab[leaf] & LAB_PENDING)) continue;
    int c;
    slot_of[leaf] = (uint32_t)map_insert(m, curr[2 * leaf], curr[2 * leaf + 1], ((uint64_t)K << 32) | leaf, 0, K,
                                          curr, &c);
    claims += (uint64_t)c;
  }
#pragma omp parallel for schedule(static)
  for (uint64_t leaf = L0; leaf < N; leaf++) {
    const uint8_t L = lab[leaf];
    if (!(L & LAB_PENDING)) continue;
    if (L != LAB_PENDING) {
      lab[leaf] = OR_FIXED_DUPL;
    } else {
      const uint64_t v = m->s[slot_of[leaf]].val;
      lab[leaf] = v == (((uint64_t)K << 32) | leaf) ? OR_FIRST_OCUR : OR_SHIFT_DUPL;
    }
  }

  /* 3. first-occurrence forest, bottom level first; then its inserts */
  unsigned D = 0;
  while ((N >> D) > 1) D++; /* floor(log2 N): level of the deepest leaves */
  for (int l = (int)D - 1; l >= 0 && C >= 2; l--) {
    uint64_t lo, hi;
    level_range(C, (unsigned)l, &lo, &hi);
#pragma omp parallel for schedule(static)
    for (uint64_t v = lo; v < hi; v++) {
      const uint64_t a = 2 * v + 1, b = 2 * v + 2;
      if (lab[a] == OR_FIRST_OCUR && lab[b] == OR_FIRST_OCUR) {
        uint8_t dg[16];
        or_hash(d->hash_kind, &curr[2 * a], 32, dg); /* children are adjacent: 2v+1, 2v+2 */
        memcpy(&curr[2 * v], dg, 16);
        lab[v] = OR_FIRST_OCUR;
      } else {
        lab[v] = OR_DONE;
      }
    }
  }
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : claims)
  for (uint64_t v = 0; v < (C >= 2 ? C - 1 : 0); v++) {
    if (lab[v] != OR_FIRST_OCUR) continue;
    int c;
    map_insert(m, curr[2 * v], curr[2 * v + 1], ((uint64_t)K << 32) | v, 1, K, curr, &c);
    claims += (uint64_t)c;
  }
  m->count += claims;

  /* 4. duplicate forest */
  for (int l = (int)D - 1; l >= 0 && C >= 2; l--) {
    uint64_t lo, hi;
    level_range(C, (unsigned)l, &lo, &hi);
#pragma omp parallel for schedule(static)
    for (uint64_t v = lo; v < hi; v++) {
      const uint64_t a = 2 * v + 1, b = 2 * v + 2;
      const uint8_t la = lab[a], lb = lab[b];
      if (la != lb) {
        lab[v] = OR_DONE;
      } else if (la == OR_FIXED_DUPL) {
        lab[v] = OR_FIXED_DUPL;
      } else if (la == OR_SHIFT_DUPL) {
        uint8_t dg[16];
        or_hash(d->hash_kind, &curr[2 * a], 32, dg);
        uint64_t k0, k1;
        memcpy(&k0, dg, 8);
        memcpy(&k1, dg + 8, 8);
        const uint64_t s = map_find(m, k0, k1, K, curr);
        if (s != SLOT_EMPTY) {
          curr[2 * v] = k0;
          curr[2 * v + 1] = k1;
          slot_of[v] = (uint32_t)s;
          lab[v] = OR_SHIFT_DUPL;
        } else {
          lab[v] = OR_DONE;
        }
      }
      /* both FIRST or both DONE: the label of pass 3 stays */
    }
  }

  /* 5. roots in node order: per-thread counts over contiguous node ranges */
  const int nt = omp_get_max_threads();
  uint64_t* cnt = (uint64_t*)calloc((size_t)3 * (nt + 1), 8);
  if (!cnt) return OR_ENOMEM;
#pragma omp parallel num_threads(nt)
  {
    const int t = omp_get_thread_num(), T = omp_get_num_threads();
    const uint64_t v0 = N * t / T, v1 = N * (t + 1) / T;
    uint64_t f = 0, fc = 0, s = 0;
    for (uint64_t v = v0; v < v1; v++) {
      if (!is_root(lab, v)) continue;
      if (lab[v] == OR_FIRST_OCUR) {
        f++;
        fc += or_subtree_leaves((uint32_t)v, (uint32_t)C, NULL);
      } else {
        s++;
      }
    }
    cnt[3 * (t + 1)] = f;
    cnt[3 * (t + 1) + 1] = fc;
    cnt[3 * (t + 1) + 2] = s;
  }
  for (int t = 0; t < nt; t++)
    for (int q = 0; q < 3; q++) cnt[3 * (t + 1) + q] += cnt[3 * t + q];
  const uint64_t F = cnt[3 * nt], DQ = cnt[3 * nt + 1], S = cnt[3 * nt + 2];
  uint32_t* f_list = (uint32_t*)malloc((F + 1) * 4);
  uint64_t* f_off = (uint64_t*)malloc((F + 1) * 8);
  uint32_t* s_node = (uint32_t*)malloc((S + 1) * 4);
  uint32_t* s_prev = (uint32_t*)malloc((S + 1) * 4);
  uint32_t* s_tree = (uint32_t*)malloc((S + 1) * 4);
  uint32_t* ss_node = (uint32_t*)malloc((S + 1) * 4);
  uint32_t* ss_prev = (uint32_t*)malloc((S + 1) * 4);
  uint64_t* hist = (uint64_t*)calloc((size_t)(nt + 1) * (K + 1), 8);
  if (!f_list || !f_off || !s_node || !s_prev || !s_tree || !ss_node || !ss_prev || !hist) {
    free(cnt); free(f_list); free(f_off); free(s_node); free(s_prev); free(s_tree); free(ss_node); free(ss_prev);
    free(hist);
    return OR_ENOMEM;
  }
#pragma omp parallel num_threads(nt)
  {
    const int t = omp_get_thread_num(), T = omp_get_num_threads();
    const uint64_t v0 = N * t / T, v1 = N * (t + 1) / T;
    uint64_t f = cnt[3 * t], fc = cnt[3 * t + 1], s = cnt[3 * t + 2];
    for (uint64_t v = v0; v < v1; v++) {
      if (!is_root(lab, v)) continue;
      if (lab[v] == OR_FIRST_OCUR) {
        f_list[f] = (uint32_t)v;
        f_off[f] = fc;
        f++;
        fc += or_subtree_leaves((uint32_t)v, (uint32_t)C, NULL);
      } else {
        /* collect_diff :572-585,690-696: the source is the map entry of the digest */
        const uint64_t val = m->s[slot_of[v]].val;
        s_node[s] = (uint32_t)v;
        s_prev[s] = (uint32_t)val;
        s_tree[s] = (uint32_t)(val >> 32);
        s++;
      }
    }
  }
  f_off[F] = DQ;
  /* Stable counting sort of the shifted roots by source checkpoint. */
#pragma omp parallel num_threads(nt)
  {
    const int t = omp_get_thread_num(), T = omp_get_num_threads();
    const uint64_t e0 = S * t / T, e1 = S * (t + 1) / T;
    uint64_t* h = hist + (uint64_t)t * (K + 1);
    for (uint64_t e = e0; e < e1; e++) h[s_tree[e]]++;
#pragma omp barrier
#pragma omp single
    {
      uint64_t run = 0;
      for (uint64_t bin = 0; bin <= K; bin++) {
        uint64_t tot = 0;
        for (int q = 0; q < T; q++) {
          const uint64_t c = hist[(uint64_t)q * (K + 1) + bin];
          hist[(uint64_t)q * (K + 1) + bin] = run;
          run += c;
          tot += c;
        }
        hist[(uint64_t)nt * (K + 1) + bin] = tot;
      }
    }
    for (uint64_t e = e0; e < e1; e++) {
      const uint64_t p = h[s_tree[e]]++;
      ss_node[p] = s_node[e];
      ss_prev[p] = s_prev[e];
    }
  }
  uint64_t P = 0;
  const uint64_t* tot = hist + (uint64_t)nt * (K + 1);
  for (uint64_t bin = 0; bin <= K; bin++) P += tot[bin] ? 1 : 0;

  /* 6. collect_diff layout */
  const uint64_t meta = 4 * F + 8 * P + 8 * S;
  const uint64_t blen = 32 + meta + DQ * cs;
  uint8_t* buf = (uint8_t*)malloc(blen);
  int rc = OR_OK;
  if (!buf) {
    rc = OR_ENOMEM;
  } else {
    put32(buf + 0, d->baseline_id);
    put32(buf + 4, K);
    put64(buf + 8, len);
    put32(buf + 16, d->chunk_size);
    put32(buf + 20, (uint32_t)F);
    put32(buf + 24, (uint32_t)P);
    put32(buf + 28, (uint32_t)S);
    uint8_t* pf = buf + 32;
    uint8_t* pp = pf + 4 * F;
    uint8_t* ps = pp + 8 * P;
    uint8_t* pd = ps + 8 * S;
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < F; i++) put32(pf + 4 * i, f_list[i]);
    for (uint64_t bin = 0, q = 0; bin <= K; bin++) {
      if (!tot[bin]) continue;
      put32(pp + 8 * q, (uint32_t)bin);
      put32(pp + 8 * q + 4, (uint32_t)tot[bin]);
      q++;
    }
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < S; i++) {
      put32(ps + 8 * i, ss_node[i]);
      put32(ps + 8 * i + 4, ss_prev[i]);
    }
    /* Data section: data chunk q lies in region r with f_off[r] <= q <
     * f_off[r+1]; chunks are split over the threads, not regions (one region
     * can hold the whole buffer). */
#pragma omp parallel for schedule(static, 4096)
    for (uint64_t q = 0; q < DQ; q++) {
      uint64_t lo = 0, hi = F; /* f_off[lo] <= q < f_off[hi] */
      while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (f_off[mid] <= q) lo = mid;
        else hi = mid;
      }
      const uint64_t leaf = subtree_leaf(f_list[lo], C, q - f_off[lo]);
      const uint64_t src = (leaf - L0) * cs;
      const uint64_t w = src + cs > len ? len - src : cs;
      uint8_t* p = pd + q * cs;
      memcpy(p, data + src, w);
      if (w < cs) memset(p + w, 0, cs - w);
    }
  }
  free(cnt); free(f_list); free(f_off); free(s_node); free(s_prev); free(s_tree); free(ss_node); free(ss_prev);
  free(hist);
  if (rc != OR_OK) {
    free(buf);
    return rc;
  }
  /* the current tree becomes the previous one (:121) */
  d->prev = curr;
  d->curr = prev;
  d->nfirst = F;
  d->nshift = S;
  d->current_id = K + 1;
  *out = buf;
  *out_len = blen;
  return OR_OK;
}
