// dd_kernels.hip -- CDNA4 (gfx950) kernels of the tree-approach deduplicator.
//
// Checkpoint pipeline, per job (reference: src/tree_approach.cpp:261-710 and
// the serial oracle include/reference_impl.hpp:15-132):
//   begin          zero the job counters
//   leaf_hash      hash every chunk, compare with the previous leaf digest
//   insert_leaves  digest -> (leaf, ckpt) map inserts; the lowest leaf of a
//                  checkpoint wins through a 64-bit atomicMin
//   pass1 (tiles)  first-occurrence forest: FIRST iff both children FIRST,
//                  parent digest = hash(child digests)
//   insert_interior  map inserts of the FIRST interior nodes
//   pass2 (tiles)  fixed / shifted / done labels with map lookups
//   roots_*        roots in node order, region sizes, block offsets
//   sbin_*         shifted roots grouped by source checkpoint (stable)
//   write_meta     header, first-occurrence list, prior table, shifted list
//   gather         first-occurrence chunks into the data section
#include <hip/hip_runtime.h>

#include "dd_kernels.hpp"

namespace dd {

// ---------------------------------------------------------------------------
// Small wave / block helpers (64-wide wavefronts).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ unsigned long long lanemask_lt() { return (1ull << lane_id()) - 1ull; }

__device__ __forceinline__ uint32_t find_job(const uint32_t* prefix, uint32_t njobs, uint32_t blk) {
  uint32_t lo = 0, hi = njobs;  // prefix[lo] <= blk < prefix[hi]
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (prefix[mid] <= blk) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T n = __shfl_up(v, o, 64);
    if ((int)lane_id() >= o) v += n;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Exclusive scan over the block; NW = number of waves.  `total` gets the sum.
template <int NW, typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* sh, T& total) {
  T inc = wave_incl_scan(v);
  const uint32_t w = threadIdx.x >> 6;
  if (lane_id() == 63) sh[w] = inc;
  __syncthreads();
  if (threadIdx.x < 64) {
    T x = threadIdx.x < NW ? sh[threadIdx.x] : T(0);
    T xi = wave_incl_scan(x);
    if (threadIdx.x < NW) sh[threadIdx.x] = xi - x;
    if (threadIdx.x == NW - 1) sh[NW] = xi;
  }
  __syncthreads();
  T res = sh[w] + inc - v;
  total = sh[NW];
  __syncthreads();
  return res;
}

__device__ __forceinline__ void store_u32(uint8_t* p, uint32_t v) { *reinterpret_cast<uint32_t*>(p) = v; }

// ---------------------------------------------------------------------------
// begin: zero the counters of every job (or only the flags, for a re-run of
// the map inserts after the map has grown).
// ---------------------------------------------------------------------------
__global__ void k_begin(const JobDev* jobs, uint32_t njobs, int rerun) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= njobs) return;
  JobCounters* c = jobs[j].ctr;
  c->map_count = jobs[j].map_count_at_start;
  c->flags = 0;
  if (rerun) return;  // keep `pending` from leaf_hash when re-running the inserts
  c->pending = 0;
  c->nfirst = c->nshift = c->nprior = 0;
  c->dchunks = 0;
  c->diff_len = 0;
  c->unresolved = 0;
}

// ---------------------------------------------------------------------------
// leaf_hash: one lane per chunk.  Leaf digests are written in place over the
// previous checkpoint's digests, only when they change.
// Reference: "Checkpoint K: Leaves", src/tree_approach.cpp:305-351 and
// include/reference_impl.hpp:43-71.
// ---------------------------------------------------------------------------
template <int Kind, bool Aligned16>
__global__ __launch_bounds__(256) void k_leaf_hash(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobDev& J = jobs[j];
  const uint64_t k = (uint64_t)(blockIdx.x - prefix[j]) * blockDim.x + threadIdx.x;
  bool pending = false;
  if (k < J.C) {
    const uint64_t cs = J.chunk_size;
    const uint64_t off = k * cs;
    const uint64_t nb = (k == J.C - 1) ? J.len - off : cs;
    const uint8_t* p = J.data + off;
    Digest d;
    if (Kind == HASH_MD5) {
      d = Aligned16 ? md5_range<true>(p, nb) : md5_range<false>(p, nb);
    } else {
      d = Aligned16 ? murmur3_aligned(p, nb) : murmur3_bytes(p, nb);
    }
    const uint64_t leaf = J.C - 1 + k;
    Digest prev{0, 0};
    if (!J.is_baseline) prev = J.tree[leaf];
    uint8_t lab;
    if (digest_eq(d, prev)) {
      // A fixed leaf of an incremental checkpoint: its digest is already in
      // the map (every leaf digest of the previous checkpoint was inserted),
      // so the insert the serial reference performs is a no-op.  In a
      // baseline the previous digest is all-zero and the insert still happens.
      lab = J.is_baseline ? (uint8_t)(FIXED_DUPL | PENDING) : (uint8_t)FIXED_DUPL;
      pending = J.is_baseline;
    } else {
      lab = PENDING;
      pending = true;
    }
    if (pending) J.tree[leaf] = d;
    J.labels[leaf] = lab;
  }
  unsigned long long b = __ballot(pending);
  if (lane_id() == 0 && b) atomicAdd(&J.ctr->pending, (unsigned long long)__popcll(b));
}

// ---------------------------------------------------------------------------
// Map insertion shared by leaves and interior nodes.
//
// A slot is claimed by a 64-bit CAS of its value word from EMPTY to
// (ckpt << 32 | node).  The key of an entry claimed during this checkpoint is
// read from the tree array (tree[node]), written by an earlier kernel, so no
// lane ever waits for another lane's key store; the claiming lane stores the
// key into the slot for later kernels.  Entries of the same checkpoint and
// digest then agree on one node: the lowest (leaves, UseMax=false) or the
// highest (interior nodes, UseMax=true), which is the first one the serial
// reference visits (include/reference_impl.hpp:43-82).
// ---------------------------------------------------------------------------
template <bool UseMax>
__device__ inline bool map_insert(const JobDev& J, const Digest& k, uint32_t node, uint64_t* slot_out, bool* claimed) {
  const MapView& m = J.map;
  const unsigned long long mine = pack_node(J.ckpt_id, node);
  uint64_t s = map_start(m, k) & m.mask;
  *claimed = false;
  for (uint64_t it = 0; it <= m.mask; it++) {
    Slot* sl = &m.slots[s];
    unsigned long long v = __hip_atomic_load(&sl->val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (v == SLOT_EMPTY) {
      unsigned long long old = atomicCAS(&sl->val, SLOT_EMPTY, mine);
      if (old == SLOT_EMPTY) {
        sl->k0 = k.k0;
        sl->k1 = k.k1;
        *slot_out = s;
        *claimed = true;
        return true;
      }
      v = old;
    }
    const uint32_t vt = (uint32_t)(v >> 32);
    Digest e;
    if (vt == J.ckpt_id) {
      e = J.tree[(uint32_t)v];
    } else {
      e.k0 = sl->k0;
      e.k1 = sl->k1;
    }
    if (digest_eq(e, k)) {
      if (vt == J.ckpt_id) {
        if (UseMax) atomicMax(&sl->val, mine);
        else atomicMin(&sl->val, mine);
      }
      *slot_out = s;
      return true;
    }
    s = (s + 1) & m.mask;
  }
  return false;
}

__device__ __forceinline__ bool map_guard_tripped(const JobDev& J) {
  // Worst case of this checkpoint: every pending leaf and one interior node
  // per pending leaf become new entries.
  return J.map_count_at_start + 2 * J.ctr->pending > J.map_limit;
}

__global__ __launch_bounds__(256) void k_insert_leaves(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobDev& J = jobs[j];
  if (map_guard_tripped(J)) {
    if (blockIdx.x == prefix[j] && threadIdx.x == 0) atomicOr(&J.ctr->flags, FLAG_NEED_GROW);
    return;
  }
  const uint64_t k = (uint64_t)(blockIdx.x - prefix[j]) * blockDim.x + threadIdx.x;
  bool claimed = false;
  if (k < J.C) {
    const uint64_t leaf = J.C - 1 + k;
    if (J.labels[leaf] & PENDING) {
      uint64_t s;
      if (map_insert<false>(J, J.tree[leaf], (uint32_t)leaf, &s, &claimed)) {
        J.slot_of[leaf] = (uint32_t)s;
      } else {
        atomicOr(&J.ctr->flags, FLAG_MAP_FULL);
      }
    }
  }
  unsigned long long b = __ballot(claimed);
  if (lane_id() == 0 && b) atomicAdd(&J.ctr->map_count, (unsigned long long)__popcll(b));
}

__global__ __launch_bounds__(256) void k_insert_interior(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobDev& J = jobs[j];
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;
  const uint64_t x = (uint64_t)(blockIdx.x - prefix[j]) * blockDim.x + threadIdx.x;
  bool claimed = false;
  if (x + 1 < J.C && J.labels[x] == FIRST_OCUR) {
    uint64_t s;
    if (!map_insert<true>(J, J.tree[x], (uint32_t)x, &s, &claimed)) atomicOr(&J.ctr->flags, FLAG_MAP_FULL);
  }
  unsigned long long b = __ballot(claimed);
  if (lane_id() == 0 && b) atomicAdd(&J.ctr->map_count, (unsigned long long)__popcll(b));
}

// ---------------------------------------------------------------------------
// Tree tiles.  A tile is a subtree of kTileLevels levels above a bottom level
// Lb; group g of a job with depth D has Lb = D - 8g.  Nodes on the bottom
// level are loaded; nodes above are computed from their children in LDS,
// except leaves one level above the deepest (chunk counts that are not a power
// of two), which are loaded too.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int tile_groups(uint32_t D) { return D == 0 ? 1 : (int)((D + kTileLevels - 1) / kTileLevels); }

// Label of a leaf after the map inserts: FIXED, FIRST when the map names this
// leaf as the checkpoint's first occurrence, SHIFT otherwise.  Written back.
__device__ __forceinline__ uint8_t resolve_leaf(const JobDev& J, uint64_t leaf) {
  uint8_t raw = J.labels[leaf];
  if (!(raw & PENDING)) return raw;
  uint8_t lab;
  if (raw == (uint8_t)(FIXED_DUPL | PENDING)) {
    lab = FIXED_DUPL;
  } else {
    unsigned long long v = J.map.slots[J.slot_of[leaf]].val;
    lab = (v == pack_node(J.ckpt_id, (uint32_t)leaf)) ? FIRST_OCUR : SHIFT_DUPL;
  }
  J.labels[leaf] = lab;
  return lab;
}

// Pass 1, the first-occurrence forest (include/reference_impl.hpp:73-82).
__global__ __launch_bounds__(256) void k_pass1(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, int g) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobDev& J = jobs[j];
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;
  if (g >= tile_groups(J.D)) return;
  const int Lb = (int)J.D - kTileLevels * g;
  const int Lr = Lb > kTileLevels ? Lb - kTileLevels : 0;
  const int H = Lb - Lr;
  const uint64_t r = ((1ull << Lr) - 1) + (blockIdx.x - prefix[j]);
  __shared__ uint8_t lab[2][256];
  __shared__ Digest dig[2][256];
  const uint32_t i = threadIdx.x;
  const uint32_t width = 1u << H;
  if (i < width) {
    const uint64_t x = ((r + 1) << H) - 1 + i;
    uint8_t L = 0xff;
    Digest d{0, 0};
    if (x < J.N) {
      L = (g == 0) ? resolve_leaf(J, x) : J.labels[x];
      if (L == FIRST_OCUR) d = J.tree[x];
    }
    lab[0][i] = L;
    dig[0][i] = d;
  }
  int cur = 0;
  for (int st = 1; st <= H; st++) {
    __syncthreads();
    const uint32_t w = width >> st;
    if (i < w) {
      const uint64_t x = ((r + 1) << (H - st)) - 1 + i;
      uint8_t L;
      Digest d{0, 0};
      if (x >= J.C - 1) {  // a leaf one level above the deepest
        L = resolve_leaf(J, x);
        if (L == FIRST_OCUR) d = J.tree[x];
      } else {
        const uint8_t a = lab[cur][2 * i], b = lab[cur][2 * i + 1];
        if (a == FIRST_OCUR && b == FIRST_OCUR) {
          L = FIRST_OCUR;
          d = hash_pair(J.hash_kind, dig[cur][2 * i], dig[cur][2 * i + 1]);
          J.tree[x] = d;
        } else {
          L = DONE;
        }
        J.labels[x] = L;
      }
      lab[cur ^ 1][i] = L;
      dig[cur ^ 1][i] = d;
    }
    cur ^= 1;
  }
}

// Pass 2, the duplicate forest (include/reference_impl.hpp:83-104).
__global__ __launch_bounds__(256) void k_pass2(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, int g) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobDev& J = jobs[j];
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;
  if (g >= tile_groups(J.D)) return;
  const int Lb = (int)J.D - kTileLevels * g;
  const int Lr = Lb > kTileLevels ? Lb - kTileLevels : 0;
  const int H = Lb - Lr;
  const uint64_t r = ((1ull << Lr) - 1) + (blockIdx.x - prefix[j]);
  __shared__ uint8_t lab[2][256];
  __shared__ Digest dig[2][256];
  const uint32_t i = threadIdx.x;
  const uint32_t width = 1u << H;
  if (i < width) {
    const uint64_t x = ((r + 1) << H) - 1 + i;
    uint8_t L = 0xff;
    Digest d{0, 0};
    if (x < J.N) {
      L = J.labels[x];
      if (L == SHIFT_DUPL) d = J.tree[x];
    }
    lab[0][i] = L;
    dig[0][i] = d;
  }
  int cur = 0;
  for (int st = 1; st <= H; st++) {
    __syncthreads();
    const uint32_t w = width >> st;
    if (i < w) {
      const uint64_t x = ((r + 1) << (H - st)) - 1 + i;
      uint8_t L;
      Digest d{0, 0};
      if (x >= J.C - 1) {
        L = J.labels[x];
        if (L == SHIFT_DUPL) d = J.tree[x];
      } else {
        const uint8_t a = lab[cur][2 * i], b = lab[cur][2 * i + 1];
        if (a != b) {
          L = DONE;
        } else if (a == SHIFT_DUPL) {
          d = hash_pair(J.hash_kind, dig[cur][2 * i], dig[cur][2 * i + 1]);
          const uint64_t s = map_find(J.map, d);
          if (s != ~0ull) {
            L = SHIFT_DUPL;
            J.tree[x] = d;
            J.slot_of[x] = (uint32_t)s;
          } else {
            L = DONE;
          }
        } else {
          L = a;  // both FIXED, both FIRST or both DONE
        }
        J.labels[x] = L;
      }
      lab[cur ^ 1][i] = L;
      dig[cur ^ 1][i] = d;
    }
    cur ^= 1;
  }
}

// ---------------------------------------------------------------------------
// Roots.  Node v is a root when it is FIRST or SHIFT and its parent is DONE;
// node 0 is a root when it is FIRST (include/reference_impl.hpp:85-106).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint8_t root_kind(const uint8_t* labels, uint64_t v) {
  const uint8_t L = labels[v];
  if (L != FIRST_OCUR && L != SHIFT_DUPL) return 0xff;
  if (v == 0) return L == FIRST_OCUR ? L : 0xff;
  return labels[(v - 1) >> 1] == DONE ? L : 0xff;
}

__device__ __forceinline__ uint64_t leaf_count(const TreeGeom& g, uint64_t v) {
  LeafRuns r = subtree_runs(g, v);
  return r.n0 + r.n1;
}

__global__ __launch_bounds__(256) void k_roots_count(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobDev& J = jobs[j];
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;
  const uint32_t b = blockIdx.x - prefix[j];
  const TreeGeom g = make_geom(J.C);
  const uint64_t base = (uint64_t)b * kNodesPerBlock + threadIdx.x * 16;
  uint32_t f = 0, fc = 0, s = 0;
  for (int t = 0; t < 16; t++) {
    const uint64_t v = base + t;
    if (v >= J.N) break;
    const uint8_t rk = root_kind(J.labels, v);
    if (rk == FIRST_OCUR) {
      f++;
      fc += (uint32_t)leaf_count(g, v);
    } else if (rk == SHIFT_DUPL) {
      s++;
    }
  }
  __shared__ uint32_t red[3][4];
  f = wave_sum(f);
  fc = wave_sum(fc);
  s = wave_sum(s);
  if (lane_id() == 0) {
    red[0][threadIdx.x >> 6] = f;
    red[1][threadIdx.x >> 6] = fc;
    red[2][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    J.blk_f[b] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    J.blk_fc[b] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    J.blk_s[b] = red[2][0] + red[2][1] + red[2][2] + red[2][3];
  }
}

// One workgroup per job: exclusive scans of the per-block counts.
__global__ __launch_bounds__(1024) void k_roots_scan(const JobDev* jobs) {
  const JobDev& J = jobs[blockIdx.x];
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;
  __shared__ uint32_t sh[17];
  uint32_t carry_f = 0, carry_fc = 0, carry_s = 0;
  for (uint32_t base = 0; base < J.nblk; base += 1024) {
    const uint32_t i = base + threadIdx.x;
    uint32_t f = i < J.nblk ? J.blk_f[i] : 0;
    uint32_t fc = i < J.nblk ? J.blk_fc[i] : 0;
    uint32_t s = i < J.nblk ? J.blk_s[i] : 0;
    uint32_t tf, tfc, ts;
    uint32_t ef = block_excl_scan<16>(f, sh, tf);
    uint32_t efc = block_excl_scan<16>(fc, sh, tfc);
    uint32_t es = block_excl_scan<16>(s, sh, ts);
    if (i < J.nblk) {
      J.blk_f[i] = carry_f + ef;
      J.blk_fc[i] = carry_fc + efc;
      J.blk_s[i] = carry_s + es;
    }
    carry_f += tf;
    carry_fc += tfc;
    carry_s += ts;
  }
  if (threadIdx.x == 0) {
    J.ctr->nfirst = carry_f;
    J.ctr->dchunks = carry_fc;
    J.ctr->nshift = carry_s;
  }
}

__global__ __launch_bounds__(256) void k_roots_write(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobDev& J = jobs[j];
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;
  const uint32_t b = blockIdx.x - prefix[j];
  const TreeGeom g = make_geom(J.C);
  const uint64_t base = (uint64_t)b * kNodesPerBlock + threadIdx.x * 16;
  uint8_t kinds[16];
  uint32_t f = 0, fc = 0, s = 0;
#pragma unroll
  for (int t = 0; t < 16; t++) {
    const uint64_t v = base + t;
    uint8_t rk = 0xff;
    if (v < J.N) rk = root_kind(J.labels, v);
    kinds[t] = rk;
    if (rk == FIRST_OCUR) {
      f++;
      fc += (uint32_t)leaf_count(g, v);
    } else if (rk == SHIFT_DUPL) {
      s++;
    }
  }
  __shared__ uint32_t sh[5];
  uint32_t tot;
  uint32_t pf = block_excl_scan<4>(f, sh, tot) + J.blk_f[b];
  uint32_t pfc = block_excl_scan<4>(fc, sh, tot) + J.blk_fc[b];
  uint32_t ps = block_excl_scan<4>(s, sh, tot) + J.blk_s[b];
#pragma unroll
  for (int t = 0; t < 16; t++) {
    const uint64_t v = base + t;
    if (kinds[t] == FIRST_OCUR) {
      J.f_list[pf] = (uint32_t)v;
      J.f_off[pf] = pfc;
      pf++;
      pfc += (uint32_t)leaf_count(g, v);
    } else if (kinds[t] == SHIFT_DUPL) {
      const unsigned long long val = J.map.slots[J.slot_of[v]].val;
      J.s_node[ps] = (uint32_t)v;
      J.s_prev[ps] = (uint32_t)val;
      J.s_tree[ps] = (uint32_t)(val >> 32);
      ps++;
    }
  }
}

// ---------------------------------------------------------------------------
// Shifted roots grouped by source checkpoint (a stable counting sort; the
// reference uses Kokkos::BinSort, src/tree_approach.cpp:670-687).
// s_hist is bin-major: entry (bin, tile) at bin * ntiles + tile.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t stiles_of(const JobDev& J) { return (J.ctr->nshift + kSTile - 1) / kSTile; }

__global__ __launch_bounds__(256) void k_sbin_count(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobDev& J = jobs[j];
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;
  const uint32_t t = blockIdx.x - prefix[j];
  const uint32_t nt = stiles_of(J);
  if (t >= nt) return;
  const uint32_t nbins = J.ckpt_id + 1;
  for (uint32_t bn = threadIdx.x; bn < nbins; bn += blockDim.x) J.s_hist[(uint64_t)bn * nt + t] = 0;
  __syncthreads();
  const uint32_t S = J.ctr->nshift;
  for (uint32_t e = t * kSTile + threadIdx.x; e < S && e < (t + 1) * kSTile; e += blockDim.x) {
    uint32_t bn = J.s_tree[e];
    if (bn >= nbins) {
      atomicOr(&J.ctr->flags, FLAG_BAD_SOURCE);
      bn = nbins - 1;
    }
    atomicAdd(&J.s_hist[(uint64_t)bn * nt + t], 1u);
  }
}

// One workgroup per job: scan of the histogram and the prior table.
__global__ __launch_bounds__(1024) void k_sbin_scan(const JobDev* jobs) {
  const JobDev& J = jobs[blockIdx.x];
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;
  __shared__ uint32_t sh[17];
  const uint32_t nt = stiles_of(J);
  const uint32_t nbins = J.ckpt_id + 1;
  uint32_t nprior = 0;
  if (nt > 0) {
    // Per-bin totals -> prior table entries in ascending checkpoint order.
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nbins; base += 1024) {
      const uint32_t bn = base + threadIdx.x;
      uint32_t cnt = 0;
      if (bn < nbins)
        for (uint32_t t = 0; t < nt; t++) cnt += J.s_hist[(uint64_t)bn * nt + t];
      uint32_t tot;
      uint32_t pos = block_excl_scan<16>(cnt ? 1u : 0u, sh, tot) + carry;
      if (bn < nbins && cnt) {
        J.ptab[2 * pos] = bn;
        J.ptab[2 * pos + 1] = cnt;
      }
      carry += tot;
    }
    nprior = carry;
    // Exclusive scan of the whole histogram in bin-major order.
    const uint64_t total = (uint64_t)nbins * nt;
    uint32_t run = 0;
    for (uint64_t base = 0; base < total; base += 1024) {
      const uint64_t i = base + threadIdx.x;
      uint32_t v = i < total ? J.s_hist[i] : 0;
      uint32_t tot;
      uint32_t e = block_excl_scan<16>(v, sh, tot);
      if (i < total) J.s_hist[i] = run + e;
      run += tot;
    }
  }
  if (threadIdx.x == 0) {
    J.ctr->nprior = nprior;
    const uint64_t meta = 4ull * J.ctr->nfirst + 8ull * nprior + 8ull * J.ctr->nshift;
    J.ctr->diff_len = 32 + meta + J.ctr->dchunks * (uint64_t)J.chunk_size;
  }
}

// One wavefront per tile walks its entries in order; lanes with the same bin
// are ranked with a ballot so equal bins keep node order.
constexpr uint32_t kLdsBins = 8192;
__global__ __launch_bounds__(64) void k_sbin_scatter(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobDev& J = jobs[j];
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;
  const uint32_t t = blockIdx.x - prefix[j];
  const uint32_t nt = stiles_of(J);
  if (t >= nt) return;
  const uint32_t nbins = J.ckpt_id + 1;
  const uint32_t S = J.ctr->nshift;
  const uint32_t e0 = t * kSTile, e1 = min(S, (t + 1) * kSTile);
  __shared__ uint32_t base[kLdsBins];
  for (uint32_t lo = 0; lo < nbins; lo += kLdsBins) {
    const uint32_t hi = min(nbins, lo + kLdsBins);
    for (uint32_t bn = lo + threadIdx.x; bn < hi; bn += 64) base[bn - lo] = J.s_hist[(uint64_t)bn * nt + t];
    __syncthreads();
    for (uint32_t e = e0; e < e1; e += 64) {
      const uint32_t idx = e + threadIdx.x;
      uint32_t bn = idx < e1 ? J.s_tree[idx] : 0xffffffffu;
      if (bn != 0xffffffffu && bn >= nbins) bn = nbins - 1;
      bool todo = idx < e1 && bn >= lo && bn < hi;
      while (true) {
        unsigned long long act = __ballot(todo);
        if (!act) break;
        const uint32_t leader = __ffsll((long long)act) - 1;
        const uint32_t lb = __shfl(bn, leader, 64);
        unsigned long long same = __ballot(todo && bn == lb);
        const uint32_t cur = base[lb - lo];
        if (todo && bn == lb) {
          const uint32_t pos = cur + __popcll(same & lanemask_lt());
          J.ss_node[pos] = J.s_node[idx];
          J.ss_prev[pos] = J.s_prev[idx];
          todo = false;
        }
        __syncthreads();
        if (threadIdx.x == leader) base[lb - lo] = cur + __popcll(same);
        __syncthreads();
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Diff metadata: header_t, first-occurrence roots, prior table, shifted
// entries (src/tree_approach.cpp:593,622-706).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_write_meta(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobDev& J = jobs[j];
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;
  const uint32_t F = J.ctr->nfirst, P = J.ctr->nprior, S = J.ctr->nshift;
  uint8_t* out = J.diff;
  const uint32_t tid = (blockIdx.x - prefix[j]) * blockDim.x + threadIdx.x;
  const uint32_t nth = (prefix[j + 1] - prefix[j]) * blockDim.x;
  if (tid == 0) {
    store_u32(out + 0, J.baseline_id);
    store_u32(out + 4, J.ckpt_id);
    store_u32(out + 8, (uint32_t)J.len);
    store_u32(out + 12, (uint32_t)(J.len >> 32));
    store_u32(out + 16, J.chunk_size);
    store_u32(out + 20, F);
    store_u32(out + 24, P);
    store_u32(out + 28, S);
  }
  uint8_t* pf = out + 32;
  uint8_t* pp = pf + 4ull * F;
  uint8_t* ps = pp + 8ull * P;
  for (uint32_t i = tid; i < F; i += nth) store_u32(pf + 4ull * i, J.f_list[i]);
  for (uint32_t i = tid; i < P; i += nth) {
    store_u32(pp + 8ull * i, J.ptab[2 * i]);
    store_u32(pp + 8ull * i + 4, J.ptab[2 * i + 1]);
  }
  for (uint32_t i = tid; i < S; i += nth) {
    store_u32(ps + 8ull * i, J.ss_node[i]);
    store_u32(ps + 8ull * i + 4, J.ss_prev[i]);
  }
}

// ---------------------------------------------------------------------------
// Gather: data chunk q of the data section comes from first region r with
// f_off[r] <= q < f_off[r+1]; its index inside the region picks the leaf in
// left-to-right order.  Copies 4-byte words (the data section is only 4-byte
// aligned: it follows 4-byte first-occurrence ids).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t region_of(const uint32_t* f_off, uint32_t F, uint64_t q) {
  uint32_t lo = 0, hi = F;  // f_off[lo] <= q < f_off[hi] (f_off[F] = total)
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (f_off[mid] <= q) lo = mid;
    else hi = mid;
  }
  return lo;
}

typedef uint32_t u32x4a4 __attribute__((ext_vector_type(4), aligned(4)));

// Copies `n` bytes of chunk data from `src` (16-byte aligned when Vec) to the
// 4-byte aligned `dst`, zero-filling the bytes at and past `avail`.
template <bool Vec>
__device__ __forceinline__ void copy_piece(uint8_t* dst, const uint8_t* src, uint64_t avail) {
  if (Vec) {
    if (avail >= 16) {
      const uint4 v = *reinterpret_cast<const uint4*>(src);
      *reinterpret_cast<u32x4a4*>(dst) = u32x4a4{v.x, v.y, v.z, v.w};
      return;
    }
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      uint32_t x = 0;
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const uint64_t o = 4 * k + b;
        if (o < avail) x |= (uint32_t)src[o] << (8 * b);
      }
      w[k] = x;
    }
    *reinterpret_cast<u32x4a4*>(dst) = u32x4a4{w[0], w[1], w[2], w[3]};
  } else {
    *dst = avail > 0 ? *src : 0;
  }
}

// Groups of 256 data chunks: each lane finds the source chunk of one, then the
// workgroup copies the group with 16-byte pieces (1-byte pieces when the
// chunk size is not a multiple of 16 or the input is not 16-byte aligned).
template <bool Vec>
__global__ __launch_bounds__(256) void k_gather(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobDev& J = jobs[j];
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;
  const uint32_t F = J.ctr->nfirst, P = J.ctr->nprior, S = J.ctr->nshift;
  const uint64_t Dq = J.ctr->dchunks;
  if (Dq == 0) return;
  const TreeGeom g = make_geom(J.C);
  const uint64_t cs = J.chunk_size;
  uint8_t* dst_base = J.diff + 32 + 4ull * F + 8ull * P + 8ull * S;
  __shared__ uint64_t src[256];
  const uint32_t nblk = prefix[j + 1] - prefix[j];
  const uint64_t ngroups = (Dq + 255) / 256;
  const uint32_t piece = Vec ? 16 : 1;
  const uint64_t ppc = cs / piece;  // pieces per chunk
  const bool pow2 = (ppc & (ppc - 1)) == 0;
  const uint32_t sh = pow2 ? (uint32_t)__builtin_ctzll(ppc) : 0;
  for (uint64_t grp = blockIdx.x - prefix[j]; grp < ngroups; grp += nblk) {
    const uint64_t q0 = grp * 256;
    const uint64_t q = q0 + threadIdx.x;
    if (q < Dq) {
      const uint32_t r = region_of(J.f_off, F, q);
      const LeafRuns runs = subtree_runs(g, J.f_list[r]);
      src[threadIdx.x] = run_leaf(runs, q - J.f_off[r]) - (J.C - 1);
    }
    __syncthreads();
    const uint64_t nq = min((uint64_t)256, Dq - q0);
    const uint64_t total = nq * ppc;
    for (uint64_t w = threadIdx.x; w < total; w += 256) {
      const uint64_t i = pow2 ? (w >> sh) : (w / ppc);
      const uint64_t pi = w - i * ppc;
      const uint64_t so = src[i] * cs + pi * piece;
      const uint64_t avail = so < J.len ? J.len - so : 0;
      copy_piece<Vec>(dst_base + (q0 + i) * cs + pi * piece, J.data + so, avail);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Map maintenance.
// ---------------------------------------------------------------------------
__global__ void k_map_clear(Slot* slots, uint64_t cap) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += stride) {
    uint4* p = reinterpret_cast<uint4*>(&slots[i]);
    p[0] = make_uint4(0, 0, 0, 0);
    p[1] = make_uint4(0xffffffffu, 0xffffffffu, 0, 0);
  }
}

// Copies every entry into a larger table; keys are unique, so an entry only
// needs an empty slot.
__global__ void k_map_rehash(const Slot* old_slots, uint64_t old_cap, MapView m, uint32_t* flags) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < old_cap; i += stride) {
    const Slot& o = old_slots[i];
    if (o.val == SLOT_EMPTY) continue;
    const Digest k{o.k0, o.k1};
    uint64_t s = map_start(m, k) & m.mask;
    bool done = false;
    for (uint64_t it = 0; it <= m.mask; it++) {
      if (atomicCAS(&m.slots[s].val, SLOT_EMPTY, o.val) == SLOT_EMPTY) {
        m.slots[s].k0 = k.k0;
        m.slots[s].k1 = k.k1;
        done = true;
        break;
      }
      s = (s + 1) & m.mask;
    }
    if (!done) atomicOr(flags, FLAG_MAP_FULL);
  }
}

// ---------------------------------------------------------------------------
// Restart (src/tree_approach.cpp:712-1128).  ref[c] holds the pending source
// of chunk c as (ckpt << 32 | leaf node), or ~0 once the chunk is written.
// Checkpoints are visited newest to oldest; for each, `cover` maps the leaves
// of its first-occurrence and shifted regions (stamped with a generation so
// the array is never cleared).
// ---------------------------------------------------------------------------
constexpr unsigned long long kRefDone = ~0ull;

__global__ void k_restart_init(uint64_t* ref, uint64_t C, uint32_t ckpt, JobCounters* ctr) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i0 == 0) {
    ctr->unresolved = 0;
    ctr->flags = 0;
  }
  for (uint64_t c = i0; c < C; c += stride) ref[c] = ((uint64_t)ckpt << 32) | (C - 1 + c);
}

__device__ __forceinline__ uint32_t load_u32(const uint8_t* p) { return *reinterpret_cast<const uint32_t*>(p); }

__global__ __launch_bounds__(256) void k_restart_cover(const RestartCkpt* rcp, TreeGeom g, CoverEntry* cover,
                                                       uint32_t stamp, uint64_t nfl, uint64_t nsl, uint32_t cs) {
  const RestartCkpt rc = *rcp;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nfl + nsl; e += stride) {
    if (e < nfl) {
      const uint32_t r = region_of(rc.f_off, rc.nfirst, e);
      const uint32_t root = load_u32(rc.diff + rc.first_off + 4ull * r);
      const uint64_t leaf = run_leaf(subtree_runs(g, root), e - rc.f_off[r]);
      CoverEntry ce;
      ce.stamp = stamp;
      ce.src_ckpt = kCoverFirst;
      ce.payload = e * (uint64_t)cs;
      cover[leaf - (g.C - 1)] = ce;
    } else {
      const uint64_t q = e - nfl;
      const uint32_t r = region_of(rc.s_off, rc.nshift, q);
      const uint32_t node = load_u32(rc.diff + rc.shift_off + 8ull * r);
      const uint32_t prev = load_u32(rc.diff + rc.shift_off + 8ull * r + 4);
      const uint64_t idx = q - rc.s_off[r];
      const uint64_t leaf = run_leaf(subtree_runs(g, node), idx);
      const uint64_t src = run_leaf(subtree_runs(g, prev), idx);
      CoverEntry ce;
      ce.stamp = stamp;
      ce.src_ckpt = rc.s_tree[r];
      ce.payload = src;
      cover[leaf - (g.C - 1)] = ce;
    }
  }
}

// Each lane resolves one chunk; chunks whose data is in this checkpoint are
// then copied by the whole workgroup, 4-byte words across lanes.
__global__ __launch_bounds__(256) void k_restart_resolve(const RestartCkpt* rcp, TreeGeom g, const CoverEntry* cover,
                                                         uint32_t stamp, uint64_t* ref, uint8_t* out, uint64_t len,
                                                         uint32_t cs, JobCounters* ctr) {
  const RestartCkpt rc = *rcp;
  __shared__ uint64_t src_off[256];
  const uint32_t cur = rc.ckpt_id;
  const uint64_t c0 = (uint64_t)blockIdx.x * 256;
  const uint64_t c = c0 + threadIdx.x;
  uint64_t so = ~0ull;
  if (c < g.C) {
    const uint64_t r = ref[c];
    if (r != kRefDone && (uint32_t)(r >> 32) == cur) {
      const uint64_t x = (uint32_t)r;
      const CoverEntry e = cover[x - (g.C - 1)];
      if (e.stamp == stamp) {
        if (e.src_ckpt == kCoverFirst) {
          so = e.payload;
        } else if (e.src_ckpt == cur) {
          const CoverEntry e2 = cover[e.payload - (g.C - 1)];
          if (e2.stamp == stamp && e2.src_ckpt == kCoverFirst) so = e2.payload;
          else atomicOr(&ctr->flags, FLAG_BAD_SOURCE);
        } else {
          ref[c] = ((uint64_t)e.src_ckpt << 32) | e.payload;
        }
      } else {
        ref[c] = ((uint64_t)(cur - 1) << 32) | x;  // fixed: same leaf, one checkpoint back
      }
      if (so != ~0ull) ref[c] = kRefDone;
    }
  }
  src_off[threadIdx.x] = so;
  __syncthreads();
  const uint8_t* data = rc.diff + rc.data_off;
  const uint64_t nchunks = min((uint64_t)256, g.C - c0);
  if ((cs & 3) == 0) {
    const uint64_t wpc = cs / 4;
    for (uint64_t w = threadIdx.x; w < nchunks * wpc; w += 256) {
      const uint64_t i = w / wpc, wi = w - i * wpc;
      const uint64_t s = src_off[i];
      if (s == ~0ull) continue;
      const uint64_t dst = (c0 + i) * cs + 4 * wi;
      if (dst + 4 <= len) {
        *reinterpret_cast<uint32_t*>(out + dst) = load_u32(data + s + 4 * wi);
      } else if (dst < len) {
        for (uint64_t b = dst; b < len; b++) out[b] = data[s + 4 * wi + (b - dst)];
      }
    }
  } else {
    for (uint64_t w = threadIdx.x; w < nchunks * cs; w += 256) {
      const uint64_t i = w / cs, bi = w - i * cs;
      const uint64_t s = src_off[i];
      if (s == ~0ull) continue;
      const uint64_t dst = (c0 + i) * cs + bi;
      if (dst < len) out[dst] = data[s + bi];
    }
  }
}

__global__ void k_restart_check(const uint64_t* ref, uint64_t C, JobCounters* ctr) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint32_t n = 0;
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < C; c += stride) n += ref[c] != kRefDone;
  n = wave_sum(n);
  if (lane_id() == 0 && n) atomicAdd(&ctr->unresolved, n);
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
#define DD_LAUNCH(kern, grid, block, lds, s, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds, s, __VA_ARGS__)

void launch_begin(const JobDev* jobs, const uint32_t*, uint32_t njobs, bool rerun, hipStream_t s) {
  DD_LAUNCH(k_begin, (njobs + 63) / 64, 64, 0, s, jobs, njobs, rerun ? 1 : 0);
}

void launch_leaf_hash(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, bool al16, int kind,
                      hipStream_t s) {
  if (grid == 0) return;
  if (kind == HASH_MD5) {
    if (al16) DD_LAUNCH((k_leaf_hash<HASH_MD5, true>), grid, 256, 0, s, jobs, prefix, njobs);
    else DD_LAUNCH((k_leaf_hash<HASH_MD5, false>), grid, 256, 0, s, jobs, prefix, njobs);
  } else {
    if (al16) DD_LAUNCH((k_leaf_hash<HASH_MURMUR3, true>), grid, 256, 0, s, jobs, prefix, njobs);
    else DD_LAUNCH((k_leaf_hash<HASH_MURMUR3, false>), grid, 256, 0, s, jobs, prefix, njobs);
  }
}

void launch_insert_leaves(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s) {
  if (grid) DD_LAUNCH(k_insert_leaves, grid, 256, 0, s, jobs, prefix, njobs);
}
void launch_pass1(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, int g, hipStream_t s) {
  if (grid) DD_LAUNCH(k_pass1, grid, 256, 0, s, jobs, prefix, njobs, g);
}
void launch_insert_interior(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s) {
  if (grid) DD_LAUNCH(k_insert_interior, grid, 256, 0, s, jobs, prefix, njobs);
}
void launch_pass2(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, int g, hipStream_t s) {
  if (grid) DD_LAUNCH(k_pass2, grid, 256, 0, s, jobs, prefix, njobs, g);
}
void launch_roots_count(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s) {
  if (grid) DD_LAUNCH(k_roots_count, grid, 256, 0, s, jobs, prefix, njobs);
}
void launch_roots_scan(const JobDev* jobs, uint32_t njobs, hipStream_t s) {
  DD_LAUNCH(k_roots_scan, njobs, 1024, 0, s, jobs);
}
void launch_roots_write(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s) {
  if (grid) DD_LAUNCH(k_roots_write, grid, 256, 0, s, jobs, prefix, njobs);
}
void launch_sbin_count(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s) {
  if (grid) DD_LAUNCH(k_sbin_count, grid, 256, 0, s, jobs, prefix, njobs);
}
void launch_sbin_scan(const JobDev* jobs, uint32_t njobs, hipStream_t s) { DD_LAUNCH(k_sbin_scan, njobs, 1024, 0, s, jobs); }
void launch_sbin_scatter(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s) {
  if (grid) DD_LAUNCH(k_sbin_scatter, grid, 64, 0, s, jobs, prefix, njobs);
}
void launch_write_meta(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s) {
  if (grid) DD_LAUNCH(k_write_meta, grid, 256, 0, s, jobs, prefix, njobs);
}
void launch_gather(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, bool vec16,
                   hipStream_t s) {
  if (!grid) return;
  if (vec16) DD_LAUNCH(k_gather<true>, grid, 256, 0, s, jobs, prefix, njobs);
  else DD_LAUNCH(k_gather<false>, grid, 256, 0, s, jobs, prefix, njobs);
}

void launch_map_clear(Slot* slots, uint64_t cap, hipStream_t s) {
  uint64_t blocks = (cap + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  DD_LAUNCH(k_map_clear, (uint32_t)blocks, 256, 0, s, slots, cap);
}

void launch_map_rehash(const Slot* old_slots, uint64_t old_cap, MapView nmap, uint32_t* flags, hipStream_t s) {
  uint64_t blocks = (old_cap + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  DD_LAUNCH(k_map_rehash, (uint32_t)blocks, 256, 0, s, old_slots, old_cap, nmap, flags);
}

void launch_restart_init(uint64_t* ref, uint64_t C, uint32_t ckpt, JobCounters* ctr, hipStream_t s) {
  uint64_t blocks = (C + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  DD_LAUNCH(k_restart_init, (uint32_t)blocks, 256, 0, s, ref, C, ckpt, ctr);
}

void launch_restart_cover(const RestartCkpt* rc, TreeGeom g, CoverEntry* cover, uint32_t stamp, uint64_t nfl,
                          uint64_t nsl, uint32_t cs, hipStream_t s) {
  uint64_t n = nfl + nsl;
  if (n == 0) return;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  DD_LAUNCH(k_restart_cover, (uint32_t)blocks, 256, 0, s, rc, g, cover, stamp, nfl, nsl, cs);
}

void launch_restart_resolve(const RestartCkpt* rc, TreeGeom g, const CoverEntry* cover, uint32_t stamp, uint64_t* ref,
                            uint8_t* out, uint64_t len, uint32_t cs, JobCounters* ctr, hipStream_t s) {
  uint64_t blocks = (g.C + 255) / 256;
  DD_LAUNCH(k_restart_resolve, (uint32_t)blocks, 256, 0, s, rc, g, cover, stamp, ref, out, len, cs, ctr);
}

void launch_restart_check(const uint64_t* ref, uint64_t C, JobCounters* ctr, hipStream_t s) {
  uint64_t blocks = (C + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  DD_LAUNCH(k_restart_check, (uint32_t)blocks, 256, 0, s, ref, C, ctr);
}

}  // namespace dd
