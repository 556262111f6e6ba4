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
//   gather         first-occurrence chunks into the data section, then the
//                  header, first-occurrence list, prior table, shifted list
#include <hip/hip_ext.h>
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

__device__ __forceinline__ void store_u32(GlobalPtr<uint8_t> p, uint32_t v) { *(GlobalPtr<uint32_t>)p = v; }

// Sum over the block (every lane must call it).
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* sh) {
  v = wave_sum(v);
  const uint32_t w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane_id() == 0) sh[w] = v;
  __syncthreads();
  T t = 0;
  for (uint32_t i = 0; i < nw; i++) t += sh[i];
  __syncthreads();
  return t;
}

// ---------------------------------------------------------------------------
// begin: zero the counters of every job (or only the flags, for a re-run of
// the map inserts after the map has grown).
// ---------------------------------------------------------------------------
__global__ void k_begin(const JobG* __restrict__ jobs, uint32_t njobs, int rerun) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= njobs) return;
  const GlobalPtr<JobCounters> c = jobs[j].ctr;
  c->map_count = jobs[j].map_count_at_start;
  c->flags = 0;
  if (rerun) return;  // keep the changed-leaf lists of leaf_hash when re-running the inserts
  for (uint32_t i = 0; i < 32; i++) {
    jobs[j].pend_cnt[i * kCntStride] = 0;
    jobs[j].dirty_cnt[i * kCntStride] = 0;
  }
  c->pending = 0;
  c->nfirst = c->nshift = c->nprior = 0;
  c->dchunks = 0;
  c->diff_len = 0;
}

// ---------------------------------------------------------------------------
typedef uint32_t u32x4a4 __attribute__((ext_vector_type(4), aligned(4)));

// One piece of chunk data at `src` (16 bytes, 16-byte aligned, when Vec;
// else one byte, in .x), zero at and past `avail` bytes.
template <bool Vec>
__device__ __forceinline__ uint4 load_piece(GlobalPtr<const uint8_t> src, uint64_t avail) {
  if (!Vec) return make_uint4(avail > 0 ? *src : 0u, 0, 0, 0);
  if (avail >= 16) return gload16((GlobalPtr<const void>)src);
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
  return make_uint4(w[0], w[1], w[2], w[3]);
}

// Stores a piece to the 4-byte aligned `dst`.
template <bool Vec>
__device__ __forceinline__ void store_piece(GlobalPtr<uint8_t> dst, const uint4& v) {
  if (Vec) *(GlobalPtr<u32x4a4>)dst = u32x4a4{v.x, v.y, v.z, v.w};
  else *dst = (uint8_t)v.x;
}

// Data-section copy of a tree baseline (JobT::write_data): byte 0 of chunk
// k's place when the root is the diff's only region, after the 32-byte
// header and the root's 4-byte node id.
__device__ __forceinline__ GlobalPtr<uint8_t> root_data_dst(const JobG& J, const TreeGeom& g, uint64_t k) {
  return J.diff + 36 + root_rank(g, k) * (uint64_t)J.chunk_size;
}

// The partial last chunk's copy, zero padded to a whole chunk as the gather
// writes it, by the wavefront that holds chunk nfull (16 bytes per lane per
// step; the buffer is 16-byte aligned and the chunk size a multiple of 16).
__device__ __forceinline__ void copy_partial_chunk(const JobG& J, const TreeGeom& g, uint32_t lane) {
  const uint64_t cs = J.chunk_size;
  const uint64_t off = J.nfull * cs;
  const uint64_t nb = J.len - off;
  const GlobalPtr<uint8_t> dst = root_data_dst(J, g, J.nfull);
  for (uint64_t o = (uint64_t)lane * 16; o < cs; o += 1024)
    store_piece<true>(dst + o, load_piece<true>(J.data + off + o, o < nb ? nb - o : 0));
}

// leaf_hash: one lane per chunk.  Leaf digests are written in place over the
// previous checkpoint's digests, only when they change.
// Reference: "Checkpoint K: Leaves", src/tree_approach.cpp:305-351 and
// include/reference_impl.hpp:43-71.
// ---------------------------------------------------------------------------
// Work lists are split into 32 sub-lists picked by the producer's block or
// tile index: one counter hit by every workgroup would serialize (32K
// same-address atomics per GiB at 128-byte chunks).
constexpr uint32_t kPendSpread = 32;

// Group-0 tile of a leaf (leaves sit on the deepest level D or on D-1).
__device__ __forceinline__ uint32_t leaf_tile(uint32_t D, uint64_t leaf) {
  const uint32_t Lr0 = D > (uint32_t)kTileLevels ? D - kTileLevels : 0;
  const uint32_t lv = ilog2_64(leaf + 1);
  return (uint32_t)(((leaf + 1) >> (lv - Lr0)) - (1ull << Lr0));
}

__device__ __forceinline__ void list_append(GlobalPtr<uint32_t> cnt, GlobalPtr<uint32_t> list, uint32_t cap,
                                            uint32_t item) {
  const uint32_t r = item % kPendSpread;
  const uint32_t k = g_add(&cnt[r * kCntStride], 1u);
  list[(uint64_t)r * cap + k] = item;
}

// The leaves of a wavefront (64 consecutive chunks k, k >> 6 = w) that need a
// map insert are appended to sub-list w % 32 of pend_list (with their digest
// d in pend_dig while the sub-list is within dig_cap): one atomic per
// wavefront with a changed leaf, no block barrier, so a block's waves retire
// independently.  A baseline visits every leaf and tile densely and keeps no
// lists.  Every lane of the wavefront calls this.
__device__ __forceinline__ void record_pending(const JobG& J, uint64_t k, bool pending, const Digest& d) {
  if (J.is_baseline) return;  // uniform: a workgroup belongs to one job
  const unsigned long long b = __ballot(pending);
  if (!b) return;
  const uint32_t r = (uint32_t)(k >> 6) % kPendSpread;
  const uint32_t leader = __ffsll((long long)b) - 1;
  uint32_t base = 0;
  if (lane_id() == leader) base = g_add(&J.pend_cnt[r * kCntStride], (uint32_t)__popcll(b));
  base = __shfl(base, leader, 64);
  const uint32_t pos = base + __popcll(b & lanemask_lt());
  if (pending) {
    J.pend_list[(uint64_t)r * J.pend_cap + pos] = (uint32_t)k;
    if (pos < J.dig_cap) store_digest(J.pend_dig + (uint64_t)r * J.dig_cap + pos, d);
  }
  // Mark the group-0 tiles of changed leaves: the first marker of a tile in
  // this checkpoint appends it to the dirty list.  The list method has no
  // tiles.
  uint32_t T = pending ? leaf_tile(J.D, J.C - 1 + k) : 0;
  bool todo = pending && !J.list_mode;
  while (true) {
    const unsigned long long act = __ballot(todo);
    if (!act) break;
    const uint32_t leader = __ffsll((long long)act) - 1;
    const uint32_t lt = __shfl(T, leader, 64);
    if (lane_id() == leader) {
      if (g_exch(&J.tile_epoch[lt], J.epoch) != J.epoch) list_append(J.dirty_cnt, J.dirty_list, J.dirty_cap, lt);
    }
    if (todo && T == lt) todo = false;
  }
}

// Item i of a 32-way split work list: sub-list counts are prefix-summed in
// LDS by the caller (pre[r] = items before sub-list r, pre[32] = total).
// Lane r of the first wavefront loads count r and the wavefront scans them
// with shuffles: one load round trip per workgroup (a loop of one thread over
// the counts would wait for each load in turn).  Every caller runs at least
// 64 threads.
__device__ __forceinline__ void load_list_prefix(GlobalPtr<const uint32_t> cnt, uint32_t* pre) {
  static_assert(kPendSpread <= 64, "one count per lane of a wavefront");
  if (threadIdx.x < 64) {
    const uint32_t l = threadIdx.x;
    uint32_t v = l < kPendSpread ? cnt[l * kCntStride] : 0;
#pragma unroll
    for (uint32_t o = 1; o < kPendSpread; o <<= 1) {
      const uint32_t u = __shfl_up(v, o, 64);
      if (l >= o) v += u;
    }
    if (l < kPendSpread) pre[l + 1] = v;
    if (l == 0) pre[0] = 0;
  }
  __syncthreads();
}

// Sub-list of item i: the last r in [0, 32) with pre[r] <= i (pre is
// non-decreasing and pre[0] = 0), by binary search.
__device__ __forceinline__ uint32_t list_item(GlobalPtr<const uint32_t> list, uint32_t cap, const uint32_t* pre,
                                              uint32_t i) {
  uint32_t r = 0;
#pragma unroll
  for (uint32_t step = kPendSpread / 2; step > 0; step >>= 1)
    if (pre[r + step] <= i) r += step;
  return list[(uint64_t)r * cap + (i - pre[r])];
}

// Sub-list-major walk of a 32-way split work list by P workgroups that take
// B items per step.  With P >= 32, workgroup w shares sub-list w % 32 with
// the other workgroups of its residue and takes items q*B .. q*B+B-1 of it,
// then every Q*B further (q = w / 32, Q workgroups share the sub-list); with
// P < 32 it takes sub-lists w, w + P, ... whole.  A worker's first item does
// not depend on the sub-list counts, so it can be loaded together with them.
struct ListWalk {
  uint32_t r, r_step;  // first sub-list, step to the next one
  uint32_t j0, j_step;  // first item (of the workgroup's B) in the sub-list, step
};

__device__ __forceinline__ ListWalk list_walk(uint32_t w, uint32_t P, uint32_t B) {
  ListWalk L;
  if (P >= kPendSpread) {
    L.r = w % kPendSpread;
    L.r_step = kPendSpread;
    L.j0 = (w / kPendSpread) * B;
    L.j_step = ((P - 1 - L.r) / kPendSpread + 1) * B;
  } else {
    L.r = w;
    L.r_step = P;
    L.j0 = 0;
    L.j_step = B;
  }
  return L;
}

__device__ __forceinline__ uint64_t pending_total(const JobG& J) {
  uint64_t t = 0;
#pragma unroll
  for (uint32_t i = 0; i < kPendSpread; i++) t += J.pend_cnt[i * kCntStride];
  return t;
}

// Pipe (Murmur3 with 16-byte aligned data; MD5 ignores it): LEAF_DIRECT
// hashes 64 bytes per group of loads, LEAF_DEEP keeps two 128-byte groups in
// flight per lane (long chunks), LEAF_SHORT loads a whole chunk of at most
// 128 bytes before hashing it.
template <int Kind, bool Aligned16, int Pipe>
__global__ __launch_bounds__(256) void k_leaf_hash(const JobG* __restrict__ jobs, const uint32_t* __restrict__ prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobG& J = jobs[j];
  const uint64_t k = (uint64_t)(blockIdx.x - prefix[j]) * blockDim.x + threadIdx.x;
  bool pending = false;
  Digest d{0, 0};
  if (k < J.C) {
    const uint64_t cs = J.chunk_size;
    const uint64_t off = k * cs;
    const uint64_t nb = (k == J.C - 1) ? J.len - off : cs;
    const uint8_t* p = (const uint8_t*)(J.data + off);
    const uint64_t leaf = J.C - 1 + k;
    // The previous digest is loaded before the chunk: its latency hides
    // behind the chunk loads instead of following the hash.
    Digest prev{0, 0};
    if (!J.is_baseline) prev = load_digest(J.tree + leaf);
    if (Kind == HASH_MD5) {
      d = Aligned16 ? md5_range<true>(p, nb) : md5_range<false>(p, nb);
    } else {
      if (!Aligned16) d = murmur3_bytes(p, nb);
      else if (Pipe == LEAF_DEEP) d = murmur3_aligned_deep(p, nb);
      else if (Pipe == LEAF_SHORT && nb == cs) d = murmur3_short(p, (uint32_t)(cs / 16));
      else d = murmur3_aligned(p, nb);  // direct, or the partial last chunk
    }
    uint8_t lab;
    if (digest_eq(d, prev)) {
      // A fixed leaf of an incremental checkpoint: its digest is already in
      // the map (every leaf digest of the previous checkpoint was inserted),
      // so the insert the serial reference performs is a no-op.  In a tree
      // baseline the previous digest is all-zero and the insert still
      // happens; the list method inserts only chunks whose digest differs
      // from the list's (src/list_approach.cpp:50-60).
      const bool ins = J.is_baseline && !J.list_mode;
      lab = ins ? (uint8_t)(FIXED_DUPL | PENDING) : (uint8_t)FIXED_DUPL;
      pending = ins;
    } else {
      lab = PENDING;
      pending = true;
    }
    if (pending) store_digest(J.tree + leaf, d);
    if (pending || J.write_all_labels) J.labels[leaf] = lab;
  }
  record_pending(J, k, pending, d);
}

// Shared tail of both leaf kernels: compare with the previous digest, record
// the label, write the new digest when it changed.
__device__ __forceinline__ bool leaf_finish(const JobG& J, uint64_t k, const Digest& d, const Digest& prev) {
  const uint64_t leaf = J.C - 1 + k;
  bool pending;
  uint8_t lab;
  if (digest_eq(d, prev)) {
    const bool ins = J.is_baseline && !J.list_mode;
    lab = ins ? (uint8_t)(FIXED_DUPL | PENDING) : (uint8_t)FIXED_DUPL;
    pending = ins;
  } else {
    lab = PENDING;
    pending = true;
  }
  if (pending) store_digest(J.tree + leaf, d);
  if (pending || J.write_all_labels) J.labels[leaf] = lab;
  return pending;
}

// leaf_hash, LDS-staged: each wavefront owns 64 consecutive chunks.  Per
// round it loads SEG bytes of every chunk with contiguous 16-byte loads
// (1 KiB per instruction when SEG >= 128), stores them to LDS with a 16-byte
// pad per chunk (conflict-free 128-bit reads), and each lane then hashes the
// next SEG bytes of its own chunk out of LDS.  Needs chunk % SEG == 0 and a
// 16-byte aligned buffer; the partial last chunk is hashed from global memory.
// PF: the loads of round r+1 go to registers before round r is hashed, and
// are stored to LDS at the start of the next round (long chunks: a wavefront
// keeps a round of loads in flight while it hashes).  Copy: the variant for
// tables with a data-section copy (JobT::write_data); it holds more
// registers, so the other tables launch the one without.
template <int Kind, int SEG, bool PF, bool Copy>
__global__ __launch_bounds__(256) void k_leaf_hash_lds(const JobG* __restrict__ jobs, const uint32_t* __restrict__ prefix, uint32_t njobs) {
  constexpr int PITCH = SEG + 16;
  constexpr int VPC = SEG / 16;        // 16-byte vectors per chunk per round
  constexpr int LOADS = VPC;           // 64 chunks x VPC vectors spread over 64 lanes
  __shared__ __attribute__((aligned(16))) uint8_t lds[4][64 * PITCH];
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobG& J = jobs[j];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint64_t k0 = ((uint64_t)(blockIdx.x - prefix[j]) * 4 + wave) * 64;  // first chunk of this wave
  const uint64_t k = k0 + lane;
  const uint64_t cs = J.chunk_size;
  const uint64_t full = J.nfull;  // chunks that are complete
  uint8_t* L = lds[wave];
  const GlobalPtr<const uint8_t> base = J.data + k0 * cs;
  const uint32_t rounds = (uint32_t)(cs / SEG);
  const bool wr = Copy && J.write_data != 0;
  const TreeGeom g = make_geom(J.C);
  Digest prev{0, 0};  // loaded first: its latency hides behind the chunk loads
  if (k < J.C && !J.is_baseline) prev = load_digest(J.tree + (J.C - 1 + k));
  Murmur m;
  uint32_t st[4];
  md5_init(st);
  // Cooperative load of round r: vector v covers chunk v / VPC, piece v % VPC
  // (chunks past the complete ones load zeros and are never hashed).
  uint4 nx[PF ? LOADS : 1];
  auto load_round = [&](uint32_t r) {
#pragma unroll
    for (int i = 0; i < LOADS; i++) {
      const uint32_t v = i * 64 + lane;
      const uint32_t c = v / VPC, piece = v % VPC;
      nx[i] = k0 + c < full ? gload16(base + (uint64_t)c * cs + (uint64_t)r * SEG + piece * 16) : make_uint4(0, 0, 0, 0);
    }
  };
  // A baseline's data-section copy of what the wavefront loaded: chunk c's
  // piece goes where the gather would put it (contiguous 16-byte stores).
  auto copy_out = [&](uint32_t c, uint32_t piece, uint32_t r, const uint4& x) {
    store_piece<true>(root_data_dst(J, g, k0 + c) + (uint64_t)r * SEG + piece * 16, x);
  };
  if (PF && rounds > 0) load_round(0);
  for (uint32_t r = 0; r < rounds; r++) {
    if (PF) {
#pragma unroll
      for (int i = 0; i < LOADS; i++) {
        const uint32_t v = i * 64 + lane;
        *reinterpret_cast<uint4*>(L + (v / VPC) * PITCH + (v % VPC) * 16) = nx[i];
        if (wr && k0 + v / VPC < full) copy_out(v / VPC, v % VPC, r, nx[i]);
      }
      __syncthreads();
      if (r + 1 < rounds) load_round(r + 1);
    } else {
#pragma unroll
      for (int i = 0; i < LOADS; i++) {
        const uint32_t v = i * 64 + lane;
        const uint32_t c = v / VPC, piece = v % VPC;
        if (k0 + c < full) {
          const uint4 x = gload16(base + (uint64_t)c * cs + (uint64_t)r * SEG + piece * 16);
          *reinterpret_cast<uint4*>(L + c * PITCH + piece * 16) = x;
          if (wr) copy_out(c, piece, r, x);
        }
      }
      __syncthreads();
    }
    if (k < full) {
      const uint8_t* my = L + lane * PITCH;
      if (Kind == HASH_MD5) {
#pragma unroll
        for (int b = 0; b < SEG / 64; b++) {
          uint32_t X[16];
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const uint4 w = *reinterpret_cast<const uint4*>(my + 64 * b + 16 * q);
            X[4 * q] = w.x; X[4 * q + 1] = w.y; X[4 * q + 2] = w.z; X[4 * q + 3] = w.w;
          }
          md5_compress(st, X);
        }
      } else {
#pragma unroll
        for (int b = 0; b < VPC; b++) {
          const uint4 w = *reinterpret_cast<const uint4*>(my + 16 * b);
          m.block(w.x | ((uint64_t)w.y << 32), w.z | ((uint64_t)w.w << 32));
        }
      }
    }
    __syncthreads();
  }
  bool pending = false;
  Digest d{0, 0};
  if (k < full) {
    if (Kind == HASH_MD5) {
      uint32_t X[16] = {0x80u, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
      const uint64_t bits = cs * 8;
      X[14] = (uint32_t)bits;
      X[15] = (uint32_t)(bits >> 32);
      md5_compress(st, X);
      d = md5_digest(st);
    } else {
      d = m.finish(cs);
    }
    pending = leaf_finish(J, k, d, prev);
  } else if (k < J.C) {  // the partial last chunk
    const uint64_t off = k * cs;
    const uint8_t* p = (const uint8_t*)(J.data + off);
    const uint64_t nb = J.len - off;
    d = (Kind == HASH_MD5) ? md5_range<true>(p, nb) : murmur3_aligned(p, nb);
    pending = leaf_finish(J, k, d, prev);
  }
  if (wr && full < J.C && k0 <= full && full - k0 < 64) copy_partial_chunk(J, g, lane);
  record_pending(J, k, pending, d);
}

// leaf_hash, direct-to-LDS: each wavefront owns 64 consecutive chunks and an
// 8 KiB LDS region.  Per 128-byte segment of its chunks it issues eight
// global_load_lds_dwordx4 (each one contiguous 1 KiB, so every cache line is
// requested once; the lane-per-chunk loads above request each line once per
// 16-byte piece) that write LDS without passing through registers, waits for
// them, and each lane reads its own segment back and hashes it.  The source
// address of each lane is swizzled so that the eight 16-byte reads of a lane
// go to distinct LDS banks within every group of 16 lanes.  The loads of the
// next segment are issued before the current one is hashed.  No block
// barrier: the region is private to the wavefront.  Needs chunk % 128 == 0
// and a 16-byte aligned buffer; the partial last chunk is hashed from global
// memory.
typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) const void global_void_t;
constexpr int kWaitVm0 = 0x0F70;    // s_waitcnt vmcnt(0)
constexpr int kWaitLgkm0 = 0xC07F;  // s_waitcnt lgkmcnt(0)

// LDS slot (16-byte unit) of piece p of chunk j (0..7) in 1 KiB block i of a
// wavefront's region: each block's slots are a permutation of 0..63, and for
// every p the chunks of an even and an odd block fill 16 distinct slots mod 16.
__device__ __forceinline__ uint32_t dma_slot(uint32_t i, uint32_t j, uint32_t p) {
  return ((i & 1) ? 8 * (p ^ 1) : 8 * p) + ((j + p) & 7);
}

// WPG wavefronts per workgroup (4: 256 chunks; 1: 64 chunks).  A wavefront
// has one 8 KiB round of loads in flight in LDS, so the LDS of a CU (160 KB)
// bounds the bytes in flight; with long chunks (many rounds per wavefront,
// few wavefronts) one-wavefront workgroups let 20 rather than 16 wavefronts
// share a CU.  Copy: as in k_leaf_hash_lds.
template <int Kind, int WPG, bool Copy>
__global__ __launch_bounds__(64 * WPG) void k_leaf_hash_dma(const JobG* __restrict__ jobs, const uint32_t* __restrict__ prefix, uint32_t njobs) {
  __shared__ __attribute__((aligned(16))) uint8_t lds[WPG][8192];
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobG& J = jobs[j];
  const uint32_t wave = WPG == 1 ? 0 : threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint64_t k0 = ((uint64_t)(blockIdx.x - prefix[j]) * WPG + wave) * 64;  // first chunk of this wave
  const uint64_t k = k0 + lane;
  const uint64_t cs = J.chunk_size;
  const uint64_t full = J.nfull;  // chunks that are complete
  Digest prev{0, 0};  // loaded first: its latency hides behind the chunk loads
  if (k < J.C && !J.is_baseline) prev = load_digest(J.tree + (J.C - 1 + k));
  uint8_t* L = lds[wave];
  // This lane's load of block i fills slot `lane` of that block: piece p of
  // chunk jj with dma_slot(i, jj, p) == lane.  Chunks past the complete ones
  // load from the wave's first chunk instead (their slots are never read).
  const uint32_t pe = lane >> 3, po = pe ^ 1;
  const uint32_t je = ((lane & 7) - pe) & 7, jo = ((lane & 7) - po) & 7;
  GlobalPtr<const uint8_t> src[8];
#pragma unroll
  for (uint32_t i = 0; i < 8; i++) {
    const uint32_t p = (i & 1) ? po : pe, jj = (i & 1) ? jo : je;
    const uint64_t c = k0 + i * 8 + jj;
    src[i] = J.data + (c < full ? c : k0) * cs + p * 16;
  }
  const bool have = k0 < full;  // uniform: some chunk of this wave is complete
  const uint32_t rounds = (uint32_t)(cs >> 7);  // chunk_size is a multiple of 128
  // A baseline's data-section copy: slot `lane` of block i is piece p of
  // chunk c (as loaded above), stored from there to the chunk's place, so
  // each store instruction covers 128 contiguous bytes of 8 chunks.
  const bool wr = Copy && J.write_data != 0;
  const TreeGeom g = make_geom(J.C);
  const uint32_t bi = lane >> 3, bj = lane & 7;  // this lane's chunk: block bi, position bj
  Murmur m;
  uint32_t st[4];
  md5_init(st);
  if (have) {
#pragma unroll
    for (uint32_t i = 0; i < 8; i++)
      __builtin_amdgcn_global_load_lds((global_void_t*)src[i], (lds_void_t*)(L + i * 1024), 16, 0, 0);
  }
  for (uint32_t r = 0; have && r < rounds; r++) {
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    uint4 v[8];
#pragma unroll
    for (uint32_t p = 0; p < 8; p++) v[p] = *reinterpret_cast<const uint4*>(L + bi * 1024 + dma_slot(bi, bj, p) * 16);
    if (wr) {
      // One slot at a time (read, store): the compiler barrier keeps the
      // reads from being hoisted together, which would hold 32 more
      // registers and cost the kernel a wavefront per SIMD.
#pragma unroll
      for (uint32_t i = 0; i < 8; i++) {
        const uint4 w = *reinterpret_cast<const uint4*>(L + i * 1024 + lane * 16);
        const uint32_t p = (i & 1) ? po : pe, jj = (i & 1) ? jo : je;
        const uint64_t c = k0 + i * 8 + jj;
        if (c < full) store_piece<true>(root_data_dst(J, g, c) + (uint64_t)r * 128 + p * 16, w);
        __asm__ volatile("" ::: "memory");
      }
    }
    __builtin_amdgcn_s_waitcnt(kWaitLgkm0);  // the reads are done before the region is refilled
    if (r + 1 < rounds) {
#pragma unroll
      for (uint32_t i = 0; i < 8; i++)
        __builtin_amdgcn_global_load_lds((global_void_t*)(src[i] + (uint64_t)(r + 1) * 128), (lds_void_t*)(L + i * 1024),
                                         16, 0, 0);
    }
    if (k < full) {
      if (Kind == HASH_MD5) {
#pragma unroll
        for (int b = 0; b < 2; b++) {
          const uint32_t X[16] = {v[4 * b].x,     v[4 * b].y,     v[4 * b].z,     v[4 * b].w,
                                  v[4 * b + 1].x, v[4 * b + 1].y, v[4 * b + 1].z, v[4 * b + 1].w,
                                  v[4 * b + 2].x, v[4 * b + 2].y, v[4 * b + 2].z, v[4 * b + 2].w,
                                  v[4 * b + 3].x, v[4 * b + 3].y, v[4 * b + 3].z, v[4 * b + 3].w};
          md5_compress(st, X);
        }
      } else {
#pragma unroll
        for (int p = 0; p < 8; p++) m.block(v[p].x | ((uint64_t)v[p].y << 32), v[p].z | ((uint64_t)v[p].w << 32));
      }
    }
  }
  bool pending = false;
  Digest d{0, 0};
  if (k < full) {
    if (Kind == HASH_MD5) {
      uint32_t X[16] = {0x80u, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
      const uint64_t bits = cs * 8;
      X[14] = (uint32_t)bits;
      X[15] = (uint32_t)(bits >> 32);
      md5_compress(st, X);
      d = md5_digest(st);
    } else {
      d = m.finish(cs);
    }
    pending = leaf_finish(J, k, d, prev);
  } else if (k < J.C) {  // the partial last chunk
    const uint64_t off = k * cs;
    const uint8_t* p = (const uint8_t*)(J.data + off);
    const uint64_t nb = J.len - off;
    d = (Kind == HASH_MD5) ? md5_range<true>(p, nb) : murmur3_aligned(p, nb);
    pending = leaf_finish(J, k, d, prev);
  }
  if (wr && full < J.C && k0 <= full && full - k0 < 64) copy_partial_chunk(J, g, lane);
  record_pending(J, k, pending, d);
}

// ---------------------------------------------------------------------------
// Map insertion shared by leaves and interior nodes.
//
// A slot is claimed by a 64-bit CAS of its value word from EMPTY to
// (ckpt << 32 | node).  The key of an entry claimed during this checkpoint is
// read from the tree array (tree[node]), written by an earlier kernel, so no
// lane ever waits for another lane's key store; the claiming lane stores the
// key into the slot for later kernels.  Entries of the same checkpoint and
// digest then agree on one node: the lowest (leaves, atomicMin) or the
// highest (interior nodes, atomicMax), which is the first one the serial
// reference visits (include/reference_impl.hpp:43-82).
// ---------------------------------------------------------------------------
struct InsertResult {
  bool ok;                  // found or inserted
  bool claimed;             // this call created the entry
  uint64_t slot;
  unsigned long long prev;  // entry value before this call's atomicMin/Max (EMPTY when claimed)
};

// Leaf = true: leaf inserts (lowest node wins); false: interior inserts
// (highest node wins).  During the leaf kernel every entry of this
// checkpoint was claimed by a leaf of the same kernel, so its key is read from
// the tree; during the interior kernel, leaf entries already have their key
// in the slot and only interior entries need the tree.
// The job fields an insert needs, read once by the caller (the compiler would
// reload them from the job table after every atomic).
struct InsertCtx {
  MapG m;
  GlobalPtr<const Digest> tree;
  uint64_t C;
  uint32_t ckpt;
};

__device__ __forceinline__ InsertCtx insert_ctx(const JobG& J) { return InsertCtx{J.map, J.tree, J.C, J.ckpt_id}; }

// Slot s of the map: its key {k0, k1} (one 16-byte load) and its value.
__device__ __forceinline__ void load_slot(const MapG& m, uint64_t s, uint4& key, unsigned long long& val) {
  const GlobalPtr<const Slot> sl = &m.slots[s];
  key = gload16((GlobalPtr<const void>)sl);
  val = sl->val;
}

// Settles an insert at slot s whose value is v (not EMPTY): the entry's key
// comes from the tree when a lane of this kernel may have claimed it (its
// slot key may not be published yet), else from the slot (`slot_key` when it
// was loaded with the value, `have_key`, else loaded here).  Equal keys:
// lowers (leaves) or raises (interior nodes) an entry of this checkpoint and
// reports it found.
template <bool Leaf>
__device__ __forceinline__ bool map_settle(const InsertCtx& J, const Digest& k, uint64_t s, unsigned long long v,
                                           const Digest& slot_key, bool have_key, unsigned long long mine,
                                           InsertResult& r) {
  const uint32_t vt = (uint32_t)(v >> 32), vn = (uint32_t)v;
  const bool this_kernel = vt == J.ckpt && (Leaf || vn < J.C - 1);
  Digest e;
  if (this_kernel) {
    e = load_digest(J.tree + vn);
  } else if (have_key) {
    e = slot_key;
  } else {
    e.k0 = J.m.slots[s].k0;
    e.k1 = J.m.slots[s].k1;
  }
  if (!digest_eq(e, k)) return false;
  // An entry of this checkpoint only moves toward `mine` (down for leaves,
  // up for interior nodes).  When the value seen is already past `mine`, the
  // atomic would leave the entry as it is and return a value past `mine` as
  // well, which the caller treats the same way, so it is skipped: a digest
  // repeated all over a buffer (zero chunks) would otherwise send one atomic
  // per run of it to one slot.
  unsigned long long prev = v;
  if (vt == J.ckpt && (Leaf ? mine < v : mine > v))
    prev = Leaf ? g_min(&J.m.slots[s].val, mine) : g_max(&J.m.slots[s].val, mine);
  r = InsertResult{true, false, s, prev};
  return true;
}

// Claims the EMPTY slot s for `mine`; on success publishes the key.  Returns
// the slot's value before the CAS.
__device__ __forceinline__ unsigned long long map_claim(const MapG& m, uint64_t s, const Digest& k,
                                                        unsigned long long mine) {
  const GlobalPtr<Slot> sl = &m.slots[s];
  const unsigned long long old = g_cas(&sl->val, SLOT_EMPTY, mine);
  if (old == SLOT_EMPTY) {
    sl->k0 = k.k0;
    sl->k1 = k.k1;
  }
  return old;
}

// Linear probing.  The rounds only load: a lane that reaches an EMPTY slot
// stops there, and the lanes of a wavefront that stopped claim their slots
// with one CAS each after the wavefront's probing ends.  A claim inside the
// rounds would hold every lane of the wavefront for the CAS's round trip in
// each round where any lane claims (most rounds, with 64 lanes).  A failed
// CAS returns the value of the lane that won, settled like a loaded one; if
// that is another key, probing resumes after the slot.  Every insert loads
// before it claims, also in a baseline that starts from an empty map: a
// blind CAS at the home slot (the atomic reaches the line without keeping
// it in the cache) was slower whenever the slot was taken, and the baseline
// inserts took 0.58 + 0.76 ms with it against 0.56 + 0.69 ms without.
//
// Round widths.  Narrow (baselines, whose inserts are bound by random-line
// traffic and want many waves in flight): an aligned pair of slots per round
// (the first round the rest of the home slot's pair), so no round reaches
// into a second 128-byte line (a round starting at a line's last slot did, for
// a quarter of the inserts: 0.56 + 0.70 ms for the baseline inserts instead
// of 0.52 + 0.63 ms).
// Wide (incremental checkpoints, whose few inserts are bound by each
// wavefront's longest chain, and each round costs the slowest of its lanes'
// random loads): the rest of the home slot's 128-byte line, then two lines
// (eight slots) per round.  All slots of a round are loaded (keys and values)
// before any is settled.
constexpr uint32_t kLineSlots = 4;  // 32-byte slots per 128-byte line

template <bool Leaf, bool Wide>
__device__ inline InsertResult map_insert(const InsertCtx& J, const Digest& k, uint32_t node) {
  constexpr uint32_t RW = Wide ? 2 * kLineSlots : 2;  // slots per round after the first
  const MapG& m = J.m;
  const unsigned long long mine = pack_node(J.ckpt, node);
  uint64_t s = map_start(m, k);
  InsertResult r;
  const Digest none{0, 0};
  uint64_t probed = 0;  // slots examined (bounds the walk of a full bucket)
  while (probed <= m.bmask) {
    uint64_t cand = ~0ull;  // EMPTY slot to claim
    // Round: slots base+lo .. base+hi-1, wrapping inside the bucket (a
    // bucket holds at least 1024 slots, so a pair or a line never straddles
    // its end).
    uint64_t base = Wide ? (s & ~(uint64_t)(kLineSlots - 1)) : (s & ~1ull);
    uint32_t lo = Wide ? (uint32_t)(s & (kLineSlots - 1)) : (uint32_t)(s & 1);
    uint32_t hi = Wide ? kLineSlots : RW;
    for (; probed <= m.bmask; probed += hi - lo, base = map_next(m, base, hi), lo = 0, hi = RW) {
      uint4 kk[RW];
      unsigned long long vv[RW];
#pragma unroll
      for (uint32_t e = 0; e < RW; e++)
        if (e >= lo && e < hi) load_slot(m, map_next(m, base, e), kk[e], vv[e]);
#pragma unroll
      for (uint32_t e = 0; e < RW; e++) {
        if (e < lo || e >= hi) continue;
        const uint64_t se = map_next(m, base, e);
        if (vv[e] == SLOT_EMPTY) {
          cand = se;
          break;
        }
        const Digest key{kk[e].x | ((uint64_t)kk[e].y << 32), kk[e].z | ((uint64_t)kk[e].w << 32)};
        if (map_settle<Leaf>(J, k, se, vv[e], key, true, mine, r)) return r;
      }
      if (cand != ~0ull) break;
    }
    if (cand == ~0ull) break;
    const unsigned long long v = map_claim(m, cand, k, mine);
    if (v == SLOT_EMPTY) return InsertResult{true, true, cand, SLOT_EMPTY};
    // The loaded key predates the claim: settle with the slot's key.
    if (map_settle<Leaf>(J, k, cand, v, none, false, mine, r)) return r;
    s = map_next(m, cand, 1);
    probed += 1;
  }
  return InsertResult{false, false, 0, 0};
}

// One lane per entry of the pending-leaf list (grid-stride).  Labels settle
// without a later map read: a changed leaf stays PENDING
// (first-occurrence candidate) unless it meets an older entry or a lower leaf
// of this checkpoint (then it labels itself SHIFT), and a leaf that lowers an
// entry demotes the previous lowest leaf to SHIFT.  Every leaf that is ever
// the lowest is demoted by the next lower one, so exactly the serial
// reference's first occurrence stays PENDING (include/reference_impl.hpp:49-67).
// FIXED|PENDING leaves (baseline, zero digest) take part in the ordering but
// keep their label.
//
// Runs of equal digests are compressed within a wavefront first: a lane whose
// left neighbour holds the same digest at a lower leaf is not the first
// occurrence, so only the run's first lane touches the map and the others
// label themselves SHIFT with the first lane's slot.  Long runs of one digest
// (zero-filled regions) would otherwise serialize on one slot's atomics.
// Insert mode of a job (one of JOBS_DENSE, JOBS_LIST_NARROW, JOBS_LIST_WIDE);
// each insert kernel variant takes the jobs of its mode, so a variant holds
// only the code (and registers) of its own path.
__device__ __forceinline__ uint32_t insert_mode(const JobG& J) {
  return J.is_baseline ? JOBS_DENSE : J.wide_probe ? JOBS_LIST_WIDE : JOBS_LIST_NARROW;
}

template <uint32_t Mode>
__global__ __launch_bounds__(256) void k_insert_leaves(const JobG* __restrict__ jobs, const uint32_t* __restrict__ prefix, uint32_t njobs) {
  constexpr bool Wide = Mode == JOBS_LIST_WIDE;
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobG& J = jobs[j];
  if (insert_mode(J) != Mode) return;  // another variant's job
  const uint32_t wg = blockIdx.x - prefix[j], nwg = prefix[j + 1] - prefix[j];
  uint32_t claims = 0;
  __shared__ uint32_t sh[4];
  __shared__ uint32_t pre[kPendSpread + 1];
  __shared__ uint32_t sflags;
  // A baseline inserts every leaf: lane k takes chunk k directly (the list
  // would hold every chunk too).
  constexpr bool dense = Mode == JOBS_DENSE;
  // The flags, the pending-list counts and (incremental) each lane's first
  // list entry with its digest, all in one round trip.  Other workgroups of
  // this kernel may raise MAP_FULL while it runs: the decision is made once,
  // uniform over the workgroup.
  if (threadIdx.x == 0) sflags = J.ctr->flags;
  ListWalk L{};
  uint32_t spec_chunk = 0;
  Digest spec_dig{0, 0};
  if (!dense) {
    L = list_walk(wg, nwg, blockDim.x);
    const uint32_t j0 = L.j0 + threadIdx.x;
    if (j0 < J.pend_cap) spec_chunk = J.pend_list[(uint64_t)L.r * J.pend_cap + j0];
    if (j0 < J.dig_cap) spec_dig = load_digest(J.pend_dig + (uint64_t)L.r * J.dig_cap + j0);
    load_list_prefix(J.pend_cnt, pre);
  } else {
    __syncthreads();
  }
  // A bulk job's map is built by k_bulk_build (such a job has no workgroups
  // here; a fallback runs the job again without it).
  if ((sflags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) || (dense && J.bulk)) {
    if (threadIdx.x == 0) J.blk_claim[wg] = 0;
    return;
  }
  // Worst case of this checkpoint: one entry per changed leaf plus one per
  // first-occurrence interior node (fewer than the changed leaves).  A
  // baseline keeps no list and starts from an empty map.
  const uint32_t nitems = dense ? (uint32_t)J.C : pre[kPendSpread];
  if (J.map_count_at_start + 2ull * (dense ? 0u : nitems) > J.map_limit) {
    if (threadIdx.x == 0) {
      g_or(&J.ctr->flags, FLAG_NEED_GROW);
      J.blk_claim[wg] = 0;
    }
    return;
  }
  const uint32_t K = J.ckpt_id;
  const InsertCtx ic = insert_ctx(J);
  const GlobalPtr<uint8_t> labels = J.labels;
  const GlobalPtr<uint64_t> slot_of = J.slot_of;
  const GlobalPtr<uint32_t> flags = &J.ctr->flags;
  const bool by_slot = J.low_root != 0;  // low-root kernels rewrite entries through their slots
  const uint32_t lane = lane_id();
  // One item per lane: leaf `leaf` with label `own` and digest `k`.  Called
  // by every lane of a wavefront together (uniform loop bounds), since it
  // shuffles; lanes past the list are inactive.
  auto insert_item = [&](bool act, uint64_t leaf, uint8_t own, const Digest& k) {
    const unsigned long long pk0 = __shfl_up(k.k0, 1, 64), pk1 = __shfl_up(k.k1, 1, 64);
    const unsigned long long pleaf = __shfl_up((unsigned long long)leaf, 1, 64);
    const int pact = __shfl_up((int)act, 1, 64);
    const bool head = !(act && lane > 0 && pact && pk0 == k.k0 && pk1 == k.k1 && pleaf < leaf);
    bool ok = true;
    uint64_t ref = 0;
    if (act && head) {
      const InsertResult r = map_insert<true, Wide>(ic, k, (uint32_t)leaf);
      ok = r.ok;
      ref = kRefSlot | r.slot;
      if (!r.ok) {
        g_or(flags, FLAG_MAP_FULL);
      } else if (r.claimed) {
        claims++;
      } else {
        const uint32_t pt = (uint32_t)(r.prev >> 32), pn = (uint32_t)r.prev;
        // An entry of an earlier checkpoint is final: the shifted-root
        // lookup then needs no map read.
        if (pt != K && !by_slot && !ref_is_slot(r.prev)) ref = r.prev;
        bool self_shift = true;
        if (pt == K) {
          if (pn > leaf) {
            self_shift = false;  // this leaf is now the lowest; the previous lowest is not
            // Outside a baseline the previous lowest leaf (a changed leaf of
            // this checkpoint) is PENDING or already SHIFT, so it is demoted
            // without reading it; a baseline's may be FIXED|PENDING, which
            // stays.
            if (!dense) labels[pn] = SHIFT_DUPL;
            else if (labels[pn] == PENDING) labels[pn] = SHIFT_DUPL;
          } else if (pn == leaf) {
            self_shift = false;  // re-run after a map growth: own entry
          }
        }
        if (self_shift && own == PENDING) labels[leaf] = SHIFT_DUPL;
      }
    }
    // Each lane's run head: the highest head lane at or below it.
    const unsigned long long heads = __ballot(head);
    const uint32_t hl = 63 - __clzll(heads & ((2ull << lane) - 1ull));
    const unsigned long long href = __shfl((unsigned long long)ref, hl, 64);
    const int hok = __shfl((int)ok, hl, 64);
    if (act && !head && hok) {
      slot_of[leaf] = href;
      if (own == PENDING) labels[leaf] = SHIFT_DUPL;  // a lower leaf of this run holds the digest
    } else if (act && head && ok) {
      slot_of[leaf] = ref;
    }
  };
  if (dense) {
    for (uint32_t base = wg * blockDim.x; base < nitems; base += nwg * blockDim.x) {
      const uint32_t it = base + threadIdx.x;
      bool act = it < nitems;
      const uint64_t leaf = ic.C - 1 + it;
      uint8_t own = 0;
      Digest k{0, 0};
      if (act) {
        own = labels[leaf];
        // A list baseline leaves all-zero digests out (they equal the fresh
        // list's entry); every other leaf of a dense pass needs its insert.
        if (own == FIXED_DUPL) act = false;
        else k = load_digest(ic.tree + leaf);
      }
      insert_item(act, leaf, own, k);
    }
  } else {
    // Every pending leaf of an incremental checkpoint was labelled PENDING
    // by the leaf kernel (a re-run after a map growth may find it SHIFT
    // already, and writing SHIFT again changes nothing).
    bool first = true;
    for (uint32_t r = L.r; r < kPendSpread; r += L.r_step) {
      const uint32_t cnt = pre[r + 1] - pre[r];
      for (uint32_t jb = L.j0; jb < cnt; jb += L.j_step, first = false) {
        const uint32_t jj = jb + threadIdx.x;
        const bool act = jj < cnt;
        uint64_t leaf = 0;
        Digest k{0, 0};
        if (act) {
          const uint64_t e = (uint64_t)r * J.pend_cap + jj;
          leaf = ic.C - 1 + (first ? spec_chunk : J.pend_list[e]);
          if (jj >= J.dig_cap) k = load_digest(ic.tree + leaf);
          else k = first ? spec_dig : load_digest(J.pend_dig + (uint64_t)r * J.dig_cap + jj);
        }
        insert_item(act, leaf, PENDING, k);
      }
      first = false;
    }
  }
  const uint32_t n = block_sum<uint32_t>(claims, sh);
  if (threadIdx.x == 0) J.blk_claim[wg] = n;
}

// ---------------------------------------------------------------------------
// Low-root first occurrences (TreeLowRootDeduplicator; dedup_low_root_ref,
// include/reference_impl.hpp:265-472, restated as a deterministic rule).
// After the leaf inserts, whose first occurrences follow the low-offset rule,
// each new digest's first occurrence moves to the leaf whose highest new
// ancestor has the smallest node id, ties to the lowest leaf.  New: a changed
// leaf whose digest has no entry from an earlier checkpoint (PENDING, or
// SHIFT with an entry of this checkpoint); an interior node whose children
// are both new.  The selection key (top << 32 | leaf) is lowered with an
// atomic in the slot's spare word.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool lr_skip(const JobG& J) {
  return !J.low_root || (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL));
}

// Leaves: new flags; each group's low-offset first occurrence (the one
// PENDING leaf of its digest) resets the group's key to ~0.
__global__ __launch_bounds__(256) void k_lr_leaves(const JobG* __restrict__ jobs) {
  const JobG& J = jobs[blockIdx.y];
  if (lr_skip(J)) return;
  for (uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x; k < J.C; k += (uint64_t)gridDim.x * 256) {
    const uint64_t leaf = J.C - 1 + k;
    const uint8_t lab = J.labels[leaf];
    uint8_t nw = 0;
    if (lab == PENDING) {
      nw = 1;
      *(GlobalPtr<uint64_t>)&J.map.slots[(uint32_t)J.slot_of[leaf]].pad = ~0ull;
    } else if (lab == SHIFT_DUPL) {
      nw = (uint32_t)(J.map.slots[(uint32_t)J.slot_of[leaf]].val >> 32) == J.ckpt_id ? 1 : 0;
    }
    J.lr_new[leaf] = nw;
  }
}

// Interior nodes of level d (all of them, deepest level first).
__global__ __launch_bounds__(256) void k_lr_level(const JobG* __restrict__ jobs, uint32_t d) {
  const JobG& J = jobs[blockIdx.y];
  if (lr_skip(J) || J.C < 2) return;
  const uint64_t lo = (1ull << d) - 1;
  const uint64_t hi = min((2ull << d) - 1, J.C - 1);  // interior nodes only
  for (uint64_t v = lo + (uint64_t)blockIdx.x * 256 + threadIdx.x; v < hi; v += (uint64_t)gridDim.x * 256)
    J.lr_new[v] = J.lr_new[2 * v + 1] & J.lr_new[2 * v + 2];
}

// Each new leaf lowers its group's key with (highest new ancestor, leaf).
__global__ __launch_bounds__(256) void k_lr_key(const JobG* __restrict__ jobs) {
  const JobG& J = jobs[blockIdx.y];
  if (lr_skip(J)) return;
  for (uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x; k < J.C; k += (uint64_t)gridDim.x * 256) {
    const uint64_t leaf = J.C - 1 + k;
    if (!J.lr_new[leaf]) continue;
    uint64_t t = leaf;
    while (t > 0 && J.lr_new[(t - 1) >> 1]) t = (t - 1) >> 1;
    const unsigned long long key = ((unsigned long long)t << 32) | (uint32_t)leaf;
    const GlobalPtr<unsigned long long> pad = (GlobalPtr<unsigned long long>)&J.map.slots[(uint32_t)J.slot_of[leaf]].pad;
    if (key < *pad) g_min(pad, key);  // a key already below cannot change
  }
}

// The group's winner becomes its first occurrence (and the map entry's
// node); the other new leaves of the group are shifted duplicates of it.
__global__ __launch_bounds__(256) void k_lr_select(const JobG* __restrict__ jobs) {
  const JobG& J = jobs[blockIdx.y];
  if (lr_skip(J)) return;
  for (uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x; k < J.C; k += (uint64_t)gridDim.x * 256) {
    const uint64_t leaf = J.C - 1 + k;
    if (!J.lr_new[leaf]) continue;
    const GlobalPtr<Slot> sl = &J.map.slots[(uint32_t)J.slot_of[leaf]];
    const unsigned long long key = *(GlobalPtr<unsigned long long>)&sl->pad;
    if ((uint32_t)key == (uint32_t)leaf) {
      J.labels[leaf] = PENDING;
      *(GlobalPtr<unsigned long long>)&sl->val = pack_node(J.ckpt_id, (uint32_t)leaf);
    } else {
      J.labels[leaf] = SHIFT_DUPL;
    }
  }
}

// ---------------------------------------------------------------------------
// List method (src/list_approach.cpp:23-62): a changed chunk is a first
// occurrence when its insert created the entry or lowered it to this chunk
// (the serial loop's first chunk of a digest), and a shifted duplicate
// otherwise.  insert_leaves leaves exactly those first occurrences PENDING.
// Interior labels are DONE from the baseline on, so the root scan emits
// every FIRST or SHIFT leaf.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_list_resolve(const JobG* __restrict__ jobs, const uint32_t* __restrict__ prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobG& J = jobs[j];
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;  // written before this kernel: uniform
  const uint32_t wg = blockIdx.x - prefix[j], nwg = prefix[j + 1] - prefix[j];
  const uint64_t C = J.C;
  const GlobalPtr<uint8_t> labels = J.labels;
  if (J.is_baseline) {
    for (uint64_t i = (uint64_t)wg * blockDim.x + threadIdx.x; i < C; i += (uint64_t)nwg * blockDim.x) {
      if (labels[C - 1 + i] == PENDING) labels[C - 1 + i] = FIRST_OCUR;
      if (i + 1 < C) labels[i] = DONE;
    }
    return;
  }
  __shared__ uint32_t pre[kPendSpread + 1];
  load_list_prefix(J.pend_cnt, pre);
  const uint32_t n = pre[kPendSpread];
  for (uint32_t it = wg * blockDim.x + threadIdx.x; it < n; it += nwg * blockDim.x) {
    const uint64_t leaf = C - 1 + list_item(J.pend_list, J.pend_cap, pre, it);
    if (labels[leaf] == PENDING) labels[leaf] = FIRST_OCUR;
  }
}

__device__ __forceinline__ int tile_groups(uint32_t D) { return D == 0 ? 1 : (int)((D + kTileLevels - 1) / kTileLevels); }

// Tile geometry shared by the tile kernels: global tile index T -> group,
// bottom/top levels and root node.
struct TileRef {
  int g, Lb, Lr, H;
  uint64_t r;
};

__device__ __forceinline__ TileRef tile_ref(const JobG& J, int g, uint64_t t) {
  TileRef tr;
  tr.g = g;
  tr.Lb = (int)J.D - kTileLevels * g;
  tr.Lr = tr.Lb > kTileLevels ? tr.Lb - kTileLevels : 0;
  tr.H = tr.Lb - tr.Lr;
  tr.r = ((1ull << tr.Lr) - 1) + t;
  return tr;
}

// Interior node i (level-major, 0 = tile root) of a tile with root r.
__device__ __forceinline__ uint64_t tile_node(uint64_t r, uint32_t i) {
  const uint32_t q = 31 - __clz(i + 1);
  return ((r + 1) << q) - 1 + (i + 1 - (1u << q));
}

// Map inserts of the first-occurrence interior nodes that pass 1 listed: one
// lane per list entry (grid-stride).
template <uint32_t Mode>
__global__ __launch_bounds__(256) void k_insert_interior(const JobG* __restrict__ jobs, const uint32_t* __restrict__ prefix, uint32_t njobs) {
  constexpr bool Wide = Mode == JOBS_LIST_WIDE;
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobG& J = jobs[j];
  if (insert_mode(J) != Mode) return;  // another variant's job
  const uint32_t wg = blockIdx.x - prefix[j], nwg = prefix[j + 1] - prefix[j];
  __shared__ uint32_t sh[4];
  __shared__ uint32_t pre[kPendSpread + 1];
  __shared__ uint32_t sflags;
  uint32_t claims = 0;
  // The flags and (incremental) the dirty-list counts in one round trip.
  if (threadIdx.x == 0) sflags = J.ctr->flags;
  if (Mode != JOBS_DENSE) load_list_prefix(J.dirty_cnt, pre);
  else __syncthreads();
  // A bulk job's map is built by k_bulk_build.
  const bool bulk_built = Mode == JOBS_DENSE && J.bulk;
  if (!(sflags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) && !bulk_built) {
    const InsertCtx ic = insert_ctx(J);
    const GlobalPtr<uint32_t> flags = &J.ctr->flags;
    if (Mode == JOBS_DENSE) {
      // A baseline visits every interior node 0 .. C-2 by label (most of
      // them are first occurrences).
      const GlobalPtr<const uint8_t> labels = J.labels;
      for (uint64_t x = (uint64_t)wg * blockDim.x + threadIdx.x; x + 1 < J.C; x += (uint64_t)nwg * blockDim.x) {
        if (labels[x] != FIRST_OCUR) continue;
        const InsertResult r = map_insert<false, Wide>(ic, load_digest(ic.tree + x), (uint32_t)x);
        if (!r.ok) g_or(flags, FLAG_MAP_FULL);
        claims += r.claimed ? 1 : 0;
      }
    } else {
      // An incremental checkpoint walks the tiles pass 1 visited (the dirty
      // group-0 tiles, then every tile of the groups above), one wavefront
      // per tile, one lane per entry of the tile's fint region.  The count
      // and the region's first 64 entries are loaded together (the region
      // holds kTileInterior >= 64 entries, so the loads stay in it).
      const uint32_t ndirty = pre[kPendSpread];
      const uint32_t ntiles0 = J.tile_base[1];  // group-0 tiles (tile_base[1] = ntiles without a group 1)
      const uint32_t nitems = ndirty + (J.ntiles - ntiles0);
      const uint32_t lane = lane_id(), nw = blockDim.x >> 6;
      for (uint32_t it = wg * nw + (threadIdx.x >> 6); it < nitems; it += nwg * nw) {
        const uint32_t T = it < ndirty ? list_item(J.dirty_list, J.dirty_cap, pre, it) : ntiles0 + (it - ndirty);
        const GlobalPtr<const uint32_t> region = J.fint_list + (uint64_t)T * kTileInterior;
        const uint32_t cnt = J.fint_n[T];
        uint32_t x0 = region[lane];
        for (uint32_t q = lane; q < cnt; q += 64) {
          const uint32_t x = q < 64 ? x0 : region[q];
          const InsertResult r = map_insert<false, Wide>(ic, load_digest(ic.tree + x), x);
          if (!r.ok) g_or(flags, FLAG_MAP_FULL);
          claims += r.claimed ? 1 : 0;
        }
      }
    }
  }
  const uint32_t n = block_sum<uint32_t>(claims, sh);
  if (threadIdx.x == 0) J.int_claim[wg] = n;
}

// ---------------------------------------------------------------------------
// Bulk map build of a baseline.  A baseline starts from an empty map, so
// instead of a clear of the whole map and one random read-modify-write of a
// 128-byte line per insert, its entries are partitioned by map bucket and
// every bucket is built in LDS and written once, its slots in order.  Per
// part (the leaves after the leaf kernel; the first-occurrence interior
// nodes after pass 1):
// - k_bulk_hist: each partition workgroup counts its tile's records per
//   bucket in LDS and writes its row of the (tile x bucket) count matrix;
// - k_bulk_colscan: per bucket, exclusive prefixes down the tiles (in place),
//   the bucket's total, and totals per group of kColBuckets buckets;
// - k_bulk_scatter: each partition workgroup turns its row into LDS cursors
//   (bucket start + its prefix) and writes every record (digest, node) at a
//   cursor bumped with an LDS atomic.  No global atomics: device-scope
//   atomics from the eight XCDs meet at memory, and a count with one per
//   record took longer than the random-line inserts it replaces.
// - k_bulk_build, one workgroup per bucket: the bucket's records go into LDS
//   and are inserted into an LDS image of the bucket with the map's own
//   probing (home slot, then the next slots, wrapping inside the bucket).
//   Equal digests share a slot, whose entry is the one the dense inserts
//   keep: the lowest leaf, and among interior nodes the highest
//   (include/reference_impl.hpp:49-82).  The leaf part labels every leaf of
//   a digest but the lowest SHIFT (pass 1 needs the leaf labels); the final
//   part writes the bucket's slots and the map slot of each such leaf.
// A bucket with more leaf records than a workgroup holds (a digest repeated
// thousands of times, as zero chunks are; FLAG_BULK_OFF) or a final bucket
// that does not fit stops the job (FLAG_MAP_FULL), and the host runs it
// again with the dense inserts.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t bulk_bucket(const MapG& m, const Digest& d) {
  return (uint32_t)(map_start(m, d) >> kBucketLog);
}

__device__ __forceinline__ bool bulk_skip(const JobG& J, uint32_t flags) {
  return !J.bulk || (flags & (FLAG_BULK_OFF | FLAG_NEED_GROW | FLAG_MAP_FULL | FLAG_BULK_DUPS));
}

// Batch `b` of a partition lane: items w * kPartTile + (b * kPartBatch + k) *
// kPartThreads + lane (k < kPartBatch) of the part, consecutive lanes on
// consecutive items.  Bit k of the result: item k is a record (every leaf;
// the first-occurrence interior nodes), with its digest in d[k].  The labels
// and then the digests are loaded together.
constexpr uint32_t kPartBatch = 8;
template <bool Int>
__device__ __forceinline__ uint32_t bulk_load(const JobG& J, uint32_t w, uint32_t b, Digest d[kPartBatch],
                                              uint64_t item[kPartBatch]) {
  const uint64_t n = Int ? J.C - 1 : J.C;
  const uint64_t node0 = Int ? 0 : J.C - 1;
  uint8_t lab[kPartBatch];
#pragma unroll
  for (uint32_t k = 0; k < kPartBatch; k++) {
    item[k] = (uint64_t)w * kPartTile + (uint64_t)(b * kPartBatch + k) * kPartThreads + threadIdx.x;
    lab[k] = FIRST_OCUR;
    if (Int) lab[k] = J.labels[item[k] < n ? item[k] : 0];
  }
  uint32_t valid = 0;
#pragma unroll
  for (uint32_t k = 0; k < kPartBatch; k++)
    valid |= (item[k] < n && lab[k] == FIRST_OCUR ? 1u : 0u) << k;
#pragma unroll
  for (uint32_t k = 0; k < kPartBatch; k++)
    d[k] = load_digest(J.tree + (((valid >> k) & 1u) ? node0 + item[k] : 0));
  return valid;
}

// Part p's arrays: the count matrix (tiles x buckets), the bucket totals and
// the group totals.
__device__ __forceinline__ GlobalPtr<uint32_t> bulk_mat(const JobG& J, bool Int) {
  return J.bk_mat + (Int ? (uint64_t)J.ntiles_b * J.nbuckets : 0);
}
__device__ __forceinline__ GlobalPtr<uint32_t> bulk_tot(const JobG& J, bool Int) {
  return J.bk_tot + (Int ? J.nbuckets : 0);
}
__device__ __forceinline__ GlobalPtr<uint32_t> bulk_grp(const JobG& J, bool Int) {
  return J.bk_grp + (Int ? (J.nbuckets + kColBuckets - 1) / kColBuckets : 0);
}

// One workgroup per tile: the tile's records per bucket -> its matrix row.
template <bool Int>
__global__ __launch_bounds__(kPartThreads) void k_bulk_hist(const JobG* __restrict__ jobs, const uint32_t* __restrict__ prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobG& J = jobs[j];
  __shared__ uint32_t hist[kBulkMaxBuckets];
  __shared__ uint32_t sflags;
  if (threadIdx.x == 0) sflags = J.ctr->flags;
  const uint32_t nb = J.nbuckets;
  for (uint32_t b = threadIdx.x; b < nb; b += kPartThreads) hist[b] = 0;
  __syncthreads();
  if (bulk_skip(J, sflags)) return;
  const uint32_t w = blockIdx.x - prefix[j];
  for (uint32_t b = 0; b < kPartItems / kPartBatch; b++) {
    Digest d[kPartBatch];
    uint64_t item[kPartBatch];
    const uint32_t valid = bulk_load<Int>(J, w, b, d, item);
#pragma unroll
    for (uint32_t k = 0; k < kPartBatch; k++)
      if ((valid >> k) & 1u) atomicAdd(&hist[bulk_bucket(J.map, d[k])], 1u);
  }
  __syncthreads();
  const GlobalPtr<uint32_t> row = bulk_mat(J, Int) + (uint64_t)w * nb;
  for (uint32_t b = threadIdx.x; b < nb; b += kPartThreads) row[b] = hist[b];
}

// One workgroup per group of kColBuckets buckets; lane (segment s, bucket
// bl) holds rows s*RS .. s*RS+RS-1 of its bucket's column (RS <= 16, all
// loaded together).  The column becomes exclusive prefixes down the tiles;
// the bucket's total and the group's total are written; the leaf part
// switches the job to the dense inserts when a bucket has more than
// bk_cap_l leaf records.
template <bool Int>
__global__ __launch_bounds__(kColBuckets * kColSegs) void k_bulk_colscan(const JobG* __restrict__ jobs, const uint32_t* __restrict__ prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobG& J = jobs[j];
  __shared__ uint32_t seg[kColSegs][kColBuckets];
  __shared__ uint32_t sh[kColSegs + 1];
  __shared__ uint32_t sflags;
  if (threadIdx.x == 0) sflags = J.ctr->flags;
  __syncthreads();
  if (bulk_skip(J, sflags)) return;
  const uint32_t g = blockIdx.x - prefix[j];
  const uint32_t nb = J.nbuckets, W = J.ntiles_b;
  const uint32_t bl = threadIdx.x % kColBuckets, s = threadIdx.x / kColBuckets;
  const uint32_t b = g * kColBuckets + bl;
  const bool inb = b < nb;
  const uint32_t RS = (W + kColSegs - 1) / kColSegs;  // at most 16 (kBulkMaxTiles)
  const GlobalPtr<uint32_t> mat = bulk_mat(J, Int);
  uint32_t v[16];
#pragma unroll
  for (uint32_t r = 0; r < 16; r++) {
    const uint32_t w = s * RS + r;
    v[r] = (inb && r < RS && w < W) ? mat[(uint64_t)w * nb + b] : 0u;
  }
  uint32_t run = 0;
#pragma unroll
  for (uint32_t r = 0; r < 16; r++) {
    const uint32_t x = v[r];
    v[r] = run;
    run += x;
  }
  seg[s][bl] = run;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
  for (uint32_t q = 0; q < kColSegs; q++) {
    const uint32_t x = seg[q][bl];
    pre += q < s ? x : 0u;
    tot += x;
  }
#pragma unroll
  for (uint32_t r = 0; r < 16; r++) {
    const uint32_t w = s * RS + r;
    if (inb && r < RS && w < W) mat[(uint64_t)w * nb + b] = pre + v[r];
  }
  if (s == 0 && inb) {
    bulk_tot(J, Int)[b] = tot;
    if (!Int && tot > J.bk_cap_l) g_or(&J.ctr->flags, FLAG_BULK_OFF | FLAG_MAP_FULL);  // later stages stop
  }
  const uint32_t gt = block_sum<uint32_t>(s == 0 && inb ? tot : 0u, sh);
  if (threadIdx.x == 0) bulk_grp(J, Int)[g] = gt;
}

// First record of bucket b within its part: the totals of the groups before
// its group and of the buckets before it in its group, summed by one
// wavefront (every lane calls it).
__device__ __forceinline__ uint32_t bulk_start(const JobG& J, bool Int, uint32_t b) {
  const uint32_t g = b / kColBuckets;
  const GlobalPtr<const uint32_t> grp = bulk_grp(J, Int), tot = bulk_tot(J, Int);
  uint32_t s = 0;
  for (uint32_t i = lane_id(); i < g; i += 64) s += grp[i];
  const uint32_t i = g * kColBuckets + lane_id();
  if (i < b) s += tot[i];
  return wave_sum(s);
}

// One workgroup per tile, the same items as k_bulk_hist: LDS cursors (bucket
// start + the tile's prefix in the bucket), then one LDS atomic per record.
template <bool Int>
__global__ __launch_bounds__(kPartThreads) void k_bulk_scatter(const JobG* __restrict__ jobs, const uint32_t* __restrict__ prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobG& J = jobs[j];
  __shared__ uint32_t cur[kBulkMaxBuckets];
  __shared__ uint32_t sh[kPartThreads / 64 + 1];
  __shared__ uint32_t sflags;
  if (threadIdx.x == 0) sflags = J.ctr->flags;
  const uint32_t nb = J.nbuckets;
  const uint32_t w = blockIdx.x - prefix[j];
  // Bucket starts: an exclusive scan of the bucket totals, kPerLane buckets
  // per lane, loaded with the tile's row.
  constexpr uint32_t kPerLane = kBulkMaxBuckets / kPartThreads;
  const GlobalPtr<const uint32_t> tot = bulk_tot(J, Int);
  const GlobalPtr<const uint32_t> row = bulk_mat(J, Int) + (uint64_t)w * nb;
  uint32_t t[kPerLane], r[kPerLane];
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t i = 0; i < kPerLane; i++) {
    const uint32_t b = threadIdx.x * kPerLane + i;
    t[i] = b < nb ? tot[b] : 0u;
    r[i] = b < nb ? row[b] : 0u;
  }
#pragma unroll
  for (uint32_t i = 0; i < kPerLane; i++) sum += t[i];
  uint32_t total;
  uint32_t base = block_excl_scan<kPartThreads / 64>(sum, sh, total);
#pragma unroll
  for (uint32_t i = 0; i < kPerLane; i++) {
    const uint32_t b = threadIdx.x * kPerLane + i;
    if (b < nb) cur[b] = base + r[i];
    base += t[i];
  }
  __syncthreads();
  if (bulk_skip(J, sflags)) return;
  const uint64_t roff = Int ? J.C : 0;  // first record of the part
  const uint64_t node0 = Int ? 0 : J.C - 1;
  for (uint32_t b = 0; b < kPartItems / kPartBatch; b++) {
    Digest d[kPartBatch];
    uint64_t item[kPartBatch];
    const uint32_t valid = bulk_load<Int>(J, w, b, d, item);
#pragma unroll
    for (uint32_t k = 0; k < kPartBatch; k++) {
      if ((valid >> k) & 1u) {
        const uint64_t pos = roff + atomicAdd(&cur[bulk_bucket(J.map, d[k])], 1u);
        // One 32-byte record (digest, node): a single sector per record
        // where records land in lines that other workgroups also write.
        const GlobalPtr<u32x4_t> rec = (GlobalPtr<u32x4_t>)(J.bk_rec + 2 * pos);
        rec[0] = u32x4_t{(uint32_t)d[k].k0, (uint32_t)(d[k].k0 >> 32), (uint32_t)d[k].k1, (uint32_t)(d[k].k1 >> 32)};
        rec[1] = u32x4_t{(uint32_t)(node0 + item[k]), 0u, 0u, 0u};
      }
    }
  }
}

// One workgroup per bucket.  Leaf part (Final = false): the bucket's leaf
// records; every leaf whose digest has a lower leaf in the bucket is labelled
// SHIFT (a FIXED|PENDING leaf keeps its label, as in the dense inserts).  Final
// part: the bucket's leaf and interior records; the bucket's slots are
// written in order (EMPTY where no record landed), each shifted leaf gets its
// digest's slot, and the entry count goes to bk_claim.
//
// Probing in LDS: a record claims an empty slot of the LDS image with a CAS
// (the slot then names it, its digest's first record); a slot naming a
// record with the same digest is the digest's slot.  Each record lowers the
// rank kept for its slot's first record with an atomicMin: ranks are unique
// (a leaf's index, below every interior node, which rank by descending id),
// so the record whose rank is the minimum is the entry.
template <bool Final>
__global__ __launch_bounds__(kBuildThreads) void k_bulk_build(const JobG* __restrict__ jobs, const uint32_t* __restrict__ prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobG& J = jobs[j];
  // LDS holds 32 bits of each record's digest (the full digests stay in the
  // registers of the lane that loaded the record, and in the record array
  // for the rare equal keys): about 50 KB, three workgroups per CU.  The
  // probing arrays are reused as the staging area of the slot writes.
  constexpr uint32_t kHalf = kBucketSlots / 2;
  constexpr uint32_t kPool = (kBucketSlots + 3 * kBulkRec) * 4;
  static_assert(kPool >= kHalf * 20, "a half bucket of entries fits in the probing arrays");
  __shared__ __attribute__((aligned(16))) uint8_t pool[kPool];
  uint32_t* const owner = reinterpret_cast<uint32_t*>(pool);  // record that claimed the slot, ~0 = empty
  uint32_t* const best = owner + kBucketSlots;                 // per first record: lowest rank of its digest
  uint32_t* const rkey = best + kBulkRec;                      // digest bits 0..31 of each record
  uint32_t* const rnode = rkey + kBulkRec;
  uint4* const sdig = reinterpret_cast<uint4*>(pool);          // staging: entries of half a bucket
  uint32_t* const snode = reinterpret_cast<uint32_t*>(pool + kHalf * 16);
  __shared__ uint16_t rslot[kBulkRec];
  __shared__ uint32_t occ[kBucketSlots / 32];  // bit q: slot q holds an entry
  __shared__ uint32_t sh[kBuildThreads / 64 + 1];
  __shared__ uint32_t sflags, sbase[2], sn[2];
  const uint32_t b = blockIdx.x - prefix[j];
  const uint64_t bs = J.map.bmask + 1;
  const GlobalPtr<Slot> slots = J.map.slots + (uint64_t)b * bs;
  // The flags and the bucket's record ranges: wavefront p (0 = leaves, 1 =
  // interior nodes) finds its part's range.
  if (threadIdx.x == 0) sflags = J.ctr->flags;
  const uint32_t w = threadIdx.x >> 6;
  if (w < (Final ? 2u : 1u)) {
    const bool Int = w == 1;
    const uint32_t st = bulk_start(J, Int, b);
    if (lane_id() == 0) {
      sbase[w] = (uint32_t)((Int ? J.C : 0) + st);
      sn[w] = bulk_tot(J, Int)[b];
    }
  }
  __syncthreads();
  const uint32_t flags = sflags;
  if (!J.bulk || (flags & (FLAG_NEED_GROW | FLAG_MAP_FULL | FLAG_BULK_DUPS))) return;
  // Without the duplicate search (bulk_dedup 0: a job whose baselines had no
  // repeated leaf digest) the leaf part has nothing to do.
  if (!Final && !J.bulk_dedup) return;
  const uint32_t nl = sn[0], ni = Final ? sn[1] : 0u, n = nl + ni;
  if (n > J.bk_cap_b || n > kBulkRec || n >= bs) {
    // Does not fit: the host runs the job again without the bulk build.
    if (threadIdx.x == 0) g_or(&J.ctr->flags, FLAG_MAP_FULL);
    return;
  }
  // Record r's entry in the record array.
  auto rec_of = [&](uint32_t r) -> GlobalPtr<const u32x4_t> {
    const uint64_t e = r < nl ? (uint64_t)sbase[0] + r : (uint64_t)sbase[1] + (r - nl);
    return (GlobalPtr<const u32x4_t>)(J.bk_rec + 2 * e);
  };
  // Every record of the lane is loaded before any is stored to LDS; lane
  // threadIdx.x keeps records threadIdx.x + i * kBuildThreads in registers.
  constexpr uint32_t kRecPerLane = (kBulkRec + kBuildThreads - 1) / kBuildThreads;
  u32x4_t rd[kRecPerLane];
  {
    uint32_t rn[kRecPerLane];
#pragma unroll
    for (uint32_t i = 0; i < kRecPerLane; i++) {
      const uint32_t r = threadIdx.x + i * kBuildThreads;
      // Past the bucket's records: the array's first record (in bounds).
      const GlobalPtr<const u32x4_t> rec = r < n ? rec_of(r) : (GlobalPtr<const u32x4_t>)J.bk_rec;
      rd[i] = rec[0];
      rn[i] = rec[1].x;
    }
#pragma unroll
    for (uint32_t i = 0; i < kRecPerLane; i++) {
      const uint32_t r = threadIdx.x + i * kBuildThreads;
      if (r < n) {
        rkey[r] = rd[i].x;
        rnode[r] = rn[i];
        best[r] = 0xffffffffu;
      }
    }
  }
  for (uint64_t q = threadIdx.x; q < bs; q += blockDim.x) owner[q] = 0xffffffffu;
  __syncthreads();
  const uint32_t C1 = (uint32_t)(J.C - 1);
  const uint32_t bm = (uint32_t)J.map.bmask;
  auto rank_of = [&](uint32_t r) -> uint32_t {
    return r < nl ? rnode[r] - C1 : 0x80000000u | (0x7fffffffu - rnode[r]);
  };
#pragma unroll
  for (uint32_t i = 0; i < kRecPerLane; i++) {
    const uint32_t r = threadIdx.x + i * kBuildThreads;
    if (r >= n) continue;
    const u32x4_t dv = rd[i];
    const Digest d{dv.x | ((uint64_t)dv.y << 32), dv.z | ((uint64_t)dv.w << 32)};
    uint32_t q = (uint32_t)(map_start(J.map, d) & bm);
    uint32_t first = r;
    for (uint64_t it = 0; it < bs; it++, q = (q + 1) & bm) {
      uint32_t cur = __hip_atomic_load(&owner[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (cur == 0xffffffffu) {
        cur = atomicCAS(&owner[q], 0xffffffffu, r);
        if (cur == 0xffffffffu) break;  // claimed: this record is the digest's first
      }
      if (rkey[cur] == dv.x) {  // equal keys: compare the whole digests
        const u32x4_t ov = rec_of(cur)[0];
        if (ov.x == dv.x && ov.y == dv.y && ov.z == dv.z && ov.w == dv.w) {
          first = cur;
          break;
        }
      }
    }
    rslot[r] = (uint16_t)q;
    atomicMin(&best[first], rank_of(r));
  }
  __syncthreads();
  if (!Final) {
    for (uint32_t r = threadIdx.x; r < nl; r += blockDim.x) {
      const uint32_t q = rslot[r];
      if (best[owner[q]] != rank_of(r) && J.labels[rnode[r]] == PENDING) J.labels[rnode[r]] = SHIFT_DUPL;
    }
    return;
  }
  // The entry of each slot: its digest's lowest-ranked record.  A leaf that
  // is not its digest's entry is a repeated leaf digest: without the leaf
  // part's duplicate search pass 1 took it for a first occurrence, so the
  // host runs the job again with the search (FLAG_BULK_DUPS).
  bool win[kRecPerLane];
  bool dup = false;
#pragma unroll
  for (uint32_t i = 0; i < kRecPerLane; i++) {
    const uint32_t r = threadIdx.x + i * kBuildThreads;
    win[i] = r < n && best[owner[rslot[r]]] == rank_of(r);
    dup = dup || (r < nl && !win[i]);
  }
  if (!J.bulk_dedup && __syncthreads_or(dup)) {
    if (threadIdx.x == 0) g_or(&J.ctr->flags, FLAG_BULK_DUPS | FLAG_MAP_FULL);  // later stages stop
    return;
  }
  // Slots, written in slot order half a bucket at a time: each digest's
  // entry is staged in LDS by the lane that holds its record (the probing
  // arrays are free by then), EMPTY where no record landed.
  const uint32_t K = J.ckpt_id;
  uint32_t node[kRecPerLane], slot[kRecPerLane];
  uint32_t used = 0;
#pragma unroll
  for (uint32_t i = 0; i < kRecPerLane; i++) {
    const uint32_t r = threadIdx.x + i * kBuildThreads;
    node[i] = r < n ? rnode[r] : 0u;
    slot[i] = r < n ? rslot[r] : 0u;
    if (r < nl && !win[i]) {
      // A shifted leaf: its digest's entry, final (the build decides it):
      // this checkpoint and the winner's node, decoded from its rank.
      const uint32_t wr = best[owner[slot[i]]];
      const uint32_t wn = wr < 0x80000000u ? wr + C1 : 0x7fffffffu - (wr & 0x7fffffffu);
      J.slot_of[node[i]] = K != 0xffffffffu ? (((uint64_t)K << 32) | wn) : (kRefSlot | ((uint64_t)b * bs + slot[i]));
    }
    used += (r < n && win[i]) ? 1u : 0u;
  }
  for (uint32_t x = threadIdx.x; x < kBucketSlots / 32; x += blockDim.x) occ[x] = 0;
  __syncthreads();
  for (uint32_t q = threadIdx.x; q < bs; q += blockDim.x)
    if (owner[q] != 0xffffffffu) atomicOr(&occ[q >> 5], 1u << (q & 31));
  __syncthreads();  // every read of the probing arrays is done
  for (uint32_t h0 = 0; h0 < bs; h0 += kHalf) {
#pragma unroll
    for (uint32_t i = 0; i < kRecPerLane; i++) {
      const uint32_t r = threadIdx.x + i * kBuildThreads;
      if (r < n && win[i] && slot[i] >= h0 && slot[i] < h0 + kHalf) {
        sdig[slot[i] - h0] = make_uint4(rd[i].x, rd[i].y, rd[i].z, rd[i].w);
        snode[slot[i] - h0] = node[i];
      }
    }
    __syncthreads();
    const uint32_t hn = (uint32_t)min((uint64_t)kHalf, bs - h0);
    for (uint32_t q = threadIdx.x; q < hn; q += blockDim.x) {
      const GlobalPtr<u32x4_t> p = (GlobalPtr<u32x4_t>)&slots[h0 + q];
      if ((occ[(h0 + q) >> 5] >> ((h0 + q) & 31)) & 1u) {
        const uint4 dg = sdig[q];
        p[0] = u32x4_t{dg.x, dg.y, dg.z, dg.w};
        p[1] = u32x4_t{snode[q], K, 0, 0};
      } else {
        p[0] = u32x4_t{0, 0, 0, 0};
        p[1] = u32x4_t{0xffffffffu, 0xffffffffu, 0, 0};
      }
    }
    __syncthreads();  // the staging area is refilled for the next half
  }
  const uint32_t tot = block_sum<uint32_t>(used, sh);
  if (threadIdx.x == 0) J.bk_claim[b] = tot;
}

// ---------------------------------------------------------------------------
// Tree tiles.  A tile is a subtree of kTileLevels levels above a bottom level
// Lb; group g of a job with depth D has Lb = D - 8g.  Nodes on the bottom
// level are loaded; nodes above are computed from their children in LDS,
// except leaves one level above the deepest (chunk counts that are not a power
// of two), which are loaded too.
// ---------------------------------------------------------------------------

// Label of a leaf after the map inserts: a leaf still PENDING is the
// checkpoint's first occurrence of its digest; FIXED|PENDING is FIXED.
__device__ __forceinline__ uint8_t resolved_label(uint8_t raw) {
  return (raw & PENDING) ? (raw == PENDING ? (uint8_t)FIRST_OCUR : (uint8_t)FIXED_DUPL) : raw;
}

// A group-1 tile's bottom nodes are group-0 tile roots.  When no leaf of
// such a group-0 tile changed, the tile was not visited this checkpoint and
// its stored labels are stale: every node of it is FIXED.  Group-0 tile
// roots start at heap node 2^(D-8) - 1.
__device__ __forceinline__ bool check_child_tiles(const JobG& J, int g) { return g == 1 && !J.is_baseline; }
__device__ __forceinline__ uint64_t child_tile_base(const JobG& J) { return (1ull << (J.D - kTileLevels)) - 1; }

__device__ __forceinline__ uint32_t group_items(const JobG& J, int g, uint32_t* pre) {
  if (g == 0 && !J.is_baseline) {
    load_list_prefix(J.dirty_cnt, pre);
    return pre[kPendSpread];
  }
  const int Lb = (int)J.D - kTileLevels * g;
  return 1u << (Lb > kTileLevels ? Lb - kTileLevels : 0);
}

// Both passes process a tile with one wavefront (a 64-lane workgroup) so that
// many tiles are in flight (six wavefronts per SIMD: at most 80 registers).  Lane i owns bottom positions 4i .. 4i+3 of the
// tile, the two level-1 nodes above them and the level-2 node above those,
// so the first two levels are computed in registers; the levels above (at
// most 63 nodes) go through LDS, one barrier per level.  Every global load
// of a tile (labels, then the digests they select) is issued before the
// level steps and every global store after them: loads and stores share one
// in-order completion counter, so a load issued after a store waits for the
// store to be acknowledged, which would cost a memory round trip per level.
// Leaves one level above the deepest (chunk counts that are not a power of
// two) sit on level 1 of group-0 tiles and are loaded with the bottom nodes.
// Job fields are read once into locals: the compiler cannot tell the job
// table apart from the arrays the kernel stores to, and would reload them.
constexpr uint32_t kTileThreads = 64;
constexpr uint32_t kTopNodes = 63;  // tile nodes on levels 3 .. 8 above the bottom

// One lane's share of a tile of height H (H <= 8) with root r: bottom nodes
// xb .. xb+3 (positions 4i .. 4i+3), level-1 nodes x1, x1+1 (positions
// 2i, 2i+1), level-2 node x2 (position i); positions past the level's width
// do not exist.  LDS position q < 2^(H-2) - 1 holds tile node q in
// level-major order from the tile root (levels 3 .. H above the bottom).
// Idx: the node index type, 32 bits when the tree has fewer than 2^31 nodes
// (no heap index of a tile position then passes 2^32), so that index
// arithmetic and label addresses take one instruction instead of two.
template <class Idx>
struct TileLane {
  int H;
  Idx r;
  uint32_t width;  // bottom positions, 2^H
  Idx xb, x1, x2;  // first heap node of this lane on levels 0, 1, 2
  bool vb[4];      // bottom node exists
  bool v1[2], v2;  // level-1 / level-2 position exists
  uint32_t nq;     // LDS positions (levels >= 3)
};

template <class Idx>
__device__ __forceinline__ TileLane<Idx> tile_lane(const TileRef& tr, uint64_t N, uint32_t i) {
  TileLane<Idx> L;
  L.H = tr.H;
  L.r = (Idx)tr.r;
  L.width = 1u << tr.H;
  L.xb = ((L.r + 1) << tr.H) - 1 + 4 * i;
  L.x1 = tr.H >= 1 ? ((L.r + 1) << (tr.H - 1)) - 1 + 2 * i : 0;
  L.x2 = tr.H >= 2 ? ((L.r + 1) << (tr.H - 2)) - 1 + i : 0;
#pragma unroll
  for (int k = 0; k < 4; k++) L.vb[k] = 4 * i + k < L.width && L.xb + k < (Idx)N;
#pragma unroll
  for (int k = 0; k < 2; k++) L.v1[k] = tr.H >= 1 && 2 * i + k < (L.width >> 1);
  L.v2 = tr.H >= 2 && i < (L.width >> 2);
  L.nq = tr.H >= 3 ? (1u << (tr.H - 2)) - 1 : 0;
  return L;
}

template <class Idx>
__device__ __forceinline__ Idx tile_node_t(Idx r, uint32_t i) {
  const uint32_t q = 31 - __clz(i + 1);
  return ((r + 1) << q) - 1 + (i + 1 - (1u << q));
}

// Labels of a lane's bottom nodes (0xff where absent; a clean child tile's
// root reads as the pass's clean label: DONE in pass 1, FIXED in pass 2,
// chosen by the caller through `clean`) and of its level-1 leaves (vu).  The
// loads are straight-line, from clamped addresses, with the results selected
// afterwards, so that all of them are in flight together.
template <class Idx>
__device__ __forceinline__ void tile_labels(const JobG& J, int g, const TileLane<Idx>& T, uint64_t C, uint8_t raw[4],
                                            uint8_t rawu[2], bool vu[2], uint8_t clean) {
  const GlobalPtr<const uint8_t> labels = J.labels;
  uint8_t b[4], u[2];
#pragma unroll
  for (int k = 0; k < 4; k++) b[k] = labels[T.vb[k] ? T.xb + k : 0];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    vu[k] = T.v1[k] && T.x1 + k >= (Idx)(C - 1);
    u[k] = labels[vu[k] ? T.x1 + k : 0];
  }
  bool cl[4] = {false, false, false, false};
  if (check_child_tiles(J, g)) {  // uniform
    const Idx base = (Idx)child_tile_base(J);
    uint32_t e[4];
#pragma unroll
    for (int k = 0; k < 4; k++) e[k] = J.tile_epoch[T.vb[k] ? T.xb + k - base : 0];
#pragma unroll
    for (int k = 0; k < 4; k++) cl[k] = e[k] != J.epoch;
  }
#pragma unroll
  for (int k = 0; k < 4; k++) raw[k] = !T.vb[k] ? (uint8_t)0xff : cl[k] ? clean : b[k];
#pragma unroll
  for (int k = 0; k < 2; k++) rawu[k] = vu[k] ? u[k] : (uint8_t)0xff;
}

// Digests of the nodes labelled `want` that a lane combines: a pair of
// bottom nodes when both are labelled `want` (their parent's digest is the
// hash of the two), a level-1 leaf when its level-1 sibling may be `want`
// too (a leaf so labelled, or a node whose two bottom children are); zero
// elsewhere.  A node whose sibling differs never reaches a parent digest.
// The loads are issued only when some lane of the wavefront needs one, all
// together from clamped addresses: a tile of isolated changes (most tiles
// when changes are spread thin) skips the round trip.
template <class Idx>
__device__ __forceinline__ void tile_digests(GlobalPtr<const Digest> tree, const TileLane<Idx>& T, const uint8_t bl[4],
                                             const uint8_t ul[2], const bool vu[2], uint8_t want, Digest bd[4],
                                             Digest ud[2]) {
  bool nb[4], nu[2], m1[2];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const bool pair = bl[2 * k] == want && bl[2 * k + 1] == want;
    nb[2 * k] = nb[2 * k + 1] = pair;
    m1[k] = vu[k] ? ul[k] == want : (T.v1[k] && pair);
  }
#pragma unroll
  for (int k = 0; k < 2; k++) nu[k] = vu[k] && ul[k] == want && m1[k ^ 1];
  const bool any = nb[0] || nb[2] || nu[0] || nu[1];
  if (!__ballot(any)) {
#pragma unroll
    for (int k = 0; k < 4; k++) bd[k] = Digest{0, 0};
#pragma unroll
    for (int k = 0; k < 2; k++) ud[k] = Digest{0, 0};
    return;
  }
#pragma unroll
  for (int k = 0; k < 4; k++) bd[k] = load_digest(tree + (nb[k] ? T.xb + k : 0));
#pragma unroll
  for (int k = 0; k < 2; k++) ud[k] = load_digest(tree + (nu[k] ? T.x1 + k : 0));
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (!nb[k]) bd[k] = Digest{0, 0};
#pragma unroll
  for (int k = 0; k < 2; k++)
    if (!nu[k]) ud[k] = Digest{0, 0};
}

// Tiles of one workgroup of a tile pass: iterations n = 0 .. niter-1 take
// item wg + n * nwg of the group.  The tile ids of 64 iterations are loaded
// together into LDS (one per lane of the one-wavefront workgroup), so a
// tile's id is an LDS read rather than a global load that would wait behind
// the previous tile's stores.
__device__ __forceinline__ uint32_t tile_item(const JobG& J, bool listed, const uint32_t* pre, uint32_t it) {
  return listed ? list_item(J.dirty_list, J.dirty_cap, pre, it) : it;
}

__device__ __forceinline__ uint32_t tile_iterations(uint32_t wg, uint32_t nwg, uint32_t nitems) {
  return wg < nitems ? (nitems - wg + nwg - 1) / nwg : 0u;
}

// Tile id of iteration n (every lane calls it with the same n, iterations in
// increasing order from 0): refills `ids` when n starts a block of 64.  A
// group visited whole needs no list.
__device__ __forceinline__ uint64_t tile_at(const JobG& J, bool listed, const uint32_t* pre, uint32_t* ids,
                                            uint32_t wg, uint32_t nwg, uint32_t niter, uint32_t n) {
  if (!listed) return wg + (uint64_t)n * nwg;
  if ((n & 63) == 0) {
    __syncthreads();  // the previous block's ids are read
    const uint32_t m = n + threadIdx.x;
    if (m < niter) ids[threadIdx.x] = tile_item(J, listed, pre, wg + m * nwg);
    __syncthreads();
  }
  return ids[n & 63];
}

// Level-2 positions 2j and 2j+1 of a tile (lanes 2j, 2j+1) are the children
// of level-3 node j: true when some such pair has both bits set in `m`.
__device__ __forceinline__ bool sibling_pair(uint64_t m) { return (m & (m >> 1) & 0x5555555555555555ull) != 0; }

// Label of LDS position q (levels 3 .. H, level-major from the tile root) when
// no node above level 2 is shifted: FIXED when every level-2 node below it is
// FIXED, FIRST when every one is FIRST, else DONE (two children with equal
// labels pass theirs on, different ones make DONE).  `mx`, `mr`: the level-2
// FIXED and FIRST masks (bit = lane = level-2 position).
__device__ __forceinline__ uint8_t upper_label(uint32_t q, int H, uint64_t mx, uint64_t mr) {
  const uint32_t l = 31 - __clz(q + 1);  // level below the tile root
  const uint32_t k = q + 1 - (1u << l);  // position within that level
  const uint32_t span = 1u << (H - (int)l - 2);  // level-2 nodes below it
  const uint32_t start = k * span;
  const uint64_t lm = span >= 64 ? ~0ull : ((1ull << span) - 1);
  if (((mx >> start) & lm) == lm) return FIXED_DUPL;
  if (((mr >> start) & lm) == lm) return FIRST_OCUR;
  return DONE;
}

// Pass 1, the first-occurrence forest (include/reference_impl.hpp:73-82).
// Group 0 visits the dirty tiles only; the groups above visit every tile.
// In an incremental checkpoint the first-occurrence interior nodes of a tile
// are appended to the fint list for the interior map inserts (a baseline's
// interior inserts visit every node instead).  Group-0 leaves left PENDING
// by the inserts are resolved and written back here.
template <class Idx>
__global__ __launch_bounds__(kTileThreads) void k_pass1(const JobG* __restrict__ jobs, const uint32_t* __restrict__ prefix, uint32_t njobs,
                                                        int g) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobG& J = jobs[j];
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;
  if (g >= tile_groups(J.D)) return;
  const uint32_t wg = blockIdx.x - prefix[j], nwg = prefix[j + 1] - prefix[j];
  __shared__ uint32_t pre[kPendSpread + 1];
  __shared__ uint8_t tlab[kTopNodes + 64];
  __shared__ Digest tdig[kTopNodes + 64];
  const uint32_t nitems = group_items(J, g, pre);
  const uint32_t i = threadIdx.x;
  const bool listing = !J.is_baseline;
  const uint64_t N = J.N, C = J.C;
  const GlobalPtr<uint8_t> labels = J.labels;
  const GlobalPtr<Digest> tree = J.tree;
  const int kind = J.hash_kind;
  const GlobalPtr<uint32_t> fint_list = J.fint_list;
  const GlobalPtr<uint32_t> fint_n = J.fint_n;
  const uint32_t tbase = J.tile_base[g];
  const bool listed = g == 0 && listing;
  __shared__ uint32_t ids[64];
  const uint32_t niter = tile_iterations(wg, nwg, nitems);
  for (uint32_t n = 0; n < niter; n++) {
    const uint64_t t = tile_at(J, listed, pre, ids, wg, nwg, niter, n);
    const TileLane<Idx> T = tile_lane<Idx>(tile_ref(J, g, t), N, i);
    // Loads: labels of the bottom nodes and of the level-1 leaves ...
    uint8_t raw[4], rawu[2], bl[4], ul[2];
    bool vu[2];
    tile_labels(J, g, T, C, raw, rawu, vu, DONE);
#pragma unroll
    for (int k = 0; k < 4; k++) bl[k] = g == 0 && T.vb[k] ? resolved_label(raw[k]) : raw[k];
#pragma unroll
    for (int k = 0; k < 2; k++) ul[k] = vu[k] ? resolved_label(rawu[k]) : (uint8_t)0xff;
    // ... then the digests of the first-occurrence pairs among them.
    Digest bd[4], ud[2];
    tile_digests(tree, T, bl, ul, vu, FIRST_OCUR, bd, ud);
    // Levels 1 and 2 in registers.
    uint8_t l1[2];
    Digest d1[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
      l1[k] = 0xff;
      d1[k] = Digest{0, 0};
      if (vu[k]) {
        l1[k] = ul[k];
        d1[k] = ud[k];
      } else if (T.v1[k]) {
        const bool both = bl[2 * k] == FIRST_OCUR && bl[2 * k + 1] == FIRST_OCUR;
        l1[k] = both ? (uint8_t)FIRST_OCUR : (uint8_t)DONE;
        if (both) d1[k] = hash_pair(kind, bd[2 * k], bd[2 * k + 1]);
      }
    }
    uint8_t l2 = 0xff;
    Digest d2{0, 0};
    if (T.v2) {
      const bool both = l1[0] == FIRST_OCUR && l1[1] == FIRST_OCUR;
      l2 = both ? (uint8_t)FIRST_OCUR : (uint8_t)DONE;
      if (both) d2 = hash_pair(kind, d1[0], d1[1]);
    }
    // Levels 3 .. H in LDS (level-2 results at positions nq .. 2 nq).  A node
    // is FIRST only when both children are: with no two sibling level-2
    // nodes FIRST every node above level 2 is DONE, and the LDS levels (a
    // barrier each) are skipped.
    const bool deep = T.nq && sibling_pair(__ballot(T.v2 && l2 == FIRST_OCUR));  // uniform
    if (deep) {
      if (T.v2) {
        tlab[T.nq + i] = l2;
        tdig[T.nq + i] = d2;
      }
      for (int st = 3; st <= T.H; st++) {
        __syncthreads();
        const uint32_t q0 = (1u << (T.H - st)) - 1, qc0 = (1u << (T.H - st + 1)) - 1;
        if (i < (T.width >> st)) {
          const uint32_t qc = qc0 + 2 * i;
          const bool both = tlab[qc] == FIRST_OCUR && tlab[qc + 1] == FIRST_OCUR;
          tlab[q0 + i] = both ? (uint8_t)FIRST_OCUR : (uint8_t)DONE;
          tdig[q0 + i] = both ? hash_pair(kind, tdig[qc], tdig[qc + 1]) : Digest{0, 0};
        }
      }
      __syncthreads();
    }
    uint8_t lq = 0xff;
    Digest dq{0, 0};
    Idx xq = 0;
    if (i < T.nq) {
      lq = deep ? tlab[i] : (uint8_t)DONE;
      if (deep) dq = tdig[i];
      xq = tile_node_t<Idx>(T.r, i);
    }
    // Positions of the first-occurrence interior nodes in the tile's fint
    // region (global tile id tbase + t).
    const bool f1a = T.v1[0] && !vu[0] && l1[0] == FIRST_OCUR, f1b = T.v1[1] && !vu[1] && l1[1] == FIRST_OCUR;
    const bool f2 = T.v2 && l2 == FIRST_OCUR, fq = i < T.nq && lq == FIRST_OCUR;
    uint64_t fpos = 0;
    if (listing) {
      const uint32_t cnt = (uint32_t)f1a + (uint32_t)f1b + (uint32_t)f2 + (uint32_t)fq;
      const uint32_t incl = wave_incl_scan(cnt);
      const uint64_t tid = tbase + t;
      fpos = tid * kTileInterior + incl - cnt;
      if (i == 63) fint_n[tid] = incl;  // the tile's total
    }
    // Stores.
    if (g == 0) {
#pragma unroll
      for (int k = 0; k < 4; k++)
        if (T.vb[k] && (raw[k] & PENDING)) labels[T.xb + k] = bl[k];
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const Idx x = T.x1 + k;
      if (vu[k]) {
        if (rawu[k] & PENDING) labels[x] = ul[k];
      } else if (T.v1[k]) {
        labels[x] = l1[k];
        if (l1[k] == FIRST_OCUR) store_digest(tree + x, d1[k]);
      }
    }
    if (T.v2) {
      labels[T.x2] = l2;
      if (f2) store_digest(tree + T.x2, d2);
    }
    if (i < T.nq) {
      labels[xq] = lq;
      if (fq) store_digest(tree + xq, dq);
    }
    if (listing) {
      if (f1a) fint_list[fpos++] = (uint32_t)T.x1;
      if (f1b) fint_list[fpos++] = (uint32_t)(T.x1 + 1);
      if (f2) fint_list[fpos++] = (uint32_t)T.x2;
      if (fq) fint_list[fpos] = (uint32_t)xq;
    }
    __syncthreads();  // the LDS levels are reused by the next tile
  }
}

// Pass 2, the duplicate forest (include/reference_impl.hpp:83-104): a parent
// of two shifted children is shifted when its digest is in the map; two
// children with different labels make a DONE parent, equal labels carry up.
// Lookups go level by level (a parent is looked up only when both children
// turned out shifted); a lane's two level-1 lookups are in flight together.
__device__ __forceinline__ uint8_t dup_label(uint8_t la, uint8_t lb, bool found) {
  if (la != lb) return DONE;
  if (la != SHIFT_DUPL) return la;  // both FIXED, both FIRST or both DONE
  return found ? (uint8_t)SHIFT_DUPL : (uint8_t)DONE;
}

// Lookups of up to NK keys: out[n] = the matching slot, outv[n] = its value
// (~0 for both when need[n] is false or the key is absent).  Each
// probe round loads two consecutive slots of every key still searching
// before comparing any of them, so a lookup that ends within two slots of
// its home slot (most of them at the map's load factor, found or not) costs
// one memory round trip.
template <int NK>
__device__ __forceinline__ void map_find_n(const MapG& m, const Digest* key, const bool* need, uint64_t* out,
                                           uint64_t* outv) {
  uint64_t s[NK];
  bool todo[NK];
#pragma unroll
  for (int n = 0; n < NK; n++) {
    out[n] = ~0ull;
    outv[n] = ~0ull;
    todo[n] = need[n];
    s[n] = map_start(m, key[n]);
  }
  for (uint64_t it = 0; it <= m.bmask; it += 2) {
    uint4 k[NK][2];
    unsigned long long v[NK][2];
#pragma unroll
    for (int n = 0; n < NK; n++) {
      if (!todo[n]) continue;
      load_slot(m, s[n], k[n][0], v[n][0]);
      load_slot(m, map_next(m, s[n], 1), k[n][1], v[n][1]);
    }
    bool any = false;
#pragma unroll
    for (int n = 0; n < NK; n++) {
      if (!todo[n]) continue;
#pragma unroll
      for (int e = 0; e < 2; e++) {
        if (!todo[n]) break;
        const unsigned long long val = v[n][e];
        const uint64_t k0 = k[n][e].x | ((uint64_t)k[n][e].y << 32), k1 = k[n][e].z | ((uint64_t)k[n][e].w << 32);
        if (val == SLOT_EMPTY) {
          todo[n] = false;
        } else if (k0 == key[n].k0 && k1 == key[n].k1) {
          out[n] = map_next(m, s[n], e);
          outv[n] = val;
          todo[n] = false;
        }
      }
      s[n] = map_next(m, s[n], 2);
      any |= todo[n];
    }
    if (!any) break;
  }
}

template <class Idx>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_pass2(const JobG* __restrict__ jobs, const uint32_t* __restrict__ prefix, uint32_t njobs,
                                                        int g) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobG& J = jobs[j];
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;
  if (g >= tile_groups(J.D)) return;
  const uint32_t wg = blockIdx.x - prefix[j], nwg = prefix[j + 1] - prefix[j];
  __shared__ uint32_t pre[kPendSpread + 1];
  __shared__ uint8_t tlab[kTopNodes + 64];
  __shared__ Digest tdig[kTopNodes + 64];
  __shared__ uint64_t tref[kTopNodes];
  const uint32_t nitems = group_items(J, g, pre);
  const uint32_t i = threadIdx.x;
  const uint64_t N = J.N, C = J.C;
  const GlobalPtr<uint8_t> labels = J.labels;
  const GlobalPtr<Digest> tree = J.tree;
  const GlobalPtr<uint64_t> slot_of = J.slot_of;
  // A shifted node's map reference: the entry's value (final: every insert
  // of the checkpoint is done), or its slot for the reserved value.
  auto ref_of = [](uint64_t slot, uint64_t val) { return ref_is_slot(val) ? kRefSlot | slot : val; };
  const int kind = J.hash_kind;
  const MapG map = J.map;
  const bool listed = g == 0 && !J.is_baseline;
  __shared__ uint32_t ids[64];
  const uint32_t niter = tile_iterations(wg, nwg, nitems);
  for (uint32_t n = 0; n < niter; n++) {
    const TileLane<Idx> T = tile_lane<Idx>(tile_ref(J, g, tile_at(J, listed, pre, ids, wg, nwg, niter, n)), N, i);
    // Loads: labels (leaves were resolved by pass 1), then the digests of
    // the shifted pairs.
    uint8_t bl[4], ul[2];
    bool vu[2];
    tile_labels(J, g, T, C, bl, ul, vu, FIXED_DUPL);
    Digest bd[4], ud[2];
    tile_digests(tree, T, bl, ul, vu, SHIFT_DUPL, bd, ud);
    // Level 1: both lookups of the lane at once.
    bool look[2];
    Digest d1[2];
    uint64_t f1[2], e1[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
      look[k] = !vu[k] && T.v1[k] && bl[2 * k] == SHIFT_DUPL && bl[2 * k + 1] == SHIFT_DUPL;
      d1[k] = vu[k] ? ud[k] : look[k] ? hash_pair(kind, bd[2 * k], bd[2 * k + 1]) : Digest{0, 0};
    }
    map_find_n<2>(map, d1, look, f1, e1);
    uint8_t l1[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
      l1[k] = vu[k] ? ul[k] : T.v1[k] ? dup_label(bl[2 * k], bl[2 * k + 1], f1[k] != ~0ull) : (uint8_t)0xff;
      if (!vu[k] && l1[k] != SHIFT_DUPL) d1[k] = Digest{0, 0};
      if (!vu[k] && l1[k] == SHIFT_DUPL) slot_of[T.x1 + k] = ref_of(f1[k], e1[k]);  // stored at once
    }
    // Level 2.
    uint8_t l2 = 0xff;
    Digest d2{0, 0};
    uint64_t f2 = ~0ull, e2 = ~0ull;
    if (T.v2) {
      const bool lk = l1[0] == SHIFT_DUPL && l1[1] == SHIFT_DUPL;
      if (lk) d2 = hash_pair(kind, d1[0], d1[1]);
      map_find_n<1>(map, &d2, &lk, &f2, &e2);
      l2 = dup_label(l1[0], l1[1], f2 != ~0ull);
      if (l2 != SHIFT_DUPL) d2 = Digest{0, 0};
      if (l2 == SHIFT_DUPL) slot_of[T.x2] = ref_of(f2, e2);
    }
    // Levels 3 .. H in LDS (level-2 results at positions nq .. 2 nq).  With
    // no two sibling level-2 nodes shifted no node above level 2 is looked
    // up or shifted, and its label follows from the level-2 labels below it
    // (upper_label): the LDS levels run only for tiles with such a pair.
    const bool deep = T.nq && sibling_pair(__ballot(T.v2 && l2 == SHIFT_DUPL));  // uniform
    if (deep) {
      if (T.v2) {
        tlab[T.nq + i] = l2;
        tdig[T.nq + i] = d2;
      }
      for (int st = 3; st <= T.H; st++) {
        __syncthreads();
        const uint32_t q0 = (1u << (T.H - st)) - 1, qc0 = (1u << (T.H - st + 1)) - 1;
        if (i < (T.width >> st)) {
          const uint32_t qc = qc0 + 2 * i;
          const uint8_t la = tlab[qc], lb = tlab[qc + 1];
          const bool lk = la == SHIFT_DUPL && lb == SHIFT_DUPL;
          Digest d{0, 0};
          uint64_t f = ~0ull, fv = ~0ull;
          if (lk) {
            d = hash_pair(kind, tdig[qc], tdig[qc + 1]);
            map_find_n<1>(map, &d, &lk, &f, &fv);
          }
          const uint8_t L = dup_label(la, lb, f != ~0ull);
          tlab[q0 + i] = L;
          tdig[q0 + i] = L == SHIFT_DUPL ? d : Digest{0, 0};
          tref[q0 + i] = ref_of(f, fv);
        }
      }
      __syncthreads();
    }
    // Stores: labels of the interior nodes, digest (and, above level 2, map
    // reference) of the shifted ones.
#pragma unroll
    for (int k = 0; k < 2; k++) {
      if (T.v1[k] && !vu[k]) {
        const Idx x = T.x1 + k;
        labels[x] = l1[k];
        if (l1[k] == SHIFT_DUPL) store_digest(tree + x, d1[k]);
      }
    }
    if (T.v2) {
      labels[T.x2] = l2;
      if (l2 == SHIFT_DUPL) store_digest(tree + T.x2, d2);
    }
    const uint64_t mx2 = __ballot(T.v2 && l2 == FIXED_DUPL), mr2 = __ballot(T.v2 && l2 == FIRST_OCUR);
    if (i < T.nq) {
      const Idx x = tile_node_t<Idx>(T.r, i);
      const uint8_t L = deep ? tlab[i] : upper_label(i, T.H, mx2, mr2);
      labels[x] = L;
      if (L == SHIFT_DUPL) {
        store_digest(tree + x, tdig[i]);
        slot_of[x] = tref[i];
      }
    }
    __syncthreads();  // the LDS levels are reused by the next tile
  }
}

// ---------------------------------------------------------------------------
// Roots.  Node v is a root when it is FIRST or SHIFT and its parent is DONE;
// node 0 is a root when it is FIRST (include/reference_impl.hpp:85-106).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t leaf_count(const TreeGeom& g, uint64_t v) {
  LeafRuns r = subtree_runs(g, v);
  return r.n0 + r.n1;
}

// Node blocks of kNodesPerBlock ids; a workgroup owns a contiguous run of
// them so that its counts are contiguous in node order.
__device__ __forceinline__ void node_block_range(const JobG& J, uint32_t wg, uint32_t nwg, uint32_t& b0,
                                                 uint32_t& b1) {
  const uint32_t per = (J.nblk + nwg - 1) / nwg;
  b0 = min(J.nblk, wg * per);
  b1 = min(J.nblk, b0 + per);
}

// Bytes of x that are zero: bit 7 of each such byte (exact, no borrow
// between bytes).
__device__ __forceinline__ uint32_t zero_bytes(uint32_t x) {
  return ~(((x & 0x7f7f7f7fu) + 0x7f7f7f7fu) | x) & 0x80808080u;
}

// Bit 7 of byte t -> bit t (t < 4).
__device__ __forceinline__ uint32_t byte_bits(uint32_t m) {
  return ((m >> 7) & 1u) | ((m >> 14) & 2u) | ((m >> 21) & 4u) | ((m >> 28) & 8u);
}

// Root kinds of the 16 nodes starting at v0 (a multiple of 16) as a bit mask
// per kind (bit t: node v0 + t), from their labels `w` and the labels of
// their parents (v - 1) / 2, which are v0/2 - 1 .. v0/2 + 7: the last byte of
// `lo` for t = 0, then byte (t - 1) / 2 of `pw`, gathered into four dwords
// matching `w` with byte permutes.  All 16 nodes are tested together, a
// byte per node: the node is FIRST or SHIFT and its parent DONE.
__device__ __forceinline__ void scan16(const JobG& J, uint64_t v0, const uint4 w, uint64_t lo, uint64_t pw,
                                       uint32_t& fmask, uint32_t& smask) {
  fmask = smask = 0;
  if (v0 >= J.N) return;
  const uint32_t lohi = (uint32_t)(lo >> 32), pl = (uint32_t)pw, ph = (uint32_t)(pw >> 32);
  // __builtin_amdgcn_perm(s0, s1, sel): byte k of the result is byte sel_k of
  // the 8 bytes {s1 (0..3), s0 (4..7)}.
  const uint32_t par[4] = {__builtin_amdgcn_perm(pl, lohi, 0x05040403u),  // p_first, pw0, pw0, pw1
                           __builtin_amdgcn_perm(pl, pl, 0x03020201u),    // pw1, pw2, pw2, pw3
                           __builtin_amdgcn_perm(ph, pl, 0x05040403u),    // pw3, pw4, pw4, pw5
                           __builtin_amdgcn_perm(ph, ph, 0x03020201u)};   // pw5, pw6, pw6, pw7
  const uint32_t lab[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const uint32_t pdone = zero_bytes(par[d] ^ (DONE * 0x01010101u));
    fmask |= byte_bits(zero_bytes(lab[d] ^ (FIRST_OCUR * 0x01010101u)) & pdone) << (4 * d);
    smask |= byte_bits(zero_bytes(lab[d] ^ (SHIFT_DUPL * 0x01010101u)) & pdone) << (4 * d);
  }
  if (v0 == 0) {
    // A tree emits node 0 only as a first occurrence (or, with the
    // root-shift option, also as a shifted root); in a list job node 0 is
    // a candidate only when it is the single chunk.
    const uint8_t L = (uint8_t)w.x;
    fmask = (fmask & ~1u) | (L == FIRST_OCUR ? 1u : 0u);
    smask = (smask & ~1u) | (L == SHIFT_DUPL && (J.list_mode || J.root_shift) ? 1u : 0u);
  }
  if (J.N - v0 < 16) {
    const uint32_t keep = (1u << (uint32_t)(J.N - v0)) - 1;
    fmask &= keep;
    smask &= keep;
  }
}

// Chunks under the first-occurrence roots of a mask.
__device__ __forceinline__ uint32_t first_chunks(const TreeGeom& g, uint64_t v0, uint32_t fmask) {
  uint32_t fc = 0;
  while (fmask) {
    const int t = __builtin_ctz(fmask);
    fmask &= fmask - 1;
    fc += (uint32_t)leaf_count(g, v0 + t);
  }
  return fc;
}

// The root scan runs at most kRootWG workgroups per job; each owns a
// contiguous run of node blocks and loads the labels of kRootBatch of them at
// once.
constexpr uint32_t kRootBatch = 4;

// Labels of node blocks b .. b+kRootBatch-1 and the parent labels of each
// group of 16 (FIXED bytes past the run or past the last node; the labels
// array is padded to whole node blocks).  The parent words are loaded for
// every group so that all loads of the batch are in flight together (they
// are contiguous: 2 KiB per node block).
struct LabelGroup {
  uint4 w;      // labels v0 .. v0+15
  uint64_t lo;  // labels v0/2 - 8 .. v0/2 - 1 (zero for v0 = 0)
  uint64_t pw;  // labels v0/2 .. v0/2 + 7
};

__device__ __forceinline__ void load_root_batch(const JobG& J, uint32_t b, uint32_t b1, LabelGroup g[kRootBatch]) {
  const GlobalPtr<const uint8_t> labels = J.labels;
#pragma unroll
  for (uint32_t q = 0; q < kRootBatch; q++) {
    const bool in = b + q < b1;
    const uint64_t v0 = (uint64_t)(in ? b + q : b) * kNodesPerBlock + threadIdx.x * 16, h = v0 >> 1;
    g[q].w = gload16((GlobalPtr<const void>)(labels + v0));
    g[q].lo = *(GlobalPtr<const uint64_t>)(labels + (v0 >= 16 ? h - 8 : 0));
    g[q].pw = *(GlobalPtr<const uint64_t>)(labels + h);
    if (v0 < 16) g[q].lo = 0;
    if (!in || v0 >= J.N) g[q].w = make_uint4(0x01010101u, 0x01010101u, 0x01010101u, 0x01010101u);
  }
}

// Sums of three counts over the block (256 threads); sh holds 12 entries.
__device__ __forceinline__ void block_sum3(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t* sh) {
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  const uint32_t w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane_id() == 0) {
    sh[w] = a;
    sh[4 + w] = b;
    sh[8 + w] = c;
  }
  __syncthreads();
  a = b = c = 0;
  for (uint32_t i = 0; i < nw; i++) {
    a += sh[i];
    b += sh[4 + i];
    c += sh[8 + i];
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_roots_count(const JobG* __restrict__ jobs, const uint32_t* __restrict__ prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobG& J = jobs[j];
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;
  const uint32_t wg = blockIdx.x - prefix[j], nwg = prefix[j + 1] - prefix[j];
  uint32_t b0, b1;
  node_block_range(J, wg, nwg, b0, b1);
  const TreeGeom g = make_geom(J.C);
  uint32_t f = 0, fc = 0, s = 0;
  for (uint32_t b = b0; b < b1; b += kRootBatch) {
    LabelGroup lg[kRootBatch];
    load_root_batch(J, b, b1, lg);
#pragma unroll
    for (uint32_t q = 0; q < kRootBatch; q++) {
      const uint64_t v0 = (uint64_t)(b + q) * kNodesPerBlock + threadIdx.x * 16;
      uint32_t fm, sm;
      scan16(J, v0, lg[q].w, lg[q].lo, lg[q].pw, fm, sm);
      f += __popc(fm);
      fc += first_chunks(g, v0, fm);
      s += __popc(sm);
      // Groups with a label other than FIXED (past the nodes the loads put
      // FIXED bytes): the root writer resets them.
      const uint4 w = lg[q].w;
      const bool dirty = ((w.x ^ 0x01010101u) | (w.y ^ 0x01010101u) | (w.z ^ 0x01010101u) | (w.w ^ 0x01010101u)) != 0;
      const uint64_t dm = __ballot(dirty);
      if (b + q < b1) {  // uniform
        J.rmask[(uint64_t)(b + q) * (kNodesPerBlock / 16) + threadIdx.x] = fm | (sm << 16);
        if (lane_id() == 0) J.lmask[(uint64_t)(b + q) * (kNodesPerBlock / 16 / 64) + (threadIdx.x >> 6)] = dm;
      }
    }
  }
  __shared__ uint32_t sh[12];
  block_sum3(f, fc, s, sh);
  if (threadIdx.x == 0) {
    J.blk_f[wg] = f;
    J.blk_fc[wg] = fc;
    J.blk_s[wg] = s;
  }
}

// Writes the roots in node order, from the root masks k_roots_count stored:
// first roots (node, chunk offset of its region) and the nodes of the
// shifted roots (their sources are looked up by k_sbin_count).  Each
// workgroup starts from the sums of the counts of the workgroups before it;
// the last one also writes the totals.  Per node block the three counts are
// scanned together, packed (a block holds at most 4096 roots of each kind,
// and its first regions at most 2^32 - 1 chunks).  The masks are the only
// loads after the prefix, so the stores never hold up a load.
__global__ __launch_bounds__(256) void k_roots_write(const JobG* __restrict__ jobs, const uint32_t* __restrict__ prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobG& J = jobs[j];
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;
  const uint32_t wg = blockIdx.x - prefix[j], nwg = prefix[j + 1] - prefix[j];
  uint32_t b0, b1;
  node_block_range(J, wg, nwg, b0, b1);
  const TreeGeom g = make_geom(J.C);
  __shared__ uint32_t sh[12];
  __shared__ uint64_t sh64[5];
  uint32_t run_f = 0, run_fc = 0, run_s = 0;
  for (uint32_t i = threadIdx.x; i < wg; i += blockDim.x) {
    run_f += J.blk_f[i];
    run_fc += J.blk_fc[i];
    run_s += J.blk_s[i];
  }
  block_sum3(run_f, run_fc, run_s, sh);
  if (wg == nwg - 1 && threadIdx.x == 0) {
    J.ctr->nfirst = run_f + J.blk_f[wg];
    J.ctr->dchunks = run_fc + J.blk_fc[wg];
    J.ctr->nshift = run_s + J.blk_s[wg];
  }
  // Label hygiene of a tree checkpoint: every group of 16 nodes the count
  // found with a label other than FIXED is reset to FIXED (16 bytes), which
  // is what pass 2 computes for a subtree whose leaves are all fixed.  A
  // later checkpoint that does not visit a tile then reads FIXED labels
  // there, and its leaf kernel writes the labels of changed leaves only.  The
  // count has read every label (and parent label) of the job by now; the
  // list method keeps its labels.
  const bool reset = !J.list_mode;
  const GlobalPtr<u32x4_t> lab16 = (GlobalPtr<u32x4_t>)J.labels;
  for (uint32_t b = b0; b < b1; b += kRootBatch) {
    uint32_t rm[kRootBatch];
    uint64_t lm[kRootBatch];
#pragma unroll
    for (uint32_t q = 0; q < kRootBatch; q++) {
      const uint64_t bq = b + q < b1 ? b + q : b;
      rm[q] = J.rmask[bq * (kNodesPerBlock / 16) + threadIdx.x];
      lm[q] = reset ? J.lmask[bq * (kNodesPerBlock / 16 / 64) + (threadIdx.x >> 6)] : 0ull;
    }
#pragma unroll
    for (uint32_t q = 0; q < kRootBatch; q++) {
      if (b + q >= b1) break;  // uniform
      const uint64_t v0 = (uint64_t)(b + q) * kNodesPerBlock + threadIdx.x * 16;
      if ((lm[q] >> lane_id()) & 1ull) lab16[v0 / 16] = u32x4_t{0x01010101u, 0x01010101u, 0x01010101u, 0x01010101u};
      uint32_t fm = rm[q] & 0xffffu, sm = rm[q] >> 16;
      if (!__syncthreads_or((fm | sm) != 0)) continue;  // no root in this node block (uniform)
      const uint32_t f = __popc(fm), fc = first_chunks(g, v0, fm), s = __popc(sm);
      uint64_t tot;
      const uint64_t ex = block_excl_scan<4>((uint64_t)f | ((uint64_t)s << 16) | ((uint64_t)fc << 32), sh64, tot);
      uint32_t pf = run_f + (uint32_t)(ex & 0xffff), ps = run_s + (uint32_t)((ex >> 16) & 0xffff);
      uint32_t pfc = run_fc + (uint32_t)(ex >> 32);
      run_f += (uint32_t)(tot & 0xffff);
      run_s += (uint32_t)((tot >> 16) & 0xffff);
      run_fc += (uint32_t)(tot >> 32);
      // Stores.
      while (fm) {
        const int t = __builtin_ctz(fm);
        fm &= fm - 1;
        const uint64_t v = v0 + t;
        J.f_list[pf] = (uint32_t)v;
        J.f_off[pf] = pfc;
        pf++;
        pfc += (uint32_t)leaf_count(g, v);
      }
      while (sm) {
        const int t = __builtin_ctz(sm);
        sm &= sm - 1;
        J.s_node[ps++] = (uint32_t)(v0 + t);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Shifted roots grouped by source checkpoint (a stable counting sort; the
// reference uses Kokkos::BinSort, src/tree_approach.cpp:670-687).
// s_hist is bin-major: entry (bin, tile) at bin * ntiles + tile.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t stiles_of(const JobG& J) { return (J.ctr->nshift + kSTile - 1) / kSTile; }

// Workgroups stride over the sort tiles (kSTile shifted roots each, kSTile /
// 256 per lane).  Each lane loads its roots' nodes, then their map slots, then
// the map entries (three rounds of loads for the whole tile) before any store
// of the tile: a load issued after a store waits for it.
__global__ __launch_bounds__(256) void k_sbin_count(const JobG* __restrict__ jobs, const uint32_t* __restrict__ prefix, uint32_t njobs) {
  constexpr uint32_t Q = kSTile / 256;
  static_assert(kSTile % 256 == 0, "whole rounds of 256 roots per tile");
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobG& J = jobs[j];
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;
  const uint32_t wg = blockIdx.x - prefix[j], nwg = prefix[j + 1] - prefix[j];
  const uint32_t nt = stiles_of(J);
  const uint32_t nbins = J.ckpt_id + 1;
  const uint32_t S = J.ctr->nshift;
  for (uint32_t t = wg; t < nt; t += nwg) {
    const uint32_t e0 = t * kSTile, e1 = min(S, e0 + kSTile);
    // The roots' sources: the map entries their digests matched, as the
    // leaf inserts, pass 2 or the bulk build recorded them: the entry's value
    // when it was final, else its slot (an entry of this checkpoint met by a
    // leaf insert, which a later insert could still lower), read now.
    uint32_t node[Q];
    unsigned long long val[Q];
#pragma unroll
    for (uint32_t q = 0; q < Q; q++) {
      const uint32_t e = e0 + q * 256 + threadIdx.x;
      node[q] = e < e1 ? J.s_node[e] : 0;
    }
#pragma unroll
    for (uint32_t q = 0; q < Q; q++) {
      const uint32_t e = e0 + q * 256 + threadIdx.x;
      val[q] = e < e1 ? J.slot_of[node[q]] : 0;
    }
#pragma unroll
    for (uint32_t q = 0; q < Q; q++) {
      const uint32_t e = e0 + q * 256 + threadIdx.x;
      if (e < e1 && ref_is_slot(val[q])) val[q] = J.map.slots[(uint32_t)val[q]].val;
    }
    __syncthreads();  // the previous tile's atomics are done
    for (uint32_t bn = threadIdx.x; bn < nbins; bn += blockDim.x) J.s_hist[(uint64_t)bn * nt + t] = 0;
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < Q; q++) {
      const uint32_t e = e0 + q * 256 + threadIdx.x;
      bool todo = e < e1;
      uint32_t bn = 0;
      if (todo) {
        J.s_prev[e] = (uint32_t)val[q];
        J.s_tree[e] = (uint32_t)(val[q] >> 32);
        bn = (uint32_t)(val[q] >> 32);
        if (bn >= nbins) {
          g_or(&J.ctr->flags, FLAG_BAD_SOURCE);
          bn = nbins - 1;
        }
      }
      // one atomic per distinct bin of the wavefront (shifted roots mostly
      // share one or two source checkpoints)
      while (true) {
        const unsigned long long act = __ballot(todo);
        if (!act) break;
        const uint32_t leader = __ffsll((long long)act) - 1;
        const uint32_t lb = __shfl(bn, leader, 64);
        const unsigned long long same = __ballot(todo && bn == lb);
        if (lane_id() == leader) g_add(&J.s_hist[(uint64_t)lb * nt + t], (uint32_t)__popcll(same));
        if (todo && bn == lb) todo = false;
      }
    }
  }
}

// Copies the final counters into host-mapped memory (read by the host once
// the stream is synchronized; no copy command needed).
__device__ __forceinline__ void export_counters(const JobG& J) {
  static_assert(sizeof(JobCounters) % 8 == 0, "counters copied as 8-byte words");
  const GlobalPtr<const uint64_t> src = (GlobalPtr<const uint64_t>)J.ctr;
  const GlobalPtr<uint64_t> dst = (GlobalPtr<uint64_t>)J.host_ctr;
  uint64_t w[sizeof(JobCounters) / 8];
#pragma unroll
  for (uint32_t i = 0; i < sizeof(JobCounters) / 8; i++) w[i] = src[i];
#pragma unroll
  for (uint32_t i = 0; i < sizeof(JobCounters) / 8; i++) dst[i] = w[i];
}

// One workgroup per job: new map entries of the checkpoint, scan of the
// shifted-root histogram, the prior table, the diff length; and, once the
// checkpoint needs no re-run, the next checkpoint's work-list counters are
// zeroed (checkpoints alternate between two counter sets, so no kernel has
// to re-arm them after the pipeline).  Two rounds of
// loads: the claim counts with the counter fields, then the histogram, which
// (at most 1024 entries, the usual case) is scanned and turned into per-bin
// totals in LDS.
__global__ __launch_bounds__(1024) void k_sbin_scan(const JobG* __restrict__ jobs) {
  const JobG& J = jobs[blockIdx.x];
  __shared__ uint32_t sh[17];
  __shared__ unsigned long long shl[16];
  __shared__ uint32_t hs[1025];
  __shared__ uint32_t sflags, snshift, snfirst, spend;
  __shared__ unsigned long long sdchunks;
  if (threadIdx.x == 0) {
    sflags = J.ctr->flags;
    snshift = J.ctr->nshift;
    snfirst = J.ctr->nfirst;
    sdchunks = J.ctr->dchunks;
    spend = J.is_baseline ? 0u : (uint32_t)pending_total(J);
  }
  unsigned long long claims = 0;
  for (uint32_t i = threadIdx.x; i < J.nwg_ins_leaf; i += blockDim.x) claims += J.blk_claim[i];
  for (uint32_t i = threadIdx.x; i < J.nwg_ins_int; i += blockDim.x) claims += J.int_claim[i];
  __syncthreads();  // sflags
  if (J.bulk && !(sflags & (FLAG_NEED_GROW | FLAG_MAP_FULL)))
    for (uint32_t i = threadIdx.x; i < J.nbuckets; i += blockDim.x) claims += J.bk_claim[i];
  claims = block_sum(claims, shl);  // its barriers publish the fields above
  const uint64_t pending = J.is_baseline ? J.C : spend;
  if (sflags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) {
    if (threadIdx.x == 0) {
      J.ctr->map_count = J.map_count_at_start + claims;
      J.ctr->pending = pending;
    }
    return;
  }
  // The work-list counters of the next checkpoint (the set this checkpoint
  // does not use) start at zero.
  if (threadIdx.x < kPendSpread) {
    J.pend_cnt_next[threadIdx.x * kCntStride] = 0;
    J.dirty_cnt_next[threadIdx.x * kCntStride] = 0;
  }
  const uint32_t S = snshift;
  const uint32_t nt = (S + kSTile - 1) / kSTile;
  const uint32_t nbins = J.ckpt_id + 1;
  uint32_t nprior = 0;
  if (nt > 0) {
    const uint64_t total = (uint64_t)nbins * nt;
    const bool in_lds = total <= blockDim.x;  // uniform
    // Exclusive scan of the histogram in bin-major order.
    uint32_t run = 0;
    for (uint64_t base = 0; base < total; base += blockDim.x) {
      const uint64_t i = base + threadIdx.x;
      const uint32_t v = i < total ? J.s_hist[i] : 0;
      uint32_t tot;
      const uint32_t e = block_excl_scan<16>(v, sh, tot);
      if (i < total) {
        J.s_hist[i] = run + e;
        if (in_lds) hs[i] = e;
      }
      run += tot;
    }
    __syncthreads();
    // Per-bin totals from the scanned bases -> prior table, ascending.
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nbins; base += blockDim.x) {
      const uint32_t bn = base + threadIdx.x;
      uint32_t cnt = 0;
      if (bn < nbins) {
        const uint64_t i0 = (uint64_t)bn * nt, i1 = (uint64_t)(bn + 1) * nt;
        const uint32_t lo = in_lds ? hs[i0] : J.s_hist[i0];
        const uint32_t hi = bn + 1 < nbins ? (in_lds ? hs[i1] : J.s_hist[i1]) : S;
        cnt = hi - lo;
      }
      uint32_t tot;
      const uint32_t pos = block_excl_scan<16>(cnt ? 1u : 0u, sh, tot) + carry;
      if (bn < nbins && cnt) {
        J.ptab[2 * pos] = bn;
        J.ptab[2 * pos + 1] = cnt;
      }
      carry += tot;
    }
    nprior = carry;
  }
  if (threadIdx.x == 0) {
    J.ctr->nprior = nprior;
    const uint64_t meta = 4ull * snfirst + 8ull * nprior + 8ull * S;
    J.ctr->diff_len = 32 + meta + sdchunks * (uint64_t)J.chunk_size;
    J.ctr->map_count = J.map_count_at_start + claims;
    J.ctr->pending = pending;
  }
}

// One wavefront per tile walks its entries in order; lanes with the same bin
// are ranked with a ballot so equal bins keep node order.  The tile's entries
// (kSTile / 64 per lane) are loaded into registers before the ranking loop,
// which then only touches LDS and issues stores.
constexpr uint32_t kLdsBins = 8192;
constexpr uint32_t kSChunks = kSTile / 64;
__global__ __launch_bounds__(64) void k_sbin_scatter(const JobG* __restrict__ jobs, const uint32_t* __restrict__ prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobG& J = jobs[j];
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;
  const uint32_t wg = blockIdx.x - prefix[j], nwg = prefix[j + 1] - prefix[j];
  const uint32_t nt = stiles_of(J);
  const uint32_t nbins = J.ckpt_id + 1;
  const uint32_t S = J.ctr->nshift;
  const GlobalPtr<const uint32_t> s_tree = J.s_tree;
  const GlobalPtr<const uint32_t> s_node = J.s_node;
  const GlobalPtr<const uint32_t> s_prev = J.s_prev;
  const GlobalPtr<uint32_t> ss_node = J.ss_node;
  const GlobalPtr<uint32_t> ss_prev = J.ss_prev;
  __shared__ uint32_t base[kLdsBins];
  for (uint32_t t = wg; t < nt; t += nwg) {
    const uint32_t e0 = t * kSTile, e1 = min(S, (t + 1) * kSTile);
    uint32_t tb[kSChunks], tn[kSChunks], tp[kSChunks];
#pragma unroll
    for (uint32_t q = 0; q < kSChunks; q++) {
      const uint32_t idx = e0 + q * 64 + threadIdx.x;
      const bool in = idx < e1;
      uint32_t bn = in ? s_tree[idx] : 0xffffffffu;
      if (in && bn >= nbins) bn = nbins - 1;
      tb[q] = bn;
      tn[q] = in ? s_node[idx] : 0u;
      tp[q] = in ? s_prev[idx] : 0u;
    }
    for (uint32_t lo = 0; lo < nbins; lo += kLdsBins) {
      const uint32_t hi = min(nbins, lo + kLdsBins);
      for (uint32_t bn = lo + threadIdx.x; bn < hi; bn += 64) base[bn - lo] = J.s_hist[(uint64_t)bn * nt + t];
      __syncthreads();
#pragma unroll
      for (uint32_t q = 0; q < kSChunks; q++) {
        const uint32_t bn = tb[q];
        bool todo = bn != 0xffffffffu && bn >= lo && bn < hi;
        while (true) {
          const unsigned long long act = __ballot(todo);
          if (!act) break;
          const uint32_t leader = __ffsll((long long)act) - 1;
          const uint32_t lb = __shfl(bn, leader, 64);
          const unsigned long long same = __ballot(todo && bn == lb);
          const uint32_t cur = base[lb - lo];
          if (todo && bn == lb) {
            const uint32_t pos = cur + __popcll(same & lanemask_lt());
            ss_node[pos] = tn[q];
            ss_prev[pos] = tp[q];
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
}

// ---------------------------------------------------------------------------
// Diff metadata: header_t, first-occurrence roots, prior table, shifted
// entries (src/tree_approach.cpp:593,622-706).
// ---------------------------------------------------------------------------
// Written by the gather workgroups (thread tid of nth), after their copies:
// loads and stores share one in-order counter, so the stores go last.
__device__ __forceinline__ void write_meta(const JobG& J, uint32_t F, uint32_t P, uint32_t S, uint32_t tid,
                                           uint32_t nth) {
  const GlobalPtr<uint8_t> out = J.diff;
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
  const GlobalPtr<uint8_t> pf = out + 32;
  const GlobalPtr<uint8_t> pp = pf + 4ull * F;
  const GlobalPtr<uint8_t> ps = pp + 8ull * P;
  const uint32_t bias = J.node_bias;  // list diffs name chunks, not leaves
  for (uint32_t i = tid; i < F; i += nth) store_u32(pf + 4ull * i, J.f_list[i] - bias);
  for (uint32_t i = tid; i < P; i += nth) {
    store_u32(pp + 8ull * i, J.ptab[2 * i]);
    store_u32(pp + 8ull * i + 4, J.ptab[2 * i + 1]);
  }
  for (uint32_t i = tid; i < S; i += nth) {
    store_u32(ps + 8ull * i, J.ss_node[i] - bias);
    store_u32(ps + 8ull * i + 4, J.ss_prev[i] - bias);
  }
}

// ---------------------------------------------------------------------------
// Gather: data chunk q of the data section comes from first region r with
// f_off[r] <= q < f_off[r+1]; its index inside the region picks the leaf in
// left-to-right order.  Copies 4-byte words (the data section is only 4-byte
// aligned: it follows 4-byte first-occurrence ids).
// ---------------------------------------------------------------------------
template <class Ptr>  // a global pointer (gather) or a generic one (restart)
__device__ __forceinline__ uint32_t region_of(Ptr f_off, uint32_t F, uint64_t q) {
  uint32_t lo = 0, hi = F;  // f_off[lo] <= q < f_off[hi] (f_off[F] = total)
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (f_off[mid] <= q) lo = mid;
    else hi = mid;
  }
  return lo;
}

// The data section: groups of 256 chunks per workgroup iteration.
template <bool Vec>
__device__ __forceinline__ void gather_data(const JobG& J, uint32_t F, uint32_t P, uint32_t S, uint64_t Dq,
                                            uint32_t wg, uint32_t nblk) {
  if (Dq == 0) return;
  const TreeGeom g = make_geom(J.C);
  const uint64_t cs = J.chunk_size;
  GlobalPtr<uint8_t> dst_base = J.diff + 32 + 4ull * F + 8ull * P + 8ull * S;
  __shared__ uint64_t src[256];
  const uint64_t ngroups = (Dq + 255) / 256;
  const uint32_t piece = Vec ? 16 : 1;
  const uint64_t ppc = cs / piece;  // pieces per chunk
  const bool pow2 = (ppc & (ppc - 1)) == 0;
  const uint32_t sh = pow2 ? (uint32_t)__builtin_ctzll(ppc) : 0;
  for (uint64_t grp = wg; grp < ngroups; grp += nblk) {
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
    // One piece per lane per step: loading several pieces before storing
    // them saves a few microseconds per incremental checkpoint but slows a
    // baseline's whole-buffer copy by 5 %.
    for (uint64_t w = threadIdx.x; w < total; w += 256) {
      const uint64_t i = pow2 ? (w >> sh) : (w / ppc);
      const uint64_t pi = w - i * ppc;
      const uint64_t so = src[i] * cs + pi * piece;
      store_piece<Vec>(dst_base + (q0 + i) * cs + pi * piece, load_piece<Vec>(J.data + so, so < J.len ? J.len - so : 0));
    }
    __syncthreads();
  }
}

// Groups of 256 data chunks: each lane finds the source chunk of one, then the
// workgroup copies the group with 16-byte pieces (1-byte pieces when the
// chunk size is not a multiple of 16 or the input is not 16-byte aligned).
template <bool Vec>
__global__ __launch_bounds__(256) void k_gather(const JobG* __restrict__ jobs, const uint32_t* __restrict__ prefix, uint32_t njobs) {
  const uint32_t j = find_job(prefix, njobs, blockIdx.x);
  const JobG& J = jobs[j];
  // The counters are final: the job's first workgroup copies them to the
  // host-mapped block (its slow stores overlap the copy below), also for a
  // checkpoint that must re-run.
  if (blockIdx.x == prefix[j] && threadIdx.x == 0) export_counters(J);
  if (J.ctr->flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return;
  const uint32_t F = J.ctr->nfirst, P = J.ctr->nprior, S = J.ctr->nshift;
  const uint64_t Dq = J.ctr->dchunks;
  const uint32_t wg = blockIdx.x - prefix[j], nblk = prefix[j + 1] - prefix[j];
  // The leaf kernel already copied the data section when the root is the
  // only region (the one region that holds all C chunks).
  const bool copied = J.write_data && F == 1 && P == 0 && S == 0 && Dq == J.C;
  if (!copied) gather_data<Vec>(J, F, P, S, Dq, wg, nblk);
  // Stores only from here on: the metadata.
  write_meta(J, F, P, S, wg * blockDim.x + threadIdx.x, nblk * blockDim.x);
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
    uint64_t s = map_start(m, k);
    bool done = false;
    for (uint64_t it = 0; it <= m.bmask; it++) {
      if (atomicCAS(&m.slots[s].val, SLOT_EMPTY, o.val) == SLOT_EMPTY) {
        m.slots[s].k0 = k.k0;
        m.slots[s].k1 = k.k1;
        done = true;
        break;
      }
      s = map_next(m, s, 1);
    }
    if (!done) atomicOr(flags, FLAG_MAP_FULL);
  }
}

// ---------------------------------------------------------------------------
// Restart (src/tree_approach.cpp:712-1128, include/reference_impl.hpp's
// restart semantics as restated by the oracle).
//
// The reference walks the chain newest to oldest and, for every older
// checkpoint, passes over all C chunks to see which still need data from it
// (:1083-1118).  Here the walk touches only what each checkpoint carries:
// - A chunk that no newer checkpoint re-sourced takes its data from the
//   newest checkpoint whose regions cover it.  So each step enumerates the
//   leaves of its own regions (k_restart_regions) and claims those that no
//   newer step covered; the cover table doubles as the claim record (its
//   stamps grow by one per step and per restart, so no clearing is needed).
// - A claimed leaf of a shifted region names a source (checkpoint, chunk);
//   the chunk goes on a pending list, and the step of the source checkpoint
//   resolves it from that step's cover entries (k_restart_pending).
// Work per restart is the regions' leaves plus the pending chunks, not
// C x chain length; the data moves once, 256 chunks per workgroup round in
// 16-byte pieces.
// ---------------------------------------------------------------------------
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));  // a cover entry {stamp, region}

template <int PIECE>
__device__ __forceinline__ void copy_chunk_piece(uint8_t* __restrict__ out, uint64_t dst, const uint8_t* src,
                                                 uint64_t len) {
  GlobalPtr<uint8_t> o = (GlobalPtr<uint8_t>)out;
  GlobalPtr<const uint8_t> in = (GlobalPtr<const uint8_t>)src;
  if (dst + PIECE <= len) {
    if (PIECE == 16) {
      // the source is 4-byte aligned (data sections follow 4-byte ids)
      const u32x4a4 v = *(GlobalPtr<const u32x4a4>)in;
      *(GlobalPtr<u32x4_t>)(o + dst) = u32x4_t{v.x, v.y, v.z, v.w};
    } else if (PIECE == 4) {
      *(GlobalPtr<uint32_t>)(o + dst) = *(GlobalPtr<const uint32_t>)in;
    } else {
      o[dst] = in[0];
    }
  } else {
    for (uint64_t b = dst; b < len; b++) o[b] = in[b - dst];
  }
}

// Copies the chunks staged in LDS by one workgroup round: chunk i of the
// round goes from data + src[i] to out + dst[i] * cs (src[i] == ~0: none).
template <int PIECE>
__device__ __forceinline__ void copy_round(const uint64_t* s_src, const uint32_t* s_dst, const uint8_t* data,
                                           uint8_t* out, uint64_t len, uint32_t cs) {
  const uint32_t ppc = cs / PIECE;  // pieces per chunk
  const bool pow2 = (ppc & (ppc - 1)) == 0;
  const uint32_t sh = pow2 ? __builtin_ctz(ppc) : 0;
  const uint32_t total = 256 * ppc;
  if (PIECE == 16) {
    // Four 16-byte pieces per lane loaded before any is stored (a load
    // issued after a store waits for it).
    GlobalPtr<uint8_t> o = (GlobalPtr<uint8_t>)out;
    constexpr uint32_t U = 4;
    for (uint32_t w0 = threadIdx.x; w0 < total; w0 += 256 * U) {
      u32x4_t v[U];
      uint64_t dst[U];
      bool full[U];
#pragma unroll
      for (uint32_t u = 0; u < U; u++) {
        const uint32_t w = w0 + 256 * u;
        full[u] = false;
        dst[u] = 0;
        if (w >= total) continue;
        const uint32_t i = pow2 ? (w >> sh) : (w / ppc);
        const uint64_t so = s_src[i];
        if (so == ~0ull) continue;
        const uint32_t pi = w - i * ppc;
        dst[u] = (uint64_t)s_dst[i] * cs + (uint64_t)pi * 16;
        const GlobalPtr<const uint8_t> in = (GlobalPtr<const uint8_t>)(data + so + (uint64_t)pi * 16);
        if (dst[u] + 16 <= len) {
          const u32x4a4 x = *(GlobalPtr<const u32x4a4>)in;  // 4-byte aligned source
          v[u] = u32x4_t{x.x, x.y, x.z, x.w};
          full[u] = true;
        } else {
          copy_chunk_piece<16>(out, dst[u], data + so + (uint64_t)pi * 16, len);  // the tail of the buffer
        }
      }
#pragma unroll
      for (uint32_t u = 0; u < U; u++)
        if (full[u]) *(GlobalPtr<u32x4_t>)(o + dst[u]) = v[u];
    }
    return;
  }
  for (uint32_t w = threadIdx.x; w < total; w += 256) {
    const uint32_t i = pow2 ? (w >> sh) : (w / ppc);
    const uint64_t so = s_src[i];
    if (so == ~0ull) continue;
    const uint32_t pi = w - i * ppc;
    copy_chunk_piece<PIECE>(out, (uint64_t)s_dst[i] * cs + (uint64_t)pi * PIECE, data + so + (uint64_t)pi * PIECE, len);
  }
}

// Position of leaf `leaf` in the left-to-right enumeration of a subtree.
__device__ __forceinline__ uint64_t run_index(const LeafRuns& r, uint64_t leaf) {
  return (leaf >= r.a0 && leaf < r.a0 + r.n0) ? leaf - r.a0 : r.n0 + (leaf - r.a1);
}

__device__ __forceinline__ uint32_t rs_load32(const uint8_t* p) {
  return *(GlobalPtr<const uint32_t>)(GlobalPtr<const uint8_t>)p;
}

// Data offset (in the data section) of chunk `c` of first-occurrence region r.
__device__ __forceinline__ uint64_t first_src(const RestartCkpt& rc, const TreeGeom& g, uint32_t r, uint64_t c,
                                              uint32_t cs) {
  const uint64_t root = (uint64_t)rs_load32(rc.diff + rc.first_off + 4ull * r) + rc.node_bias;
  const uint64_t idx = run_index(subtree_runs(g, root), c + g.C - 1);
  return ((uint64_t)rc.f_off[r] + idx) * cs;
}

// Source chunk of chunk `c` in shifted region r: the leaf at the same
// position of the region's source subtree.
__device__ __forceinline__ uint64_t shift_src(const RestartCkpt& rc, const TreeGeom& g, uint32_t r, uint64_t c) {
  const uint64_t node = (uint64_t)rs_load32(rc.diff + rc.shift_off + 8ull * r) + rc.node_bias;
  const uint64_t prev = (uint64_t)rs_load32(rc.diff + rc.shift_off + 8ull * r + 4) + rc.node_bias;
  const uint64_t idx = run_index(subtree_runs(g, node), c + g.C - 1);
  return run_leaf(subtree_runs(g, prev), idx) - (g.C - 1);
}

// Appends the workgroup's pushed chunks to the pending list (capacity `cap`)
// with one atomic per workgroup round (a counter hit once per wavefront
// serializes hundreds of same-address atomics per step).  Each chunk is
// claimed once per restart, so the list cannot pass C entries unless a
// diff's regions overlap; such a diff is flagged.  `sh` holds 5 words of LDS.
__device__ __forceinline__ void pend_push(bool push, uint32_t c, uint32_t* pend, uint64_t cap, RestartCtr* ctr,
                                          uint32_t* sh) {
  const unsigned long long m = __ballot(push);
  const uint32_t w = threadIdx.x >> 6;
  if (lane_id() == 0) sh[w] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t n = sh[0] + sh[1] + sh[2] + sh[3];
    sh[4] = n ? g_add((GlobalPtr<uint32_t>)&ctr->npend, n) : 0u;
  }
  __syncthreads();
  if (push) {
    uint32_t before = 0;
    for (uint32_t i = 0; i < w; i++) before += sh[i];
    const uint64_t at = (uint64_t)sh[4] + before + __popcll(m & lanemask_lt());
    if (at < cap) ((GlobalPtr<uint32_t>)pend)[at] = c;
    else g_or((GlobalPtr<uint32_t>)&ctr->flags, FLAG_BAD_SOURCE);
  }
}

// Source checkpoint of entry r of a shifted list (cumulative prior-table
// counts, src/tree_approach.cpp:868-898): the first j with r < pb_end[j].
template <class Ptr>
__device__ __forceinline__ uint32_t prior_search(Ptr pb_end, Ptr pb_ck, uint32_t npb, uint32_t r) {
  uint32_t lo = 0, hi = npb;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pb_end[mid] <= r) lo = mid + 1;
    else hi = mid;
  }
  return pb_ck[lo < npb ? lo : npb - 1];
}
__device__ __forceinline__ uint32_t prior_of(const RestartCkpt& rc, uint32_t r) {
  return prior_search((GlobalPtr<const uint32_t>)rc.pb_end, (GlobalPtr<const uint32_t>)rc.pb_ck, rc.npb, r);
}

// The prior boundaries of a checkpoint in LDS (up to kPriorLds of them).
constexpr uint32_t kPriorLds = 128;
struct PriorLds {
  uint32_t end[kPriorLds], ck[kPriorLds];
};
__device__ __forceinline__ void prior_to_lds(const RestartCkpt& rc, PriorLds& p) {
  if (rc.npb > kPriorLds) return;
  for (uint32_t i = threadIdx.x; i < rc.npb; i += blockDim.x) {
    p.end[i] = ((GlobalPtr<const uint32_t>)rc.pb_end)[i];
    p.ck[i] = ((GlobalPtr<const uint32_t>)rc.pb_ck)[i];
  }
}
__device__ __forceinline__ uint32_t prior_from(const RestartCkpt& rc, const PriorLds& p, uint32_t r) {
  if (rc.npb <= kPriorLds) return prior_search(p.end, p.ck, rc.npb, r);
  return prior_of(rc, r);
}

__device__ __forceinline__ uint64_t leaves_of(const TreeGeom& g, uint64_t v) {
  const LeafRuns r = subtree_runs(g, v);
  return r.n0 + r.n1;
}

// Parse, step 1: each block of 1024 entries of a list counts its regions'
// leaves, checks the node ids, and writes the block-local exclusive prefix
// into f_off / s_off and its total into blk_sum.
__global__ __launch_bounds__(256) void k_rs_count(RestartCkpt* __restrict__ desc, const RestartList* __restrict__ lists,
                                                  uint32_t nlists, const uint32_t* __restrict__ list_blk,
                                                  unsigned long long* __restrict__ blk_sum, TreeGeom g,
                                                  RestartCtr* __restrict__ ctr) {
  __shared__ unsigned long long sh[5];
  const uint32_t L = find_job(list_blk, nlists, blockIdx.x);
  const RestartList li = lists[L];
  const RestartCkpt& rc = desc[li.ckpt];
  const uint64_t e0 = (uint64_t)(blockIdx.x - list_blk[L]) * kParseBlock + threadIdx.x * 4;
  uint64_t cnt[4];
  bool bad = false;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const uint64_t e = e0 + j;
    cnt[j] = 0;
    if (e >= li.n) continue;
    if (!li.shift) {
      const uint64_t node = (uint64_t)rs_load32(rc.diff + rc.first_off + 4 * e) + rc.node_bias;
      if (node >= g.N) bad = true;
      else cnt[j] = leaves_of(g, node);
    } else {
      const uint64_t node = (uint64_t)rs_load32(rc.diff + rc.shift_off + 8 * e) + rc.node_bias;
      const uint64_t prev = (uint64_t)rs_load32(rc.diff + rc.shift_off + 8 * e + 4) + rc.node_bias;
      if (node >= g.N || prev >= g.N) {
        bad = true;
      } else {
        cnt[j] = leaves_of(g, node);
        if (leaves_of(g, prev) != cnt[j]) bad = true;
      }
    }
  }
  const unsigned long long t = cnt[0] + cnt[1] + cnt[2] + cnt[3];
  unsigned long long total;
  unsigned long long x = block_excl_scan<4>(t, sh, total);
  GlobalPtr<uint32_t> off = (GlobalPtr<uint32_t>)(li.shift ? rc.s_off : rc.f_off);
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (e0 + j < li.n) off[e0 + j] = (uint32_t)x;
    x += cnt[j];
  }
  if (threadIdx.x == 0) ((GlobalPtr<unsigned long long>)blk_sum)[blockIdx.x] = total;
  if (bad) g_or((GlobalPtr<uint32_t>)&ctr->flags, FLAG_BAD_DIFF);
}

// Parse, step 2: one workgroup per list scans its block totals into block
// offsets, stores the list's leaf total (nfl / nsl and the prefix's last
// entry) and checks it against the chunk count and the data section.
__global__ __launch_bounds__(256) void k_rs_scan(RestartCkpt* __restrict__ desc, const RestartList* __restrict__ lists,
                                                 unsigned long long* __restrict__ blk_sum, TreeGeom g, uint32_t cs,
                                                 RestartCtr* __restrict__ ctr) {
  __shared__ unsigned long long sh[5];
  const RestartList li = lists[blockIdx.x];
  RestartCkpt& rc = desc[li.ckpt];
  const uint32_t nb = (li.n + kParseBlock - 1) / kParseBlock;
  GlobalPtr<unsigned long long> bs = (GlobalPtr<unsigned long long>)blk_sum + li.blk0;
  unsigned long long run = 0;
  for (uint32_t b0 = 0; b0 < nb; b0 += 256) {
    const uint32_t b = b0 + threadIdx.x;
    const unsigned long long v = b < nb ? bs[b] : 0ull;
    unsigned long long tot;
    const unsigned long long x = block_excl_scan<4>(v, sh, tot);
    if (b < nb) bs[b] = run + x;
    run += tot;
  }
  if (threadIdx.x == 0) {
    bool bad = run > g.C;
    if (!li.shift) {
      bad = bad || run * cs > rc.data_bytes;
      ((GlobalPtr<uint32_t>)rc.f_off)[li.n] = (uint32_t)run;
      rc.nfl = run;
    } else {
      ((GlobalPtr<uint32_t>)rc.s_off)[li.n] = (uint32_t)run;
      rc.nsl = run;
    }
    if (bad) g_or((GlobalPtr<uint32_t>)&ctr->flags, FLAG_BAD_DIFF);
  }
}

// Parse, step 3: block offsets added to the block-local prefixes.
__global__ __launch_bounds__(256) void k_rs_add(const RestartCkpt* __restrict__ desc,
                                                const RestartList* __restrict__ lists, uint32_t nlists,
                                                const uint32_t* __restrict__ list_blk,
                                                const unsigned long long* __restrict__ blk_sum) {
  const uint32_t L = find_job(list_blk, nlists, blockIdx.x);
  const uint32_t add = (uint32_t)((GlobalPtr<const unsigned long long>)blk_sum)[blockIdx.x];
  if (add == 0) return;
  const RestartList li = lists[L];
  const RestartCkpt& rc = desc[li.ckpt];
  GlobalPtr<uint32_t> off = (GlobalPtr<uint32_t>)(li.shift ? rc.s_off : rc.f_off);
  const uint64_t e0 = (uint64_t)(blockIdx.x - list_blk[L]) * kParseBlock + threadIdx.x * 4;
#pragma unroll
  for (int j = 0; j < 4; j++)
    if (e0 + j < li.n) off[e0 + j] += add;
}

// Parse, step 4: for every round of 256 leaves of a list, the region that
// holds its first leaf (the walk then finds each leaf's region among the next
// 256 entries in LDS, not by a search in global memory).
__global__ __launch_bounds__(256) void k_rs_index(const RestartCkpt* __restrict__ desc,
                                                  const RestartList* __restrict__ lists,
                                                  RestartCtr* __restrict__ ctr) {
  if (((GlobalPtr<const uint32_t>)&ctr->flags)[0] & FLAG_BAD_DIFF) return;
  const RestartList li = lists[blockIdx.y];
  const RestartCkpt& rc = desc[li.ckpt];
  const uint64_t nl = li.shift ? rc.nsl : rc.nfl;
  const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (t * 256 >= nl) return;
  const GlobalPtr<const uint32_t> off = (GlobalPtr<const uint32_t>)(li.shift ? rc.s_off : rc.f_off);
  const GlobalPtr<uint32_t> idx = (GlobalPtr<uint32_t>)(li.shift ? rc.s_idx : rc.f_idx);
  idx[t] = region_of(off, li.n, t * 256);
}

template <int PIECE>
__global__ __launch_bounds__(256) void k_restart_regions(const RestartCkpt* __restrict__ rcp, TreeGeom g,
                                                         uint2* __restrict__ cover, uint32_t base, uint32_t stamp,
                                                         uint64_t* __restrict__ ref, uint32_t* __restrict__ pend,
                                                         RestartCtr* __restrict__ ctr, uint8_t* __restrict__ out,
                                                         uint64_t len, uint32_t cs) {
  const RestartCkpt rc = *rcp;
  if (((GlobalPtr<const uint32_t>)&ctr->flags)[0] & FLAG_BAD_DIFF) return;
  __shared__ uint64_t s_src[256];
  __shared__ uint32_t s_dst[256];
  __shared__ uint32_t s_off[257];
  __shared__ uint32_t s_cnt[5];
  __shared__ PriorLds s_pb;
  prior_to_lds(rc, s_pb);
  // Rounds of 256 leaves: first the first-occurrence list's, then the shifted list's.
  const uint64_t rf = (rc.nfl + 255) / 256, rs = (rc.nsl + 255) / 256;
  const GlobalPtr<u32x2_t> cv = (GlobalPtr<u32x2_t>)cover;
  uint32_t claimed = 0;
  for (uint64_t t = blockIdx.x; t < rf + rs; t += gridDim.x) {
    const bool sh = t >= rf;
    const uint64_t tl = sh ? t - rf : t;
    const uint64_t nl = sh ? rc.nsl : rc.nfl;
    const uint32_t nreg = sh ? rc.nshift : rc.nfirst;
    const GlobalPtr<const uint32_t> off = (GlobalPtr<const uint32_t>)(sh ? rc.s_off : rc.f_off);
    const uint32_t r0 = ((GlobalPtr<const uint32_t>)(sh ? rc.s_idx : rc.f_idx))[tl];
    // Prefix entries r0 .. r0+256 (past the list: above every leaf index).
    for (uint32_t i = threadIdx.x; i < 257; i += 256) s_off[i] = (uint64_t)r0 + i <= nreg ? off[r0 + i] : 0xffffffffu;
    __syncthreads();
    const uint64_t e = tl * 256 + threadIdx.x;
    uint64_t so = ~0ull;
    uint32_t c = 0;
    bool push = false;
    if (e < nl) {
      uint32_t lo = 0, hi = 256;  // s_off[lo] <= e < s_off[hi]
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_off[mid] <= e) lo = mid;
        else hi = mid;
      }
      const uint32_t r = r0 + lo;
      const uint64_t idx = e - s_off[lo];
      uint64_t leaf, src = 0;
      uint32_t info, src_ck = 0;
      if (!sh) {
        const uint64_t root = (uint64_t)rs_load32(rc.diff + rc.first_off + 4ull * r) + rc.node_bias;
        leaf = run_leaf(subtree_runs(g, root), idx);
        info = r;
      } else {
        const uint64_t node = (uint64_t)rs_load32(rc.diff + rc.shift_off + 8ull * r) + rc.node_bias;
        const uint64_t prev = (uint64_t)rs_load32(rc.diff + rc.shift_off + 8ull * r + 4) + rc.node_bias;
        leaf = run_leaf(subtree_runs(g, node), idx);
        src = run_leaf(subtree_runs(g, prev), idx) - (g.C - 1);
        src_ck = prior_from(rc, s_pb, r);
        info = r | kCoverShift;
      }
      c = (uint32_t)(leaf - (g.C - 1));
      const u32x2_t old = cv[c];
      cv[c] = u32x2_t{stamp, info};
      if (old.x <= base) {  // no newer checkpoint of this restart covered the chunk
        claimed++;
        if (!sh) {
          so = e * cs;  // first-occurrence data is stored in region-leaf order
        } else {
          push = true;
          ((GlobalPtr<uint64_t>)ref)[c] =
              ((uint64_t)src_ck << 32) | (src_ck == rc.ckpt_id ? kRefSameCkpt : 0u) | (uint32_t)src;
        }
      }
    }
    pend_push(push, c, pend, g.C, ctr, s_cnt);
    s_src[threadIdx.x] = so;
    s_dst[threadIdx.x] = c;
    __syncthreads();
    copy_round<PIECE>(s_src, s_dst, rc.diff + rc.data_off, out, len, cs);
    __syncthreads();
  }
  const uint32_t tot = wave_sum(claimed);
  if (lane_id() == 0 && tot) g_add((GlobalPtr<unsigned long long>)&ctr->claimed, (unsigned long long)tot);
}

// Step of a checkpoint whose diff holds one first-occurrence region and no
// shifted one (a baseline without repeated chunks).  The region's leaves are
// at most two runs of consecutive chunks, stored in that order in the data
// section, so a round of 256 leaves reads one contiguous range; no region
// search and no root load per leaf.  The round's pieces are loaded together
// with its cover entries, before the claims are known (a claimed chunk's
// pieces are read and dropped), so a round costs one memory round trip and
// the stores.  Cover entries and claims as in k_restart_regions (region 0);
// the walk's last step (no_cover) writes no cover entries: no older step
// reads them, and its pending step tests membership in region 0 instead.
__global__ __launch_bounds__(256) void k_restart_whole(const RestartCkpt* __restrict__ rcp, TreeGeom g,
                                                       uint2* __restrict__ cover, uint32_t base, uint32_t stamp,
                                                       RestartCtr* __restrict__ ctr, uint8_t* __restrict__ out,
                                                       uint64_t len, uint32_t cs) {
  const RestartCkpt rc = *rcp;
  if (((GlobalPtr<const uint32_t>)&ctr->flags)[0] & FLAG_BAD_DIFF) return;
  __shared__ uint32_t s_dst[256];  // chunk of round leaf i; ~0 when a newer checkpoint covers it
  const uint64_t root = (uint64_t)rs_load32(rc.diff + rc.first_off) + rc.node_bias;
  const LeafRuns runs = subtree_runs(g, root);
  const uint64_t nl = min(rc.nfl, runs.n0 + runs.n1);  // equal once the parse accepted the diff
  const GlobalPtr<u32x2_t> cv = (GlobalPtr<u32x2_t>)cover;
  const GlobalPtr<uint8_t> o = (GlobalPtr<uint8_t>)out;
  const uint8_t* data = rc.diff + rc.data_off;
  const uint32_t ppc = cs / 16;
  const bool pow2 = (ppc & (ppc - 1)) == 0;
  const uint32_t sh = pow2 ? __builtin_ctz(ppc) : 0;
  const bool keep = rc.no_cover == 0;
  constexpr uint32_t U = 8;
  uint32_t claimed = 0;
  for (uint64_t t = blockIdx.x; t * 256 < nl; t += gridDim.x) {
    const uint64_t e = t * 256 + threadIdx.x;
    const uint64_t leaf = e < nl ? run_leaf(runs, e) : 0;
    const bool valid = e < nl && leaf >= g.C - 1 && leaf < g.N;
    const uint32_t cc = valid ? (uint32_t)(leaf - (g.C - 1)) : 0u;
    const u32x2_t old = valid ? (u32x2_t)cv[cc] : u32x2_t{0xffffffffu, 0u};
    const uint32_t nq = (uint32_t)min((uint64_t)256, nl - t * 256);
    const uint32_t total = nq * ppc;
    const uint8_t* src0 = data + t * 256 * (uint64_t)cs;
    u32x4_t v[U];
    auto load_pass = [&](uint32_t w0) {
#pragma unroll
      for (uint32_t u = 0; u < U; u++) {
        const uint32_t w = w0 + 256 * u;
        v[u] = u32x4_t{0u, 0u, 0u, 0u};
        if (w < total) {
          const u32x4a4 x = *(GlobalPtr<const u32x4a4>)(src0 + (uint64_t)w * 16);  // 4-byte aligned source
          v[u] = u32x4_t{x.x, x.y, x.z, x.w};
        }
      }
    };
    auto store_pass = [&](uint32_t w0) {
#pragma unroll
      for (uint32_t u = 0; u < U; u++) {
        const uint32_t w = w0 + 256 * u;
        if (w >= total) continue;
        const uint32_t i = pow2 ? (w >> sh) : (w / ppc);
        const uint32_t cd = s_dst[i];
        if (cd == 0xffffffffu) continue;
        const uint64_t dst = (uint64_t)cd * cs + (uint64_t)(w - i * ppc) * 16;
        if (dst + 16 <= len) *(GlobalPtr<u32x4_t>)(o + dst) = v[u];
        else copy_chunk_piece<16>(out, dst, src0 + (uint64_t)w * 16, len);  // the tail of the buffer
      }
    };
    // Passes of 256 x U pieces (one for chunks of at most 128 bytes); the
    // first one's loads are in flight together with the cover entry.
    const uint32_t npass = (total + 256 * U - 1) / (256 * U);
    for (uint32_t p = 0; p < npass; p++) {  // uniform
      const uint32_t w0 = threadIdx.x + p * 256 * U;
      load_pass(w0);
      if (p == 0) {
        const bool own = valid && old.x <= base;  // no newer checkpoint of this restart covered the chunk
        if (valid && keep) cv[cc] = u32x2_t{stamp, 0u};
        claimed += own ? 1u : 0u;
        s_dst[threadIdx.x] = own ? cc : 0xffffffffu;
        __syncthreads();
      }
      store_pass(w0);
    }
    __syncthreads();
  }
  const uint32_t tot = wave_sum(claimed);
  if (lane_id() == 0 && tot) g_add((GlobalPtr<unsigned long long>)&ctr->claimed, (unsigned long long)tot);
}

// Cover entry of chunk x for a pending step: the table's, or for a step that
// wrote none (no_cover: one first-occurrence region, r0 its leaf runs)
// membership in that region.
__device__ __forceinline__ u32x2_t pending_cover(const RestartCkpt& rc, const TreeGeom& g,
                                                 GlobalPtr<const u32x2_t> cv, uint32_t stamp, const LeafRuns& r0,
                                                 uint64_t x) {
  if (!rc.no_cover) return cv[x];
  const uint64_t leaf = x + g.C - 1;
  const bool in = (leaf >= r0.a0 && leaf - r0.a0 < r0.n0) || (leaf >= r0.a1 && leaf - r0.a1 < r0.n1);
  return in ? u32x2_t{stamp, 0u} : u32x2_t{0u, 0u};
}

template <int PIECE>
__global__ __launch_bounds__(256) void k_restart_pending(const RestartCkpt* __restrict__ rcp, TreeGeom g,
                                                         const uint2* __restrict__ cover, uint32_t stamp,
                                                         uint64_t* __restrict__ ref, const uint32_t* __restrict__ pend,
                                                         RestartCtr* __restrict__ ctr, uint8_t* __restrict__ out,
                                                         uint64_t len, uint32_t cs) {
  const RestartCkpt rc = *rcp;
  if (((GlobalPtr<const uint32_t>)&ctr->flags)[0] & FLAG_BAD_DIFF) return;
  __shared__ uint64_t s_src[256];
  __shared__ uint32_t s_dst[256];
  __shared__ PriorLds s_pb;
  prior_to_lds(rc, s_pb);
  __syncthreads();
  const uint64_t n = min((uint64_t)((GlobalPtr<const uint32_t>)&ctr->npend)[0], g.C);
  const uint32_t cur = rc.ckpt_id;
  const GlobalPtr<const u32x2_t> cv = (GlobalPtr<const u32x2_t>)cover;
  const GlobalPtr<uint64_t> rf = (GlobalPtr<uint64_t>)ref;
  LeafRuns r0{0, 0, 0, 0};
  if (rc.no_cover) r0 = subtree_runs(g, (uint64_t)rs_load32(rc.diff + rc.first_off) + rc.node_bias);
  uint32_t done = 0, bad = 0;
  for (uint64_t i0 = (uint64_t)blockIdx.x * 256; i0 < n; i0 += (uint64_t)gridDim.x * 256) {
    const uint64_t i = i0 + threadIdx.x;
    uint64_t so = ~0ull;
    uint32_t c = 0;
    if (i < n) {
      c = ((GlobalPtr<const uint32_t>)pend)[i];
      const uint64_t r = rf[c];
      if (r != kRefDone && (uint32_t)(r >> 32) == cur) {
        const bool same = ((uint32_t)r & kRefSameCkpt) != 0;
        const uint64_t x = (uint32_t)r & ~kRefSameCkpt;
        const u32x2_t e = pending_cover(rc, g, cv, stamp, r0, x);
        if (e.x == stamp && !(e.y & kCoverShift)) {
          so = first_src(rc, g, e.y, x, cs);
        } else if (e.x == stamp && !same) {
          const uint32_t sr = e.y & ~kCoverShift;
          const uint32_t s2 = prior_from(rc, s_pb, sr);
          const uint64_t x2 = shift_src(rc, g, sr, x);
          if (s2 < cur) {
            rf[c] = ((uint64_t)s2 << 32) | (uint32_t)x2;
          } else {  // a source in this checkpoint must be a first occurrence here
            const u32x2_t e2 = pending_cover(rc, g, cv, stamp, r0, x2);
            if (e2.x == stamp && !(e2.y & kCoverShift)) so = first_src(rc, g, e2.y, x2, cs);
            else bad = 1;
          }
        } else if (same) {
          bad = 1;
        } else if (cur > 0) {
          rf[c] = ((uint64_t)(cur - 1) << 32) | (uint32_t)x;  // not here: same chunk, one checkpoint back
        }
        if (so != ~0ull) {
          rf[c] = kRefDone;
          done++;
        }
      }
    }
    s_src[threadIdx.x] = so;
    s_dst[threadIdx.x] = c;
    __syncthreads();
    copy_round<PIECE>(s_src, s_dst, rc.diff + rc.data_off, out, len, cs);
    __syncthreads();
  }
  const uint32_t tot = wave_sum(done);
  if (lane_id() == 0 && tot) g_add((GlobalPtr<uint32_t>)&ctr->ndone, tot);
  if (bad) g_or((GlobalPtr<uint32_t>)&ctr->flags, FLAG_BAD_SOURCE);
}

// Copies n bytes from src to d with the launch's threads of one range.
__device__ __forceinline__ void fetch_copy(GlobalPtr<const uint8_t> src, GlobalPtr<uint8_t> d, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  if (((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(d) | n) & 3) == 0) {
    for (uint64_t w = (uint64_t)blockIdx.x * 256 + threadIdx.x; w < n / 4; w += stride)
      ((GlobalPtr<uint32_t>)d)[w] = ((GlobalPtr<const uint32_t>)src)[w];
  } else {
    for (uint64_t b = (uint64_t)blockIdx.x * 256 + threadIdx.x; b < n; b += stride) d[b] = src[b];
  }
}

// Copies up to kFetchRanges device byte ranges into dst (blockIdx.y: range);
// in table mode, each diff's header and prior table.
__global__ __launch_bounds__(256) void k_fetch(FetchArgs a, uint8_t* __restrict__ dst) {
  const uint32_t r = blockIdx.y;
  if (r >= a.n) return;
  const GlobalPtr<const uint8_t> src = (GlobalPtr<const uint8_t>)a.src[r];
  const GlobalPtr<uint8_t> d = (GlobalPtr<uint8_t>)dst + a.dst_off[r];
  if (!a.tables) {
    fetch_copy(src, d, a.len[r]);
    return;
  }
  // Header words 5 and 6: num_first_ocur, num_prior_chkpts (byte loads: a
  // diff need not be 4-byte aligned).
  auto word = [&](uint32_t i) {
    return (uint32_t)src[4 * i] | ((uint32_t)src[4 * i + 1] << 8) | ((uint32_t)src[4 * i + 2] << 16) |
           ((uint32_t)src[4 * i + 3] << 24);
  };
  const uint64_t at = 32 + 4ull * word(5);
  uint64_t tb = 8ull * word(6);
  tb = min(tb, a.len[r] - 32);
  tb = at < a.lim[r] ? min(tb, a.lim[r] - at) : 0;
  fetch_copy(src, d, 32);
  fetch_copy(src + at, d + 32, tb);
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    ((GlobalPtr<uint32_t>)p)[i] = v;
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
#define DD_LAUNCH(kern, grid, block, lds, s, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds, s, __VA_ARGS__)

// The kernels read the host's job table through global-address-space
// pointers (same layout).
static const JobG* gview(const JobDev* j) { return reinterpret_cast<const JobG*>(j); }

void launch_begin(const JobDev* jobs_h, const uint32_t*, uint32_t njobs, bool rerun, hipStream_t s) {
  const JobG* jobs = gview(jobs_h);
  DD_LAUNCH(k_begin, (njobs + 63) / 64, 64, 0, s, jobs, njobs, rerun ? 1 : 0);
}

// With timing events the leaf kernel is launched with them attached to its
// own dispatch (hipExtLaunchKernelGGL) instead of separate event records
// around it, which cost the host tens of microseconds per checkpoint.
#define DD_LEAF_LAUNCH_T(kern, threads)                                                                    \
  do {                                                                                                     \
    if (ev_start) hipExtLaunchKernelGGL(kern, dim3(grid), dim3(threads), 0, s, ev_start, ev_stop, 0, jobs, \
                                        prefix, njobs);                                                    \
    else DD_LAUNCH(kern, grid, threads, 0, s, jobs, prefix, njobs);                                        \
    return;                                                                                                \
  } while (0)
#define DD_LEAF_LAUNCH(kern) DD_LEAF_LAUNCH_T(kern, 256)

void launch_leaf(const JobDev* jobs_h, const uint32_t* prefix, uint32_t njobs, uint32_t grid, int mapping, bool al16,
                 int kind, hipStream_t s, hipEvent_t ev_start, hipEvent_t ev_stop, bool copy) {
  const JobG* jobs = gview(jobs_h);
  if (grid == 0) return;
  if (!al16) mapping = LEAF_DIRECT;  // the other mappings need 16-byte aligned buffers
  copy = copy && leaf_writes_data(mapping);
  if (kind == HASH_MD5) {
    switch (mapping) {
      case LEAF_LDS64:
        if (copy) DD_LEAF_LAUNCH((k_leaf_hash_lds<HASH_MD5, 64, false, true>));
        DD_LEAF_LAUNCH((k_leaf_hash_lds<HASH_MD5, 64, false, false>));
      case LEAF_LDS128:
        if (copy) DD_LEAF_LAUNCH((k_leaf_hash_lds<HASH_MD5, 128, false, true>));
        DD_LEAF_LAUNCH((k_leaf_hash_lds<HASH_MD5, 128, false, false>));
      case LEAF_LDS256:
        if (copy) DD_LEAF_LAUNCH((k_leaf_hash_lds<HASH_MD5, 256, false, true>));
        DD_LEAF_LAUNCH((k_leaf_hash_lds<HASH_MD5, 256, false, false>));
      case LEAF_LDS128P:
        if (copy) DD_LEAF_LAUNCH((k_leaf_hash_lds<HASH_MD5, 128, true, true>));
        DD_LEAF_LAUNCH((k_leaf_hash_lds<HASH_MD5, 128, true, false>));
      case LEAF_DMA:
        if (copy) DD_LEAF_LAUNCH((k_leaf_hash_dma<HASH_MD5, 4, true>));
        DD_LEAF_LAUNCH((k_leaf_hash_dma<HASH_MD5, 4, false>));
      case LEAF_DMA1:
        if (copy) DD_LEAF_LAUNCH_T((k_leaf_hash_dma<HASH_MD5, 1, true>), 64);
        DD_LEAF_LAUNCH_T((k_leaf_hash_dma<HASH_MD5, 1, false>), 64);
      default:
        if (al16) DD_LEAF_LAUNCH((k_leaf_hash<HASH_MD5, true, LEAF_DIRECT>));
        DD_LEAF_LAUNCH((k_leaf_hash<HASH_MD5, false, LEAF_DIRECT>));
    }
  }
  switch (mapping) {
    case LEAF_DEEP: DD_LEAF_LAUNCH((k_leaf_hash<HASH_MURMUR3, true, LEAF_DEEP>));
    case LEAF_SHORT: DD_LEAF_LAUNCH((k_leaf_hash<HASH_MURMUR3, true, LEAF_SHORT>));
    case LEAF_LDS64:
      if (copy) DD_LEAF_LAUNCH((k_leaf_hash_lds<HASH_MURMUR3, 64, false, true>));
      DD_LEAF_LAUNCH((k_leaf_hash_lds<HASH_MURMUR3, 64, false, false>));
    case LEAF_LDS128:
      if (copy) DD_LEAF_LAUNCH((k_leaf_hash_lds<HASH_MURMUR3, 128, false, true>));
      DD_LEAF_LAUNCH((k_leaf_hash_lds<HASH_MURMUR3, 128, false, false>));
    case LEAF_LDS256:
      if (copy) DD_LEAF_LAUNCH((k_leaf_hash_lds<HASH_MURMUR3, 256, false, true>));
      DD_LEAF_LAUNCH((k_leaf_hash_lds<HASH_MURMUR3, 256, false, false>));
    case LEAF_LDS128P:
      if (copy) DD_LEAF_LAUNCH((k_leaf_hash_lds<HASH_MURMUR3, 128, true, true>));
      DD_LEAF_LAUNCH((k_leaf_hash_lds<HASH_MURMUR3, 128, true, false>));
    case LEAF_DMA:
      if (copy) DD_LEAF_LAUNCH((k_leaf_hash_dma<HASH_MURMUR3, 4, true>));
      DD_LEAF_LAUNCH((k_leaf_hash_dma<HASH_MURMUR3, 4, false>));
    case LEAF_DMA1:
      if (copy) DD_LEAF_LAUNCH_T((k_leaf_hash_dma<HASH_MURMUR3, 1, true>), 64);
      DD_LEAF_LAUNCH_T((k_leaf_hash_dma<HASH_MURMUR3, 1, false>), 64);
    default:
      if (al16) DD_LEAF_LAUNCH((k_leaf_hash<HASH_MURMUR3, true, LEAF_DIRECT>));
      DD_LEAF_LAUNCH((k_leaf_hash<HASH_MURMUR3, false, LEAF_DIRECT>));
  }
}
#undef DD_LEAF_LAUNCH
#undef DD_LEAF_LAUNCH_T

void launch_insert_leaves(const JobDev* jobs_h, const uint32_t* prefix, uint32_t njobs, uint32_t grid, uint32_t kinds,
                          hipStream_t s) {
  const JobG* jobs = gview(jobs_h);
  if (!grid) return;
  if (kinds & JOBS_DENSE) DD_LAUNCH(k_insert_leaves<JOBS_DENSE>, grid, 256, 0, s, jobs, prefix, njobs);
  if (kinds & JOBS_LIST_NARROW) DD_LAUNCH(k_insert_leaves<JOBS_LIST_NARROW>, grid, 256, 0, s, jobs, prefix, njobs);
  if (kinds & JOBS_LIST_WIDE) DD_LAUNCH(k_insert_leaves<JOBS_LIST_WIDE>, grid, 256, 0, s, jobs, prefix, njobs);
}
void launch_list_resolve(const JobDev* jobs_h, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s) {
  const JobG* jobs = gview(jobs_h);
  if (grid) DD_LAUNCH(k_list_resolve, grid, 256, 0, s, jobs, prefix, njobs);
}
void launch_low_root(const JobDev* jobs_h, uint32_t njobs, uint64_t max_c, uint32_t max_depth, hipStream_t s) {
  const JobG* jobs = gview(jobs_h);
  uint64_t bx = (max_c + 255) / 256;
  if (bx > 1024) bx = 1024;
  const dim3 leaf_grid((uint32_t)bx, njobs);
  hipLaunchKernelGGL(k_lr_leaves, leaf_grid, dim3(256), 0, s, jobs);
  for (int d = (int)max_depth - 1; d >= 0; d--) {
    uint64_t n = 1ull << d, b = (n + 255) / 256;
    if (b > 1024) b = 1024;
    hipLaunchKernelGGL(k_lr_level, dim3((uint32_t)b, njobs), dim3(256), 0, s, jobs, (uint32_t)d);
  }
  hipLaunchKernelGGL(k_lr_key, leaf_grid, dim3(256), 0, s, jobs);
  hipLaunchKernelGGL(k_lr_select, leaf_grid, dim3(256), 0, s, jobs);
}
void launch_pass1(const JobDev* jobs_h, const uint32_t* prefix, uint32_t njobs, uint32_t grid, int g, bool narrow,
                  hipStream_t s) {
  const JobG* jobs = gview(jobs_h);
  if (!grid) return;
  if (narrow) DD_LAUNCH(k_pass1<uint32_t>, grid, kTileThreads, 0, s, jobs, prefix, njobs, g);
  else DD_LAUNCH(k_pass1<uint64_t>, grid, kTileThreads, 0, s, jobs, prefix, njobs, g);
}
void launch_insert_interior(const JobDev* jobs_h, const uint32_t* prefix, uint32_t njobs, uint32_t grid,
                            uint32_t kinds, hipStream_t s) {
  const JobG* jobs = gview(jobs_h);
  if (!grid) return;
  if (kinds & JOBS_DENSE) DD_LAUNCH(k_insert_interior<JOBS_DENSE>, grid, 256, 0, s, jobs, prefix, njobs);
  if (kinds & JOBS_LIST_NARROW) DD_LAUNCH(k_insert_interior<JOBS_LIST_NARROW>, grid, 256, 0, s, jobs, prefix, njobs);
  if (kinds & JOBS_LIST_WIDE) DD_LAUNCH(k_insert_interior<JOBS_LIST_WIDE>, grid, 256, 0, s, jobs, prefix, njobs);
}
void launch_pass2(const JobDev* jobs_h, const uint32_t* prefix, uint32_t njobs, uint32_t grid, int g, bool narrow,
                  hipStream_t s) {
  const JobG* jobs = gview(jobs_h);
  if (!grid) return;
  if (narrow) DD_LAUNCH(k_pass2<uint32_t>, grid, kTileThreads, 0, s, jobs, prefix, njobs, g);
  else DD_LAUNCH(k_pass2<uint64_t>, grid, kTileThreads, 0, s, jobs, prefix, njobs, g);
}
void launch_roots_count(const JobDev* jobs_h, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s) {
  const JobG* jobs = gview(jobs_h);
  if (grid) DD_LAUNCH(k_roots_count, grid, 256, 0, s, jobs, prefix, njobs);
}
void launch_roots_write(const JobDev* jobs_h, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s) {
  const JobG* jobs = gview(jobs_h);
  if (grid) DD_LAUNCH(k_roots_write, grid, 256, 0, s, jobs, prefix, njobs);
}
void launch_sbin_count(const JobDev* jobs_h, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s) {
  const JobG* jobs = gview(jobs_h);
  if (grid) DD_LAUNCH(k_sbin_count, grid, 256, 0, s, jobs, prefix, njobs);
}
void launch_sbin_scan(const JobDev* jobs_h, uint32_t njobs, hipStream_t s) {
  DD_LAUNCH(k_sbin_scan, njobs, 1024, 0, s, gview(jobs_h));
}
void launch_sbin_scatter(const JobDev* jobs_h, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s) {
  const JobG* jobs = gview(jobs_h);
  if (grid) DD_LAUNCH(k_sbin_scatter, grid, 64, 0, s, jobs, prefix, njobs);
}
void launch_gather(const JobDev* jobs_h, const uint32_t* prefix, uint32_t njobs, uint32_t grid, bool vec16,
                   hipStream_t s) {
  const JobG* jobs = gview(jobs_h);
  if (!grid) return;
  if (vec16) DD_LAUNCH(k_gather<true>, grid, 256, 0, s, jobs, prefix, njobs);
  else DD_LAUNCH(k_gather<false>, grid, 256, 0, s, jobs, prefix, njobs);
}

void launch_bulk(const JobDev* jobs_h, uint32_t njobs, bool final_part, bool dedup, const uint32_t* ppre,
                 uint32_t pgrid, const uint32_t* spre, uint32_t sgrid, const uint32_t* bpre, uint32_t bgrid,
                 hipStream_t s) {
  const JobG* jobs = gview(jobs_h);
  if (!pgrid || !sgrid || !bgrid) return;
  if (!final_part) {
    DD_LAUNCH(k_bulk_hist<false>, pgrid, kPartThreads, 0, s, jobs, ppre, njobs);
    DD_LAUNCH(k_bulk_colscan<false>, sgrid, kColBuckets * kColSegs, 0, s, jobs, spre, njobs);
    DD_LAUNCH(k_bulk_scatter<false>, pgrid, kPartThreads, 0, s, jobs, ppre, njobs);
    if (dedup) DD_LAUNCH(k_bulk_build<false>, bgrid, kBuildThreads, 0, s, jobs, bpre, njobs);
  } else {
    DD_LAUNCH(k_bulk_hist<true>, pgrid, kPartThreads, 0, s, jobs, ppre, njobs);
    DD_LAUNCH(k_bulk_colscan<true>, sgrid, kColBuckets * kColSegs, 0, s, jobs, spre, njobs);
    DD_LAUNCH(k_bulk_scatter<true>, pgrid, kPartThreads, 0, s, jobs, ppre, njobs);
    DD_LAUNCH(k_bulk_build<true>, bgrid, kBuildThreads, 0, s, jobs, bpre, njobs);
  }
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

void launch_fetch(const FetchArgs& a, uint64_t max_len, uint8_t* dst, hipStream_t s) {
  if (a.n == 0) return;
  uint64_t bx = (max_len + 4095) / 4096;
  if (bx > 256) bx = 256;
  if (bx == 0) bx = 1;
  hipLaunchKernelGGL(k_fetch, dim3((uint32_t)bx, a.n), dim3(256), 0, s, a, dst);
}

void launch_restart_parse(RestartCkpt* desc, const RestartList* lists, uint32_t nlists, const uint32_t* list_blk,
                          uint32_t nblk, uint32_t* blk_sum, TreeGeom g, uint32_t cs, RestartCtr* ctr,
                          hipStream_t s) {
  unsigned long long* bs = reinterpret_cast<unsigned long long*>(blk_sum);
  if (nblk) DD_LAUNCH(k_rs_count, nblk, 256, 0, s, desc, lists, nlists, list_blk, bs, g, ctr);
  DD_LAUNCH(k_rs_scan, nlists, 256, 0, s, desc, lists, bs, g, cs, ctr);
  if (nblk) DD_LAUNCH(k_rs_add, nblk, 256, 0, s, desc, lists, nlists, list_blk, bs);
  const uint32_t rounds = (uint32_t)((g.C + 255) / 256);  // rounds of 256 leaves per list, at most
  DD_LAUNCH(k_rs_index, dim3((rounds + 255) / 256, nlists), dim3(256), 0, s, desc, lists, ctr);
}

void launch_restart_regions(const RestartCkpt* rc, TreeGeom g, uint2* cover, uint32_t base, uint32_t stamp,
                            uint64_t* ref, uint32_t* pend, RestartCtr* ctr, uint8_t* out, uint64_t len, uint32_t cs,
                            int piece, hipStream_t s) {
  // The leaf count is computed on the device: a grid-stride launch of at
  // most 2048 workgroups (8 per CU).
  uint64_t blocks = (g.C + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  const uint32_t b = (uint32_t)blocks;
  if (piece == 16) DD_LAUNCH(k_restart_regions<16>, b, 256, 0, s, rc, g, cover, base, stamp, ref, pend, ctr, out, len, cs);
  else if (piece == 4) DD_LAUNCH(k_restart_regions<4>, b, 256, 0, s, rc, g, cover, base, stamp, ref, pend, ctr, out, len, cs);
  else DD_LAUNCH(k_restart_regions<1>, b, 256, 0, s, rc, g, cover, base, stamp, ref, pend, ctr, out, len, cs);
}

void launch_restart_whole(const RestartCkpt* rc, TreeGeom g, uint2* cover, uint32_t base, uint32_t stamp,
                          RestartCtr* ctr, uint8_t* out, uint64_t len, uint32_t cs, hipStream_t s) {
  uint64_t blocks = (g.C + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  DD_LAUNCH(k_restart_whole, (uint32_t)blocks, 256, 0, s, rc, g, cover, base, stamp, ctr, out, len, cs);
}

void launch_restart_pending(const RestartCkpt* rc, TreeGeom g, const uint2* cover, uint32_t stamp, uint64_t* ref,
                            const uint32_t* pend, RestartCtr* ctr, uint8_t* out, uint64_t len, uint32_t cs, int piece,
                            hipStream_t s) {
  // The list length is only known on the device: a grid-stride launch sized
  // for a tenth of the chunks (workgroups past the list return at once).
  uint64_t blocks = (g.C / 10 + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  if (blocks == 0) blocks = 1;
  const uint32_t b = (uint32_t)blocks;
  if (piece == 16) DD_LAUNCH(k_restart_pending<16>, b, 256, 0, s, rc, g, cover, stamp, ref, pend, ctr, out, len, cs);
  else if (piece == 4) DD_LAUNCH(k_restart_pending<4>, b, 256, 0, s, rc, g, cover, stamp, ref, pend, ctr, out, len, cs);
  else DD_LAUNCH(k_restart_pending<1>, b, 256, 0, s, rc, g, cover, stamp, ref, pend, ctr, out, len, cs);
}

void launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  DD_LAUNCH(k_fill_u32, (uint32_t)blocks, 256, 0, s, p, n, v);
}

}  // namespace dd
