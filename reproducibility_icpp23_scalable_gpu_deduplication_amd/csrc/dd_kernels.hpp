// dd_kernels.hpp -- host-visible declarations of the deduplication kernels.
//
// Every kernel works on a table of jobs (one job = one checkpoint buffer with
// its own tree, digest map and output diff).  A launch covers all jobs at once:
// `prefix[j]` is the first workgroup of job j and prefix[njobs] the grid size,
// so the single-buffer path is a table of one job and the multi-buffer path
// runs the same kernels over many jobs in one launch per stage.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dd_device.hpp"

namespace dd {

constexpr int kTileLevels = 8;          // levels per tree tile (256 bottom nodes)
constexpr int kNodesPerBlock = 4096;    // root scan: 256 threads x 16 nodes
constexpr uint32_t kRootWG = 1024;      // root-scan workgroups per job
constexpr int kSTile = 1024;            // shifted-root sort tile
constexpr uint32_t kMaxStageWG = 4096;  // workgroups per job of a grid-stride stage
constexpr uint32_t kMaxTileWG = 8192;   // one-wavefront workgroups per job of a tile pass
constexpr uint32_t kTileInterior = 255;  // interior nodes of a tile of 8 levels
// Work-list counters (32 per list) sit kCntStride words apart, one per
// 256-byte block: counters that share a cache line serialize their atomics
// in one L2 channel (one atomic per wavefront with a changed chunk).
constexpr uint32_t kCntStride = 64;

// Per-job counters written by the device and read back by the host.
struct JobCounters {
  unsigned long long pending;     // leaves that needed a map insert
  unsigned long long map_count;   // entries in the digest map
  uint32_t flags;                 // FLAG_*
  uint32_t nfirst;                // first-occurrence roots
  uint32_t nshift;                // shifted-duplicate roots
  uint32_t nprior;                // prior-checkpoint table entries
  uint64_t dchunks;               // chunks in first-occurrence regions
  uint64_t diff_len;              // bytes of the diff
  uint32_t unresolved;            // restart: chunks left without a source
  uint32_t pad;
};

// One job.  The host fills it as JobDev (plain pointers); kernels read the
// same bytes as JobG (global-address-space pointers, see dd_device.hpp).
template <template <class> class P>
struct JobT {
  // input
  P<const uint8_t> data;
  uint64_t len;
  uint32_t chunk_size;
  int32_t hash_kind;
  uint64_t C, N;
  uint64_t nfull;  // complete chunks: len / chunk_size (C or C-1)
  uint32_t D;
  uint32_t ckpt_id;
  uint32_t baseline_id;
  uint32_t is_baseline;
  // Method: 0 = tree, 1 = list (ListDeduplicator, src/list_approach.cpp:23-256):
  // leaves only, every interior label DONE, so each changed chunk is its own
  // root.  Diff node ids of a list job are chunk indices (leaf - node_bias).
  uint32_t list_mode;
  uint32_t node_bias;
  // 0: every leaf label is FIXED on entry (the previous tree checkpoint of
  // the handle completed and reset the labels of the tiles it visited), so
  // the leaf kernel writes the labels of changed leaves only.
  uint32_t write_all_labels;
  // 1: the map inserts probe in wide rounds (few inserts expected, each
  // wavefront waiting for its longest chain); 0: narrow rounds (baselines
  // and change-heavy checkpoints, bound by random-line traffic).
  uint32_t wide_probe;
  // persistent state
  P<Digest> tree;        // N digests, heap order
  P<uint8_t> labels;     // N labels (padded for 16-byte reads)
  P<uint32_t> slot_of;   // N: map slot of pending leaves and shifted interior nodes
  MapT<P> map;
  uint64_t map_count_at_start;  // entries before this checkpoint
  uint64_t map_limit;           // max entries the map may hold
  P<JobCounters> ctr;       // device counters
  P<JobCounters> host_ctr;  // host-mapped copy written at the end of the pipeline
  // Work lists.  Each is split into 32 sub-lists (picked by the producing
  // leaf block or tile index modulo 32) with one counter each, so no single
  // counter is hit by every workgroup; consumers prefix-sum the 32 counts.
  P<uint32_t> pend_cnt;    // 32 counters, kCntStride words apart
  P<uint32_t> pend_cnt_next;   // the other set (the next checkpoint's), zeroed by k_sbin_scan
  P<uint32_t> pend_list;   // 32 x pend_cap: chunk indices of the leaves that need a map insert
  uint32_t pend_cap;
  // Digests of the first dig_cap entries of every pending sub-list (32 x
  // dig_cap, beside pend_list), so the leaf inserts read an entry and its
  // digest in one round trip; entries past dig_cap are read from the tree.
  uint32_t dig_cap;
  P<Digest> pend_dig;
  // First-occurrence interior nodes found by pass 1, per tile (global tile
  // id T over all groups): fint_n[T] entries at fint_list[T * kTileInterior].
  // Pass 1 writes fint_n for every tile it visits, so no clearing is needed.
  P<uint32_t> fint_list;
  P<uint32_t> fint_n;
  uint32_t epoch;        // stamp of this checkpoint (never 0)
  P<uint32_t> tile_epoch;  // per group-0 tile: epoch of the last checkpoint that changed one of its leaves
  P<uint32_t> dirty_cnt;   // 32 counters, kCntStride words apart
  P<uint32_t> dirty_cnt_next;  // the other set (the next checkpoint's), zeroed by k_sbin_scan
  P<uint32_t> dirty_list;  // 32 x dirty_cap: group-0 tiles with a changed leaf
  uint32_t dirty_cap;
  uint32_t tile_base[4];  // first tile index of each group
  uint32_t ntiles;        // tiles over all groups
  P<uint32_t> blk_claim;    // per insert_leaves workgroup: new map entries
  P<uint32_t> int_claim;    // per insert_interior workgroup: new map entries
  uint32_t nwg_ins_leaf, nwg_ins_int;  // workgroups of the two insert stages
  // root scan
  P<uint32_t> blk_f;    // per root-scan workgroup: first roots (then their exclusive prefix)
  P<uint32_t> blk_fc;   // per root-scan workgroup: chunks under first roots (prefix)
  P<uint32_t> blk_s;    // per root-scan workgroup: shifted roots (prefix)
  uint32_t nblk;      // node blocks = ceil(N / kNodesPerBlock)
  P<uint32_t> rmask;  // per group of 16 nodes: first-root bits | shifted-root bits << 16
  P<uint32_t> f_list;   // first roots, ascending
  P<uint32_t> f_off;    // chunk offset of each first region in the data section
  P<uint32_t> s_node;   // shifted roots, ascending
  P<uint32_t> s_prev;   // source node of each shifted root
  P<uint32_t> s_tree;   // source checkpoint of each shifted root
  P<uint32_t> ss_node;  // shifted roots grouped by source checkpoint
  P<uint32_t> ss_prev;
  P<uint32_t> s_hist;   // (ckpt_id+1) x ceil(C / kSTile) histogram / bases
  P<uint32_t> ptab;     // prior table {ckpt, count} x (ckpt_id+1)
  // output
  P<uint8_t> diff;
  uint64_t diff_cap;
};
using JobDev = JobT<HostPtr>;
using JobG = JobT<GlobalPtr>;
static_assert(sizeof(JobDev) == sizeof(JobG), "one layout for both pointer kinds");

// Restart descriptors -------------------------------------------------------
struct RestartCkpt {
  const uint8_t* diff;      // device copy of the diff
  const uint32_t* f_off;    // exclusive prefix of first-region leaf counts (nfirst+1)
  const uint32_t* s_off;    // exclusive prefix of shifted-region leaf counts (nshift+1)
  const uint32_t* s_tree;   // source checkpoint per shifted entry
  uint32_t nfirst, nshift;
  uint32_t ckpt_id;
  uint64_t first_off;       // byte offset of the first-occurrence list
  uint64_t shift_off;       // byte offset of the shifted list
  uint64_t data_off;        // byte offset of the data section
  uint32_t node_bias;       // added to every node id of the diff (C-1 for list diffs: chunk -> leaf)
  uint32_t pad;
};

struct CoverEntry {
  uint32_t stamp;     // generation that wrote the entry
  uint32_t src_ckpt;  // kCoverFirst for data in this checkpoint
  uint64_t payload;   // data byte offset, or source leaf node
};
constexpr uint32_t kCoverFirst = 0xffffffffu;

// Launch helpers (dd_kernels.hip) -------------------------------------------
void launch_begin(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, bool rerun, hipStream_t s);
// Chunk-to-lane mappings of the leaf kernel.  DIRECT: one lane per chunk,
// 16-byte loads, 64 bytes in flight per lane; DEEP: the same with 256 bytes
// in flight (Murmur3); LDS<seg>: each wavefront loads `seg` bytes of 64
// chunks with contiguous 16-byte loads into LDS per round, then each lane
// hashes its chunk from LDS (needs chunk_size % seg == 0); SHORT: one lane
// per chunk of at most 128 bytes, the whole chunk in flight (Murmur3); DMA:
// each wavefront loads 128-byte segments of its 64 chunks straight into LDS
// with contiguous loads (global_load_lds) and each lane hashes its chunk
// from LDS (needs chunk_size % 128 == 0).  All but DIRECT need 16-byte
// aligned buffers; one workgroup covers 256 chunks in every case.
enum LeafMapping : int {
  LEAF_DIRECT = 0,
  LEAF_DEEP = 1,
  LEAF_LDS64 = 2,
  LEAF_LDS128 = 3,
  LEAF_LDS256 = 4,
  LEAF_SHORT = 5,
  LEAF_DMA = 6,
  LEAF_NMAP = 7
};
// ev_start/ev_stop (optional) are recorded at the start and the end of the
// leaf kernel's own dispatch.
void launch_leaf(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, int mapping,
                 bool all_aligned16, int hash_kind, hipStream_t s, hipEvent_t ev_start = nullptr,
                 hipEvent_t ev_stop = nullptr);
// Workgroups of the leaf stage for a job of C chunks: one per 256 chunks in
// every mapping.
inline uint64_t leaf_workgroups(int, uint64_t C) { return (C + 255) / 256; }
// Insert modes of the jobs in a table (bit mask): the insert stages launch
// one kernel variant per mode present.  Dense: a baseline (every leaf and
// interior node, narrow probe rounds); list, narrow or wide rounds: an
// incremental checkpoint (its work lists; wide_probe picks the width).
enum : uint32_t { JOBS_DENSE = 1, JOBS_LIST_NARROW = 2, JOBS_LIST_WIDE = 4 };
void launch_insert_leaves(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, uint32_t kinds,
                          hipStream_t s);
// List method: the leaves left PENDING by insert_leaves are first
// occurrences; a baseline also sets every interior label to DONE.
void launch_list_resolve(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s);
void launch_pass1(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, int group, hipStream_t s);
void launch_insert_interior(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, uint32_t kinds,
                            hipStream_t s);
void launch_pass2(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, int group, hipStream_t s);
void launch_roots_count(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s);
void launch_roots_write(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s);
void launch_sbin_count(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s);
void launch_sbin_scan(const JobDev* jobs, uint32_t njobs, hipStream_t s);
void launch_sbin_scatter(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s);
void launch_gather(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, bool vec16,
                   hipStream_t s);

void launch_map_clear(Slot* slots, uint64_t cap, hipStream_t s);
void launch_map_rehash(const Slot* old_slots, uint64_t old_cap, MapView nmap, uint32_t* flags, hipStream_t s);

void launch_restart_init(uint64_t* ref, uint64_t C, uint32_t ckpt, JobCounters* ctr, hipStream_t s);
void launch_restart_cover(const RestartCkpt* rc, TreeGeom g, CoverEntry* cover, uint32_t stamp, uint64_t nfirst_leaves,
                          uint64_t nshift_leaves, uint32_t chunk_size, hipStream_t s);
void launch_restart_resolve(const RestartCkpt* rc, TreeGeom g, const CoverEntry* cover, uint32_t stamp,
                            uint64_t* ref, uint8_t* out, uint64_t len, uint32_t chunk_size, JobCounters* ctr,
                            hipStream_t s);
void launch_restart_check(const uint64_t* ref, uint64_t C, JobCounters* ctr, hipStream_t s);

}  // namespace dd
