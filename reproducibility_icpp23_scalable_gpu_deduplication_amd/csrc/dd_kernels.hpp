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
// Bulk map build of a baseline: partition workgroups of kPartTile records
// (a row of per-bucket counts each), column scans over groups of
// kColBuckets buckets, one build workgroup per bucket holding at most
// kBulkRec records.  A job has at most kBulkMaxBuckets buckets (LDS
// counters of a partition workgroup: maps of at most 2^25 slots) and
// kBulkMaxTiles tiles (16 rows for each of the kColSegs segments of a
// column scan).
constexpr uint32_t kPartThreads = 1024;
constexpr uint32_t kPartItems = 32;  // records per lane
constexpr uint32_t kPartTile = kPartThreads * kPartItems;
constexpr uint32_t kBulkMaxBuckets = 8192;
constexpr uint32_t kColBuckets = 64;
constexpr uint32_t kColSegs = 16;
constexpr uint32_t kBulkMaxTiles = 16 * kColSegs;
constexpr uint32_t kBuildThreads = 512;
constexpr uint32_t kBulkRec = 2400;

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
  // the handle completed and its root writer reset every other label), so
  // the leaf kernel writes the labels of changed leaves only.
  uint32_t write_all_labels;
  // 1: the map inserts probe in wide rounds (few inserts expected, each
  // wavefront waiting for its longest chain); 0: narrow rounds (baselines
  // and change-heavy checkpoints, bound by random-line traffic).
  uint32_t wide_probe;
  // 1: node 0 is emitted as a shifted root when the whole buffer equals an
  // earlier checkpoint (DD_OPT_EMIT_ROOT_SHIFT); 0: only as a first
  // occurrence, as the reference does (include/reference_impl.hpp:105-106).
  uint32_t root_shift;
  // 1: low-root first occurrences (DD_METHOD_TREE_LOW_ROOT); lr_new holds N
  // new-node flags of this checkpoint.
  uint32_t low_root;
  // Bulk map build of a baseline (k_bulk_*): 1 for a tree-method baseline
  // whose map is built bucket by bucket instead of by the dense inserts.
  // The leaves, and after pass 1 the first-occurrence interior nodes, are
  // partitioned by map bucket into record arrays (digest, node); each bucket
  // is then built in LDS and written out whole.
  uint32_t bulk;
  // 1: the leaf part searches each bucket's leaf records for repeated
  // digests and labels them SHIFT before pass 1; 0: the job is expected to
  // have none (its last baseline had none), and the final build checks.
  uint32_t bulk_dedup;
  uint32_t nbuckets;    // map buckets (capacity / kBucketSlots, at least 1)
  uint32_t ntiles_b;    // partition workgroups: ceil(C / kPartTile)
  uint32_t bk_cap_l;    // leaf records per bucket above which the job falls back to the dense inserts
  uint32_t bk_cap_b;    // records per bucket above which the final build gives up (FLAG_MAP_FULL)
  P<uint32_t> bk_mat;   // per part (leaves, then interior nodes): ntiles_b x nbuckets counts, then column prefixes
  P<uint32_t> bk_tot;   // per part: records per bucket
  P<uint32_t> bk_grp;   // per part: records per group of kColBuckets buckets
  P<uint32_t> bk_claim;  // per bucket: entries written by the final build
  // Records of 32 bytes (digest; node id and zero padding): C leaves, then
  // C-1 interior nodes.
  P<u32x4_t> bk_rec;
  // persistent state
  P<Digest> tree;        // N digests, heap order
  P<uint8_t> labels;     // N labels (padded for 16-byte reads)
  // N: the map entry of a pending leaf or shifted node: its value (source
  // checkpoint << 32 | source node) when it could not change any more when
  // the node met it (an entry of an earlier checkpoint, or one met after this
  // checkpoint's inserts), else kRefSlot | its slot (low-root jobs: always
  // the slot, which their kernels rewrite entries through).
  P<uint64_t> slot_of;
  P<uint8_t> lr_new;     // N (low-root jobs only)
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
  // Per 64 groups of 16 nodes (a wavefront of the root count): bit g set when
  // group g holds a label other than FIXED, which the root writer resets.
  P<uint64_t> lmask;
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
  // 1: a tree baseline whose leaf kernel also copies every chunk into the
  // data section of the diff, at the place it takes when the root is the
  // only region (every chunk a first occurrence); the gather then copies no
  // data unless the labels found another layout, which overwrites it.
  uint32_t write_data;
  uint32_t pad_;
};
// slot_of entries that name a map slot rather than an entry's value.
constexpr uint64_t kRefSlot = 0xffffffff00000000ull;
__host__ __device__ inline bool ref_is_slot(uint64_t ref) { return (ref >> 32) == 0xffffffffu; }

using JobDev = JobT<HostPtr>;
using JobG = JobT<GlobalPtr>;
static_assert(sizeof(JobDev) == sizeof(JobG), "one layout for both pointer kinds");

// Restart descriptors -------------------------------------------------------
// One checkpoint of the chain a restart walks (newest first).  The host fills
// everything but f_off, s_off, f_idx, s_idx, nfl and nsl, which the parse
// kernels compute from the diff's region lists.
struct RestartCkpt {
  const uint8_t* diff;      // device copy of the diff
  uint32_t* f_off;          // exclusive prefix of first-region leaf counts (nfirst+1)
  uint32_t* s_off;          // exclusive prefix of shifted-region leaf counts (nshift+1)
  uint32_t* f_idx;          // per round of 256 first-region leaves: the region of its first leaf
  uint32_t* s_idx;          // the same for the shifted regions (ceil(C/256) entries each)
  const uint32_t* pb_end;   // source checkpoints of the shifted list: entries below pb_end[j]
  const uint32_t* pb_ck;    //   (and at or above pb_end[j-1]) come from checkpoint pb_ck[j]
  uint32_t npb;
  uint32_t nfirst, nshift;
  uint32_t ckpt_id;
  uint32_t node_bias;       // added to every node id of the diff (C-1 for list diffs: chunk -> leaf)
  uint32_t no_cover;        // 1: the walk's last step, by k_restart_whole: it writes no cover entries
  uint64_t first_off;       // byte offset of the first-occurrence list
  uint64_t shift_off;       // byte offset of the shifted list
  uint64_t data_off;        // byte offset of the data section
  uint64_t data_bytes;      // bytes of the data section
  uint64_t nfl, nsl;        // leaves in first-occurrence / shifted regions
};

// A region list of the chain (first-occurrence or shifted list of one
// checkpoint) in the parse kernels: `blk0` is its first block of 1024 entries
// in the launch; blk_sum[blk0 + b] holds block b's leaf count, then its
// offset.
struct RestartList {
  uint32_t ckpt;   // descriptor index
  uint32_t shift;  // 0 first-occurrence list, 1 shifted list
  uint32_t n;      // entries
  uint32_t blk0;
};
constexpr uint32_t kParseBlock = 1024;  // region entries per parse block (256 lanes x 4)

// Counters of one restart (device), zeroed by the host's upload.
struct RestartCtr {
  unsigned long long claimed;  // chunks whose newest covering checkpoint was found
  uint32_t npend;              // entries of the pending list (chunks that name an older source)
  uint32_t ndone;              // pending entries resolved
  uint32_t flags;              // FLAG_BAD_SOURCE, FLAG_BAD_DIFF
  uint32_t pad[3];
};

// Cover table entry, one per chunk: {stamp of the checkpoint step that wrote
// it, region}.  Region: bit 31 set for a shifted region, low 31 bits its index
// in the checkpoint's first-occurrence or shifted list.
constexpr uint32_t kCoverShift = 0x80000000u;
// Pending source of a chunk: ckpt << 32 | kRefSameCkpt? | source chunk.
// kRefSameCkpt marks a source in the same checkpoint as the shifted region
// that named it, which must be a first occurrence there.
constexpr uint64_t kRefDone = ~0ull;
constexpr uint32_t kRefSameCkpt = 0x80000000u;

// Up to kFetchRanges device byte ranges copied into one buffer by one launch
// (headers and metadata of the diffs of a chain).  With `tables` set, range r
// is a diff of lim[r] bytes: its 32-byte header goes to dst_off[r] and its
// prior table (8 x num_prior_chkpts bytes after the first-occurrence list),
// cut to len[r] - 32 bytes and to the diff's end, right after it.
constexpr int kFetchRanges = 32;
struct FetchArgs {
  uint32_t n;
  uint32_t tables;
  const uint8_t* src[kFetchRanges];
  uint64_t len[kFetchRanges];
  uint64_t dst_off[kFetchRanges];
  uint64_t lim[kFetchRanges];
};

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
// from LDS (needs chunk_size % 128 == 0); DMA1: DMA with one-wavefront
// workgroups.  All but DIRECT need 16-byte aligned buffers; one workgroup
// covers 256 chunks (64 for DMA1).
enum LeafMapping : int {
  LEAF_DIRECT = 0,
  LEAF_DEEP = 1,
  LEAF_LDS64 = 2,
  LEAF_LDS128 = 3,
  LEAF_LDS256 = 4,
  LEAF_SHORT = 5,
  LEAF_DMA = 6,
  LEAF_DMA1 = 7,
  LEAF_LDS128P = 8,
  LEAF_NMAP = 9
};
// ev_start/ev_stop (optional) are recorded at the start and the end of the
// leaf kernel's own dispatch.  `copy`: some job of the table has write_data
// set (the mapping must be one with leaf_writes_data).
void launch_leaf(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, int mapping,
                 bool all_aligned16, int hash_kind, hipStream_t s, hipEvent_t ev_start = nullptr,
                 hipEvent_t ev_stop = nullptr, bool copy = false);
// Workgroups of the leaf stage for a job of C chunks.
inline uint64_t leaf_workgroups(int mapping, uint64_t C) { return mapping == LEAF_DMA1 ? (C + 63) / 64 : (C + 255) / 256; }
// Mappings with a kernel variant that copies the chunks into a baseline's
// data section (JobT::write_data): each stages whole rounds of 64 chunks, so
// its stores are contiguous.
inline bool leaf_writes_data(int mapping) {
  return mapping == LEAF_DMA || mapping == LEAF_DMA1 || mapping == LEAF_LDS64 || mapping == LEAF_LDS128 ||
         mapping == LEAF_LDS256 || mapping == LEAF_LDS128P;
}
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
// Low-root method: after the leaf inserts, moves each new digest's first
// occurrence to the leaf whose highest new ancestor has the smallest node id
// (jobs without low_root are skipped).  `max_c` / `max_depth`: over the jobs.
void launch_low_root(const JobDev* jobs, uint32_t njobs, uint64_t max_c, uint32_t max_depth, hipStream_t s);
// `narrow`: every job's tree has fewer than 2^31 nodes (32-bit node
// indices in the tile passes).
void launch_pass1(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, int group, bool narrow,
                  hipStream_t s);
void launch_insert_interior(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, uint32_t kinds,
                            hipStream_t s);
void launch_pass2(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, int group, bool narrow,
                  hipStream_t s);
void launch_roots_count(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s);
void launch_roots_write(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s);
void launch_sbin_count(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s);
void launch_sbin_scan(const JobDev* jobs, uint32_t njobs, hipStream_t s);
void launch_sbin_scatter(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, hipStream_t s);
void launch_gather(const JobDev* jobs, const uint32_t* prefix, uint32_t njobs, uint32_t grid, bool vec16,
                   hipStream_t s);

// Bulk map build of the bulk jobs of a table (JobT::bulk).  Leaf part,
// between the leaf kernel and pass 1: counts, column scan, partition, then
// (`dedup`: some job has bulk_dedup set) the per-bucket duplicate search
// that labels every leaf but the lowest of each digest SHIFT.  Final part,
// after pass 1: the same for the first-occurrence interior nodes, then the
// per-bucket build that writes the map.  `pgrid`/`ppre`: the partition
// tiles, `sgrid`/`spre`: the column-scan groups, `bgrid`/`bpre`: one
// workgroup per bucket.
void launch_bulk(const JobDev* jobs, uint32_t njobs, bool final_part, bool dedup, const uint32_t* ppre,
                 uint32_t pgrid, const uint32_t* spre, uint32_t sgrid, const uint32_t* bpre, uint32_t bgrid,
                 hipStream_t s);
void launch_map_clear(Slot* slots, uint64_t cap, hipStream_t s);
void launch_map_rehash(const Slot* old_slots, uint64_t old_cap, MapView nmap, uint32_t* flags, hipStream_t s);

void launch_fetch(const FetchArgs& a, uint64_t max_len, uint8_t* dst, hipStream_t s);
// Parse of every region list of a chain: leaf counts and their exclusive
// prefixes per list (RestartCkpt f_off/s_off, nfl/nsl), with the checks that
// keep the walk in bounds (node ids, equal sizes of a shifted region and its
// source, totals within the chunk count and the data section).  A failed
// check sets FLAG_BAD_DIFF and the walk kernels then do nothing.
void launch_restart_parse(RestartCkpt* desc, const RestartList* lists, uint32_t nlists, const uint32_t* list_blk,
                          uint32_t nblk, uint32_t* blk_sum, TreeGeom g, uint32_t chunk_size, RestartCtr* ctr,
                          hipStream_t s);
// Restart step of one checkpoint (newest first): every leaf of its regions
// gets a cover entry stamped `stamp`; a leaf whose cover entry is older than
// this restart (stamp <= base) has this checkpoint as its newest cover: data
// of a first-occurrence region is copied to `out`, a shifted region puts the
// chunk on the pending list with its source.  `piece`: 16, 4 or 1 byte
// copies (chosen by the host from the alignments).
void launch_restart_regions(const RestartCkpt* rc, TreeGeom g, uint2* cover, uint32_t base, uint32_t stamp,
                            uint64_t* ref, uint32_t* pend, RestartCtr* ctr, uint8_t* out, uint64_t len,
                            uint32_t chunk_size, int piece, hipStream_t s);
// Resolves the pending chunks whose source is in this checkpoint (read from
// the cover entries the step wrote), copying their data or moving them to an
// older checkpoint.
// Step of a checkpoint with one first-occurrence region and no shifted one,
// 16-byte pieces (same results as launch_restart_regions).
void launch_restart_whole(const RestartCkpt* rc, TreeGeom g, uint2* cover, uint32_t base, uint32_t stamp,
                          RestartCtr* ctr, uint8_t* out, uint64_t len, uint32_t cs, hipStream_t s);
void launch_restart_pending(const RestartCkpt* rc, TreeGeom g, const uint2* cover, uint32_t stamp, uint64_t* ref,
                            const uint32_t* pend, RestartCtr* ctr, uint8_t* out, uint64_t len, uint32_t chunk_size,
                            int piece, hipStream_t s);
void launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);

}  // namespace dd
