// dd_engine.hpp -- host-side orchestration shared by the C ABI entry points:
// device buffers, the per-checkpoint kernel sequence over a job table, digest
// map growth, and the restart chain walk.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/dedup_c.h"
#include "dd_kernels.hpp"

namespace dd {

// A grow-only device allocation.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf();
  // Grows to at least n bytes; a fresh allocation is zero-filled on stream
  // `zs` when `zero` is set (ordered before later work on that stream).
  dd_status ensure(size_t n, bool zero = false, hipStream_t zs = nullptr);
  void release();
  template <typename T>
  T* as() const { return reinterpret_cast<T*>(p); }
};

// Persistent per-stream deduplication state: the previous tree's leaf
// digests, the digest map and the scratch arrays of one checkpoint buffer.
struct StreamState {
  uint32_t chunk_size = 0;
  int hash_kind = 0;
  int method = DD_METHOD_TREE;  // DD_METHOD_TREE, DD_METHOD_LIST or DD_METHOD_TREE_LOW_ROOT
  uint32_t options = 0;         // DD_OPT_* flags
  uint32_t current_id = 0, baseline_id = 0;
  bool have_tree = false;
  bool ctr_armed = false;  // the device counters are zero and hold the current map count
  bool labels_fixed = false;  // every leaf label is FIXED (the last tree checkpoint completed)
  // A checkpoint started changing the persistent state (tree, labels, map)
  // and did not complete: the next checkpoint must be a baseline.
  bool need_baseline = false;
  uint64_t data_len = 0, C = 0, N = 0;
  int64_t last_map_ckpt = -1;  // checkpoint id of the last insert into the map

  DevBuf tree, labels, slot_of, lr_new;
  DevBuf map;
  uint64_t map_cap = 0, map_count = 0;
  uint64_t base_map_count = 0;    // map entries after the chain's baseline
  uint32_t ckpts_since_base = 0;  // completed incremental checkpoints since it
  DevBuf blk, f_list, f_off, s_node, s_prev, s_tree, ss_node, ss_prev, s_hist, ptab;
  DevBuf leafblk, pend_list, pend_dig, tiles, tile_epoch, lists;
  uint32_t epoch = 0;     // stamp of the current checkpoint (tile_epoch)
  uint32_t cnt_par = 1;   // work-list counter set of the last completed checkpoint (the next uses the other)
  uint64_t last_pending = ~0ull;  // pending leaves of the last completed incremental checkpoint (~0: none yet)
  uint64_t tile_epoch_n = 0;
  DevBuf diff;  // internal diff buffer (host-output path)
  DevBuf ctr;   // JobCounters
  // Bulk map build of a baseline: count matrices, bucket and group totals,
  // entry counts and the records (digest, node) of the leaves and interior
  // nodes.
  DevBuf bulk;
  uint64_t bk_off[6] = {0, 0, 0, 0, 0, 0};
  uint64_t bk_nb = 0, bk_nt = 0;
  bool bulk_now = false;  // the baseline being prepared uses the bulk build
  bool bulk_dups = false;  // a baseline of this handle had a repeated leaf digest
  bool bulk_heavy = false;  // ... one repeated so often that a bucket overflowed: no bulk build
  // The last baseline's diff had another layout than the root as its only
  // region: the next baseline's leaf kernel copies no data (JobT::write_data).
  bool data_miss = false;

  uint64_t nfirst = 0, nshift = 0, last_diff_len = 0;
  dd_ckpt_stats stats{};  // of the last completed checkpoint
  double timers[4] = {0, 0, 0, 0};
  double restart_timers[2] = {0, 0};

  // Whether a checkpoint call with `make_baseline` runs as a baseline.
  bool runs_baseline(int make_baseline) const {
    return make_baseline || !have_tree || current_id == 0 || need_baseline;
  }
  // The checks of prepare() that change nothing: a call they reject leaves
  // the handle as it was.
  dd_status validate(uint64_t len, bool baseline, const uint8_t* out, uint64_t out_cap) const;
  // Sizes the state for a checkpoint of `len` bytes (fresh map and zero tree
  // on a baseline) and fills `job`.  `out`/`out_cap` are the diff buffer.
  // From its first change of the persistent state until commit(), the
  // handle needs a baseline.
  dd_status prepare(const uint8_t* data, uint64_t len, bool baseline, uint8_t* out, uint64_t out_cap, JobDev& job,
                    hipStream_t s);
  // Called after the checkpoint's counters are back on the host;
  // `leaf_copy`: its leaf kernel wrote the data section.
  void commit(const JobCounters& c, bool baseline, bool leaf_copy);
  // Grows the map so that `need` entries stay under the load limit.
  dd_status grow_map(uint64_t need, hipStream_t s);
  // Bulk map build: whether a baseline of Cn chunks and map capacity `cap`
  // uses it, its arrays, and its job fields.
  bool uses_bulk(uint64_t Cn, uint64_t cap) const;
  dd_status prepare_bulk(uint64_t Cn, uint64_t cap, hipStream_t s);
  void fill_bulk(JobDev& job) const;
};

// Runs the checkpoint kernels over a table of jobs.  Owns the device copies
// of the job table and of the per-stage workgroup prefixes.
class Pipeline {
 public:
  ~Pipeline();
  // Uploads the job table and launches every stage on `s`.  `host_ctr`
  // receives the counters (pinned, njobs entries) once `s` is synchronized.
  // `stage_ev`, when given, holds DD_NUM_STAGES + 1 events recorded around
  // the stages; `ev_dedup` is recorded after the duplicate forest.
  // `begin` launches k_begin first; without it the counters must be armed
  // (the work-list counter set of this checkpoint's parity zeroed by the
  // previous completed checkpoint's k_sbin_scan, or a zero-filled allocation).
  dd_status run(const std::vector<JobDev>& jobs, hipStream_t s, bool rerun, bool begin, hipEvent_t ev_dedup = nullptr,
                hipEvent_t* stage_ev = nullptr);
  // The two halves of run(): upload the job table (a copy on `s`), then
  // enqueue the kernels.  launch_all() enqueues nothing but kernels and event
  // records, so it can be captured into a hipGraph.
  dd_status upload(const std::vector<JobDev>& jobs, hipStream_t s);
  // `part`: PART_ALL, or PART_HEAD (counter reset and leaf kernel) and
  // PART_TAIL (the stages after the leaf kernel) in two calls.
  enum { PART_ALL = 0, PART_HEAD = 1, PART_TAIL = 2 };
  dd_status launch_all(hipStream_t s, bool rerun, bool begin, hipEvent_t ev_dedup = nullptr,
                       hipEvent_t* stage_ev = nullptr, int part = PART_ALL);
  // Identifies the launch sequence of the last upload (grids, kernel
  // variants, table address): a captured graph is valid while it is equal.
  std::vector<uint64_t> signature() const;
  // Waits for the stream; the kernels have written every job's counters to
  // host_ctr[j] (host-mapped pinned memory).
  dd_status fetch_counters(const std::vector<JobDev>& jobs, hipStream_t s);
  JobCounters* host_ctr = nullptr;
  int profile_level_ = 0;

 private:
  uint32_t njobs_ = 0;
  DevBuf dev_table_;           // job table followed by the stage prefixes
  uint8_t* host_table_ = nullptr;  // pinned staging of dev_table_
  size_t host_table_cap_ = 0, host_ctr_cap_ = 0;
  JobDev* dev_jobs_ = nullptr;
  uint32_t* dev_prefix_ = nullptr;
  std::vector<uint32_t> off_;  // offset of each stage's prefix in dev_prefix_
  std::vector<uint32_t> grid_;
  int max_groups_ = 0;
  bool aligned16_ = true, vec16_ = true;
  int hash_kind_ = 0;
  bool list_mode_ = false;  // list-method jobs: no tree passes
  bool low_root_ = false;   // some job uses the low-root rule
  uint64_t max_c_ = 0;      // chunks of the largest job
  uint32_t kinds_ = 0;      // insert modes (JOBS_DENSE | JOBS_LIST_NARROW | JOBS_LIST_WIDE) in the table
  bool bulk_dedup_ = false;  // some bulk job runs the leaf duplicate search
  int leaf_map_ = LEAF_DIRECT;  // leaf kernel mapping (LeafMapping), chosen by leaf_mapping_for
  bool copy_ = false;  // some job of the table has write_data: the leaf kernel's copy variant
  // Comparison of the tuner's candidate mappings on this pipeline's own
  // incremental checkpoints: each launch of a table with no baseline job
  // runs the candidate with the fewest samples, timed with events on the
  // leaf kernel's dispatch.
  struct LeafTrial {
    std::vector<uint64_t> key;
    std::vector<int> cand;
    std::vector<std::vector<float>> ms;
    int current = -1;  // candidate of the launch in flight
  } trial_;
  hipEvent_t trial_ev_[2] = {nullptr, nullptr};
  hipEvent_t trial_used_[2] = {nullptr, nullptr};  // the events on the leaf dispatch in flight
  static constexpr size_t kTrialRounds = 3;
  // The same for the incremental stages' batch-wide workgroup budget, once
  // the leaf mapping is settled: each candidate's whole-checkpoint time
  // (events before the first and after the last launch).
  struct BudgetTrial {
    std::vector<uint64_t> key;
    std::vector<std::vector<float>> ms;
    int current = -1;
  } btrial_;
  hipEvent_t bev_[2] = {nullptr, nullptr};
  bool btrial_end_ = false;  // the end event of the launch in flight is recorded
  uint64_t budget_ = 8192;

 public:
  int leaf_mapping() const { return leaf_map_; }
  // Job j of the last uploaded table had its chunks copied into the data
  // section by the leaf kernel.
  bool job_copied(size_t j) const {
    return j < njobs_ && reinterpret_cast<const JobDev*>(host_table_)[j].write_data != 0;
  }
  uint64_t device_bytes() const { return dev_table_.bytes; }
  // After the stream has completed a launch_all: records the trial launch's
  // leaf time and settles the mapping once every candidate has its rounds
  // (then the same for the stage budget).
  void trial_done();
  // Records the end of a checkpoint's launches for a budget trial; launch_all
  // does it itself for PART_ALL, the caller after a graph's launch.
  dd_status trial_mark_end(hipStream_t s);
  uint64_t stage_budget() const { return budget_; }
};

uint64_t max_diff_size(uint64_t len, uint32_t chunk_size, uint32_t ckpt_id);

// Leaf kernel mapping for a job table (dd_tune.cpp): the fastest mapping
// measured on this device for the table's hash, alignment, chunk sizes, job
// count and size class, timed on first use and cached for the process.
// `key` receives the table's tuning key; `cand`, while the key is not
// settled, the mappings still to be compared on checkpoints (fastest on the
// tuner's sample first).
int leaf_mapping_for(const std::vector<JobDev>& jobs, bool aligned16, int hash_kind, hipStream_t s,
                     std::vector<uint64_t>* key = nullptr, std::vector<int>* cand = nullptr);
// The mapping for a table whose baselines ask for the data-section copy
// (JobT::write_data): `current` when it is forced, makes the copy (DMA
// becomes DMA1, whose copy variant runs more wavefronts per SIMD), or the
// asking jobs hold less than half of the table's bytes; else DMA1, or LDS64
// where chunks are not a multiple of 128 bytes (a copy in the leaf kernel
// saves the gather a second read of the buffer, which costs more than the
// mapping's difference in hashing speed).
int leaf_mapping_with_copy(const std::vector<JobDev>& jobs, bool aligned16, int hash_kind, int current);
// Settles a key's mapping after its comparison on checkpoints.
void leaf_mapping_settle(const std::vector<uint64_t>& key, int mapping);
// Candidate batch-wide workgroup budgets of the incremental stages, the
// default first.
constexpr uint64_t kBudgetCand[] = {8192, 16384, 4096};
constexpr size_t kBudgetCands = sizeof(kBudgetCand) / sizeof(kBudgetCand[0]);
// The budget a key settled on (the default while `*settled` is false).
uint64_t stage_budget_for(const std::vector<uint64_t>& key, bool* settled);
void stage_budget_settle(const std::vector<uint64_t>& key, uint64_t budget);

}  // namespace dd

struct dd_dedup {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  dd::StreamState st;
  dd::Pipeline pipe;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  int profile = 0;  // 0 off, 1 leaf-hash events only, 2 every stage
  hipEvent_t pev[DD_NUM_STAGES + 1] = {};
  double stage_ms[DD_NUM_STAGES] = {};
  // restart buffers: cover table (8 B/chunk), pending sources (8 B/chunk),
  // pending list (4 B/chunk), header staging, arena (metadata staging, upload
  // block, host diff copies) and pinned host staging
  dd::DevBuf cover, ref, pend, rs_hdr, rs_arena;
  uint32_t stamp = 0;  // last stamp written into `cover`
  uint8_t* rs_host = nullptr;
  uint64_t rs_host_cap = 0;
};

namespace dd {
// Device bytes a handle holds now (its state, scratch and restart buffers).
uint64_t resident_bytes(const dd_dedup* h);
}  // namespace dd

struct dd_batch {
  int device = 0;
  std::vector<dd_dedup*> hs;
  hipStream_t stream = nullptr;
  dd::Pipeline pipe, pipe_rerun;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint64_t> sig;
  int profile = 0;  // as dd_dedup::profile, for the batch's sequence (level 2 runs without the graph)
  hipEvent_t pev[DD_NUM_STAGES + 1] = {};
  double stage_ms[DD_NUM_STAGES] = {};
};
