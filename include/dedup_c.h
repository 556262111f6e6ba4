/*
 * dedup_c.h -- C ABI of the MI355X tree-approach incremental-checkpoint
 * deduplicator (libdedup_hip.so).
 *
 * Plain pointers and sizes only.  Device pointers are HIP device addresses on
 * the handle's device; host buffers are ordinary process memory.  Every entry
 * point returns a dd_status; none prints.
 *
 * Each function names the reference interface it replaces
 * (TauferLab ICPP'23 GPU deduplication; paths relative to that repository).
 *
 * Threading: calls on one handle must be serialized.  Different handles may be
 * used concurrently from different host threads; each handle owns one HIP
 * stream and its own device-resident tree and digest map.
 */
#ifndef DEDUP_C_H
#define DEDUP_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum dd_status {
  DD_OK = 0,
  DD_EINVAL = 1,    /* bad argument (length mismatch, zero length, bad handle) */
  DD_ENOMEM = 2,    /* host or device allocation failed */
  DD_ECORRUPT = 3,  /* a diff in the chain is inconsistent or the chain is short */
  DD_EHIP = 4,      /* a HIP runtime call failed */
  DD_ESTATE = 5,    /* call order violates the checkpoint chain (see dd_checkpoint) */
  DD_ELIMIT = 6     /* the request exceeds the fixed scratch of an allocation-free entry point */
} dd_status;

typedef enum dd_hash_kind {
  DD_HASH_MURMUR3 = 0, /* include/kokkos_murmur3.hpp:213-292 incl. the truncating rotl64 (:30-31) */
  DD_HASH_MD5 = 1      /* include/kokkos_md5.hpp:287-323 */
} dd_hash_kind;

/* Deduplication method of a handle (the reference's DedupMode,
 * include/utils.hpp:40-49, restricted to the approaches built here). */
typedef enum dd_method {
  DD_METHOD_TREE = 0, /* TreeDeduplicator, src/tree_approach.cpp: low-offset
                         first occurrences */
  DD_METHOD_LIST = 1, /* ListDeduplicator, src/list_approach.cpp:23-256: one
                         digest per chunk, no tree; diff node ids are chunk
                         indices */
  DD_METHOD_TREE_LOW_ROOT = 2 /* TreeLowRootDeduplicator: the tree method with
                         low-root first occurrences, the rule of
                         dedup_low_root_ref (include/reference_impl.hpp:265-472)
                         made deterministic: among the changed leaves whose
                         digest has no entry from an earlier checkpoint, the
                         first occurrence of a digest is the leaf whose highest
                         ancestor with only such leaves below has the smallest
                         node id, ties to the lowest leaf.  Tree diffs and
                         restarts.  (The reference's TreeLowRootDeduplicator
                         runs the low-offset labeling,
                         src/tree_low_root_approach.cpp:393-399.) */
} dd_method;

/* header_t, include/utils.hpp:29-37 (32 bytes, little-endian). */
typedef struct dd_header {
  uint32_t ref_id;
  uint32_t chkpt_id;
  uint64_t datalen;
  uint32_t chunk_size;
  uint32_t num_first_ocur;
  uint32_t num_prior_chkpts;
  uint32_t num_shift_dupl;
} dd_header;

/* A host byte buffer; buffers returned by the library are released with
 * dd_buf_free (Kokkos::View<uint8_t*>::HostMirror in the reference). */
typedef struct dd_buf {
  uint8_t* data;
  uint64_t len;
} dd_buf;

typedef struct dd_dedup dd_dedup;

/* TreeDeduplicator(uint32_t bytes_per_chunk), include/tree_approach.hpp:16-205,
 * with the hash chosen at run time instead of at compile time
 * (include/hash_functions.hpp:16-20).  `hip_stream` may be NULL: the handle
 * then creates its own stream on the current device. */
dd_status dd_create(uint32_t chunk_size, int hash_kind, void* hip_stream, dd_dedup** out);
/* dd_create with a method: DD_METHOD_TREE is dd_create; DD_METHOD_LIST is
 * ListDeduplicator(uint32_t bytes_per_chunk), src/list_approach.cpp:5-9.
 * Every other entry point takes handles of either method; a list handle's
 * diffs use the list layout (src/list_approach.cpp:64-256) and its restarts
 * read list diffs.  The chunk count of a list chain is fixed between
 * baselines (DD_ESTATE otherwise), as for the tree. */
dd_status dd_create_method(uint32_t chunk_size, int hash_kind, int method, void* hip_stream, dd_dedup** out);
/* Options of a handle (no counterpart in the reference), applied from its
 * next checkpoint on.  DD_OPT_EMIT_ROOT_SHIFT: when a whole buffer equals an
 * earlier checkpoint its root (node 0) is a shifted duplicate, which the
 * reference's labeling does not emit (include/reference_impl.hpp:105-106:
 * node 0 becomes a root only as a first occurrence), so the diff carries no
 * region and a restart of that checkpoint returns its predecessor's bytes;
 * with the option node 0 is emitted as a shifted root and the checkpoint
 * restarts to its own bytes.  The reference's restart reads such an entry
 * like any other (src/tree_approach.cpp:887-925).  Default 0: the
 * reference's bytes. */
#define DD_OPT_EMIT_ROOT_SHIFT 1u
/* DD_OPT_DENSE_BASELINE: a baseline builds its digest map with one map
 * insert per leaf and per first-occurrence interior node, as an incremental
 * checkpoint does, instead of partitioning the entries by map bucket and
 * writing each bucket once.  The diff bytes are the same either way. */
#define DD_OPT_DENSE_BASELINE 2u
dd_status dd_set_options(dd_dedup* h, uint32_t options);
dd_status dd_get_options(dd_dedup* h, uint32_t* options);
/* The method a handle was created with. */
dd_status dd_get_method(dd_dedup* h, int* method);
dd_status dd_destroy(dd_dedup* h);

/* TreeDeduplicator::checkpoint(uint8_t* data_ptr, size_t len, HostMirror& diff_h,
 * bool make_baseline), src/tree_approach.cpp:1740-1749: deduplicates the device
 * buffer against the handle's previous checkpoint, returns the diff on the
 * host in *out (header + metadata + data), and advances the checkpoint id.
 * An incremental checkpoint (make_baseline == 0 after the first call) must
 * keep the buffer length of the chain; otherwise DD_ESTATE.
 * Errors (every checkpoint entry point, batched ones included): a call
 * rejected by its argument checks (DD_EINVAL, DD_ESTATE) leaves the handle as
 * it was.  A call that fails after it started to change the handle's device
 * state (DD_ENOMEM, DD_ECORRUPT, DD_EHIP) leaves the handle needing a
 * baseline: its next checkpoint runs as one whatever make_baseline says. */
dd_status dd_checkpoint(dd_dedup* h, const void* dev_data, uint64_t len, int make_baseline,
                        dd_buf* out);

/* TreeDeduplicator::checkpoint(header_t&, uint8_t*, size_t, HostMirror&, bool),
 * src/tree_approach.cpp:1577-1689: as dd_checkpoint but also returns the
 * header and does NOT advance the checkpoint id (callers advance it with
 * dd_set_current_id, as the reference's wrapper overloads do). */
dd_status dd_checkpoint_header(dd_dedup* h, dd_header* header, const void* dev_data, uint64_t len,
                               int make_baseline, dd_buf* out);

/* Device-resident variant of dd_checkpoint (the reference's collect_diff
 * output buffer_d before its D2H copy, src/tree_approach.cpp:1668-1684): the
 * diff is written to `dev_out` (capacity `out_cap`, at least
 * dd_max_diff_size(len, chunk_size, current id)) and its length to *out_len.
 * Bytes of `dev_out` past *out_len, within that bound, may be overwritten
 * (a baseline's leaf kernel copies every chunk into the data section before
 * the labels decide the diff's layout).  Advances the checkpoint id. */
dd_status dd_checkpoint_device(dd_dedup* h, const void* dev_data, uint64_t len, int make_baseline,
                               void* dev_out, uint64_t out_cap, uint64_t* out_len);

/* Upper bound of a diff's size: header, metadata for every chunk, the prior
 * checkpoint table (ckpt_id+1 entries) and every chunk's data. */
uint64_t dd_max_diff_size(uint64_t len, uint32_t chunk_size, uint32_t ckpt_id);

/* num_first_ocur() / num_shift_dupl() of the last checkpoint,
 * include/tree_approach.hpp:53-59. */
dd_status dd_counts(dd_dedup* h, uint64_t* num_first_ocur, uint64_t* num_shift_dupl);

/* BaseDeduplicator::current_id / baseline_id, include/deduplicator_interface.hpp:16-18. */
dd_status dd_get_ids(dd_dedup* h, uint32_t* current_id, uint32_t* baseline_id);
dd_status dd_set_current_id(dd_dedup* h, uint32_t current_id);

/* BaseDeduplicator::timers[4] (setup, dedup, gather, copy-to-host, seconds) and
 * restart_timers[2] (copy, total), include/deduplicator_interface.hpp:22-23.
 * The GPU time is split into dedup and gather only at profiling level 2
 * (dd_set_profiling(h, 2): the split costs an event between the stages);
 * otherwise timers[1] holds the whole GPU time and timers[2] is 0. */
dd_status dd_get_timers(dd_dedup* h, double timers[4], double restart_timers[2]);

/* TreeDeduplicator::restart(uint8_t* data_ptr, size_t len,
 * std::vector<HostMirror>& chkpts, std::string& logname, uint32_t chkpt_id),
 * src/tree_approach.cpp:1809-1816 -> restart_chkpt :712-1128: rebuilds
 * checkpoint `chkpt_id` from host diffs indexed by checkpoint id into the
 * device buffer `dev_out`, whose length must equal the header's datalen
 * (DD_EINVAL otherwise). */
dd_status dd_restart(dd_dedup* h, const dd_buf* chain, uint32_t nchain, uint32_t chkpt_id,
                     void* dev_out, uint64_t len);

/* Same as dd_restart with the diffs already in device memory
 * (dev_chain[i] holds chain_lens[i] bytes; both arrays are host arrays). */
dd_status dd_restart_device(dd_dedup* h, const void* const* dev_chain, const uint64_t* chain_lens,
                            uint32_t nchain, uint32_t chkpt_id, void* dev_out, uint64_t len);

/* Allocation-free restart (no counterpart in the reference): for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, checks and bytes
 * as dd_restart_device, but every device buffer the walk needs (cover table,
 * pending chunks, staging, parse arrays) comes from one static device array
 * of DD_NOALLOC_SCRATCH_BYTES per device, which the library holds until the
 * process ends; the call allocates nothing.  Limits: at most
 * DD_NOALLOC_MAX_CHUNKS chunks (ceil(len / chunk_size); 1 GiB at 128-byte
 * chunks fits) and DD_NOALLOC_ARENA_BYTES of per-chain arrays (about 4 bytes
 * per region entry plus 8 bytes per 256 chunks for each checkpoint walked);
 * past either limit the call returns DD_ELIMIT before any kernel runs.
 * Calls on one device share its array, so the library serializes them (a
 * process-wide lock per device); calls from different processes use their
 * own copies. */
#define DD_NOALLOC_SCRATCH_BYTES (256ull << 20)
#define DD_NOALLOC_MAX_CHUNKS (10ull << 20)
#define DD_NOALLOC_ARENA_BYTES (DD_NOALLOC_SCRATCH_BYTES - 4096ull - 20ull * DD_NOALLOC_MAX_CHUNKS)
dd_status dd_restart_device_noalloc(dd_dedup* h, const void* const* dev_chain, const uint64_t* chain_lens,
                                    uint32_t nchain, uint32_t chkpt_id, void* dev_out, uint64_t len);

/* Batched checkpointing (no counterpart in the reference, whose multi-rank
 * driver src/dedup_files_mpi.cpp:54-120 runs one TreeDeduplicator per MPI
 * rank and checkpoints each buffer on its own).  A batch groups handles; one
 * batch call makes one checkpoint of every handle (same semantics and bytes as
 * dd_checkpoint_device on each), running each stage as a single launch over
 * all buffers.  With graphs enabled the launch sequence is captured once per
 * buffer-shape set and replayed.  The handles must not be used elsewhere
 * while a batch call runs; the batch has its own HIP stream.  All handles of
 * a batch share the device, the hash and the method (DD_EINVAL otherwise). */
typedef struct dd_batch dd_batch;
dd_status dd_batch_create(dd_dedup* const* handles, uint32_t n, dd_batch** out);
dd_status dd_batch_destroy(dd_batch* b);
dd_status dd_batch_use_graph(dd_batch* b, int enable);
/* Arrays have one entry per handle.  dev_out[i] must hold at least
 * dd_max_diff_size(lens[i], chunk_size, current id of handle i) bytes; as in
 * dd_checkpoint_device, bytes past out_lens[i] within that bound may be
 * overwritten. */
dd_status dd_batch_checkpoint_device(dd_batch* b, const void* const* dev_data, const uint64_t* lens,
                                     const int* make_baseline, void* const* dev_out, const uint64_t* out_caps,
                                     uint64_t* out_lens);

/* Serving (no counterpart in the reference): a request loop over many
 * checkpoint streams (one dd_dedup handle per simulated rank).  Requests are
 * admitted in arrival order while an execution slot is free, the rank has no
 * request in flight, and the request's device-memory footprint fits the HBM
 * budget; a request that does not fit waits for completions and holds back
 * the requests behind it.  With nothing in flight the first eligible request
 * always starts, so a request larger than the budget runs alone.  Admitted
 * work runs concurrently on the execution slots: checkpoint requests admitted
 * together run as one batched job table (dd_server_set_batching), a restart
 * or a lone checkpoint on its handle's HIP stream.  Results are the same bytes
 * dd_checkpoint_device / dd_restart_device return.  While a handle has a
 * request queued or in flight, the caller must not use it directly. */
typedef struct dd_server dd_server;
typedef struct dd_request_result {
  int32_t status;      /* dd_status of the request */
  uint32_t kind;       /* 0 checkpoint, 1 restart */
  uint64_t out_len;    /* diff length (checkpoint) or restored bytes (restart) */
  double t_submit;     /* seconds, steady clock */
  double t_start;      /* admitted and started */
  double t_end;        /* finished */
  uint32_t group;      /* checkpoint requests run in one batched job table with it (1: alone) */
  uint32_t pad;
} dd_request_result;
/* `footprint` of each request: its input and output buffers plus, on a
 * handle's first request, the handle's state estimate. */
dd_status dd_server_create(uint64_t hbm_budget_bytes, uint32_t slots, dd_server** out);
dd_status dd_server_destroy(dd_server* s); /* waits for every request */
dd_status dd_server_submit_checkpoint(dd_server* s, dd_dedup* h, const void* dev_data, uint64_t len,
                                      int make_baseline, void* dev_out, uint64_t out_cap, uint64_t* ticket);
dd_status dd_server_submit_restart(dd_server* s, dd_dedup* h, const void* const* dev_chain,
                                   const uint64_t* chain_lens, uint32_t nchain, uint32_t chkpt_id,
                                   void* dev_out, uint64_t len, uint64_t* ticket);
/* Blocks until the request is done; each ticket can be waited once (a second
 * wait, also one from another thread while the first blocks, returns
 * DD_EINVAL).  The budget charges each handle, after each of its requests,
 * the device bytes it holds (state, scratch and restart buffers); a handle's
 * first checkpoint is admitted with an estimate of that state. */
dd_status dd_server_wait(dd_server* s, uint64_t ticket, dd_request_result* out);
/* Checkpoint requests admitted together into one batched job table (no
 * counterpart in the reference, whose ranks each run their own checkpoint):
 * a slot that admits a checkpoint also admits the queued checkpoints of other
 * idle handles of the same device, hash and method, in queue order while they
 * fit the budget, at most max_group (default 64; 1 runs every request alone). */
dd_status dd_server_set_batching(dd_server* s, uint32_t max_group);
/* Bytes currently reserved against the budget, and the peak so far. */
dd_status dd_server_usage(dd_server* s, uint64_t* in_use, uint64_t* peak);
double dd_now(void); /* the steady clock of dd_request_result */

/* Diagnostics (no counterpart in the reference; its equivalent is the
 * Kokkos::Profiling regions of src/tree_approach.cpp).  Level 1 times the
 * leaf-hash kernel only, with HIP events attached to its own dispatch (no
 * extra work on the stream); level 2 records events between every two stages
 * (events in the middle of the stream cost tens of microseconds of host time
 * per checkpoint).  Level 0 turns them off.  dd_get_stage_times returns the last checkpoint's GPU
 * milliseconds for {leaf hash, leaf inserts, first-occurrence forest,
 * interior inserts, duplicate forest, roots + grouping, metadata + gather};
 * stages not recorded read 0. */
#define DD_NUM_STAGES 7
dd_status dd_set_profiling(dd_dedup* h, int level);
dd_status dd_get_stage_times(dd_dedup* h, double ms[DD_NUM_STAGES]);
/* Leaf kernel mapping (no counterpart in the reference) of the last
 * checkpoint of a handle or batch: 0 one lane per chunk, 1 the same with
 * deeper load pipelining, 2/3/4 LDS-staged in 64/128/256-byte segments, 5 one
 * lane per chunk of at most 128 bytes with the whole chunk in flight, 6
 * 128-byte segments loaded straight into LDS with contiguous loads, 7 the
 * same with one-wavefront workgroups, 8 LDS-staged 128-byte segments with
 * the next segment's loads in flight.  The
 * library picks it by timing the valid mappings once per process for each
 * hash, alignment and chunk-size combination (DD_LEAF_KERNEL=direct|deep|
 * lds64|lds128|lds256|short|dma|dma1|lds128p forces one). */
dd_status dd_get_leaf_mapping(dd_dedup* h, int* mapping);
/* Profiling of a batch's stage sequence (no counterpart in the reference):
 * as dd_set_profiling / dd_get_stage_times for the last batch call; the
 * events are recorded on the batch's stream (inside its graph when graphs
 * are enabled). */
dd_status dd_batch_set_profiling(dd_batch* b, int level);
dd_status dd_batch_get_stage_times(dd_batch* b, double ms[DD_NUM_STAGES]);
/* Sizes of a handle's last completed checkpoint (no counterpart in the
 * reference; what the algorithmic byte counts of a roofline need). */
typedef struct dd_ckpt_stats {
  uint64_t chunks;        /* chunks of the buffer */
  uint64_t changed;       /* leaves whose digest was rewritten: changed leaves, every leaf at a baseline */
  uint64_t first_chunks;  /* chunks copied into the diff's data section */
  uint64_t diff_bytes;    /* bytes of the diff */
  uint32_t baseline;      /* 1 when it ran as a baseline */
  uint32_t leaf_copy;     /* 1 when the leaf kernel also wrote the data section (a baseline) */
} dd_ckpt_stats;
dd_status dd_get_ckpt_stats(dd_dedup* h, dd_ckpt_stats* out);
/* Device bytes the handle holds now: tree, digest map, scratch arrays and
 * restart buffers (no counterpart in the reference; what dd_server charges a
 * handle against its budget). */
dd_status dd_get_device_bytes(dd_dedup* h, uint64_t* bytes);
dd_status dd_batch_get_leaf_mapping(dd_batch* b, int* mapping);
/* Workgroup budget (no counterpart in the reference) that the last
 * checkpoint of a handle or batch gave each incremental stage over all its
 * buffers: a stage whose per-buffer workgroups add up to more is scaled down
 * to it.  Compared on the first incremental checkpoints of a kind of table
 * (8192, 16384 or 4096 workgroups of 256 lanes, four times that for the
 * tree-tile passes) and then fixed for the process; DD_STAGE_BUDGET forces
 * one. */
dd_status dd_get_stage_budget(dd_dedup* h, uint64_t* workgroups);
dd_status dd_batch_get_stage_budget(dd_batch* b, uint64_t* workgroups);

void dd_buf_free(dd_buf* b);
const char* dd_strerror(int status);

#ifdef __cplusplus
}
#endif
#endif
