/*
 * dedup_oracle.h -- CPU oracle for the tree-approach incremental checkpoint.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never links it.
 *
 * It is a serial restatement (plain C11) of the reference's
 *   - chunk hashes     include/kokkos_murmur3.hpp:213-292 (with the truncating
 *                      rotl64 of :30-31) and include/kokkos_md5.hpp:48-323
 *   - serial labeler   include/reference_impl.hpp:15-132 (dedup_low_offset_ref)
 *   - diff serializer  src/tree_approach.cpp:511-710 (collect_diff), emitted in
 *                      one canonical order (first-occurrence roots ascending,
 *                      prior-checkpoint table ascending, shifted roots grouped
 *                      by source checkpoint ascending then node ascending)
 *   - restart          src/tree_approach.cpp:712-1128 (restart_chkpt)
 *   - orchestration    src/tree_approach.cpp:1577-1689 and :1740-1749
 *
 * Chunk counts that are not a power of two: the reference's subtree leaf range
 * (include/kokkos_merkle_tree.hpp:102-142) is wrong there.  Regions are
 * instead enumerated in in-order (left-to-right) leaf order: the run of the
 * subtree's leaves on the deepest level followed by its run on the level
 * above.  For power-of-two chunk counts this is the reference's contiguous
 * range, byte for byte.
 */
#ifndef DEDUP_ORACLE_H
#define DEDUP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_HASH_MURMUR3 = 0, OR_HASH_MD5 = 1 };

enum {
  OR_OK = 0,
  OR_EINVAL = 1,
  OR_ENOMEM = 2,
  OR_ECORRUPT = 3
};

/* Labels, include/utils.hpp:51-57. */
enum { OR_FIRST_OCUR = 0, OR_FIXED_DUPL = 1, OR_SHIFT_DUPL = 2, OR_FIRST_DUPL = 3, OR_DONE = 4 };

typedef struct or_dedup or_dedup;

/* Hash `len` bytes with the reference's chunk hash (seed 0). */
void or_hash(int kind, const void* data, uint64_t len, uint8_t out[16]);

or_dedup* or_create(uint32_t chunk_size, int hash_kind);
void or_destroy(or_dedup* d);
/* Options of this build (no counterpart in the reference).
 * OR_OPT_EMIT_ROOT_SHIFT: emit node 0 as a shifted root when it is one.
 * OR_OPT_LOW_ROOT: the low-root first-occurrence rule of
 * TreeLowRootDeduplicator (include/reference_impl.hpp:265-472), restated as a
 * deterministic rule (dedup_oracle.c, low_root_select). */
enum { OR_OPT_EMIT_ROOT_SHIFT = 1, OR_OPT_LOW_ROOT = 2 };
void or_set_options(or_dedup* d, uint32_t options);

/* TreeDeduplicator::checkpoint(data, len, diff_h, make_baseline)
 * (src/tree_approach.cpp:1740-1749): builds the diff and advances the
 * checkpoint id.  *out is malloc'd; release it with or_free. */
int or_checkpoint(or_dedup* d, const uint8_t* data, uint64_t len, int make_baseline,
                  uint8_t** out, uint64_t* out_len);

uint32_t or_current_id(const or_dedup* d);
/* num_first_ocur() / num_shift_dupl() of the last checkpoint
 * (include/tree_approach.hpp:53-59). */
void or_counts(const or_dedup* d, uint64_t* nfirst, uint64_t* nshift);
/* Labels (2C-1 bytes) and leaf digests (16*C bytes) of the last checkpoint. */
int or_last_labels(const or_dedup* d, uint8_t* out, uint64_t n);
int or_last_leaf_digests(const or_dedup* d, uint8_t* out, uint64_t n);
uint64_t or_map_size(const or_dedup* d);

/* TreeDeduplicator::restart(...) on a chain of diffs indexed by checkpoint id
 * (src/tree_approach.cpp:712-1128).  out_len must equal the header datalen. */
int or_restart(const uint8_t* const* chain, const uint64_t* lens, uint32_t nchain,
               uint32_t id, uint8_t* out, uint64_t out_len);

/* In-order leaf node ids of the subtree rooted at `node` in a heap of 2C-1
 * nodes.  Returns the count; writes them to out when out != NULL. */
uint64_t or_subtree_leaves(uint32_t node, uint32_t num_chunks, uint32_t* out);

void or_free(void* p);

/* List approach (ListDeduplicator, src/list_approach.cpp:23-256), serial:
 * one digest per chunk compared with the previous checkpoint's list; a
 * changed chunk is inserted into the digest map as (chunk, ckpt) and is a
 * first occurrence when the insert creates the entry, a shifted duplicate
 * otherwise.  The diff lists chunk indices: first occurrences ascending, the
 * prior table ascending, shifted entries grouped by source checkpoint then
 * chunk (the order a one-thread run of the reference produces).  An
 * incremental checkpoint keeps the chain's length (OR_EINVAL otherwise). */
typedef struct or_list or_list;
or_list* or_list_create(uint32_t chunk_size, int hash_kind);
void or_list_destroy(or_list* d);
int or_list_checkpoint(or_list* d, const uint8_t* data, uint64_t len, int make_baseline,
                       uint8_t** out, uint64_t* out_len);
void or_list_counts(const or_list* d, uint64_t* nfirst, uint64_t* nshift);
uint32_t or_list_current_id(const or_list* d);
/* ListDeduplicator::restart_chkpt (src/list_approach.cpp:258-512): the
 * per-chunk (node, ckpt) walk from checkpoint `id` down to its baseline.
 * Chunks no checkpoint of the chain names are an error (OR_ECORRUPT). */
int or_list_restart(const uint8_t* const* chain, const uint64_t* lens, uint32_t nchain,
                    uint32_t id, uint8_t* out, uint64_t out_len);

/* Parallel (OpenMP) restatement of or_checkpoint (dedup_omp.c): the same diff
 * bytes, computed on omp_get_max_threads() host threads.  CPU baseline only. */
typedef struct om_dedup om_dedup;
om_dedup* om_create(uint32_t chunk_size, int hash_kind);
void om_destroy(om_dedup* d);
int om_checkpoint(om_dedup* d, const uint8_t* data, uint64_t len, int make_baseline,
                  uint8_t** out, uint64_t* out_len);
void om_counts(const om_dedup* d, uint64_t* nfirst, uint64_t* nshift);
uint32_t om_current_id(const om_dedup* d);
/* Host threads om_checkpoint uses. */
int om_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
