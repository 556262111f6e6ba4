// tree_deduplicator.hpp -- C++ host interface mirroring the reference's
// TreeDeduplicator (include/tree_approach.hpp:16-205, include/
// deduplicator_interface.hpp:14-184), TreeLowRootDeduplicator
// (include/tree_approach.hpp:207-218) and ListDeduplicator
// (include/list_approach.hpp:21-200), implemented over the C ABI in
// dedup_c.h.  Kokkos views are replaced by raw device pointers plus lengths
// (device data) and std::vector<uint8_t> (host diffs).  Errors raise
// std::runtime_error where the reference would print or abort.
#pragma once

#include <stdint.h>

#include <string>
#include <utility>
#include <vector>

#include "dedup_c.h"

namespace dd {

using header_t = dd_header;
using HostBuf = std::vector<uint8_t>;

enum class HashKind : int { Murmur3 = DD_HASH_MURMUR3, MD5 = DD_HASH_MD5 };

class TreeDeduplicator {
 public:
  // TreeDeduplicator(uint32_t bytes_per_chunk), src/tree_approach.cpp.
  explicit TreeDeduplicator(uint32_t chunk_size, HashKind kind = HashKind::Murmur3, void* hip_stream = nullptr);
  ~TreeDeduplicator();
  TreeDeduplicator(const TreeDeduplicator&) = delete;
  TreeDeduplicator& operator=(const TreeDeduplicator&) = delete;

  // src/tree_approach.cpp:1577 -- does not advance current_id.
  void checkpoint(header_t& header, uint8_t* data_ptr, size_t data_len, HostBuf& diff_h, bool make_baseline);
  // src/tree_approach.cpp:1705 -- writes `filename` and the logs, advances current_id.
  void checkpoint(uint8_t* data_ptr, size_t len, std::string& filename, std::string& logname, bool make_baseline);
  // src/tree_approach.cpp:1740 -- advances current_id.
  void checkpoint(uint8_t* data_ptr, size_t len, HostBuf& diff_h, bool make_baseline);
  // src/tree_approach.cpp:1765 -- writes the logs, advances current_id.
  void checkpoint(uint8_t* data_ptr, size_t len, HostBuf& diff_h, std::string& logname, bool make_baseline);

  // src/tree_approach.cpp:1809 -- rebuilds checkpoint `chkpt_id` into the
  // device buffer data_ptr (len must equal the header's datalen).
  void restart(uint8_t* data_ptr, size_t len, std::vector<HostBuf>& chkpts, std::string& logname, uint32_t chkpt_id);
  // src/tree_approach.cpp:1827 -- reads "<name>.hashtree.incr_chkpt" files
  // (list handles: "<name>.hashlist.incr_chkpt").
  void restart(uint8_t* data_ptr, size_t len, std::vector<std::string>& chkpt_filenames, std::string& logname,
               uint32_t chkpt_id);

  // Checkpoints one buffer per deduplicator in one call (dd_batch_*: one
  // launch per stage over all buffers) and writes each diff to filenames[i]
  // and its logs under lognames[i], as checkpoint(ptr, len, filename,
  // logname, make_baseline) does for one; advances every current_id.  The
  // deduplicators share the device, hash and method.  No counterpart in the
  // reference, whose multi-rank driver (src/dedup_files_mpi.cpp:54-120) runs
  // one deduplicator per MPI rank.
  static void checkpoint_batch(const std::vector<TreeDeduplicator*>& dedups, const std::vector<uint8_t*>& data,
                               const std::vector<size_t>& lens, std::vector<std::string>& filenames,
                               std::vector<std::string>& lognames, bool make_baseline);

  // include/tree_approach.hpp:53-59
  size_t num_first_ocur();
  size_t num_shift_dupl();

  // src/tree_approach.cpp:1853-1915
  void write_chkpt_log(header_t& header, HostBuf& diff_h, std::string& logname);
  void write_restart_log(uint32_t select_chkpt, std::string& logname);

  uint32_t current_id() const;
  uint32_t baseline_id() const;
  void set_current_id(uint32_t id);
  dd_dedup* handle() const { return h_; }

 protected:
  // A handle of method `method` (dd_method); `approach` names it in the logs
  // and `suffix` is appended to checkpoint file names.
  TreeDeduplicator(uint32_t chunk_size, HashKind kind, void* hip_stream, int method, const char* approach,
                   const char* suffix);

 private:
  dd_dedup* h_ = nullptr;
  uint32_t chunk_size_;
  int method_ = DD_METHOD_TREE;
  std::string approach_ = "Tree", suffix_ = ".hashtree.incr_chkpt";
  uint64_t data_len_ = 0;
  // Tree: (data incl. header, metadata), src/tree_approach.cpp:706-709;
  // list: (data, header + metadata), src/list_approach.cpp:249-254.
  std::pair<uint64_t, uint64_t> datasizes_{0, 0};
};

// TreeLowRootDeduplicator(uint32_t), src/tree_low_root_approach.cpp:8-12:
// the tree method with low-root first occurrences (DD_METHOD_TREE_LOW_ROOT,
// the rule of dedup_low_root_ref, include/reference_impl.hpp:265-472, made
// deterministic; the reference's own class runs the low-offset labeling,
// :393-399).  Tree diffs, ".hashtree.incr_chkpt" files and "Tree" log rows.
class TreeLowRootDeduplicator : public TreeDeduplicator {
 public:
  explicit TreeLowRootDeduplicator(uint32_t chunk_size, HashKind kind = HashKind::Murmur3, void* hip_stream = nullptr)
      : TreeDeduplicator(chunk_size, kind, hip_stream, DD_METHOD_TREE_LOW_ROOT, "Tree", ".hashtree.incr_chkpt") {}
};

// ListDeduplicator(uint32_t bytes_per_chunk), src/list_approach.cpp:5-9: the
// same surface; diffs in the list layout, "<name>.hashlist.incr_chkpt" files
// (src/list_approach.cpp:996-1010) and "List" log rows.
class ListDeduplicator : public TreeDeduplicator {
 public:
  explicit ListDeduplicator(uint32_t chunk_size, HashKind kind = HashKind::Murmur3, void* hip_stream = nullptr)
      : TreeDeduplicator(chunk_size, kind, hip_stream, DD_METHOD_LIST, "List", ".hashlist.incr_chkpt") {}
};

// src/utils.cpp:38-124: header, first-occurrence metadata and per-source-
// checkpoint shifted metadata sizes of one diff, as one CSV row.
std::string metadata_breakdown(const header_t& header, const HostBuf& diff, uint32_t num_chkpts);

}  // namespace dd
