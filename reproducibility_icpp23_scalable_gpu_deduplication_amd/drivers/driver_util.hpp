// driver_util.hpp -- shared pieces of the file drivers (dedup_chkpt_files,
// restart_chkpt_files): mode flags, file reading, the restored data's digest.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "tree_deduplicator.hpp"

namespace dd_drivers {

// The reference's mode flags (src/utils.cpp:15-37).  Its drivers hand every
// tree mode to TreeDeduplicator (src/dedup_files.cpp:61-63) and the list
// mode to ListDeduplicator (:59-60).  Here the low-root flags select the
// low-root method (TreeLowRootDeduplicator, which labels with the low-root
// rule); the other tree flags the tree method.  The full and basic methods
// are not part of this library.
enum class Mode { Unknown, Tree, TreeLowRoot, List, Other };

inline Mode get_mode(int argc, char** argv) {
  static const char* tree_flags[] = {"--run-tree-chkpt", "--run-tree-low-offset-ref-chkpt",
                                     "--run-tree-low-offset-chkpt"};
  static const char* low_root_flags[] = {"--run-tree-low-root-ref-chkpt", "--run-tree-low-root-chkpt"};
  static const char* other_flags[] = {"--run-full-chkpt", "--run-basic-chkpt"};
  for (int i = 0; i < argc; i++) {
    for (const char* f : tree_flags)
      if (!strcmp(argv[i], f)) return Mode::Tree;
    for (const char* f : low_root_flags)
      if (!strcmp(argv[i], f)) return Mode::TreeLowRoot;
    if (!strcmp(argv[i], "--run-list-chkpt")) return Mode::List;
    for (const char* f : other_flags)
      if (!strcmp(argv[i], f)) return Mode::Other;
  }
  return Mode::Unknown;
}

inline void print_mode_help() {
  printf("Modes (the low-root modes pick low-root first occurrences, the other tree modes lowest offset):\n");
  printf("List checkpoint:                                   --run-list-chkpt\n");
  printf("Tree checkpoint:                                   --run-tree-chkpt\n");
  printf("Tree checkpoint lowest offset reference algorithm: --run-tree-low-offset-ref-chkpt\n");
  printf("Tree checkpoint lowest offset algorithm:           --run-tree-low-offset-chkpt\n");
  printf("Tree checkpoint lowest root reference algorithm:   --run-tree-low-root-ref-chkpt\n");
  printf("Tree checkpoint lowest root algorithm:             --run-tree-low-root-chkpt\n");
}

// The deduplicator of a mode, the suffix of its checkpoint files
// (src/dedup_files.cpp:94-99) and the label of its restart digest
// (src/restart_chkpt_files.cpp:165,194).
inline std::unique_ptr<dd::TreeDeduplicator> make_dedup(Mode m, uint32_t chunk_size) {
  if (m == Mode::List) return std::unique_ptr<dd::TreeDeduplicator>(new dd::ListDeduplicator(chunk_size));
  if (m == Mode::TreeLowRoot) return std::unique_ptr<dd::TreeDeduplicator>(new dd::TreeLowRootDeduplicator(chunk_size));
  return std::unique_ptr<dd::TreeDeduplicator>(new dd::TreeDeduplicator(chunk_size));
}
inline const char* file_suffix(Mode m) { return m == Mode::List ? ".hashlist.incr_chkpt" : ".hashtree.incr_chkpt"; }
inline const char* digest_label(Mode m) { return m == Mode::List ? "Hashlist digest:       " : "Hashtree digest:       "; }

inline std::vector<uint8_t> read_file(const std::string& name) {
  std::ifstream f;
  f.exceptions(std::ifstream::failbit | std::ifstream::badbit);
  f.open(name, std::ifstream::in | std::ifstream::binary | std::ifstream::ate);
  const std::streamsize n = f.tellg();
  f.seekg(0);
  std::vector<uint8_t> buf((size_t)n);
  if (n) f.read(reinterpret_cast<char*>(buf.data()), n);
  return buf;
}

inline std::string basename_of(const std::string& path) {
  const size_t p = path.rfind('/');
  return p == std::string::npos ? path : path.substr(p + 1);
}

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// MurmurHash3_x64_128 (seed 0) with the reference's 64-bit rotate that
// returns only its low 32 bits (include/kokkos_murmur3.hpp:30-31,213-292):
// the digest the reference's restart driver prints for the restored data.
namespace detail {
inline uint64_t rot(uint64_t x, int r) { return (uint64_t)(uint32_t)((x << r) | (x >> (64 - r))); }
inline uint64_t fmix(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}
inline uint64_t le64(const uint8_t* p, size_t n) {
  uint64_t v = 0;
  for (size_t i = 0; i < n; i++) v |= (uint64_t)p[i] << (8 * i);
  return v;
}
}  // namespace detail

inline void murmur3_digest(const uint8_t* data, size_t len, uint8_t out[16]) {
  using namespace detail;
  const uint64_t c1 = 0x87c37b91114253d5ULL, c2 = 0x4cf5ad432745937fULL;
  uint64_t h1 = 0, h2 = 0;
  const size_t nblocks = len / 16;
  for (size_t i = 0; i < nblocks; i++) {
    uint64_t k1 = le64(data + 16 * i, 8), k2 = le64(data + 16 * i + 8, 8);
    k1 *= c1; k1 = rot(k1, 31); k1 *= c2; h1 ^= k1;
    h1 = rot(h1, 27); h1 += h2; h1 = h1 * 5 + 0x52dce729;
    k2 *= c2; k2 = rot(k2, 33); k2 *= c1; h2 ^= k2;
    h2 = rot(h2, 31); h2 += h1; h2 = h2 * 5 + 0x38495ab5;
  }
  const uint8_t* tail = data + 16 * nblocks;
  const size_t rem = len & 15;
  if (rem > 8) {
    uint64_t k2 = le64(tail + 8, rem - 8);
    k2 *= c2; k2 = rot(k2, 33); k2 *= c1; h2 ^= k2;
  }
  if (rem) {
    uint64_t k1 = le64(tail, rem < 8 ? rem : 8);
    k1 *= c1; k1 = rot(k1, 31); k1 *= c2; h1 ^= k1;
  }
  h1 ^= len; h2 ^= len;
  h1 += h2; h2 += h1;
  h1 = fmix(h1); h2 = fmix(h2);
  h1 += h2; h2 += h1;
  for (int i = 0; i < 8; i++) {
    out[i] = (uint8_t)(h1 >> (8 * i));
    out[8 + i] = (uint8_t)(h2 >> (8 * i));
  }
}

// Hex with a space after every 4 bytes, as the reference prints digests
// (include/hash_functions.hpp:47-63).
inline std::string digest_string(const uint8_t d[16]) {
  static const char hexchars[] = "0123456789ABCDEF";
  std::string s;
  for (int k = 0; k < 16; k++) {
    s.push_back(hexchars[d[k] >> 4]);
    s.push_back(hexchars[d[k] & 0xF]);
    if (k % 4 == 3) s.push_back(' ');
  }
  return s;
}

}  // namespace dd_drivers
