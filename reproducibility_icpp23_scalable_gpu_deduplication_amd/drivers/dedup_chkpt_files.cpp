// dedup_chkpt_files -- deduplicates a sequence of checkpoint files with the
// tree or list approach (the reference's src/dedup_files.cpp, tree and list
// modes).
//
//   dedup_chkpt_files chunk_size num_chkpts --run-tree-chkpt file0 file1 ...
//   dedup_chkpt_files chunk_size num_chkpts --run-list-chkpt file0 file1 ...
//
// Each file is read, copied to the GPU and checkpointed (the first one as the
// baseline); the diff goes to "<file>.hashtree.incr_chkpt" (list:
// "<file>.hashlist.incr_chkpt") and the CSV logs
// "<basename>.chunk_size.<chunk>.{csv,size.csv,timing.csv}" are appended in
// the working directory (src/tree_approach.cpp:1705-1738,1853-1915).
#include <stdio.h>
#include <stdlib.h>

#include <exception>
#include <string>
#include <vector>

#include "driver_util.hpp"

using namespace dd_drivers;

int main(int argc, char** argv) {
  if (argc < 5) {
    fprintf(stderr, "usage: %s chunk_size num_chkpts <mode> file...\n", argv[0]);
    print_mode_help();
    return 2;
  }
  const uint32_t chunk_size = (uint32_t)strtoul(argv[1], nullptr, 10);
  const uint32_t num_chkpts = (uint32_t)strtoul(argv[2], nullptr, 10);
  const Mode mode = get_mode(argc, argv);
  if (mode != Mode::Tree && mode != Mode::TreeLowRoot && mode != Mode::List) {
    fprintf(stderr, "ERROR: %s mode\n", mode == Mode::Other ? "unsupported (not tree or list)" : "incorrect");
    print_mode_help();
    return 2;
  }
  const int first = 4;  // argv[3] is the mode (src/dedup_files.cpp:36-45)
  if (chunk_size == 0 || argc < first + (int)num_chkpts) {
    fprintf(stderr, "ERROR: need chunk_size > 0 and %u checkpoint files\n", num_chkpts);
    return 2;
  }
  try {
    std::unique_ptr<dd::TreeDeduplicator> dedup = make_dedup(mode, chunk_size);
    uint8_t* dev = nullptr;
    size_t dev_cap = 0;
    for (uint32_t i = 0; i < num_chkpts; i++) {
      const std::string path = argv[first + i];
      const std::vector<uint8_t> data = read_file(path);
      if (data.empty()) throw std::runtime_error(path + ": empty checkpoint file");
      if (data.size() > dev_cap) {
        if (dev) hip_check(hipFree(dev), "hipFree");
        hip_check(hipMalloc(&dev, data.size()), "hipMalloc");
        dev_cap = data.size();
      }
      hip_check(hipMemcpy(dev, data.data(), data.size(), hipMemcpyHostToDevice), "hipMemcpy");
      std::string filename = path + file_suffix(mode);
      std::string logname = basename_of(path);
      dedup->checkpoint(dev, data.size(), filename, logname, i == 0);
      printf("%s: %zu bytes -> %s (first-occurrence regions %zu, shifted regions %zu)\n", path.c_str(),
             data.size(), filename.c_str(), dedup->num_first_ocur(), dedup->num_shift_dupl());
    }
    if (dev) hip_check(hipFree(dev), "hipFree");
  } catch (const std::exception& e) {
    fprintf(stderr, "ERROR: %s\n", e.what());
    return 1;
  }
  return 0;
}
