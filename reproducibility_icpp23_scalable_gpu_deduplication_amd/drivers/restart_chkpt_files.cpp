// restart_chkpt_files -- rebuilds one checkpoint from its chain of
// ".hashtree.incr_chkpt" (list mode: ".hashlist.incr_chkpt") files (the
// reference's src/restart_chkpt_files.cpp, tree and list modes) and prints
// the digest of the restored data.
//
//   restart_chkpt_files restart_id num_chkpts num_tests chunk_size --run-tree-chkpt file0 file1 ...
//   restart_chkpt_files restart_id num_chkpts num_tests chunk_size --run-list-chkpt file0 file1 ...
//
// The restored length is the original checkpoint file's size, or the diff
// header's data length when that file is absent.  The restart runs
// num_tests times (restart timing is appended to
// "<basename>.chunk_size.<chunk>.restart_timing.csv").
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <exception>
#include <string>
#include <vector>

#include "driver_util.hpp"

using namespace dd_drivers;

int main(int argc, char** argv) {
  if (argc < 7) {
    fprintf(stderr, "usage: %s restart_id num_chkpts num_tests chunk_size <mode> file...\n", argv[0]);
    print_mode_help();
    return 2;
  }
  const uint32_t restart_id = (uint32_t)strtoul(argv[1], nullptr, 10);
  const uint32_t num_chkpts = (uint32_t)strtoul(argv[2], nullptr, 10);
  const uint32_t num_tests = (uint32_t)strtoul(argv[3], nullptr, 10);
  const uint32_t chunk_size = (uint32_t)strtoul(argv[4], nullptr, 10);
  const Mode mode = get_mode(argc, argv);
  if (mode != Mode::Tree && mode != Mode::TreeLowRoot && mode != Mode::List) {
    fprintf(stderr, "ERROR: %s mode\n", mode == Mode::Other ? "unsupported (not tree or list)" : "incorrect");
    print_mode_help();
    return 2;
  }
  const int first = 6;  // argv[5] is the mode (src/restart_chkpt_files.cpp:38-53)
  if (chunk_size == 0 || restart_id >= num_chkpts || argc < first + (int)num_chkpts) {
    fprintf(stderr, "ERROR: need chunk_size > 0, restart_id < num_chkpts and %u checkpoint files\n", num_chkpts);
    return 2;
  }
  try {
    std::vector<dd::HostBuf> chain;
    const std::string suffix = file_suffix(mode);
    for (uint32_t i = 0; i < num_chkpts; i++) chain.push_back(read_file(std::string(argv[first + i]) + suffix));
    const std::string path = argv[first + restart_id];
    size_t len = 0;
    try {
      len = read_file(path).size();
    } catch (const std::exception&) {
      const dd::HostBuf& d = chain[restart_id];
      if (d.size() < 32) throw std::runtime_error(path + suffix + ": truncated header");
      uint64_t n;
      memcpy(&n, d.data() + 8, 8);  // header_t::datalen
      len = (size_t)n;
    }
    uint8_t* dev = nullptr;
    hip_check(hipMalloc(&dev, len ? len : 1), "hipMalloc");
    std::string logname = basename_of(path);
    std::vector<uint8_t> host(len);
    for (uint32_t t = 0; t < (num_tests ? num_tests : 1); t++) {
      hip_check(hipMemset(dev, 0, len), "hipMemset");
      std::unique_ptr<dd::TreeDeduplicator> dedup = make_dedup(mode, chunk_size);
      dedup->restart(dev, len, chain, logname, restart_id);
      hip_check(hipMemcpy(host.data(), dev, len, hipMemcpyDeviceToHost), "hipMemcpy");
      uint8_t dig[16];
      murmur3_digest(host.data(), len, dig);
      printf("%s%s\n", digest_label(mode), digest_string(dig).c_str());
    }
    hip_check(hipFree(dev), "hipFree");
  } catch (const std::exception& e) {
    fprintf(stderr, "ERROR: %s\n", e.what());
    return 1;
  }
  return 0;
}
