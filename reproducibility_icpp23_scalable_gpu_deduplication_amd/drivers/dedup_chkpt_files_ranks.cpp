// dedup_chkpt_files_ranks -- deduplicates the checkpoint files of many ranks
// (the reference's src/dedup_files_mpi.cpp), one process per GPU.
//
//   dedup_chkpt_files_ranks chunk_size num_chkpts <mode> [--rank R --ranks N] file...
//
// A file belongs to the rank named by the first "Rank<digits>" in its base
// name (the reference matches the substring "Rank<r>", so Rank1 also takes
// Rank10's files; here the number is read whole).  Every rank lists
// num_chkpts files, in checkpoint order.  This process takes the ranks r with
// r % N == R (N, R from --ranks/--rank, else from the RANK / WORLD_SIZE,
// OMPI_COMM_WORLD_RANK / OMPI_COMM_WORLD_SIZE or PMI_RANK / PMI_SIZE
// environment variables, else all ranks), keeps one deduplicator per rank and
// checkpoints checkpoint k of all of them in one batch call (one launch per
// stage over every rank's buffer).  Outputs are the single-rank driver's:
// "<file>.hashtree.incr_chkpt" (list mode "<file>.hashlist.incr_chkpt") and
// the CSV logs under each file's base name.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <exception>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "driver_util.hpp"

using namespace dd_drivers;

namespace {

// Rank named by the first "Rank<digits>" of the base name, or -1.
long rank_of(const std::string& path) {
  const std::string b = basename_of(path);
  size_t p = 0;
  while ((p = b.find("Rank", p)) != std::string::npos) {
    size_t q = p + 4;
    long r = 0;
    bool digits = false;
    while (q < b.size() && b[q] >= '0' && b[q] <= '9') {
      r = r * 10 + (b[q] - '0');
      q++;
      digits = true;
    }
    if (digits) return r;
    p += 4;
  }
  return -1;
}

long env_long(const char* a, const char* b, const char* c) {
  for (const char* n : {a, b, c}) {
    const char* v = getenv(n);
    if (v && *v) return strtol(v, nullptr, 10);
  }
  return -1;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 5) {
    fprintf(stderr, "usage: %s chunk_size num_chkpts <mode> [--rank R --ranks N] file...\n", argv[0]);
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
  long me = env_long("RANK", "OMPI_COMM_WORLD_RANK", "PMI_RANK");
  long procs = env_long("WORLD_SIZE", "OMPI_COMM_WORLD_SIZE", "PMI_SIZE");
  std::vector<std::string> files;
  for (int i = 4; i < argc; i++) {
    if (!strcmp(argv[i], "--rank") && i + 1 < argc) {
      me = strtol(argv[++i], nullptr, 10);
    } else if (!strcmp(argv[i], "--ranks") && i + 1 < argc) {
      procs = strtol(argv[++i], nullptr, 10);
    } else {
      files.push_back(argv[i]);
    }
  }
  if (procs <= 0 || me < 0) {
    procs = 1;
    me = 0;
  }
  if (chunk_size == 0 || num_chkpts == 0 || me >= procs) {
    fprintf(stderr, "ERROR: need chunk_size > 0, num_chkpts > 0 and 0 <= rank < ranks\n");
    return 2;
  }
  // Files per rank, in command-line order.
  std::map<long, std::vector<std::string>> by_rank;
  for (const std::string& f : files) {
    const long r = rank_of(f);
    if (r < 0) {
      fprintf(stderr, "ERROR: %s names no rank (Rank<digits>)\n", f.c_str());
      return 2;
    }
    if (r % procs == me) by_rank[r].push_back(f);
  }
  if (by_rank.empty()) {
    fprintf(stderr, "ERROR: no files for rank %ld of %ld\n", me, procs);
    return 2;
  }
  for (const auto& kv : by_rank) {
    if (kv.second.size() != num_chkpts) {
      fprintf(stderr, "ERROR: rank %ld has %zu files, expected %u\n", kv.first, kv.second.size(), num_chkpts);
      return 2;
    }
  }
  printf("process %ld of %ld: %zu ranks\n", me, procs, by_rank.size());
  try {
    std::vector<std::unique_ptr<dd::TreeDeduplicator>> dedups;
    std::vector<dd::TreeDeduplicator*> ptrs;
    std::vector<const std::vector<std::string>*> lists;
    for (const auto& kv : by_rank) {
      dedups.push_back(make_dedup(mode, chunk_size));
      ptrs.push_back(dedups.back().get());
      lists.push_back(&kv.second);
    }
    const size_t n = ptrs.size();
    std::vector<uint8_t*> dev(n, nullptr);
    std::vector<size_t> cap(n, 0), lens(n, 0);
    for (uint32_t k = 0; k < num_chkpts; k++) {
      std::vector<std::string> filenames(n), lognames(n);
      for (size_t i = 0; i < n; i++) {
        const std::string& path = (*lists[i])[k];
        const std::vector<uint8_t> data = read_file(path);
        if (data.empty()) throw std::runtime_error(path + ": empty checkpoint file");
        if (data.size() > cap[i]) {
          if (dev[i]) hip_check(hipFree(dev[i]), "hipFree");
          hip_check(hipMalloc(&dev[i], data.size()), "hipMalloc");
          cap[i] = data.size();
        }
        hip_check(hipMemcpy(dev[i], data.data(), data.size(), hipMemcpyHostToDevice), "hipMemcpy");
        lens[i] = data.size();
        filenames[i] = path + file_suffix(mode);
        lognames[i] = basename_of(path);
      }
      dd::TreeDeduplicator::checkpoint_batch(ptrs, dev, lens, filenames, lognames, k == 0);
      for (size_t i = 0; i < n; i++)
        printf("%s: %zu bytes -> %s (first-occurrence regions %zu, shifted regions %zu)\n", (*lists[i])[k].c_str(),
               lens[i], filenames[i].c_str(), ptrs[i]->num_first_ocur(), ptrs[i]->num_shift_dupl());
    }
    for (uint8_t* p : dev)
      if (p) hip_check(hipFree(p), "hipFree");
  } catch (const std::exception& e) {
    fprintf(stderr, "ERROR: %s\n", e.what());
    return 1;
  }
  return 0;
}
