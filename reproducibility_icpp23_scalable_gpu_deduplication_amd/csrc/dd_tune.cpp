// dd_tune.cpp -- autotuner of the leaf kernel's chunk-to-lane mapping.
//
// The leaf kernel reads every byte of every checkpoint, so its mapping decides
// the checkpoint's bandwidth: one lane per chunk with 64 or 256 bytes of loads
// in flight, or wavefront-cooperative loads staged through LDS in 64/128/256
// byte segments (LeafMapping, dd_kernels.hpp).  Which is fastest depends on
// the chunk size and the hash (MI355X, 1 GiB: lane-per-chunk 4.9 TB/s at
// 128 B, LDS-staged 2.9 TB/s at 4096 B).  The first upload of a job table with
// a new (device, hash, alignment, chunk sizes) key times every valid mapping
// on the largest job's first 256 MiB and keeps the fastest for the process.
// The timed launches run as baselines on scratch tree, label and list arrays,
// so the handles' state is not touched.  The mappings are timed round-robin
// (each round launches every mapping once) and each keeps its median, so a
// drift of the device's clock or of other work affects all alike.  One
// thread tunes a key; callers that need the same key meanwhile wait for it.  DD_LEAF_KERNEL=direct|deep|lds64|
// lds128|lds256|short|dma forces a mapping (when it is valid for the table).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "dd_engine.hpp"

namespace dd {
namespace {

using Key = std::tuple<int, int, int, uint32_t, uint32_t>;  // device, hash, aligned16, min chunk, max chunk

std::mutex g_mu;
std::condition_variable g_cv;
constexpr int kTuning = -1;  // g_choice value while one thread tunes the key
std::map<Key, int> g_choice;

bool mapping_valid(int m, bool al16, int kind, const std::vector<JobDev>& jobs) {
  if (m == LEAF_DIRECT) return true;
  if (!al16) return false;
  if (m == LEAF_DEEP) return kind == HASH_MURMUR3;
  if (m == LEAF_SHORT) {
    if (kind != HASH_MURMUR3) return false;
    for (const JobDev& J : jobs)
      if (J.chunk_size > 128) return false;
    return true;
  }
  const uint32_t seg = m == LEAF_LDS64 ? 64 : (m == LEAF_LDS128 || m == LEAF_DMA || m == LEAF_DMA1) ? 128 : 256;
  for (const JobDev& J : jobs)
    if (J.chunk_size % seg) return false;
  return true;
}

int forced_mapping() {
  const char* e = std::getenv("DD_LEAF_KERNEL");
  if (!e) return -1;
  static const char* names[LEAF_NMAP] = {"direct", "deep", "lds64", "lds128", "lds256", "short", "dma", "dma1"};
  for (int m = 0; m < LEAF_NMAP; m++)
    if (!strcmp(e, names[m])) return m;
  return -1;
}

// Time of one launch of a mapping on the scratch job, in milliseconds;
// negative on a HIP error.
double time_mapping(const JobDev* dj, const uint32_t* dprefix, uint32_t grid, int m, bool al16, int kind, hipStream_t s,
                    hipEvent_t e0, hipEvent_t e1) {
  launch_begin(dj, nullptr, 1, false, s);
  if (hipEventRecord(e0, s) != hipSuccess) return -1;
  launch_leaf(dj, dprefix, 1, grid, m, al16, kind, s);
  if (hipEventRecord(e1, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) return -1;
  float ms = 0;
  if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1;
  return ms;
}

}  // namespace

int leaf_mapping_for(const std::vector<JobDev>& jobs, bool al16, int kind, hipStream_t s) {
  const int f = forced_mapping();
  if (f >= 0 && mapping_valid(f, al16, kind, jobs)) return f;
  if (!al16) return LEAF_DIRECT;
  uint32_t cmin = ~0u, cmax = 0;
  const JobDev* R = &jobs[0];
  for (const JobDev& J : jobs) {
    cmin = std::min(cmin, J.chunk_size);
    cmax = std::max(cmax, J.chunk_size);
    if (J.C > R->C) R = &J;
  }
  int dev = 0;
  (void)hipGetDevice(&dev);
  const Key key{dev, kind, al16 ? 1 : 0, cmin, cmax};
  {
    std::unique_lock<std::mutex> lk(g_mu);
    g_cv.wait(lk, [&] {
      auto it = g_choice.find(key);
      return it == g_choice.end() || it->second != kTuning;
    });
    auto it = g_choice.find(key);
    if (it != g_choice.end()) return it->second;
    g_choice[key] = kTuning;  // this thread tunes the key
  }
  // Publishes the result (or, when tuning could not run, removes the marker
  // so a later call tries again) and wakes the waiting callers.
  auto publish = [&](int m, bool keep) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (keep) g_choice[key] = m;
    else g_choice.erase(key);
    g_cv.notify_all();
    return m;
  };
  // Scratch baseline job over the first (at most) 256 MiB of the largest job.
  const uint64_t cs = R->chunk_size;
  uint64_t tl = std::min<uint64_t>(R->len, 256ull << 20);
  if (tl >= cs) tl = tl / cs * cs;
  const uint64_t C = (tl + cs - 1) / cs, N = 2 * C - 1;
  const uint64_t pend_cap = ((C + 63) / 64 + 31) / 32 * 64;  // unused: a baseline keeps no pending list
  DevBuf tree, labels, ctr, plist, table;
  int best = LEAF_DIRECT;
  if (tree.ensure(16 * N) != DD_OK || labels.ensure(N + 64) != DD_OK ||
      ctr.ensure(sizeof(JobCounters) + 2 * 32 * kCntStride * 4) != DD_OK || plist.ensure(32 * pend_cap * 4) != DD_OK ||
      table.ensure(sizeof(JobDev) + LEAF_NMAP * 2 * 4) != DD_OK)
    return publish(best, false);  // no memory to tune: keep the default, do not cache
  JobDev t;
  memset(&t, 0, sizeof t);
  t.data = R->data;
  t.len = tl;
  t.chunk_size = (uint32_t)cs;
  t.hash_kind = kind;
  t.C = C;
  t.N = N;
  t.nfull = tl / cs;
  t.D = make_geom(C).D;
  t.is_baseline = 1;
  t.write_all_labels = 1;
  t.epoch = 1;
  t.tree = tree.as<Digest>();
  t.labels = labels.as<uint8_t>();
  t.ctr = ctr.as<JobCounters>();
  t.pend_cnt = reinterpret_cast<uint32_t*>(ctr.as<uint8_t>() + sizeof(JobCounters));
  t.dirty_cnt = t.pend_cnt + 32 * kCntStride;
  t.pend_cnt_next = t.pend_cnt;  // a baseline keeps no lists
  t.dirty_cnt_next = t.dirty_cnt;
  t.pend_list = plist.as<uint32_t>();
  t.pend_cap = (uint32_t)pend_cap;
  // One workgroup prefix {0, grid} per mapping: the grid-stride mappings
  // read their stride from it.
  uint8_t host[sizeof(JobDev) + LEAF_NMAP * 8];
  memcpy(host, &t, sizeof t);
  uint32_t grids[LEAF_NMAP];
  for (int m = 0; m < LEAF_NMAP; m++) {
    grids[m] = (uint32_t)leaf_workgroups(m, C);
    const uint32_t pre[2] = {0, grids[m]};
    memcpy(host + sizeof(JobDev) + 8 * m, pre, sizeof pre);
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipMemcpyAsync(table.p, host, sizeof host, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess || hipEventCreateWithFlags(&e0, hipEventDisableSystemFence) != hipSuccess ||
      hipEventCreateWithFlags(&e1, hipEventDisableSystemFence) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return publish(best, false);
  }
  const JobDev* dj = table.as<JobDev>();
  const uint32_t* dprefix = reinterpret_cast<const uint32_t*>(table.as<uint8_t>() + sizeof(JobDev));
  // One untimed launch per mapping, then kRounds rounds of one timed launch
  // each; the fastest median wins.
  constexpr int kRounds = 7;
  std::vector<double> times[LEAF_NMAP];
  bool ok = true;
  for (int r = 0; r <= kRounds && ok; r++) {
    for (int m = 0; m < LEAF_NMAP && ok; m++) {
      if (!mapping_valid(m, al16, kind, jobs)) continue;
      const double ms = time_mapping(dj, dprefix + 2 * m, grids[m], m, al16, kind, s, e0, e1);
      if (ms < 0) ok = false;
      else if (r > 0) times[m].push_back(ms);
    }
  }
  double best_ms = -1;
  for (int m = 0; m < LEAF_NMAP && ok; m++) {
    if (times[m].empty()) continue;
    std::sort(times[m].begin(), times[m].end());
    const double med = times[m][times[m].size() / 2];
    if (best_ms < 0 || med < best_ms) {
      best_ms = med;
      best = m;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipGetLastError();
  return publish(best, ok && best_ms >= 0);
}

}  // namespace dd
