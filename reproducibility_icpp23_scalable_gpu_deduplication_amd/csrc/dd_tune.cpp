// dd_tune.cpp -- autotuner of the leaf kernel's chunk-to-lane mapping.
//
// The leaf kernel reads every byte of every checkpoint, so its mapping decides
// the checkpoint's bandwidth: one lane per chunk with 64 or 256 bytes of loads
// in flight, or wavefront-cooperative loads staged through LDS in 64/128/256
// byte segments (LeafMapping, dd_kernels.hpp).  Which is fastest depends on
// the chunk size, the hash and the number of jobs in the table (MI355X:
// lane-per-chunk 5.9 TB/s for one 1 GiB buffer at 128 B, direct-to-LDS 5.6
// TB/s over a batch of 64 buffers).  The first upload of a job table with a
// new (device, hash, alignment, chunk sizes, one job or several) key times
// every valid mapping on a sample of the table (every job, at most 64, over
// the first bytes of its buffer, at most 1 GiB in all) and keeps the fastest
// for the process.  The timed launches are incremental checkpoints of
// unchanged data (after one baseline launch) on scratch tree, label and list
// arrays, so the handles' state is not touched.  The sample has no changed
// chunks; the mappings within 15 % of the fastest on it (at most 3) are then
// compared on the first incremental checkpoints of tables of the key, timed
// on the leaf kernel's own dispatch (Pipeline::trial_done), and the fastest
// median of three launches each is settled for the process.  The mappings are timed round-robin
// (each round launches every mapping once) and each keeps its median, so a
// drift of the device's clock or of other work affects all alike.  One
// thread tunes a key; callers that need the same key meanwhile wait for it.
// DD_LEAF_KERNEL=direct|deep|lds64|lds128|lds256|short|dma|dma1|lds128p forces a
// mapping (when it is valid for the table).
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

// device, hash, aligned16, min/max chunk, several jobs, size class of the table
using Key = std::tuple<int, int, int, uint32_t, uint32_t, int, int>;

// A key's result: the fastest mapping on the sample and the mappings close to
// it, which the first checkpoints of the table's kind then compare on their
// own data (LeafTrial) until one is settled.
struct Entry {
  int best = LEAF_DIRECT;
  std::vector<int> cand;
  bool settled = false;
};

std::mutex g_mu;
std::condition_variable g_cv;
std::map<Key, Entry> g_choice;
std::map<Key, bool> g_tuning;  // keys one thread is tuning

std::vector<uint64_t> key_vec(const Key& k) {
  return {(uint64_t)std::get<0>(k), (uint64_t)std::get<1>(k), (uint64_t)std::get<2>(k), std::get<3>(k), std::get<4>(k),
          (uint64_t)std::get<5>(k), (uint64_t)std::get<6>(k)};
}

Key key_of(const std::vector<uint64_t>& v) {
  return Key{(int)v[0], (int)v[1], (int)v[2], (uint32_t)v[3], (uint32_t)v[4], (int)v[5], (int)v[6]};
}

// Candidates within kNear of the fastest median, at most kMaxCand, fastest
// first.
constexpr double kNear = 1.15;
constexpr size_t kMaxCand = 3;

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
  const uint32_t seg = m == LEAF_LDS64 ? 64 : (m == LEAF_LDS128 || m == LEAF_LDS128P || m == LEAF_DMA || m == LEAF_DMA1) ? 128 : 256;
  for (const JobDev& J : jobs)
    if (J.chunk_size % seg) return false;
  return true;
}

int forced_mapping() {
  const char* e = std::getenv("DD_LEAF_KERNEL");
  if (!e) return -1;
  static const char* names[LEAF_NMAP] = {"direct", "deep", "lds64", "lds128", "lds256", "short", "dma", "dma1", "lds128p"};
  for (int m = 0; m < LEAF_NMAP; m++)
    if (!strcmp(e, names[m])) return m;
  return -1;
}

// Time of one launch of a mapping on the scratch job, in milliseconds;
// negative on a HIP error.
double time_mapping(const JobDev* dj, const uint32_t* dprefix, uint32_t njobs, uint32_t grid, int m, bool al16, int kind,
                    hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  launch_begin(dj, nullptr, njobs, false, s);
  if (hipEventRecord(e0, s) != hipSuccess) return -1;
  launch_leaf(dj, dprefix, njobs, grid, m, al16, kind, s);
  if (hipEventRecord(e1, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) return -1;
  float ms = 0;
  if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1;
  return ms;
}

}  // namespace

int leaf_mapping_with_copy(const std::vector<JobDev>& jobs, bool al16, int kind, int current) {
  if (forced_mapping() >= 0) return current;
  // DMA's copy variant holds more registers than DMA1's (a wavefront per
  // SIMD fewer); both read the chunks the same way.
  if (current == LEAF_DMA) return LEAF_DMA1;
  if (leaf_writes_data(current)) return current;
  uint64_t copy = 0, total = 0;
  for (const JobDev& J : jobs) {
    total += J.len;
    if (J.write_data) copy += J.len;
  }
  if (2 * copy < total) return current;
  for (int m : {LEAF_DMA1, LEAF_LDS64})
    if (mapping_valid(m, al16, kind, jobs)) return m;
  return current;
}

int leaf_mapping_for(const std::vector<JobDev>& jobs, bool al16, int kind, hipStream_t s, std::vector<uint64_t>* key_out,
                     std::vector<int>* cand_out) {
  if (cand_out) cand_out->clear();
  const int f = forced_mapping();
  if (f >= 0 && mapping_valid(f, al16, kind, jobs)) return f;
  if (!al16) return LEAF_DIRECT;
  uint32_t cmin = ~0u, cmax = 0;
  uint64_t total = 0;
  for (const JobDev& J : jobs) {
    cmin = std::min(cmin, J.chunk_size);
    cmax = std::max(cmax, J.chunk_size);
    total += J.len;
  }
  const int size_class = total < (64ull << 20) ? 0 : total < (512ull << 20) ? 1 : total < (4ull << 30) ? 2 : 3;
  int dev = 0;
  (void)hipGetDevice(&dev);
  const Key key{dev, kind, al16 ? 1 : 0, cmin, cmax, jobs.size() > 1 ? 1 : 0, size_class};
  if (key_out) *key_out = key_vec(key);
  {
    std::unique_lock<std::mutex> lk(g_mu);
    g_cv.wait(lk, [&] { return !g_tuning.count(key); });
    auto it = g_choice.find(key);
    if (it != g_choice.end()) {
      if (cand_out && !it->second.settled) *cand_out = it->second.cand;
      return it->second.best;
    }
    g_tuning[key] = true;  // this thread tunes the key
  }
  // Publishes the result (or, when tuning could not run, removes the marker
  // so a later call tries again) and wakes the waiting callers.
  auto publish = [&](int m, bool keep, const std::vector<int>& cand) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_tuning.erase(key);
    if (keep) {
      Entry& e = g_choice[key];
      e.best = m;
      e.cand = cand;
      e.settled = cand.size() < 2;
      if (cand_out && !e.settled) *cand_out = cand;
    }
    g_cv.notify_all();
    return m;
  };
  const std::vector<int> none;
  // Scratch jobs: every job of the table (at most kSampleJobs, the largest)
  // over the first bytes of its buffer, scaled so that the sample holds at
  // most kSampleBytes, each at least one chunk.  A table of many small jobs
  // and one of a single large job favour different mappings.
  constexpr uint64_t kSampleBytes = 1ull << 30;
  constexpr size_t kSampleJobs = 64;
  std::vector<const JobDev*> src;
  for (const JobDev& J : jobs) src.push_back(&J);
  std::sort(src.begin(), src.end(), [](const JobDev* a, const JobDev* b) { return a->len > b->len; });
  if (src.size() > kSampleJobs) src.resize(kSampleJobs);
  uint64_t stotal = 0;
  for (const JobDev* J : src) stotal += J->len;
  const double scale = stotal > kSampleBytes ? (double)kSampleBytes / (double)stotal : 1.0;
  const size_t nj = src.size();
  std::vector<JobDev> sj(nj);
  std::vector<uint64_t> tree_off(nj), lab_off(nj), list_off(nj);
  uint64_t tree_bytes = 0, lab_bytes = 0, list_bytes = 0;
  for (size_t j = 0; j < nj; j++) {
    const JobDev& R = *src[j];
    const uint64_t cs = R.chunk_size;
    uint64_t tl = R.len;
    if (scale < 1.0) tl = std::max<uint64_t>(std::min<uint64_t>(cs, R.len), (uint64_t)((double)R.len * scale) / cs * cs);
    const uint64_t C = (tl + cs - 1) / cs, N = 2 * C - 1;
    JobDev& t = sj[j];
    memset(&t, 0, sizeof t);
    t.data = R.data;
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
    // List mode: a leaf that a timed launch finds changed (none, as the data
    // does not change) would be listed without marking tiles.
    t.list_mode = 1;
    t.pend_cap = (uint32_t)(((C + 63) / 64 + 31) / 32 * 64);
    tree_off[j] = tree_bytes;
    lab_off[j] = lab_bytes;
    list_off[j] = list_bytes;
    tree_bytes += 16 * N;
    lab_bytes += (N + 64 + 255) / 256 * 256;
    list_bytes += 32ull * t.pend_cap * 4;
  }
  // Counters: one block per job (JobCounters, then 2 x 32 list counts).
  const size_t ctr_stride = (sizeof(JobCounters) + 2 * 32 * kCntStride * 4 + 255) / 256 * 256;
  DevBuf tree, labels, ctr, plist, table;
  int best = LEAF_DIRECT;
  const size_t jobs_bytes = (nj * sizeof(JobDev) + 255) / 256 * 256;
  const size_t table_bytes = jobs_bytes + LEAF_NMAP * (nj + 1) * 4;
  if (tree.ensure(tree_bytes) != DD_OK || labels.ensure(lab_bytes) != DD_OK || ctr.ensure(nj * ctr_stride) != DD_OK ||
      plist.ensure(list_bytes) != DD_OK || table.ensure(table_bytes) != DD_OK)
    return publish(best, false, none);  // no memory to tune: keep the default, do not cache
  std::vector<uint8_t> host(table_bytes, 0);
  std::vector<uint32_t> grids(LEAF_NMAP);
  for (size_t j = 0; j < nj; j++) {
    JobDev& t = sj[j];
    t.tree = reinterpret_cast<Digest*>(tree.as<uint8_t>() + tree_off[j]);
    t.labels = labels.as<uint8_t>() + lab_off[j];
    uint8_t* cb = ctr.as<uint8_t>() + j * ctr_stride;
    t.ctr = reinterpret_cast<JobCounters*>(cb);
    t.pend_cnt = reinterpret_cast<uint32_t*>(cb + sizeof(JobCounters));
    t.dirty_cnt = t.pend_cnt + 32 * kCntStride;
    t.pend_cnt_next = t.pend_cnt;
    t.dirty_cnt_next = t.dirty_cnt;
    t.pend_list = reinterpret_cast<uint32_t*>(plist.as<uint8_t>() + list_off[j]);
  }
  // One workgroup prefix over the jobs per mapping.
  for (int m = 0; m < LEAF_NMAP; m++) {
    uint32_t* pre = reinterpret_cast<uint32_t*>(host.data() + jobs_bytes) + m * (nj + 1);
    uint64_t acc = 0;
    for (size_t j = 0; j < nj; j++) {
      pre[j] = (uint32_t)acc;
      acc += leaf_workgroups(m, sj[j].C);
    }
    pre[nj] = (uint32_t)acc;
    grids[m] = (uint32_t)acc;
  }
  auto upload = [&]() {
    memcpy(host.data(), sj.data(), nj * sizeof(JobDev));
    return hipMemcpyAsync(table.p, host.data(), table_bytes, hipMemcpyHostToDevice, s) == hipSuccess &&
           hipStreamSynchronize(s) == hipSuccess;
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!upload() || hipEventCreateWithFlags(&e0, hipEventDisableSystemFence) != hipSuccess ||
      hipEventCreateWithFlags(&e1, hipEventDisableSystemFence) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return publish(best, false, none);
  }
  const JobDev* dj = table.as<JobDev>();
  const uint32_t* dprefix = reinterpret_cast<const uint32_t*>(table.as<uint8_t>() + jobs_bytes);
  // One baseline launch stores every leaf digest; the timed launches are then
  // incremental checkpoints of unchanged data (read the chunks and the
  // previous digests, store nothing), the common case of a checkpoint chain.
  launch_begin(dj, nullptr, (uint32_t)nj, false, s);
  launch_leaf(dj, dprefix + LEAF_DIRECT * (nj + 1), (uint32_t)nj, grids[LEAF_DIRECT], LEAF_DIRECT, al16, kind, s);
  for (JobDev& t : sj) {
    t.is_baseline = 0;
    t.write_all_labels = 0;
  }
  if (!upload()) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return publish(best, false, none);
  }
  // One untimed launch per mapping, then kRounds rounds of one timed launch
  // each; the fastest median wins.
  constexpr int kRounds = 7;
  std::vector<double> times[LEAF_NMAP];
  bool ok = true;
  for (int r = 0; r <= kRounds && ok; r++) {
    for (int m = 0; m < LEAF_NMAP && ok; m++) {
      if (!mapping_valid(m, al16, kind, jobs)) continue;
      const double ms = time_mapping(dj, dprefix + m * (nj + 1), (uint32_t)nj, grids[m], m, al16, kind, s, e0, e1);
      if (ms < 0) ok = false;
      else if (r > 0) times[m].push_back(ms);
    }
  }
  double best_ms = -1;
  double med[LEAF_NMAP];
  for (int m = 0; m < LEAF_NMAP && ok; m++) {
    med[m] = -1;
    if (times[m].empty()) continue;
    std::sort(times[m].begin(), times[m].end());
    med[m] = times[m][times[m].size() / 2];
    if (best_ms < 0 || med[m] < best_ms) {
      best_ms = med[m];
      best = m;
    }
  }
  std::vector<int> cand;
  for (int m = 0; m < LEAF_NMAP && ok && best_ms > 0; m++)
    if (med[m] >= 0 && med[m] <= kNear * best_ms) cand.push_back(m);
  std::sort(cand.begin(), cand.end(), [&](int a, int b) { return med[a] < med[b]; });
  if (cand.size() > kMaxCand) cand.resize(kMaxCand);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipGetLastError();
  return publish(best, ok && best_ms >= 0, cand);
}

// Stage budgets settled per key (absent: still to be compared).
std::map<Key, uint64_t> g_budget;

uint64_t stage_budget_for(const std::vector<uint64_t>& key, bool* settled) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_budget.find(key_of(key));
  *settled = it != g_budget.end();
  return *settled ? it->second : kBudgetCand[0];
}

void stage_budget_settle(const std::vector<uint64_t>& key, uint64_t b) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_budget.emplace(key_of(key), b);
}

void leaf_mapping_settle(const std::vector<uint64_t>& key, int m) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_choice.find(key_of(key));
  if (it == g_choice.end() || it->second.settled) return;
  it->second.best = m;
  it->second.settled = true;
}

}  // namespace dd
