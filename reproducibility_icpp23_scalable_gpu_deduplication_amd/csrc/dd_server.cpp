// dd_server.cpp -- request loop over many checkpoint streams (include/dedup_c.h,
// dd_server_*).  Host threads (execution slots) run admitted work
// concurrently.  A slot that admits a checkpoint request also admits the
// checkpoint requests queued behind it for other idle handles (same device,
// hash and method, up to the group limit) and runs them as one batched job
// table (dd_batch: one launch per stage over all of them); a restart, or a
// checkpoint with no companion, runs alone on its handle's stream.
// Admission is in arrival order, one request in flight per handle, bounded by
// the slots and an HBM budget.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "dd_engine.hpp"

namespace {

double steady_now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
  uint64_t ticket = 0;
  uint32_t kind = 0;  // 0 checkpoint, 1 restart
  dd_dedup* h = nullptr;
  // checkpoint
  const void* data = nullptr;
  uint64_t len = 0;
  int make_baseline = 0;
  void* out = nullptr;
  uint64_t out_cap = 0;
  // restart
  std::vector<const void*> chain;
  std::vector<uint64_t> lens;
  uint32_t id = 0;
  // accounting
  uint64_t footprint = 0;
  dd_request_result res{};
  bool done = false;
  bool waited = false;
};

// Checkpoint requests that may share a job table with `a`.
bool groupable(const Request* a, const Request* b) {
  return b->kind == 0 && a->kind == 0 && a->h->device == b->h->device && a->h->st.hash_kind == b->h->st.hash_kind &&
         a->h->st.method == b->h->st.method;
}

// Device bytes a handle keeps for a buffer of `len` bytes: tree (16 B/node),
// labels and map references (9 B/node), the digest map (2N slots of 32 B), the
// per-chunk scratch arrays (root lists, work lists with their digests, tile
// regions: about 42 B/chunk, counted as 44) and the bulk map build's records
// (32 B/node), count matrices (8 B per tile of 32768 chunks and map bucket of
// 4096 slots) and bucket totals; plus 128 KiB for the arrays whose size does
// not follow the buffer's (per-workgroup claim counts, counter sets, the
// smallest work lists).
uint64_t state_bytes(uint64_t len, uint32_t chunk) {
  const uint64_t C = (len + chunk - 1) / chunk, N = 2 * C - 1;
  uint64_t cap = 1024;  // the map's capacity rule: a power of two up to 4096 slots, else whole buckets of 4096
  while (cap < 2 * N && cap < 4096) cap <<= 1;
  if (2 * N > 4096) cap = (2 * N + 4095) / 4096 * 4096;
  const uint64_t tiles = (C + 32767) / 32768, buckets = (cap + 4095) / 4096;
  return 25 * N + 32 * cap + 44 * C + 32 * N + 8 * tiles * buckets + 16 * buckets + (128u << 10);
}

}  // namespace

struct dd_server {
  uint64_t budget = 0, in_use = 0, peak = 0;
  uint32_t slots = 0;
  uint32_t max_group = 64;  // checkpoint requests per batched job table (1: each alone)
  int device = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;  // arrival order
  std::unordered_map<uint64_t, Request*> by_ticket;
  std::unordered_set<dd_dedup*> busy;
  std::unordered_map<dd_dedup*, uint64_t> charged;  // state bytes charged per handle
  uint64_t next_ticket = 1;
  bool stopping = false;
  std::vector<std::thread> workers;

  // Requests that start now together (empty: none may start).  The first
  // queued request of an idle handle leads; requests of a busy handle are
  // skipped (they wait for their predecessor).  A request that does not fit
  // the budget stops the scan, so large requests are not starved; when
  // nothing is in flight the leader always starts, even if it (with the
  // device state the handles keep charged) exceeds the budget, so the loop
  // never stalls.  A checkpoint leader takes along the queued checkpoint
  // requests of other idle handles that may share its job table, in queue
  // order, while they fit the budget.  A handle whose request is passed over
  // (a restart, another hash) has none of its later requests taken, so each
  // handle's requests run in submission order.
  std::vector<Request*> pick() {
    std::vector<Request*> g;
    std::unordered_set<dd_dedup*> passed;
    for (auto it = queue.begin(); it != queue.end();) {
      Request* r = *it;
      if (busy.count(r->h) || passed.count(r->h)) {
        ++it;
        continue;
      }
      if (!g.empty() && !groupable(g[0], r)) {
        passed.insert(r->h);
        ++it;
        continue;
      }
      const bool new_state = !charged.count(r->h);
      // A handle's first request is charged an estimate of its state; after
      // every request the charge becomes the bytes it actually holds.
      const uint64_t state = new_state ? (r->kind == 0 ? state_bytes(r->len, r->h->st.chunk_size) : 0) : 0;
      const uint64_t need = r->footprint + state;
      if (in_use + need > budget && !busy.empty()) break;
      it = queue.erase(it);
      if (new_state) charged[r->h] = state;
      in_use += need;
      if (in_use > peak) peak = in_use;
      busy.insert(r->h);
      g.push_back(r);
      if (r->kind != 0 || g.size() >= max_group) break;
    }
    return g;
  }

  // One request alone, on its handle's stream.
  static void run_one(Request* r) {
    if (r->kind == 0) {
      uint64_t n = 0;
      r->res.status = dd_checkpoint_device(r->h, r->data, r->len, r->make_baseline, r->out, r->out_cap, &n);
      r->res.out_len = n;
    } else {
      r->res.status =
          dd_restart_device(r->h, r->chain.data(), r->lens.data(), (uint32_t)r->chain.size(), r->id, r->out, r->len);
      r->res.out_len = r->res.status == DD_OK ? r->len : 0;
    }
    r->res.group = 1;
  }

  // Checkpoint requests of distinct idle handles as one batched job table on
  // the slot's batch (created on first use, its handle list replaced per
  // group).  Requests that the batch would reject (invalid arguments or
  // state) are answered alone first, so one bad request does not fail the
  // others.
  static void run_group(std::vector<Request*>& g, dd_batch*& b) {
    std::vector<Request*> ok;
    for (Request* r : g) {
      const dd::StreamState& st = r->h->st;
      const dd_status v = st.validate(r->len, st.runs_baseline(r->make_baseline), static_cast<const uint8_t*>(r->out),
                                      r->out_cap);
      if (v == DD_OK) {
        ok.push_back(r);
      } else {
        r->res.status = v;
        r->res.out_len = 0;
        r->res.group = 1;
      }
    }
    if (ok.size() == 1) run_one(ok[0]);
    if (ok.size() < 2) return;
    const size_t n = ok.size();
    std::vector<dd_dedup*> hs(n);
    std::vector<const void*> data(n);
    std::vector<uint64_t> lens(n), caps(n), outl(n, 0);
    std::vector<int> base(n);
    std::vector<void*> outs(n);
    for (size_t i = 0; i < n; i++) {
      hs[i] = ok[i]->h;
      data[i] = ok[i]->data;
      lens[i] = ok[i]->len;
      base[i] = ok[i]->make_baseline;
      outs[i] = ok[i]->out;
      caps[i] = ok[i]->out_cap;
    }
    dd_status st = DD_OK;
    if (b && b->device != hs[0]->device) {
      (void)dd_batch_destroy(b);
      b = nullptr;
    }
    if (!b) st = dd_batch_create(hs.data(), (uint32_t)n, &b);
    if (st == DD_OK) {
      b->hs = hs;
      st = dd_batch_checkpoint_device(b, data.data(), lens.data(), base.data(), outs.data(), caps.data(), outl.data());
    }
    for (size_t i = 0; i < n; i++) {
      ok[i]->res.status = st;
      ok[i]->res.out_len = st == DD_OK ? outl[i] : 0;
      ok[i]->res.group = (uint32_t)n;
    }
  }

  void worker() {
    (void)hipSetDevice(device);
    dd_batch* batch = nullptr;  // this slot's batched job table
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      std::vector<Request*> g;
      cv_work.wait(lk, [&] { return !(g = pick()).empty() || (stopping && queue.empty()); });
      if (g.empty()) break;  // stopping and nothing left to start
      lk.unlock();
      const double t0 = steady_now();
      for (Request* r : g) r->res.t_start = t0;
      if (g.size() == 1) run_one(g[0]);
      else run_group(g, batch);
      const double t1 = steady_now();
      std::vector<uint64_t> resident(g.size());
      for (size_t i = 0; i < g.size(); i++) {
        g[i]->res.t_end = t1;
        resident[i] = dd::resident_bytes(g[i]->h);  // the handle is not used elsewhere meanwhile
      }
      lk.lock();
      for (size_t i = 0; i < g.size(); i++) {
        Request* r = g[i];
        busy.erase(r->h);
        in_use -= r->footprint;
        // The handle's buffers stay allocated after the request (its state,
        // and restart buffers that grow with the chain): charge what it holds
        // now.
        uint64_t& ch = charged[r->h];
        in_use = in_use - ch + resident[i];
        ch = resident[i];
        if (in_use > peak) peak = in_use;
        r->done = true;
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
    lk.unlock();
    if (batch) (void)dd_batch_destroy(batch);
  }

  dd_status submit(Request* r, uint64_t* ticket) {
    std::lock_guard<std::mutex> lk(mu);
    if (stopping) return DD_ESTATE;
    r->ticket = next_ticket++;
    r->res.t_submit = steady_now();
    r->res.kind = r->kind;
    by_ticket[r->ticket] = r;
    queue.push_back(r);
    *ticket = r->ticket;
    cv_work.notify_all();
    return DD_OK;
  }
};

extern "C" {

double dd_now(void) { return steady_now(); }

dd_status dd_server_create(uint64_t hbm_budget_bytes, uint32_t slots, dd_server** out) {
  if (!out || slots == 0 || hbm_budget_bytes == 0) return DD_EINVAL;
  *out = nullptr;
  dd_server* s = new (std::nothrow) dd_server;
  if (!s) return DD_ENOMEM;
  s->budget = hbm_budget_bytes;
  s->slots = slots;
  if (hipGetDevice(&s->device) != hipSuccess) {
    delete s;
    return DD_EHIP;
  }
  for (uint32_t i = 0; i < slots; i++) s->workers.emplace_back([s] { s->worker(); });
  *out = s;
  return DD_OK;
}

dd_status dd_server_destroy(dd_server* s) {
  if (!s) return DD_EINVAL;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
    s->cv_work.notify_all();
  }
  for (auto& t : s->workers) t.join();
  for (auto& kv : s->by_ticket) delete kv.second;
  delete s;
  return DD_OK;
}

dd_status dd_server_submit_checkpoint(dd_server* s, dd_dedup* h, const void* dev_data, uint64_t len, int make_baseline,
                                      void* dev_out, uint64_t out_cap, uint64_t* ticket) {
  if (!s || !h || !dev_data || !dev_out || !ticket) return DD_EINVAL;
  Request* r = new (std::nothrow) Request;
  if (!r) return DD_ENOMEM;
  r->kind = 0;
  r->h = h;
  r->data = dev_data;
  r->len = len;
  r->make_baseline = make_baseline;
  r->out = dev_out;
  r->out_cap = out_cap;
  r->footprint = len + out_cap;
  const dd_status st = s->submit(r, ticket);
  if (st != DD_OK) delete r;
  return st;
}

dd_status dd_server_submit_restart(dd_server* s, dd_dedup* h, const void* const* dev_chain, const uint64_t* lens,
                                   uint32_t nchain, uint32_t chkpt_id, void* dev_out, uint64_t len, uint64_t* ticket) {
  if (!s || !h || !dev_chain || !lens || !dev_out || !ticket) return DD_EINVAL;
  Request* r = new (std::nothrow) Request;
  if (!r) return DD_ENOMEM;
  r->kind = 1;
  r->h = h;
  r->chain.assign(dev_chain, dev_chain + nchain);
  r->lens.assign(lens, lens + nchain);
  r->id = chkpt_id;
  r->out = dev_out;
  r->len = len;
  // output buffer plus the restart scratch (chunk sources and cover table,
  // 24 B per chunk at the smallest chunk size the chain may use)
  r->footprint = len + 24 * (len / 16 + 1);
  const dd_status st = s->submit(r, ticket);
  if (st != DD_OK) delete r;
  return st;
}

dd_status dd_server_wait(dd_server* s, uint64_t ticket, dd_request_result* out) {
  if (!s || !out) return DD_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->by_ticket.find(ticket);
  if (it == s->by_ticket.end()) return DD_EINVAL;
  Request* r = it->second;
  if (r->waited) return DD_EINVAL;  // one waiter per ticket: another thread waits for it
  r->waited = true;
  s->cv_done.wait(lk, [&] { return r->done; });
  *out = r->res;
  s->by_ticket.erase(it);
  delete r;
  return DD_OK;
}

dd_status dd_server_set_batching(dd_server* s, uint32_t max_group) {
  if (!s || max_group == 0) return DD_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  s->max_group = max_group;
  return DD_OK;
}

dd_status dd_server_usage(dd_server* s, uint64_t* in_use, uint64_t* peak) {
  if (!s) return DD_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  if (in_use) *in_use = s->in_use;
  if (peak) *peak = s->peak;
  return DD_OK;
}

}  // extern "C"
