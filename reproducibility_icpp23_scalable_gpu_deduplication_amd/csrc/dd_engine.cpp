// dd_engine.cpp -- host orchestration and the C ABI (include/dedup_c.h).
//
// Reference orchestration: TreeDeduplicator::checkpoint
// (src/tree_approach.cpp:1577-1689) and restart_chkpt (:712-1128).
#include "dd_engine.hpp"

#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <mutex>
#include <new>

#define DD_HIP(x)                         \
  do {                                    \
    if ((x) != hipSuccess) return DD_EHIP; \
  } while (0)
#define DD_TRY(x)                    \
  do {                               \
    dd_status _rc = (x);             \
    if (_rc != DD_OK) return _rc;    \
  } while (0)

namespace dd {

// ---------------------------------------------------------------------------
// DevBuf
// ---------------------------------------------------------------------------
DevBuf::~DevBuf() { release(); }

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

dd_status DevBuf::ensure(size_t n, bool zero, hipStream_t zs) {
  if (n == 0) n = 16;
  if (n <= bytes) return DD_OK;
  release();
  if (hipMalloc(&p, n) != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    return DD_ENOMEM;
  }
  bytes = n;
  if (zero && hipMemsetAsync(p, 0, n, zs) != hipSuccess) return DD_EHIP;
  return DD_OK;
}

static uint64_t next_pow2(uint64_t x) {
  uint64_t c = 1;
  while (c < x) c <<= 1;
  return c;
}

static uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

uint64_t resident_bytes(const dd_dedup* h) {
  const StreamState& st = h->st;
  const DevBuf* bufs[] = {&st.tree, &st.labels, &st.slot_of, &st.lr_new, &st.map, &st.blk, &st.f_list, &st.f_off, &st.s_node,
                          &st.s_prev, &st.s_tree, &st.ss_node, &st.ss_prev, &st.s_hist, &st.ptab, &st.leafblk,
                          &st.pend_list, &st.pend_dig, &st.tiles, &st.tile_epoch, &st.lists, &st.diff, &st.ctr, &st.bulk,
                          &h->cover, &h->ref, &h->pend, &h->rs_hdr, &h->rs_arena};
  uint64_t n = h->pipe.device_bytes();
  for (const DevBuf* b : bufs) n += b->bytes;
  return n;
}

uint64_t max_diff_size(uint64_t len, uint32_t chunk_size, uint32_t ckpt_id) {
  if (chunk_size == 0) return 0;
  const uint64_t C = ceil_div(len, chunk_size);
  // first-occurrence and shifted roots are disjoint subtrees: 4F + 8S <= 8C
  return 32 + 8 * C + 8 * ((uint64_t)ckpt_id + 1) + C * (uint64_t)chunk_size;
}

static MapView make_view(void* slots, uint64_t cap) {
  MapView m;
  m.slots = reinterpret_cast<Slot*>(slots);
  const uint64_t bs = std::min<uint64_t>(cap, kBucketSlots);
  m.bmask = bs - 1;
  m.nbuckets = cap / bs;
  return m;
}

// Map capacity for `slots` slots: a power of two of at least 1024 up to one
// bucket, else a whole number of buckets.
static uint64_t map_capacity(uint64_t slots) {
  if (slots <= kBucketSlots) return next_pow2(std::max<uint64_t>(slots, 1024));
  return ceil_div(slots, kBucketSlots) * kBucketSlots;
}

static uint64_t map_limit_of(uint64_t cap) { return cap / 2 + cap / 4; }

// Checkpoints of growth a map growth makes room for.
constexpr uint64_t kGrowHorizon = 16;

// Inserts per job table above which every job probes in narrow rounds.
constexpr uint64_t kWideInsertLimit = 256 * 1024;

// Tree tiles: group g of a tree of depth D covers levels D-8g down to
// max(D-8(g+1), 0); it has one tile per node of its top level.
static int groups_of(uint32_t D) { return D == 0 ? 1 : (int)((D + kTileLevels - 1) / kTileLevels); }

static uint64_t tiles_in_group(uint32_t D, int g) {
  if (g >= groups_of(D)) return 0;
  const int Lb = (int)D - kTileLevels * g;
  const int Lr = Lb > kTileLevels ? Lb - kTileLevels : 0;
  return 1ull << Lr;
}

// ---------------------------------------------------------------------------
// StreamState
// ---------------------------------------------------------------------------
dd_status StreamState::grow_map(uint64_t need, hipStream_t s) {
  // Room for `need` at half load, and for the entries that kGrowHorizon more
  // checkpoints would add at the rate the chain has added them since its
  // baseline, under the load limit: a chain that adds many entries per
  // checkpoint grows in fewer, larger steps (each growth rewrites the whole
  // table).
  uint64_t cap = map_capacity(2 * need);
  if (ckpts_since_base > 0 && map_count > base_map_count) {
    const uint64_t rate = (map_count - base_map_count) / ckpts_since_base;
    cap = std::max(cap, map_capacity((need + kGrowHorizon * rate) / 3 * 4));
  }
  if (cap <= map_cap) cap = map_capacity(map_cap * 2);
  DevBuf fresh;
  DD_TRY(fresh.ensure(cap * sizeof(Slot)));
  launch_map_clear(fresh.as<Slot>(), cap, s);
  DevBuf flag;
  DD_TRY(flag.ensure(4, true, s));
  launch_map_rehash(map.as<Slot>(), map_cap, make_view(fresh.p, cap), flag.as<uint32_t>(), s);
  uint32_t f = 0;
  DD_HIP(hipMemcpyAsync(&f, flag.p, 4, hipMemcpyDeviceToHost, s));
  DD_HIP(hipStreamSynchronize(s));
  if (f) return DD_ENOMEM;
  std::swap(map.p, fresh.p);
  std::swap(map.bytes, fresh.bytes);
  map_cap = cap;
  return DD_OK;
}

// Leaf records per bucket above which a bulk job falls back to the dense
// inserts (a bucket holds 1024 leaves on average at most), and records per
// bucket the final build takes (leaves and interior nodes, 2048 on average
// at most).  DD_BULK_CAPS="leaf,build" overrides both (at most kBulkRec).
static void bulk_caps(uint32_t* cap_l, uint32_t* cap_b) {
  *cap_l = 1600;
  *cap_b = kBulkRec;
  if (const char* e = std::getenv("DD_BULK_CAPS")) {
    unsigned a = 0, b = 0;
    if (sscanf(e, "%u,%u", &a, &b) == 2) {
      *cap_l = std::min<uint32_t>(a, kBulkRec);
      *cap_b = std::min<uint32_t>(b, kBulkRec);
    }
  }
}

bool StreamState::uses_bulk(uint64_t Cn, uint64_t cap) const {
  if (method != DD_METHOD_TREE || (options & DD_OPT_DENSE_BASELINE) || bulk_heavy) return false;
  return (cap >> kBucketLog) <= kBulkMaxBuckets && ceil_div(Cn, kPartTile) <= kBulkMaxTiles;
}

// Sizes the bulk build's arrays.  None needs clearing: the partition
// kernels write every count they later read.
dd_status StreamState::prepare_bulk(uint64_t Cn, uint64_t cap, hipStream_t) {
  const uint64_t nb = std::max<uint64_t>(cap >> kBucketLog, 1);
  const uint64_t nt = ceil_div(Cn, kPartTile);
  const uint64_t ng = ceil_div(nb, kColBuckets);
  const uint64_t nrec = 2 * Cn - 1;
  auto a256 = [](uint64_t x) { return (x + 255) & ~255ull; };
  bk_off[0] = 0;                                         // count matrices: leaves, interior nodes
  bk_off[1] = bk_off[0] + a256(2 * nt * nb * 4);         // bucket totals
  bk_off[2] = bk_off[1] + a256(2 * nb * 4);              // group totals
  bk_off[3] = bk_off[2] + a256(2 * ng * 4);              // entries per bucket
  bk_off[4] = bk_off[3] + a256(nb * 4);                  // records
  const uint64_t total = bk_off[4] + nrec * 32;
  DD_TRY(bulk.ensure(total));
  bk_nb = nb;
  bk_nt = nt;
  return DD_OK;
}

void StreamState::fill_bulk(JobDev& job) const {
  uint8_t* base = bulk.as<uint8_t>();
  job.bulk = 1;
  job.nbuckets = (uint32_t)bk_nb;
  job.ntiles_b = (uint32_t)bk_nt;
  bulk_caps(&job.bk_cap_l, &job.bk_cap_b);
  job.bk_mat = reinterpret_cast<uint32_t*>(base + bk_off[0]);
  job.bk_tot = reinterpret_cast<uint32_t*>(base + bk_off[1]);
  job.bk_grp = reinterpret_cast<uint32_t*>(base + bk_off[2]);
  job.bk_claim = reinterpret_cast<uint32_t*>(base + bk_off[3]);
  job.bk_rec = reinterpret_cast<u32x4_t*>(base + bk_off[4]);
  job.bulk_dedup = bulk_dups ? 1 : 0;
}

dd_status StreamState::validate(uint64_t len, bool baseline, const uint8_t* out, uint64_t out_cap) const {
  const uint64_t Cn = ceil_div(len, chunk_size);
  if (len == 0 || Cn > 0x7fffffffull) return DD_EINVAL;
  if (!baseline) {
    if (len != data_len) return DD_ESTATE;
    if ((int64_t)current_id <= last_map_ckpt) return DD_ESTATE;
  }
  if (out && out_cap < max_diff_size(len, chunk_size, current_id)) return DD_EINVAL;
  return DD_OK;
}

dd_status StreamState::prepare(const uint8_t* data, uint64_t len, bool baseline, uint8_t* out, uint64_t out_cap,
                               JobDev& job, hipStream_t s) {
  DD_TRY(validate(len, baseline, out, out_cap));
  const uint64_t cs = chunk_size;
  const uint64_t Cn = ceil_div(len, cs);
  const uint64_t Nn = 2 * Cn - 1;
  const uint32_t K = current_id;
  need_baseline = true;  // until commit()
  if (baseline) {
    if (method == DD_METHOD_TREE_LOW_ROOT) DD_TRY(lr_new.ensure(Nn));
    DD_TRY(tree.ensure(Nn * sizeof(Digest)));
    DD_TRY(labels.ensure(ceil_div(Nn, kNodesPerBlock) * kNodesPerBlock + 64));  // root scans read 16 at a time
    DD_TRY(slot_of.ensure(Nn * 8));
    const uint64_t cap = map_capacity(2 * Nn);
    if (map.bytes > 4 * cap * sizeof(Slot)) map.release();
    DD_TRY(map.ensure(cap * sizeof(Slot)));
    map_cap = cap;
    // The bulk build writes every slot (or clears its buckets when it falls
    // back to the dense inserts); the dense inserts need a cleared map.
    bulk_now = uses_bulk(Cn, cap);
    if (bulk_now) DD_TRY(prepare_bulk(Cn, cap, s));
    else launch_map_clear(map.as<Slot>(), cap, s);
    map_count = 0;
    have_tree = true;
    data_len = len;
    C = Cn;
    N = Nn;
    last_map_ckpt = -1;
  }
  const uint64_t nblk = ceil_div(N, kNodesPerBlock);
  // Per root-scan workgroup (at most one per node block): three counts; then
  // the root masks of every group of 16 nodes.
  // Block counts, root masks, then (8-byte aligned) the label masks.
  const uint64_t lmask_off = ((3 + kNodesPerBlock / 16) * nblk * 4 + 7) & ~7ull;
  DD_TRY(blk.ensure(lmask_off + nblk * (kNodesPerBlock / 16 / 64) * 8));
  DD_TRY(f_list.ensure((C + 1) * 4));
  DD_TRY(f_off.ensure((C + 1) * 4));
  DD_TRY(s_node.ensure(C * 4));
  DD_TRY(s_prev.ensure(C * 4));
  DD_TRY(s_tree.ensure(C * 4));
  DD_TRY(ss_node.ensure(C * 4));
  DD_TRY(ss_prev.ensure(C * 4));
  // Buffers that grow with the checkpoint id are sized for ids in steps of 64:
  // an allocation per checkpoint (a free and a malloc of megabytes) would
  // cost more host time than the whole incremental checkpoint's kernels.
  const uint64_t Kcap = ((uint64_t)K | 63) + 1;  // ids 0 .. Kcap-1
  DD_TRY(s_hist.ensure(Kcap * ceil_div(C, kSTile) * 4));
  DD_TRY(ptab.ensure(Kcap * 8));
  // Two sets of 32 pending-list counters, one per checkpoint parity.
  DD_TRY(ctr.ensure(sizeof(JobCounters) + 2 * 32 * kCntStride * 4, true, s));
  const uint64_t nlb = ceil_div(C, 256);
  // blk_claim (one per insert_leaves workgroup) then int_claim (one per
  // insert_interior workgroup); a baseline launches one of each per leaf block.
  const uint64_t nclaim = std::max<uint64_t>(nlb, kMaxStageWG);
  DD_TRY(leafblk.ensure(2 * nclaim * 4));
  // Pending leaves: sub-list r takes wavefronts w = r (mod 32), 64 leaves each.
  const uint64_t pend_cap = ceil_div(ceil_div(C, 64), 32) * 64;
  DD_TRY(pend_list.ensure(32 * pend_cap * 4));
  // Digests beside the first eighth of each sub-list (every entry up to
  // 12.5 % changed chunks): 2 bytes per chunk instead of 16.
  const uint64_t dig_cap = std::min<uint64_t>(pend_cap, ceil_div(ceil_div(pend_cap, 8), 64) * 64);
  DD_TRY(pend_dig.ensure(32 * dig_cap * sizeof(Digest)));
  const TreeGeom gm = make_geom(C);
  uint64_t ntiles = 0;
  uint32_t tile_base[4] = {0, 0, 0, 0};
  for (int g = 0; g < 4; g++) {
    tile_base[g] = (uint32_t)ntiles;
    ntiles += tiles_in_group(gm.D, g);
  }
  // First-occurrence interior nodes: a region of kTileInterior entries per
  // tile, then one count per tile.
  DD_TRY(tiles.ensure(ntiles * (kTileInterior + 1) * 4));
  // Per group-0 tile stamp of the last checkpoint that changed one of its
  // leaves; zero-filled when (re)allocated, epochs start at 1.
  const uint64_t ntiles0 = tiles_in_group(gm.D, 0);
  if (tile_epoch.bytes < ntiles0 * 4 || epoch == 0xffffffffu) {
    tile_epoch.release();
    DD_TRY(tile_epoch.ensure(ntiles0 * 4, true, s));
    epoch = 0;
  }
  epoch += 1;
  const uint64_t dirty_cap = ceil_div(ntiles0, 32);
  DD_TRY(lists.ensure((2 * 32 * kCntStride + 32 * dirty_cap) * 4, true, s));
  if (!out) {
    DD_TRY(diff.ensure(max_diff_size(len, chunk_size, (uint32_t)(Kcap - 1))));
    out = diff.as<uint8_t>();
    out_cap = diff.bytes;
  }

  memset(&job, 0, sizeof job);
  job.data = data;
  job.len = len;
  job.chunk_size = chunk_size;
  job.hash_kind = hash_kind;
  job.C = C;
  job.N = N;
  job.nfull = len / chunk_size;
  job.D = make_geom(C).D;
  job.ckpt_id = K;
  job.baseline_id = baseline ? K : baseline_id;
  job.is_baseline = baseline ? 1 : 0;
  job.list_mode = method == DD_METHOD_LIST ? 1 : 0;
  job.node_bias = job.list_mode ? (uint32_t)(C - 1) : 0;
  job.write_all_labels = (baseline || job.list_mode || !labels_fixed) ? 1 : 0;
  // Wide probe rounds while the changed leaves stay under 1/8 of the chunks
  // (the last incremental checkpoint's count predicts this one's; a handle
  // with none yet starts wide): a wavefront then waits for its longest
  // chain.  Above that the inserts are bound by random-line traffic, where
  // narrow rounds keep more waves in flight (an MD5 field with a fifth of
  // its chunks changed: 7.8 ms per checkpoint instead of 8.1-20.8 ms).
  job.wide_probe = (!baseline && (last_pending == ~0ull || last_pending * 8 < Cn)) ? 1 : 0;
  // DD_INSERT_ROUNDS=narrow|wide forces the width of incremental probe rounds.
  static const char* rounds = std::getenv("DD_INSERT_ROUNDS");
  if (!baseline && rounds) job.wide_probe = strcmp(rounds, "wide") == 0 ? 1 : 0;
  job.root_shift = (options & DD_OPT_EMIT_ROOT_SHIFT) ? 1 : 0;
  if (baseline && bulk_now) fill_bulk(job);
  job.low_root = method == DD_METHOD_TREE_LOW_ROOT ? 1 : 0;
  job.lr_new = job.low_root ? lr_new.as<uint8_t>() : nullptr;
  labels_fixed = false;  // until this checkpoint completes
  job.tree = tree.as<Digest>();
  job.labels = labels.as<uint8_t>();
  job.slot_of = slot_of.as<uint64_t>();
  job.map = make_view(map.p, map_cap);
  job.map_count_at_start = map_count;
  job.map_limit = map_limit_of(map_cap);
  job.ctr = ctr.as<JobCounters>();
  job.host_ctr = nullptr;  // set by the pipeline
  job.blk_claim = leafblk.as<uint32_t>();
  job.int_claim = job.blk_claim + nclaim;
  // The counter sets alternate with the checkpoint's parity: this checkpoint
  // uses one, its k_sbin_scan zeroes the other for the next checkpoint.
  const uint32_t par = cnt_par ^ 1u;  // adopted by commit() once the checkpoint completes
  uint32_t* const pend_sets = reinterpret_cast<uint32_t*>(ctr.as<uint8_t>() + sizeof(JobCounters));
  job.pend_cnt = pend_sets + par * 32 * kCntStride;
  job.pend_cnt_next = pend_sets + (1 - par) * 32 * kCntStride;
  job.pend_list = pend_list.as<uint32_t>();
  job.pend_cap = (uint32_t)pend_cap;
  job.dig_cap = (uint32_t)dig_cap;
  job.pend_dig = pend_dig.as<Digest>();
  job.fint_list = tiles.as<uint32_t>();
  job.fint_n = job.fint_list + ntiles * kTileInterior;
  for (int g = 0; g < 4; g++) job.tile_base[g] = tile_base[g];
  job.ntiles = (uint32_t)ntiles;
  job.epoch = epoch;
  job.tile_epoch = tile_epoch.as<uint32_t>();
  job.dirty_cnt = lists.as<uint32_t>() + par * 32 * kCntStride;
  job.dirty_cnt_next = lists.as<uint32_t>() + (1 - par) * 32 * kCntStride;
  job.dirty_cap = (uint32_t)dirty_cap;
  job.dirty_list = lists.as<uint32_t>() + 2 * 32 * kCntStride;
  job.blk_f = blk.as<uint32_t>();
  job.blk_fc = job.blk_f + nblk;
  job.blk_s = job.blk_fc + nblk;
  job.nblk = (uint32_t)nblk;
  job.rmask = job.blk_s + nblk;
  job.lmask = reinterpret_cast<uint64_t*>(blk.as<uint8_t>() + lmask_off);
  job.f_list = f_list.as<uint32_t>();
  job.f_off = f_off.as<uint32_t>();
  job.s_node = s_node.as<uint32_t>();
  job.s_prev = s_prev.as<uint32_t>();
  job.s_tree = s_tree.as<uint32_t>();
  job.ss_node = ss_node.as<uint32_t>();
  job.ss_prev = ss_prev.as<uint32_t>();
  job.s_hist = s_hist.as<uint32_t>();
  job.ptab = ptab.as<uint32_t>();
  job.diff = out;
  job.diff_cap = out_cap;
  // A tree baseline's leaf kernel copies the chunks into the data section
  // when the handle's last baseline had every chunk as a first occurrence
  // (the pipeline keeps the request only with a leaf mapping that stages
  // the chunks; out_cap >= max_diff_size covers the whole copy).  Only for
  // chunks of at most 128 bytes: a round of a wavefront's 64 chunks is then
  // one contiguous range, read and written whole; with longer chunks the
  // round's pieces lie a chunk apart, and the copy's stores in that pattern
  // cost more than the gather's contiguous copy (1 GiB chains: +2-5 % at
  // 128 B, -1 % at 512 B, -4 to -5 % at 4096 B).  DD_LEAF_COPY=0 turns the
  // copy off (the gather then copies every chunk).
  static const bool copy_on = !(std::getenv("DD_LEAF_COPY") && strcmp(std::getenv("DD_LEAF_COPY"), "0") == 0);
  job.write_data =
      (copy_on && baseline && method != DD_METHOD_LIST && !data_miss && chunk_size <= 128) ? 1 : 0;
  return DD_OK;
}

void StreamState::commit(const JobCounters& c, bool baseline, bool leaf_copy) {
  if (baseline) baseline_id = current_id;
  need_baseline = false;
  // The root writer of a completed tree checkpoint reset every label that
  // was not FIXED; the list method keeps its labels.
  labels_fixed = method != DD_METHOD_LIST;
  cnt_par ^= 1u;  // this checkpoint's k_sbin_scan zeroed the other counter set
  if (!baseline) last_pending = c.pending;
  stats.chunks = C;
  stats.changed = baseline ? C : c.pending;
  stats.first_chunks = c.dchunks;
  stats.diff_bytes = c.diff_len;
  stats.baseline = baseline ? 1 : 0;
  stats.leaf_copy = leaf_copy ? 1 : 0;
  nfirst = c.nfirst;
  nshift = c.nshift;
  map_count = c.map_count;
  if (baseline) {
    base_map_count = map_count;
    ckpts_since_base = 0;
  } else {
    ckpts_since_base++;
  }
  last_diff_len = c.diff_len;
  last_map_ckpt = current_id;
  // The root alone as the first-occurrence region: the data section is every
  // chunk in the root's leaf order.
  if (baseline) data_miss = !(c.nfirst == 1 && c.nprior == 0 && c.nshift == 0 && c.dchunks == C);
}

// ---------------------------------------------------------------------------
// Pipeline
// ---------------------------------------------------------------------------
// S_BULK: partition tiles of the bulk build, S_BSCAN: its column-scan groups,
// S_BUCKET: its map buckets (none for jobs without the bulk build).
enum Stage { S_LEAF = 0, S_INSL, S_TILES, S_NODEBLK, S_STILE, S_GATHER, S_BULK, S_BSCAN, S_BUCKET, S_TILE0 };

// Batch-wide workgroup budget of an incremental grid-stride stage of
// 256-lane workgroups: DD_STAGE_BUDGET when set (0 otherwise), else the one
// the tables of a tuning key settled on (stage_budget_for); 8192 is 32 per CU
// of the 256 CUs, four rounds of the eight such workgroups a CU holds at once.
static uint64_t forced_stage_budget() {
  static const uint64_t b = [] {
    const char* e = std::getenv("DD_STAGE_BUDGET");
    const long long v = e ? std::atoll(e) : 0;
    return v > 0 ? (uint64_t)v : (uint64_t)0;
  }();
  return b;
}


Pipeline::~Pipeline() {
  if (host_table_) (void)hipHostFree(host_table_);
  if (host_ctr) (void)hipHostFree(host_ctr);
  for (hipEvent_t e : trial_ev_)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : bev_)
    if (e) (void)hipEventDestroy(e);
}

dd_status Pipeline::upload(const std::vector<JobDev>& jobs, hipStream_t s) {
  const size_t nj = jobs.size();
  if (nj == 0) return DD_EINVAL;
  njobs_ = (uint32_t)nj;
  max_groups_ = 0;
  aligned16_ = vec16_ = true;
  hash_kind_ = jobs[0].hash_kind;
  list_mode_ = jobs[0].list_mode != 0;
  low_root_ = false;
  max_c_ = 0;
  kinds_ = 0;
  bulk_dedup_ = false;
  for (const JobDev& J : jobs) {
    if (J.hash_kind != hash_kind_ || (J.list_mode != 0) != list_mode_) return DD_EINVAL;
    kinds_ |= J.is_baseline ? (J.bulk ? 0u : JOBS_DENSE) : J.wide_probe ? JOBS_LIST_WIDE : JOBS_LIST_NARROW;
    bulk_dedup_ = bulk_dedup_ || (J.bulk && J.bulk_dedup);
    low_root_ = low_root_ || J.low_root;
    max_c_ = std::max<uint64_t>(max_c_, J.C);
    max_groups_ = std::max(max_groups_, groups_of(J.D));
    const bool al = (reinterpret_cast<uintptr_t>(J.data) & 15) == 0 && (J.chunk_size & 15) == 0;
    aligned16_ = aligned16_ && al;
    vec16_ = vec16_ && al;
  }
  std::vector<uint64_t> key;
  bool incremental = true;  // a baseline's launches are not comparable with an incremental's
  for (const JobDev& J : jobs) incremental = incremental && !J.is_baseline;
  {
    std::vector<int> cand;
    leaf_map_ = leaf_mapping_for(jobs, aligned16_, hash_kind_, s, &key, &cand);
    trial_.current = -1;
    if (cand.size() < 2) {
      trial_ = LeafTrial{};
    } else {
      if (trial_.key != key || trial_.cand != cand) trial_ = LeafTrial{key, cand, std::vector<std::vector<float>>(cand.size()), -1};
      if (incremental) {
        size_t i = 0;
        for (size_t k = 1; k < cand.size(); k++)
          if (trial_.ms[k].size() < trial_.ms[i].size()) i = k;
        trial_.current = (int)i;
        leaf_map_ = cand[i];
      }
    }
  }
  if (!incremental) leaf_map_ = leaf_mapping_with_copy(jobs, aligned16_, hash_kind_, leaf_map_);
  const bool copy_data = leaf_writes_data(leaf_map_);
  copy_ = false;
  const size_t nstages = S_TILE0 + max_groups_;
  const size_t need_prefix = nstages * (nj + 1);
  const size_t jobs_bytes = (nj * sizeof(JobDev) + 255) & ~size_t(255);
  const size_t table_bytes = jobs_bytes + need_prefix * 4;
  if (table_bytes > host_table_cap_) {
    if (host_table_) (void)hipHostFree(host_table_);
    DD_HIP(hipHostMalloc((void**)&host_table_, table_bytes, hipHostMallocDefault));
    host_table_cap_ = table_bytes;
  }
  if (nj > host_ctr_cap_) {
    if (host_ctr) (void)hipHostFree(host_ctr);
    DD_HIP(hipHostMalloc((void**)&host_ctr, nj * sizeof(JobCounters), hipHostMallocDefault));
    host_ctr_cap_ = nj;
  }
  JobDev* hjobs = reinterpret_cast<JobDev*>(host_table_);
  uint32_t* hprefix = reinterpret_cast<uint32_t*>(host_table_ + jobs_bytes);
  off_.assign(nstages, 0);
  grid_.assign(nstages, 0);
  // Workgroups of stage st for job J before the batch-wide budget.
  auto stage_wgs = [&](size_t st, const JobDev& J) -> uint64_t {
    uint64_t n = 0;
    switch (st) {
      case S_LEAF: n = leaf_workgroups(leaf_map_, J.C); break;
      // A baseline visits every leaf block and tile (one workgroup each);
      // an incremental checkpoint strides over its work lists.
      // The insert stages take one list entry per lane: a baseline lists
      // (nearly) every leaf and interior node.
      // A bulk baseline has no dense inserts (a fallback runs it again
      // without the bulk build).
      case S_INSL:
      case S_TILES:
        n = J.is_baseline ? (J.bulk ? 0 : ceil_div(J.C, 256)) : std::min<uint64_t>(ceil_div(J.C, 256), kMaxStageWG);
        break;
      case S_BULK: n = J.bulk ? J.ntiles_b : 0; break;
      case S_BSCAN: n = J.bulk ? ceil_div(J.nbuckets, kColBuckets) : 0; break;
      case S_BUCKET: n = J.bulk ? J.nbuckets : 0; break;
      case S_NODEBLK: n = std::min<uint64_t>(J.nblk, kRootWG); break;
      case S_STILE: n = std::min<uint64_t>(ceil_div(J.C, kSTile), 512); break;  // grid-stride over sort tiles
      case S_GATHER: n = std::min<uint64_t>(ceil_div(J.C, 256), 4096); break;
      default: {
        const uint64_t tg = tiles_in_group(J.D, (int)st - S_TILE0);
        n = J.is_baseline ? tg : std::min<uint64_t>(tg, kMaxTileWG);
        break;
      }
    }
    return n;
  };
  // The budget: forced, settled for the key, or (an incremental table whose
  // stages the smallest candidate would cap, while the key is unsettled and
  // no leaf-mapping trial runs) the candidate with the fewest timed launches.
  budget_ = kBudgetCand[0];
  btrial_.current = -1;
  if (forced_stage_budget()) {
    budget_ = forced_stage_budget();
  } else if (!key.empty()) {
    bool settled = true;
    budget_ = stage_budget_for(key, &settled);
    bool binds = false;
    for (size_t st = 1; st < nstages && !binds; st++) {
      uint64_t incr = 0;
      for (const JobDev& J : jobs)
        if (!J.is_baseline) incr += stage_wgs(st, J);
      binds = incr > (st >= S_TILE0 ? 4 : 1) * kBudgetCand[kBudgetCands - 1];
    }
    if (!settled && incremental && binds && trial_.cand.empty()) {
      if (btrial_.key != key) btrial_ = BudgetTrial{key, std::vector<std::vector<float>>(kBudgetCands), -1};
      size_t i = 0;
      for (size_t k = 1; k < kBudgetCands; k++)
        if (btrial_.ms[k].size() < btrial_.ms[i].size()) i = k;
      btrial_.current = (int)i;
      budget_ = kBudgetCand[i];
    }
  }
  std::vector<uint64_t> nwg(nj);
  const uint64_t budget = budget_;
  for (size_t st = 0; st < nstages; st++) {
    off_[st] = (uint32_t)(st * (nj + 1));
    uint64_t incr = 0;  // workgroups of the incremental jobs
    for (size_t j = 0; j < nj; j++) {
      const JobDev& J = jobs[j];
      const uint64_t n = stage_wgs(st, J);
      nwg[j] = n;
      if (!J.is_baseline && st != S_LEAF) incr += n;
    }
    // The incremental stages stride over work lists, so any workgroup count
    // gives the same result; their per-job caps add up over a batch of many
    // jobs to far more workgroups than the device holds, most of which only
    // find that the lists are done.  Over a batch-wide budget (four times for
    // the one-wavefront tile passes), each incremental job keeps its share in
    // proportion, at least one workgroup.
    const uint64_t cap = st >= S_TILE0 ? 4 * budget : budget;
    uint64_t acc = 0;
    for (size_t j = 0; j < nj; j++) {
      uint64_t n = nwg[j];
      if (incr > cap && !jobs[j].is_baseline && st != S_LEAF) n = std::max<uint64_t>(1, n * cap / incr);
      hprefix[off_[st] + j] = (uint32_t)acc;
      acc += n;
    }
    if (acc > 0x7fffffffull) return DD_EINVAL;
    hprefix[off_[st] + nj] = (uint32_t)acc;
    grid_[st] = (uint32_t)acc;
  }
  for (size_t j = 0; j < nj; j++) {
    hjobs[j] = jobs[j];
    hjobs[j].host_ctr = &host_ctr[j];
    hjobs[j].write_data = copy_data ? jobs[j].write_data : 0;
    copy_ = copy_ || hjobs[j].write_data;
    hjobs[j].nwg_ins_leaf = hprefix[off_[S_INSL] + j + 1] - hprefix[off_[S_INSL] + j];
    // A list job launches no interior inserts, so it has no interior claims
    // for sbin_scan to add to its entry count.
    hjobs[j].nwg_ins_int = jobs[j].list_mode ? 0 : hprefix[off_[S_TILES] + j + 1] - hprefix[off_[S_TILES] + j];
  }
  DD_TRY(dev_table_.ensure(table_bytes));
  dev_jobs_ = dev_table_.as<JobDev>();
  dev_prefix_ = reinterpret_cast<uint32_t*>(dev_table_.as<uint8_t>() + jobs_bytes);
  DD_HIP(hipMemcpyAsync(dev_table_.p, host_table_, table_bytes, hipMemcpyHostToDevice, s));
  return DD_OK;
}

std::vector<uint64_t> Pipeline::signature() const {
  std::vector<uint64_t> sig;
  // The graph holds the stages after the leaf kernel: the leaf mapping and
  // grid are not part of it.
  sig.push_back(njobs_);
  sig.push_back((uint64_t)max_groups_);
  sig.push_back((uint64_t)aligned16_ | ((uint64_t)vec16_ << 1));
  sig.push_back((uint64_t)hash_kind_ | ((uint64_t)list_mode_ << 8) | ((uint64_t)kinds_ << 16) |
                ((uint64_t)low_root_ << 24) | ((uint64_t)bulk_dedup_ << 25));
  sig.push_back(low_root_ ? max_c_ : 0);
  sig.push_back(reinterpret_cast<uintptr_t>(dev_table_.p));
  for (uint32_t o : off_) sig.push_back(o);
  for (size_t st = 0; st < grid_.size(); st++)
    if (st != S_LEAF) sig.push_back(grid_[st]);
  return sig;
}

void Pipeline::trial_done() {
  if (btrial_.current >= 0) {
    float ms = 0;
    const size_t i = (size_t)btrial_.current;
    btrial_.current = -1;
    if (btrial_end_ && hipEventElapsedTime(&ms, bev_[0], bev_[1]) == hipSuccess && ms > 0) {
      btrial_.ms[i].push_back(ms);
      bool all = true;
      for (const auto& v : btrial_.ms) all = all && v.size() >= kTrialRounds;
      if (all) {
        size_t best = 0;
        float best_ms = 0;
        for (size_t k = 0; k < kBudgetCands; k++) {
          std::vector<float> v = btrial_.ms[k];
          std::sort(v.begin(), v.end());
          if (k == 0 || v[v.size() / 2] < best_ms) {
            best_ms = v[v.size() / 2];
            best = k;
          }
        }
        stage_budget_settle(btrial_.key, kBudgetCand[best]);
        btrial_ = BudgetTrial{};
      }
    } else {
      (void)hipGetLastError();
    }
    btrial_end_ = false;
  }
  if (trial_.current < 0) return;
  float ms = 0;
  const size_t i = (size_t)trial_.current;
  trial_.current = -1;
  if (hipEventElapsedTime(&ms, trial_used_[0], trial_used_[1]) != hipSuccess || ms <= 0) {
    (void)hipGetLastError();
    return;
  }
  trial_.ms[i].push_back(ms);
  for (const auto& v : trial_.ms)
    if (v.size() < kTrialRounds) return;
  size_t best = 0;
  float best_ms = 0;
  for (size_t k = 0; k < trial_.cand.size(); k++) {
    std::vector<float> v = trial_.ms[k];
    std::sort(v.begin(), v.end());
    if (k == 0 || v[v.size() / 2] < best_ms) {
      best_ms = v[v.size() / 2];
      best = k;
    }
  }
  leaf_mapping_settle(trial_.key, trial_.cand[best]);
  leaf_map_ = trial_.cand[best];
  trial_ = LeafTrial{};
}

dd_status Pipeline::run(const std::vector<JobDev>& jobs, hipStream_t s, bool rerun, bool begin, hipEvent_t ev_dedup,
                        hipEvent_t* stage_ev) {
  if (jobs.empty()) return DD_EINVAL;
  DD_TRY(upload(jobs, s));
  return launch_all(s, rerun, begin, ev_dedup, stage_ev);
}

dd_status Pipeline::launch_all(hipStream_t s, bool rerun, bool begin, hipEvent_t ev_dedup, hipEvent_t* stage_ev,
                               int part) {
  const uint32_t nj = njobs_;
  const JobDev* J = dev_jobs_;
  auto P = [&](int st) { return dev_prefix_ + off_[st]; };
  // Level 2 records an event between every two stages; level 1 times only
  // the leaf kernel, with events attached to its dispatch.
  auto mark = [&](int i) -> dd_status {
    if (stage_ev && profile_level_ >= 2) DD_HIP(hipEventRecord(stage_ev[i], s));
    return DD_OK;
  };
  const bool leaf_events = stage_ev && profile_level_ == 1;
  if (part != PART_TAIL) {
    if (rerun) btrial_.current = -1;  // a re-run's launches are not a whole checkpoint
    if (btrial_.current >= 0) {
      for (hipEvent_t& e : bev_)
        if (!e) DD_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
      DD_HIP(hipEventRecord(bev_[0], s));
      btrial_end_ = false;
    }
    if (rerun || begin) launch_begin(J, nullptr, nj, rerun, s);
    DD_TRY(mark(0));
    if (rerun) trial_.current = -1;  // no leaf kernel in a re-run
    if (!rerun) {
      hipEvent_t e0 = leaf_events ? stage_ev[0] : nullptr, e1 = leaf_events ? stage_ev[1] : nullptr;
      if (trial_.current >= 0 && !e0) {
        for (hipEvent_t& e : trial_ev_)
          if (!e) DD_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        e0 = trial_ev_[0];
        e1 = trial_ev_[1];
      }
      trial_used_[0] = e0;
      trial_used_[1] = e1;
      launch_leaf(J, P(S_LEAF), nj, grid_[S_LEAF], leaf_map_, aligned16_, hash_kind_, s, e0, e1, copy_);
    }
    DD_TRY(mark(1));
    if (part == PART_HEAD) {
      DD_HIP(hipGetLastError());
      return DD_OK;
    }
  }
  launch_bulk(J, nj, false, bulk_dedup_, P(S_BULK), grid_[S_BULK], P(S_BSCAN), grid_[S_BSCAN], P(S_BUCKET),
              grid_[S_BUCKET], s);
  launch_insert_leaves(J, P(S_INSL), nj, grid_[S_INSL], kinds_, s);
  if (low_root_) launch_low_root(J, nj, max_c_, (uint32_t)ilog2_64(2 * max_c_ - 1), s);
  DD_TRY(mark(2));
  if (list_mode_) {
    // The list method has no tree: its chunks are classified by the inserts.
    launch_list_resolve(J, P(S_INSL), nj, grid_[S_INSL], s);
    DD_TRY(mark(3));
    DD_TRY(mark(4));
  } else {
    const bool narrow = 2 * max_c_ - 1 < (1ull << 31);  // every tree below 2^31 nodes
    for (int g = 0; g < max_groups_; g++) launch_pass1(J, P(S_TILE0 + g), nj, grid_[S_TILE0 + g], g, narrow, s);
    DD_TRY(mark(3));
    launch_insert_interior(J, P(S_TILES), nj, grid_[S_TILES], kinds_, s);
    launch_bulk(J, nj, true, bulk_dedup_, P(S_BULK), grid_[S_BULK], P(S_BSCAN), grid_[S_BSCAN], P(S_BUCKET),
                grid_[S_BUCKET], s);
    DD_TRY(mark(4));
    for (int g = 0; g < max_groups_; g++) launch_pass2(J, P(S_TILE0 + g), nj, grid_[S_TILE0 + g], g, narrow, s);
  }
  DD_TRY(mark(5));
  if (ev_dedup) DD_HIP(hipEventRecord(ev_dedup, s));
  launch_roots_count(J, P(S_NODEBLK), nj, grid_[S_NODEBLK], s);
  launch_roots_write(J, P(S_NODEBLK), nj, grid_[S_NODEBLK], s);
  launch_sbin_count(J, P(S_STILE), nj, grid_[S_STILE], s);
  launch_sbin_scan(J, nj, s);
  launch_sbin_scatter(J, P(S_STILE), nj, grid_[S_STILE], s);
  DD_TRY(mark(6));
  launch_gather(J, P(S_GATHER), nj, grid_[S_GATHER], vec16_, s);
  DD_TRY(mark(7));
  if (part == PART_ALL) DD_TRY(trial_mark_end(s));
  DD_HIP(hipGetLastError());
  return DD_OK;
}

dd_status Pipeline::trial_mark_end(hipStream_t s) {
  if (btrial_.current >= 0 && bev_[1]) {
    DD_HIP(hipEventRecord(bev_[1], s));
    btrial_end_ = true;
  }
  return DD_OK;
}

dd_status Pipeline::fetch_counters(const std::vector<JobDev>&, hipStream_t s) {
  DD_HIP(hipStreamSynchronize(s));
  return DD_OK;
}

}  // namespace dd

// ===========================================================================
// C ABI
// ===========================================================================
using namespace dd;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// A bulk baseline run without the leaf duplicate search that found a repeated
// leaf digest runs again from the leaf kernel with the search (and so do the
// handle's later baselines).
dd_status rerun_bulk_dedup(StreamState& st, JobDev& job, Pipeline& p, hipStream_t s, JobCounters* c) {
  st.bulk_dups = true;
  job.bulk_dedup = 1;
  DD_TRY(p.run(std::vector<JobDev>(1, job), s, false, true));
  DD_TRY(p.fetch_counters(std::vector<JobDev>(1, job), s));
  *c = p.host_ctr[0];
  return DD_OK;
}

// A bulk baseline that stopped (a bucket with more leaf records than the
// build holds, or a final bucket that does not fit) runs again from the leaf
// kernel with the dense inserts, on a cleared map; after a bucket of repeated
// leaf digests the handle's later baselines go straight to the dense inserts.
dd_status rerun_dense_baseline(StreamState& st, JobDev& job, Pipeline& p, hipStream_t s, JobCounters* c) {
  if (c->flags & FLAG_BULK_OFF) st.bulk_heavy = true;
  job.bulk = 0;
  launch_map_clear(st.map.as<Slot>(), st.map_cap, s);
  DD_TRY(p.run(std::vector<JobDev>(1, job), s, false, true));
  DD_TRY(p.fetch_counters(std::vector<JobDev>(1, job), s));
  *c = p.host_ctr[0];
  return DD_OK;
}

// One checkpoint of a handle: runs the pipeline, grows the map and re-runs the
// inserts when the map would pass its load limit, and commits the state.
dd_status run_checkpoint(dd_dedup* h, const void* dev_data, uint64_t len, int make_baseline, uint8_t* out,
                         uint64_t out_cap, JobCounters* result) {
  const double t0 = now_s();
  DD_HIP(hipSetDevice(h->device));
  StreamState& st = h->st;
  const bool baseline = st.runs_baseline(make_baseline);
  JobDev job;
  DD_TRY(st.prepare(reinterpret_cast<const uint8_t*>(dev_data), len, baseline, out, out_cap, job, h->stream));
  std::vector<JobDev> jobs(1, job);
  st.timers[0] = now_s() - t0;
  // The counters are armed by the previous successful checkpoint (k_sbin_scan);
  // a baseline resets the entry count, so it starts with k_begin.
  const bool begin = baseline || !st.ctr_armed;
  st.ctr_armed = false;
  // The upload (and, on a table's first use, the leaf mapping's tuning) runs
  // before the timers' first event.
  DD_TRY(h->pipe.upload(jobs, h->stream));
  DD_HIP(hipEventRecord(h->ev[0], h->stream));
  h->pipe.profile_level_ = h->profile;
  // Events in the middle of the stream cost the host tens of microseconds per
  // checkpoint: the dedup/gather split is recorded only at profiling level 2,
  // and level 1 attaches its two events to the leaf kernel's dispatch.
  DD_TRY(h->pipe.launch_all(h->stream, false, begin, h->profile >= 2 ? h->ev[1] : nullptr,
                            h->profile ? h->pev : nullptr));
  DD_HIP(hipEventRecord(h->ev[2], h->stream));
  DD_TRY(h->pipe.fetch_counters(jobs, h->stream));
  h->pipe.trial_done();
  JobCounters c = h->pipe.host_ctr[0];
  if (jobs[0].bulk && (c.flags & FLAG_BULK_DUPS)) {
    DD_TRY(rerun_bulk_dedup(st, jobs[0], h->pipe, h->stream, &c));
    DD_HIP(hipEventRecord(h->ev[2], h->stream));
  }
  if (jobs[0].bulk && (c.flags & (FLAG_NEED_GROW | FLAG_MAP_FULL))) {
    DD_TRY(rerun_dense_baseline(st, jobs[0], h->pipe, h->stream, &c));
    DD_HIP(hipEventRecord(h->ev[2], h->stream));
  }
  for (int attempt = 0; (c.flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) && attempt < 4; attempt++) {
    const uint64_t need = c.map_count + 2 * c.pending;
    DD_TRY(st.grow_map(need, h->stream));
    jobs[0].map = make_view(st.map.p, st.map_cap);
    jobs[0].map_count_at_start = c.map_count;
    jobs[0].map_limit = map_limit_of(st.map_cap);
    DD_TRY(h->pipe.run(jobs, h->stream, true, true));
    DD_HIP(hipEventRecord(h->ev[2], h->stream));
    DD_TRY(h->pipe.fetch_counters(jobs, h->stream));
    c = h->pipe.host_ctr[0];
  }
  if (c.flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return DD_ENOMEM;
  if (c.flags & FLAG_BAD_SOURCE) return DD_ECORRUPT;
  float ms1 = 0, ms2 = 0;
  if (h->profile >= 2) {
    (void)hipEventElapsedTime(&ms1, h->ev[0], h->ev[1]);
    (void)hipEventElapsedTime(&ms2, h->ev[1], h->ev[2]);
  } else {
    (void)hipEventElapsedTime(&ms1, h->ev[0], h->ev[2]);
  }
  st.timers[1] = ms1 * 1e-3;
  st.timers[2] = ms2 * 1e-3;
  if (h->profile) {
    for (int i = 0; i < DD_NUM_STAGES; i++) {
      float ms = 0;
      if (h->profile >= 2 || i == 0) (void)hipEventElapsedTime(&ms, h->pev[i], h->pev[i + 1]);
      h->stage_ms[i] = ms;
    }
  }
  st.commit(c, baseline, h->pipe.job_copied(0));
  st.ctr_armed = true;  // k_sbin_scan zeroed the next checkpoint's counter set
  *result = c;
  return DD_OK;
}

void fill_header(const dd_dedup* h, const JobCounters& c, uint64_t len, dd_header* hd) {
  hd->ref_id = h->st.baseline_id;
  hd->chkpt_id = h->st.current_id;
  hd->datalen = len;
  hd->chunk_size = h->st.chunk_size;
  hd->num_first_ocur = c.nfirst;
  hd->num_prior_chkpts = c.nprior;
  hd->num_shift_dupl = c.nshift;
}

dd_status checkpoint_to_host(dd_dedup* h, dd_header* header, const void* dev_data, uint64_t len, int make_baseline,
                             dd_buf* out) {
  if (!h || !out || !dev_data) return DD_EINVAL;
  out->data = nullptr;
  out->len = 0;
  JobCounters c;
  DD_TRY(run_checkpoint(h, dev_data, len, make_baseline, nullptr, 0, &c));
  const double t0 = now_s();
  uint8_t* buf = static_cast<uint8_t*>(std::malloc(c.diff_len));
  if (!buf) return DD_ENOMEM;
  if (hipMemcpyAsync(buf, h->st.diff.p, c.diff_len, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess) {
    std::free(buf);
    return DD_EHIP;
  }
  h->st.timers[3] = now_s() - t0;
  out->data = buf;
  out->len = c.diff_len;
  if (header) fill_header(h, c, len, header);
  return DD_OK;
}

// ---------------------------------------------------------------------------
// Restart
// ---------------------------------------------------------------------------
uint32_t rd32(const uint8_t* p) {
  uint32_t v;
  memcpy(&v, p, 4);
  return v;
}

uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

// Device memory and host staging of one restart: the handle's own buffers,
// or (for callers that cannot allocate) the library's static scratch.
class RestartSpace {
 public:
  virtual ~RestartSpace() = default;
  // Device staging for a header round trip of `want` bytes; *cap is what it
  // holds (at least kHdrWindow headers).
  virtual dd_status headers(uint64_t want, uint8_t** dev, uint64_t* cap) = 0;
  // Cover table, pending sources and pending list for C chunks, and the
  // stamp counter of that cover table.  A cover table that is new (or whose
  // contents are unknown) reads as all zero, with the counter at zero.
  virtual dd_status tables(uint64_t C, uint2** cover, uint64_t** ref, uint32_t** pend, uint32_t** stamp) = 0;
  // Device arena of at least `bytes` (called after tables()).
  virtual dd_status arena(uint64_t bytes, uint8_t** dev) = 0;
  // Host staging of at least `bytes`; pinned when `pinned()`.
  virtual dd_status host(uint64_t bytes, uint8_t** p) = 0;
  virtual bool pinned() const = 0;
};

constexpr uint32_t kHdrWindow = 64;

// The handle's restart buffers: grow-only device buffers and pinned staging.
class HandleSpace : public RestartSpace {
 public:
  HandleSpace(dd_dedup* h) : h_(h) {}
  dd_status headers(uint64_t want, uint8_t** dev, uint64_t* cap) override {
    // Grows to at most 1 MiB: windows of very long chains fetch headers only.
    DD_TRY(h_->rs_hdr.ensure(std::max<uint64_t>(std::min<uint64_t>(want, 1 << 20), kHdrWindow * 32)));
    *dev = h_->rs_hdr.as<uint8_t>();
    *cap = h_->rs_hdr.bytes;
    return DD_OK;
  }
  dd_status tables(uint64_t C, uint2** cover, uint64_t** ref, uint32_t** pend, uint32_t** stamp) override {
    if (h_->cover.bytes < C * sizeof(uint2)) {
      h_->cover.release();
      DD_TRY(h_->cover.ensure(C * sizeof(uint2), true, h_->stream));
      h_->stamp = 0;
    }
    DD_TRY(h_->ref.ensure(C * 8));
    DD_TRY(h_->pend.ensure(C * 4));
    *cover = h_->cover.as<uint2>();
    *ref = h_->ref.as<uint64_t>();
    *pend = h_->pend.as<uint32_t>();
    *stamp = &h_->stamp;
    return DD_OK;
  }
  dd_status arena(uint64_t bytes, uint8_t** dev) override {
    // Grows by at least half its size: restarts of chains that grow by one
    // diff at a time would otherwise reallocate on every call.
    if (bytes > h_->rs_arena.bytes) DD_TRY(h_->rs_arena.ensure(std::max<uint64_t>(bytes, h_->rs_arena.bytes * 3 / 2)));
    *dev = h_->rs_arena.as<uint8_t>();
    return DD_OK;
  }
  dd_status host(uint64_t bytes, uint8_t** p) override {
    if (bytes > h_->rs_host_cap) {
      // the previous restart's copies from the staging have completed (it synchronized)
      if (h_->rs_host) (void)hipHostFree(h_->rs_host);
      h_->rs_host = nullptr;
      h_->rs_host_cap = 0;
      const uint64_t n = std::max<uint64_t>(bytes, 64 << 10);
      DD_HIP(hipHostMalloc((void**)&h_->rs_host, n, hipHostMallocDefault));
      h_->rs_host_cap = n;
    }
    *p = h_->rs_host;
    return DD_OK;
  }
  bool pinned() const override { return true; }

 private:
  dd_dedup* h_;
};

// The allocation-free restart's buffers: one static device array per device,
// laid out independently of the request, so that no part of a cover table
// ever holds another request's pending list or arena bytes:
// [headers 4 KiB][cover 8 B x max chunks][pending sources 8 B x max][pending
// list 4 B x max][arena].  Host staging is ordinary (pageable) memory.
static __device__ __align__(256) uint8_t scratch[DD_NOALLOC_SCRATCH_BYTES];
constexpr uint64_t kNaCover = 4096;
constexpr uint64_t kNaRef = kNaCover + 8 * DD_NOALLOC_MAX_CHUNKS;
constexpr uint64_t kNaPend = kNaRef + 8 * DD_NOALLOC_MAX_CHUNKS;
constexpr uint64_t kNaArena = kNaPend + 4 * DD_NOALLOC_MAX_CHUNKS;
static_assert(kNaArena + DD_NOALLOC_ARENA_BYTES == DD_NOALLOC_SCRATCH_BYTES, "scratch layout");
static_assert(kHdrWindow * 32 <= kNaCover, "header staging");
constexpr int kNaDevices = 64;

// Per device: the device address of `scratch` (from hipGetSymbolAddress,
// with that device current), the last stamp written into its cover table,
// and the lock that serializes the calls using it.
struct NoAllocDevice {
  std::mutex mu;
  uint8_t* ptr = nullptr;
  uint32_t stamp = 0;
};
NoAllocDevice na_dev[kNaDevices];

class NoAllocSpace : public RestartSpace {
 public:
  explicit NoAllocSpace(NoAllocDevice& d) : d_(d) {}
  dd_status headers(uint64_t, uint8_t** dev, uint64_t* cap) override {
    *dev = d_.ptr;
    *cap = kNaCover;
    return DD_OK;
  }
  dd_status tables(uint64_t C, uint2** cover, uint64_t** ref, uint32_t** pend, uint32_t** stamp) override {
    if (C > DD_NOALLOC_MAX_CHUNKS) return DD_ELIMIT;
    *cover = reinterpret_cast<uint2*>(d_.ptr + kNaCover);
    *ref = reinterpret_cast<uint64_t*>(d_.ptr + kNaRef);
    *pend = reinterpret_cast<uint32_t*>(d_.ptr + kNaPend);
    *stamp = &d_.stamp;
    return DD_OK;
  }
  dd_status arena(uint64_t bytes, uint8_t** dev) override {
    if (bytes > DD_NOALLOC_ARENA_BYTES) return DD_ELIMIT;
    *dev = d_.ptr + kNaArena;
    return DD_OK;
  }
  dd_status host(uint64_t bytes, uint8_t** p) override {
    if (host_.size() < bytes) host_.resize(bytes);
    *p = host_.data();
    return DD_OK;
  }
  bool pinned() const override { return false; }

 private:
  NoAllocDevice& d_;
  std::vector<uint8_t> host_;
};

// Validated headers of the checkpoints a restart visits: chkpt_id down to the
// reference id named by each visited header (src/tree_approach.cpp:943).
// `hdr_of(i, hdr)` returns diff i's header.
template <typename HdrOf>
dd_status walk_headers(uint32_t chkpt_id, uint64_t len, HdrOf hdr_of, std::vector<uint32_t>& ids,
                       std::vector<dd_header>& hdrs) {
  int64_t idx = chkpt_id, ref = 0;
  bool first = true;
  while (first || idx >= ref) {
    dd_header hd;
    DD_TRY(hdr_of((uint32_t)idx, hd));
    if (first) {
      if (hd.datalen != len) return DD_EINVAL;
      if (hd.chunk_size == 0 || ceil_div(len, hd.chunk_size) > 0x7fffffffull) return DD_ECORRUPT;
    } else if (hd.datalen != len || hd.chunk_size != hdrs[0].chunk_size) {
      return DD_ECORRUPT;
    }
    if (hd.chkpt_id != (uint32_t)idx || hd.ref_id > hd.chkpt_id) return DD_ECORRUPT;
    ids.push_back((uint32_t)idx);
    hdrs.push_back(hd);
    ref = hd.ref_id;
    first = false;
    idx--;
  }
  return DD_OK;
}

// Copies device byte ranges into one device buffer with k_fetch launches
// (kFetchRanges per launch), then into host memory; waits for the copy.
// With `lim` (diff lengths), each range is a diff whose header and prior
// table are fetched into n[i] bytes.
dd_status fetch_to_host(const std::vector<const uint8_t*>& src, const std::vector<uint64_t>& n,
                        const std::vector<uint64_t>& off, uint8_t* dev_dst, uint8_t* host_dst, uint64_t total,
                        hipStream_t s, const std::vector<uint64_t>* lim = nullptr) {
  for (size_t b = 0; b < src.size(); b += kFetchRanges) {
    FetchArgs a;
    memset(&a, 0, sizeof a);
    a.tables = lim ? 1u : 0u;
    uint64_t mx = 0;
    for (size_t i = b; i < src.size() && i < b + kFetchRanges; i++) {
      a.src[a.n] = src[i];
      a.len[a.n] = n[i];
      a.dst_off[a.n] = off[i];
      a.lim[a.n] = lim ? (*lim)[i] : 0;
      mx = std::max(mx, n[i]);
      a.n++;
    }
    launch_fetch(a, mx, dev_dst, s);
  }
  DD_HIP(hipGetLastError());
  if (total) DD_HIP(hipMemcpyAsync(host_dst, dev_dst, total, hipMemcpyDeviceToHost, s));
  DD_HIP(hipStreamSynchronize(s));
  return DD_OK;
}

// Restart of checkpoint chkpt_id from diffs in host memory (`host_chain`) or
// device memory (`dev_chain`); exactly one of them is given.
dd_status restart_run(RestartSpace& sp, hipStream_t s, int method, const dd_buf* host_chain,
                      const void* const* dev_chain, const uint64_t* dev_lens, uint32_t nchain, uint32_t chkpt_id,
                      void* dev_out, uint64_t len, double timers[2]) {
  const double t0 = now_s();
  if (chkpt_id >= nchain) return DD_EINVAL;
  auto total_of = [&](uint32_t i) -> uint64_t { return host_chain ? host_chain[i].len : dev_lens[i]; };
  auto valid = [&](uint32_t i) -> bool {
    return (host_chain ? host_chain[i].data != nullptr : dev_chain[i] != nullptr) && total_of(i) >= 32;
  };
  // 1. Headers of the visited checkpoints (device diffs: read in windows of
  // kHdrWindow ids per round trip, with their prior tables where the staging
  // holds them).
  std::vector<uint32_t> ids;
  std::vector<dd_header> hdrs;
  std::vector<std::vector<uint8_t>> tabs;  // device diffs: prior table of id i, when fetched whole
  std::vector<char> have_tab;
  if (host_chain) {
    DD_TRY(walk_headers(chkpt_id, len, [&](uint32_t i, dd_header& hd) -> dd_status {
      if (!valid(i)) return DD_ECORRUPT;
      memcpy(&hd, host_chain[i].data, 32);
      return DD_OK;
    }, ids, hdrs));
  } else {
    std::vector<dd_header> known((size_t)chkpt_id + 1);
    std::vector<char> have((size_t)chkpt_id + 1, 0);
    tabs.resize((size_t)chkpt_id + 1);
    have_tab.assign((size_t)chkpt_id + 1, 0);
    DD_TRY(walk_headers(chkpt_id, len, [&](uint32_t i, dd_header& hd) -> dd_status {
      if (!valid(i)) return DD_ECORRUPT;
      if (!have[i]) {
        const uint32_t lo = i + 1 >= kHdrWindow ? i + 1 - kHdrWindow : 0;
        std::vector<const uint8_t*> src;
        std::vector<uint64_t> n, off, lim;
        std::vector<uint32_t> got;
        uint64_t total = 0;
        for (uint32_t k = lo; k <= i; k++) {
          if (have[k] || !valid(k)) continue;
          src.push_back(static_cast<const uint8_t*>(dev_chain[k]));
          // A valid diff k lists at most k + 1 prior checkpoints.
          n.push_back(32 + 8ull * (k + 1));
          off.push_back(total);
          lim.push_back(dev_lens[k]);
          total += align256(n.back());
          got.push_back(k);
        }
        uint8_t* stage = nullptr;
        uint64_t cap = 0;
        DD_TRY(sp.headers(total, &stage, &cap));
        const bool with_tab = total <= cap;
        if (!with_tab) {  // headers only
          for (size_t k = 0; k < got.size(); k++) {
            n[k] = 32;
            off[k] = 32ull * k;
          }
          total = 32ull * got.size();
        }
        uint8_t* hst = nullptr;
        DD_TRY(sp.host(total, &hst));
        DD_TRY(fetch_to_host(src, n, off, stage, hst, total, s, with_tab ? &lim : nullptr));
        for (size_t k = 0; k < got.size(); k++) {
          dd_header& kh = known[got[k]];
          memcpy(&kh, hst + off[k], 32);
          have[got[k]] = 1;
          // The table is whole when it fitted its staging and the diff.
          const uint64_t at = 32 + 4ull * kh.num_first_ocur, tb = 8ull * kh.num_prior_chkpts;
          if (with_tab && tb <= n[k] - 32 && at + tb <= dev_lens[got[k]]) {
            tabs[got[k]].assign(hst + off[k] + 32, hst + off[k] + 32 + tb);
            have_tab[got[k]] = 1;
          }
        }
      }
      hd = known[i];
      return DD_OK;
    }, ids, hdrs));
  }
  const uint32_t cs = hdrs[0].chunk_size;
  const uint64_t C = ceil_div(len, cs);
  const TreeGeom g = make_geom(C);
  const uint32_t bias = method == DD_METHOD_LIST ? (uint32_t)(C - 1) : 0;
  const size_t M = ids.size();
  uint2* cover = nullptr;
  uint64_t* ref = nullptr;
  uint32_t* pend = nullptr;
  uint32_t* stamp = nullptr;
  DD_TRY(sp.tables(C, &cover, &ref, &pend, &stamp));
  // 2. Layout.  Region lists are parsed on the device; the host reads only
  // the prior tables (8 bytes per source checkpoint).
  std::vector<uint64_t> meta_bytes(M), tab_bytes(M), tab_off(M);
  uint64_t tab_total = 0, nblk = 0;
  std::vector<RestartList> lists(2 * M);
  std::vector<uint32_t> list_blk(2 * M + 1);
  for (size_t k = 0; k < M; k++) {
    const dd_header& hd = hdrs[k];
    meta_bytes[k] = 4ull * hd.num_first_ocur + 8ull * hd.num_prior_chkpts + 8ull * hd.num_shift_dupl;
    if (32 + meta_bytes[k] > total_of(ids[k])) return DD_ECORRUPT;
    if (hd.num_first_ocur > C || hd.num_shift_dupl > C) return DD_ECORRUPT;  // regions are disjoint subtrees
    tab_bytes[k] = 8ull * hd.num_prior_chkpts;
    tab_off[k] = tab_total;
    tab_total += tab_bytes[k];
    for (int sh = 0; sh < 2; sh++) {
      RestartList& L = lists[2 * k + sh];
      L.ckpt = (uint32_t)k;
      L.shift = (uint32_t)sh;
      L.n = sh ? hd.num_shift_dupl : hd.num_first_ocur;
      L.blk0 = (uint32_t)nblk;
      list_blk[2 * k + sh] = (uint32_t)nblk;
      nblk += ceil_div(L.n, kParseBlock);
    }
  }
  list_blk[2 * M] = (uint32_t)nblk;
  if (nblk > 0x7fffffffull) return DD_ECORRUPT;
  // Prior tables -> per checkpoint, the end of each source checkpoint's run
  // of the shifted list (cumulative counts in checkpoint order).
  std::vector<std::vector<uint32_t>> pb_end(M), pb_ck(M);
  uint64_t pb_total = 0;
  {
    uint8_t* stage = nullptr;
    uint8_t* hst = nullptr;
    // Device diffs: the tables the header round trips did not bring, in one
    // more (none when the walk has no prior table: a baseline alone).
    std::vector<const uint8_t*> src;
    std::vector<uint64_t> n, off;
    uint64_t miss = 0;
    for (size_t k = 0; dev_chain && k < M; k++) {
      if (!tab_bytes[k] || have_tab[ids[k]]) continue;
      src.push_back(static_cast<const uint8_t*>(dev_chain[ids[k]]) + 32 + 4ull * hdrs[k].num_first_ocur);
      n.push_back(tab_bytes[k]);
      off.push_back(miss);
      tab_off[k] = miss;
      miss += tab_bytes[k];
    }
    if (miss) {
      DD_TRY(sp.arena(align256(miss) + 4096, &stage));
      DD_TRY(sp.host(miss, &hst));
      DD_TRY(fetch_to_host(src, n, off, stage, hst, miss, s));
    }
    for (size_t k = 0; k < M; k++) {
      const dd_header& hd = hdrs[k];
      const uint8_t* tb = !tab_bytes[k]             ? nullptr
                          : !dev_chain              ? host_chain[ids[k]].data + 32 + 4ull * hd.num_first_ocur
                          : have_tab[ids[k]]        ? tabs[ids[k]].data()
                                                    : hst + tab_off[k];
      std::vector<uint64_t> cnt((size_t)hd.chkpt_id + 1, 0);
      for (uint32_t i = 0; i < hd.num_prior_chkpts; i++) {
        const uint32_t ck = rd32(tb + 8ull * i), n = rd32(tb + 8ull * i + 4);
        if (ck > hd.chkpt_id) return DD_ECORRUPT;
        cnt[ck] = n;
      }
      uint64_t cum = 0;
      for (uint32_t t = 0; t <= hd.chkpt_id; t++) {
        if (!cnt[t]) continue;
        cum += cnt[t];
        if (cum > hd.num_shift_dupl) return DD_ECORRUPT;
        pb_end[k].push_back((uint32_t)cum);
        pb_ck[k].push_back(t);
      }
      if (cum != hd.num_shift_dupl) return DD_ECORRUPT;
      pb_total += align256(8ull * pb_end[k].size() + 8);
    }
  }
  // Arena: [upload block: counters, descriptors, lists, list prefixes, prior
  // boundaries][block sums][f_off and s_off of every checkpoint][host diff copies].
  const uint64_t up_desc = sizeof(RestartCtr);
  const uint64_t up_lists = up_desc + align256(M * sizeof(RestartCkpt));
  const uint64_t up_lblk = up_lists + align256(2 * M * sizeof(RestartList));
  const uint64_t up_pb = up_lblk + align256((2 * M + 1) * 4);
  const uint64_t upload = up_pb + pb_total;
  const uint64_t blk_at = align256(upload);
  uint64_t off_at = blk_at + align256(8 * nblk + 8);
  uint64_t end = off_at;
  const uint64_t idx_bytes = align256(4 * ceil_div(C, 256));
  for (size_t k = 0; k < M; k++)
    end += align256(4ull * (hdrs[k].num_first_ocur + 1)) + align256(4ull * (hdrs[k].num_shift_dupl + 1)) +
           2 * idx_bytes;
  const uint64_t copies_at = end;
  if (host_chain)
    for (size_t k = 0; k < M; k++) end += align256(total_of(ids[k]));
  uint8_t* ar = nullptr;
  DD_TRY(sp.arena(end, &ar));
  uint8_t* hst = nullptr;
  DD_TRY(sp.host(upload, &hst));
  const double c0 = now_s();
  // 3. Diff bytes on the device (host diffs are copied into the arena).
  std::vector<const uint8_t*> diff(M);
  {
    uint64_t off = copies_at;
    for (size_t k = 0; k < M; k++) {
      if (dev_chain) {
        diff[k] = static_cast<const uint8_t*>(dev_chain[ids[k]]);
      } else {
        DD_HIP(hipMemcpyAsync(ar + off, host_chain[ids[k]].data, total_of(ids[k]), hipMemcpyHostToDevice, s));
        diff[k] = ar + off;
        off += align256(total_of(ids[k]));
      }
    }
  }
  // Copy width: 16-byte pieces need 16-byte chunks and output and 4-byte
  // aligned data sections (16-byte loads at 4-byte alignment).
  bool src4 = true;
  for (size_t k = 0; k < M; k++) src4 = src4 && ((reinterpret_cast<uintptr_t>(diff[k]) + 32 + meta_bytes[k]) & 3) == 0;
  const uintptr_t op = reinterpret_cast<uintptr_t>(dev_out);
  const int piece = (src4 && cs % 16 == 0 && op % 16 == 0) ? 16 : (src4 && cs % 4 == 0 && op % 4 == 0) ? 4 : 1;
  // A diff with one first-occurrence region and no shifted one takes the
  // whole-region step.
  auto whole_step = [&](size_t k) { return piece == 16 && hdrs[k].num_first_ocur == 1 && hdrs[k].num_shift_dupl == 0; };
  // 4. Upload block.
  memset(hst, 0, up_lists);
  RestartCkpt* desc = reinterpret_cast<RestartCkpt*>(hst + up_desc);
  {
    uint64_t pb = up_pb;
    for (size_t k = 0; k < M; k++) {
      const dd_header& hd = hdrs[k];
      RestartCkpt& rc = desc[k];
      rc.diff = diff[k];
      rc.f_off = reinterpret_cast<uint32_t*>(ar + off_at);
      off_at += align256(4ull * (hd.num_first_ocur + 1));
      rc.s_off = reinterpret_cast<uint32_t*>(ar + off_at);
      off_at += align256(4ull * (hd.num_shift_dupl + 1));
      rc.f_idx = reinterpret_cast<uint32_t*>(ar + off_at);
      off_at += idx_bytes;
      rc.s_idx = reinterpret_cast<uint32_t*>(ar + off_at);
      off_at += idx_bytes;
      const size_t np = pb_end[k].size();
      if (np) {
        memcpy(hst + pb, pb_end[k].data(), 4 * np);
        memcpy(hst + pb + 4 * np, pb_ck[k].data(), 4 * np);
      }
      rc.pb_end = reinterpret_cast<const uint32_t*>(ar + pb);
      rc.pb_ck = reinterpret_cast<const uint32_t*>(ar + pb + 4 * np);
      rc.npb = (uint32_t)np;
      pb += align256(8ull * np + 8);
      rc.nfirst = hd.num_first_ocur;
      rc.nshift = hd.num_shift_dupl;
      rc.ckpt_id = hd.chkpt_id;
      rc.node_bias = bias;
      // The walk's last step takes k_restart_whole without cover writes.
      rc.no_cover = (k + 1 == M && whole_step(k)) ? 1u : 0u;
      rc.first_off = 32;
      rc.shift_off = 32 + 4ull * hd.num_first_ocur + 8ull * hd.num_prior_chkpts;
      rc.data_off = 32 + meta_bytes[k];
      rc.data_bytes = total_of(ids[k]) - 32 - meta_bytes[k];
    }
  }
  memcpy(hst + up_lists, lists.data(), lists.size() * sizeof(RestartList));
  memcpy(hst + up_lblk, list_blk.data(), list_blk.size() * 4);
  DD_HIP(hipMemcpyAsync(ar, hst, upload, hipMemcpyHostToDevice, s));
  RestartCtr* ctr = reinterpret_cast<RestartCtr*>(ar);
  RestartCkpt* ddesc = reinterpret_cast<RestartCkpt*>(ar + up_desc);
  launch_restart_parse(ddesc, reinterpret_cast<const RestartList*>(ar + up_lists), (uint32_t)(2 * M),
                       reinterpret_cast<const uint32_t*>(ar + up_lblk), (uint32_t)nblk,
                       reinterpret_cast<uint32_t*>(ar + blk_at), g, cs, ctr, s);
  // 5. The walk, newest first: two launches per checkpoint.
  if ((uint64_t)*stamp + M + 1 > 0xfffffff0ull) {  // stamps would wrap: start over from a cleared table
    launch_fill_u32(reinterpret_cast<uint32_t*>(cover), 2 * C, 0, s);
    *stamp = 0;
  }
  const uint32_t base = *stamp;
  uint8_t* out = static_cast<uint8_t*>(dev_out);
  bool any_shift = false;  // the pending list stays empty until a checkpoint has shifted entries
  for (size_t k = 0; k < M; k++) {
    const uint32_t st = base + 1 + (uint32_t)k;
    if (whole_step(k))
      launch_restart_whole(ddesc + k, g, cover, base, st, ctr, out, len, cs, s);
    else
      launch_restart_regions(ddesc + k, g, cover, base, st, ref, pend, ctr, out, len, cs, piece, s);
    any_shift = any_shift || hdrs[k].num_shift_dupl > 0;
    if (any_shift) launch_restart_pending(ddesc + k, g, cover, st, ref, pend, ctr, out, len, cs, piece, s);
  }
  *stamp = base + (uint32_t)M;
  DD_HIP(hipGetLastError());
  RestartCtr res;
  DD_HIP(hipMemcpyAsync(hst, ctr, sizeof(RestartCtr), hipMemcpyDeviceToHost, s));
  DD_HIP(hipStreamSynchronize(s));
  memcpy(&res, hst, sizeof res);
  timers[0] = c0 - t0;
  timers[1] = now_s() - t0;
  if (res.flags & FLAG_BAD_DIFF) return DD_ECORRUPT;
  const uint64_t unresolved = (C - std::min<uint64_t>(C, res.claimed)) + (res.npend - res.ndone);
  if (unresolved || (res.flags & FLAG_BAD_SOURCE)) return DD_ECORRUPT;
  return DD_OK;
}

}  // namespace

extern "C" {

dd_status dd_create(uint32_t chunk_size, int hash_kind, void* hip_stream, dd_dedup** out) {
  return dd_create_method(chunk_size, hash_kind, DD_METHOD_TREE, hip_stream, out);
}

dd_status dd_create_method(uint32_t chunk_size, int hash_kind, int method, void* hip_stream, dd_dedup** out) {
  if (!out) return DD_EINVAL;
  *out = nullptr;
  if (chunk_size == 0 || (hash_kind != DD_HASH_MURMUR3 && hash_kind != DD_HASH_MD5)) return DD_EINVAL;
  if (method != DD_METHOD_TREE && method != DD_METHOD_LIST && method != DD_METHOD_TREE_LOW_ROOT) return DD_EINVAL;
  dd_dedup* h = new (std::nothrow) dd_dedup;
  if (!h) return DD_ENOMEM;
  h->st.method = method;
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    return DD_EHIP;
  }
  h->st.chunk_size = chunk_size;
  h->st.hash_kind = hash_kind;
  if (hip_stream) {
    h->stream = reinterpret_cast<hipStream_t>(hip_stream);
  } else {
    if (hipStreamCreate(&h->stream) != hipSuccess) {
      delete h;
      return DD_EHIP;
    }
    h->own_stream = true;
  }
  // Timing-only events: without the system-scope fence of a default event,
  // a record in the middle of the pipeline does not write back and
  // invalidate the caches (which cost ~65 us per checkpoint at 1 GiB).  They
  // are read after the stream has been synchronized.
  for (auto& e : h->ev) {
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) {
      dd_destroy(h);
      return DD_EHIP;
    }
  }
  for (auto& e : h->pev) {
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) {
      dd_destroy(h);
      return DD_EHIP;
    }
  }
  *out = h;
  return DD_OK;
}

dd_status dd_set_options(dd_dedup* h, uint32_t options) {
  if (!h || (options & ~(uint32_t)(DD_OPT_EMIT_ROOT_SHIFT | DD_OPT_DENSE_BASELINE))) return DD_EINVAL;
  h->st.options = options;
  return DD_OK;
}

dd_status dd_get_options(dd_dedup* h, uint32_t* options) {
  if (!h || !options) return DD_EINVAL;
  *options = h->st.options;
  return DD_OK;
}

dd_status dd_get_method(dd_dedup* h, int* method) {
  if (!h || !method) return DD_EINVAL;
  *method = h->st.method;
  return DD_OK;
}

dd_status dd_destroy(dd_dedup* h) {
  if (!h) return DD_EINVAL;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (auto& e : h->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : h->pev)
    if (e) (void)hipEventDestroy(e);
  if (h->rs_host) (void)hipHostFree(h->rs_host);
  if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return DD_OK;
}

dd_status dd_checkpoint(dd_dedup* h, const void* dev_data, uint64_t len, int make_baseline, dd_buf* out) {
  DD_TRY(checkpoint_to_host(h, nullptr, dev_data, len, make_baseline, out));
  h->st.current_id += 1;
  return DD_OK;
}

dd_status dd_checkpoint_header(dd_dedup* h, dd_header* header, const void* dev_data, uint64_t len, int make_baseline,
                               dd_buf* out) {
  if (!header) return DD_EINVAL;
  return checkpoint_to_host(h, header, dev_data, len, make_baseline, out);
}

dd_status dd_checkpoint_device(dd_dedup* h, const void* dev_data, uint64_t len, int make_baseline, void* dev_out,
                               uint64_t out_cap, uint64_t* out_len) {
  if (!h || !dev_data || !dev_out || !out_len) return DD_EINVAL;
  JobCounters c;
  DD_TRY(run_checkpoint(h, dev_data, len, make_baseline, reinterpret_cast<uint8_t*>(dev_out), out_cap, &c));
  h->st.timers[3] = 0;
  *out_len = c.diff_len;
  h->st.current_id += 1;
  return DD_OK;
}

uint64_t dd_max_diff_size(uint64_t len, uint32_t chunk_size, uint32_t ckpt_id) {
  return max_diff_size(len, chunk_size, ckpt_id);
}

dd_status dd_counts(dd_dedup* h, uint64_t* nf, uint64_t* ns) {
  if (!h) return DD_EINVAL;
  if (nf) *nf = h->st.nfirst;
  if (ns) *ns = h->st.nshift;
  return DD_OK;
}

dd_status dd_get_ids(dd_dedup* h, uint32_t* cur, uint32_t* base) {
  if (!h) return DD_EINVAL;
  if (cur) *cur = h->st.current_id;
  if (base) *base = h->st.baseline_id;
  return DD_OK;
}

dd_status dd_set_current_id(dd_dedup* h, uint32_t id) {
  if (!h) return DD_EINVAL;
  h->st.current_id = id;
  return DD_OK;
}

dd_status dd_get_timers(dd_dedup* h, double t[4], double rt[2]) {
  if (!h) return DD_EINVAL;
  if (t) memcpy(t, h->st.timers, sizeof h->st.timers);
  if (rt) memcpy(rt, h->st.restart_timers, sizeof h->st.restart_timers);
  return DD_OK;
}

dd_status dd_restart(dd_dedup* h, const dd_buf* chain, uint32_t nchain, uint32_t chkpt_id, void* dev_out,
                     uint64_t len) {
  if (!h || !chain || !dev_out) return DD_EINVAL;
  if (hipSetDevice(h->device) != hipSuccess) return DD_EHIP;
  HandleSpace sp(h);
  return restart_run(sp, h->stream, h->st.method, chain, nullptr, nullptr, nchain, chkpt_id, dev_out, len,
                     h->st.restart_timers);
}

dd_status dd_restart_device(dd_dedup* h, const void* const* dev_chain, const uint64_t* lens, uint32_t nchain,
                            uint32_t chkpt_id, void* dev_out, uint64_t len) {
  if (!h || !dev_chain || !lens || !dev_out) return DD_EINVAL;
  if (hipSetDevice(h->device) != hipSuccess) return DD_EHIP;
  HandleSpace sp(h);
  return restart_run(sp, h->stream, h->st.method, nullptr, dev_chain, lens, nchain, chkpt_id, dev_out, len,
                     h->st.restart_timers);
}

dd_status dd_restart_device_noalloc(dd_dedup* h, const void* const* dev_chain, const uint64_t* lens,
                                    uint32_t nchain, uint32_t chkpt_id, void* dev_out, uint64_t len) {
  if (!h || !dev_chain || !lens || !dev_out) return DD_EINVAL;
  if (h->device < 0 || h->device >= kNaDevices) return DD_ELIMIT;
  if (hipSetDevice(h->device) != hipSuccess) return DD_EHIP;
  NoAllocDevice& d = na_dev[h->device];
  std::lock_guard<std::mutex> lk(d.mu);
  if (!d.ptr) {
    // Set-up, once per device with the device current: the array's device
    // address (the symbol's host-side name is not one), then a cleared
    // cover table.
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return DD_EHIP;
    }
    launch_fill_u32(reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(p) + kNaCover), 2 * DD_NOALLOC_MAX_CHUNKS, 0,
                    h->stream);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) return DD_EHIP;
    d.ptr = static_cast<uint8_t*>(p);
    d.stamp = 0;
  }
  NoAllocSpace sp(d);
  return restart_run(sp, h->stream, h->st.method, nullptr, dev_chain, lens, nchain, chkpt_id, dev_out, len,
                     h->st.restart_timers);
}

dd_status dd_batch_create(dd_dedup* const* handles, uint32_t n, dd_batch** out) {
  if (!handles || n == 0 || !out) return DD_EINVAL;
  *out = nullptr;
  for (uint32_t i = 0; i < n; i++) {
    if (!handles[i]) return DD_EINVAL;
    if (handles[i]->device != handles[0]->device || handles[i]->st.hash_kind != handles[0]->st.hash_kind ||
        handles[i]->st.method != handles[0]->st.method)
      return DD_EINVAL;
    for (uint32_t k = 0; k < i; k++)
      if (handles[k] == handles[i]) return DD_EINVAL;
  }
  dd_batch* b = new (std::nothrow) dd_batch;
  if (!b) return DD_ENOMEM;
  b->device = handles[0]->device;
  b->hs.assign(handles, handles + n);
  if (hipSetDevice(b->device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess) {
    delete b;
    return DD_EHIP;
  }
  for (auto& e : b->pev) {
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) {
      dd_batch_destroy(b);
      return DD_EHIP;
    }
  }
  *out = b;
  return DD_OK;
}

dd_status dd_batch_destroy(dd_batch* b) {
  if (!b) return DD_EINVAL;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto& e : b->pev)
    if (e) (void)hipEventDestroy(e);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return DD_OK;
}

dd_status dd_batch_use_graph(dd_batch* b, int enable) {
  if (!b) return DD_EINVAL;
  b->use_graph = enable != 0;
  return DD_OK;
}

dd_status dd_batch_checkpoint_device(dd_batch* b, const void* const* dev_data, const uint64_t* lens,
                                     const int* make_baseline, void* const* dev_out, const uint64_t* out_caps,
                                     uint64_t* out_lens) {
  if (!b || !dev_data || !lens || !make_baseline || !dev_out || !out_caps || !out_lens) return DD_EINVAL;
  DD_HIP(hipSetDevice(b->device));
  const size_t n = b->hs.size();
  std::vector<JobDev> jobs(n);
  std::vector<char> base(n);
  bool begin = false;  // k_begin for every job unless all counters are armed
  // Every job is checked before any handle's state changes, so a rejected
  // call leaves all of them as they were.
  for (size_t i = 0; i < n; i++) {
    if (!dev_data[i] || !dev_out[i]) return DD_EINVAL;
    StreamState& st = b->hs[i]->st;
    base[i] = st.runs_baseline(make_baseline[i]);
    DD_TRY(st.validate(lens[i], base[i], static_cast<const uint8_t*>(dev_out[i]), out_caps[i]));
  }
  for (size_t i = 0; i < n; i++) {
    StreamState& st = b->hs[i]->st;
    begin = begin || base[i] || !st.ctr_armed;
    DD_TRY(st.prepare(static_cast<const uint8_t*>(dev_data[i]), lens[i], base[i], static_cast<uint8_t*>(dev_out[i]),
                      out_caps[i], jobs[i], b->stream));
  }
  for (size_t i = 0; i < n; i++) b->hs[i]->st.ctr_armed = false;
  // Probe-round width over the whole table: wide rounds suit a few inserts
  // (a wavefront waits for its longest chain), narrow rounds many (bound by
  // random-line traffic, more waves in flight).  The 64-buffer workload, 1.6M
  // inserts per call, runs 3.6 % faster with narrow rounds for every job.
  if (!std::getenv("DD_INSERT_ROUNDS")) {
    uint64_t expect = 0;
    for (size_t i = 0; i < n; i++) {
      const StreamState& st = b->hs[i]->st;
      if (!base[i] && st.last_pending != ~0ull) expect += st.last_pending;
    }
    if (expect > kWideInsertLimit)
      for (JobDev& J : jobs) J.wide_probe = 0;
  }
  DD_TRY(b->pipe.upload(jobs, b->stream));
  b->pipe.profile_level_ = b->profile;
  hipEvent_t* pev = b->profile ? b->pev : nullptr;
  // The counter reset and the leaf kernel are launched directly (events
  // attach to the leaf kernel's dispatch: events recorded inside a graph
  // cannot be timed); the stages after it are replayed from the graph.
  // Profiling level 2 records events between all stages and runs without
  // the graph.
  if (b->use_graph && b->profile < 2) {
    std::vector<uint64_t> sig = b->pipe.signature();
    DD_TRY(b->pipe.launch_all(b->stream, false, begin, nullptr, pev, Pipeline::PART_HEAD));
    if (!b->exec || sig != b->sig) {
      if (b->exec) (void)hipGraphExecDestroy(b->exec);
      if (b->graph) (void)hipGraphDestroy(b->graph);
      b->exec = nullptr;
      b->graph = nullptr;
      DD_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      const dd_status rc = b->pipe.launch_all(b->stream, false, begin, nullptr, nullptr, Pipeline::PART_TAIL);
      const hipError_t ec = hipStreamEndCapture(b->stream, &b->graph);
      if (rc != DD_OK) return rc;
      if (ec != hipSuccess) return DD_EHIP;
      DD_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->sig = sig;
    }
    DD_HIP(hipGraphLaunch(b->exec, b->stream));
    DD_TRY(b->pipe.trial_mark_end(b->stream));
  } else {
    DD_TRY(b->pipe.launch_all(b->stream, false, begin, nullptr, pev));
  }
  DD_HIP(hipStreamSynchronize(b->stream));
  b->pipe.trial_done();
  if (b->profile) {
    for (int i = 0; i < DD_NUM_STAGES; i++) {
      float ms = 0;
      if (b->profile >= 2 || i == 0) (void)hipEventElapsedTime(&ms, b->pev[i], b->pev[i + 1]);
      b->stage_ms[i] = ms;
    }
  }
  std::vector<JobCounters> c(b->pipe.host_ctr, b->pipe.host_ctr + n);
  for (size_t i = 0; i < n; i++) {
    if (jobs[i].bulk && (c[i].flags & FLAG_BULK_DUPS))
      DD_TRY(rerun_bulk_dedup(b->hs[i]->st, jobs[i], b->pipe_rerun, b->stream, &c[i]));
    if (jobs[i].bulk && (c[i].flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)))
      DD_TRY(rerun_dense_baseline(b->hs[i]->st, jobs[i], b->pipe_rerun, b->stream, &c[i]));
  }
  for (int attempt = 0; attempt < 4; attempt++) {
    std::vector<size_t> redo;
    for (size_t i = 0; i < n; i++)
      if (c[i].flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) redo.push_back(i);
    if (redo.empty()) break;
    std::vector<JobDev> sub;
    for (size_t i : redo) {
      StreamState& st = b->hs[i]->st;
      DD_TRY(st.grow_map(c[i].map_count + 2 * c[i].pending, b->stream));
      jobs[i].map = make_view(st.map.p, st.map_cap);
      jobs[i].map_count_at_start = c[i].map_count;
      jobs[i].map_limit = map_limit_of(st.map_cap);
      sub.push_back(jobs[i]);
    }
    DD_TRY(b->pipe_rerun.run(sub, b->stream, true, true));
    DD_HIP(hipStreamSynchronize(b->stream));
    for (size_t k = 0; k < redo.size(); k++) c[redo[k]] = b->pipe_rerun.host_ctr[k];
  }
  for (size_t i = 0; i < n; i++) {
    if (c[i].flags & (FLAG_NEED_GROW | FLAG_MAP_FULL)) return DD_ENOMEM;
    if (c[i].flags & FLAG_BAD_SOURCE) return DD_ECORRUPT;
  }
  for (size_t i = 0; i < n; i++) {
    StreamState& st = b->hs[i]->st;
    st.commit(c[i], base[i], b->pipe.job_copied(i));
    st.ctr_armed = true;
    out_lens[i] = c[i].diff_len;
    st.current_id += 1;
  }
  return DD_OK;
}

dd_status dd_get_leaf_mapping(dd_dedup* h, int* mapping) {
  if (!h || !mapping) return DD_EINVAL;
  *mapping = h->pipe.leaf_mapping();
  return DD_OK;
}

dd_status dd_get_stage_budget(dd_dedup* h, uint64_t* workgroups) {
  if (!h || !workgroups) return DD_EINVAL;
  *workgroups = h->pipe.stage_budget();
  return DD_OK;
}

dd_status dd_batch_get_stage_budget(dd_batch* b, uint64_t* workgroups) {
  if (!b || !workgroups) return DD_EINVAL;
  *workgroups = b->pipe.stage_budget();
  return DD_OK;
}

dd_status dd_batch_get_leaf_mapping(dd_batch* b, int* mapping) {
  if (!b || !mapping) return DD_EINVAL;
  *mapping = b->pipe.leaf_mapping();
  return DD_OK;
}

dd_status dd_set_profiling(dd_dedup* h, int enable) {
  if (!h) return DD_EINVAL;
  h->profile = enable < 0 ? 0 : (enable > 2 ? 2 : enable);
  return DD_OK;
}

dd_status dd_get_stage_times(dd_dedup* h, double ms[DD_NUM_STAGES]) {
  if (!h || !ms) return DD_EINVAL;
  memcpy(ms, h->stage_ms, sizeof h->stage_ms);
  return DD_OK;
}

dd_status dd_get_ckpt_stats(dd_dedup* h, dd_ckpt_stats* out) {
  if (!h || !out) return DD_EINVAL;
  *out = h->st.stats;
  return DD_OK;
}

dd_status dd_get_device_bytes(dd_dedup* h, uint64_t* bytes) {
  if (!h || !bytes) return DD_EINVAL;
  *bytes = resident_bytes(h);
  return DD_OK;
}

dd_status dd_batch_set_profiling(dd_batch* b, int level) {
  if (!b) return DD_EINVAL;
  b->profile = level < 0 ? 0 : (level > 2 ? 2 : level);
  return DD_OK;
}

dd_status dd_batch_get_stage_times(dd_batch* b, double ms[DD_NUM_STAGES]) {
  if (!b || !ms) return DD_EINVAL;
  memcpy(ms, b->stage_ms, sizeof b->stage_ms);
  return DD_OK;
}

void dd_buf_free(dd_buf* b) {
  if (!b) return;
  std::free(b->data);
  b->data = nullptr;
  b->len = 0;
}

const char* dd_strerror(int s) {
  switch (s) {
    case DD_OK: return "ok";
    case DD_EINVAL: return "invalid argument";
    case DD_ENOMEM: return "out of memory";
    case DD_ECORRUPT: return "corrupt or incomplete checkpoint chain";
    case DD_EHIP: return "HIP runtime error";
    case DD_ESTATE: return "call violates the checkpoint chain";
    case DD_ELIMIT: return "request exceeds the allocation-free scratch";
    default: return "unknown status";
  }
}

}  // extern "C"
