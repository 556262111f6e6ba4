#!/usr/bin/env python3
"""Throughput of tree-approach incremental checkpointing on one MI355X.

Headline workload (`value`): the batched form.  64 checkpoint buffers (one per
simulated rank) with sizes log-uniform in [4 MiB, 512 MiB] (7.6 GiB in all),
128-byte chunks, Murmur3-x64-128, a 10-checkpoint chain each; every later
checkpoint overwrites r/3 of the chunks with fresh bytes and 2r/3 with copies
of other chunks (r uniform in [0.1 %, 5 %] per buffer) and zero-fills 1 MiB.
All chains are generated in HBM before timing.  A step is one batched call
(dd_batch_checkpoint_device, replayed from a captured hipGraph): one
checkpoint of all 64 buffers, inputs resident in HBM, diffs assembled in HBM;
steps walk the chains in order, so every 10th step is a baseline.  `value` =
bytes deduplicated / wall time over the timed steps.  `plain_gbs` is the same
chains checkpointed by 64 sequential dd_checkpoint_device calls per step.

Bit-exactness is checked in the same run, untimed: every batched diff equals
the plain path's, the diffs of a sample of buffers equal the OpenMP oracle's
(that sample is also the CPU baseline's workload), and restarts of the last
checkpoint of those buffers return their data.

`roofline` covers the dominant kernel, the leaf hash (every byte of every
buffer is read by it), timed with HIP events recorded on the batch's stream
around it during the timed steps.  Algorithmic bytes per launch: for each
buffer, C * (chunk + 16) (chunk and previous leaf digest read) plus 17 bytes
(new digest and label) per changed leaf, or C * (chunk + 33) at a baseline.
`whole_step` gives the fraction of the whole checkpoint step by
B = N_in + 32 C + 2 D + 32 + M per buffer (D first-occurrence data bytes, M
metadata bytes).

Other configurations measured beside it (each with its own object): the plain
1 GiB chain (`plain`, with the host-output rate that includes the PCIe copy),
its restart (`restart`), the chunk x change-rate sweep (`sweep`), the 4 GiB
MD5 float field (`field_md5`), the list method (`list`), the request loop
(`serving`, p50/p99), and the 64 MiB CPU reference path (`c1_cpu_reference`).
"""
import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "input GB/s deduped (batched vs plain), bit-exact; p99 latency in serving loop"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed batched steps (one checkpoint of every buffer)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--batch-buffers", type=int, default=64)
    ap.add_argument("--nckpt", type=int, default=10)
    ap.add_argument("--chunk", type=int, default=128)
    ap.add_argument("--size", type=int, default=1 << 30, help="buffer of the plain chain")
    ap.add_argument("--frac", type=float, default=0.01, help="change fraction of the plain chain")
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--only", choices=["batched", "plain", "restart", "sweep", "field", "list", "serve"],
                    help="run one section only (for profiling)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-field", action="store_true")
    ap.add_argument("--no-list", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--field-values", type=int, default=1 << 30)
    ap.add_argument("--cpu-sample-bytes", type=int, default=3 << 30,
                    help="batched-workload bytes (per checkpoint) the CPU baseline runs")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run 5 device-to-device copies of BYTES (scales PMC byte counters) and exit")
    ap.add_argument("--serve-ranks", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=200.0, help="Poisson arrivals, requests/s")
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-budget-gib", type=float, default=200.0)
    return ap.parse_args()


def log(obj):
    print(json.dumps(obj), file=sys.stderr, flush=True)


def traffic_from_profiles(kernel, workload):
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def leaf_bytes(st, chunk):
    """Algorithmic bytes of one leaf-kernel pass over a buffer (ckpt_stats).
    A baseline's leaf kernel that also wrote the diff's data section
    (leaf_copy) moved a chunk more per chunk."""
    C = st["chunks"]
    if st["baseline"]:
        return C * (chunk + 33) + (C * chunk if st.get("leaf_copy") else 0)
    return C * (chunk + 16) + 17 * st["changed"]


def ckpt_bytes(st, length, chunk):
    """B = N_in + 32 C + 2 D + 32 + M of one checkpoint (ckpt_stats)."""
    D = st["first_chunks"] * chunk
    M = st["diff_bytes"] - 32 - D
    return length + 32 * st["chunks"] + 2 * D + 32 + M


def roofline(achieved_gbs, traffic, kernel, alg_bytes, launch_ms):
    return {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel,
            "algorithmic_bytes_per_launch": int(alg_bytes), "avg_launch_ms": round(launch_ms, 4)}


def host_cpu():
    import platform

    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            cpu = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), cpu)
    except OSError:
        pass
    return cpu


# ---------------------------------------------------------------------------
# Batched form (headline)
# ---------------------------------------------------------------------------
def run_batched(torch, dd, W, O, args, seed=1931):
    import numpy as np

    nbuf, nckpt, chunk = args.batch_buffers, args.nckpt, args.chunk
    sizes = W.batch_sizes(seed, nbuf)
    rates = np.random.default_rng(seed).uniform(0.001, 0.05, nbuf)
    chains = [W.chain_uniform_changes_gpu(torch, seed + 1 + i, sz, chunk, nckpt, r / 3, 2 * r / 3)
              for i, (sz, r) in enumerate(zip(sizes, rates))]
    torch.cuda.synchronize()
    caps = [dd.max_diff_size(sz, chunk, 4 * nckpt + args.steps + args.warmup) for sz in sizes]
    outs_p = [torch.empty(c, dtype=torch.uint8, device="cuda") for c in caps]
    outs_b = [torch.empty(c, dtype=torch.uint8, device="cuda") for c in caps]
    plain = [dd.TreeDeduplicator(chunk) for _ in sizes]
    hs = [dd.TreeDeduplicator(chunk) for _ in sizes]
    batch = dd.Batch(hs, graph=True)
    ptrs = [[c[k].data_ptr() for c in chains] for k in range(nckpt)]
    # the CPU baseline's sample: buffers in index order up to the sample size
    sample, acc = [], 0
    for i, sz in enumerate(sizes):
        if acc >= args.cpu_sample_bytes or args.no_cpu:
            break
        sample.append(i)
        acc += sz

    def plain_step(k):
        return [g.checkpoint_device(ptrs[k][i], sizes[i], outs_p[i].data_ptr(), caps[i], k == 0)
                for i, g in enumerate(plain)]

    def batch_step(k):
        return batch.checkpoint_device(ptrs[k], sizes, [k == 0] * nbuf, [o.data_ptr() for o in outs_b], caps)

    # Untimed chain: both paths, every diff compared byte for byte; per
    # checkpoint index the byte counts of the roofline; the sample's diffs kept.
    mismatches = 0
    alg_leaf, alg_step = [0] * nckpt, [0] * nckpt
    kept = {i: [] for i in sample}
    for k in range(nckpt):
        lp = plain_step(k)
        lb = batch_step(k)
        torch.cuda.synchronize()
        for i in range(nbuf):
            if lp[i] != lb[i] or not torch.equal(outs_p[i][:lp[i]], outs_b[i][:lb[i]]):
                mismatches += 1
            st = hs[i].ckpt_stats()
            alg_leaf[k] += leaf_bytes(st, chunk)
            alg_step[k] += ckpt_bytes(st, sizes[i], chunk)
            if i in kept:
                kept[i].append(outs_b[i][:lb[i]].clone())
    log({"batched_verify": {"mismatches": mismatches}})
    # Timed: warmup, then `steps` batched calls with the leaf kernel's events.
    # The batch handles continue their chains (checkpoint index = step % nckpt).
    batch.set_profiling(1)
    gc.collect()  # the cyclic collector is off while timing (main)
    step = nckpt
    for _ in range(args.warmup):
        batch_step(step % nckpt)
        step += 1
    torch.cuda.synchronize()
    leaf_ms, idx = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        k = step % nckpt
        batch_step(k)
        leaf_ms.append(batch.stage_times()["leaf_hash"])
        idx.append(k)
        step += 1
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    batch.set_profiling(0)
    # Plain path: 64 sequential calls per step, same number of steps.
    gc.collect()
    pstep = nckpt
    for _ in range(min(args.warmup, nckpt)):
        plain_step(pstep % nckpt)
        pstep += 1
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        plain_step(pstep % nckpt)
        pstep += 1
    torch.cuda.synchronize()
    el_plain = time.perf_counter() - t1
    total = sum(sizes)
    res = dict(buffers=nbuf, total_bytes=total, checkpoints=nckpt, chunk_size=chunk, steps=args.steps,
               gbs=total * args.steps / el / 1e9, ms_per_step=el * 1e3 / args.steps,
               plain_gbs=total * args.steps / el_plain / 1e9, plain_ms_per_step=el_plain * 1e3 / args.steps,
               mismatches=mismatches, leaf_mapping=batch.leaf_mapping(), stage_budget=batch.stage_budget())
    lm = sum(leaf_ms) / len(leaf_ms)
    ab = sum(alg_leaf[k] for k in idx) / len(idx)
    res["leaf_ms"], res["leaf_alg_bytes"] = lm, ab
    res["leaf_gbs"] = ab / (lm * 1e-3) / 1e9
    res["step_alg_bytes"] = sum(alg_step[k] for k in idx) / len(idx)
    res["step_gbs"] = res["step_alg_bytes"] / (el / args.steps) / 1e9
    # Oracle and restart checks of the sample (and the CPU baseline's timing).
    oracle_bad, restart_bad = 0, 0
    cpu_time, cpu_bytes = 0.0, 0
    for i in sample:
        host = [chains[i][k].cpu().numpy() for k in range(nckpt)]
        o = O.ParallelOracleDeduplicator(chunk)
        tc = time.perf_counter()
        want = [o.checkpoint_array(x, k == 0) for k, x in enumerate(host)]
        cpu_time += time.perf_counter() - tc
        cpu_bytes += nckpt * sizes[i]
        for k in range(nckpt):
            got = kept[i][k].cpu().numpy()
            if got.size != len(want[k]) or not np.array_equal(got, np.frombuffer(want[k], np.uint8)):
                oracle_bad += 1
        out = torch.empty(sizes[i], dtype=torch.uint8, device="cuda")
        hs[i].restart_device([d.data_ptr() for d in kept[i]], [d.numel() for d in kept[i]], nckpt - 1,
                             out.data_ptr(), sizes[i])
        torch.cuda.synchronize()
        restart_bad += 0 if torch.equal(out, chains[i][nckpt - 1]) else 1
        del host, want, out
    res.update(oracle_sample_buffers=len(sample), oracle_sample_bytes=acc, oracle_mismatches=oracle_bad,
               restart_mismatches=restart_bad,
               bit_exact=mismatches == 0 and oracle_bad == 0 and restart_bad == 0)
    res["cpu"] = dict(gbs=cpu_bytes / cpu_time / 1e9 if cpu_time else None, seconds=cpu_time, bytes=cpu_bytes, buffers=len(sample),
                      threads=O.ParallelOracleDeduplicator.threads())
    del chains, outs_p, outs_b, plain, hs, batch, kept
    torch.cuda.empty_cache()
    return res


# ---------------------------------------------------------------------------
# Plain method, one buffer
# ---------------------------------------------------------------------------
def run_plain(torch, dd, W, O, size, chunk, frac, nckpt, steps, warmup, kind=0, cls=None, check=True,
              restart=False, d2h=False, seed=1931):
    """The plain 1 GiB chain: timed dd_checkpoint_device steps (leaf kernel
    timed with events attached to its dispatch), an untimed pass that checks
    every diff against the OpenMP oracle and collects the byte counts, and
    optionally the restart and host-output (PCIe) measurements."""
    import numpy as np

    chain = W.chain_runs_gpu(torch, seed, size, chunk, nckpt, frac)
    torch.cuda.synchronize()
    g = (cls or dd.TreeDeduplicator)(chunk, kind)
    cap = dd.max_diff_size(size, chunk, warmup + steps + 4 * nckpt + 1)
    out = torch.empty(cap, dtype=torch.uint8, device="cuda")
    res = {}
    # Untimed pass: diffs kept (for the oracle and the restart), byte counts.
    diffs, alg_leaf, alg_step = [], [], []
    for k in range(nckpt):
        n = g.checkpoint_device(chain[k].data_ptr(), size, out.data_ptr(), cap, k == 0)
        st = g.ckpt_stats()
        alg_leaf.append(leaf_bytes(st, chunk))
        alg_step.append(ckpt_bytes(st, size, chunk))
        diffs.append(out[:n].clone())
    torch.cuda.synchronize()
    g.set_profiling(1)
    gc.collect()
    i = nckpt
    for _ in range(warmup):
        g.checkpoint_device(chain[i % nckpt].data_ptr(), size, out.data_ptr(), cap, i % nckpt == 0)
        i += 1
    torch.cuda.synchronize()
    leaf, idx, nbytes = [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        k = i % nckpt
        nbytes.append(g.checkpoint_device(chain[k].data_ptr(), size, out.data_ptr(), cap, k == 0))
        leaf.append(g.stage_times()["leaf_hash"])
        idx.append(k)
        i += 1
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    # Stage breakdown of one more chain (events between stages: untimed).
    g.set_profiling(2)
    i += (nckpt - i % nckpt) % nckpt
    stages = []
    for _ in range(nckpt):
        k = i % nckpt
        g.checkpoint_device(chain[k].data_ptr(), size, out.data_ptr(), cap, k == 0)
        stages.append((k, g.stage_times()))
        i += 1
    g.set_profiling(0)
    lm = sum(leaf) / len(leaf)
    ab = sum(alg_leaf[k] for k in idx) / len(idx)
    sb = sum(alg_step[k] for k in idx) / len(idx)
    base = [sum(s.values()) for k, s in stages if k == 0]
    incr = [sum(s.values()) for k, s in stages if k != 0]
    res.update(gbs=steps * size / el / 1e9, ms_per_step=el * 1e3 / steps, leaf_ms=lm, leaf_alg_bytes=ab,
               leaf_gbs=ab / (lm * 1e-3) / 1e9, step_alg_bytes=sb, step_gbs=sb / (el / steps) / 1e9,
               stage_ms={key: sum(s[key] for _, s in stages) / len(stages) for key in stages[0][1]},
               gpu_ms_baseline=sum(base) / len(base) if base else None,
               gpu_ms_incremental=sum(incr) / len(incr) if incr else None,
               diff_bytes_avg=sum(nbytes) / len(nbytes), leaf_mapping=g.leaf_mapping())
    if check:
        o = (O.ParallelOracleDeduplicator if cls is None or cls is dd.TreeDeduplicator
             else O.OracleListDeduplicator)(chunk, kind)
        bad = 0
        for k in range(nckpt):
            want = o.checkpoint_array(chain[k].cpu().numpy(), k == 0)
            got = diffs[k].cpu().numpy()
            bad += 0 if got.size == len(want) and np.array_equal(got, np.frombuffer(want, np.uint8)) else 1
        res["oracle_mismatches"] = bad
    if restart:
        res["restart"] = measure_restart(torch, dd, g, chain, diffs, size)
    if d2h:
        # dd_checkpoint: the diff copied to host memory (the reference's
        # timers[3]), the same chain on a fresh handle.
        h2 = (cls or dd.TreeDeduplicator)(chunk, kind)
        for k in range(nckpt):
            h2.checkpoint(chain[k].data_ptr(), size, k == 0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        host_bytes = 0
        for j in range(steps):
            k = j % nckpt
            host_bytes += len(h2.checkpoint(chain[k].data_ptr(), size, k == 0))
        el2 = time.perf_counter() - t0
        res["d2h_inclusive"] = dict(gbs=round(steps * size / el2 / 1e9, 2), ms_per_step=round(el2 * 1e3 / steps, 4),
                                    diff_bytes_copied=host_bytes,
                                    note="dd_checkpoint: the same steps with the diff copied to pageable host "
                                         "memory inside the call (PCIe included); never the headline")
        del h2
    del chain, out, g, diffs
    torch.cuda.empty_cache()
    return res


def measure_restart(torch, dd, g, chain, diffs, size, reps=10):
    """Restart of the newest checkpoint (the whole chain walked) and of the
    baseline from device-resident diffs, each timed as whole calls (host work
    included); B_rst = sum of the walked diffs' bytes + datalen."""
    out = torch.empty(size, dtype=torch.uint8, device="cuda")
    res = {}
    ptrs = [d.data_ptr() for d in diffs]
    lens = [d.numel() for d in diffs]
    for name, k in (("newest", len(diffs) - 1), ("baseline", 0)):
        times = []
        ok = True
        for rep in range(reps + 1):
            out.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            g.restart_device(ptrs[:k + 1], lens[:k + 1], k, out.data_ptr(), size)
            torch.cuda.synchronize()
            if rep:
                times.append(time.perf_counter() - t0)
            else:
                ok = torch.equal(out, chain[k])
        t = sorted(times)[len(times) // 2]
        brst = sum(lens[:k + 1]) + size
        res[name] = dict(ckpt_id=k, ms_median=round(t * 1e3, 4), b_rst=brst,
                         achieved_gbs=round(brst / t / 1e9, 1), frac=round(brst / t / 1e9 / HBM_PEAK_GBS, 4),
                         byte_identical=ok)
    # the allocation-free entry point on the newest checkpoint
    times = []
    ok = True
    k = len(diffs) - 1
    for rep in range(reps + 1):
        out.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rc = g.restart_device_noalloc(ptrs, lens, k, out.data_ptr(), size)
        torch.cuda.synchronize()
        if rep:
            times.append(time.perf_counter() - t0)
        else:
            ok = rc == dd.DD_OK and torch.equal(out, chain[k])
    res["noalloc_newest_ms_median"] = round(sorted(times)[len(times) // 2] * 1e3, 4)
    res["noalloc_byte_identical"] = ok
    del out
    return res


# ---------------------------------------------------------------------------
# MD5 float field
# ---------------------------------------------------------------------------
def run_field(torch, dd, W, n_values, nckpt, chunk, seed=1931):
    """MD5 on a float32 simulation-like field: 2^30 values (4 GiB), 64-byte
    chunks, 20 checkpoints; 60 % of the 4 KiB blocks are exact zeros, each
    version perturbs 5 % of the nonzero values by N(0,1e-3) and flips 1 % of
    the blocks between zero and nonzero.  Each version is generated in HBM
    from the previous one (untimed); each checkpoint call is timed on its own
    (the call synchronizes), with HIP events around the leaf-hash kernel,
    after an untimed baseline of version 0 that allocates the handle's state
    and tunes the leaf mapping.
    Checkpoints 0, nckpt/2 and nckpt-1 are restarted from the device-resident
    chain and compared byte for byte with the saved versions."""
    size = n_values * 4
    g = dd.TreeDeduplicator(chunk, dd.HASH_MD5)
    g.set_profiling(1)
    cap = dd.max_diff_size(size, chunk, nckpt)
    out = torch.empty(cap, dtype=torch.uint8, device="cuda")
    keep_ids = sorted({0, nckpt // 2, nckpt - 1})
    saved, diffs, times, leaf, diff_bytes = {}, [], [], [], []
    for k, v in enumerate(W.float_field_versions_gpu(torch, seed, n_values, nckpt)):
        if k in keep_ids:
            saved[k] = v.clone()
        if k == 0:
            # Untimed warm-up baseline on the same handle: allocates its state
            # and runs the leaf-mapping tuner (once per process); the chain
            # then starts again at checkpoint id 0.
            g.checkpoint_device(v.data_ptr(), size, out.data_ptr(), cap, True)
            g.current_id = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n = g.checkpoint_device(v.data_ptr(), size, out.data_ptr(), cap, k == 0)
        times.append(time.perf_counter() - t0)
        leaf.append(g.stage_times()["leaf_hash"])
        diff_bytes.append(n)
        diffs.append(out[:n].clone())
    torch.cuda.synchronize()
    ok = True
    for k in keep_ids:
        o = torch.empty(size, dtype=torch.uint8, device="cuda")
        g.restart_device([d.data_ptr() for d in diffs[:k + 1]], [d.numel() for d in diffs[:k + 1]], k,
                         o.data_ptr(), size)
        ok = ok and torch.equal(o, saved[k])
    leaf_ms = sum(leaf) / len(leaf)
    res = dict(workload=f"field_{size >> 20}MiB_c{chunk}_md5", checkpoints=nckpt, buffer_bytes=size,
               chunk_size=chunk, gbs=round(nckpt * size / sum(times) / 1e9, 2),
               ms_per_checkpoint=round(sum(times) * 1e3 / nckpt, 3),
               ms_baseline=round(times[0] * 1e3, 3),
               ms_incremental_avg=round(sum(times[1:]) * 1e3 / max(1, nckpt - 1), 3),
               leaf_kernel_ms=round(leaf_ms, 4), leaf_kernel_input_gbs=round(size / (leaf_ms * 1e-3) / 1e9, 1),
               bound="valu (MD5: two compressions per 64-byte chunk)",
               diff_bytes_total=sum(diff_bytes), restarts_verified=keep_ids, restart_bit_exact=ok)
    del diffs, saved, out, g
    torch.cuda.empty_cache()
    return res


# ---------------------------------------------------------------------------
# Serving
# ---------------------------------------------------------------------------
def _pct(xs, q):
    import numpy as np

    return round(float(np.percentile(np.asarray(xs), q)) * 1e3, 3) if xs else None


def run_serving(torch, dd, W, ranks, rate, seconds, budget, slots, chunk, kind, seed=2024, versions=10):
    """Request loop over `ranks` simulated ranks (dd_server).  Each rank has a
    buffer of log-uniform size in [16 MiB, 1 GiB] and its own handle, whose
    tree and digest map stay in HBM between requests.  Checkpoint data comes
    from a pool of `versions` 1 GiB versions (the 64 MiB chain's recipe: 1 %
    fresh chunks, 2 % copied chunks, one zeroed 1 MiB region per 64 MiB); rank r
    reads version n % versions at its own offset, and re-baselines every
    `versions` checkpoints.  Arrivals are Poisson at `rate` req/s for
    `seconds` (85 % checkpoints, 15 % restarts of a random checkpoint of the
    rank's current chain); a second, saturated pass submits a fixed number of
    requests at once, and a third does the same with batching off (every
    request alone on its handle's stream).  Checkpoints admitted together run
    as one batched job table (at most 64).  Every restart is compared byte for
    byte with the data it restores.  Latency = completion - submission (steady
    clock of the library)."""
    import queue
    import threading

    import numpy as np

    rng = np.random.default_rng(seed)
    sizes = W.batch_sizes(seed, ranks, 16 << 20, 1 << 30, 4096)
    pool_size = max(sizes)
    pool = W.chain_uniform_changes_gpu(torch, seed, pool_size, chunk, versions, 0.01, 0.02, 1 << 20,
                                       zero_count=max(1, pool_size >> 26))
    offs = [int(rng.integers(0, (pool_size - sz) // 4096 + 1)) * 4096 for sz in sizes]
    streams = [torch.cuda.Stream() for _ in range(ranks)]
    handles = [dd.TreeDeduplicator(chunk, kind, stream=st.cuda_stream) for st in streams]
    kmax = 1 << 16
    caps = [dd.max_diff_size(sz, chunk, kmax) for sz in sizes]
    outs = [[torch.empty(caps[r], dtype=torch.uint8, device="cuda") for _ in range(versions)] for r in range(ranks)]
    torch.cuda.synchronize()
    srv = dd.Server(budget, slots)
    nck = [0] * ranks  # checkpoints submitted per rank (= next checkpoint id)

    def ckpt_args(r):
        n = nck[r]
        assert n < kmax
        nck[r] += 1
        v = n % versions
        return (handles[r], pool[v].data_ptr() + offs[r], sizes[r], outs[r][v].data_ptr(), caps[r], v == 0)

    def restart_args(r):
        last = nck[r] - 1
        base = last - last % versions
        j = int(rng.integers(base, last + 1))
        ptrs = [0] * (j + 1)
        lens = [0] * (j + 1)
        for i in range(base, j + 1):
            ptrs[i] = outs[r][i % versions].data_ptr()
            lens[i] = caps[r]
        o = torch.empty(sizes[r], dtype=torch.uint8, device="cuda")
        return j, o, (handles[r], ptrs, lens, j, o.data_ptr(), sizes[r])

    # Warm-up (untimed): the first checkpoint of every rank allocates its state.
    for t in [srv.submit_checkpoint(*ckpt_args(r)[:5], make_baseline=True) for r in range(ranks)]:
        assert srv.wait(t)["status"] == dd.DD_OK
    torch.cuda.synchronize()

    def run_pass(arrivals):
        """arrivals: list of (time offset s or None, rank, is_restart)."""
        done = queue.Queue()
        stats = dict(lat_ck=[], lat_rs=[], wait=[], groups=[], ck_bytes=0, rs_bytes=0, bad=0, errors=0, t0=None,
                     t1=0.0, error_status=[])

        def collector(nreq):
            for _ in range(nreq):
                item = done.get()
                res = srv.wait(item[0])
                if res["status"] != dd.DD_OK:
                    stats["errors"] += 1
                    stats["error_status"].append(dd.strerror(res["status"]))
                    continue
                lat = res["t_end"] - res["t_submit"]
                stats["wait"].append(res["t_start"] - res["t_submit"])
                stats["t1"] = max(stats["t1"], res["t_end"])
                if item[1] is None:
                    stats["lat_ck"].append(lat)
                    stats["ck_bytes"] += item[2]
                    stats["groups"].append(res["group"])
                else:
                    r, j, o = item[1], item[2], item[3]
                    v = j % versions
                    if not torch.equal(o, pool[v][offs[r]:offs[r] + sizes[r]]):
                        stats["bad"] += 1
                    stats["lat_rs"].append(lat)
                    stats["rs_bytes"] += sizes[r]

        th = threading.Thread(target=collector, args=(len(arrivals),))
        th.start()
        start = time.perf_counter()
        stats["t0"] = srv.now()
        for when, r, is_rs in arrivals:
            if when is not None:
                d = start + when - time.perf_counter()
                if d > 0:
                    time.sleep(d)
            if is_rs and nck[r] > 0:
                j, o, a = restart_args(r)
                done.put((srv.submit_restart(*a), r, j, o))
            else:
                a = ckpt_args(r)
                done.put((srv.submit_checkpoint(*a[:5], make_baseline=a[5]), None, sizes[r]))
        th.join()
        el = stats["t1"] - stats["t0"]
        lat_all = stats["lat_ck"] + stats["lat_rs"]
        return dict(
            requests=len(arrivals), checkpoints=len(stats["lat_ck"]), restarts=len(stats["lat_rs"]),
            elapsed_s=round(el, 3), req_per_s=round(len(lat_all) / el, 1),
            checkpoint_gbs=round(stats["ck_bytes"] / el / 1e9, 2),
            served_gbs=round((stats["ck_bytes"] + stats["rs_bytes"]) / el / 1e9, 2),
            p50_ms=_pct(lat_all, 50), p99_ms=_pct(lat_all, 99),
            checkpoint_p50_ms=_pct(stats["lat_ck"], 50), checkpoint_p99_ms=_pct(stats["lat_ck"], 99),
            restart_p50_ms=_pct(stats["lat_rs"], 50), restart_p99_ms=_pct(stats["lat_rs"], 99),
            queue_wait_p50_ms=_pct(stats["wait"], 50), queue_wait_p99_ms=_pct(stats["wait"], 99),
            restarts_verified=len(stats["lat_rs"]) - stats["bad"], restart_mismatches=stats["bad"],
            checkpoint_group_mean=round(float(np.mean(stats["groups"])), 2) if stats["groups"] else None,
            checkpoint_group_max=max(stats["groups"]) if stats["groups"] else None,
            errors=stats["errors"], error_status=sorted(set(stats["error_status"])))

    n = int(rate * seconds)
    gaps = rng.exponential(1.0 / rate, n)
    poisson = [(float(t), int(rng.integers(0, ranks)), bool(rng.random() < 0.15)) for t in np.cumsum(gaps)]
    out = dict(ranks=ranks, buffer_bytes_total=sum(sizes), buffer_bytes_min=min(sizes), buffer_bytes_max=max(sizes),
               chunk_size=chunk, slots=slots, hbm_budget_bytes=budget, rebaseline_every=versions)
    out["poisson"] = dict(rate_req_per_s=rate, **run_pass(poisson))
    log({"serving_poisson": out["poisson"]})
    nsat = 12 * ranks
    sat = [(None, int(rng.integers(0, ranks)), bool(rng.random() < 0.15)) for _ in range(nsat)]
    out["saturated"] = run_pass(sat)
    log({"serving_saturated": out["saturated"]})
    # The same saturated load with every request alone on its handle's stream.
    srv.set_batching(1)
    sat = [(None, int(rng.integers(0, ranks)), bool(rng.random() < 0.15)) for _ in range(nsat)]
    out["saturated_unbatched"] = run_pass(sat)
    log({"serving_saturated_unbatched": out["saturated_unbatched"]})
    in_use, peak = srv.usage()
    out["hbm_reserved_peak_bytes"] = peak
    srv.close()
    del pool, outs, handles
    torch.cuda.empty_cache()
    return out


# ---------------------------------------------------------------------------
# CPU reference path (64 MiB, the recipe's own data)
# ---------------------------------------------------------------------------
def run_c1_cpu(W, O, chunk=128, size=64 << 20, nckpt=10):
    """The 64 MiB chain on the CPU: 1 % fresh chunks, 2 % copied chunks and a
    zeroed 1 MiB region per checkpoint.  The serial oracle (one core) and the
    OpenMP restatement on the same chain; the serial oracle's restart of the
    newest checkpoint is compared with the data."""
    import numpy as np

    chain = W.chain_uniform_changes(1931, size, chunk, nckpt)
    o = O.OracleDeduplicator(chunk)
    t0 = time.perf_counter()
    diffs = [o.checkpoint_array(x, k == 0) for k, x in enumerate(chain)]
    ser = time.perf_counter() - t0
    p = O.ParallelOracleDeduplicator(chunk)
    t0 = time.perf_counter()
    pdiffs = [p.checkpoint_array(x, k == 0) for k, x in enumerate(chain)]
    par = time.perf_counter() - t0
    rs = O.restart(diffs, nckpt - 1, size)
    return dict(workload=f"c1_{size >> 20}MiB_c{chunk}_murmur3", checkpoints=nckpt,
                serial_oracle_gbs=round(nckpt * size / ser / 1e9, 4),
                openmp_gbs=round(nckpt * size / par / 1e9, 4), openmp_threads=O.ParallelOracleDeduplicator.threads(),
                diff_plus_metadata_bytes=sum(len(d) for d in diffs), openmp_equals_serial=pdiffs == diffs,
                restart_byte_identical=bool(np.array_equal(np.frombuffer(rs, np.uint8), chain[-1])))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (buffers of different ranks are "
              "independent, so N GPUs run N copies of this workload).", file=sys.stderr)
        sys.exit(2)
    import torch

    import reproducibility_icpp23_scalable_gpu_deduplication_amd as dd
    from reproducibility_icpp23_scalable_gpu_deduplication_amd import workloads as W

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    # The cyclic garbage collector runs only where called (before each timed
    # loop): a full collection inside a timed loop stalled one step of the
    # batched leg by ~45 ms.
    gc.disable()

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration:
        src = torch.randint(0, 256, (args.copy_calibration,), dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": args.copy_calibration, "copies": 5}), flush=True)
        return
    only = args.only
    chunk = args.chunk
    plain_wl = f"plain_{args.size >> 20}MiB_c{chunk}_{args.frac * 100:g}pct_murmur3"
    if only:
        # one section alone (for profiling); prints its object only
        if only == "batched":
            r = {k: v for k, v in run_batched(torch, dd, W, O, args).items() if k != "cpu"}
        elif only == "plain":
            r = run_plain(torch, dd, W, O, args.size, chunk, args.frac, args.nckpt, args.plain_steps, args.warmup,
                          check=False)
        elif only == "restart":
            r = run_plain(torch, dd, W, O, args.size, chunk, args.frac, args.nckpt, 1, 0, check=False, restart=True)
        elif only == "list":
            r = run_plain(torch, dd, W, O, args.size, chunk, args.frac, args.nckpt, args.plain_steps, args.warmup,
                          cls=dd.ListDeduplicator, check=False)
        elif only == "field":
            r = run_field(torch, dd, W, args.field_values, 20, 64)
        elif only == "serve":
            r = run_serving(torch, dd, W, args.serve_ranks, args.serve_rate, args.serve_seconds,
                            int(args.serve_budget_gib * (1 << 30)), args.serve_slots, chunk, dd.HASH_MURMUR3)
        else:
            r = [dict(chunk=c, frac=f, **{k: v for k, v in run_plain(torch, dd, W, O, args.size, c, f, args.nckpt,
                                                                      args.nckpt, args.nckpt, check=False).items()
                                          if k in ("gbs", "leaf_gbs")})
                 for c in (128, 512, 4096) for f in (0.001, 0.01, 0.1)]
        print(json.dumps({only: r}, default=float), flush=True)
        return

    b = run_batched(torch, dd, W, O, args)
    log({"batched": {k: v for k, v in b.items() if k != "cpu"}})
    total = b["total_bytes"]
    line = {
        "metric": METRIC if b["bit_exact"] else METRIC.replace(", bit-exact", "") + " (bit-exactness check FAILED)",
        "value": round(b["gbs"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(b["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": f"batched_{b['buffers']}buf_{total / 2**30:.1f}GiB_c{chunk}_murmur3",
            "buffers": b["buffers"],
            "buffer_sizes": "log-uniform in [4 MiB, 512 MiB]",
            "total_bytes": total,
            "chunk_size": chunk,
            "hash": "murmur3",
            "checkpoints_in_chain": args.nckpt,
            "change_pattern": "per buffer r ~ U[0.1 %, 5 %]: r/3 fresh chunks, 2r/3 copied chunks, 1 MiB zeroed",
            "step": "one dd_batch_checkpoint_device call (hipGraph replay) over all buffers",
            "parallelism": "single GPU",
            "leaf_mapping": b["leaf_mapping"],
            "stage_budget": b["stage_budget"],
        },
        "bit_exact": b["bit_exact"],
        "bit_exact_checks": {"batched_vs_plain_mismatches": b["mismatches"],
                             "oracle_sample_buffers": b["oracle_sample_buffers"],
                             "oracle_mismatches": b["oracle_mismatches"],
                             "restart_mismatches": b["restart_mismatches"]},
        "plain_gbs": round(b["plain_gbs"], 2),
        "plain_ms_per_step": round(b["plain_ms_per_step"], 4),
        "speedup_vs_plain": round(b["gbs"] / b["plain_gbs"], 3),
        "roofline": dict(roofline(b["leaf_gbs"], traffic_from_profiles("leaf_hash", f"batched_{b['buffers']}buf_c{chunk}"),
                                  "k_leaf_hash (all buffers in one launch)", b["leaf_alg_bytes"], b["leaf_ms"]),
                         whole_step={"bytes_per_step": int(b["step_alg_bytes"]),
                                     "achieved": round(b["step_gbs"], 1),
                                     "frac": round(b["step_gbs"] / HBM_PEAK_GBS, 4),
                                     "definition": "B = N_in + 32 C + 2 D + 32 + M per buffer"}),
    }
    if not args.no_cpu:
        c = b["cpu"]
        line["cpu_baseline"] = {
            "value": round(c["gbs"], 3), "unit": "GB/s", "cores": c["threads"], "kind": "port",
            "sample": f"OpenMP restatement of the checkpoint path (oracle/dedup_omp.c, {c['threads']} threads, "
                      f"{host_cpu()}) over {c['buffers']} of the {b['buffers']} buffers of this workload "
                      f"({b['oracle_sample_bytes'] / 2**30:.2f} GiB), all {args.nckpt} checkpoints each, same bytes; "
                      f"{c['seconds']:.1f} s of CPU work"}
    p = run_plain(torch, dd, W, O, args.size, chunk, args.frac, args.nckpt, args.plain_steps, args.warmup,
                  restart=True, d2h=True)
    traffic = traffic_from_profiles("leaf_hash", plain_wl)
    line["plain"] = {
        "workload": plain_wl, "gbs": round(p["gbs"], 2), "ms_per_step": round(p["ms_per_step"], 4),
        "oracle_mismatches": p["oracle_mismatches"], "leaf_mapping": p["leaf_mapping"],
        "roofline": dict(roofline(p["leaf_gbs"], traffic, "k_leaf_hash", p["leaf_alg_bytes"], p["leaf_ms"]),
                         counter_frac=(round(traffic / (p["leaf_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                       if traffic else None),
                         whole_step={"bytes_per_step": int(p["step_alg_bytes"]), "achieved": round(p["step_gbs"], 1),
                                     "frac": round(p["step_gbs"] / HBM_PEAK_GBS, 4)}),
        "stage_ms_untimed_chain": {k: round(v, 4) for k, v in p["stage_ms"].items()},
        "gpu_ms_baseline_ckpt": p["gpu_ms_baseline"] and round(p["gpu_ms_baseline"], 4),
        "gpu_ms_incremental_ckpt": p["gpu_ms_incremental"] and round(p["gpu_ms_incremental"], 4),
        "diff_bytes_avg": int(p["diff_bytes_avg"]),
        "d2h_inclusive": p["d2h_inclusive"],
    }
    line["restart"] = dict(workload=plain_wl, **p["restart"])
    line["bit_exact"] = line["bit_exact"] and p["oracle_mismatches"] == 0 and p["restart"]["newest"]["byte_identical"]
    log({"plain": line["plain"], "restart": line["restart"]})
    if not args.no_sweep:
        sweep = []
        for c in (128, 512, 4096):
            for f in (0.001, 0.01, 0.1):
                s = run_plain(torch, dd, W, O, args.size, c, f, args.nckpt, args.nckpt, args.nckpt, check=False)
                sweep.append(dict(chunk=c, frac=f, gbs=round(s["gbs"], 2), leaf_gbs=round(s["leaf_gbs"], 1),
                                  whole_step_gbs=round(s["step_gbs"], 1), leaf_mapping=s["leaf_mapping"],
                                  ms_baseline=round(s["gpu_ms_baseline"], 4),
                                  ms_incremental=round(s["gpu_ms_incremental"], 4)))
                log({"sweep_point": sweep[-1]})
        line["sweep"] = sweep
    if not args.no_field:
        line["field_md5"] = run_field(torch, dd, W, args.field_values, 20, 64)
        log({"field_md5": line["field_md5"]})
    if not args.no_list:
        lr = run_plain(torch, dd, W, O, args.size, chunk, args.frac, args.nckpt, args.plain_steps, args.warmup,
                       cls=dd.ListDeduplicator, check=False)
        line["list"] = dict(gbs=round(lr["gbs"], 3), ms_per_step=round(lr["ms_per_step"], 4),
                            leaf_gbs=round(lr["leaf_gbs"], 1), leaf_mapping=lr["leaf_mapping"],
                            gpu_ms_baseline_ckpt=round(lr["gpu_ms_baseline"], 4),
                            gpu_ms_incremental_ckpt=round(lr["gpu_ms_incremental"], 4),
                            diff_bytes_avg=int(lr["diff_bytes_avg"]),
                            diff_bytes_vs_tree=round(lr["diff_bytes_avg"] / p["diff_bytes_avg"], 3))
        log({"list": line["list"]})
    if not args.no_serve:
        line["serving"] = run_serving(torch, dd, W, args.serve_ranks, args.serve_rate, args.serve_seconds,
                                      int(args.serve_budget_gib * (1 << 30)), args.serve_slots, chunk,
                                      dd.HASH_MURMUR3)
        line["p99_ms"] = line["serving"]["poisson"]["p99_ms"]
        line["p50_ms"] = line["serving"]["poisson"]["p50_ms"]
    if not args.no_cpu:
        line["c1_cpu_reference"] = run_c1_cpu(W, O)
        log({"c1_cpu_reference": line["c1_cpu_reference"]})
    if not line["bit_exact"] and "FAILED" not in line["metric"]:
        line["metric"] = METRIC.replace(", bit-exact", "") + " (bit-exactness check FAILED)"
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
