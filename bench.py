#!/usr/bin/env python3
"""Throughput of tree-approach incremental checkpointing on one MI355X.

Workload (default): a 10-checkpoint chain of a 1 GiB uint8 buffer, 128-byte
chunks, Murmur3-x64-128; checkpoint 0 is uniform random bytes and each later
checkpoint changes 1 % of the chunks as runs of geometric length (mean 16
chunks), half fresh bytes and half copies of earlier runs.  The whole chain
is generated in HBM before timing.

A step is one checkpoint call (dd_checkpoint_device: input resident in HBM,
diff assembled in HBM); steps walk the chain in order, so every 10th step is
a baseline.  `value` = bytes deduplicated / wall time over the timed steps.

`roofline` covers the dominant kernel (leaf hashing), timed with HIP events
on the library's stream during the timed steps; its algorithmic bytes per
launch are C * (chunk + 33): the chunk, the previous 16-byte leaf digest, the
new digest and the label.  `cpu_baseline` times the serial CPU oracle (one
core) on the same recipe at 256 MiB.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=1 << 30)
    ap.add_argument("--chunk", type=int, default=128)
    ap.add_argument("--frac", type=float, default=0.01)
    ap.add_argument("--nckpt", type=int, default=10)
    ap.add_argument("--hash", choices=["murmur3", "md5"], default="murmur3")
    ap.add_argument("--sweep", action="store_true", help="also run chunk {128,512,4096} x change {0.1,1,10}%%")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-size", type=int, default=512 << 20)
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-buffer batched comparison")
    ap.add_argument("--no-list", action="store_true", help="skip the list-approach measurement")
    ap.add_argument("--method", choices=["tree", "list"], default="tree",
                    help="method of the main measurement (list: for profiling the list kernels alone)")
    ap.add_argument("--batch-buffers", type=int, default=64)
    ap.add_argument("--field", action="store_true",
                    help="also run the MD5 float-field chain (2^30 float32 values, 64 B chunks, 20 checkpoints)")
    ap.add_argument("--field-values", type=int, default=1 << 30)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run 5 device-to-device copies of BYTES (scales PMC byte counters) and exit")
    ap.add_argument("--no-serve", action="store_true", help="skip the request-loop (serving) measurement")
    ap.add_argument("--serve-ranks", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=200.0, help="Poisson arrivals, requests/s")
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-budget-gib", type=float, default=200.0)
    return ap.parse_args()


def traffic_from_profiles(kernel, workload):
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def run_config(torch, dd, W, size, chunk, frac, nckpt, steps, warmup, kind, seed=1931, cls=None):
    chain = W.chain_runs_gpu(torch, seed, size, chunk, nckpt, frac)
    torch.cuda.synchronize()
    g = (cls or dd.TreeDeduplicator)(chunk, kind)
    cap = dd.max_diff_size(size, chunk, warmup + steps + 2 * nckpt + 1)
    out = torch.empty(cap, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    diff_bytes = []

    def step(i):
        k = i % nckpt
        n = g.checkpoint_device(chain[k].data_ptr(), size, out.data_ptr(), cap, k == 0)
        return k, n

    # Timed steps record HIP events around the leaf-hash kernel only.
    g.set_profiling(1)
    for i in range(warmup):
        step(i)
    torch.cuda.synchronize()
    leaf = []
    t0 = time.perf_counter()
    for i in range(warmup, warmup + steps):
        k, n = step(i)
        diff_bytes.append(n)
        leaf.append(g.stage_times()["leaf_hash"])
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    # One more chain, untimed, with events around every stage (breakdown only).
    g.set_profiling(2)
    first = warmup + steps
    first += (nckpt - first % nckpt) % nckpt  # start at a baseline
    stages = []
    for i in range(first, first + nckpt):
        k, _ = step(i)
        stages.append((k, g.stage_times()))
    leaf_ms = sum(leaf) / len(leaf)
    C = (size + chunk - 1) // chunk
    leaf_bytes = C * (chunk + 33)
    avg = {key: sum(s[key] for _, s in stages) / len(stages) for key in stages[0][1]}
    base = [sum(s.values()) for k, s in stages if k == 0]
    incr = [sum(s.values()) for k, s in stages if k != 0]
    res = dict(
        gbs=steps * size / el / 1e9,
        ms_per_step=el * 1e3 / steps,
        leaf_ms=leaf_ms,
        leaf_gbs=leaf_bytes / (leaf_ms * 1e-3) / 1e9,
        leaf_bytes=leaf_bytes,
        stage_ms=avg,
        gpu_ms_baseline=(sum(base) / len(base)) if base else None,
        gpu_ms_incremental=(sum(incr) / len(incr)) if incr else None,
        diff_bytes_avg=sum(diff_bytes) / len(diff_bytes),
        leaf_mapping=g.leaf_mapping(),
    )
    del chain, out, g
    torch.cuda.empty_cache()
    return res


def run_batched(torch, dd, W, nbuf, nckpt, chunk, kind, seed=1931):
    """64 checkpoint buffers of log-uniform sizes in [4 MiB, 512 MiB] (one per
    simulated rank), each with its own 10-checkpoint chain (1/3 fresh and 2/3
    copied chunks at a per-buffer change rate uniform in [0.1 %, 5 %], plus a
    1 MiB zero region).  Plain: 64 sequential dd_checkpoint_device calls per
    checkpoint index.  Batched: one dd_batch call (hipGraph replay) per
    checkpoint index.  Outputs are compared byte for byte."""
    import numpy as np

    sizes = W.batch_sizes(seed, nbuf)
    rates = np.random.default_rng(seed).uniform(0.001, 0.05, nbuf)
    chains = [W.chain_uniform_changes_gpu(torch, seed + 1 + i, sz, chunk, nckpt, r / 3, 2 * r / 3)
              for i, (sz, r) in enumerate(zip(sizes, rates))]
    torch.cuda.synchronize()
    caps = [dd.max_diff_size(sz, chunk, 4 * nckpt) for sz in sizes]
    outs_p = [torch.empty(c, dtype=torch.uint8, device="cuda") for c in caps]
    outs_b = [torch.empty(c, dtype=torch.uint8, device="cuda") for c in caps]
    plain = [dd.TreeDeduplicator(chunk, kind) for _ in sizes]
    hs = [dd.TreeDeduplicator(chunk, kind) for _ in sizes]
    batch = dd.Batch(hs, graph=True)
    ptrs = [[c[k].data_ptr() for c in chains] for k in range(nckpt)]

    def plain_step(k):
        return [g.checkpoint_device(ptrs[k][i], sizes[i], outs_p[i].data_ptr(), caps[i], k == 0)
                for i, g in enumerate(plain)]

    def batch_step(k):
        return batch.checkpoint_device(ptrs[k], sizes, [k == 0] * nbuf, [o.data_ptr() for o in outs_b], caps)

    # Untimed chain: both paths, every diff compared byte for byte.
    mismatches = 0
    for k in range(nckpt):
        lp = plain_step(k)
        lb = batch_step(k)
        torch.cuda.synchronize()
        for i in range(nbuf):
            if lp[i] != lb[i] or not torch.equal(outs_p[i][:lp[i]], outs_b[i][:lb[i]]):
                mismatches += 1
    res = {}
    for name, fn in (("plain", plain_step), ("batched", batch_step)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(nckpt):
            fn(k)
        torch.cuda.synchronize()
        res[name] = time.perf_counter() - t0
    total = sum(sizes) * nckpt
    out = dict(
        buffers=nbuf, total_bytes=sum(sizes), checkpoints=nckpt, chunk_size=chunk,
        plain_gbs=round(total / res["plain"] / 1e9, 2), batched_gbs=round(total / res["batched"] / 1e9, 2),
        speedup=round(res["plain"] / res["batched"], 3), bit_exact=mismatches == 0, mismatches=mismatches,
        leaf_mapping=batch.leaf_mapping(),
        ms_per_checkpoint_index=dict(plain=round(res["plain"] * 1e3 / nckpt, 3),
                                     batched=round(res["batched"] * 1e3 / nckpt, 3)))
    del chains, outs_p, outs_b, plain, hs, batch
    torch.cuda.empty_cache()
    return out


def run_field(torch, dd, W, n_values, nckpt, chunk, seed=1931):
    """MD5 on a float32 simulation-like field: 2^30 values (4 GiB), 64-byte
    chunks, 20 checkpoints; 60 % of the 4 KiB blocks are exact zeros, each
    version perturbs 5 % of the nonzero values by N(0,1e-3) and flips 1 % of
    the blocks between zero and nonzero.  Each version is generated in HBM
    from the previous one (untimed); each checkpoint call is timed on its own
    (the call synchronizes), with HIP events around the leaf-hash kernel.
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
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n = g.checkpoint_device(v.data_ptr(), size, out.data_ptr(), cap, k == 0)
        times.append(time.perf_counter() - t0)
        leaf.append(g.stage_times()["leaf_hash"])
        diff_bytes.append(n)
        diffs.append(out[:n].clone())
        print(json.dumps({"field_ckpt": k, "ms": round(times[-1] * 1e3, 3), "diff_bytes": n,
                          "first_shift": g.counts()}), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    ok = True
    for k in keep_ids:
        o = torch.empty(size, dtype=torch.uint8, device="cuda")
        g.restart_device([d.data_ptr() for d in diffs[:k + 1]], [d.numel() for d in diffs[:k + 1]], k,
                         o.data_ptr(), size)
        ok = ok and torch.equal(o, saved[k])
    C = size // chunk
    leaf_ms = sum(leaf) / len(leaf)
    res = dict(workload=f"field_{size >> 20}MiB_c{chunk}_md5", checkpoints=nckpt, buffer_bytes=size,
               chunk_size=chunk, gbs=round(nckpt * size / sum(times) / 1e9, 2),
               ms_per_checkpoint=round(sum(times) * 1e3 / nckpt, 3),
               ms_baseline=round(times[0] * 1e3, 3),
               ms_incremental_avg=round(sum(times[1:]) * 1e3 / max(1, nckpt - 1), 3),
               leaf_kernel_ms=round(leaf_ms, 4), leaf_kernel_gbs=round(C * (chunk + 33) / (leaf_ms * 1e-3) / 1e9, 1),
               diff_bytes_total=sum(diff_bytes), restarts_verified=keep_ids, restart_bit_exact=ok)
    del diffs, saved, out, g
    torch.cuda.empty_cache()
    return res


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
    requests at once.  Every restart is compared byte for byte with the data
    it restores.  Latency = completion - submission (steady clock of the
    library)."""
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
        stats = dict(lat_ck=[], lat_rs=[], wait=[], ck_bytes=0, rs_bytes=0, bad=0, errors=0, t0=None, t1=0.0)

        def collector(nreq):
            for _ in range(nreq):
                item = done.get()
                res = srv.wait(item[0])
                if res["status"] != dd.DD_OK:
                    stats["errors"] += 1
                    continue
                lat = res["t_end"] - res["t_submit"]
                stats["wait"].append(res["t_start"] - res["t_submit"])
                stats["t1"] = max(stats["t1"], res["t_end"])
                if item[1] is None:
                    stats["lat_ck"].append(lat)
                    stats["ck_bytes"] += item[2]
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
            errors=stats["errors"])

    n = int(rate * seconds)
    gaps = rng.exponential(1.0 / rate, n)
    poisson = [(float(t), int(rng.integers(0, ranks)), bool(rng.random() < 0.15)) for t in np.cumsum(gaps)]
    out = dict(ranks=ranks, buffer_bytes_total=sum(sizes), buffer_bytes_min=min(sizes), buffer_bytes_max=max(sizes),
               chunk_size=chunk, slots=slots, hbm_budget_bytes=budget, rebaseline_every=versions)
    out["poisson"] = dict(rate_req_per_s=rate, **run_pass(poisson))
    print(json.dumps({"serving_poisson": out["poisson"]}), file=sys.stderr, flush=True)
    nsat = 12 * ranks
    sat = [(None, int(rng.integers(0, ranks)), bool(rng.random() < 0.15)) for _ in range(nsat)]
    out["saturated"] = run_pass(sat)
    print(json.dumps({"serving_saturated": out["saturated"]}), file=sys.stderr, flush=True)
    in_use, peak = srv.usage()
    out["hbm_reserved_peak_bytes"] = peak
    srv.close()
    del pool, outs, handles
    torch.cuda.empty_cache()
    return out


def cpu_baseline(W, size, chunk, frac, nckpt, serial_size=64 << 20):
    """The OpenMP restatement of the checkpoint path (oracle/dedup_omp.c, the
    serial oracle's bytes on all OpenMP threads) on a `size`-byte chain of the
    benchmark's recipe; the single-thread serial oracle on a `serial_size`
    chain beside it."""
    import platform

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    def rate(cls, n):
        chain = W.chain_runs(1931, n, chunk, nckpt, frac)
        o = cls(chunk)
        t0 = time.perf_counter()
        for i, x in enumerate(chain):
            o.checkpoint_array(x, i == 0)
        return nckpt * n / (time.perf_counter() - t0) / 1e9

    par = rate(O.ParallelOracleDeduplicator, size)
    ser = rate(O.OracleDeduplicator, serial_size)
    threads = O.ParallelOracleDeduplicator.threads()
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            cpu = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), cpu)
    except OSError:
        pass
    return dict(value=round(par, 4), unit="GB/s", cores=threads, kind="port",
                sample=f"OpenMP restatement of the checkpoint path, {threads} threads ({cpu}), "
                       f"{nckpt}-checkpoint chain of {size >> 20} MiB, chunk {chunk} B, {frac * 100:g}% changed "
                       f"per checkpoint (same recipe as the GPU workload)",
                serial_oracle_gbs=round(ser, 4),
                serial_sample=f"single-thread serial oracle, {nckpt}-checkpoint chain of {serial_size >> 20} MiB")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (buffers of different ranks are "
              "independent, so N GPUs run N copies of this workload).", file=sys.stderr)
        sys.exit(2)
    import torch

    import reproducibility_icpp23_scalable_gpu_deduplication_amd as dd
    from reproducibility_icpp23_scalable_gpu_deduplication_amd import workloads as W

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
    kind = dd.HASH_MD5 if args.hash == "md5" else dd.HASH_MURMUR3
    main_cls = dd.ListDeduplicator if args.method == "list" else dd.TreeDeduplicator
    r = run_config(torch, dd, W, args.size, args.chunk, args.frac, args.nckpt, args.steps, args.warmup, kind,
                   cls=main_cls)
    workload = f"plain_{args.size >> 20}MiB_c{args.chunk}_{args.frac * 100:g}pct_{args.hash}"
    traffic = traffic_from_profiles("leaf_hash", workload)
    line = {
        "metric": f"input GB/s deduped (plain {args.method} approach), bit-exact",
        "value": round(r["gbs"], 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "buffer_bytes": args.size,
            "chunk_size": args.chunk,
            "hash": args.hash,
            "checkpoints_in_chain": args.nckpt,
            "change_fraction": args.frac,
            "change_pattern": "geometric runs (mean 16 chunks), half fresh, half copies",
            "parallelism": "single GPU",
            "leaf_mapping": r["leaf_mapping"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(r["leaf_gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(r["leaf_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "k_leaf_hash",
            "algorithmic_bytes_per_launch": r["leaf_bytes"],
            "avg_launch_ms": round(r["leaf_ms"], 4),
        },
        "stage_ms_untimed_chain": {k: round(v, 4) for k, v in r["stage_ms"].items()},
        "gpu_ms_baseline_ckpt": r["gpu_ms_baseline"] and round(r["gpu_ms_baseline"], 4),
        "gpu_ms_incremental_ckpt": r["gpu_ms_incremental"] and round(r["gpu_ms_incremental"], 4),
        "diff_bytes_avg": int(r["diff_bytes_avg"]),
    }
    if args.sweep:
        sweep = []
        for chunk in (128, 512, 4096):
            for frac in (0.001, 0.01, 0.1):
                s = run_config(torch, dd, W, args.size, chunk, frac, args.nckpt, args.nckpt, args.nckpt, kind)
                sweep.append(dict(chunk=chunk, frac=frac, gbs=round(s["gbs"], 2), leaf_gbs=round(s["leaf_gbs"], 1),
                                  leaf_mapping=s["leaf_mapping"], ms_baseline=round(s["gpu_ms_baseline"], 4),
                                  ms_incremental=round(s["gpu_ms_incremental"], 4)))
                print(json.dumps({"sweep_point": sweep[-1]}), file=sys.stderr, flush=True)
        line["sweep"] = sweep
    if not args.no_list and args.method == "tree":
        # The list approach (ListDeduplicator) on the same chain: one map
        # entry and one metadata record per changed chunk, no tree passes.
        lr = run_config(torch, dd, W, args.size, args.chunk, args.frac, args.nckpt, args.steps, args.warmup, kind,
                        cls=dd.ListDeduplicator)
        line["list"] = dict(gbs=round(lr["gbs"], 3), ms_per_step=round(lr["ms_per_step"], 4),
                            leaf_gbs=round(lr["leaf_gbs"], 1), leaf_mapping=lr["leaf_mapping"],
                            gpu_ms_baseline_ckpt=round(lr["gpu_ms_baseline"], 4),
                            gpu_ms_incremental_ckpt=round(lr["gpu_ms_incremental"], 4),
                            diff_bytes_avg=int(lr["diff_bytes_avg"]),
                            diff_bytes_vs_tree=round(lr["diff_bytes_avg"] / r["diff_bytes_avg"], 3))
        print(json.dumps({"list": line["list"]}), file=sys.stderr, flush=True)
    if not args.no_batched:
        line["batched"] = run_batched(torch, dd, W, args.batch_buffers, args.nckpt, args.chunk, kind)
        print(json.dumps({"batched": line["batched"]}), file=sys.stderr, flush=True)
    if args.field:
        line["field_md5"] = run_field(torch, dd, W, args.field_values, 20, 64)
        print(json.dumps({"field_md5": line["field_md5"]}), file=sys.stderr, flush=True)
    if not args.no_serve:
        line["serving"] = run_serving(torch, dd, W, args.serve_ranks, args.serve_rate, args.serve_seconds,
                                      int(args.serve_budget_gib * (1 << 30)), args.serve_slots, args.chunk, kind)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(W, args.cpu_size, args.chunk, args.frac, args.nckpt)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
