"""Bulk map build of a baseline (k_bulk_*) against the oracle and against the
dense inserts (DD_OPT_DENSE_BASELINE), bit-exact, with both of its fallbacks
forced: a bucket with more leaves than the build holds (the job switches to
the dense inserts on the device) and a final bucket that does not fit (the
host runs the baseline again with the dense inserts).  Every checkpoint also
restarts byte-identical, so the maps the bulk build wrote serve the later
incremental checkpoints' lookups and inserts correctly."""
import numpy as np
import pytest

import oracle as O
import reproducibility_icpp23_scalable_gpu_deduplication_amd as dd
from reproducibility_icpp23_scalable_gpu_deduplication_amd import workloads as W
from gpu_util import first_mismatch, gpu_restart, oracle_chain, to_dev

pytestmark = pytest.mark.gpu


def gpu_chain_opts(torch, chunk, datas, kind, options, baselines=(0,)):
    g = dd.TreeDeduplicator(chunk, kind)
    g.options = options
    diffs, counts = [], []
    for i, x in enumerate(datas):
        t = to_dev(torch, x)
        diffs.append(g.checkpoint(t.data_ptr(), t.numel(), i in baselines))
        counts.append(g.counts())
        del t
    return g, diffs, counts


def check(torch, chunk, datas, kind=dd.HASH_MURMUR3, baselines=(0,), restart_ids=None):
    od, oc = oracle_chain(chunk, datas, kind, baselines)
    for opts in (0, dd.OPT_DENSE_BASELINE):
        g, gd, gc = gpu_chain_opts(torch, chunk, datas, kind, opts, baselines)
        for k, (a, b) in enumerate(zip(gd, od)):
            assert a == b, f"options {opts} ckpt {k}: {first_mismatch(a, b)}"
        assert gc == oc
        ids = range(len(datas)) if restart_ids is None else restart_ids
        for k in ids:
            want = datas[k].tobytes() if isinstance(datas[k], np.ndarray) else bytes(datas[k])
            assert gpu_restart(torch, g, gd, k, len(want)) == want, f"options {opts} restart {k}"


def doubled(seed, half, chunk, nckpt):
    # Baseline = X X: the second half's leaves are shifted duplicates of the
    # first half's, and its subtrees are found among the interior entries the
    # final build wrote.
    rng = np.random.default_rng(seed)
    x = rng.integers(0, 256, half, dtype=np.uint8)
    base = np.concatenate([x, x])
    out = [base]
    for k in range(1, nckpt):
        v = out[-1].copy()
        idx = rng.integers(0, base.size // chunk, 64)
        for i in idx:
            v[i * chunk:(i + 1) * chunk] = rng.integers(0, 256, chunk, dtype=np.uint8)
        out.append(v)
    return out


@pytest.mark.parametrize("C,cs", [(1, 16), (3, 8), (100, 16), (257, 64), (3000, 32), (1 << 14, 128),
                                  ((1 << 16) + 77, 128)])
def test_uniform_chains(torch_gpu, C, cs):
    # Map capacities from 1024 slots (one bucket smaller than 4096) to 2^19
    # (128 buckets); partial last chunks.
    datas = W.chain_uniform_changes(C + cs, C * cs + cs // 3, cs, 5, fresh_frac=0.02, copy_frac=0.05,
                                    zero_bytes=4 * cs)
    check(torch_gpu, cs, datas)


@pytest.mark.parametrize("kind", [dd.HASH_MURMUR3, dd.HASH_MD5])
def test_doubled_baseline(torch_gpu, kind):
    check(torch_gpu, 64, doubled(5, 1 << 19, 64, 4), kind)


def test_zero_runs_fall_back_to_dense(torch_gpu):
    # 60 % zero blocks: tens of thousands of equal leaf digests in one bucket,
    # more than the build holds, so the job takes the dense inserts.
    datas = W.chain_float_field(3, 1 << 18, 4)
    check(torch_gpu, 64, datas, dd.HASH_MD5, restart_ids=(0, 3))


def test_rebaselines(torch_gpu):
    datas = W.chain_uniform_changes(9, (1 << 15) * 128, 128, 7, fresh_frac=0.05, copy_frac=0.05, zero_bytes=4096)
    check(torch_gpu, 128, datas, baselines=(0, 3, 5))


@pytest.mark.parametrize("caps", ["4,2400", "1600,8"])
def test_forced_fallbacks(torch_gpu, caps, monkeypatch):
    # "4,...": every bucket with more than four leaves switches its job to the
    # dense inserts on the device; "...,8": the final build overflows and the
    # host runs the baseline again with the dense inserts.
    monkeypatch.setenv("DD_BULK_CAPS", caps)
    datas = W.chain_uniform_changes(21, (1 << 15) * 128 + 5, 128, 4, fresh_frac=0.02, copy_frac=0.05,
                                    zero_bytes=4096)
    check(torch_gpu, 128, datas)
    check(torch_gpu, 64, doubled(6, 1 << 17, 64, 3))


@pytest.mark.parametrize("caps", [None, "1600,8"])
def test_batch_with_bulk_baselines(torch_gpu, caps, monkeypatch):
    # Batched baselines (one launch per stage over every job's buckets),
    # every few checkpoints, with graph replay; with the final build forced
    # to overflow, each job is run again alone.
    if caps:
        monkeypatch.setenv("DD_BULK_CAPS", caps)
    torch = torch_gpu
    sizes = [1 << 20, 3 << 18, 777 * 128, 5000 * 128 + 40, 1 << 16]
    chains = [W.chain_uniform_changes(41 + i, s, 128, 6, fresh_frac=0.02, copy_frac=0.03, zero_bytes=4096)
              for i, s in enumerate(sizes)]
    baselines = (0, 3)
    handles = [dd.TreeDeduplicator(128) for _ in chains]
    b = dd.Batch(handles, graph=True)
    diffs = [[] for _ in chains]
    for k in range(len(chains[0])):
        bufs = [torch.from_numpy(c[k]).to("cuda") for c in chains]
        caps_ = [dd.max_diff_size(c[k].nbytes, 128, k) for c in chains]
        outs = [torch.empty(cap, dtype=torch.uint8, device="cuda") for cap in caps_]
        torch.cuda.synchronize()
        lens = b.checkpoint_device([t.data_ptr() for t in bufs], [t.numel() for t in bufs],
                                   [k in baselines] * len(chains), [o.data_ptr() for o in outs], caps_)
        torch.cuda.synchronize()
        for i in range(len(chains)):
            diffs[i].append(outs[i][:lens[i]].cpu().numpy().tobytes())
    for i, c in enumerate(chains):
        od, _ = oracle_chain(128, c, baselines=baselines)
        for k in range(len(c)):
            assert diffs[i][k] == od[k], f"buffer {i} ckpt {k}: {first_mismatch(diffs[i][k], od[k])}"
        assert gpu_restart(torch, handles[i], diffs[i], len(c) - 1, c[-1].nbytes) == c[-1].tobytes()
