"""Batched checkpointing (dd_batch_*) is byte-identical to one handle per
buffer on the plain path and to the oracle, with and without hipGraph replay."""
import numpy as np
import pytest

import oracle as O
import reproducibility_icpp23_scalable_gpu_deduplication_amd as dd
from reproducibility_icpp23_scalable_gpu_deduplication_amd import workloads as W
from gpu_util import first_mismatch, gpu_restart, oracle_chain

pytestmark = pytest.mark.gpu


def run_batch(torch, chains, chunk, graph, baselines=(0,), cls=dd.TreeDeduplicator):
    n = len(chains)
    handles = [cls(chunk) for _ in range(n)]
    b = dd.Batch(handles, graph=graph)
    nck = len(chains[0])
    diffs = [[] for _ in range(n)]
    for k in range(nck):
        bufs = [torch.from_numpy(c[k]).to("cuda") for c in chains]
        caps = [dd.max_diff_size(c[k].nbytes, chunk, k) for c in chains]
        outs = [torch.empty(cap, dtype=torch.uint8, device="cuda") for cap in caps]
        torch.cuda.synchronize()
        lens = b.checkpoint_device([t.data_ptr() for t in bufs], [t.numel() for t in bufs], [k in baselines] * n,
                                   [o.data_ptr() for o in outs], caps)
        torch.cuda.synchronize()
        for i in range(n):
            diffs[i].append(outs[i][:lens[i]].cpu().numpy().tobytes())
    return handles, diffs


@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_oracle_uneven_sizes(torch_gpu, graph):
    sizes = [1 << 20, 3 << 18, 777 * 128, 5000 * 128 + 40, 1 << 16]
    chains = [W.chain_uniform_changes(11 + i, s, 128, 5, fresh_frac=0.02, copy_frac=0.03, zero_bytes=4096)
              for i, s in enumerate(sizes)]
    handles, diffs = run_batch(torch_gpu, chains, 128, graph)
    for i, c in enumerate(chains):
        od, _ = oracle_chain(128, c)
        for k in range(len(c)):
            assert diffs[i][k] == od[k], f"buffer {i} ckpt {k}: {first_mismatch(diffs[i][k], od[k])}"
        for k in (0, len(c) - 1):
            assert gpu_restart(torch_gpu, handles[i], diffs[i], k, c[k].nbytes) == c[k].tobytes()


def test_batch_map_growth_rerun(torch_gpu):
    # Fresh random data each checkpoint forces map growth and the re-run of
    # the inserts for the grown buffers only.
    rng = np.random.default_rng(4)
    chains = [[rng.integers(0, 256, s, dtype=np.uint8) for _ in range(4)] for s in (1 << 18, 3 << 16, 1 << 17)]
    _, diffs = run_batch(torch_gpu, chains, 128, True)
    for i, c in enumerate(chains):
        od, _ = oracle_chain(128, c)
        assert diffs[i] == od


def test_batch_rejected_call_changes_no_handle(torch_gpu):
    # The second job's output buffer is too small: the whole call is
    # rejected before any handle's state changes, and the next call continues
    # both chains.
    torch = torch_gpu
    chains = [W.chain_uniform_changes(31 + i, s, 128, 3, fresh_frac=0.03, copy_frac=0.03, zero_bytes=4096)
              for i, s in enumerate((1 << 18, 5000 * 128))]
    handles = [dd.TreeDeduplicator(128) for _ in chains]
    b = dd.Batch(handles, graph=True)
    diffs = [[] for _ in chains]

    def call(k, caps):
        bufs = [torch.from_numpy(c[k]).to("cuda") for c in chains]
        outs = [torch.empty(max(cap, 1), dtype=torch.uint8, device="cuda") for cap in caps]
        torch.cuda.synchronize()
        lens = b.checkpoint_device([t.data_ptr() for t in bufs], [t.numel() for t in bufs], [k == 0] * len(chains),
                                   [o.data_ptr() for o in outs], caps)
        torch.cuda.synchronize()
        for i in range(len(chains)):
            diffs[i].append(outs[i][:lens[i]].cpu().numpy().tobytes())

    full = lambda k: [dd.max_diff_size(c[k].nbytes, 128, k) for c in chains]
    call(0, full(0))
    with pytest.raises(dd.DedupError) as e:
        call(1, [full(1)[0], 16])
    assert e.value.status == dd.DD_EINVAL
    call(1, full(1))
    call(2, full(2))
    for i, c in enumerate(chains):
        od, _ = oracle_chain(128, c)
        assert diffs[i] == od, f"buffer {i}"


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("cls", [dd.TreeDeduplicator, dd.ListDeduplicator])
def test_batch_mixing_baselines_and_incrementals(torch_gpu, graph, cls):
    # Handle i re-baselines at checkpoint index i + 1, so calls 1-3 mix one
    # baseline with incremental checkpoints; the last handle gets fresh bytes
    # every checkpoint, so from its second incremental checkpoint on it
    # probes in narrow rounds while the others probe wide (every insert
    # kernel variant runs in one call).
    torch = torch_gpu
    sizes = (1 << 18, 777 * 128, 5000 * 128 + 40)
    chains = [W.chain_uniform_changes(61 + i, s, 128, 5, fresh_frac=0.03, copy_frac=0.04, zero_bytes=4096)
              for i, s in enumerate(sizes)]
    rng = np.random.default_rng(67)
    chains.append([rng.integers(0, 256, 3 << 16, dtype=np.uint8) for _ in range(5)])
    base = [{0, i + 1} for i in range(len(sizes))] + [{0}]
    handles = [cls(128) for _ in chains]
    b = dd.Batch(handles, graph=graph)
    diffs = [[] for _ in chains]
    for k in range(5):
        bufs = [torch.from_numpy(c[k]).to("cuda") for c in chains]
        caps = [dd.max_diff_size(c[k].nbytes, 128, k) for c in chains]
        outs = [torch.empty(cap, dtype=torch.uint8, device="cuda") for cap in caps]
        torch.cuda.synchronize()
        lens = b.checkpoint_device([t.data_ptr() for t in bufs], [t.numel() for t in bufs],
                                   [k in base[i] for i in range(len(chains))], [o.data_ptr() for o in outs], caps)
        torch.cuda.synchronize()
        for i in range(len(chains)):
            diffs[i].append(outs[i][:lens[i]].cpu().numpy().tobytes())
    ocls = O.OracleListDeduplicator if cls is dd.ListDeduplicator else O.OracleDeduplicator
    for i, c in enumerate(chains):
        od, _ = oracle_chain(128, c, baselines=base[i], cls=ocls)
        for k in range(len(c)):
            assert diffs[i][k] == od[k], f"buffer {i} ckpt {k}: {first_mismatch(diffs[i][k], od[k])}"
        assert gpu_restart(torch, handles[i], diffs[i], 4, c[4].nbytes) == c[4].tobytes()


def test_batch_equals_plain_path(torch_gpu):
    torch = torch_gpu
    sizes = [1 << 19, 9 << 16]
    chains = [W.chain_runs(21 + i, s, 128, 4, 0.05) for i, s in enumerate(sizes)]
    _, bdiffs = run_batch(torch, chains, 128, True)
    for i, c in enumerate(chains):
        g = dd.TreeDeduplicator(128)
        for k, x in enumerate(c):
            t = torch.from_numpy(x).to("cuda")
            torch.cuda.synchronize()
            assert g.checkpoint(t.data_ptr(), t.numel(), k == 0) == bdiffs[i][k]


def test_batch_tuning_trials_keep_bytes(torch_gpu):
    # 64 buffers of 4 MiB: the incremental stages' workgroups add up past the
    # smallest candidate budget, so the first incremental calls compare leaf
    # mappings and then stage budgets (timed on the calls themselves) before
    # both settle.  Every call's diffs equal the plain path's throughout, and
    # the settled budget is one of the candidates and stays.
    torch = torch_gpu
    nb, size, cs, nk = 64, (4 << 20) + 1024, 128, 26
    gen = torch.Generator(device="cuda")
    gen.manual_seed(77)
    versions = [torch.randint(0, 256, (nb, size), dtype=torch.uint8, device="cuda", generator=gen)]
    for v in range(1, 4):  # each version changes 2 % of the chunks of every buffer
        x = versions[-1].clone()
        C = size // cs
        idx = torch.randint(0, C, (nb, C // 50), device="cuda", generator=gen)
        rows = torch.arange(nb, device="cuda").unsqueeze(1).expand_as(idx)
        x.view(nb, -1, cs)[rows, idx] = torch.randint(0, 256, (cs,), dtype=torch.uint8, device="cuda", generator=gen)
        versions.append(x)
    cap = dd.max_diff_size(size, cs, nk + 1)
    outs_b = torch.empty((nb, cap), dtype=torch.uint8, device="cuda")
    outs_p = torch.empty((nb, cap), dtype=torch.uint8, device="cuda")
    hb = [dd.TreeDeduplicator(cs) for _ in range(nb)]
    hp = [dd.TreeDeduplicator(cs) for _ in range(nb)]
    batch = dd.Batch(hb, graph=True)
    torch.cuda.synchronize()
    budgets = []
    for k in range(nk):
        v = versions[k % len(versions)]
        ptrs = [v[i].data_ptr() for i in range(nb)]
        lb = batch.checkpoint_device(ptrs, [size] * nb, [k == 0] * nb, [outs_b[i].data_ptr() for i in range(nb)],
                                     [cap] * nb)
        budgets.append(batch.stage_budget())
        for i in range(nb):
            lp = hp[i].checkpoint_device(ptrs[i], size, outs_p[i].data_ptr(), cap, k == 0)
            assert lp == lb[i], (k, i)
        torch.cuda.synchronize()
        for i in range(nb):
            assert torch.equal(outs_b[i, :lb[i]], outs_p[i, :lb[i]]), (k, i)
    assert set(budgets) <= {4096, 8192, 16384}
    assert len(set(budgets[-4:])) == 1, budgets
    batch.close()
