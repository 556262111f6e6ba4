"""List method (ListDeduplicator, src/list_approach.cpp) on the GPU vs the
list oracle, bit-exact, through the C ABI; every checkpoint restarts
byte-identical on the GPU."""
import random

import numpy as np
import pytest

import oracle as O
import reproducibility_icpp23_scalable_gpu_deduplication_amd as dd
from reproducibility_icpp23_scalable_gpu_deduplication_amd import workloads as W
from gpu_util import first_mismatch, gpu_chain, gpu_restart, oracle_chain, to_dev
from test_gpu_batch import run_batch
from test_oracle import _make_chain, diff_bytes

pytestmark = pytest.mark.gpu


def check_list_chain(torch, chunk, datas, kind=dd.HASH_MURMUR3, baselines=(0,), restart_ids=None):
    g, gd, gc = gpu_chain(torch, chunk, datas, kind, baselines, cls=dd.ListDeduplicator)
    od, oc = oracle_chain(chunk, datas, kind, baselines, cls=O.OracleListDeduplicator)
    for k, (a, b) in enumerate(zip(gd, od)):
        assert a == b, f"ckpt {k}: {first_mismatch(a, b)}"
    assert gc == oc
    ids = range(len(datas)) if restart_ids is None else restart_ids
    for k in ids:
        want = datas[k].tobytes() if isinstance(datas[k], np.ndarray) else bytes(datas[k])
        got = gpu_restart(torch, g, gd, k, len(want))
        assert got == want, f"restart {k}: {first_mismatch(got, want)}"
    return g, gd


def test_method_and_hand_derived_chain(torch_gpu):
    g, gd = check_list_chain(torch_gpu, 1, [b"ABCA", b"ABDB", b"DDDD"])
    assert g.method == dd.METHOD_LIST and dd.TreeDeduplicator(1).method == dd.METHOD_TREE
    assert dd.TreeLowRootDeduplicator(1).method == dd.METHOD_TREE_LOW_ROOT
    assert gd[0] == diff_bytes(0, 0, 4, 1, [0, 1, 2], [(0, 1)], [(3, 0)], b"ABC")
    assert gd[2] == diff_bytes(0, 2, 4, 1, [], [(1, 3)], [(0, 2), (1, 2), (3, 2)], b"")


def test_single_chunk_and_partial_chunk(torch_gpu):
    check_list_chain(torch_gpu, 4, [b"AAAA", b"BBBB", b"AAAA"])
    check_list_chain(torch_gpu, 2, [b"abcdabe", b"cdcdxye", b"abcdxyq"])


@pytest.mark.parametrize("seed", range(16))
def test_random_small_chains(torch_gpu, seed):
    rng = random.Random(7300 + seed)
    C = rng.choice([1, 2, 3, 7, 16, 33, 100, 257, 1000, 4097])
    cs = rng.choice([1, 3, 4, 16, 64, 128])
    tail = rng.randrange(0, cs) if cs > 1 else 0
    kind = rng.choice([dd.HASH_MURMUR3, dd.HASH_MD5])
    datas = _make_chain(rng, C, cs, 6, tail)
    check_list_chain(torch_gpu, cs, datas, kind, baselines=(0,) if seed % 3 else (0, 3))


@pytest.mark.parametrize("C,cs", [(1 << 12, 128), (5000, 128), (1 << 14, 16), (3 << 11, 512), (777, 4096)])
def test_medium_chains_with_dups(torch_gpu, C, cs):
    datas = W.chain_uniform_changes(C + cs, C * cs, cs, 6, fresh_frac=0.02, copy_frac=0.05, zero_bytes=4 * cs)
    check_list_chain(torch_gpu, cs, datas)


def test_long_runs_of_equal_chunks(torch_gpu):
    # Zero-filled regions: long runs of one digest through the leaf inserts'
    # run compression, toggling between checkpoints.
    rng = np.random.default_rng(77)
    C, cs = 1 << 13, 64
    x = np.zeros(C * cs, dtype=np.uint8)
    x[: C * cs // 8] = rng.integers(0, 256, C * cs // 8, dtype=np.uint8)
    datas = [x.copy()]
    for _ in range(4):
        y = datas[-1].copy()
        for st in rng.integers(0, C - 64, 6):
            ln = int(rng.integers(1, 64))
            y[st * cs:(st + ln) * cs] = 0 if rng.random() < 0.5 else rng.integers(0, 256, ln * cs, dtype=np.uint8)
        datas.append(y)
    check_list_chain(torch_gpu, cs, datas, baselines=(0, 3))


def test_md5_float_field(torch_gpu):
    check_list_chain(torch_gpu, 64, W.chain_float_field(3, 1 << 16, 5), dd.HASH_MD5)


def test_fresh_data_every_checkpoint_grows_map(torch_gpu):
    # The reference's test/list_chkpt.cpp: new random bytes every checkpoint
    # (the map grows and the inserts re-run).
    rng = np.random.default_rng(1931)
    datas = [rng.integers(0, 256, 1 << 20, dtype=np.uint8) for _ in range(6)]
    check_list_chain(torch_gpu, 128, datas)


def test_config1_shape_64mib(torch_gpu):
    datas = W.chain_uniform_changes(1931, 64 << 20, 128, 10)
    check_list_chain(torch_gpu, 128, datas, restart_ids=[0, 4, 9])


@pytest.mark.parametrize("mapping", ["direct", "deep", "lds128", None])
def test_leaf_mappings(torch_gpu, mapping, monkeypatch):
    if mapping:
        monkeypatch.setenv("DD_LEAF_KERNEL", mapping)
    datas = W.chain_uniform_changes(3, 3000 * 256, 256, 4, fresh_frac=0.02, copy_frac=0.04, zero_bytes=4096)
    check_list_chain(torch_gpu, 256, datas, restart_ids=[3])


def test_device_output_and_device_restart(torch_gpu):
    torch = torch_gpu
    datas = W.chain_uniform_changes(8, 1 << 20, 128, 4, zero_bytes=8192)
    g1 = dd.ListDeduplicator(128)
    g2 = dd.ListDeduplicator(128)
    cap = dd.max_diff_size(1 << 20, 128, 3)
    out = torch.empty(cap, dtype=torch.uint8, device="cuda")
    dev_diffs = []
    for i, x in enumerate(datas):
        t = to_dev(torch, x)
        host = g1.checkpoint(t.data_ptr(), t.numel(), i == 0)
        n = g2.checkpoint_device(t.data_ptr(), t.numel(), out.data_ptr(), cap, i == 0)
        torch.cuda.synchronize()
        assert out[:n].cpu().numpy().tobytes() == host
        dev_diffs.append(out[:n].clone())
    torch.cuda.synchronize()
    res = torch.zeros(1 << 20, dtype=torch.uint8, device="cuda")
    for k in range(len(datas)):
        g2.restart_device([d.data_ptr() for d in dev_diffs], [d.numel() for d in dev_diffs], k, res.data_ptr(),
                          1 << 20)
        torch.cuda.synchronize()
        assert res.cpu().numpy().tobytes() == datas[k].tobytes()


def test_length_change_is_rejected(torch_gpu):
    torch = torch_gpu
    g = dd.ListDeduplicator(16)
    t = to_dev(torch, bytes(range(256)) * 4)
    g.checkpoint(t.data_ptr(), 1024, True)
    with pytest.raises(dd.DedupError) as e:
        g.checkpoint(t.data_ptr(), 512, False)
    assert e.value.status == dd.DD_ESTATE


@pytest.mark.parametrize("graph", [False, True])
def test_batch_of_list_handles(torch_gpu, graph):
    sizes = [1 << 20, 777 * 128, 5000 * 128 + 40]
    chains = [W.chain_uniform_changes(41 + i, s, 128, 4, fresh_frac=0.02, copy_frac=0.03, zero_bytes=4096)
              for i, s in enumerate(sizes)]
    handles, diffs = run_batch(torch_gpu, chains, 128, graph, cls=dd.ListDeduplicator)
    for i, c in enumerate(chains):
        od, _ = oracle_chain(128, c, cls=O.OracleListDeduplicator)
        for k in range(len(c)):
            assert diffs[i][k] == od[k], f"buffer {i} ckpt {k}: {first_mismatch(diffs[i][k], od[k])}"
        assert gpu_restart(torch_gpu, handles[i], diffs[i], len(c) - 1, c[-1].nbytes) == c[-1].tobytes()


def test_batch_rejects_mixed_methods(torch_gpu):
    with pytest.raises(dd.DedupError) as e:
        dd.Batch([dd.TreeDeduplicator(128), dd.ListDeduplicator(128)])
    assert e.value.status == dd.DD_EINVAL


def test_server_runs_list_handles(torch_gpu):
    torch = torch_gpu
    chain = W.chain_uniform_changes(5, 1 << 18, 128, 3, zero_bytes=4096)
    od, _ = oracle_chain(128, chain, cls=O.OracleListDeduplicator)
    srv = dd.Server(1 << 30, slots=2)
    g = dd.ListDeduplicator(128)
    for k, x in enumerate(chain):
        t = to_dev(torch, x)
        cap = dd.max_diff_size(x.nbytes, 128, k)
        o = torch.empty(cap, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        res = srv.wait(srv.submit_checkpoint(g, t.data_ptr(), t.numel(), o.data_ptr(), cap, k == 0))
        assert res["status"] == dd.DD_OK
        assert o[:res["out_len"]].cpu().numpy().tobytes() == od[k]
    srv.close()
