"""A tree baseline's leaf kernel copies every chunk into the diff's data
section (the layout of a diff whose only region is the root) and the gather
skips the data when the labels agree; when they do not (repeated chunks), the
gather overwrites the section.  Every diff is compared with the oracle's bytes
and restarted, over the copy kernels (DMA at chunks that are multiples of 128
bytes, LDS64 at 64 bytes), a chunk size without a copy variant, partial last
chunks, trees of every shape (C a power of two, one above, odd) and baselines
that alternate between the two layouts, so the handle's prediction of the
layout is wrong in both directions.  The device-output path's guard bytes past
dd_max_diff_size stay untouched."""
import numpy as np
import pytest

import reproducibility_icpp23_scalable_gpu_deduplication_amd as dd
from gpu_util import first_mismatch, gpu_chain, gpu_restart, oracle_chain, to_dev

pytestmark = pytest.mark.gpu


def unique(rng, n):
    return rng.integers(0, 256, n, dtype=np.uint8)


def repeated(rng, n, cs):
    # The first chunk appears again at the end: the baseline's leaves are
    # not all distinct, so the root is not its only region.
    x = unique(rng, n)
    if n >= 2 * cs:
        x[n // cs * cs - cs:n // cs * cs] = x[:cs]
    return x


def check(torch, cs, datas, kind, baselines):
    g, gd, gc = gpu_chain(torch, cs, datas, kind, baselines)
    od, oc = oracle_chain(cs, datas, kind, baselines)
    for k, (a, b) in enumerate(zip(gd, od)):
        assert a == b, f"ckpt {k}: {first_mismatch(a, b)}"
    assert gc == oc
    for k in range(len(datas)):
        assert gpu_restart(torch, g, gd, k, datas[k].nbytes) == datas[k].tobytes(), f"restart {k}"
    return g


@pytest.mark.parametrize("kind", [dd.HASH_MURMUR3, dd.HASH_MD5])
@pytest.mark.parametrize("C,cs,tail", [(1, 128, 0), (3, 128, 40), (64, 128, 0), (65, 64, 0), (1 << 12, 128, 0),
                                       ((1 << 12) + 1, 128, 100), (3001, 64, 17), (777, 4096, 1000),
                                       (1000, 48, 5)])
def test_alternating_baselines(torch_gpu, kind, C, cs, tail):
    rng = np.random.default_rng(C * 7 + cs)
    n = C * cs + tail if tail else C * cs
    # unique, repeated (predicted unique: copied, then overwritten), unique
    # (predicted repeated: no copy), unique (copied), then an incremental
    # checkpoint over the copied baseline.
    datas = [unique(rng, n), repeated(rng, n, cs), unique(rng, n), unique(rng, n)]
    last = datas[-1].copy()
    last[:cs] ^= 0xFF
    datas.append(last)
    check(torch_gpu, cs, datas, kind, baselines=(0, 1, 2, 3))


def test_device_output_guard(torch_gpu):
    # The copy writes the whole data section before the labels are known:
    # it must stay inside dd_max_diff_size, also for a diff much shorter
    # than that (repeated chunks).
    rng = np.random.default_rng(3)
    cs, C = 128, 5000
    for data in (unique(rng, C * cs + 77), np.tile(unique(rng, 64 * cs), C // 64 + 1)[:C * cs]):
        g = dd.TreeDeduplicator(cs)
        t = to_dev(torch_gpu, data)
        cap = dd.max_diff_size(data.nbytes, cs, 0)
        guard = 4096
        out = torch_gpu.full((cap + guard,), 0xA5, dtype=torch_gpu.uint8, device="cuda")
        torch_gpu.cuda.synchronize()
        n = g.checkpoint_device(t.data_ptr(), t.numel(), out.data_ptr(), cap, True)
        torch_gpu.cuda.synchronize()
        host = out.cpu().numpy()
        assert (host[cap:] == 0xA5).all()
        od, _ = oracle_chain(cs, [data])
        assert host[:n].tobytes() == od[0]


def test_batch_mixed_layouts(torch_gpu):
    # One batched call: buffers with and without repeated chunks, with and
    # without partial last chunks, all baselines, then an incremental call.
    rng = np.random.default_rng(9)
    cs = 128
    sizes = [(1 << 20), (3 << 16) + 50, 777 * cs, (1 << 14) + 1]
    first = [unique(rng, s) if i % 2 == 0 else repeated(rng, s, cs) for i, s in enumerate(sizes)]
    second = [x.copy() for x in first]
    for x in second:
        x[cs:2 * cs] ^= 0x5A
    handles = [dd.TreeDeduplicator(cs) for _ in sizes]
    b = dd.Batch(handles, graph=True)
    for k, datas in enumerate((first, second)):
        bufs = [to_dev(torch_gpu, x) for x in datas]
        caps = [dd.max_diff_size(x.nbytes, cs, k) for x in datas]
        outs = [torch_gpu.empty(c, dtype=torch_gpu.uint8, device="cuda") for c in caps]
        torch_gpu.cuda.synchronize()
        lens = b.checkpoint_device([t.data_ptr() for t in bufs], [t.numel() for t in bufs], [k == 0] * len(sizes),
                                   [o.data_ptr() for o in outs], caps)
        torch_gpu.cuda.synchronize()
        for i in range(len(sizes)):
            od, _ = oracle_chain(cs, [first[i], second[i]][:k + 1])
            got = outs[i][:lens[i]].cpu().numpy().tobytes()
            assert got == od[k], f"buffer {i} ckpt {k}: {first_mismatch(got, od[k])}"


@pytest.mark.parametrize("cs,copies", [(128, True), (64, True), (48, False), (512, False)])
def test_copy_follows_last_baseline(torch_gpu, cs, copies):
    # The handle asks for the copy while its last baseline had the root as
    # its only region; ckpt_stats reports whether the leaf kernel made it.
    rng = np.random.default_rng(cs)
    n = 4096 * cs
    datas = [unique(rng, n), repeated(rng, n, cs), unique(rng, n), unique(rng, n)]
    inc = datas[-1].copy()
    inc[:cs] ^= 1
    datas.append(inc)
    want = [copies, copies, False, copies, False]
    od, _ = oracle_chain(cs, datas, baselines=(0, 1, 2, 3))
    g = dd.TreeDeduplicator(cs)
    for k, x in enumerate(datas):
        t = to_dev(torch_gpu, x)
        d = g.checkpoint(t.data_ptr(), t.numel(), k < 4)
        assert d == od[k], f"ckpt {k}: {first_mismatch(d, od[k])}"
        assert bool(g.ckpt_stats()["leaf_copy"]) == want[k], f"ckpt {k}"
