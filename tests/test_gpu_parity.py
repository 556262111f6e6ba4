"""HIP path vs the CPU oracle, bit-exact, through the C ABI.

Every diff the GPU returns must equal the oracle's byte for byte, and every
checkpoint of every chain must restart byte-identical (restart on the GPU).
"""
import random
import struct

import numpy as np
import pytest

import oracle as O
import reproducibility_icpp23_scalable_gpu_deduplication_amd as dd
from reproducibility_icpp23_scalable_gpu_deduplication_amd import workloads as W
from gpu_util import first_mismatch, gpu_chain, gpu_restart, oracle_chain, to_dev
from test_oracle import HAND_DIFFS, KATS, _make_chain

pytestmark = pytest.mark.gpu


def check_chain(torch, chunk, datas, kind=dd.HASH_MURMUR3, baselines=(0,), restart_ids=None):
    g, gd, gc = gpu_chain(torch, chunk, datas, kind, baselines)
    od, oc = oracle_chain(chunk, datas, kind, baselines)
    for k, (a, b) in enumerate(zip(gd, od)):
        assert a == b, f"ckpt {k}: {first_mismatch(a, b)}"
    assert gc == oc
    ids = range(len(datas)) if restart_ids is None else restart_ids
    for k in ids:
        want = datas[k].tobytes() if isinstance(datas[k], np.ndarray) else bytes(datas[k])
        got = gpu_restart(torch, g, gd, k, len(want))
        assert got == want, f"restart {k}: {first_mismatch(got, want)}"
    return g, gd


@pytest.mark.parametrize("name", sorted(KATS))
def test_kats(torch_gpu, name):
    a, b, ef, es = KATS[name]
    g, gd = check_chain(torch_gpu, 1, [a.encode(), b.encode()])
    assert g.counts() == (ef, es)
    if name in HAND_DIFFS:
        assert gd[1] == HAND_DIFFS[name]


@pytest.mark.parametrize("seed", range(24))
def test_random_small_chains(torch_gpu, seed):
    rng = random.Random(1000 + seed)
    C = rng.choice([1, 2, 3, 5, 8, 13, 32, 33, 100, 255, 256, 257])
    cs = rng.choice([1, 3, 4, 16, 32, 64])
    tail = rng.randrange(0, cs) if cs > 1 else 0
    kind = rng.choice([dd.HASH_MURMUR3, dd.HASH_MD5])
    datas = _make_chain(rng, C, cs, 6, tail)
    check_chain(torch_gpu, cs, datas, kind)


@pytest.mark.parametrize("C,cs", [(1 << 12, 128), (5000, 128), (1 << 14, 16), (3 << 11, 512), (777, 4096)])
def test_medium_chains_with_dups(torch_gpu, C, cs):
    datas = W.chain_uniform_changes(C + cs, C * cs, cs, 6, fresh_frac=0.02, copy_frac=0.05, zero_bytes=4 * cs)
    check_chain(torch_gpu, cs, datas)


def test_md5_float_field(torch_gpu):
    # 60 % zero 4 KiB blocks: intra-checkpoint duplicates from the baseline on.
    datas = W.chain_float_field(3, 1 << 16, 5)
    check_chain(torch_gpu, 64, datas, dd.HASH_MD5)


@pytest.mark.parametrize("C,cs,kind", [(1 << 13, 64, dd.HASH_MURMUR3), (6001, 32, dd.HASH_MURMUR3),
                                       (3 << 10, 64, dd.HASH_MD5)])
def test_long_runs_of_equal_chunks(torch_gpu, C, cs, kind):
    # Mostly-zero buffers with a few repeated patterns: long runs of one
    # digest at the baseline (run compression in the leaf inserts and in the
    # duplicate-forest lookups), runs that toggle between checkpoints, and a
    # repeated non-zero pattern written at several offsets.
    rng = np.random.default_rng(C + cs)
    n = C * cs
    x = np.zeros(n, dtype=np.uint8)
    pat = rng.integers(0, 256, 4 * cs, dtype=np.uint8)
    for off in rng.integers(0, C - 8, 20):
        x[off * cs:off * cs + pat.size] = pat
    x[: n // 8] = rng.integers(0, 256, n // 8, dtype=np.uint8)
    datas = [x.copy()]
    for _ in range(4):
        y = datas[-1].copy()
        for st in rng.integers(0, C - 64, 6):
            ln = int(rng.integers(1, 64))
            y[st * cs:(st + ln) * cs] = 0 if rng.random() < 0.5 else rng.integers(0, 256, ln * cs, dtype=np.uint8)
        for off in rng.integers(0, C - 8, 5):
            y[off * cs:off * cs + pat.size] = pat
        datas.append(y)
    check_chain(torch_gpu, cs, datas, kind, baselines=(0, 3))


def test_partial_last_chunk(torch_gpu):
    rng = np.random.default_rng(9)
    base = rng.integers(0, 256, 128 * 37 + 77, dtype=np.uint8)
    d1 = base.copy()
    d1[-50:] = 0
    d2 = d1.copy()
    d2[:128] = d1[128 * 5:128 * 6]
    check_chain(torch_gpu, 128, [base, d1, d2])


def test_fresh_data_every_checkpoint_grows_map(torch_gpu):
    # The reference's round-trip test (test/tree_chkpt.cpp): new random bytes
    # each checkpoint, so every checkpoint inserts about 2C digests and the
    # map must grow.
    rng = np.random.default_rng(2)
    datas = [rng.integers(0, 256, 1 << 20, dtype=np.uint8) for _ in range(6)]
    check_chain(torch_gpu, 128, datas)


def test_rebaseline_mid_chain(torch_gpu):
    datas = W.chain_uniform_changes(5, 1 << 18, 64, 8, zero_bytes=4096)
    check_chain(torch_gpu, 64, datas, baselines=(0, 3))


def test_whole_buffer_repeat_matches_oracle(torch_gpu):
    # Root 0 SHIFT is not emitted (include/reference_impl.hpp:105-106).
    datas = [b"ABCDEFGH", b"IJKLMNOP", b"ABCDEFGH"]
    g, gd, _ = gpu_chain(torch_gpu, 1, datas)
    od, _ = oracle_chain(1, datas)
    assert gd == od
    assert gpu_restart(torch_gpu, g, gd, 2, 8) == b"IJKLMNOP"


def test_whole_buffer_repeat_root_shift_option(torch_gpu):
    # With DD_OPT_EMIT_ROOT_SHIFT node 0 is emitted as a shifted root: the
    # diff equals the oracle's with the same option and the repeated
    # checkpoint restarts to its own bytes (chunk 1 and a 64 KiB chain).
    rng = np.random.default_rng(3)
    x0, x1 = (rng.integers(0, 256, 1 << 16, dtype=np.uint8) for _ in range(2))
    for chunk, datas in ((1, [b"ABCDEFGH", b"IJKLMNOP", b"ABCDEFGH"]), (64, [x0, x1, x0.copy()])):
        g = dd.TreeDeduplicator(chunk)
        g.options = dd.OPT_EMIT_ROOT_SHIFT
        assert g.options == dd.OPT_EMIT_ROOT_SHIFT
        o = O.OracleDeduplicator(chunk, options=O.OPT_EMIT_ROOT_SHIFT)
        gd = []
        for i, x in enumerate(datas):
            t = to_dev(torch_gpu, x)
            gd.append(g.checkpoint(t.data_ptr(), t.numel(), i == 0))
            want = o.checkpoint(bytes(x), i == 0)
            assert gd[-1] == want, f"chunk {chunk} checkpoint {i}"
        assert struct.unpack_from("<IIQIIII", gd[2])[4:] == (0, 1, 1)
        assert gpu_restart(torch_gpu, g, gd, 2, len(datas[2])) == bytes(datas[2])
    with pytest.raises(dd.DedupError):
        g.options = 1 << 8  # no such option


def test_config1_shape_64mib(torch_gpu):
    # 64 MiB, 128 B chunks, 2^19 leaves, 10 checkpoints: 1 % fresh chunks,
    # 2 % copied chunks, one 1 MiB zero region per checkpoint.
    datas = W.chain_uniform_changes(1931, 64 << 20, 128, 10)
    check_chain(torch_gpu, 128, datas, restart_ids=[0, 4, 9])


def test_more_dirty_tiles_than_tile_workgroups(torch_gpu):
    # 2^22 + 1000 leaves of 16 B (a chunk count that is not a power of two,
    # so group-0 tiles hold leaves one level above the deepest), 4
    # checkpoints with 20 % fresh and 10 % copied chunks scattered over the
    # buffer: nearly all of the ~16K group-0 tiles are dirty, more than the
    # 8192 workgroups of a tile pass, so workgroups walk several tiles and
    # the interior inserts walk ~16K per-tile regions.
    C, cs = (1 << 22) + 1000, 16
    datas = W.chain_uniform_changes(77, C * cs, cs, 4, fresh_frac=0.2, copy_frac=0.1, zero_bytes=1 << 16)
    check_chain(torch_gpu, cs, datas, restart_ids=[0, 3])


def test_device_output_matches_host_output(torch_gpu):
    torch = torch_gpu
    datas = W.chain_uniform_changes(8, 1 << 20, 128, 4, zero_bytes=8192)
    g1 = dd.TreeDeduplicator(128)
    g2 = dd.TreeDeduplicator(128)
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


def test_header_overload_does_not_advance(torch_gpu):
    torch = torch_gpu
    g = dd.TreeDeduplicator(1)
    t = to_dev(torch, b"ABCDEFGH")
    hd, diff = g.checkpoint_header(t.data_ptr(), 8, True)
    assert g.current_id == 0
    assert hd == dict(ref_id=0, chkpt_id=0, datalen=8, chunk_size=1, num_first_ocur=1, num_prior_chkpts=0,
                      num_shift_dupl=0)
    assert diff[:32] == struct.pack("<IIQIIII", 0, 0, 8, 1, 1, 0, 0)
    g.current_id = 1
    t2 = to_dev(torch, b"IJABCDEK")
    hd2, diff2 = g.checkpoint_header(t2.data_ptr(), 8, False)
    assert diff2 == HAND_DIFFS["04"]
    assert hd2["num_shift_dupl"] == 3 and g.current_id == 1


def test_rejected_call_leaves_the_chain_intact(torch_gpu):
    # A call rejected by its checks (here a baseline request whose output
    # buffer is too small) changes nothing: the next incremental checkpoint
    # still deduplicates against the chain before it.
    torch = torch_gpu
    datas = W.chain_uniform_changes(77, 1 << 18, 128, 4, fresh_frac=0.03, copy_frac=0.03, zero_bytes=4096)
    g = dd.TreeDeduplicator(128)
    got = []
    for k in (0, 1):
        t = to_dev(torch, datas[k])
        got.append(g.checkpoint(t.data_ptr(), t.numel(), k == 0))
    t = to_dev(torch, datas[2])
    small = torch.empty(64, dtype=torch.uint8, device="cuda")
    with pytest.raises(dd.DedupError) as e:
        g.checkpoint_device(t.data_ptr(), t.numel(), small.data_ptr(), small.numel(), True)
    assert e.value.status == dd.DD_EINVAL
    assert g.current_id == 2
    for k in (2, 3):
        t = to_dev(torch, datas[k])
        got.append(g.checkpoint(t.data_ptr(), t.numel(), False))
    od, _ = oracle_chain(128, datas)
    for k in range(4):
        assert got[k] == od[k], f"ckpt {k}: {first_mismatch(got[k], od[k])}"


def test_error_paths(torch_gpu):
    torch = torch_gpu
    g = dd.TreeDeduplicator(16)
    t = to_dev(torch, bytes(range(256)) * 4)
    d0 = g.checkpoint(t.data_ptr(), 1024, True)
    # an incremental checkpoint must keep the chain's length
    with pytest.raises(dd.DedupError) as e:
        g.checkpoint(t.data_ptr(), 512, False)
    assert e.value.status == dd.DD_ESTATE
    out = torch.zeros(1024, dtype=torch.uint8, device="cuda")
    assert g.restart_rc([d0], 0, out.data_ptr(), 1000) == dd.DD_EINVAL
    assert g.restart_rc([d0[:40]], 0, out.data_ptr(), 1024) == dd.DD_ECORRUPT
    bad = bytearray(d0)
    struct.pack_into("<I", bad, 32, 1 << 30)
    assert g.restart_rc([bytes(bad)], 0, out.data_ptr(), 1024) == dd.DD_ECORRUPT
    assert g.restart_rc([d0], 3, out.data_ptr(), 1024) == dd.DD_EINVAL
    # a missing middle checkpoint makes the chain incomplete
    d1 = g.checkpoint(t.data_ptr(), 1024, False)
    assert g.restart_rc([d0[:10], d1], 1, out.data_ptr(), 1024) == dd.DD_ECORRUPT
    g.restart([d0, d1], 1, out.data_ptr(), 1024)
    torch.cuda.synchronize()
    assert out.cpu().numpy().tobytes() == bytes(range(256)) * 4


@pytest.mark.parametrize("mapping", ["direct", "deep", "lds64", "lds128", "lds256", "short", "dma", "dma1", "lds128p",
                                     None])
def test_leaf_mappings_match_oracle(torch_gpu, mapping, monkeypatch):
    # Every leaf kernel mapping (forced through DD_LEAF_KERNEL) and the tuned
    # choice (None) give the oracle's diffs; 256-byte chunks admit all but
    # the short mapping, which takes chunks of at most 128 bytes (Murmur3);
    # dma takes multiples of 128 bytes (both hashes).
    if mapping:
        monkeypatch.setenv("DD_LEAF_KERNEL", mapping)
    else:
        monkeypatch.delenv("DD_LEAF_KERNEL", raising=False)
    cases = ((256, dd.HASH_MURMUR3), (256, dd.HASH_MD5))
    if mapping in ("short", None):
        cases += ((128, dd.HASH_MURMUR3), (48, dd.HASH_MURMUR3), (16, dd.HASH_MURMUR3))
    if mapping in ("dma", "dma1", "lds128p", None):
        cases += ((128, dd.HASH_MURMUR3), (128, dd.HASH_MD5), (4096, dd.HASH_MURMUR3))
    for cs, kind in cases:
        datas = W.chain_uniform_changes(17 + cs, 3000 * cs + 7, cs, 4, fresh_frac=0.03, copy_frac=0.05,
                                        zero_bytes=8 * cs)
        g, _ = check_chain(torch_gpu, cs, datas, kind, restart_ids=[0, 3])
        got = g.leaf_mapping()
        valid = (not (mapping in ("deep", "short") and kind == dd.HASH_MD5) and not (mapping == "short" and cs > 128)
                 and not (mapping in ("dma", "dma1", "lds128p") and cs % 128))
        if mapping and valid:
            assert got == mapping
        else:
            assert got in dd.LEAF_MAPPINGS


def test_leaf_mapping_trials_keep_bytes(torch_gpu, monkeypatch):
    # With no forced mapping, the first incremental checkpoints of a key run
    # the tuner's near-best candidate mappings in turn until one is settled;
    # every checkpoint's diff is the oracle's meanwhile, and once settled the
    # reported mapping stays the same.  A buffer size of its own (37 MiB) and
    # 384-byte chunks give a key no other test uses.
    monkeypatch.delenv("DD_LEAF_KERNEL", raising=False)
    cs = 384
    datas = W.chain_uniform_changes(4242, (37 << 20) + 5 * cs, cs, 14, fresh_frac=0.01, copy_frac=0.02,
                                    zero_bytes=64 * cs)
    g, gd, _ = gpu_chain(torch_gpu, cs, datas)
    od, _ = oracle_chain(cs, datas)
    for k in range(len(datas)):
        assert gd[k] == od[k], f"checkpoint {k}: {first_mismatch(gd[k], od[k])}"
    maps = []
    for k, x in enumerate(datas):  # a second chain on a fresh handle: the key is settled by now
        if k == 0:
            h = dd.TreeDeduplicator(cs)
        t = to_dev(torch_gpu, x)
        assert h.checkpoint(t.data_ptr(), t.numel(), k == 0) == od[k]
        maps.append(h.leaf_mapping())
    assert len(set(maps[1:])) == 1, maps
