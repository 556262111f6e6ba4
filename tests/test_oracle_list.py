"""The list-approach oracle (ListDeduplicator, src/list_approach.cpp) pinned
by hand-derived diffs, by the tree oracle's leaf labels (the two approaches
classify chunks the same way; the tree oracle is pinned by the reference's
known-answer tests), and by restart round trips (the reference's
test/list_chkpt.cpp checks restarts of every checkpoint)."""
import random
import struct

import numpy as np
import pytest

import oracle as O
from reproducibility_icpp23_scalable_gpu_deduplication_amd import workloads as W
from test_oracle import _make_chain, diff_bytes


def run_list(chunk, datas, kind=O.MURMUR3, baselines=(0,)):
    o = O.OracleListDeduplicator(chunk, kind)
    diffs, counts = [], []
    for k, x in enumerate(datas):
        diffs.append(o.checkpoint_array(x, k in baselines) if isinstance(x, np.ndarray) else o.checkpoint(x, k in baselines))
        counts.append(o.counts())
    return diffs, counts


def parse(d):
    ref, ck, dl, cs, nf, npr, ns = struct.unpack_from("<IIQIIII", d)
    p = 32
    first = list(struct.unpack_from(f"<{nf}I", d, p))
    p += 4 * nf
    table = [struct.unpack_from("<II", d, p + 8 * i) for i in range(npr)]
    p += 8 * npr
    shifts = [struct.unpack_from("<II", d, p + 8 * i) for i in range(ns)]
    p += 8 * ns
    return dict(ref=ref, ckpt=ck, datalen=dl, chunk=cs, first=first, table=table, shifts=shifts, data=d[p:])


# Derived by hand from dedup_data / collect_diff (src/list_approach.cpp:23-256)
# on one-byte chunks, in the serial (one-thread) order.
def test_hand_derived_chain():
    d = run_list(1, [b"ABCA", b"ABDB", b"DDDD"])[0]
    # ckpt 0: every chunk differs from the zero list; chunk 3 repeats chunk 0.
    assert d[0] == diff_bytes(0, 0, 4, 1, [0, 1, 2], [(0, 1)], [(3, 0)], b"ABC")
    # ckpt 1: chunks 0, 1 unchanged; 2 is new; 3 = "B" first seen at (1, ckpt 0).
    assert d[1] == diff_bytes(0, 1, 4, 1, [2], [(0, 1)], [(3, 1)], b"D")
    # ckpt 2: chunk 2 unchanged; 0, 1, 3 = "D" first seen at (2, ckpt 1).
    assert d[2] == diff_bytes(0, 2, 4, 1, [], [(1, 3)], [(0, 2), (1, 2), (3, 2)], b"")


def test_hand_derived_partial_chunk_and_mixed_sources():
    # 2-byte chunks, 7 bytes: chunks "ab" "cd" "ab" "e" (short).
    d = run_list(2, [b"abcdabe", b"cdcdxye", b"abcdxyq"])[0]
    assert d[0] == diff_bytes(0, 0, 7, 2, [0, 1, 3], [(0, 1)], [(2, 0)], b"abcde\x00")
    # ckpt 1: chunk 0 "cd" seen at (1, 0); chunk 2 "xy" new; chunk 3 unchanged.
    assert d[1] == diff_bytes(0, 1, 7, 2, [2], [(0, 1)], [(0, 1)], b"xy")
    # ckpt 2: chunk 0 "ab" from ckpt 0 (chunk 0), chunk 3 "q" new; sources
    # grouped by checkpoint.
    assert d[2] == diff_bytes(0, 2, 7, 2, [3], [(0, 1)], [(0, 0)], b"q\x00")


def test_single_chunk_shift_is_emitted():
    d = run_list(4, [b"AAAA", b"BBBB", b"AAAA"])[0]
    assert d[2] == diff_bytes(0, 2, 4, 4, [], [(0, 1)], [(0, 0)], b"")
    assert O.restart_list(d, 2, 4) == b"AAAA"


def _tree_leaf_sets(chunk, datas, kind, baselines):
    o = O.OracleDeduplicator(chunk, kind)
    out = []
    for k, x in enumerate(datas):
        o.checkpoint(x, k in baselines)
        C = -(-len(x) // chunk)
        lab = np.frombuffer(o.labels(C), np.uint8)[C - 1:]
        out.append((set(np.nonzero(lab == O.FIRST_OCUR)[0].tolist()), set(np.nonzero(lab == O.SHIFT_DUPL)[0].tolist())))
    return out


@pytest.mark.parametrize("seed", range(30))
def test_classification_equals_tree_leaf_labels_and_restarts(seed):
    rng = random.Random(9100 + seed)
    C = rng.choice([1, 2, 3, 7, 16, 33, 100, 257, 1000])
    cs = rng.choice([1, 3, 4, 16, 64])
    tail = rng.randrange(0, cs) if cs > 1 else 0
    kind = rng.choice([O.MURMUR3, O.MD5])
    datas = _make_chain(rng, C, cs, 7, tail)
    baselines = (0,) if seed % 3 else (0, 4)
    diffs, counts = run_list(cs, datas, kind, baselines)
    tree = _tree_leaf_sets(cs, datas, kind, baselines)
    for k, d in enumerate(diffs):
        p = parse(d)
        assert p["first"] == sorted(p["first"])
        assert set(p["first"]) == tree[k][0], k
        assert {n for n, _ in p["shifts"]} == tree[k][1], k
        assert counts[k] == (len(p["first"]), len(p["shifts"]))
        # Prior table ascending with nonzero counts; shifted entries grouped
        # by source checkpoint in table order, chunks ascending in a group.
        assert [c for c, _ in p["table"]] == sorted(c for c, _ in p["table"])
        assert all(n > 0 for _, n in p["table"]) and sum(n for _, n in p["table"]) == len(p["shifts"])
        pos = 0
        for c, n in p["table"]:
            grp = [s for s, _ in p["shifts"][pos:pos + n]]
            assert grp == sorted(grp)
            pos += n
        assert len(p["data"]) == cs * len(p["first"])
        assert p["ref"] == max(b for b in baselines if b <= k) and p["ckpt"] == k
    for k in range(len(datas)):
        assert O.restart_list(diffs, k, len(datas[k])) == datas[k], k


def test_shift_sources_are_first_occurrences_with_equal_data():
    datas = W.chain_uniform_changes(11, 3000 * 32, 32, 6, fresh_frac=0.03, copy_frac=0.06, zero_bytes=2048)
    diffs, _ = run_list(32, datas)
    firsts = {}
    for k, d in enumerate(diffs):
        p = parse(d)
        for i, f in enumerate(p["first"]):
            firsts[(f, k)] = p["data"][32 * i:32 * (i + 1)]
        pos = 0
        for c, n in p["table"]:
            for node, prev in p["shifts"][pos:pos + n]:
                assert (prev, c) in firsts
                assert bytes(datas[k][32 * node:32 * node + 32]) == firsts[(prev, c)]
            pos += n


def test_medium_chain_restarts():
    datas = W.chain_uniform_changes(5, (1 << 14) * 64 + 17, 64, 6, fresh_frac=0.02, copy_frac=0.05, zero_bytes=8192)
    diffs, _ = run_list(64, datas)
    for k in (0, 3, 5):
        assert O.restart_list(diffs, k, len(datas[k])) == bytes(datas[k])


def test_errors():
    o = O.OracleListDeduplicator(8)
    o.checkpoint(b"x" * 64, True)
    with pytest.raises(RuntimeError):
        o.checkpoint(b"x" * 72)  # incremental checkpoint with another length
    d = run_list(8, [b"a" * 64, b"b" * 64])[0]
    assert O.restart_list_rc(d, 1, 63) != 0  # length differs from the header
    assert O.restart_list_rc([d[0], d[1][:40]], 1, 64) != 0  # truncated metadata
