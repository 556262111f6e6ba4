"""The OpenMP restatement (oracle/dedup_omp.c, the CPU baseline) returns the
serial oracle's diff bytes and counts on every chain: KATs, random chains with
intra-buffer duplicates and partial chunks, re-baselines, long runs of equal
chunks, both hashes."""
import random

import numpy as np
import pytest

import oracle as O
from reproducibility_icpp23_scalable_gpu_deduplication_amd import workloads as W
from test_oracle import KATS, _make_chain


def both(chunk, datas, kind=O.MURMUR3, baselines=(0,)):
    s = O.OracleDeduplicator(chunk, kind)
    p = O.ParallelOracleDeduplicator(chunk, kind)
    for k, x in enumerate(datas):
        mk = k in baselines
        if isinstance(x, np.ndarray):
            a, b = s.checkpoint_array(x, mk), p.checkpoint_array(x, mk)
        else:
            a, b = s.checkpoint(x, mk), p.checkpoint(x, mk)
        assert a == b, f"ckpt {k}: diffs differ ({len(a)} vs {len(b)} bytes)"
        assert s.counts() == p.counts()


@pytest.mark.parametrize("name", sorted(KATS))
def test_kats(name):
    a, b, _, _ = KATS[name]
    both(1, [a.encode(), b.encode()])


@pytest.mark.parametrize("seed", range(40))
def test_random_chains(seed):
    rng = random.Random(5000 + seed)
    C = rng.choice([1, 2, 3, 5, 8, 13, 32, 33, 100, 255, 256, 257, 1000])
    cs = rng.choice([1, 3, 4, 16, 32, 64])
    tail = rng.randrange(0, cs) if cs > 1 else 0
    kind = rng.choice([O.MURMUR3, O.MD5])
    datas = _make_chain(rng, C, cs, 6, tail)
    both(cs, datas, kind, baselines=(0,) if seed % 3 else (0, 3))


@pytest.mark.parametrize("C,cs", [(1 << 12, 128), (5000, 64), (3 << 11, 32)])
def test_medium_chains_with_dups(C, cs):
    datas = W.chain_uniform_changes(C, C * cs, cs, 6, fresh_frac=0.02, copy_frac=0.05, zero_bytes=4 * cs)
    both(cs, datas)


def test_mostly_zero_buffers():
    rng = np.random.default_rng(3)
    C, cs = 6001, 32
    x = np.zeros(C * cs, dtype=np.uint8)
    x[: C * cs // 8] = rng.integers(0, 256, C * cs // 8, dtype=np.uint8)
    datas = [x.copy()]
    for _ in range(4):
        y = datas[-1].copy()
        for st in rng.integers(0, C - 64, 6):
            ln = int(rng.integers(1, 64))
            y[st * cs:(st + ln) * cs] = 0 if rng.random() < 0.5 else rng.integers(0, 256, ln * cs, dtype=np.uint8)
        datas.append(y)
    both(cs, datas)


def test_float_field_md5():
    both(64, W.chain_float_field(3, 1 << 15, 4), O.MD5)
