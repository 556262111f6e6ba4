"""The data generator (bin/data_generation) and the multi-rank driver
(bin/dedup_chkpt_files_ranks) on the GPU: every generator mode perturbs the
previous checkpoint exactly as its definition says, series are
reproducible, and the multi-rank driver's files equal the oracle's diffs per
rank, in one process or split over processes."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import reproducibility_icpp23_scalable_gpu_deduplication_amd as dd
from reproducibility_icpp23_scalable_gpu_deduplication_amd import workloads as W
from gpu_util import first_mismatch, oracle_chain

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(dd.LIB_PATH), "bin")


def run(args, cwd, env=None):
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, f"{args[0]} failed: {r.stderr}"
    return r.stdout


def generate(tmp_path, length, ck, mode, n, prefix="gen"):
    out = run([os.path.join(BIN, "data_generation"), str(length), str(ck), mode, str(n), str(tmp_path / prefix)],
              tmp_path)
    files = [np.fromfile(tmp_path / f"{prefix}.{i}.chkpt", dtype=np.uint8) for i in range(ck)]
    offs = [int(ln.split(":")[1]) for ln in out.splitlines() if "ffset:" in ln]
    return files, offs


LEN = (1 << 20) + 13


@pytest.mark.parametrize("mode,n", [("R", 0), ("B", 4096), ("C", 5000), ("I", 0), ("S", 300), ("Z", 7777),
                                    ("W", 4096)])
def test_generator_modes(torch_gpu, tmp_path, mode, n):
    files, offs = generate(tmp_path, LEN, 4, mode, n)
    assert all(f.size == LEN for f in files)
    # Checkpoint 0 is random bytes: every value occurs about LEN/256 times.
    hist = np.bincount(files[0], minlength=256)
    assert hist.min() > LEN / 256 * 0.8 and hist.max() < LEN / 256 * 1.2
    for i in range(1, 4):
        a, b = files[i - 1], files[i]
        diff = np.nonzero(a != b)[0]
        if mode == "R":
            assert diff.size > LEN * 0.99
        elif mode == "B":
            assert np.array_equal(a[:n], b[:n]) and diff.min() >= n and diff.size > (LEN - n) * 0.99
        elif mode == "C":
            o = offs[i - 1]
            assert diff.size > n * 0.99 and diff.min() >= o and diff.max() < o + n
        elif mode == "I":
            assert diff.size == 0
        elif mode == "S":
            assert 0 < diff.size <= n
        elif mode == "Z":
            o = offs[i - 1]
            assert not b[o:o + n].any() and np.array_equal(np.delete(a, np.s_[o:o + n]), np.delete(b, np.s_[o:o + n]))
        elif mode == "W":
            o = offs[i - 1]
            assert o % 128 == 0 and o >= n and o + n <= LEN
            assert np.array_equal(b[:n], a[o:o + n]) and np.array_equal(b[o:o + n], a[:n])
            rest = np.ones(LEN, bool)
            rest[:n] = rest[o:o + n] = False
            assert np.array_equal(a[rest], b[rest])


def test_generator_is_reproducible(torch_gpu, tmp_path):
    f1, _ = generate(tmp_path, LEN, 3, "S", 1000, "one")
    f2, _ = generate(tmp_path, LEN, 3, "S", 1000, "two")
    assert all(np.array_equal(a, b) for a, b in zip(f1, f2))


def test_generated_series_through_file_drivers(torch_gpu, tmp_path):
    # The reference's workflow: generate, deduplicate the files, restart.
    files, _ = generate(tmp_path, 3 << 20, 5, "C", 64 << 10)
    names = [str(tmp_path / f"gen.{i}.chkpt") for i in range(5)]
    run([os.path.join(BIN, "dedup_chkpt_files"), "128", "5", "--run-tree-chkpt", *names], tmp_path)
    od, _ = oracle_chain(128, files)
    for k, nm in enumerate(names):
        assert open(nm + ".hashtree.incr_chkpt", "rb").read() == od[k]
    d = O.hash_bytes(files[4].tobytes())
    want = "Hashtree digest:       " + "".join(f"{b:02X}" + (" " if i % 4 == 3 else "") for i, b in enumerate(d))
    assert want in run([os.path.join(BIN, "restart_chkpt_files"), "4", "5", "1", "128", "--run-tree-chkpt", *names],
                       tmp_path)


def rank_files(tmp_path, ranks, nck, chunk):
    chains, paths = {}, {}
    for r in ranks:
        size = (1 << 19) + 4096 * r + 17 * r
        chains[r] = W.chain_uniform_changes(100 + r, size, chunk, nck, fresh_frac=0.02, copy_frac=0.04, zero_bytes=4096)
        paths[r] = []
        for k, x in enumerate(chains[r]):
            p = tmp_path / f"field.Rank{r}.{k}.dat"
            x.tofile(p)
            paths[r].append(str(p))
    return chains, paths


@pytest.mark.parametrize("mode", ["tree", "list"])
def test_rank_driver_one_process(torch_gpu, tmp_path, mode):
    chunk, nck, ranks = 128, 4, [0, 1, 2, 10]
    chains, paths = rank_files(tmp_path, ranks, nck, chunk)
    flag, suffix, ocls = {"tree": ("--run-tree-chkpt", ".hashtree.incr_chkpt", O.OracleDeduplicator),
                          "list": ("--run-list-chkpt", ".hashlist.incr_chkpt", O.OracleListDeduplicator)}[mode]
    # Files interleaved across ranks on the command line.
    args = [p for k in range(nck) for r in ranks for p in [paths[r][k]]]
    env = {k: v for k, v in os.environ.items() if k not in ("RANK", "WORLD_SIZE")}
    out = run([os.path.join(BIN, "dedup_chkpt_files_ranks"), str(chunk), str(nck), flag, *args], tmp_path, env)
    assert "process 0 of 1: 4 ranks" in out
    for r in ranks:
        od, _ = oracle_chain(chunk, chains[r], cls=ocls)
        for k in range(nck):
            got = open(paths[r][k] + suffix, "rb").read()
            assert got == od[k], f"rank {r} ckpt {k}: {first_mismatch(got, od[k])}"
        rows = open(tmp_path / f"field.Rank{r}.1.dat.chunk_size.{chunk}.csv").read().splitlines()
        assert rows[1].startswith(("Tree" if mode == "tree" else "List") + ",1,")


def test_rank_driver_split_over_processes(torch_gpu, tmp_path):
    chunk, nck, ranks = 128, 3, [0, 1, 10, 11]
    chains, paths = rank_files(tmp_path, ranks, nck, chunk)
    args = [p for r in ranks for p in paths[r]]
    drv = os.path.join(BIN, "dedup_chkpt_files_ranks")
    out0 = run([drv, str(chunk), str(nck), "--run-tree-chkpt", "--rank", "0", "--ranks", "2", *args], tmp_path)
    env = dict(os.environ, RANK="1", WORLD_SIZE="2")
    out1 = run([drv, str(chunk), str(nck), "--run-tree-chkpt", *args], tmp_path, env)
    assert "process 0 of 2: 2 ranks" in out0 and "process 1 of 2: 2 ranks" in out1
    assert "Rank10" in out0 and "Rank1." not in out0 and "Rank11" in out1
    for r in ranks:
        od, _ = oracle_chain(chunk, chains[r])
        for k in range(nck):
            assert open(paths[r][k] + ".hashtree.incr_chkpt", "rb").read() == od[k], (r, k)
