"""Checkpoint tools that need no GPU for these checks: the data generator's
window offsets follow the reference's random-engine calls, and the
multi-rank driver rejects bad rank layouts before using a GPU."""
import os
import subprocess

import pytest

import reproducibility_icpp23_scalable_gpu_deduplication_amd as dd

BIN = os.path.join(os.path.dirname(dd.LIB_PATH), "bin")

# The reference's offset draws, include/data_generation.hpp:50-52 (C and Z:
# one uniform draw over [0, len - n] per checkpoint) and :85-91 (W: redraw
# until the window is 128-aligned, past [0, n) and inside the buffer), with
# std::default_random_engine seeded 1931 (src/data_generation.cpp:48).
PROBE = r"""
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
int main(int argc, char** argv) {
  uint64_t len = strtoull(argv[1], 0, 0), n = strtoull(argv[3], 0, 0);
  unsigned ck = atoi(argv[2]);
  char mode = argv[4][0];
  std::default_random_engine generator(1931);
  for (unsigned i = 1; i < ck; i++) {
    std::uniform_int_distribution<uint64_t> distribution(0, len - n);
    if (mode == 'W') {
      uint64_t b = distribution(generator);
      while ((b < n) || (b + n > len) || ((b / 128) * 128 != b)) b = distribution(generator);
      printf("B offset: %lu\n", (unsigned long)b);
    } else {
      printf("Offset: %lu\n", (unsigned long)distribution(generator));
    }
  }
}
"""


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    d = tmp_path_factory.mktemp("probe")
    src, exe = d / "probe.cpp", d / "probe"
    src.write_text(PROBE)
    subprocess.run(["g++", "-O1", "-o", str(exe), str(src)], check=True)
    return str(exe)


@pytest.mark.parametrize("length,ck,n,mode", [(1 << 30, 6, 4096, "C"), (1 << 20, 9, 100, "Z"),
                                                (65536, 5, 1000, "W"), (123457, 4, 7, "C")])
def test_generator_offsets_follow_reference_draws(probe, tmp_path, length, ck, n, mode):
    want = subprocess.run([probe, str(length), str(ck), str(n), mode], capture_output=True, text=True,
                          check=True).stdout.split("\n")
    out = subprocess.run([os.path.join(BIN, "data_generation"), str(length), str(ck), mode, str(n),
                          str(tmp_path / "x"), "--offsets-only"], capture_output=True, text=True, check=True).stdout
    got = [ln for ln in out.split("\n") if "ffset:" in ln]
    assert got == [w for w in want if w]
    assert not list(tmp_path.iterdir())  # nothing written


def test_generator_rejects_bad_arguments(tmp_path):
    gen = os.path.join(BIN, "data_generation")
    for args in (["100", "3", "Q", "1"], ["100", "3", "W", "60"], ["0", "3", "R", "1"], ["100", "3", "C", "101"]):
        r = subprocess.run([gen, *args, str(tmp_path / "x")], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2, args


def test_rank_driver_rejects_bad_layouts(tmp_path):
    drv = os.path.join(BIN, "dedup_chkpt_files_ranks")
    cases = [
        (["128", "2", "--run-tree-chkpt", "a.Rank0.x", "b.Rank1.x"], "expected 2"),
        (["128", "1", "--run-tree-chkpt", "a.Rank0.x", "b.norank"], "names no rank"),
        (["128", "1", "--run-tree-chkpt", "--rank", "1", "--ranks", "2", "a.Rank0.x"], "no files"),
        (["128", "1", "--run-basic-chkpt", "a.Rank0.x"], "unsupported"),
    ]
    for args, msg in cases:
        r = subprocess.run([drv, *args], cwd=tmp_path, capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and msg in r.stderr, (args, r.stderr)
