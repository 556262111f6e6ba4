"""File drivers (bin/dedup_chkpt_files, bin/restart_chkpt_files): the
.hashtree.incr_chkpt (list mode: .hashlist.incr_chkpt) files are the
oracle's diffs byte for byte, and the restart driver prints the oracle's
Murmur3 digest of the original data."""
import os
import subprocess

import pytest

import oracle as O
import reproducibility_icpp23_scalable_gpu_deduplication_amd as dd
from reproducibility_icpp23_scalable_gpu_deduplication_amd import workloads as W
from gpu_util import first_mismatch, oracle_chain

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(dd.LIB_PATH), "bin")


MODES = {
    "tree": ("--run-tree-chkpt", "--run-tree-low-offset-chkpt", ".hashtree.incr_chkpt", "Hashtree digest:       ",
             O.OracleDeduplicator),
    "list": ("--run-list-chkpt", "--run-list-chkpt", ".hashlist.incr_chkpt", "Hashlist digest:       ",
             O.OracleListDeduplicator),
}


def digest_line(label, data):
    d = O.hash_bytes(data)
    hexs = "".join(f"{b:02X}" + (" " if i % 4 == 3 else "") for i, b in enumerate(d))
    return label + hexs


def run(args, cwd):
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, f"{args[0]} failed: {r.stderr}"
    return r.stdout


@pytest.mark.parametrize("mode", ["tree", "list"])
@pytest.mark.parametrize("size,chunk", [((1 << 20) + 77, 128), (3000 * 64, 64)])
def test_file_drivers_round_trip(torch_gpu, tmp_path, size, chunk, mode):
    flag, rflag, suffix, label, ocls = MODES[mode]
    chain = W.chain_uniform_changes(5 + chunk, size, chunk, 5, fresh_frac=0.02, copy_frac=0.04, zero_bytes=4096)
    files = []
    for k, x in enumerate(chain):
        p = tmp_path / f"ckpt{k}.dat"
        x.tofile(p)
        files.append(str(p))
    run([os.path.join(BIN, "dedup_chkpt_files"), str(chunk), str(len(files)), flag, *files], tmp_path)
    od, _ = oracle_chain(chunk, chain, cls=ocls)
    for k, f in enumerate(files):
        got = open(f + suffix, "rb").read()
        assert got == od[k], f"ckpt {k}: {first_mismatch(got, od[k])}"
    rows = open(tmp_path / f"ckpt0.dat.chunk_size.{chunk}.csv").read().splitlines()
    assert rows[0].startswith("Approach,") and rows[1].startswith(("Tree," if mode == "tree" else "List,") + "0,")
    for k in (0, 2, len(files) - 1):
        out = run([os.path.join(BIN, "restart_chkpt_files"), str(k), str(len(files)), "2", str(chunk),
                   rflag, *files], tmp_path)
        assert out.count(digest_line(label, chain[k].tobytes())) == 2, out
    # Without the original data file the restored length comes from the diff header.
    os.remove(files[3])
    out = run([os.path.join(BIN, "restart_chkpt_files"), "3", str(len(files)), "1", str(chunk), flag,
               *files], tmp_path)
    assert digest_line(label, chain[3].tobytes()) in out
