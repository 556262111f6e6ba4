"""C ABI boundary checks that need no GPU: the library loads and exports
every entry point include/dedup_c.h declares; pure-host helpers behave."""
import ctypes
import os
import re
import subprocess

import reproducibility_icpp23_scalable_gpu_deduplication_amd as dd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "dedup_c.h")).read()
    return sorted(set(re.findall(r"^\s*(?:[a-z_][a-z0-9_ ]*[\s*]+)(dd_[a-z_]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = dd.load()
    decl = declared_symbols()
    assert len(decl) >= 14
    assert sorted(dd.EXPORTED) == decl
    for name in decl:
        assert hasattr(L, name), name


def test_exports_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", dd.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for name in declared_symbols():
        assert name in names, name
    # The C++ facade (include/tree_deduplicator.hpp) is in the same library.
    demangled = subprocess.run(["nm", "-DC", "--defined-only", dd.LIB_PATH], capture_output=True, text=True,
                               check=True).stdout
    assert "dd::TreeDeduplicator::checkpoint(" in demangled
    assert "dd::TreeDeduplicator::restart(" in demangled


def test_header_struct_is_32_bytes():
    assert ctypes.sizeof(dd._lib.dd_header) == 32


def test_max_diff_size():
    # header + 8 B of metadata per chunk + prior table + every chunk's data
    assert dd.max_diff_size(8, 1, 1) == 32 + 8 * 8 + 16 + 8
    assert dd.max_diff_size(1 << 20, 128, 9) == 32 + 8 * 8192 + 80 + (1 << 20)
    assert dd.max_diff_size(130, 128, 0) == 32 + 16 + 8 + 256


def test_strerror():
    assert dd.strerror(0) == "ok"
    assert "corrupt" in dd.strerror(dd.DD_ECORRUPT)
    assert dd.strerror(99) == "unknown status"


def test_no_fallback_when_library_missing(tmp_path, monkeypatch):
    from reproducibility_icpp23_scalable_gpu_deduplication_amd import _lib

    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    try:
        _lib.load()
        raise AssertionError("load() must raise without the HIP library")
    except ImportError:
        pass


def test_file_drivers_built_and_reject_other_methods(tmp_path):
    # The drivers exit with status 2 on the full and basic methods (not built
    # here) and on an unknown mode, before using a GPU.
    import os
    import subprocess

    bindir = os.path.join(os.path.dirname(dd.LIB_PATH), "bin")
    for name in ("dedup_chkpt_files", "dedup_chkpt_files_ranks", "restart_chkpt_files", "data_generation"):
        assert os.access(os.path.join(bindir, name), os.X_OK)
    for mode, msg in (("--run-full-chkpt", "unsupported"), ("--run-basic-chkpt", "unsupported"),
                      ("--run-nothing", "incorrect")):
        r = subprocess.run([os.path.join(bindir, "dedup_chkpt_files"), "128", "1", mode, "x"],
                           cwd=tmp_path, capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and msg in r.stderr, (mode, r.stderr)


def test_method_constants_match_header():
    import re

    hdr = open(os.path.join(ROOT, "include", "dedup_c.h")).read()
    assert int(re.search(r"DD_METHOD_TREE = (\d+)", hdr).group(1)) == dd.METHOD_TREE
    assert int(re.search(r"DD_METHOD_LIST = (\d+)", hdr).group(1)) == dd.METHOD_LIST
