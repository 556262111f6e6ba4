"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module, and only as the checker.  The product path
(the HIP library) never imports it.

The oracle restates the reference's serial low-offset labeler
(include/reference_impl.hpp:15-132), its diff serializer
(src/tree_approach.cpp:511-710) and restart (src/tree_approach.cpp:712-1128),
and the list approach (src/list_approach.cpp:23-512); see oracle/dedup_oracle.h.
"""
import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

MURMUR3 = 0
MD5 = 1

FIRST_OCUR, FIXED_DUPL, SHIFT_DUPL, FIRST_DUPL, DONE = 0, 1, 2, 3, 4


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        L.or_hash.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
        L.or_hash.restype = None
        L.or_murmur3_x64_128.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                         ctypes.c_int, ctypes.c_void_p]
        L.or_murmur3_x64_128.restype = None
        L.or_md5.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
        L.or_md5.restype = None
        L.or_create.argtypes = [ctypes.c_uint32, ctypes.c_int]
        L.or_create.restype = ctypes.c_void_p
        L.or_destroy.argtypes = [ctypes.c_void_p]
        L.or_destroy.restype = None
        L.or_set_options.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.or_set_options.restype = None
        L.or_checkpoint.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                    ctypes.POINTER(u8p), ctypes.POINTER(ctypes.c_uint64)]
        L.or_checkpoint.restype = ctypes.c_int
        L.or_current_id.argtypes = [ctypes.c_void_p]
        L.or_current_id.restype = ctypes.c_uint32
        L.or_counts.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64),
                                ctypes.POINTER(ctypes.c_uint64)]
        L.or_counts.restype = None
        L.or_last_labels.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
        L.or_last_labels.restype = ctypes.c_int
        L.or_last_leaf_digests.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
        L.or_last_leaf_digests.restype = ctypes.c_int
        L.or_map_size.argtypes = [ctypes.c_void_p]
        L.or_map_size.restype = ctypes.c_uint64
        L.or_restart.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_uint64),
                                 ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64]
        L.or_restart.restype = ctypes.c_int
        L.or_subtree_leaves.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
        L.or_subtree_leaves.restype = ctypes.c_uint64
        L.or_free.argtypes = [ctypes.c_void_p]
        L.or_free.restype = None
        L.om_create.argtypes = [ctypes.c_uint32, ctypes.c_int]
        L.om_create.restype = ctypes.c_void_p
        L.om_destroy.argtypes = [ctypes.c_void_p]
        L.om_destroy.restype = None
        L.om_checkpoint.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                    ctypes.POINTER(u8p), ctypes.POINTER(ctypes.c_uint64)]
        L.om_checkpoint.restype = ctypes.c_int
        L.om_counts.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64),
                                ctypes.POINTER(ctypes.c_uint64)]
        L.om_counts.restype = None
        L.om_max_threads.argtypes = []
        L.om_max_threads.restype = ctypes.c_int
        L.or_list_create.argtypes = [ctypes.c_uint32, ctypes.c_int]
        L.or_list_create.restype = ctypes.c_void_p
        L.or_list_destroy.argtypes = [ctypes.c_void_p]
        L.or_list_destroy.restype = None
        L.or_list_checkpoint.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                         ctypes.POINTER(u8p), ctypes.POINTER(ctypes.c_uint64)]
        L.or_list_checkpoint.restype = ctypes.c_int
        L.or_list_counts.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64),
                                     ctypes.POINTER(ctypes.c_uint64)]
        L.or_list_counts.restype = None
        L.or_list_current_id.argtypes = [ctypes.c_void_p]
        L.or_list_current_id.restype = ctypes.c_uint32
        L.or_list_restart.argtypes = L.or_restart.argtypes
        L.or_list_restart.restype = ctypes.c_int
        _lib = L
    return _lib


def _buf(data):
    b = bytes(data)
    return b, ctypes.cast(ctypes.c_char_p(b), ctypes.c_void_p)


def hash_bytes(data, kind=MURMUR3):
    b = bytes(data)
    out = ctypes.create_string_buffer(16)
    lib().or_hash(kind, ctypes.c_char_p(b), len(b), out)
    return out.raw


def murmur3_x64_128(data, seed=0, quirk=True):
    b = bytes(data)
    out = ctypes.create_string_buffer(16)
    lib().or_murmur3_x64_128(ctypes.c_char_p(b), len(b), seed, 1 if quirk else 0, out)
    return out.raw


def md5(data):
    b = bytes(data)
    out = ctypes.create_string_buffer(16)
    lib().or_md5(ctypes.c_char_p(b), len(b), out)
    return out.raw


def subtree_leaves(node, num_chunks):
    L = lib()
    n = L.or_subtree_leaves(node, num_chunks, None)
    arr = (ctypes.c_uint32 * max(int(n), 1))()
    L.or_subtree_leaves(node, num_chunks, arr)
    return list(arr[: int(n)])


OPT_EMIT_ROOT_SHIFT = 1
OPT_LOW_ROOT = 2  # the low-root first-occurrence rule (TreeLowRootDeduplicator)


class OracleDeduplicator:
    """Serial CPU oracle with the TreeDeduplicator checkpoint surface.
    `options`: OPT_EMIT_ROOT_SHIFT emits a shifted node 0 (an option of this
    build; the reference emits node 0 only as a first occurrence)."""

    def __init__(self, chunk_size, hash_kind=MURMUR3, options=0):
        self._h = lib().or_create(chunk_size, hash_kind)
        if not self._h:
            raise ValueError("bad chunk size")
        self.chunk_size = chunk_size
        if options:
            lib().or_set_options(self._h, options)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().or_destroy(h)
            self._h = None

    def checkpoint(self, data, make_baseline=False):
        """Returns the diff bytes (header + metadata + data) and advances the id."""
        b = bytes(data) if not isinstance(data, bytes) else data
        out = ctypes.POINTER(ctypes.c_uint8)()
        n = ctypes.c_uint64()
        rc = lib().or_checkpoint(self._h, ctypes.c_char_p(b), len(b), 1 if make_baseline else 0,
                                 ctypes.byref(out), ctypes.byref(n))
        if rc != 0:
            raise RuntimeError(f"or_checkpoint failed: {rc}")
        res = ctypes.string_at(out, n.value)
        lib().or_free(out)
        return res

    def checkpoint_array(self, arr, make_baseline=False):
        """Same as checkpoint() on a contiguous numpy uint8 array (no copy)."""
        out = ctypes.POINTER(ctypes.c_uint8)()
        n = ctypes.c_uint64()
        rc = lib().or_checkpoint(self._h, ctypes.c_void_p(arr.ctypes.data), arr.nbytes,
                                 1 if make_baseline else 0, ctypes.byref(out), ctypes.byref(n))
        if rc != 0:
            raise RuntimeError(f"or_checkpoint failed: {rc}")
        res = ctypes.string_at(out, n.value)
        lib().or_free(out)
        return res

    @property
    def current_id(self):
        return lib().or_current_id(self._h)

    def counts(self):
        f = ctypes.c_uint64()
        s = ctypes.c_uint64()
        lib().or_counts(self._h, ctypes.byref(f), ctypes.byref(s))
        return f.value, s.value

    def num_first_ocur(self):
        return self.counts()[0]

    def num_shift_dupl(self):
        return self.counts()[1]

    def labels(self, num_chunks):
        n = 2 * num_chunks - 1
        out = ctypes.create_string_buffer(n)
        if lib().or_last_labels(self._h, out, n) != 0:
            raise RuntimeError("no labels")
        return out.raw

    def leaf_digests(self, num_chunks):
        out = ctypes.create_string_buffer(16 * num_chunks)
        if lib().or_last_leaf_digests(self._h, out, 16 * num_chunks) != 0:
            raise RuntimeError("no digests")
        return out.raw

    def map_size(self):
        return lib().or_map_size(self._h)


class ParallelOracleDeduplicator:
    """The OpenMP restatement (dedup_omp.c): the serial oracle's diff bytes on
    omp_get_max_threads() host threads (set OMP_NUM_THREADS before the first
    call).  Used as the CPU baseline of the benchmark."""

    def __init__(self, chunk_size, hash_kind=MURMUR3):
        self._h = lib().om_create(chunk_size, hash_kind)
        if not self._h:
            raise ValueError("bad chunk size")
        self.chunk_size = chunk_size

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().om_destroy(h)
            self._h = None

    def _run(self, ptr, n, make_baseline):
        out = ctypes.POINTER(ctypes.c_uint8)()
        m = ctypes.c_uint64()
        rc = lib().om_checkpoint(self._h, ptr, n, 1 if make_baseline else 0, ctypes.byref(out), ctypes.byref(m))
        if rc != 0:
            raise RuntimeError(f"om_checkpoint failed: {rc}")
        res = ctypes.string_at(out, m.value)
        lib().or_free(out)
        return res

    def checkpoint(self, data, make_baseline=False):
        b = bytes(data) if not isinstance(data, bytes) else data
        return self._run(ctypes.c_char_p(b), len(b), make_baseline)

    def checkpoint_array(self, arr, make_baseline=False):
        return self._run(ctypes.c_void_p(arr.ctypes.data), arr.nbytes, make_baseline)

    def counts(self):
        f = ctypes.c_uint64()
        s = ctypes.c_uint64()
        lib().om_counts(self._h, ctypes.byref(f), ctypes.byref(s))
        return f.value, s.value

    @staticmethod
    def threads():
        return lib().om_max_threads()


class OracleListDeduplicator:
    """Serial CPU oracle of the list approach (ListDeduplicator,
    src/list_approach.cpp) with the same checkpoint surface."""

    def __init__(self, chunk_size, hash_kind=MURMUR3):
        self._h = lib().or_list_create(chunk_size, hash_kind)
        if not self._h:
            raise ValueError("bad chunk size")
        self.chunk_size = chunk_size

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().or_list_destroy(h)
            self._h = None

    def _run(self, ptr, n, make_baseline):
        out = ctypes.POINTER(ctypes.c_uint8)()
        m = ctypes.c_uint64()
        rc = lib().or_list_checkpoint(self._h, ptr, n, 1 if make_baseline else 0, ctypes.byref(out),
                                      ctypes.byref(m))
        if rc != 0:
            raise RuntimeError(f"or_list_checkpoint failed: {rc}")
        res = ctypes.string_at(out, m.value)
        lib().or_free(out)
        return res

    def checkpoint(self, data, make_baseline=False):
        b = bytes(data) if not isinstance(data, bytes) else data
        return self._run(ctypes.c_char_p(b), len(b), make_baseline)

    def checkpoint_array(self, arr, make_baseline=False):
        return self._run(ctypes.c_void_p(arr.ctypes.data), arr.nbytes, make_baseline)

    @property
    def current_id(self):
        return lib().or_list_current_id(self._h)

    def counts(self):
        f = ctypes.c_uint64()
        s = ctypes.c_uint64()
        lib().or_list_counts(self._h, ctypes.byref(f), ctypes.byref(s))
        return f.value, s.value


def _restart_call(fn, chain, ckpt_id, datalen):
    keep = [bytes(c) for c in chain]
    ptrs = (ctypes.c_void_p * len(keep))(*[ctypes.cast(ctypes.c_char_p(b), ctypes.c_void_p) for b in keep])
    lens = (ctypes.c_uint64 * len(keep))(*[len(b) for b in keep])
    out = ctypes.create_string_buffer(datalen)
    return fn(ptrs, lens, len(keep), ckpt_id, out, datalen), out.raw


def restart_list(chain, ckpt_id, datalen):
    """Rebuild checkpoint `ckpt_id` from a chain of list-approach diffs."""
    rc, out = _restart_call(lib().or_list_restart, chain, ckpt_id, datalen)
    if rc != 0:
        raise RuntimeError(f"or_list_restart failed: {rc}")
    return out


def restart_list_rc(chain, ckpt_id, datalen):
    return _restart_call(lib().or_list_restart, chain, ckpt_id, datalen)[0]


def restart(chain, ckpt_id, datalen):
    """Rebuild checkpoint `ckpt_id` from a list of diffs indexed by id."""
    keep = [bytes(c) for c in chain]
    ptrs = (ctypes.c_void_p * len(keep))(*[ctypes.cast(ctypes.c_char_p(b), ctypes.c_void_p) for b in keep])
    lens = (ctypes.c_uint64 * len(keep))(*[len(b) for b in keep])
    out = ctypes.create_string_buffer(datalen)
    rc = lib().or_restart(ptrs, lens, len(keep), ckpt_id, out, datalen)
    if rc != 0:
        raise RuntimeError(f"or_restart failed: {rc}")
    return out.raw


def restart_rc(chain, ckpt_id, datalen):
    keep = [bytes(c) for c in chain]
    ptrs = (ctypes.c_void_p * len(keep))(*[ctypes.cast(ctypes.c_char_p(b), ctypes.c_void_p) for b in keep])
    lens = (ctypes.c_uint64 * len(keep))(*[len(b) for b in keep])
    out = ctypes.create_string_buffer(datalen)
    return lib().or_restart(ptrs, lens, len(keep), ckpt_id, out, datalen)
