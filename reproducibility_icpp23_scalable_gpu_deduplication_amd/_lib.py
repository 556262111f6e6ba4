"""ctypes binding of libdedup_hip.so (the C ABI declared in include/dedup_c.h).

There is no fallback: if the HIP library is missing or fails to load, every
entry point raises.  Device buffers are passed as raw HIP device addresses
(for example ``tensor.data_ptr()`` of a ROCm torch tensor).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdedup_hip.so")

DD_OK, DD_EINVAL, DD_ENOMEM, DD_ECORRUPT, DD_EHIP, DD_ESTATE, DD_ELIMIT = range(7)
# Limits of dd_restart_device_noalloc (include/dedup_c.h).
NOALLOC_SCRATCH_BYTES = 256 << 20
NOALLOC_MAX_CHUNKS = 10 << 20

HASH_MURMUR3 = 0
HASH_MD5 = 1
OPT_EMIT_ROOT_SHIFT = 1
OPT_DENSE_BASELINE = 2
METHOD_TREE = 0
METHOD_LIST = 1
METHOD_TREE_LOW_ROOT = 2

# Every symbol include/dedup_c.h declares.
EXPORTED = [
    "dd_create", "dd_create_method", "dd_get_method", "dd_set_options", "dd_get_options", "dd_destroy", "dd_checkpoint", "dd_checkpoint_header", "dd_checkpoint_device",
    "dd_max_diff_size", "dd_counts", "dd_get_ids", "dd_set_current_id", "dd_get_timers",
    "dd_restart", "dd_restart_device", "dd_restart_device_noalloc", "dd_batch_create", "dd_batch_destroy", "dd_batch_use_graph",
    "dd_batch_checkpoint_device", "dd_server_create", "dd_server_destroy", "dd_server_submit_checkpoint",
    "dd_server_submit_restart", "dd_server_wait", "dd_server_set_batching", "dd_server_usage", "dd_now",
    "dd_set_profiling",
    "dd_get_stage_times", "dd_get_leaf_mapping", "dd_batch_get_leaf_mapping", "dd_get_stage_budget",
    "dd_batch_get_stage_budget", "dd_batch_set_profiling",
    "dd_batch_get_stage_times", "dd_get_ckpt_stats", "dd_get_device_bytes", "dd_buf_free", "dd_strerror",
]
LEAF_MAPPINGS = ["direct", "deep", "lds64", "lds128", "lds256", "short", "dma", "dma1", "lds128p"]
STAGES = ["leaf_hash", "insert_leaves", "first_forest", "insert_interior", "dup_forest", "roots_group",
          "meta_gather"]


class DedupError(RuntimeError):
    def __init__(self, status, what):
        super().__init__(f"{what}: {strerror(status)} ({status})")
        self.status = status


class dd_buf(ctypes.Structure):
    _fields_ = [("data", ctypes.POINTER(ctypes.c_uint8)), ("len", ctypes.c_uint64)]


class dd_header(ctypes.Structure):
    _fields_ = [("ref_id", ctypes.c_uint32), ("chkpt_id", ctypes.c_uint32), ("datalen", ctypes.c_uint64),
                ("chunk_size", ctypes.c_uint32), ("num_first_ocur", ctypes.c_uint32),
                ("num_prior_chkpts", ctypes.c_uint32), ("num_shift_dupl", ctypes.c_uint32)]


class dd_ckpt_stats(ctypes.Structure):
    _fields_ = [("chunks", ctypes.c_uint64), ("changed", ctypes.c_uint64), ("first_chunks", ctypes.c_uint64),
                ("diff_bytes", ctypes.c_uint64), ("baseline", ctypes.c_uint32), ("leaf_copy", ctypes.c_uint32)]


class dd_request_result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("kind", ctypes.c_uint32), ("out_len", ctypes.c_uint64),
                ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_end", ctypes.c_double),
                ("group", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


_lib = None


def load():
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    P = ctypes.POINTER
    sig = {
        "dd_create": ([u32, i32, vp, P(vp)], i32),
        "dd_create_method": ([u32, i32, i32, vp, P(vp)], i32),
        "dd_get_method": ([vp, P(i32)], i32),
        "dd_set_options": ([vp, u32], i32),
        "dd_get_options": ([vp, P(u32)], i32),
        "dd_destroy": ([vp], i32),
        "dd_checkpoint": ([vp, vp, u64, i32, P(dd_buf)], i32),
        "dd_checkpoint_header": ([vp, P(dd_header), vp, u64, i32, P(dd_buf)], i32),
        "dd_checkpoint_device": ([vp, vp, u64, i32, vp, u64, P(u64)], i32),
        "dd_max_diff_size": ([u64, u32, u32], u64),
        "dd_counts": ([vp, P(u64), P(u64)], i32),
        "dd_get_ids": ([vp, P(u32), P(u32)], i32),
        "dd_set_current_id": ([vp, u32], i32),
        "dd_get_timers": ([vp, P(ctypes.c_double), P(ctypes.c_double)], i32),
        "dd_restart": ([vp, P(dd_buf), u32, u32, vp, u64], i32),
        "dd_restart_device": ([vp, P(vp), P(u64), u32, u32, vp, u64], i32),
        "dd_restart_device_noalloc": ([vp, P(vp), P(u64), u32, u32, vp, u64], i32),
        "dd_batch_create": ([P(vp), u32, P(vp)], i32),
        "dd_batch_destroy": ([vp], i32),
        "dd_batch_use_graph": ([vp, i32], i32),
        "dd_batch_checkpoint_device": ([vp, P(vp), P(u64), P(i32), P(vp), P(u64), P(u64)], i32),
        "dd_server_create": ([u64, u32, P(vp)], i32),
        "dd_server_destroy": ([vp], i32),
        "dd_server_submit_checkpoint": ([vp, vp, vp, u64, i32, vp, u64, P(u64)], i32),
        "dd_server_submit_restart": ([vp, vp, P(vp), P(u64), u32, u32, vp, u64, P(u64)], i32),
        "dd_server_wait": ([vp, u64, P(dd_request_result)], i32),
        "dd_server_set_batching": ([vp, u32], i32),
        "dd_server_usage": ([vp, P(u64), P(u64)], i32),
        "dd_now": ([], ctypes.c_double),
        "dd_set_profiling": ([vp, i32], i32),
        "dd_get_stage_times": ([vp, P(ctypes.c_double)], i32),
        "dd_get_leaf_mapping": ([vp, P(i32)], i32),
        "dd_batch_get_leaf_mapping": ([vp, P(i32)], i32),
        "dd_get_stage_budget": ([vp, P(u64)], i32),
        "dd_batch_get_stage_budget": ([vp, P(u64)], i32),
        "dd_batch_set_profiling": ([vp, i32], i32),
        "dd_batch_get_stage_times": ([vp, P(ctypes.c_double)], i32),
        "dd_get_ckpt_stats": ([vp, P(dd_ckpt_stats)], i32),
        "dd_get_device_bytes": ([vp, P(u64)], i32),
        "dd_buf_free": ([P(dd_buf)], None),
        "dd_strerror": ([i32], ctypes.c_char_p),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def strerror(status):
    return load().dd_strerror(status).decode()


def _check(rc, what):
    if rc != DD_OK:
        raise DedupError(rc, what)


def max_diff_size(length, chunk_size, ckpt_id):
    return load().dd_max_diff_size(length, chunk_size, ckpt_id)


class TreeDeduplicator:
    """Python view of the C ABI, shaped like the reference's TreeDeduplicator
    (include/tree_approach.hpp:16-205).  Data arguments are device addresses
    (int) plus lengths; diffs come back as ``bytes``."""

    METHOD = METHOD_TREE

    def __init__(self, chunk_size, hash_kind=HASH_MURMUR3, stream=None):
        self._h = None
        L = load()
        h = ctypes.c_void_p()
        _check(L.dd_create_method(chunk_size, hash_kind, self.METHOD, stream, ctypes.byref(h)), "dd_create_method")
        self._h = h
        self.chunk_size = chunk_size

    @property
    def options(self):
        """DD_OPT_* flags (OPT_EMIT_ROOT_SHIFT, OPT_DENSE_BASELINE)."""
        o = ctypes.c_uint32()
        _check(load().dd_get_options(self._h, ctypes.byref(o)), "dd_get_options")
        return o.value

    @options.setter
    def options(self, flags):
        _check(load().dd_set_options(self._h, int(flags)), "dd_set_options")

    @property
    def method(self):
        m = ctypes.c_int()
        _check(load().dd_get_method(self._h, ctypes.byref(m)), "dd_get_method")
        return m.value

    def close(self):
        if getattr(self, "_h", None):
            load().dd_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # -- checkpoint -------------------------------------------------------
    def checkpoint(self, dev_ptr, length, make_baseline=False):
        """dd_checkpoint: diff bytes on the host; advances the checkpoint id."""
        b = dd_buf()
        _check(load().dd_checkpoint(self._h, dev_ptr, length, int(make_baseline), ctypes.byref(b)), "dd_checkpoint")
        try:
            return ctypes.string_at(b.data, b.len)
        finally:
            load().dd_buf_free(ctypes.byref(b))

    def checkpoint_header(self, dev_ptr, length, make_baseline=False):
        """dd_checkpoint_header: (header dict, diff bytes); id not advanced."""
        b = dd_buf()
        hd = dd_header()
        _check(load().dd_checkpoint_header(self._h, ctypes.byref(hd), dev_ptr, length, int(make_baseline),
                                           ctypes.byref(b)), "dd_checkpoint_header")
        try:
            data = ctypes.string_at(b.data, b.len)
        finally:
            load().dd_buf_free(ctypes.byref(b))
        return {k: getattr(hd, k) for k, _ in dd_header._fields_}, data

    def checkpoint_device(self, dev_ptr, length, out_ptr, out_cap, make_baseline=False):
        """dd_checkpoint_device: diff left in device memory; returns its length."""
        n = ctypes.c_uint64()
        _check(load().dd_checkpoint_device(self._h, dev_ptr, length, int(make_baseline), out_ptr, out_cap,
                                           ctypes.byref(n)), "dd_checkpoint_device")
        return n.value

    # -- restart ----------------------------------------------------------
    def restart(self, chain, ckpt_id, out_ptr, length):
        keep = [bytes(c) for c in chain]
        arr = (dd_buf * len(keep))()
        for i, b in enumerate(keep):
            arr[i].data = ctypes.cast(ctypes.c_char_p(b), ctypes.POINTER(ctypes.c_uint8))
            arr[i].len = len(b)
        _check(load().dd_restart(self._h, arr, len(keep), ckpt_id, out_ptr, length), "dd_restart")

    def restart_rc(self, chain, ckpt_id, out_ptr, length):
        keep = [bytes(c) for c in chain]
        arr = (dd_buf * len(keep))()
        for i, b in enumerate(keep):
            arr[i].data = ctypes.cast(ctypes.c_char_p(b), ctypes.POINTER(ctypes.c_uint8))
            arr[i].len = len(b)
        return load().dd_restart(self._h, arr, len(keep), ckpt_id, out_ptr, length)

    def restart_device(self, dev_ptrs, lens, ckpt_id, out_ptr, length):
        n = len(dev_ptrs)
        ptrs = (ctypes.c_void_p * n)(*dev_ptrs)
        ls = (ctypes.c_uint64 * n)(*lens)
        _check(load().dd_restart_device(self._h, ptrs, ls, n, ckpt_id, out_ptr, length), "dd_restart_device")

    def restart_device_noalloc(self, dev_ptrs, lens, ckpt_id, out_ptr, length):
        """dd_restart_device_noalloc: the restart in the library's static
        scratch (no device allocation); returns the dd_status."""
        n = len(dev_ptrs)
        ptrs = (ctypes.c_void_p * n)(*dev_ptrs)
        ls = (ctypes.c_uint64 * n)(*lens)
        return load().dd_restart_device_noalloc(self._h, ptrs, ls, n, ckpt_id, out_ptr, length)

    # -- state ------------------------------------------------------------
    def counts(self):
        f, s = ctypes.c_uint64(), ctypes.c_uint64()
        _check(load().dd_counts(self._h, ctypes.byref(f), ctypes.byref(s)), "dd_counts")
        return f.value, s.value

    def num_first_ocur(self):
        return self.counts()[0]

    def num_shift_dupl(self):
        return self.counts()[1]

    @property
    def current_id(self):
        c, b = ctypes.c_uint32(), ctypes.c_uint32()
        _check(load().dd_get_ids(self._h, ctypes.byref(c), ctypes.byref(b)), "dd_get_ids")
        return c.value

    @current_id.setter
    def current_id(self, v):
        _check(load().dd_set_current_id(self._h, v), "dd_set_current_id")

    @property
    def baseline_id(self):
        c, b = ctypes.c_uint32(), ctypes.c_uint32()
        _check(load().dd_get_ids(self._h, ctypes.byref(c), ctypes.byref(b)), "dd_get_ids")
        return b.value

    def set_profiling(self, level=1):
        """0 off, 1 leaf-hash kernel events only, 2 events around every stage."""
        _check(load().dd_set_profiling(self._h, int(level)), "dd_set_profiling")

    def stage_times(self):
        """GPU milliseconds per stage of the last checkpoint (profiling on)."""
        ms = (ctypes.c_double * len(STAGES))()
        _check(load().dd_get_stage_times(self._h, ms), "dd_get_stage_times")
        return dict(zip(STAGES, list(ms)))

    def ckpt_stats(self):
        """Sizes of the last completed checkpoint: chunks, changed leaves
        (every leaf at a baseline), first-occurrence chunks copied, diff
        bytes, baseline flag, and whether the leaf kernel wrote the data
        section (leaf_copy)."""
        st = dd_ckpt_stats()
        _check(load().dd_get_ckpt_stats(self._h, ctypes.byref(st)), "dd_get_ckpt_stats")
        return {k: getattr(st, k) for k, _ in dd_ckpt_stats._fields_}

    def device_bytes(self):
        """Device bytes the handle holds (state, scratch, restart buffers)."""
        n = ctypes.c_uint64()
        _check(load().dd_get_device_bytes(self._h, ctypes.byref(n)), "dd_get_device_bytes")
        return n.value

    def leaf_mapping(self):
        """Name of the leaf kernel mapping of the last checkpoint."""
        m = ctypes.c_int()
        _check(load().dd_get_leaf_mapping(self._h, ctypes.byref(m)), "dd_get_leaf_mapping")
        return LEAF_MAPPINGS[m.value]

    def stage_budget(self):
        """Batch-wide workgroup budget of the incremental stages at the last checkpoint."""
        n = ctypes.c_uint64()
        _check(load().dd_get_stage_budget(self._h, ctypes.byref(n)), "dd_get_stage_budget")
        return n.value

    def timers(self):
        t = (ctypes.c_double * 4)()
        r = (ctypes.c_double * 2)()
        _check(load().dd_get_timers(self._h, t, r), "dd_get_timers")
        return list(t), list(r)


class TreeLowRootDeduplicator(TreeDeduplicator):
    """TreeLowRootDeduplicator (include/tree_approach.hpp:207-218): the tree
    method with low-root first occurrences (DD_METHOD_TREE_LOW_ROOT): the
    rule of dedup_low_root_ref (include/reference_impl.hpp:265-472) made
    deterministic.  Tree diffs and restarts."""

    METHOD = METHOD_TREE_LOW_ROOT


class ListDeduplicator(TreeDeduplicator):
    """ListDeduplicator (include/list_approach.hpp:21-200,
    src/list_approach.cpp): one digest per chunk, first occurrences and
    shifted duplicates named by chunk index."""

    METHOD = METHOD_LIST


class Batch:
    """dd_batch: one checkpoint of every handle per call, each stage launched
    once over all buffers (optionally replayed from a captured hipGraph)."""

    def __init__(self, dedups, graph=True):
        L = load()
        self._dedups = list(dedups)
        arr = (ctypes.c_void_p * len(self._dedups))(*[d._h.value for d in self._dedups])
        h = ctypes.c_void_p()
        _check(L.dd_batch_create(arr, len(self._dedups), ctypes.byref(h)), "dd_batch_create")
        self._h = h
        _check(L.dd_batch_use_graph(self._h, int(graph)), "dd_batch_use_graph")

    def close(self):
        if getattr(self, "_h", None):
            load().dd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def leaf_mapping(self):
        m = ctypes.c_int()
        _check(load().dd_batch_get_leaf_mapping(self._h, ctypes.byref(m)), "dd_batch_get_leaf_mapping")
        return LEAF_MAPPINGS[m.value]

    def stage_budget(self):
        n = ctypes.c_uint64()
        _check(load().dd_batch_get_stage_budget(self._h, ctypes.byref(n)), "dd_batch_get_stage_budget")
        return n.value

    def set_profiling(self, level=1):
        """0 off, 1 leaf-hash kernel events, 2 events around every stage."""
        _check(load().dd_batch_set_profiling(self._h, int(level)), "dd_batch_set_profiling")

    def stage_times(self):
        ms = (ctypes.c_double * len(STAGES))()
        _check(load().dd_batch_get_stage_times(self._h, ms), "dd_batch_get_stage_times")
        return dict(zip(STAGES, list(ms)))

    def checkpoint_device(self, dev_ptrs, lens, make_baseline, out_ptrs, out_caps):
        n = len(self._dedups)
        vp, u64 = ctypes.c_void_p, ctypes.c_uint64
        d = (vp * n)(*dev_ptrs)
        ln = (u64 * n)(*lens)
        mk = (ctypes.c_int * n)(*[int(bool(x)) for x in make_baseline])
        o = (vp * n)(*out_ptrs)
        oc = (u64 * n)(*out_caps)
        res = (u64 * n)()
        _check(load().dd_batch_checkpoint_device(self._h, d, ln, mk, o, oc, res), "dd_batch_checkpoint_device")
        return list(res)


class Server:
    """dd_server: checkpoint and restart requests over many handles, admitted
    in arrival order by execution slots and an HBM budget; checkpoints
    admitted together run as one batched job table, the rest on the handles'
    streams.  ``submit_*`` return a ticket; ``wait`` returns the request's
    result as a dict (status, kind, out_len, t_submit, t_start, t_end, group;
    times in seconds of the steady clock ``now()``, group the number of
    checkpoints batched with it)."""

    def __init__(self, hbm_budget_bytes, slots=8):
        h = ctypes.c_void_p()
        _check(load().dd_server_create(int(hbm_budget_bytes), int(slots), ctypes.byref(h)), "dd_server_create")
        self._h = h
        self._keep = {}

    def close(self):
        if getattr(self, "_h", None):
            load().dd_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        self.close()

    def submit_checkpoint(self, dedup, dev_ptr, length, out_ptr, out_cap, make_baseline=False):
        t = ctypes.c_uint64()
        _check(load().dd_server_submit_checkpoint(self._h, dedup._h, dev_ptr, length, int(bool(make_baseline)),
                                                  out_ptr, out_cap, ctypes.byref(t)), "dd_server_submit_checkpoint")
        return t.value

    def submit_restart(self, dedup, dev_ptrs, lens, ckpt_id, out_ptr, length):
        n = len(dev_ptrs)
        ptrs = (ctypes.c_void_p * n)(*dev_ptrs)
        ls = (ctypes.c_uint64 * n)(*lens)
        t = ctypes.c_uint64()
        _check(load().dd_server_submit_restart(self._h, dedup._h, ptrs, ls, n, ckpt_id, out_ptr, length,
                                               ctypes.byref(t)), "dd_server_submit_restart")
        return t.value

    def wait(self, ticket):
        r = dd_request_result()
        _check(load().dd_server_wait(self._h, ticket, ctypes.byref(r)), "dd_server_wait")
        return {k: getattr(r, k) for k, _ in dd_request_result._fields_ if k != "pad"}

    def set_batching(self, max_group):
        """Checkpoint requests admitted together run as one batched job table,
        at most ``max_group`` of them (1: every request alone)."""
        _check(load().dd_server_set_batching(self._h, int(max_group)), "dd_server_set_batching")

    def usage(self):
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        _check(load().dd_server_usage(self._h, ctypes.byref(a), ctypes.byref(b)), "dd_server_usage")
        return a.value, b.value

    @staticmethod
    def now():
        return load().dd_now()
