"""MI355X-native tree-approach incremental-checkpoint deduplicator.

The product is libdedup_hip.so: hand-written HIP kernels for gfx950 behind the
C ABI in include/dedup_c.h, plus the C++ TreeDeduplicator facade in
include/tree_deduplicator.hpp.  This package only binds that library for
Python callers (tests, bench).
"""
from ._lib import (  # noqa: F401
    DD_ECORRUPT, DD_EHIP, DD_EINVAL, DD_ELIMIT, DD_ENOMEM, DD_ESTATE, DD_OK, EXPORTED, HASH_MD5, HASH_MURMUR3,
    LEAF_MAPPINGS, LIB_PATH, METHOD_LIST, METHOD_TREE, METHOD_TREE_LOW_ROOT, NOALLOC_MAX_CHUNKS, NOALLOC_SCRATCH_BYTES, OPT_DENSE_BASELINE,
    OPT_EMIT_ROOT_SHIFT,
    Batch, DedupError, ListDeduplicator, Server, TreeDeduplicator, TreeLowRootDeduplicator, load, max_diff_size,
    strerror,
)
