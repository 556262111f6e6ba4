"""The low-root first-occurrence rule (TreeLowRootDeduplicator,
include/reference_impl.hpp:265-472) in the oracle.

This build restates the reference's low-root labeler as a deterministic rule
(oracle/dedup_oracle.c, low_root_select): among the new leaves with one digest
(changed, and no map entry from an earlier checkpoint), the first occurrence
is the leaf whose highest all-new ancestor has the smallest node id, ties to
the lowest leaf; everything else is the low-offset labeling.  Pinned here by:
- a diff derived by hand from that rule and the reference's diff layout;
- a brute force of the selection on small random chains: new leaves from the
  leaf digests and an independently kept set of the map's digests, all-new
  subtrees from their explicit leaf sets (or_subtree_leaves), and the winner
  of every group compared with the oracle's labels;
- identical diffs to the low-offset rule whenever no digest repeats among
  the new leaves (the two rules differ only in that choice);
- byte-identical restarts of every checkpoint.
"""
import random
import struct

import oracle as O
from test_oracle import _make_chain


def test_hand_derived_low_root_diff():
    # Baseline "ABCDEFGH", then "XBCDEFXX" (one byte per chunk, 8 chunks,
    # leaves 7..14).  X is new at chunks 0, 6, 7 (leaves 7, 13, 14).  Node 6
    # (leaves 13, 14) is all new; node 3 (leaves 7, 8) is not, leaf 8 being
    # unchanged.  Highest new ancestors: leaf 7 -> 7, leaves 13 and 14 -> 6,
    # so leaf 13 (chunk 6) is the first occurrence; leaves 7 and 14 shift to
    # it.  Low-offset picks leaf 7 instead.
    lr = O.OracleDeduplicator(1, options=O.OPT_LOW_ROOT)
    lo = O.OracleDeduplicator(1)
    c0 = lr.checkpoint(b"ABCDEFGH", True)
    lo.checkpoint(b"ABCDEFGH", True)
    got = lr.checkpoint(b"XBCDEFXX")
    want = (struct.pack("<IIQIIII", 0, 1, 8, 1, 1, 1, 2) + struct.pack("<I", 13) + struct.pack("<II", 1, 2)
            + struct.pack("<IIII", 7, 13, 14, 13) + b"X")
    assert got == want
    assert lo.checkpoint(b"XBCDEFXX") == (struct.pack("<IIQIIII", 0, 1, 8, 1, 1, 1, 2) + struct.pack("<I", 7)
                                          + struct.pack("<II", 1, 2) + struct.pack("<IIII", 13, 7, 14, 7) + b"X")
    assert O.restart([c0, got], 1, 8) == b"XBCDEFXX"


def _interior_digest_fn(C, leaf_digests):
    cache = {}

    def dig(v):
        if v >= C - 1:
            return leaf_digests[v - (C - 1)]
        if v not in cache:
            cache[v] = O.hash_bytes(dig(2 * v + 1) + dig(2 * v + 2))
        return cache[v]

    return dig


def brute_force_winners(C, prev_digests, digests, map_digests, baseline):
    """Leaf -> winner of its digest group among the new leaves, by the rule's
    definition with explicit leaf sets."""
    first = C - 1
    new = set()
    for i in range(C):
        leaf = first + i
        changed = baseline or digests[i] != prev_digests[i]
        if baseline and digests[i] == prev_digests[i]:
            changed = False  # equal to the all-zero previous tree: fixed
        if changed and digests[i] not in map_digests:
            new.add(leaf)
    all_new = {}
    for v in range(2 * C - 1):
        all_new[v] = all(x in new for x in O.subtree_leaves(v, C))
    groups = {}
    for leaf in sorted(new):
        t = leaf
        while t > 0 and all_new[(t - 1) // 2]:
            t = (t - 1) // 2
        groups.setdefault(digests[leaf - first], []).append((t, leaf))
    return {leaf: min(g)[1] for g in groups.values() for _, leaf in g}


def test_selection_matches_brute_force():
    rng = random.Random(20261204)
    differs = 0
    for trial in range(200):
        C = rng.randrange(1, 40)
        nckpt = rng.randrange(2, 6)
        # two-symbol chunks of 1 byte repeat often, so groups have several members
        chain = [bytes(rng.choice(b"ab" if trial % 2 else b"abcd") for _ in range(C))]
        for _ in range(1, nckpt):
            cur = bytearray(chain[-1])
            for _ in range(rng.randrange(1, max(2, C // 2))):
                cur[rng.randrange(C)] = rng.choice(b"abcdefgh")
            chain.append(bytes(cur))
        d = O.OracleDeduplicator(1, options=O.OPT_LOW_ROOT)
        map_digests = set()
        prev = [b"\0" * 16] * C
        diffs = []
        for k, x in enumerate(chain):
            diffs.append(d.checkpoint(x, k == 0))
            ld = d.leaf_digests(C)
            digests = [ld[16 * i:16 * i + 16] for i in range(C)]
            lab = d.labels(C)
            if k == 0:
                map_digests = set()
            want = brute_force_winners(C, prev, digests, map_digests, k == 0)
            lowest = {}
            for leaf, w in want.items():
                lowest[w] = min(lowest.get(w, leaf), leaf)
            differs += sum(1 for w, m in lowest.items() if m != w)  # groups where low-offset would pick another leaf
            for leaf, w in want.items():
                exp = O.FIRST_OCUR if leaf == w else O.SHIFT_DUPL
                assert lab[leaf] == exp, (trial, k, leaf, w)
            # the map gains every changed leaf's digest (every leaf at a
            # baseline) and the first-occurrence interior nodes' digests
            dig = _interior_digest_fn(C, digests)
            for i in range(C):
                if k == 0 or digests[i] != prev[i]:
                    map_digests.add(digests[i])
            for v in range(C - 1):
                if lab[v] == O.FIRST_OCUR:
                    map_digests.add(dig(v))
            prev = digests
        for k in range(nckpt):
            assert O.restart(diffs, k, C) == chain[k], (trial, k)
    assert differs > 20  # the rule was exercised where it departs from low-offset


def test_equal_to_low_offset_without_repeated_new_digests():
    rng = random.Random(77)
    compared = []
    for trial in range(60):
        C, cs = rng.randrange(1, 200), rng.choice([4, 8, 16])
        chain = _make_chain(rng, C, cs, 4)
        # keep only chains whose new leaves never repeat a digest within a
        # checkpoint: the low-offset diff then already has the low-root one
        lo, lr = O.OracleDeduplicator(cs), O.OracleDeduplicator(cs, options=O.OPT_LOW_ROOT)
        for k, x in enumerate(chain):
            a, b = lo.checkpoint(x, k == 0), lr.checkpoint(x, k == 0)
            chunks = [x[i * cs:(i + 1) * cs] for i in range((len(x) + cs - 1) // cs)]
            if len(set(chunks)) < len(chunks):
                break  # from here on the two chains' maps may differ
            assert a == b, (trial, k)
            compared.append((trial, k))
    assert len(compared) > 30


def test_random_chains_restart_with_low_root():
    rng = random.Random(5)
    for trial in range(30):
        C, cs = rng.randrange(1, 300), rng.choice([1, 2, 8, 64])
        chain = _make_chain(rng, C, cs, 5)
        d = O.OracleDeduplicator(cs, options=O.OPT_LOW_ROOT)
        diffs = [d.checkpoint(x, k == 0) for k, x in enumerate(chain)]
        for k in range(len(chain)):
            assert O.restart(diffs, k, len(chain[k])) == bytes(chain[k]), (trial, k)
