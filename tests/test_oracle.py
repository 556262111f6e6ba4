"""CPU oracle pinned against checks that do not come from the oracle itself.

- Murmur3 structure: the published SMHasher verification value of textbook
  MurmurHash3_x64_128 (0x6384BA69) with the truncation switched off.
- Murmur3 truncation: rotl64 in the reference returns uint32_t
  (include/kokkos_murmur3.hpp:30-31); checked against an independent
  big-integer restatement below (parity with the reference binary itself is
  unpinned: the reference cannot be built here).
- MD5: hashlib and the RFC 1321 appendix A.5 test suite.
- Labels / region counts: the reference's ten known-answer tests
  (test/test_case_01.cpp ... test_case_10.cpp).
- Diff bytes: hand-derived from the reference's collect_diff layout
  (src/tree_approach.cpp:511-710) for KATs 04, 05, 06 and 10.
- Restart: byte-identical rebuild of every checkpoint of random chains.
"""
import hashlib
import random
import struct

import numpy as np
import pytest

import oracle as O

KATS = {
    "01": ("ABCDEFGH", "IJKLMNOP", 1, 0),
    "02": ("ABCDEFGH", "ABCDEFGH", 0, 0),
    "03": ("ABCDEFGH", "ABCDEIJK", 2, 0),
    "04": ("ABCDEFGH", "IJABCDEK", 2, 3),
    "05": ("ABCDEFGH", "ABCIABCJ", 2, 2),
    "06": ("ABCDEFGH", "EFGBCDIJ", 1, 4),
    "07": ("ABCDEFGH", "IJKLABIH", 1, 2),
    "08": ("MNXYOPQR", "IJXYIJAB", 2, 1),
    "09": ("MNXYOPQR", "IJXYIJIB", 2, 2),
    "10": ("MNXYOPQR", "IJXYIJJI", 1, 3),
}


def diff_bytes(ref_id, ckpt_id, datalen, chunk, first, table, shifts, data):
    """Assembles a diff from its parts (layout of src/tree_approach.cpp:622-696)."""
    out = struct.pack("<IIQIIII", ref_id, ckpt_id, datalen, chunk, len(first), len(table), len(shifts))
    out += b"".join(struct.pack("<I", f) for f in first)
    out += b"".join(struct.pack("<II", c, n) for c, n in table)
    out += b"".join(struct.pack("<II", n, p) for n, p in shifts)
    return out + data


# Derived by hand from the serial semantics (include/reference_impl.hpp:43-119)
# on 8 one-byte chunks (leaves are nodes 7..14).
HAND_DIFFS = {
    # IJABCDEK: F roots {3 (IJ), 14 (K)}; shifted 4=AB->3, 5=CD->4, 13=E->11.
    "04": diff_bytes(0, 1, 8, 1, [3, 14], [(0, 3)], [(4, 3), (5, 4), (13, 11)], b"IJK"),
    # ABCIABCJ: F roots {10 (I), 14 (J)}; shifted 5=AB->3, 13=C->9.
    "05": diff_bytes(0, 1, 8, 1, [10, 14], [(0, 2)], [(5, 3), (13, 9)], b"IJ"),
    # EFGBCDIJ: F root {6 (IJ)}; shifted 3=EF->5, 5=CD->4, 9=G->13, 10=B->8.
    "06": diff_bytes(0, 1, 8, 1, [6], [(0, 4)], [(3, 5), (5, 4), (9, 13), (10, 8)], b"IJ"),
    # IJXYIJJI: F root {3 (IJ)}; shifted from checkpoint 1 itself:
    # 5=IJ->3, 13=J->8, 14=I->7.
    "10": diff_bytes(0, 1, 8, 1, [3], [(1, 3)], [(5, 3), (13, 8), (14, 7)], b"IJ"),
}


def test_murmur3_structure_smhasher_verification():
    hashes = b"".join(O.murmur3_x64_128(bytes(range(i)), seed=256 - i, quirk=False) for i in range(256))
    final = O.murmur3_x64_128(hashes, seed=0, quirk=False)
    assert struct.unpack("<I", final[:4])[0] == 0x6384BA69


M64 = (1 << 64) - 1


def _murmur_bigint(data, seed=0):
    """Independent restatement from the reference text with C conversion
    rules spelled out: (x << r | x >> (64 - r)) is computed in uint64_t,
    then returned as uint32_t and widened back on assignment."""

    def rot(x, r):
        return (((x << r) & M64) | (x >> (64 - r))) & 0xFFFFFFFF

    def fmix(k):
        k ^= k >> 33
        k = (k * 0xFF51AFD7ED558CCD) & M64
        k ^= k >> 33
        k = (k * 0xC4CEB9FE1A85EC53) & M64
        return k ^ (k >> 33)

    c1, c2 = 0x87C37B91114253D5, 0x4CF5AD432745937F
    h1 = h2 = seed
    n = len(data)
    for i in range(n // 16):
        k1, k2 = struct.unpack_from("<QQ", data, 16 * i)
        k1 = (rot((k1 * c1) & M64, 31) * c2) & M64
        h1 ^= k1
        h1 = (rot(h1, 27) + h2) & M64
        h1 = (h1 * 5 + 0x52DCE729) & M64
        k2 = (rot((k2 * c2) & M64, 33) * c1) & M64
        h2 ^= k2
        h2 = (rot(h2, 31) + h1) & M64
        h2 = (h2 * 5 + 0x38495AB5) & M64
    tail = data[16 * (n // 16):]
    if len(tail) > 8:
        k2 = int.from_bytes(tail[8:], "little")
        h2 ^= (rot((k2 * c2) & M64, 33) * c1) & M64
    if len(tail) > 0:
        k1 = int.from_bytes(tail[:8], "little")
        h1 ^= (rot((k1 * c1) & M64, 31) * c2) & M64
    h1 ^= n
    h2 ^= n
    h1 = (h1 + h2) & M64
    h2 = (h2 + h1) & M64
    h1, h2 = fmix(h1), fmix(h2)
    h1 = (h1 + h2) & M64
    h2 = (h2 + h1) & M64
    return struct.pack("<QQ", h1, h2)


def test_murmur3_truncating_rotate_matches_bigint_restatement():
    rng = random.Random(7)
    for n in list(range(0, 70)) + [127, 128, 129, 512, 4096]:
        data = bytes(rng.getrandbits(8) for _ in range(n))
        assert O.murmur3_x64_128(data) == _murmur_bigint(data), n
        assert O.hash_bytes(data) == _murmur_bigint(data)


def test_murmur3_quirk_changes_digest():
    # The truncation is real: textbook and reference digests differ.
    d = b"0123456789abcdef0123"
    assert O.murmur3_x64_128(d, quirk=True) != O.murmur3_x64_128(d, quirk=False)
    # ... except for the empty input, where no rotate runs.
    assert O.murmur3_x64_128(b"", quirk=True) == O.murmur3_x64_128(b"", quirk=False) == bytes(16)


def test_md5_rfc1321_suite_and_hashlib():
    suite = {
        b"": "d41d8cd98f00b204e9800998ecf8427e",
        b"a": "0cc175b9c0f1b6a831c399e269772661",
        b"abc": "900150983cd24fb0d6963f7d28e17f72",
        b"message digest": "f96b697d7cb7938d525a2f31aaf161d0",
        b"abcdefghijklmnopqrstuvwxyz": "c3fcd3d76192e4007dfb496cca67e13b",
        b"ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789": "d174ab98d277d9f5a5611c2c9f419d9f",
        b"1234567890" * 8: "57edf4a22be3c955ac49da2e2107b67a",
    }
    for msg, hx in suite.items():
        assert O.md5(msg).hex() == hx
        assert O.hash_bytes(msg, O.MD5).hex() == hx
    rng = random.Random(3)
    for n in list(range(0, 140)) + [4096, 10000]:
        d = bytes(rng.getrandbits(8) for _ in range(n))
        assert O.md5(d) == hashlib.md5(d).digest()


@pytest.mark.parametrize("name", sorted(KATS))
def test_kat_counts_and_restart(name):
    a, b, ef, es = KATS[name]
    d = O.OracleDeduplicator(1)
    c0 = d.checkpoint(a.encode(), True)
    c1 = d.checkpoint(b.encode(), False)
    assert d.counts() == (ef, es)
    assert O.restart([c0, c1], 0, 8) == a.encode()
    assert O.restart([c0, c1], 1, 8) == b.encode()


@pytest.mark.parametrize("name", sorted(HAND_DIFFS))
def test_kat_diff_bytes_hand_derived(name):
    a, b, _, _ = KATS[name]
    d = O.OracleDeduplicator(1)
    d.checkpoint(a.encode(), True)
    assert d.checkpoint(b.encode(), False) == HAND_DIFFS[name]


def test_baseline_diff_layout():
    # A baseline of all-distinct chunks is one first-occurrence region rooted
    # at node 0 holding every chunk (src/tree_approach.cpp:392-399).
    d = O.OracleDeduplicator(1)
    assert d.checkpoint(b"ABCDEFGH", True) == diff_bytes(0, 0, 8, 1, [0], [], [], b"ABCDEFGH")


def test_subtree_leaves_power_of_two_matches_contiguous_range():
    # include/kokkos_merkle_tree.hpp:102-142 for C = 2^k.
    for C in (1, 2, 4, 8, 64):
        N = 2 * C - 1
        for v in range(N):
            lo, hi = v, v
            while 2 * lo + 1 < N:
                lo, hi = 2 * lo + 1, 2 * hi + 2
            assert O.subtree_leaves(v, C) == list(range(lo, hi + 1))


def test_subtree_leaves_general_is_inorder_leaf_set():
    # Brute force: the left-to-right leaves of each subtree by recursion.
    for C in range(1, 40):
        N = 2 * C - 1

        def inorder(v):
            if v >= C - 1:
                return [v]
            return inorder(2 * v + 1) + inorder(2 * v + 2)

        for v in range(N):
            assert O.subtree_leaves(v, C) == inorder(v), (C, v)


def _make_chain(rng, C, cs, nckpt, tail=0):
    n = C * cs - (cs - tail if tail else 0)
    cur = bytearray(rng.getrandbits(8) for _ in range(n))
    # seed a few intra-buffer duplicates into the baseline too
    for _ in range(rng.randrange(0, 3)):
        if C >= 2:
            s, t = rng.randrange(C - 1), rng.randrange(C - 1)
            cur[t * cs:(t + 1) * cs] = cur[s * cs:(s + 1) * cs]
    chain = [bytes(cur)]
    for _ in range(nckpt - 1):
        for _ in range(rng.randrange(0, 4)):
            op = rng.randrange(3)
            if C < 2:
                op = 0
            if op == 0:
                c = rng.randrange(C)
                end = min((c + 1) * cs, n)
                for i in range(c * cs, end):
                    cur[i] = rng.getrandbits(8)
            elif op == 1:
                s, t = rng.randrange(C - 1), rng.randrange(C - 1)
                ln = rng.randrange(1, 4)
                for q in range(ln):
                    if s + q < C - 1 and t + q < C - 1:
                        cur[(t + q) * cs:(t + q + 1) * cs] = chain[-1][(s + q) * cs:(s + q + 1) * cs]
            else:
                c = rng.randrange(C - 1)
                cur[c * cs:(c + 1) * cs] = bytes(cs)
        chain.append(bytes(cur))
    return chain


@pytest.mark.parametrize("seed", range(40))
def test_roundtrip_random_chains(seed):
    rng = random.Random(seed)
    C = rng.choice([1, 2, 3, 5, 7, 8, 13, 16, 31, 32, 33])
    cs = rng.choice([1, 2, 3, 4, 8, 16])
    tail = rng.randrange(0, cs) if cs > 1 else 0
    data = _make_chain(rng, C, cs, 6, tail)
    d = O.OracleDeduplicator(cs, rng.choice([O.MURMUR3, O.MD5]))
    diffs = [d.checkpoint(x, i == 0) for i, x in enumerate(data)]
    for k in range(len(data)):
        assert O.restart(diffs, k, len(data[k])) == data[k], (seed, k)


def test_roundtrip_rebaseline_mid_chain():
    rng = random.Random(11)
    data = _make_chain(rng, 16, 4, 8)
    d = O.OracleDeduplicator(4)
    diffs = [d.checkpoint(x, i in (0, 4)) for i, x in enumerate(data)]
    assert struct.unpack_from("<I", diffs[5])[0] == 4  # ref_id after re-baseline
    for k in range(len(data)):
        assert O.restart(diffs, k, len(data[k])) == data[k]


def test_every_changed_leaf_is_covered_once():
    # Each FIRST or SHIFT leaf lies under exactly one emitted root.
    rng = random.Random(5)
    for trial in range(30):
        C = rng.choice([8, 16, 13, 21])
        data = _make_chain(rng, C, 2, 4)
        d = O.OracleDeduplicator(2)
        for i, x in enumerate(data):
            diff = d.checkpoint(x, i == 0)
            lab = d.labels(C)
            if lab[0] == O.SHIFT_DUPL:
                continue
            _, _, _, _, F, P, S = struct.unpack_from("<IIQIIII", diff)
            roots = list(struct.unpack_from(f"<{F}I", diff, 32))
            roots += [struct.unpack_from("<I", diff, 32 + 4 * F + 8 * P + 8 * k)[0] for k in range(S)]
            covered = {}
            for r in roots:
                for leaf in O.subtree_leaves(r, C):
                    covered[leaf] = covered.get(leaf, 0) + 1
            for leaf in range(C - 1, 2 * C - 1):
                want = 0 if lab[leaf] == O.FIXED_DUPL else 1
                assert covered.get(leaf, 0) == want, (trial, i, leaf)


def test_shift_root_zero_is_not_emitted():
    # The serial reference only emits node 0 when it is FIRST
    # (include/reference_impl.hpp:105-106): a whole buffer that repeats an
    # older checkpoint leaves no metadata and restarts to its predecessor.
    d = O.OracleDeduplicator(1)
    c0 = d.checkpoint(b"ABCDEFGH", True)
    c1 = d.checkpoint(b"IJKLMNOP", False)
    c2 = d.checkpoint(b"ABCDEFGH", False)
    assert struct.unpack_from("<IIQIIII", c2)[4:] == (0, 0, 0)
    assert O.restart([c0, c1, c2], 2, 8) == b"IJKLMNOP"


def test_shift_root_zero_emitted_with_the_option():
    # Option of this build: node 0 is emitted as a shifted root, so the
    # repeated buffer restarts to its own bytes.  Hand-derived diff: header
    # (ref 0, ckpt 2, datalen 8, chunk 1, F 0, P 1, S 1), prior table
    # {(0, 1)}, shifted entry (node 0, source node 0 of checkpoint 0).
    d = O.OracleDeduplicator(1, options=O.OPT_EMIT_ROOT_SHIFT)
    c0 = d.checkpoint(b"ABCDEFGH", True)
    c1 = d.checkpoint(b"IJKLMNOP", False)
    c2 = d.checkpoint(b"ABCDEFGH", False)
    assert c2 == struct.pack("<IIQIIII", 0, 2, 8, 1, 0, 1, 1) + struct.pack("<II", 0, 1) + struct.pack("<II", 0, 0)
    assert O.restart([c0, c1, c2], 2, 8) == b"ABCDEFGH"
    # the option changes nothing else: the same chain without a repeat
    plain = O.OracleDeduplicator(1)
    opt = O.OracleDeduplicator(1, options=O.OPT_EMIT_ROOT_SHIFT)
    for i, x in enumerate([b"ABCDEFGH", b"ABCDEFGX", b"XBCDEFGH", b"ABCDABCD"]):
        assert plain.checkpoint(x, i == 0) == opt.checkpoint(x, i == 0)


def test_restart_errors():
    d = O.OracleDeduplicator(4)
    c0 = d.checkpoint(bytes(range(64)), True)
    assert O.restart_rc([c0], 0, 63) == 1  # length mismatch
    assert O.restart_rc([c0[:20]], 0, 64) == 3  # truncated
    bad = bytearray(c0)
    struct.pack_into("<I", bad, 32, 10**6)  # node id out of range
    assert O.restart_rc([bytes(bad)], 0, 64) == 3


def test_map_sizes_follow_inserts():
    # Every leaf digest and every FIRST interior digest is inserted once.
    rng = np.random.default_rng(1)
    data = rng.integers(0, 256, 64 * 16, dtype=np.uint8)
    d = O.OracleDeduplicator(16)
    d.checkpoint_array(data, True)
    assert d.map_size() == 2 * 64 - 1
