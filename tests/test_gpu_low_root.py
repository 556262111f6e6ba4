"""The low-root method (DD_METHOD_TREE_LOW_ROOT, TreeLowRootDeduplicator) on
the GPU against the oracle's low-root rule (OPT_LOW_ROOT), bit-exact, through
the C ABI; every checkpoint restarted byte-identical."""
import random

import numpy as np
import pytest

import oracle as O
import reproducibility_icpp23_scalable_gpu_deduplication_amd as dd
from reproducibility_icpp23_scalable_gpu_deduplication_amd import workloads as W
from gpu_util import first_mismatch, gpu_chain, gpu_restart, oracle_chain, to_dev
from test_oracle import _make_chain

pytestmark = pytest.mark.gpu


def check(torch, chunk, datas, kind=dd.HASH_MURMUR3, baselines=(0,), restart_ids=None):
    g, gd, gc = gpu_chain(torch, chunk, datas, kind, baselines, cls=dd.TreeLowRootDeduplicator)
    assert g.method == dd.METHOD_TREE_LOW_ROOT
    o = O.OracleDeduplicator(chunk, kind, options=O.OPT_LOW_ROOT)
    for i, x in enumerate(datas):
        want = o.checkpoint_array(x, i in baselines) if isinstance(x, np.ndarray) else o.checkpoint(x, i in baselines)
        assert gd[i] == want, f"checkpoint {i}: {first_mismatch(gd[i], want)}"
        assert gc[i] is not None
    ids = range(len(datas)) if restart_ids is None else restart_ids
    for k in ids:
        want = bytes(datas[k])
        got = gpu_restart(torch, g, gd, k, len(want))
        assert got == want, f"restart {k}: {first_mismatch(got, want)}"
    return gd


def test_hand_derived(torch_gpu):
    gd = check(torch_gpu, 1, [b"ABCDEFGH", b"XBCDEFXX"])
    lo = O.OracleDeduplicator(1)
    lo.checkpoint(b"ABCDEFGH", True)
    assert gd[1] != lo.checkpoint(b"XBCDEFXX")  # the low-offset diff picks another leaf


@pytest.mark.parametrize("seed", range(12))
def test_small_chains_with_repeats(torch_gpu, seed):
    rng = random.Random(seed)
    C = rng.randrange(1, 300)
    alphabet = b"ab" if seed % 2 else b"abcdefgh"
    chain = [bytes(rng.choice(alphabet) for _ in range(C))]
    for _ in range(4):
        cur = bytearray(chain[-1])
        for _ in range(rng.randrange(1, max(2, C // 3))):
            cur[rng.randrange(C)] = rng.choice(b"abcdefghij")
        chain.append(bytes(cur))
    check(torch_gpu, 1, chain, kind=dd.HASH_MD5 if seed % 3 == 0 else dd.HASH_MURMUR3)


@pytest.mark.parametrize("seed", range(6))
def test_random_chains(torch_gpu, seed):
    rng = random.Random(100 + seed)
    C, cs = rng.randrange(1, 2000), rng.choice([1, 2, 8, 64])
    chain = _make_chain(rng, C, cs, 5)
    check(torch_gpu, cs, chain, baselines=(0, 3) if seed % 2 else (0,))


def test_zero_runs_and_copies_4mib(torch_gpu):
    # many digests repeat (zero regions, copied chunks): large groups whose
    # members sit in all-new subtrees of different heights
    datas = W.chain_uniform_changes(11, 4 << 20, 64, 6, fresh_frac=0.02, copy_frac=0.05, zero_bytes=1 << 16)
    check(torch_gpu, 64, datas, restart_ids=[0, 3, 5])


def test_batched_low_root_handles(torch_gpu):
    torch = torch_gpu
    chains = [W.chain_uniform_changes(40 + i, s, 32, 4, fresh_frac=0.03, copy_frac=0.05, zero_bytes=4096)
              for i, s in enumerate([1 << 18, 3 << 16, 5000 * 32 + 7])]
    hs = [dd.TreeLowRootDeduplicator(32) for _ in chains]
    batch = dd.Batch(hs, graph=True)
    outs = [torch.empty(dd.max_diff_size(c[0].nbytes, 32, 8), dtype=torch.uint8, device="cuda") for c in chains]
    oracles = [O.OracleDeduplicator(32, options=O.OPT_LOW_ROOT) for _ in chains]
    for k in range(4):
        ins = [to_dev(torch, c[k]) for c in chains]
        lens = batch.checkpoint_device([t.data_ptr() for t in ins], [t.numel() for t in ins], [k == 0] * len(ins),
                                       [o.data_ptr() for o in outs], [o.numel() for o in outs])
        torch.cuda.synchronize()
        for i, c in enumerate(chains):
            want = oracles[i].checkpoint_array(c[k], k == 0)
            assert outs[i][:lens[i]].cpu().numpy().tobytes() == want, f"buffer {i} checkpoint {k}"
