"""dd_restart_device_noalloc (the allocation-free restart) against
dd_restart_device through the C ABI: byte-equal restored buffers on two calls
in a row (the second on a smaller buffer, a list handle and a partial last
chunk), the DD_ELIMIT rejection past the chunk limit, and the bytes of the
chain itself (the oracle's and the input's)."""
import numpy as np
import pytest

import oracle as O
import reproducibility_icpp23_scalable_gpu_deduplication_amd as dd
from reproducibility_icpp23_scalable_gpu_deduplication_amd import workloads as W
from gpu_util import to_dev

pytestmark = pytest.mark.gpu


def device_chain(torch, cls, chunk, datas):
    g = cls(chunk)
    diffs = []
    for i, x in enumerate(datas):
        t = to_dev(torch, x)
        cap = dd.max_diff_size(t.numel(), chunk, i)
        o = torch.empty(cap, dtype=torch.uint8, device="cuda")
        n = g.checkpoint_device(t.data_ptr(), t.numel(), o.data_ptr(), cap, i == 0)
        diffs.append(o[:n].clone())
        del t, o
    torch.cuda.synchronize()
    return g, diffs


def restore_both(torch, g, diffs, k, length):
    ptrs = [d.data_ptr() for d in diffs]
    lens = [d.numel() for d in diffs]
    a = torch.full((length,), 7, dtype=torch.uint8, device="cuda")
    b = torch.full((length,), 9, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    g.restart_device(ptrs, lens, k, a.data_ptr(), length)
    rc = g.restart_device_noalloc(ptrs, lens, k, b.data_ptr(), length)
    torch.cuda.synchronize()
    assert rc == dd.DD_OK, dd.strerror(rc)
    return a.cpu().numpy(), b.cpu().numpy()


def test_noalloc_matches_restart_device_twice_second_smaller(torch_gpu):
    torch = torch_gpu
    # first call: 8 MiB, 128-byte chunks, 6 checkpoints
    big = W.chain_uniform_changes(5, 8 << 20, 128, 6, zero_bytes=1 << 16)
    g, diffs = device_chain(torch, dd.TreeDeduplicator, 128, big)
    for k in (5, 0, 3):
        a, b = restore_both(torch, g, diffs, k, 8 << 20)
        assert np.array_equal(a, b), f"checkpoint {k}: noalloc differs from restart_device"
        assert np.array_equal(b, big[k])
    # second call: a smaller list-method chain with a partial last chunk
    size = (1 << 20) + 37
    small = W.chain_uniform_changes(6, size, 64, 4, zero_bytes=4096)
    g2, diffs2 = device_chain(torch, dd.ListDeduplicator, 64, small)
    for k in (3, 1):
        a, b = restore_both(torch, g2, diffs2, k, size)
        assert np.array_equal(a, b), f"list checkpoint {k}: noalloc differs from restart_device"
        assert np.array_equal(b, small[k])
    # and the oracle's diffs, restored by the same scratch
    o = O.OracleDeduplicator(64)
    host = [o.checkpoint_array(x, i == 0) for i, x in enumerate(small)]
    dev = [to_dev(torch, np.frombuffer(h, np.uint8)) for h in host]
    out = torch.zeros(size, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    g3 = dd.TreeDeduplicator(64)
    rc = g3.restart_device_noalloc([d.data_ptr() for d in dev], [d.numel() for d in dev], 3, out.data_ptr(), size)
    torch.cuda.synchronize()
    assert rc == dd.DD_OK
    assert np.array_equal(out.cpu().numpy(), small[3])


def test_noalloc_rejects_more_chunks_than_its_scratch(torch_gpu):
    torch = torch_gpu
    chunk = 64
    length = (dd.NOALLOC_MAX_CHUNKS + 1) * chunk
    # a baseline diff header naming that length, with no regions: the call
    # must refuse it from the header alone, before any kernel reads regions
    hdr = np.zeros(32, np.uint8)
    hdr[8:16] = np.frombuffer(np.uint64(length).tobytes(), np.uint8)
    hdr[16:20] = np.frombuffer(np.uint32(chunk).tobytes(), np.uint8)
    d = to_dev(torch, hdr)
    out = torch.empty(16, dtype=torch.uint8, device="cuda")
    g = dd.TreeDeduplicator(chunk)
    rc = g.restart_device_noalloc([d.data_ptr()], [32], 0, out.data_ptr(), length)
    assert rc == dd.DD_ELIMIT
    assert g.restart_device_noalloc([d.data_ptr()], [32], 1, out.data_ptr(), length) == dd.DD_EINVAL
