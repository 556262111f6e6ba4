"""Helpers shared by the GPU tests: device buffers through torch (ROCm) and
GPU-vs-oracle chain runs."""
import numpy as np

import oracle as O
import reproducibility_icpp23_scalable_gpu_deduplication_amd as dd


def to_dev(torch, data):
    arr = np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray) else data
    t = torch.from_numpy(arr.copy()).to("cuda")
    torch.cuda.synchronize()
    return t


def gpu_chain(torch, chunk, datas, kind=dd.HASH_MURMUR3, baselines=(0,), cls=dd.TreeDeduplicator):
    g = cls(chunk, kind)
    diffs, counts = [], []
    for i, x in enumerate(datas):
        t = to_dev(torch, x)
        diffs.append(g.checkpoint(t.data_ptr(), t.numel(), i in baselines))
        counts.append(g.counts())
        del t
    return g, diffs, counts


def oracle_chain(chunk, datas, kind=O.MURMUR3, baselines=(0,), cls=O.OracleDeduplicator):
    o = cls(chunk, kind)
    diffs, counts = [], []
    for i, x in enumerate(datas):
        if isinstance(x, np.ndarray):
            diffs.append(o.checkpoint_array(x, i in baselines))
        else:
            diffs.append(o.checkpoint(x, i in baselines))
        counts.append(o.counts())
    return diffs, counts


def gpu_restart(torch, g, diffs, k, length):
    out = torch.zeros(length, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    g.restart(diffs, k, out.data_ptr(), length)
    return out.cpu().numpy().tobytes()


def first_mismatch(a, b):
    if len(a) != len(b):
        return f"length {len(a)} != {len(b)}"
    x = np.frombuffer(a, np.uint8)
    y = np.frombuffer(b, np.uint8)
    i = int(np.argmax(x != y))
    return f"first differing byte {i}: {x[i]} != {y[i]}"
