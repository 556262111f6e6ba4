"""Synthetic checkpoint chains for the tree-approach workloads.

Each generator is seeded and returns the full chain (checkpoint 0 first).
Host versions return numpy uint8 arrays; ``*_gpu`` versions build the chain
directly in device memory with torch (ROCm) so multi-GiB chains never cross
PCIe.  Changes are chunk-aligned so they can deduplicate.
"""
import numpy as np


def _pick_runs(rng, nchunks, target, mean_len):
    """Contiguous runs of geometric length (mean `mean_len`) at uniform starts
    until `target` chunks are covered; returns a list of (start, length)."""
    p = 1.0 / mean_len
    n = int(target / mean_len * 1.3) + 8
    runs, covered = [], 0
    while covered < target:
        lens = np.minimum(rng.geometric(p, n), nchunks)
        starts = rng.integers(0, nchunks - lens + 1)
        for st, ln in zip(starts.tolist(), lens.tolist()):
            runs.append((st, ln))
            covered += ln
            if covered >= target:
                break
    return runs


def chain_uniform_changes(seed, size, chunk, nckpt, fresh_frac=0.01, copy_frac=0.02, zero_bytes=1 << 20):
    """64 MiB-style chain: per checkpoint, `fresh_frac` of the chunks get fresh
    random bytes, `copy_frac` become copies of other random chunks of the same
    buffer, and one `zero_bytes`-aligned region is zero-filled."""
    rng = np.random.default_rng(seed)
    cur = rng.integers(0, 256, size, dtype=np.uint8)
    out = [cur.copy()]
    C = size // chunk
    for _ in range(1, nckpt):
        prev = out[-1]
        cur = prev.copy()
        v = cur[: C * chunk].reshape(C, chunk)
        pv = prev[: C * chunk].reshape(C, chunk)
        nf = int(round(fresh_frac * C))
        nc = int(round(copy_frac * C))
        idx = rng.permutation(C)
        fresh = idx[:nf]
        dst = idx[nf:nf + nc]
        src = rng.integers(0, C, nc)
        v[fresh] = rng.integers(0, 256, (nf, chunk), dtype=np.uint8)
        v[dst] = pv[src]
        if zero_bytes and size >= zero_bytes:
            z = int(rng.integers(0, size // zero_bytes)) * zero_bytes
            cur[z:z + zero_bytes] = 0
        out.append(cur)
    return out


def chain_runs(seed, size, chunk, nckpt, change_frac, mean_run=16):
    """1 GiB-style chain (host): per checkpoint `change_frac` of the chunks
    change as runs of geometric length; half of the runs get fresh bytes,
    half copy a run of the previous checkpoint at a random offset."""
    rng = np.random.default_rng(seed)
    cur = rng.integers(0, 256, size, dtype=np.uint8)
    out = [cur.copy()]
    C = size // chunk
    for _ in range(1, nckpt):
        prev = out[-1]
        cur = prev.copy()
        v = cur[: C * chunk].reshape(C, chunk)
        pv = prev[: C * chunk].reshape(C, chunk)
        for i, (st, ln) in enumerate(_pick_runs(rng, C, max(1, int(change_frac * C)), mean_run)):
            if i % 2 == 0:
                v[st:st + ln] = rng.integers(0, 256, (ln, chunk), dtype=np.uint8)
            else:
                s = int(rng.integers(0, C - ln + 1))
                v[st:st + ln] = pv[s:s + ln]
        out.append(cur)
    return out


def chain_runs_gpu(torch, seed, size, chunk, nckpt, change_frac, mean_run=16, device="cuda"):
    """Same recipe as chain_runs, generated on the GPU (torch RNG, so bytes
    differ from the host version for the same seed)."""
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    rng = np.random.default_rng(seed)
    C = size // chunk
    cur = torch.randint(0, 256, (size,), dtype=torch.uint8, device=device, generator=gen)
    out = [cur]
    for _ in range(1, nckpt):
        prev = out[-1]
        cur = prev.clone()
        v = cur[: C * chunk].view(C, chunk)
        pv = prev[: C * chunk].view(C, chunk)
        fresh_idx, dst_idx, src_idx = [], [], []
        for i, (st, ln) in enumerate(_pick_runs(rng, C, max(1, int(change_frac * C)), mean_run)):
            if i % 2 == 0:
                fresh_idx.append(np.arange(st, st + ln))
            else:
                s = int(rng.integers(0, C - ln + 1))
                dst_idx.append(np.arange(st, st + ln))
                src_idx.append(np.arange(s, s + ln))
        if fresh_idx:
            fi = torch.from_numpy(np.concatenate(fresh_idx)).to(device)
            v[fi] = torch.randint(0, 256, (fi.numel(), chunk), dtype=torch.uint8, device=device, generator=gen)
        if dst_idx:
            di = torch.from_numpy(np.concatenate(dst_idx)).to(device)
            si = torch.from_numpy(np.concatenate(src_idx)).to(device)
            v[di] = pv[si]
        out.append(cur)
    return out


def chain_float_field(seed, n_values, nckpt, zero_frac=0.6, block_bytes=4096, perturb_frac=0.05, toggle_frac=0.01):
    """4 GiB-style float32 field (host): `zero_frac` of the values are exact
    zeros in blocks of `block_bytes`, the rest N(0,1); per checkpoint
    `perturb_frac` of the nonzero values get +N(0,1e-3) and `toggle_frac` of
    the blocks flip between zero and nonzero."""
    rng = np.random.default_rng(seed)
    per_block = block_bytes // 4
    nblk = n_values // per_block
    zero = rng.random(nblk) < zero_frac
    x = rng.standard_normal(n_values).astype(np.float32)
    x.reshape(nblk, per_block)[zero] = 0
    out = [x.view(np.uint8).copy()]
    for _ in range(1, nckpt):
        x = x.copy()
        nz = np.flatnonzero(x)
        pick = rng.choice(nz, size=int(perturb_frac * nz.size), replace=False) if nz.size else nz
        x[pick] += rng.normal(0, 1e-3, pick.size).astype(np.float32)
        tog = rng.choice(nblk, size=max(1, int(toggle_frac * nblk)), replace=False)
        xb = x.reshape(nblk, per_block)
        for b in tog:
            if zero[b]:
                xb[b] = rng.standard_normal(per_block).astype(np.float32)
            else:
                xb[b] = 0
            zero[b] = ~zero[b]
        out.append(x.view(np.uint8).copy())
    return out


def chain_uniform_changes_gpu(torch, seed, size, chunk, nckpt, fresh_frac=0.01, copy_frac=0.02,
                              zero_bytes=1 << 20, device="cuda", zero_count=1):
    """chain_uniform_changes generated in device memory (torch RNG);
    `zero_count` aligned regions of `zero_bytes` are zero-filled per
    checkpoint (one per 64 MiB keeps the 64 MiB recipe's density)."""
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    C = size // chunk
    cur = torch.randint(0, 256, (size,), dtype=torch.uint8, device=device, generator=gen)
    out = [cur]
    for _ in range(1, nckpt):
        prev = out[-1]
        cur = prev.clone()
        v = cur[: C * chunk].view(C, chunk)
        pv = prev[: C * chunk].view(C, chunk)
        nf = int(round(fresh_frac * C))
        nc = int(round(copy_frac * C))
        perm = torch.randperm(C, device=device, generator=gen)
        if nf:
            v[perm[:nf]] = torch.randint(0, 256, (nf, chunk), dtype=torch.uint8, device=device, generator=gen)
        if nc:
            src = torch.randint(0, C, (nc,), device=device, generator=gen)
            v[perm[nf:nf + nc]] = pv[src]
        if zero_bytes and size >= zero_bytes:
            for z in torch.randint(0, size // zero_bytes, (zero_count,), generator=gen, device=device).tolist():
                cur[z * zero_bytes:(z + 1) * zero_bytes] = 0
        out.append(cur)
    return out


def batch_sizes(seed, n=64, lo=4 << 20, hi=512 << 20, align=4096):
    """`n` buffer sizes log-uniform in [lo, hi], rounded to `align` bytes."""
    rng = np.random.default_rng(seed)
    x = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    return [int(max(align, round(v / align) * align)) for v in x]


def float_field_versions_gpu(torch, seed, n_values, nckpt, zero_frac=0.6, block_bytes=4096, perturb_frac=0.05,
                             toggle_frac=0.01, device="cuda"):
    """chain_float_field generated in device memory (torch RNG), one version
    at a time: yields a uint8 view of the float32 field, then updates the
    same buffer in place for the next version.  `zero_frac` of the
    `block_bytes` blocks are exact zeros, the rest N(0,1); each later version
    adds N(0,1e-3) to a Bernoulli(`perturb_frac`) sample of the nonzero values
    and flips `toggle_frac` of the blocks between zero and fresh N(0,1)."""
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    per = block_bytes // 4
    nblk = n_values // per
    zero = torch.rand(nblk, generator=gen, device=device) < zero_frac
    x = torch.randn(n_values, generator=gen, device=device, dtype=torch.float32)
    x.view(nblk, per)[zero] = 0
    ntog = max(1, int(toggle_frac * nblk))
    for k in range(nckpt):
        if k:
            pick = (torch.rand(n_values, generator=gen, device=device) < perturb_frac) & (x != 0)
            noise = torch.randn(n_values, generator=gen, device=device, dtype=torch.float32) * 1e-3
            x += torch.where(pick, noise, torch.zeros((), device=device))
            del pick, noise
            tog = torch.randperm(nblk, generator=gen, device=device)[:ntog]
            fresh = torch.randn(ntog, per, generator=gen, device=device, dtype=torch.float32)
            x.view(nblk, per)[tog] = torch.where(zero[tog].unsqueeze(1), fresh, torch.zeros((), device=device))
            zero[tog] = ~zero[tog]
        yield x.view(torch.uint8)
