"""The benchmark configurations at their configured sizes, checked byte for
byte against the oracle.

The oracle here is the OpenMP restatement (oracle/dedup_omp.c), which
tests/test_oracle_parallel.py pins to the serial oracle's bytes; the serial
one would take hours at these sizes.  Data is generated on the GPU; each
version is copied to the host once for the oracle.  Restarts are checked
against the data they restore.

- plain method, 1 GiB, chunks {128, 512, 4096} B x change {0.1, 1, 10} %,
  geometric runs of changes, 10 checkpoints, restarts of 0, 5 and 9;
- MD5 float field, 2^30 float32 values (4 GiB), 64 B chunks, 20 checkpoints,
  restarts of 0, 10 and 19;
- batched form, 64 buffers log-uniform in [4 MiB, 512 MiB] (~8 GiB), 10
  checkpoints, every batched diff against the oracle, every buffer's last
  checkpoint restarted;
- serving form, 32 ranks with buffers log-uniform in [16 MiB, 1 GiB], Poisson
  arrivals at 200 requests/s (85 % checkpoints, 15 % restarts), 200 GiB
  budget: every restart verified, every diff of a sample of ranks equal to
  the same ranks' diffs alone and to the oracle's.
"""
import numpy as np
import pytest

import oracle as O
import reproducibility_icpp23_scalable_gpu_deduplication_amd as dd
from reproducibility_icpp23_scalable_gpu_deduplication_amd import workloads as W

pytestmark = pytest.mark.gpu


def host_bytes(t):
    return t.cpu().numpy()


def same(dev_diff, want):
    got = host_bytes(dev_diff)
    return got.size == len(want) and np.array_equal(got, np.frombuffer(want, np.uint8))


def restart_equal(torch, g, diffs, k, data):
    out = torch.empty(data.numel(), dtype=torch.uint8, device="cuda")
    g.restart_device([d.data_ptr() for d in diffs[:k + 1]], [d.numel() for d in diffs[:k + 1]], k, out.data_ptr(),
                     data.numel())
    torch.cuda.synchronize()
    return torch.equal(out, data)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("frac", [0.001, 0.01, 0.1])
@pytest.mark.parametrize("chunk", [128, 512, 4096])
def test_plain_1gib_sweep_cell(torch_gpu, chunk, frac):
    torch = torch_gpu
    size, nckpt = 1 << 30, 10
    print(f"plain 1 GiB chunk {chunk} change {frac}", flush=True)
    chain = W.chain_runs_gpu(torch, 1931, size, chunk, nckpt, frac)
    torch.cuda.synchronize()
    g = dd.TreeDeduplicator(chunk)
    o = O.ParallelOracleDeduplicator(chunk)
    cap = dd.max_diff_size(size, chunk, nckpt)
    out = torch.empty(cap, dtype=torch.uint8, device="cuda")
    diffs = []
    for k, v in enumerate(chain):
        n = g.checkpoint_device(v.data_ptr(), size, out.data_ptr(), cap, k == 0)
        torch.cuda.synchronize()
        want = o.checkpoint_array(host_bytes(v), k == 0)
        assert same(out[:n], want), f"checkpoint {k}: diff differs from the oracle"
        assert g.counts() == o.counts()
        diffs.append(out[:n].clone())
    for k in (0, 5, 9):
        assert restart_equal(torch, g, diffs, k, chain[k]), f"restart {k} differs"


@pytest.mark.timeout(900)
def test_md5_float_field_4gib(torch_gpu):
    torch = torch_gpu
    n_values, chunk, nckpt = 1 << 30, 64, 20
    size = 4 * n_values
    keep = (0, 10, 19)
    g = dd.TreeDeduplicator(chunk, dd.HASH_MD5)
    o = O.ParallelOracleDeduplicator(chunk, O.MD5)
    cap = dd.max_diff_size(size, chunk, nckpt)
    out = torch.empty(cap, dtype=torch.uint8, device="cuda")
    saved, diffs = {}, []
    for k, v in enumerate(W.float_field_versions_gpu(torch, 1931, n_values, nckpt)):
        print(f"field checkpoint {k}", flush=True)
        if k in keep:
            saved[k] = v.clone()
        n = g.checkpoint_device(v.data_ptr(), size, out.data_ptr(), cap, k == 0)
        torch.cuda.synchronize()
        want = o.checkpoint_array(host_bytes(v), k == 0)
        assert same(out[:n], want), f"checkpoint {k}: diff differs from the oracle"
        diffs.append(out[:n].clone())
    del out
    for k in keep:
        assert restart_equal(torch, g, diffs, k, saved[k]), f"restart {k} differs"


@pytest.mark.timeout(900)
def test_batched_64_buffers_8gib(torch_gpu):
    torch = torch_gpu
    seed, nbuf, nckpt, chunk = 1931, 64, 10, 128
    sizes = W.batch_sizes(seed, nbuf)
    rates = np.random.default_rng(seed).uniform(0.001, 0.05, nbuf)
    chains = [W.chain_uniform_changes_gpu(torch, seed + 1 + i, sz, chunk, nckpt, r / 3, 2 * r / 3)
              for i, (sz, r) in enumerate(zip(sizes, rates))]
    torch.cuda.synchronize()
    hs = [dd.TreeDeduplicator(chunk) for _ in sizes]
    batch = dd.Batch(hs, graph=True)
    caps = [dd.max_diff_size(sz, chunk, nckpt) for sz in sizes]
    outs = [torch.empty(c, dtype=torch.uint8, device="cuda") for c in caps]
    diffs = [[] for _ in sizes]
    for k in range(nckpt):
        print(f"batched checkpoint {k}", flush=True)
        lens = batch.checkpoint_device([c[k].data_ptr() for c in chains], sizes, [k == 0] * nbuf,
                                       [o.data_ptr() for o in outs], caps)
        torch.cuda.synchronize()
        for i in range(nbuf):
            diffs[i].append(outs[i][:lens[i]].clone())
    del outs
    for i in range(nbuf):
        o = O.ParallelOracleDeduplicator(chunk)
        for k in range(nckpt):
            want = o.checkpoint_array(host_bytes(chains[i][k]), k == 0)
            assert same(diffs[i][k], want), f"buffer {i} ({sizes[i]} B) checkpoint {k}: differs from the oracle"
        assert restart_equal(torch, hs[i], diffs[i], nckpt - 1, chains[i][nckpt - 1]), f"buffer {i}: restart differs"


@pytest.mark.timeout(900)
def test_serving_32_ranks(torch_gpu):
    torch = torch_gpu
    seed, ranks, chunk, versions = 2024, 32, 128, 10
    rng = np.random.default_rng(seed)
    sizes = W.batch_sizes(seed, ranks, 16 << 20, 1 << 30, 4096)
    pool_size = max(sizes)
    pool = W.chain_uniform_changes_gpu(torch, seed, pool_size, chunk, versions, 0.01, 0.02, 1 << 20,
                                       zero_count=max(1, pool_size >> 26))
    offs = [int(rng.integers(0, (pool_size - sz) // 4096 + 1)) * 4096 for sz in sizes]
    streams = [torch.cuda.Stream() for _ in range(ranks)]
    handles = [dd.TreeDeduplicator(chunk, stream=st.cuda_stream) for st in streams]
    caps = [dd.max_diff_size(sz, chunk, 64) for sz in sizes]
    torch.cuda.synchronize()

    def data(r, v):
        return pool[v][offs[r]:offs[r] + sizes[r]]

    srv = dd.Server(200 << 30, 8)
    ckpts = [[] for _ in range(ranks)]  # per rank: (version, output buffer, ticket)
    restarts = []  # (rank, checkpoint id, output, ticket)
    n_req = 320
    arrivals = np.cumsum(rng.exponential(1 / 200.0, n_req))
    import time

    start = time.perf_counter()
    for t, r, is_rs in zip(arrivals, rng.integers(0, ranks, n_req), rng.random(n_req) < 0.15):
        d = start + t - time.perf_counter()
        if d > 0:
            time.sleep(d)
        r = int(r)
        done = len(ckpts[r])
        if is_rs and done:
            base = (done - 1) - (done - 1) % versions
            j = int(rng.integers(base, done))
            out = torch.empty(sizes[r], dtype=torch.uint8, device="cuda")
            ptrs = [0] * (j + 1)
            lens = [0] * (j + 1)
            for i in range(base, j + 1):
                ptrs[i] = ckpts[r][i][1].data_ptr()
                lens[i] = caps[r]
            restarts.append((r, j, out, srv.submit_restart(handles[r], ptrs, lens, j, out.data_ptr(), sizes[r])))
        else:
            v = done % versions
            out = torch.empty(caps[r], dtype=torch.uint8, device="cuda")
            tk = srv.submit_checkpoint(handles[r], data(r, v).data_ptr(), sizes[r], out.data_ptr(), caps[r],
                                       make_baseline=v == 0)
            ckpts[r].append((v, out, tk))
    got = [[] for _ in range(ranks)]
    for r in range(ranks):
        for v, out, tk in ckpts[r]:
            res = srv.wait(tk)
            assert res["status"] == dd.DD_OK
            got[r].append(out[:res["out_len"]])
    for r, j, out, tk in restarts:
        assert srv.wait(tk)["status"] == dd.DD_OK
        torch.cuda.synchronize()
        assert torch.equal(out, data(r, j % versions)), f"rank {r} restart {j} differs"
    srv.close()
    assert len(restarts) > 20
    # A sample of ranks (the smallest, the largest and two more with the most
    # checkpoints): alone (a fresh handle, one call at a time) and the oracle.
    order = sorted(range(ranks), key=lambda r: sizes[r])
    busiest = sorted(range(ranks), key=lambda r: -len(ckpts[r]))
    sample = list(dict.fromkeys([order[0], order[-1]] + busiest[:2]))
    for r in sample:
        alone = dd.TreeDeduplicator(chunk)
        o = O.ParallelOracleDeduplicator(chunk)
        out = torch.empty(caps[r], dtype=torch.uint8, device="cuda")
        for i, (v, _, _) in enumerate(ckpts[r]):
            n = alone.checkpoint_device(data(r, v).data_ptr(), sizes[r], out.data_ptr(), caps[r], v == 0)
            torch.cuda.synchronize()
            assert torch.equal(out[:n], got[r][i]), f"rank {r} checkpoint {i}: mix differs from alone"
            want = o.checkpoint_array(host_bytes(data(r, v)), v == 0)
            assert same(got[r][i], want), f"rank {r} checkpoint {i}: differs from the oracle"
