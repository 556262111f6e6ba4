"""Request loop (dd_server_*): every checkpoint and restart result is the same
bytes when the request runs alone and when it runs in a mix of ranks under
the HBM budget, batched with other ranks' checkpoints or not; admission keeps
reservations within the budget."""
import numpy as np
import pytest

import reproducibility_icpp23_scalable_gpu_deduplication_amd as dd
from reproducibility_icpp23_scalable_gpu_deduplication_amd import workloads as W
from gpu_util import first_mismatch, oracle_chain

pytestmark = pytest.mark.gpu

CHUNK = 128


def rank_chains():
    sizes = [1 << 20, 3 << 18, 777 * CHUNK, 5000 * CHUNK + 40, 1 << 16, (1 << 19) + 4096]
    return [W.chain_uniform_changes(101 + i, s, CHUNK, 5, fresh_frac=0.02, copy_frac=0.03, zero_bytes=4096)
            for i, s in enumerate(sizes)]


def alone(torch, chain, baselines):
    """Plain path, one handle, nothing else running."""
    g = dd.TreeDeduplicator(CHUNK)
    out = []
    for k, x in enumerate(chain):
        t = torch.from_numpy(x).to("cuda")
        torch.cuda.synchronize()
        out.append(g.checkpoint(t.data_ptr(), t.numel(), k in baselines))
    return out


def test_server_mix_equals_alone(torch_gpu):
    torch = torch_gpu
    chains = rank_chains()
    nr, nk = len(chains), len(chains[0])
    baselines = [(0,) if r % 2 == 0 else (0, 3) for r in range(nr)]
    expect = [alone(torch, c, baselines[r]) for r, c in enumerate(chains)]
    for r, c in enumerate(chains):
        od, _ = oracle_chain(CHUNK, c, baselines=baselines[r])
        assert expect[r] == od, f"rank {r}: plain path differs from the oracle"

    handles = [dd.TreeDeduplicator(CHUNK) for _ in range(nr)]
    inputs = [[torch.from_numpy(x).to("cuda") for x in c] for c in chains]
    caps = [[dd.max_diff_size(x.nbytes, CHUNK, k) for k, x in enumerate(c)] for c in chains]
    outs = [[torch.empty(caps[r][k], dtype=torch.uint8, device="cuda") for k in range(nk)] for r in range(nr)]
    torch.cuda.synchronize()
    # The six ranks keep about 8 MB of device state charged; a 12 MiB budget
    # leaves room for two or three requests in flight, so requests wait for
    # others to finish although six execution slots are free.
    budget = 12 << 20
    srv = dd.Server(budget, slots=6)
    tickets = {}
    for k in range(nk):  # interleave the ranks: rank r checkpoint k
        for r in range(nr):
            tickets[(r, k)] = srv.submit_checkpoint(handles[r], inputs[r][k].data_ptr(), inputs[r][k].numel(),
                                                    outs[r][k].data_ptr(), caps[r][k], k in baselines[r])
    lens = {}
    for key, t in tickets.items():
        res = srv.wait(t)
        assert res["status"] == dd.DD_OK, f"rank/ckpt {key}: {dd.strerror(res['status'])}"
        assert res["kind"] == 0
        assert res["t_submit"] <= res["t_start"] <= res["t_end"]
        lens[key] = res["out_len"]
    for r in range(nr):
        for k in range(nk):
            got = outs[r][k][:lens[(r, k)]].cpu().numpy().tobytes()
            assert got == expect[r][k], f"rank {r} ckpt {k}: {first_mismatch(got, expect[r][k])}"

    # Restarts of every checkpoint, all ranks in flight together.
    rest = {}
    rtick = {}
    for k in range(nk):
        for r in range(nr):
            o = torch.zeros(chains[r][k].nbytes, dtype=torch.uint8, device="cuda")
            rest[(r, k)] = o
    torch.cuda.synchronize()
    for k in reversed(range(nk)):
        for r in range(nr):
            ptrs = [outs[r][j].data_ptr() for j in range(k + 1)]
            ls = [lens[(r, j)] for j in range(k + 1)]
            rtick[(r, k)] = srv.submit_restart(handles[r], ptrs, ls, k, rest[(r, k)].data_ptr(), chains[r][k].nbytes)
    for key, t in rtick.items():
        res = srv.wait(t)
        assert res["status"] == dd.DD_OK, f"restart {key}: {dd.strerror(res['status'])}"
        assert res["kind"] == 1 and res["out_len"] == chains[key[0]][key[1]].nbytes
    for (r, k), o in rest.items():
        assert o.cpu().numpy().tobytes() == chains[r][k].tobytes(), f"restart rank {r} ckpt {k}"
    in_use, peak = srv.usage()
    srv.close()
    assert peak <= budget
    # with nothing in flight the budget holds exactly what the handles hold
    # (their state and their restart buffers)
    assert in_use == sum(h.device_bytes() for h in handles)
    assert in_use > 0


def test_server_oversized_request_runs_alone(torch_gpu):
    # A request larger than the whole budget still runs once nothing else is
    # in flight, and an error status comes back through the result.
    torch = torch_gpu
    rng = np.random.default_rng(9)
    x = rng.integers(0, 256, 1 << 18, dtype=np.uint8)
    t = torch.from_numpy(x).to("cuda")
    cap = dd.max_diff_size(x.nbytes, CHUNK, 0)
    o = torch.empty(cap, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    srv = dd.Server(1 << 16, slots=2)
    g = dd.TreeDeduplicator(CHUNK)
    res = srv.wait(srv.submit_checkpoint(g, t.data_ptr(), t.numel(), o.data_ptr(), cap, True))
    assert res["status"] == dd.DD_OK
    assert o[:res["out_len"]].cpu().numpy().tobytes() == alone(torch, [x], (0,))[0]
    g2 = dd.TreeDeduplicator(CHUNK)
    res = srv.wait(srv.submit_checkpoint(g2, t.data_ptr(), t.numel(), o.data_ptr(), 16, True))
    assert res["status"] == dd.DD_EINVAL and res["out_len"] == 0
    in_use, peak = srv.usage()
    srv.close()
    assert peak > (1 << 16)  # the oversized request was admitted alone


def test_server_ticket_waited_once(torch_gpu):
    torch = torch_gpu
    x = np.random.default_rng(4).integers(0, 256, 1 << 16, dtype=np.uint8)
    t = torch.from_numpy(x).to("cuda")
    cap = dd.max_diff_size(x.nbytes, CHUNK, 0)
    o = torch.empty(cap, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    srv = dd.Server(1 << 30, slots=1)
    g = dd.TreeDeduplicator(CHUNK)
    tk = srv.submit_checkpoint(g, t.data_ptr(), t.numel(), o.data_ptr(), cap, True)
    assert srv.wait(tk)["status"] == dd.DD_OK
    with pytest.raises(dd.DedupError) as e:
        srv.wait(tk)
    assert e.value.status == dd.DD_EINVAL
    srv.close()


@pytest.mark.parametrize("max_group", [64, 1])
def test_server_batches_admitted_checkpoints(torch_gpu, max_group):
    # One execution slot and every request queued at once: the slot admits
    # the queued checkpoints of the other idle ranks with each one it starts
    # and runs them as one batched job table (baselines and incremental
    # checkpoints mixed).  A request the batch would reject (output buffer too
    # small) fails alone, the others of its group succeed.
    torch = torch_gpu
    chains = rank_chains()
    nr, nk = len(chains), len(chains[0])
    baselines = [(0,) if r % 2 == 0 else (0, 2) for r in range(nr)]
    expect = [alone(torch, c, baselines[r]) for r, c in enumerate(chains)]
    handles = [dd.TreeDeduplicator(CHUNK) for _ in range(nr)]
    bad = dd.TreeDeduplicator(CHUNK)
    inputs = [[torch.from_numpy(x).to("cuda") for x in c] for c in chains]
    caps = [[dd.max_diff_size(x.nbytes, CHUNK, k) for k, x in enumerate(c)] for c in chains]
    outs = [[torch.empty(caps[r][k], dtype=torch.uint8, device="cuda") for k in range(nk)] for r in range(nr)]
    torch.cuda.synchronize()
    srv = dd.Server(1 << 30, slots=1)
    srv.set_batching(max_group)
    tickets = {}
    for k in range(nk):
        for r in range(nr):
            tickets[(r, k)] = srv.submit_checkpoint(handles[r], inputs[r][k].data_ptr(), inputs[r][k].numel(),
                                                    outs[r][k].data_ptr(), caps[r][k], k in baselines[r])
        if k == 1:
            tickets[("bad", k)] = srv.submit_checkpoint(bad, inputs[0][0].data_ptr(), inputs[0][0].numel(),
                                                        outs[0][0].data_ptr(), 16, True)
    groups, lens = [], {}
    for key, t in tickets.items():
        res = srv.wait(t)
        groups.append(res["group"])
        if key[0] == "bad":
            assert res["status"] == dd.DD_EINVAL and res["out_len"] == 0
            continue
        assert res["status"] == dd.DD_OK, f"rank/ckpt {key}: {dd.strerror(res['status'])}"
        lens[key] = res["out_len"]
    srv.close()
    if max_group == 1:
        assert max(groups) == 1
    else:
        assert max(groups) > 1, groups
        assert max(groups) <= nr  # one request per rank in a group
    for r in range(nr):
        for k in range(nk):
            got = outs[r][k][:lens[(r, k)]].cpu().numpy().tobytes()
            assert got == expect[r][k], f"rank {r} ckpt {k}: {first_mismatch(got, expect[r][k])}"


def test_server_batching_keeps_rank_order(torch_gpu):
    # Queue while one slot is busy: checkpoint 1 of rank A, a restart of rank
    # B's checkpoint 0, then B's checkpoint 1 written over the buffer that
    # holds B's checkpoint 0.  A's checkpoint leads a group that may not take
    # B's checkpoint past B's restart, which reads that buffer.
    torch = torch_gpu
    ca, cb = rank_chains()[:2]
    big = torch.randint(0, 256, (64 << 20,), dtype=torch.uint8, device="cuda")
    ta = [torch.from_numpy(x).to("cuda") for x in ca[:2]]
    tb = [torch.from_numpy(x).to("cuda") for x in cb[:2]]
    ha, hb, hc = dd.TreeDeduplicator(CHUNK), dd.TreeDeduplicator(CHUNK), dd.TreeDeduplicator(CHUNK)
    cap = dd.max_diff_size(max(ca[0].nbytes, cb[0].nbytes), CHUNK, 1)
    oa = [torch.empty(cap, dtype=torch.uint8, device="cuda") for _ in range(2)]
    ob = torch.empty(cap, dtype=torch.uint8, device="cuda")
    oc = torch.empty(dd.max_diff_size(big.numel(), CHUNK, 0), dtype=torch.uint8, device="cuda")
    rb = torch.zeros(cb[0].nbytes, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    srv = dd.Server(1 << 32, slots=1)
    res = srv.wait(srv.submit_checkpoint(ha, ta[0].data_ptr(), ta[0].numel(), oa[0].data_ptr(), cap, True))
    assert res["status"] == dd.DD_OK
    lb = srv.wait(srv.submit_checkpoint(hb, tb[0].data_ptr(), tb[0].numel(), ob.data_ptr(), cap, True))
    assert lb["status"] == dd.DD_OK
    blocker = srv.submit_checkpoint(hc, big.data_ptr(), big.numel(), oc.data_ptr(), oc.numel(), True)
    t_a1 = srv.submit_checkpoint(ha, ta[1].data_ptr(), ta[1].numel(), oa[1].data_ptr(), cap, False)
    t_rb = srv.submit_restart(hb, [ob.data_ptr()], [lb["out_len"]], 0, rb.data_ptr(), cb[0].nbytes)
    t_b1 = srv.submit_checkpoint(hb, tb[1].data_ptr(), tb[1].numel(), ob.data_ptr(), cap, False)
    assert srv.wait(blocker)["status"] == dd.DD_OK
    ra, rr, r1 = srv.wait(t_a1), srv.wait(t_rb), srv.wait(t_b1)
    srv.close()
    assert ra["status"] == rr["status"] == r1["status"] == dd.DD_OK
    assert rr["t_end"] <= r1["t_start"]
    assert rb.cpu().numpy().tobytes() == cb[0].tobytes()
    assert ob[:r1["out_len"]].cpu().numpy().tobytes() == alone(torch, cb[:2], (0,))[1]
    assert oa[1][:ra["out_len"]].cpu().numpy().tobytes() == alone(torch, ca[:2], (0,))[1]
