"""The persistent launches and per-bucket tuning of the batched path vs the oracle.

With ``Tuning(persistent=1)`` the forward's sampling is one launch (cell-table
and guide-table chunks taken from a device queue, after the build of the
curves whose tables exceed LDS) and the VJP one launch (cell, guide and
bin-range items).  The batch mixes every kind of work
and has more items than resident work-groups, so that groups loop over the
queue.  Bars as in test_gpu_parity: kept counts, bins and compaction order
bit-exact; float64 values bit-exact (sequential CDF chain), float32 within
1e-5; gradients within 1e-9 (float64) / 1e-4 (float32) of the oracle per
curve.  Against the per-kind launches of the same plan the forward is
bit-identical and float32 gradients too (exact integer sums).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _mixed_batch(dtype):
    """(n_nodes, events, x, density): 300 small curves, cell-table sizes, curves
    above the LDS build (global tables, bin-range VJP) and one curve whose CDF is
    not monotone."""
    from quantom_loits_amd.workloads import density_params, eval_density

    rng = np.random.default_rng(71)
    small_n = rng.integers(16, 700, 300)
    small_e = rng.integers(50, 3000, 300)
    big_n = [2048, 4000, 5000, 9000, 13000, 4800]
    big_e = [70000, 20000, 30000, 100000, 40000, 25000]
    nodes = np.concatenate([small_n, big_n]).astype(np.int64)
    events = np.concatenate([small_e, big_e]).astype(np.int64)
    params = density_params(len(nodes), 72)
    xs, ds = [], []
    for i, n in enumerate(nodes):
        xg = np.linspace(0.001, 0.999, int(n))
        if i == len(nodes) - 1:  # negative densities: a non-monotone CDF
            d = 1.0 + 1.6 * np.sin(np.linspace(0, 12, int(n)))
        else:
            d = eval_density(xg, params[i])
        xs.append(xg)
        ds.append(d)
    perm = rng.permutation(len(nodes))  # big curves anywhere in the batch
    nodes, events = nodes[perm], events[perm]
    xs = [xs[p] for p in perm]
    ds = [ds[p] for p in perm]
    return nodes, events, np.concatenate(xs), np.concatenate(ds)


def _check_grads(got, ref, node_off, nodes, tol):
    for i, (o, n) in enumerate(zip(node_off, nodes)):
        sl = slice(int(o), int(o + n))
        scale = np.abs(ref[sl]).max()
        if scale == 0:
            assert np.abs(got[sl]).max() == 0, i
            continue
        err = np.abs(got[sl] - ref[sl]).max() / scale
        assert err <= tol, (i, int(n), err)


@pytest.mark.parametrize("dtype,window", [(torch.float32, (0.05, 0.95)), (torch.float32, None),
                                          (torch.float64, (0.05, 0.95))])
def test_persistent_launches_match_oracle_and_per_kind_launches(oracle_port, hip_device, dtype,
                                                                window):
    from quantom_loits_amd.batched import BatchedLOITS, Tuning

    nodes, events, x, d = _mixed_batch(dtype)
    n = len(nodes)
    node_off = np.concatenate([[0], np.cumsum(nodes)[:-1]]).astype(np.int64)
    rng = np.random.default_rng(73)
    base = rng.integers(0, 2 ** 40, n).astype(np.uint64)
    streams = rng.integers(0, 2 ** 33, n).astype(np.uint64)
    accept = np.ones(n, np.uint8)
    accept[5] = 0
    scan = 1 if dtype == torch.float64 else 0
    common = dict(threads=1024, chunk_events=1 << 16, scan_mode=scan)
    per_kind = BatchedLOITS(nodes, events, base, streams, accept, node_off, dtype=dtype,
                            window=window, tuning=Tuning(**common))
    pers = BatchedLOITS(nodes, events, base, streams, accept, node_off, dtype=dtype,
                        window=window, tuning=Tuning(persistent=1, **common))
    L = pers.layout
    print("persistent", dtype, window, "chunks", L["n_chunks"], "forward items",
          L["forward_items"], "vjp items", L["grid_backward"], "global", L["global_instances"],
          "ranged", L["range_instances"], "cell", L["cell_instances"], flush=True)
    assert L["persistent"] == 1 and L["forward_items"] == L["n_chunks"]
    assert L["n_chunks"] > 300 and L["global_instances"] >= 2
    xt = torch.from_numpy(x).to(dtype).cuda()
    dt = torch.from_numpy(d).to(dtype).cuda()
    x64, d64 = xt.double().cpu().numpy(), dt.double().cpu().numpy()
    ref_v, ref_b, ref_k, _, _ = oracle_port.batched_forward(
        x64, d64, node_off, nodes, events, base, streams, 17, accept=accept, window=window)
    gv = np.random.default_rng(74).standard_normal(max(len(ref_v), 1))
    ref_g = oracle_port.batched_backward(x64, d64, node_off, nodes, events, base, streams, 17,
                                         gv, accept=accept, window=window)
    kv, kk, kst, kex = per_kind.run_forward(xt, dt, 17, want_bins=True)
    kgd = torch.empty_like(dt)
    per_kind.backward_raw(xt, dt, 17, kst, torch.from_numpy(gv).to(dtype).cuda(), kgd)
    st = pers.new_state()
    for call in range(2):  # the queue counters return to zero after each launch
        values, kept, st, ex = pers.run_forward(xt, dt, 17, want_bins=True, state=st)
        assert kept.cpu().numpy().tolist() == ref_k.tolist(), call
        assert np.array_equal(ex["bins"].cpu().numpy(), ref_b), call
        v = values.double().cpu().numpy()
        if dtype == torch.float64:
            assert np.array_equal(v, ref_v), call
        else:
            assert np.max(np.abs(v - ref_v) / np.abs(ref_v)) <= 1e-5, call
        assert torch.equal(values, kv) and torch.equal(ex["bins"], kex["bins"]), call
        gd = torch.empty_like(dt)
        pers.backward_raw(xt, dt, 17, st, torch.from_numpy(gv).to(dtype).cuda(), gd)
        got = gd.double().cpu().numpy()
        _check_grads(got, ref_g, node_off, nodes, 1e-9 if dtype == torch.float64 else 1e-4)
        if dtype == torch.float32:
            assert torch.equal(gd, kgd), call
        else:  # float64 LDS and device atomics may sum in another order
            _check_grads(got, kgd.cpu().numpy(), node_off, nodes, 1e-12)
        flags = st["scratch"][:64].view(torch.int32).cpu().numpy()
        assert flags[4:8].tolist() == [0, 0, 0, 0], flags


def test_persistent_graph_replays_match_eager(hip_device):
    from quantom_loits_amd.batched import BatchedLOITS, Tuning

    nodes, events, x, d = _mixed_batch(torch.float32)
    n = len(nodes)
    node_off = np.concatenate([[0], np.cumsum(nodes)[:-1]]).astype(np.int64)
    base = np.random.default_rng(75).integers(0, 2 ** 40, n).astype(np.uint64)
    tun = Tuning(threads=1024, chunk_events=1 << 16, persistent=1)
    s = BatchedLOITS(nodes, events, base, np.zeros(n, np.uint64), None, node_off,
                     dtype=torch.float32, window=(0.05, 0.95), tuning=tun)
    xt = torch.from_numpy(x).float().cuda()
    dt = torch.from_numpy(d).float().cuda()
    g = s.capture(xt.clone(), dt.clone(), 5)
    other = BatchedLOITS.uniform(4, 64, 5000, 3, tuning=Tuning(threads=256))
    ox = torch.linspace(0, 1, 256, device="cuda")
    od = torch.rand(256, device="cuda") + 0.1
    for r in range(3):
        print("persistent graph replay", r, flush=True)
        rr = np.random.default_rng(200 + r)
        g.density.copy_(torch.from_numpy(d * rr.uniform(0.5, 2.0, d.shape)).float())
        streams = rr.integers(0, 2 ** 33, n).astype(np.uint64)
        g.set_streams(streams)
        grad = torch.from_numpy(rr.standard_normal(max(s.capacity, 1))).float().cuda()
        g.grad.copy_(grad)
        g.replay()
        other.run_forward(ox, od, 77 + r)
        torch.cuda.synchronize()
        e = BatchedLOITS(nodes, events, base, streams, None, node_off, dtype=torch.float32,
                         window=(0.05, 0.95), tuning=tun)
        ev, ek, est, _ = e.run_forward(g.x, g.density, 5)
        egd = torch.empty_like(g.density)
        e.backward_raw(g.x, g.density, 5, est, grad[: max(ev.numel(), 1)].contiguous(), egd)
        torch.cuda.synchronize()
        assert torch.equal(g.kept, ek)
        assert torch.equal(g.values(), ev)
        assert torch.equal(g.grad_density, egd)


def test_bucket_tuning_changes_chunks_not_results(oracle_port, hip_device):
    from quantom_loits_amd.batched import BatchedLOITS, Tuning

    nodes, events, x, d = _mixed_batch(torch.float32)
    n = len(nodes)
    node_off = np.concatenate([[0], np.cumsum(nodes)[:-1]]).astype(np.int64)
    base = np.random.default_rng(76).integers(0, 2 ** 40, n).astype(np.uint64)
    xt = torch.from_numpy(x).float().cuda()
    dt = torch.from_numpy(d).float().cuda()
    plain = BatchedLOITS(nodes, events, base, None, None, node_off, window=(0.05, 0.95),
                         tuning=Tuning(chunk_events=1 << 16))
    tuned = BatchedLOITS(nodes, events, base, None, None, node_off, window=(0.05, 0.95),
                         tuning=Tuning(chunk_events=1 << 16, persistent=1,
                                       buckets={10: (1 << 12, None), 12: (1 << 17, None),
                                                13: (1 << 14, 1), 14: (1 << 15, 3)}))
    assert tuned.layout["n_chunks"] != plain.layout["n_chunks"]
    pv, pk, _, _ = plain.run_forward(xt, dt, 3)
    tv, tk, tst, _ = tuned.run_forward(xt, dt, 3)
    assert torch.equal(pk, tk) and torch.equal(pv, tv)
    gv = np.random.default_rng(77).standard_normal(max(pv.numel(), 1))
    ref_g = oracle_port.batched_backward(xt.double().cpu().numpy(), dt.double().cpu().numpy(),
                                         node_off, nodes, events, base,
                                         np.zeros(n, np.uint64), 3, gv, window=(0.05, 0.95))
    gd = torch.empty_like(dt)
    tuned.backward_raw(xt, dt, 3, tst, torch.from_numpy(gv).float().cuda(), gd)
    _check_grads(gd.double().cpu().numpy(), ref_g, node_off, nodes, 1e-4)
