"""Public entry points against the oracle (GPU).

* ``LOITS(backend="hip")``: events and both xsec gradients over two successive
  calls, with the reference's seed policy (seed = torch.initial_seed(),
  Philox streams 2*sequence and 2*sequence+1, sequence post-incremented per
  call: cpp/backend.py:34-101, tests/test_native_rng.py:73-85).
* ``BoundedLOITS``: the 11-tensor state and the gradients against the oracle's
  forward/backward (cpp/bindings.cpp:24-208, openmp row offsets).
* The batched VJP that rebuilds the curve tables (``recompute_tables``)
  against the stored-table VJP and the oracle, eagerly and in a HIP graph.
* The bindings on an explicit side stream, their argument checks, and the
  soft-binning layer on NaN and far out-of-range values.
"""
import numpy as np
import pytest
import torch

from oracle.oracle import uniform_batch

pytestmark = pytest.mark.gpu


def _theory(rng, B, nx, ny, k, dtype=torch.float64):
    """Reference-layout theory outputs (cpp/loits_core.cpp:62-94) on the GPU."""
    xb = np.sort(rng.uniform(0.001, 0.999, (B, nx, k)), axis=2)
    qb = np.sort(rng.uniform(1.0, 100.0, (B, ny, k)), axis=2)
    xs = rng.uniform(0.1, 1.0, (B, ny, nx, k))
    qs = rng.uniform(0.1, 1.0, (B, nx, ny, k))
    w = rng.uniform(0.5, 1.5, (B, nx, ny))
    w /= w.sum()
    acc = rng.uniform(0, 1, (B, nx, ny)) < 0.85
    return [torch.from_numpy(a).to(dtype).cuda() for a in (xb, xs, qb, qs, w)] + \
        [torch.from_numpy(acc).cuda()]


def _np(t):
    return t.detach().cpu().numpy()


def test_public_loits_matches_oracle_over_two_calls(hip_device, oracle_port):
    from quantom_loits_amd import LOITS

    torch.manual_seed(20231)
    seed = torch.initial_seed()
    rng = np.random.default_rng(8)
    xb, xs, qb, qs, w, acc = _theory(rng, 2, 4, 3, 10)
    sampler = LOITS(backend="hip", device="cuda")
    assert sampler.impl.seed == seed and sampler.impl.sequence == 0
    E = 30000
    for seq in (0, 1):
        print("public LOITS call, sequence", seq, flush=True)
        xs_g = xs.clone().requires_grad_(True)
        qs_g = qs.clone().requires_grad_(True)
        outs = (xb, xs_g, qb, qs_g, w, acc, torch.ones(2, device="cuda"))  # 7th: total_xsec
        events = sampler(outs, E)
        ref = oracle_port.forward(_np(xb), _np(xs), _np(qb), _np(qs), _np(w), _np(acc), E, seed,
                                  seq)
        assert events.shape == ref["events"].shape
        np.testing.assert_array_equal(_np(events), ref["events"])
        g = torch.from_numpy(rng.standard_normal(ref["events"].shape)).cuda()
        events.backward(g)
        gx, gq = oracle_port.backward(_np(g), _np(xb), _np(xs), _np(qb), _np(qs), _np(acc), ref)
        np.testing.assert_array_equal(_np(xs_g.grad), gx)
        np.testing.assert_array_equal(_np(qs_g.grad), gq)
        assert sampler.impl.sequence == seq + 1
    # a no-grad call advances the sequence too (pytorch/gan.py:111)
    with torch.no_grad():
        sampler((xb, xs, qb, qs, w, acc), 100)
    assert sampler.impl.sequence == 3


def test_bounded_state_and_gradients_match_oracle(hip_device, oracle_port):
    from quantom_loits_amd.backend import BoundedLOITS

    rng = np.random.default_rng(41)
    xb, xs, qb, qs, w, acc = _theory(rng, 2, 6, 5, 10)
    E, seed, seq = 40000, 1234, 5
    b = BoundedLOITS(xb, qb, torch.float64, nmax_cap=3000, active_cap=2 * E)
    print("bounded vs oracle", flush=True)
    b.forward(xb, xs, qb, qs, w, acc, E, seed, seq)
    state = b.state()
    ref = oracle_port.forward(_np(xb), _np(xs), _np(qb), _np(qs), _np(w), _np(acc), E, seed, seq)
    names = ["events", "norm_x", "norm_q", "cdf_x", "cdf_q", "u_x", "u_q", "interval_x",
             "interval_q", "packed", "row_offsets"]
    for name, got in zip(names, state):
        r = ref[name]
        assert got.shape == tuple(r.shape), name
        np.testing.assert_array_equal(_np(got).reshape(r.shape), r, err_msg=name)
    kept = state[0].shape[0]
    g = rng.standard_normal((kept, 2))
    b.grad_events[:kept].copy_(torch.from_numpy(g))
    b.backward(xb, xs, qb, qs, acc)
    gx, gq = oracle_port.backward(g, _np(xb), _np(xs), _np(qb), _np(qs), _np(acc), ref)
    np.testing.assert_array_equal(_np(b.grad_x), gx)
    np.testing.assert_array_equal(_np(b.grad_q), gq)


def _mixed_curves(dtype):
    """A regular curve, one whose forward tables stay in global memory and one
    whose VJP splits into bin ranges, plus a tiny one."""
    from quantom_loits_amd.workloads import density_params, eval_density

    nodes = np.array([2048, 6000, 20000, 37], dtype=np.int64)
    events = np.array([150000, 60000, 80000, 3000], dtype=np.int64)
    params = density_params(len(nodes), 13)
    x = np.concatenate([np.linspace(0.001, 0.999, n) for n in nodes])
    d = np.concatenate([eval_density(np.linspace(0.001, 0.999, n), params[i])
                        for i, n in enumerate(nodes)])
    return nodes, events, torch.from_numpy(x).to(dtype).cuda(), torch.from_numpy(d).to(dtype).cuda()


@pytest.mark.parametrize("dtype,scan_mode", [(torch.float32, 0), (torch.float32, 1),
                                             (torch.float64, 0), (torch.float64, 1)])
def test_recompute_tables_vjp_matches_stored_and_oracle(hip_device, oracle_port, dtype, scan_mode):
    """backward_raw(recompute_tables=True) rebuilds every curve's tables from x
    and the density before the VJP (the reference keeps them as saved state,
    cpp/bindings.cpp:132-160); gradients must equal the stored-table VJP's and
    the oracle's, eagerly and replayed from a HIP graph."""
    from quantom_loits_amd.batched import BatchedLOITS, Tuning

    nodes, events, xt, dt = _mixed_curves(dtype)
    n = len(nodes)
    node_off = np.concatenate([[0], np.cumsum(nodes)[:-1]]).astype(np.int64)
    base = (np.arange(n, dtype=np.uint64) * np.uint64(1 << 20)).astype(np.uint64)
    streams = np.arange(n, dtype=np.uint64) * np.uint64(2)
    s = BatchedLOITS(nodes, events, base, streams, None, node_off, dtype=dtype,
                     tuning=Tuning(threads=1024, chunk_events=65536, scan_mode=scan_mode))
    L = s.layout
    print("recompute", dtype, scan_mode, "global", L["global_instances"], "ranges",
          L["range_instances"], flush=True)
    assert L["global_instances"] >= 1 and L["range_instances"] >= 1
    seed = 77
    values, kept, st, _ = s.run_forward(xt, dt, seed)
    gv = torch.from_numpy(np.random.default_rng(3).standard_normal(values.numel())).to(dtype).cuda()
    g_stored = torch.empty_like(dt)
    s.backward_raw(xt, dt, seed, st, gv, g_stored)
    g_re = torch.empty_like(dt)
    s.backward_raw(xt, dt, seed, st, gv, g_re, recompute_tables=True)
    x64, d64 = xt.double().cpu().numpy(), dt.double().cpu().numpy()
    ref = oracle_port.batched_backward(x64, d64, node_off, nodes, events, base, streams, seed,
                                       gv.double().cpu().numpy())
    tol = 1e-9 if dtype == torch.float64 else 1e-4
    for i in range(n):
        sl = slice(int(node_off[i]), int(node_off[i] + nodes[i]))
        scale = np.abs(ref[sl]).max()
        a, b = g_stored.double().cpu().numpy()[sl], g_re.double().cpu().numpy()[sl]
        e_re = np.abs(b - ref[sl]).max() / scale
        e_st = np.abs(a - b).max() / scale
        print(f"  curve {i} nodes {nodes[i]}: recompute vs oracle {e_re:.2e}, "
              f"vs stored {e_st:.2e}", flush=True)
        assert e_re <= tol and e_st <= tol

    print("recompute graph", dtype, scan_mode, flush=True)
    graph = s.capture(xt, dt, seed, recompute_tables=True)
    total = values.numel()
    for r in range(2):
        graph.grad.zero_()
        graph.grad[:total].copy_(gv * (r + 1))
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(graph.values(), values)
        got = graph.grad_density.double().cpu().numpy()
        for i in range(n):
            sl = slice(int(node_off[i]), int(node_off[i] + nodes[i]))
            err = np.abs(got[sl] - (r + 1) * ref[sl]).max() / ((r + 1) * np.abs(ref[sl]).max())
            assert err <= tol, (r, i, err)


def test_bindings_on_an_explicit_side_stream(hip_device):
    """forward/backward given a stream that is not torch's current one: their
    temporaries and host reads are ordered on that stream."""
    from quantom_loits_amd import load_extension

    ext = load_extension()
    rng = np.random.default_rng(9)
    xb, xs, qb, qs, w, acc = _theory(rng, 3, 9, 7, 10)
    E, seed, seq = 500000, 5, 2
    ref = ext.forward(xb, xs, qb, qs, w, acc, E, seed, seq, False)
    g = torch.randn(ref[0].shape, dtype=torch.float64, device="cuda")
    rgx, rgq = ext.backward(g, xb, xs, qb, qs, acc, *ref[1:], False)
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    print("side stream forward/backward", flush=True)
    got = ext.forward(xb, xs, qb, qs, w, acc, E, seed, seq, False, side.cuda_stream)
    gx, gq = ext.backward(g, xb, xs, qb, qs, acc, *got[1:], False, side.cuda_stream)
    torch.cuda.current_stream().wait_stream(side)
    for a, b in zip(got, ref):
        assert torch.equal(a, b)
    assert torch.equal(gx, rgx) and torch.equal(gq, rgq)


def test_bindings_reject_mismatched_dtypes_and_sizes(hip_device):
    from quantom_loits_amd import load_extension
    from quantom_loits_amd.backend import BoundedLOITS

    ext = load_extension()
    rng = np.random.default_rng(2)
    xb, xs, qb, qs, w, acc = _theory(rng, 1, 3, 2, 6)
    b32 = BoundedLOITS(xb, qb, torch.float32, nmax_cap=100, active_cap=200)
    with pytest.raises(ValueError, match="dtype"):
        b32.forward(xb, xs, qb, qs, w, acc, 100, 1, 0)  # float64 inputs, float32 buffers
    st = ext.forward(xb, xs, qb, qs, w, acc, 100, 1, 0, False)
    short = torch.zeros(max(st[0].shape[0] - 1, 0), 2, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="grad_events"):
        ext.backward(short, xb, xs, qb, qs, acc, *st[1:], False)
    with pytest.raises(ValueError, match="u_x"):
        ext.backward(st[0], xb, xs, qb, qs, acc, *st[1:5], st[5].double(), *st[6:], False)


def test_soft_histogram_skips_nan_and_far_values(hip_device):
    from quantom_loits_amd import load_extension

    ext = load_extension()
    gen = torch.Generator(device="cuda").manual_seed(4)
    good = torch.rand(5000, device="cuda", generator=gen) * 0.9 + 0.05
    bad = torch.tensor([float("nan"), float("inf"), -float("inf"), 1e30, -1e30, 3e9, -7e12],
                       device="cuda")
    mixed = torch.cat([good[:2000], bad, good[2000:]])
    rs_mixed = torch.tensor([0, mixed.numel()], dtype=torch.int64, device="cuda")
    rs_good = torch.tensor([0, good.numel()], dtype=torch.int64, device="cuda")
    lo, hi, bins, width = 0.05, 0.95, 64, 0.5
    h_mixed = ext.soft_histogram(mixed, rs_mixed, lo, hi, bins, width)
    h_good = ext.soft_histogram(good, rs_good, lo, hi, bins, width)
    assert torch.allclose(h_mixed, h_good, rtol=1e-5, atol=1e-6)
    gh = torch.randn(1, bins, device="cuda", generator=gen)
    g_mixed = ext.soft_histogram_backward(mixed, rs_mixed, lo, hi, bins, width, gh)
    g_good = ext.soft_histogram_backward(good, rs_good, lo, hi, bins, width, gh)
    assert torch.equal(g_mixed[2000:2000 + bad.numel()], torch.zeros_like(bad))
    assert torch.equal(torch.cat([g_mixed[:2000], g_mixed[2000 + bad.numel():]]), g_good)
