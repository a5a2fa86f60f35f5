"""The benchmark configurations at their stated sizes (GPU).

Every slot of every curve is checked, not a sample: an independent restatement
builds, from the oracle's float64 CDFs (sequential chain, as
cpp/loits_core.cpp:111-165) and the oracle's Philox words
(rng/philox.hpp:17-61), each slot's bin (the count-based interval of
cpp/loits_core.cpp:12-18 as an exact integer search, u >= C_t <=> k >=
ceil(C_t 2^24)), its float64 value (:190-195) and its keep decision
(finite, non-zero, inside the window).  Against it:

* bins of the credited configuration (parallel CDF scan) are compared slot by
  slot and the number of bin flips is printed (the bar is zero);
* kept counts per curve must be equal, and float32 values within 1e-5
  relative of the float64 values, row by row in packed order;
* the oracle itself samples and differentiates 8 curves per configuration at
  full event count (values, bins, and float32 gradients within 1e-4 normwise);
* the VJP at the headline shape is checked against float64 central
  differences (h = 1e-6) on 16 nodes of each of the 256 curves.

Configurations: B=256 N=2048 E=2^18 forward + VJP (the benchmark's step and
tuning), B=1024 N=4096 E=2^16, 4096 ragged curves with the x-window cut
through a captured HIP graph, and one GAN-style training step at B=512
N=2048 E=2^16.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SEED = 1234
WINDOW = (0.05, 0.95)


def thresholds_of(cdf):
    """Integer form of u >= C for u = k 2^-24: k >= T(C) (exact in float64)."""
    c = np.asarray(cdf, dtype=np.float64)
    s = c * 16777216.0
    t = np.where(s >= 16777216.0, 16777216.0, np.ceil(np.where(c <= 0.0, 0.0, s)))
    t = np.where(np.isnan(c), 2.0 ** 32 - 1, t)
    return t.astype(np.int64)


class SlotReference:
    """Bins, float64 values and keep flags of every slot of a batch whose
    curves draw consecutive Philox counters on one stream (counter_base[i] =
    counter_base[0] + sum of the events before curve i)."""

    def __init__(self, oracle, x64, d64, node_off, n_nodes, events, counter_base, stream, seed,
                 window, dev):
        self.oracle, self.seed, self.stream, self.window, self.dev = oracle, seed, stream, window, dev
        n = len(n_nodes)
        self.events = np.asarray(events, dtype=np.int64)
        self.out_off = np.concatenate([[0], np.cumsum(self.events)[:-1]]).astype(np.int64)
        base = np.asarray(counter_base, dtype=np.uint64)
        assert np.array_equal(base - base[0], self.out_off.astype(np.uint64))
        self.base0 = int(base[0])
        zeros = np.zeros(n, dtype=np.int64)
        _, _, _, self.norm, self.cdf = oracle.batched_forward(
            x64, d64, node_off, n_nodes, zeros, base, np.full(n, stream, np.uint64), seed)
        inst_of_node = np.repeat(np.arange(n, dtype=np.int64), n_nodes)
        # thresholds of curve i offset by i 2^25: one ascending array for all curves
        self.T = torch.from_numpy(thresholds_of(self.cdf) + (inst_of_node << 25)).to(dev)
        self.C = torch.from_numpy(self.cdf).to(dev)
        self.x = torch.from_numpy(np.asarray(x64, dtype=np.float64)).to(dev)
        self.node_off = torch.from_numpy(np.asarray(node_off, dtype=np.int64)).to(dev)
        self.n_nodes = torch.from_numpy(np.asarray(n_nodes, dtype=np.int64)).to(dev)
        self.starts = torch.from_numpy(self.out_off).to(dev)
        self.total = int(self.events.sum())

    def chunks(self, size=1 << 26):
        """Yields (first slot, curve, bin, value, keep) over slot ranges, in order."""
        for p0 in range(0, self.total, size):
            p1 = min(self.total, p0 + size)
            c0, c1 = self.base0 + p0, self.base0 + p1
            blk0, blk1 = c0 >> 2, (c1 + 3) >> 2
            words = self.oracle.philox_words(blk0, blk1 - blk0, self.stream, self.seed)
            k = (words[c0 - 4 * blk0:c1 - 4 * blk0] >> 8).astype(np.int64)
            kd = torch.from_numpy(k).to(self.dev)
            slot = torch.arange(p0, p1, device=self.dev)
            inst = torch.searchsorted(self.starts, slot, right=True) - 1
            off = self.node_off[inst]
            cnt = torch.searchsorted(self.T, kd + (inst << 25), right=True) - off
            j = torch.minimum((cnt - 1).clamp(min=0), self.n_nodes[inst] - 2)
            g = off + j
            u = kd.double() * 2.0 ** -24
            c0v, c1v, x0, x1 = self.C[g], self.C[g + 1], self.x[g], self.x[g + 1]
            v = x0 + (x1 - x0) * (1.0 / (c1v - c0v + 1e-5)) * (u - c0v)
            keep = torch.isfinite(v) & (v != 0)
            if self.window is not None:
                keep &= (v >= self.window[0]) & (v <= self.window[1])
            yield p0, inst, j, v, keep


def compare_rows(ref, values, bins=None):
    """Kept counts per curve, row-by-row float32 vs float64 values and (when
    every slot is kept) bins; returns (kept, bin flips, max rel value err)."""
    n = len(ref.events)
    kept = torch.zeros(n, dtype=torch.int64, device=ref.dev)
    row = 0
    flips = 0
    worst = 0.0
    for p0, inst, j, v, keep in ref.chunks():
        kept += torch.bincount(inst[keep], minlength=n)
        vk = v[keep]
        got = values[row:row + vk.numel()].double()
        assert got.numel() == vk.numel()
        worst = max(worst, float(((got - vk).abs() / vk.abs()).max()))
        if bins is not None:
            assert bool(keep.all())
            flips += int((bins[p0:p0 + j.numel()].long() != j).sum())
        row += vk.numel()
    return kept.cpu().numpy(), flips, worst


def oracle_curves(oracle, x64, d64, node_off, n_nodes, events, counter_base, streams, seed, idx,
                  window=None, grad_rows=None):
    """The oracle on a subset `idx` of the curves: (values, bins, kept, grad)."""
    idx = np.asarray(idx)
    nodes = np.asarray(n_nodes)[idx]
    xs = np.concatenate([x64[node_off[i]:node_off[i] + n_nodes[i]] for i in idx])
    ds = np.concatenate([d64[node_off[i]:node_off[i] + n_nodes[i]] for i in idx])
    off = np.concatenate([[0], np.cumsum(nodes)[:-1]]).astype(np.int64)
    desc = (off, nodes, np.asarray(events)[idx], np.asarray(counter_base, np.uint64)[idx],
            np.asarray(streams, np.uint64)[idx])
    v, b, k, _, _ = oracle.batched_forward(xs, ds, *desc, seed, window=window)
    grad = None
    if grad_rows is not None:
        grad = oracle.batched_backward(xs, ds, *desc, seed, grad_rows, window=window)
    return v, b, k, grad, off, nodes


def _gpu(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).cuda()


def _uniform_config(oracle, B, N, E, tuning, with_vjp, data_seed):
    from quantom_loits_amd.batched import BatchedLOITS
    from quantom_loits_amd.workloads import uniform_family

    x, d = uniform_family(B, N, seed=data_seed, dtype=np.float32)
    s = BatchedLOITS.uniform(B, N, E, 0, dtype=torch.float32, tuning=tuning)
    xt, dt = _gpu(x.reshape(-1), torch.float32), _gpu(d.reshape(-1), torch.float32)
    print(f"config B={B} N={N} E={E} tuning={tuning}", flush=True)
    values, kept, st, ex = s.run_forward(xt, dt, SEED, want_bins=True, want_tables=True)
    assert values.numel() == B * E and kept.cpu().tolist() == [E] * B
    bins = ex["bins"]
    assert int(bins.min()) >= 0 and int(bins.max()) <= N - 2
    x64, d64 = x.reshape(-1).astype(np.float64), d.reshape(-1).astype(np.float64)
    node_off = np.arange(B, dtype=np.int64) * N
    n_nodes = np.full(B, N, np.int64)
    events = np.full(B, E, np.int64)
    base = (np.arange(B, dtype=np.uint64) * np.uint64(E)).astype(np.uint64)
    ref = SlotReference(oracle, x64, d64, node_off, n_nodes, events, base, 0, SEED, None, "cuda")
    cdf_err = float((ex["cdf"] - ref.C).abs().max())
    kept_ref, flips, worst = compare_rows(ref, values, bins)
    print(f"  all {B * E} slots: bin flips {flips}, max rel err f32 vs f64 {worst:.2e}, "
          f"max |CDF - oracle CDF| {cdf_err:.1e}", flush=True)
    assert flips == 0 and worst <= 1e-5 and np.all(kept_ref == E)
    rng = np.random.default_rng(B + N)
    idx = np.unique(np.concatenate([[0, B // 2, B - 1], rng.choice(B, 5, replace=False)]))
    grad_rows = gd = None
    if with_vjp:
        gen = torch.Generator(device="cuda").manual_seed(1)
        gv = torch.randn(B * E, device="cuda", generator=gen, dtype=torch.float32)
        gd = torch.empty_like(dt)
        s.backward_raw(xt, dt, SEED, st, gv, gd)
        grad_rows = np.concatenate([gv[i * E:(i + 1) * E].double().cpu().numpy() for i in idx])
    v_o, b_o, k_o, g_o, off, nodes = oracle_curves(oracle, x64, d64, node_off, n_nodes, events,
                                                   base, np.zeros(B, np.uint64), SEED, idx,
                                                   grad_rows=grad_rows)
    got_v = torch.cat([values[i * E:(i + 1) * E] for i in idx]).double().cpu().numpy()
    got_b = torch.cat([bins[i * E:(i + 1) * E] for i in idx]).cpu().numpy()
    assert np.array_equal(got_b, b_o)
    assert np.max(np.abs(got_v - v_o) / np.abs(v_o)) <= 1e-5
    if with_vjp:
        got = gd.double().cpu().numpy()
        for q, i in enumerate(idx):
            r = g_o[off[q]:off[q] + nodes[q]]
            err = np.abs(got[i * N:(i + 1) * N] - r).max() / np.abs(r).max()
            print(f"  curve {i}: gradient normwise rel err vs oracle {err:.2e}", flush=True)
            assert err <= 1e-4
    return s, xt, dt, gd


def test_headline_forward_and_vjp_full_size(hip_device, oracle_port):
    """B=256, N=2048, E=2^18, float32, the benchmark's tuning (1024 threads,
    2^18-event chunks, guide shift 2, parallel CDF scan)."""
    from quantom_loits_amd.batched import Tuning

    _uniform_config(oracle_port, 256, 2048, 1 << 18,
                    Tuning(threads=1024, guide_shift=2, chunk_events=1 << 18, scan_mode=0),
                    with_vjp=True, data_seed=0)


def test_large_grid_forward_full_size(hip_device, oracle_port):
    """B=1024, N=4096, E=2^16, float32 (the benchmark's other uniform shape)."""
    from quantom_loits_amd.batched import Tuning

    _uniform_config(oracle_port, 1024, 4096, 1 << 16, Tuning(chunk_events=1 << 16),
                    with_vjp=True, data_seed=1)


def test_headline_vjp_matches_central_differences_full_size(hip_device):
    """d/d density of L = sum_r g_r v_r at B=256, N=2048, E=2^18 (float64, the
    sequential CDF chain), by central differences with h = 1e-6 on 16 nodes of
    every curve (one node per curve per round: curves are independent).

    An event whose bin differs between density -h and +h jumps in value, and
    a difference quotient across the jump measures the jump; such events get a
    zero cotangent on both sides of the comparison (the quotient leaves them
    out and the VJP is evaluated for the masked cotangent), so every event
    that stays in its bin is checked.
    """
    from quantom_loits_amd.batched import BatchedLOITS, Tuning
    from quantom_loits_amd.workloads import uniform_family

    B, N, E, h = 256, 2048, 1 << 18, 1e-6
    x, d = uniform_family(B, N, seed=0, dtype=np.float64)
    s = BatchedLOITS.uniform(B, N, E, 0, dtype=torch.float64,
                             tuning=Tuning(threads=1024, chunk_events=1 << 18, scan_mode=1))
    xt, dt = _gpu(x.reshape(-1), torch.float64), _gpu(d.reshape(-1), torch.float64)
    gen = torch.Generator(device="cuda").manual_seed(1)
    g = torch.randn(B * E, device="cuda", generator=gen, dtype=torch.float64)
    _, _, st, _ = s.run_forward(xt, dt, SEED)
    full = torch.empty_like(dt)
    s.backward_raw(xt, dt, SEED, st, g, full)
    scale = full.view(B, N).abs().amax(1)
    rng = np.random.default_rng(16)
    base = torch.arange(B, device="cuda") * N
    worst, masked = 0.0, 0
    for r in range(16):
        # one node per curve whose density stays positive at -h
        t = torch.from_numpy(rng.integers(1, N - 1, B)).cuda()
        t = torch.where(dt[base + t] > 10 * h, t, torch.full_like(t, N // 2))
        e = torch.zeros_like(dt)
        e[base + t] = h
        vp, _, _, exp_ = s.run_forward(xt, dt + e, SEED, want_bins=True)
        vp, bp = vp.clone(), exp_["bins"].clone()
        vm, _, _, exm = s.run_forward(xt, dt - e, SEED, want_bins=True)
        keep = bp == exm["bins"]
        masked += int((~keep).sum())
        gm = torch.where(keep, g, torch.zeros_like(g))
        fd = (gm * (vp - vm)).view(B, E).sum(1) / (2 * h)
        vjp = torch.empty_like(dt)
        s.backward_raw(xt, dt, SEED, st, gm, vjp)
        err = ((vjp[base + t] - fd).abs() / scale).max()
        worst = max(worst, float(err))
    print(f"central differences on 16 x {B} nodes: max rel err {worst:.2e}, "
          f"events masked {masked} of {16 * B * E}", flush=True)
    assert worst <= 1e-5


@pytest.mark.parametrize("persistent", [0, 1])
def test_ragged_window_graph_full_size(hip_device, oracle_port, persistent):
    """4096 curves, N log-uniform in [128, 16384], E log-uniform in
    [2^10, 2^20], x-window [0.05, 0.95], forward + pack + VJP replayed from one
    captured HIP graph (the benchmark's ragged workload), with per-kind and
    with persistent launches (the autotuner's choice for this batch): kept
    counts of all 4096 curves and every kept value against the slot
    reference, 8 curves against the oracle."""
    from quantom_loits_amd.batched import BatchedLOITS, Tuning
    from quantom_loits_amd.workloads import ragged_family

    w = ragged_family(4096, n_range=(128, 16384), e_range=(1024, 1 << 20), seed=5)
    n = len(w["n_nodes"])
    base = np.concatenate([[0], np.cumsum(w["events"])[:-1]]).astype(np.uint64)
    s = BatchedLOITS(w["n_nodes"], w["events"], base, None, None, w["node_off"],
                     dtype=torch.float32, window=WINDOW, tuning=Tuning(persistent=persistent))
    xr, dr = _gpu(w["x"], torch.float32), _gpu(w["density"], torch.float32)
    print("ragged graph: persistent", persistent, "curves", n, "events",
          int(w["events"].sum()), "global-table curves", s.layout["global_instances"],
          "bin-range curves", s.layout["range_instances"], flush=True)
    graph = s.capture(xr, dr, SEED)
    total = int(graph.kept.sum())
    gen = torch.Generator(device="cuda").manual_seed(2)
    graph.grad[:total].normal_(generator=gen)
    graph.replay()
    torch.cuda.synchronize()
    values = graph.values()
    kept = graph.kept.cpu().numpy()
    assert values.numel() == total == int(kept.sum())
    x64, d64 = w["x"].astype(np.float64), w["density"].astype(np.float64)
    ref = SlotReference(oracle_port, x64, d64, w["node_off"], w["n_nodes"], w["events"], base, 0,
                        SEED, WINDOW, "cuda")
    kept_ref, _, worst = compare_rows(ref, values)
    print(f"  kept {total}; curves with a kept-count mismatch "
          f"{int((kept_ref != kept).sum())}; max rel err f32 vs f64 {worst:.2e}", flush=True)
    assert np.array_equal(kept_ref, kept) and worst <= 1e-5
    # the oracle on 8 curves whose linear-scan cost (events x nodes) is moderate
    cost = w["events"] * w["n_nodes"]
    rng = np.random.default_rng(6)
    big = np.flatnonzero(w["n_nodes"] > 6000)
    idx = np.unique(np.concatenate([rng.choice(np.flatnonzero(cost < 3e8), 6, replace=False),
                                    big[np.argsort(w["events"][big])[:2]]]))
    rows0 = np.concatenate([[0], np.cumsum(kept)[:-1]])
    gv = graph.grad
    grad_rows = np.concatenate([gv[rows0[i]:rows0[i] + kept[i]].double().cpu().numpy()
                                for i in idx])
    v_o, _, k_o, g_o, off, nodes = oracle_curves(oracle_port, x64, d64, w["node_off"],
                                                 w["n_nodes"], w["events"], base,
                                                 np.zeros(n, np.uint64), SEED, idx, WINDOW,
                                                 grad_rows)
    assert np.array_equal(k_o, kept[idx])
    got_v = np.concatenate([values[rows0[i]:rows0[i] + kept[i]].double().cpu().numpy()
                            for i in idx])
    assert np.max(np.abs(got_v - v_o) / np.abs(v_o)) <= 1e-5
    got_g = graph.grad_density.double().cpu().numpy()
    for q, i in enumerate(idx):
        sl = slice(int(w["node_off"][i]), int(w["node_off"][i] + w["n_nodes"][i]))
        r = g_o[off[q]:off[q] + nodes[q]]
        err = np.abs(got_g[sl] - r).max() / np.abs(r).max()
        print(f"  curve {i} (N={w['n_nodes'][i]}, E={w['events'][i]}): gradient rel err "
              f"{err:.2e}", flush=True)
        assert err <= 1e-4


def test_gan_step_full_size(hip_device):
    """One GAN-style training step at B=512, N=2048, E=2^16 with the x-window
    cut: losses finite, the generator's gradients (through the HIP VJP) finite
    and non-zero, every kept value inside the window."""
    from quantom_loits_amd.training import BatchedGanTrainer

    print("gan step B=512 N=2048 E=2^16", flush=True)
    t = BatchedGanTrainer("hip", "cuda", batch=512, nodes=2048, events=1 << 16, seed=3)
    before = [p.detach().clone() for p in t.generator.parameters()]
    d_loss, g_loss = t.step()
    torch.cuda.synchronize()
    assert torch.isfinite(d_loss) and torch.isfinite(g_loss)
    for p, b in zip(t.generator.parameters(), before):
        assert p.grad is not None and torch.isfinite(p.grad).all()
        assert float((p.detach() - b).abs().max()) > 0
    with torch.no_grad():
        density = t.generator(t.z)
        vals, kept = t._sample(density)
    assert int(kept.sum()) == vals.numel() > 0
    assert float(vals.min()) >= WINDOW[0] and float(vals.max()) <= WINDOW[1]
