"""HIP path vs the oracle and the compiled reference's golden vectors (GPU).

Bars: integer outputs (Philox words, intervals, bin indices, counts, packed
slots, compaction order) bit-exact; float64 drop-in outputs bit-exact against
the reference fixtures; float32 batched values within 1e-5 relative and
gradients within 1e-4 (normwise per instance) of the float64 oracle.
"""
import numpy as np
import pytest
import torch

from conftest import reference_case
from oracle.oracle import uniform_batch

pytestmark = pytest.mark.gpu

CASES = ["manual", "b1_g3_k5", "b3_g2_k10", "b1_g1_k2", "b1_g2_k64", "b1_g1_k1024", "dyadic",
         "acc_zero", "nonfinite", "zero_bins"]


@pytest.fixture(scope="module")
def ext(hip_device):
    from quantom_loits_amd import load_extension

    return load_extension()


def test_device_is_gfx950(ext):
    arch, cus = ext.device_info(0)
    print("device", arch, cus)
    assert arch.startswith("gfx950") and cus > 0


def test_philox_words_match_compiled_reference(ext, golden_philox):
    grid, words = golden_philox["grid"], golden_philox["words"]
    for (b, s, k), w in zip(grid, words):
        if int(b) > 2 ** 63 - 5:
            continue  # first_block is int64 at the C ABI
        got = ext.philox_words(int(b), 1, int(s), int(k)).cpu().numpy().view(np.uint32)
        assert np.array_equal(got, w), (b, s, k)


def test_philox_kats(ext, golden_philox):
    # tests/test_native_rng.py:43-55 and the Random123 known answers
    got = ext.philox_words(0, 1, 0, 0).cpu().numpy().view(np.uint32)
    assert got.tolist() == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    for row in golden_philox["kat"][1:]:
        c, k, e = row[:4], row[4:6], row[6:]
        block = int(c[0]) | (int(c[1]) << 32)
        if block > 2 ** 63 - 1:
            continue
        got = ext.philox_words(block, 1, int(c[2]) | (int(c[3]) << 32),
                               int(k[0]) | (int(k[1]) << 32)).cpu().numpy().view(np.uint32)
        assert np.array_equal(got, e)


def test_philox_fill_matches_reference_fill(ext, golden_philox):
    for key in golden_philox.files:
        if key.startswith("fill_"):
            _, n, seed, stream = key.split("_")
            got = ext.philox_fill(int(n), int(seed), int(stream)).cpu().numpy()
            assert np.array_equal(got, golden_philox[key]), key


def _to_dev(arrays, dtype):
    out = []
    for a in arrays:
        if a.dtype == bool:
            out.append(torch.from_numpy(np.ascontiguousarray(a)).cuda())
        else:
            out.append(torch.from_numpy(np.ascontiguousarray(a)).to(dtype).cuda())
    return out


@pytest.mark.parametrize("name", CASES)
def test_dropin_forward_backward_bit_exact_vs_reference(ext, golden_reference, name):
    c = reference_case(golden_reference, name)
    n_events, seed, sequence = (int(v) for v in c["scalars"])
    args = _to_dev([c["x_bins"], c["xsec_x"], c["q_bins"], c["xsec_q"], c["weights"],
                    c["acceptance"]], torch.float64)
    print("case", name, flush=True)
    st = ext.forward(*args, n_events, seed, sequence, False)
    events, norm_x, norm_q, cdf_x, cdf_q, u_x, u_q, ix, iq, packed, row_offsets = st
    np.testing.assert_array_equal(events.cpu().numpy(), c["events"])
    for got, key in ((norm_x, "norm_x"), (norm_q, "norm_q"), (cdf_x, "cdf_x"), (cdf_q, "cdf_q"),
                     (packed, "packed"), (row_offsets, "row_offsets")):
        np.testing.assert_array_equal(got.cpu().numpy(), c[key], err_msg=key)
    counts = c["counts"].reshape(-1)
    nmax = u_x.shape[2]
    active = np.zeros(u_x.numel(), dtype=bool)
    for g, cnt in enumerate(counts):
        active[g * nmax: g * nmax + cnt] = True
    assert np.array_equal(ix.cpu().numpy()[active], c["interval_x"])
    assert np.array_equal(iq.cpu().numpy()[active], c["interval_q"])
    if "u_x" in c:
        assert np.array_equal(u_x.cpu().numpy(), c["u_x"])
        assert np.array_equal(u_q.cpu().numpy(), c["u_q"])
    g = torch.from_numpy(c["grad_events"]).cuda()
    gx, gq = ext.backward(g, args[0], args[1], args[2], args[3], args[5], norm_x, norm_q, cdf_x,
                          cdf_q, u_x, u_q, ix, iq, packed, row_offsets, False)
    np.testing.assert_array_equal(gx.cpu().numpy(), c["grad_xsec_x"])
    np.testing.assert_array_equal(gq.cpu().numpy(), c["grad_xsec_q"])


def test_dropin_float32_inputs_close_to_float64_reference(ext, golden_reference):
    c = reference_case(golden_reference, "b3_g2_k10")
    n_events, seed, sequence = (int(v) for v in c["scalars"])
    raw = [c["x_bins"], c["xsec_x"], c["q_bins"], c["xsec_q"], c["weights"], c["acceptance"]]
    # float32-representable inputs so that the float64 oracle sees the same data
    raw32 = [a if a.dtype == bool else a.astype(np.float32).astype(np.float64) for a in raw]
    from oracle.oracle import Oracle
    ref = Oracle("port").forward(*raw32, n_events, seed, sequence)
    st = ext.forward(*_to_dev(raw32, torch.float32), n_events, seed, sequence, False)
    assert np.array_equal(st[9].cpu().numpy(), ref["packed"])
    np.testing.assert_allclose(st[0].cpu().numpy().astype(np.float64), ref["events"], rtol=1e-6)


@pytest.mark.parametrize("kx,kq,zero_node", [(3, 29, False), (7, 2, False), (5, 12, True),
                                             (40, 3, False), (60, 10, False)])
def test_dropin_uneven_axes_bit_exact_vs_oracle(ext, oracle_port, kx, kq, zero_node):
    """Different node counts per axis: the interpolation VJP with a lane per
    node (kx + kq <= 64; two cells per wave, or one for 40 + 3), the one with a
    lane per node pair (60 + 10) and both compaction orders: drop-free curves
    (rows from the cell counts), and a node at x = 0 plus a NaN density, whose
    cell's slots are all dropped (look-back order)."""
    rng = np.random.default_rng(kx * 100 + kq)
    B, nx, ny = 2, 3, 4
    xb = np.sort(rng.uniform(0.01, 1.0, (B, nx, kx)), axis=2)
    if zero_node:
        xb[0, 1, 0] = 0.0
    qb = np.sort(rng.uniform(1.0, 50.0, (B, ny, kq)), axis=2)
    xs = rng.uniform(0.1, 1.0, (B, ny, nx, kx))
    if zero_node:
        xs[1, 2, 0, 1] = np.nan
    qs = rng.uniform(0.1, 1.0, (B, nx, ny, kq))
    w = rng.uniform(0.5, 1.5, (B, nx, ny))
    w /= w.sum()
    acc = rng.uniform(0, 1, (B, nx, ny)) < 0.9
    n_events, seed, sequence = 60000, 11, 3
    ref = oracle_port.forward(xb, xs, qb, qs, w, acc, n_events, seed, sequence)
    print("uneven axes", kx, kq, zero_node, flush=True)
    args = _to_dev([xb, xs, qb, qs, w, acc], torch.float64)
    st = ext.forward(*args, n_events, seed, sequence, False)
    events, norm_x, norm_q, cdf_x, cdf_q, u_x, u_q, ix, iq, packed, row_offsets = st
    np.testing.assert_array_equal(events.cpu().numpy(), ref["events"])
    np.testing.assert_array_equal(packed.cpu().numpy(), ref["packed"])
    np.testing.assert_array_equal(row_offsets.cpu().numpy(), ref["row_offsets"])
    if zero_node:
        assert len(ref["packed"]) < int(ref["active"])  # some slots were dropped
    g = rng.standard_normal(ref["events"].shape)
    gx_ref, gq_ref = oracle_port.backward(g, xb, xs, qb, qs, acc, ref)
    gx, gq = ext.backward(torch.from_numpy(g).cuda(), args[0], args[1], args[2], args[3], args[5],
                          norm_x, norm_q, cdf_x, cdf_q, u_x, u_q, ix, iq, packed, row_offsets,
                          False)
    np.testing.assert_array_equal(gx.cpu().numpy(), gx_ref)
    np.testing.assert_array_equal(gq.cpu().numpy(), gq_ref)


@pytest.mark.parametrize("n_events", [1000000, 10000000])
def test_dropin_gan_shape_bit_exact_vs_oracle(ext, oracle_port, n_events):
    """The reference's GAN shape on the reference-API path: 99 x 99 cells of
    10 nodes per axis, float64, a random acceptance mask.  Every curve is
    drop-free here, so this covers the one-wave sampling tiles whose slots
    span cell boundaries, rows placed from the cell counts, and the VJP with
    two cells per wave over about 100 (1e6) or 1000 (1e7) rows per cell."""
    rng = np.random.default_rng(7)
    grid, k = 99, 10
    xb = np.sort(rng.uniform(0.001, 0.999, (1, grid, k)), axis=2)
    qb = np.sort(rng.uniform(1.0, 100.0, (1, grid, k)), axis=2)
    xs = rng.uniform(0.1, 1.0, (1, grid, grid, k))
    qs = rng.uniform(0.1, 1.0, (1, grid, grid, k))
    w = rng.uniform(0.5, 1.5, (1, grid, grid))
    w /= w.sum()
    acc = rng.uniform(0, 1, (1, grid, grid)) < 0.95
    seed, sequence = 42, 5
    ref = oracle_port.forward(xb, xs, qb, qs, w, acc, n_events, seed, sequence)
    assert len(ref["packed"]) == int(ref["active"])  # drop-free
    print("gan shape", n_events, flush=True)
    args = _to_dev([xb, xs, qb, qs, w, acc], torch.float64)
    st = ext.forward(*args, n_events, seed, sequence, False)
    events, norm_x, norm_q, cdf_x, cdf_q, u_x, u_q, ix, iq, packed, row_offsets = st
    np.testing.assert_array_equal(packed.cpu().numpy(), ref["packed"])
    np.testing.assert_array_equal(events.cpu().numpy(), ref["events"])
    np.testing.assert_array_equal(row_offsets.cpu().numpy(), ref["row_offsets"])
    g = rng.standard_normal(ref["events"].shape)
    gx_ref, gq_ref = oracle_port.backward(g, xb, xs, qb, qs, acc, ref)
    gx, gq = ext.backward(torch.from_numpy(g).cuda(), args[0], args[1], args[2], args[3], args[5],
                          norm_x, norm_q, cdf_x, cdf_q, u_x, u_q, ix, iq, packed, row_offsets,
                          False)
    np.testing.assert_array_equal(gx.cpu().numpy(), gx_ref)
    np.testing.assert_array_equal(gq.cpu().numpy(), gq_ref)


def _batched_inputs(g, dtype):
    x = torch.from_numpy(np.ascontiguousarray(g["x"])).to(dtype).cuda().reshape(-1)
    d = torch.from_numpy(np.ascontiguousarray(g["density"])).to(dtype).cuda().reshape(-1)
    return x, d


@pytest.mark.parametrize("scan_mode", [1, 0])
def test_batched_float64_matches_config0_fixture(golden_config0, hip_device, scan_mode):
    from quantom_loits_amd.batched import BatchedLOITS, Tuning

    g = golden_config0
    B, N = g["x"].shape
    E = int(g["n_events"])
    print("batched f64 scan_mode", scan_mode, flush=True)
    s = BatchedLOITS.uniform(B, N, E, int(g["sequence"]), dtype=torch.float64,
                             tuning=Tuning(threads=256, chunk_events=1024, scan_mode=scan_mode))
    x, d = _batched_inputs(g, torch.float64)
    values, kept, st, ex = s.run_forward(x, d, int(g["seed"]), want_bins=True, want_tables=True)
    assert kept.cpu().tolist() == [E] * B
    assert np.array_equal(ex["bins"].cpu().numpy(), g["bins"])
    if scan_mode == 1:  # the reference's sequential chain: bit-exact
        assert np.array_equal(values.cpu().numpy(), g["values"])
        assert np.array_equal(ex["cdf"].cpu().numpy().reshape(B, N), g["cdf"])
        assert np.array_equal(ex["norm"].cpu().numpy(), g["norm"])
    else:
        np.testing.assert_allclose(values.cpu().numpy(), g["values"], rtol=1e-12)
    grad = torch.from_numpy(g["grad_values"]).cuda()
    gd = torch.empty_like(d)
    s.backward_raw(x, d, int(g["seed"]), st, grad, gd)
    ref = g["grad_density"]
    got = gd.cpu().numpy().reshape(B, N)
    err = np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)
    print("f64 grad normwise rel err", err.max())
    assert err.max() < 1e-10


@pytest.mark.parametrize("threads,chunk,fmt", [(1024, 1 << 18, 0), (256, 4096, 0), (512, 2048, 0),
                                              (1024, 1 << 18, 1), (256, 4096, 1)])
def test_batched_float32_matches_config0_fixture(golden_config0, hip_device, threads, chunk, fmt):
    """float32 against the compiled reference's config-0 slice, on cell tables
    (fmt 0) and on guide tables (fmt 1)."""
    from quantom_loits_amd.batched import BatchedLOITS, Tuning

    g = golden_config0
    B, N = g["x"].shape
    E = int(g["n_events"])
    print("batched f32", threads, chunk, fmt, flush=True)
    s = BatchedLOITS.uniform(B, N, E, int(g["sequence"]), dtype=torch.float32,
                             tuning=Tuning(threads=threads, chunk_events=chunk, table_format=fmt))
    assert s.layout["cell_instances"] == (B if fmt == 0 else 0)
    x, d = _batched_inputs(g, torch.float32)
    values, kept, st, ex = s.run_forward(x, d, int(g["seed"]), want_bins=True)
    assert np.array_equal(ex["bins"].cpu().numpy(), g["bins"])
    v = values.cpu().numpy().astype(np.float64)
    rel = np.abs(v - g["values"]) / np.abs(g["values"])
    print("f32 value rel err", rel.max())
    assert rel.max() <= 1e-5
    grad = torch.from_numpy(g["grad_values"].astype(np.float32)).cuda()
    gd = torch.empty_like(d)
    s.backward_raw(x, d, int(g["seed"]), st, grad, gd)
    ref = g["grad_density"]
    got = gd.cpu().numpy().astype(np.float64).reshape(B, N)
    err = np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)
    print("f32 grad normwise rel err", err.max())
    assert err.max() <= 1e-4


def test_batched_ragged_window_matches_oracle(oracle_port, hip_device):
    from quantom_loits_amd.batched import BatchedLOITS, Tuning
    from quantom_loits_amd.workloads import ragged_family

    w = ragged_family(12, n_range=(16, 2048), e_range=(1, 20000), seed=3, dtype=np.float64)
    n = len(w["n_nodes"])
    rng = np.random.default_rng(4)
    base = rng.integers(0, 2 ** 40, n).astype(np.uint64)  # unaligned counter bases
    streams = rng.integers(0, 2 ** 33, n).astype(np.uint64)
    accept = np.ones(n, np.uint8)
    accept[3] = 0
    window = (0.05, 0.95)
    ref_v, ref_b, ref_k, _, _ = oracle_port.batched_forward(
        w["x"], w["density"], w["node_off"], w["n_nodes"], w["events"], base, streams, 99,
        accept=accept, window=window)
    for dtype, fmt in ((torch.float64, 0), (torch.float32, 0), (torch.float32, 1)):
        print("ragged window", dtype, fmt, flush=True)
        s = BatchedLOITS(w["n_nodes"], w["events"], base, streams, accept, w["node_off"],
                         dtype=dtype, window=window,
                         tuning=Tuning(threads=256, chunk_events=4096, scan_mode=1,
                                       table_format=fmt))
        x = torch.from_numpy(w["x"]).to(dtype).cuda()
        d = torch.from_numpy(w["density"]).to(dtype).cuda()
        if dtype == torch.float32:  # the oracle must see the same float32 data
            x64, d64 = x.double().cpu().numpy(), d.double().cpu().numpy()
            ref_v, ref_b, ref_k, _, _ = oracle_port.batched_forward(
                x64, d64, w["node_off"], w["n_nodes"], w["events"], base, streams, 99,
                accept=accept, window=window)
        values, kept, st, ex = s.run_forward(x, d, 99, want_bins=True)
        assert kept.cpu().numpy().tolist() == ref_k.tolist()
        assert np.array_equal(ex["bins"].cpu().numpy(), ref_b)
        v = values.cpu().numpy().astype(np.float64)
        if dtype == torch.float64:
            assert np.array_equal(v, ref_v)
        else:
            assert np.max(np.abs(v - ref_v) / np.abs(ref_v)) <= 1e-5
        rng2 = np.random.default_rng(5)
        gv = rng2.standard_normal(len(ref_v))
        ref_g = oracle_port.batched_backward(x.double().cpu().numpy(), d.double().cpu().numpy(),
                                             w["node_off"], w["n_nodes"], w["events"], base,
                                             streams, 99, gv, accept=accept, window=window)
        gd = torch.empty_like(d)
        s.backward_raw(x, d, 99, st, torch.from_numpy(gv).to(dtype).cuda(), gd)
        got = gd.double().cpu().numpy()
        for i in range(n):
            sl = slice(int(w["node_off"][i]), int(w["node_off"][i] + w["n_nodes"][i]))
            scale = np.abs(ref_g[sl]).max()
            if scale == 0:
                assert np.abs(got[sl]).max() == 0
                continue
            err = np.abs(got[sl] - ref_g[sl]).max() / scale
            assert err <= (1e-10 if dtype == torch.float64 else 1e-4), (i, err)


@pytest.mark.parametrize("fmt", [0, 1])
def test_window_keep_classes_match_oracle(oracle_port, hip_device, fmt):
    """Window cuts whose per-bin keep classes (all kept / none kept / per value)
    meet their edge cases: window edges exactly on nodes, decreasing x, curves
    wholly inside or outside the window, bins straddling an edge by 1e-7, and
    a non-monotone CDF (negative densities; classes off).  float32, bins and
    kept counts exact, values within 1e-5 of the oracle on the same data."""
    from quantom_loits_amd.batched import BatchedLOITS, Tuning

    window = (0.25, 0.75)
    grid = np.linspace(0.0, 1.0, 257)  # holds 0.25 and 0.75 exactly
    rng = np.random.default_rng(31)
    xs = [grid, grid[::-1].copy(), np.linspace(0.3, 0.7, 200), np.linspace(2.0, 3.0, 150),
          0.25 + 1e-7 * np.arange(-60, 61), grid]
    ds = [rng.uniform(0.2, 2.0, len(v)) for v in xs]
    ds[5] = np.sin(np.linspace(0.0, 9.0, len(grid))) + 0.3  # dips below zero
    n_nodes = np.array([len(v) for v in xs], np.int64)
    events = np.full(len(xs), 20000, np.int64)
    node_off = np.concatenate([[0], np.cumsum(n_nodes)[:-1]]).astype(np.int64)
    base = np.concatenate([[0], np.cumsum(events)[:-1]]).astype(np.uint64)
    streams = np.arange(len(xs), dtype=np.uint64) * 2
    x32 = np.concatenate(xs).astype(np.float32)
    d32 = np.concatenate(ds).astype(np.float32)
    ref_v, ref_b, ref_k, _, _ = oracle_port.batched_forward(
        x32.astype(np.float64), d32.astype(np.float64), node_off, n_nodes, events, base,
        streams, 77, window=window)
    assert ref_k[3] == 0 and ref_k[2] > 0
    s = BatchedLOITS(n_nodes, events, base, streams, None, node_off, dtype=torch.float32,
                     window=window, tuning=Tuning(threads=256, chunk_events=4096, scan_mode=1,
                                                  table_format=fmt))
    if fmt == 0:
        assert s.layout["cell_instances"] == len(xs)
    values, kept, _, ex = s.run_forward(torch.from_numpy(x32).cuda(), torch.from_numpy(d32).cuda(),
                                        77, want_bins=True)
    assert kept.cpu().numpy().tolist() == ref_k.tolist()
    assert np.array_equal(ex["bins"].cpu().numpy(), ref_b)
    v = values.cpu().numpy().astype(np.float64)
    assert np.max(np.abs(v - ref_v) / np.maximum(np.abs(ref_v), 1e-30)) <= 1e-5


@pytest.mark.parametrize("dtype,fmt", [(torch.float64, 0), (torch.float32, 0), (torch.float32, 1)])
def test_batched_tiny_curves_and_event_counts_match_oracle(oracle_port, hip_device, dtype, fmt):
    """Two- and three-node curves (one and two bins) and event counts at the
    edges of the Philox blocks, segments and chunks: none, one, 63/64/65 slots
    (a partial 4-word block, one wave of blocks, one over), several chunks."""
    from quantom_loits_amd.batched import BatchedLOITS, Tuning

    n_nodes = np.array([2, 3, 2, 5, 3, 64, 2, 17], np.int64)
    events = np.array([0, 1, 3, 63, 64, 65, 4097, 2500], np.int64)
    rng = np.random.default_rng(21)
    xs, ds = [], []
    for n in n_nodes:
        xs.append(np.sort(rng.uniform(0.1, 3.0, n)))  # x > 0: relative value errors
        ds.append(rng.uniform(0.1, 2.0, n))
    x64, d64 = np.concatenate(xs), np.concatenate(ds)
    if dtype == torch.float32:
        x64, d64 = x64.astype(np.float32).astype(np.float64), d64.astype(np.float32).astype(np.float64)
    node_off = np.concatenate([[0], np.cumsum(n_nodes)[:-1]]).astype(np.int64)
    base = rng.integers(0, 2 ** 40, len(n_nodes)).astype(np.uint64)
    streams = rng.integers(0, 2 ** 33, len(n_nodes)).astype(np.uint64)
    accept = np.ones(len(n_nodes), np.uint8)
    ref_v, ref_b, ref_k, _, _ = oracle_port.batched_forward(x64, d64, node_off, n_nodes, events,
                                                            base, streams, 7, accept=accept)
    assert ref_k.tolist() == events.tolist()
    print("tiny curves", dtype, fmt, flush=True)
    s = BatchedLOITS(n_nodes, events, base, streams, accept, node_off, dtype=dtype,
                     tuning=Tuning(threads=256, chunk_events=1024, scan_mode=1, table_format=fmt))
    x = torch.from_numpy(x64).to(dtype).cuda()
    d = torch.from_numpy(d64).to(dtype).cuda()
    values, kept, st, ex = s.run_forward(x, d, 7, want_bins=True)
    assert kept.cpu().numpy().tolist() == ref_k.tolist()
    assert np.array_equal(ex["bins"].cpu().numpy(), ref_b)
    v = values.cpu().numpy().astype(np.float64)
    if dtype == torch.float64:
        assert np.array_equal(v, ref_v)
    else:
        assert np.max(np.abs(v - ref_v) / np.abs(ref_v)) <= 1e-5
    gv = np.random.default_rng(22).standard_normal(len(ref_v))
    ref_g = oracle_port.batched_backward(x64, d64, node_off, n_nodes, events, base, streams, 7,
                                         gv, accept=accept)
    gd = torch.empty_like(d)
    s.backward_raw(x, d, 7, st, torch.from_numpy(gv).to(dtype).cuda(), gd)
    got = gd.double().cpu().numpy()
    rows = np.concatenate([[0], np.cumsum(ref_k)])
    for i in range(len(n_nodes)):
        sl = slice(int(node_off[i]), int(node_off[i] + n_nodes[i]))
        if events[i] == 0:  # no events: no gradient
            assert np.abs(got[sl]).max() == 0, i
            continue
        # rounding floor: eps-sized terms |g| |x| summed over the events
        floor = 1e-12 * np.abs(gv[rows[i]:rows[i + 1]]).sum() * np.abs(x64[sl]).max()
        if n_nodes[i] == 2:  # one bin spanning C = 0..1: events do not depend on the density
            assert np.abs(got[sl]).max() <= floor and np.abs(ref_g[sl]).max() <= floor, i
            continue
        err = np.abs(got[sl] - ref_g[sl]).max() / np.abs(ref_g[sl]).max()
        assert err <= (1e-10 if dtype == torch.float64 else 1e-4), (i, err)


@pytest.mark.parametrize("dtype,fmt", [(torch.float64, 0), (torch.float32, 0), (torch.float32, 1)])
def test_batched_nonfinite_densities_match_oracle(oracle_port, hip_device, dtype, fmt):
    """All-zero, infinite and NaN densities (norm 0, inf, NaN): no finite events
    survive compaction and the density adjoint is NaN over the whole curve, as
    in the reference's cores; a neighbouring curve with a negative node is
    unaffected."""
    from quantom_loits_amd.batched import BatchedLOITS, Tuning

    n_nodes = np.array([5, 5, 5, 4], np.int64)
    events = np.array([100, 100, 100, 50], np.int64)
    x64 = np.concatenate([np.linspace(0.2, 1.0, 5)] * 3 + [np.linspace(0.2, 1.0, 4)])
    d64 = np.concatenate([np.zeros(5), [1, 2, np.inf, 1, 1], [1, np.nan, 1, 1, 1], [1, -1, 1, 1]])
    if dtype == torch.float32:
        x64 = x64.astype(np.float32).astype(np.float64)
    node_off = np.array([0, 5, 10, 15], np.int64)
    base = np.zeros(4, np.uint64)
    streams = np.arange(4, dtype=np.uint64)
    accept = np.ones(4, np.uint8)
    ref_v, ref_b, ref_k, _, _ = oracle_port.batched_forward(x64, d64, node_off, n_nodes, events,
                                                            base, streams, 3, accept=accept)
    assert ref_k.tolist() == [0, 0, 0, 50]
    print("non-finite densities", dtype, fmt, flush=True)
    s = BatchedLOITS(n_nodes, events, base, streams, accept, node_off, dtype=dtype,
                     tuning=Tuning(threads=256, chunk_events=1024, scan_mode=1, table_format=fmt))
    x = torch.from_numpy(x64).to(dtype).cuda()
    d = torch.from_numpy(d64).to(dtype).cuda()
    values, kept, st, ex = s.run_forward(x, d, 3, want_bins=True)
    assert kept.cpu().numpy().tolist() == ref_k.tolist()
    assert np.array_equal(ex["bins"].cpu().numpy(), ref_b)
    v = values.cpu().numpy().astype(np.float64)
    if dtype == torch.float64:
        assert np.array_equal(v, ref_v)
    else:
        assert np.max(np.abs(v - ref_v) / np.abs(ref_v)) <= 1e-5
    gv = np.linspace(-1.0, 1.0, len(ref_v))
    ref_g = oracle_port.batched_backward(x64, d64, node_off, n_nodes, events, base, streams, 3,
                                         gv, accept=accept)
    gd = torch.empty_like(d)
    s.backward_raw(x, d, 3, st, torch.from_numpy(gv).to(dtype).cuda(), gd)
    got = gd.double().cpu().numpy()
    assert np.array_equal(np.isnan(got), np.isnan(ref_g))
    ok = ~np.isnan(ref_g)
    scale = np.abs(ref_g[ok]).max()
    tol = 1e-10 if dtype == torch.float64 else 1e-4
    assert np.abs(got[ok] - ref_g[ok]).max() <= tol * scale


def test_batched_large_curves_global_tables_match_oracle(oracle_port, hip_device):
    """Curves too large for LDS (global tables) mixed with LDS-resident ones,
    several chunks per curve, a window cut; both precisions."""
    from quantom_loits_amd.batched import BatchedLOITS, Tuning
    from quantom_loits_amd.workloads import ragged_family

    w = ragged_family(6, n_range=(2500, 16384), e_range=(20000, 150000), seed=11,
                      dtype=np.float64)
    n = len(w["n_nodes"])
    rng = np.random.default_rng(12)
    base = rng.integers(0, 2 ** 40, n).astype(np.uint64)
    streams = rng.integers(0, 2 ** 33, n).astype(np.uint64)
    accept = np.ones(n, np.uint8)
    window = (0.05, 0.95)
    for dtype, scan in ((torch.float64, 1), (torch.float32, 0)):
        print("large curves", dtype, w["n_nodes"].tolist(), w["events"].tolist(), flush=True)
        s = BatchedLOITS(w["n_nodes"], w["events"], base, streams, accept, w["node_off"],
                         dtype=dtype, window=window,
                         tuning=Tuning(threads=1024, chunk_events=65536, scan_mode=scan))
        assert s.layout["global_instances"] > 0
        x = torch.from_numpy(w["x"]).to(dtype).cuda()
        d = torch.from_numpy(w["density"]).to(dtype).cuda()
        x64, d64 = x.double().cpu().numpy(), d.double().cpu().numpy()
        ref_v, ref_b, ref_k, _, _ = oracle_port.batched_forward(
            x64, d64, w["node_off"], w["n_nodes"], w["events"], base, streams, 7,
            accept=accept, window=window)
        values, kept, st, ex = s.run_forward(x, d, 7, want_bins=True)
        assert kept.cpu().numpy().tolist() == ref_k.tolist()
        assert np.array_equal(ex["bins"].cpu().numpy(), ref_b)
        v = values.cpu().numpy().astype(np.float64)
        if dtype == torch.float64:
            assert np.array_equal(v, ref_v)
        else:
            assert np.max(np.abs(v - ref_v) / np.abs(ref_v)) <= 1e-5
        gv = np.random.default_rng(13).standard_normal(len(ref_v))
        ref_g = oracle_port.batched_backward(x64, d64, w["node_off"], w["n_nodes"], w["events"],
                                             base, streams, 7, gv, accept=accept, window=window)
        gd = torch.empty_like(d)
        s.backward_raw(x, d, 7, st, torch.from_numpy(gv).to(dtype).cuda(), gd)
        got = gd.double().cpu().numpy()
        for i in range(n):
            sl = slice(int(w["node_off"][i]), int(w["node_off"][i] + w["n_nodes"][i]))
            scale = np.abs(ref_g[sl]).max()
            err = np.abs(got[sl] - ref_g[sl]).max() / scale
            assert err <= (1e-9 if dtype == torch.float64 else 1e-4), (i, err)


def test_public_autograd_path(hip_device):
    from quantom_loits_amd import LOITS
    from quantom_loits_amd.workloads import reference_1d_inputs, uniform_family

    torch.manual_seed(17)
    x, d = uniform_family(2, 16, seed=1, dtype=np.float64)
    outs = [torch.from_numpy(a).cuda() for a in reference_1d_inputs(x, d)]
    xsec = outs[1].clone().requires_grad_(True)
    outs[1] = xsec
    sampler = LOITS(backend="hip", device="cuda")
    events = sampler(tuple(outs), 50)
    events.square().mean().backward()
    assert events.ndim == 2 and events.shape[1] == 2 and events.shape[0] == 100
    assert xsec.grad is not None and torch.isfinite(xsec.grad).all()
    assert sampler.impl.sequence == 1


def test_batched_autograd_function(hip_device, oracle_port):
    from quantom_loits_amd.batched import BatchedLOITS, Tuning
    from quantom_loits_amd.workloads import uniform_family

    B, N, E = 3, 128, 3000
    x, d = uniform_family(B, N, seed=6, dtype=np.float64)
    s = BatchedLOITS.uniform(B, N, E, 2, dtype=torch.float64, tuning=Tuning(threads=256,
                                                                           chunk_events=1024))
    xt = torch.from_numpy(x.reshape(-1)).cuda()
    dt = torch.from_numpy(d.reshape(-1)).cuda().requires_grad_(True)
    vals = s(xt, dt, 7)
    w = torch.randn_like(vals)
    (vals * w).sum().backward()
    desc = uniform_batch(B, N, E, 2)
    ref = oracle_port.batched_backward(x.reshape(-1), d.reshape(-1), *desc, 7,
                                       w.detach().cpu().numpy())
    np.testing.assert_allclose(dt.grad.cpu().numpy(), ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


@pytest.mark.parametrize("dtype,fmt", [(torch.float64, 0), (torch.float32, 0), (torch.float32, 1)])
def test_batched_non_monotone_cdf_uses_reference_count(oracle_port, hip_device, dtype, fmt):
    """Negative densities make the CDF decrease; the reference's interval is then
    the count #{t : u >= C_t} - 1 (cpp/loits_core.cpp:12-18), not a sorted search.
    The VJP of such a curve (t = u - C_j may be negative) against the oracle too."""
    from quantom_loits_amd.batched import BatchedLOITS, Tuning

    B, N, E = 3, 300, 20000
    x = np.tile(np.linspace(0.001, 0.999, N), (B, 1))
    dens = 1.0 + 0.3 * np.sin(np.linspace(0, 12, N))[None, :] * np.array([[1.0], [4.0], [6.0]])
    assert (dens < 0).any()
    if dtype == torch.float32:
        x, dens = x.astype(np.float32).astype(np.float64), dens.astype(np.float32).astype(np.float64)
    desc = uniform_batch(B, N, E, 1)
    ref_v, ref_b, ref_k, _, _ = oracle_port.batched_forward(x.reshape(-1), dens.reshape(-1),
                                                            *desc, 5)
    s = BatchedLOITS.uniform(B, N, E, 1, dtype=dtype, tuning=Tuning(threads=256, chunk_events=8192,
                                                                    scan_mode=1,
                                                                    table_format=fmt))
    xt = torch.from_numpy(x.reshape(-1)).to(dtype).cuda()
    dt = torch.from_numpy(dens.reshape(-1)).to(dtype).cuda()
    values, kept, st, ex = s.run_forward(xt, dt, 5, want_bins=True)
    assert kept.cpu().numpy().tolist() == ref_k.tolist()
    assert np.array_equal(ex["bins"].cpu().numpy(), ref_b)
    v = values.cpu().numpy().astype(np.float64)
    if dtype == torch.float64:
        assert np.array_equal(v, ref_v)
    else:
        assert np.max(np.abs(v - ref_v) / np.abs(ref_v)) <= 1e-5
    gv = np.random.default_rng(12).standard_normal(len(ref_v))
    ref_g = oracle_port.batched_backward(x.reshape(-1), dens.reshape(-1), *desc, 5,
                                         gv).reshape(B, N)
    gd = torch.empty_like(dt)
    s.backward_raw(xt, dt, 5, st, torch.from_numpy(gv).to(dtype).cuda(), gd)
    got = gd.double().cpu().numpy().reshape(B, N)
    err = np.abs(got - ref_g).max(axis=1) / np.abs(ref_g).max(axis=1)
    print("non-monotone", dtype, fmt, "grad rel err", err, flush=True)
    assert err.max() <= (1e-9 if dtype == torch.float64 else 1e-4)


@pytest.mark.parametrize("guide_shift,dtype,fmt", [(0, torch.float64, 0), (1, torch.float64, 0),
                                                   (0, torch.float32, 0), (1, torch.float32, 1)])
def test_batched_spiky_densities_exercise_guide_overflow(oracle_port, hip_device, guide_shift,
                                                         dtype, fmt):
    """Narrow peaks on a tiny floor put most thresholds into a few crowded guide
    cells, more than level 2 of the guide (or of the cell tables) can hold: the
    bisection fallback must give the reference's bins exactly."""
    from quantom_loits_amd.batched import BatchedLOITS, Tuning

    B, N, E = 4, 2048, 40000
    xg = np.linspace(0.001, 0.999, N)
    rng = np.random.default_rng(11)
    dens = []
    for b in range(B):
        d = np.full(N, 1e-9)
        for mu in rng.uniform(0.1, 0.9, 6):
            d += np.exp(-0.5 * ((xg - mu) / rng.uniform(1e-4, 2e-3)) ** 2)
        dens.append(d)
    x = np.tile(xg, (B, 1))
    dens = np.array(dens)
    if dtype == torch.float32:
        x, dens = x.astype(np.float32).astype(np.float64), dens.astype(np.float32).astype(np.float64)
    desc = uniform_batch(B, N, E, 3)
    ref_v, ref_b, ref_k, _, _ = oracle_port.batched_forward(x.reshape(-1), dens.reshape(-1),
                                                            *desc, 77)
    s = BatchedLOITS.uniform(B, N, E, 3, dtype=dtype,
                             tuning=Tuning(threads=512, chunk_events=16384, scan_mode=1,
                                           guide_shift=guide_shift, table_format=fmt))
    print("spiky", guide_shift, dtype, fmt, "cell curves", s.layout["cell_instances"], flush=True)
    xt = torch.from_numpy(x.reshape(-1)).to(dtype).cuda()
    dt = torch.from_numpy(dens.reshape(-1)).to(dtype).cuda()
    values, kept, st, ex = s.run_forward(xt, dt, 77, want_bins=True)
    assert kept.cpu().numpy().tolist() == ref_k.tolist()
    assert np.array_equal(ex["bins"].cpu().numpy(), ref_b)
    if dtype == torch.float64:
        assert np.array_equal(values.cpu().numpy(), ref_v)
    else:
        v = values.cpu().numpy().astype(np.float64)
        assert np.max(np.abs(v - ref_v) / np.abs(ref_v)) <= 1e-5
    gv = np.random.default_rng(2).standard_normal(len(ref_v))
    ref_g = oracle_port.batched_backward(x.reshape(-1), dens.reshape(-1), *desc, 77, gv)
    gd = torch.empty_like(dt)
    s.backward_raw(xt, dt, 77, st, torch.from_numpy(gv).to(dtype).cuda(), gd)
    got = gd.double().cpu().numpy().reshape(B, N)
    ref = ref_g.reshape(B, N)
    err = np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)
    print("  grad rel err", err, flush=True)
    assert err.max() < (1e-9 if dtype == torch.float64 else 1e-4)


def test_batched_vjp_matches_float64_central_differences(hip_device):
    """d/d density_i of L = sum_r w_r v_r by central differences (h = 1e-6) on
    16 random nodes per curve, float64, no window (kept rows do not move).

    The value is x_j + dx_j (u - C_j) / (C_{j+1} - C_j + eps): at a bin
    boundary it jumps by dx_j eps / (d_j + eps), so a difference quotient
    across a bin change measures that jump, not the derivative.  A node whose
    perturbation moves any event to another bin is retried with h / 10 (twice)
    and otherwise left out; at most 2 of the 48 nodes may be left out.
    """
    from quantom_loits_amd.batched import BatchedLOITS, Tuning
    from quantom_loits_amd.workloads import uniform_family

    B, N, E, seed = 3, 256, 20000, 5
    x, d = uniform_family(B, N, seed=8, dtype=np.float64)
    s = BatchedLOITS.uniform(B, N, E, 1, dtype=torch.float64,
                             tuning=Tuning(threads=256, chunk_events=4096, scan_mode=1))
    xt = torch.from_numpy(x.reshape(-1)).cuda()
    dt = torch.from_numpy(d.reshape(-1)).cuda()
    w = torch.from_numpy(np.random.default_rng(9).standard_normal(B * E)).cuda()

    def loss(dens):
        v, _, _, ex = s.run_forward(xt, dens, seed, want_bins=True)
        return float((v * w).sum()), ex["bins"]

    values, kept, st, ex0 = s.run_forward(xt, dt, seed, want_bins=True)
    bins0 = ex0["bins"].clone()
    gd = torch.empty_like(dt)
    s.backward_raw(xt, dt, seed, st, w, gd)
    got = gd.cpu().numpy()
    rng = np.random.default_rng(10)
    skipped = 0
    for b in range(B):
        nodes = rng.choice(N, 16, replace=False) + b * N
        scale = np.abs(got[b * N:(b + 1) * N]).max()
        worst = 0.0
        for i in nodes:
            for h in (1e-6, 1e-7, 1e-8):
                dp = dt.clone()
                dp[i] += h
                dm = dt.clone()
                dm[i] -= h
                lp, bp = loss(dp)
                lm, bm = loss(dm)
                if torch.equal(bp, bins0) and torch.equal(bm, bins0):
                    worst = max(worst, abs(got[i] - (lp - lm) / (2 * h)) / scale)
                    break
            else:
                skipped += 1
        print("finite differences curve", b, "max rel err", worst, flush=True)
        assert worst <= 1e-4, (b, worst)
    assert skipped <= 2, skipped


def _big_curves(kind, rng):
    """(n_nodes, events, x, density) of curves too large for an LDS-resident VJP."""
    from quantom_loits_amd.workloads import density_params, eval_density

    if kind == "smooth":
        nodes, events = [4096, 32767, 9000], [150000, 60000, 100000]
    elif kind == "spiky":
        nodes, events = [16384, 6000], [80000, 50000]
    else:  # negative densities: a CDF that is not monotone
        nodes, events = [5000, 4500], [40000, 30000]
    xs, ds = [], []
    params = density_params(len(nodes), 21)
    for i, n in enumerate(nodes):
        xg = np.linspace(0.001, 0.999, n)
        if kind == "smooth":
            d = eval_density(xg, params[i])
        elif kind == "spiky":
            d = np.full(n, 1e-9)
            for mu in rng.uniform(0.1, 0.9, 6):
                d += np.exp(-0.5 * ((xg - mu) / rng.uniform(1e-4, 2e-3)) ** 2)
        else:
            d = 1.0 + 0.3 * (4.0 + 2 * i) * np.sin(np.linspace(0, 12, n))
        xs.append(xg)
        ds.append(d)
    return np.array(nodes), np.array(events), np.concatenate(xs), np.concatenate(ds)


@pytest.mark.parametrize("kind,dtype", [("smooth", torch.float32), ("smooth", torch.float64),
                                        ("spiky", torch.float64), ("spiky", torch.float32),
                                        ("nonmono", torch.float64)])
def test_batched_vjp_bin_ranges_match_oracle(oracle_port, hip_device, kind, dtype):
    """Curves whose VJP tables and histograms exceed LDS: the VJP splits the bins
    into ranges of at most 7168 (float32, exact integer sums) or 5120 (float64)
    bins, one work-group per chunk and range; a curve that is not monotone takes
    the global-table VJP instead."""
    from quantom_loits_amd.batched import BatchedLOITS, Tuning

    rng = np.random.default_rng(31)
    nodes, events, x, d = _big_curves(kind, rng)
    n = len(nodes)
    node_off = np.concatenate([[0], np.cumsum(nodes)[:-1]]).astype(np.int64)
    base = rng.integers(0, 2 ** 40, n).astype(np.uint64)
    streams = rng.integers(0, 2 ** 33, n).astype(np.uint64)
    accept = np.ones(n, np.uint8)
    window = (0.05, 0.95) if kind == "smooth" else None
    s = BatchedLOITS(nodes, events, base, streams, accept, node_off, dtype=dtype, window=window,
                     tuning=Tuning(threads=1024, chunk_events=65536,
                                   scan_mode=1 if dtype == torch.float64 else 0, table_format=1))
    assert s.layout["range_instances"] == n
    print("bin ranges", kind, dtype, nodes.tolist(), s.layout["grid_backward"], flush=True)
    xt = torch.from_numpy(x).to(dtype).cuda()
    dt = torch.from_numpy(d).to(dtype).cuda()
    x64, d64 = xt.double().cpu().numpy(), dt.double().cpu().numpy()
    ref_v, ref_b, ref_k, _, _ = oracle_port.batched_forward(
        x64, d64, node_off, nodes, events, base, streams, 9, accept=accept, window=window)
    values, kept, st, ex = s.run_forward(xt, dt, 9, want_bins=True)
    assert kept.cpu().numpy().tolist() == ref_k.tolist()
    assert np.array_equal(ex["bins"].cpu().numpy(), ref_b)
    gv = np.random.default_rng(32).standard_normal(len(ref_v))
    ref_g = oracle_port.batched_backward(x64, d64, node_off, nodes, events, base, streams, 9, gv,
                                         accept=accept, window=window)
    gd = torch.empty_like(dt)
    s.backward_raw(xt, dt, 9, st, torch.from_numpy(gv).to(dtype).cuda(), gd)
    # twice in a row on the same buffers: counters and partial sums reset themselves
    gd2 = torch.empty_like(dt)
    s.backward_raw(xt, dt, 9, st, torch.from_numpy(gv).to(dtype).cuda(), gd2)
    for call, g in enumerate((gd, gd2)):
        got = g.double().cpu().numpy()
        for i in range(n):
            sl = slice(int(node_off[i]), int(node_off[i] + nodes[i]))
            err = np.abs(got[sl] - ref_g[sl]).max() / np.abs(ref_g[sl]).max()
            print("  call", call, "curve", i, "nodes", nodes[i], "grad rel err", err, flush=True)
            assert err <= (1e-9 if dtype == torch.float64 else 1e-4), (call, i, err)


@pytest.mark.parametrize("nodes,fmt", [(2048, 0), (2048, 1), (9000, 0)])
def test_batched_float32_vjp_heavy_tailed_cotangents(oracle_port, hip_device, nodes, fmt):
    """Rare cotangents 1e9 times larger than the rest overflow the VJP's fixed-point
    scale (set from each segment's first rows) and take its fp64 overflow sums;
    2048 nodes: LDS-resident VJP, 9000 nodes: bin ranges."""
    from quantom_loits_amd.batched import BatchedLOITS, Tuning
    from quantom_loits_amd.workloads import uniform_family

    B, E = 3, 200000
    x, d = uniform_family(B, nodes, seed=4, dtype=np.float32)
    s = BatchedLOITS.uniform(B, nodes, E, 0, dtype=torch.float32,
                             tuning=Tuning(threads=1024, chunk_events=1 << 16, table_format=fmt))
    xt = torch.from_numpy(x.reshape(-1)).cuda()
    dt = torch.from_numpy(d.reshape(-1)).cuda()
    values, kept, st, _ = s.run_forward(xt, dt, 3)
    rng = np.random.default_rng(6)
    gv = rng.standard_normal(values.numel()).astype(np.float32)
    gv[rng.integers(0, gv.size, gv.size // 1000)] *= 1e9
    gd = torch.empty_like(dt)
    s.backward_raw(xt, dt, 3, st, torch.from_numpy(gv).cuda(), gd)
    desc = uniform_batch(B, nodes, E, 0)
    ref = oracle_port.batched_backward(x.reshape(-1).astype(np.float64),
                                       d.reshape(-1).astype(np.float64), *desc, 3,
                                       gv.astype(np.float64)).reshape(B, nodes)
    got = gd.double().cpu().numpy().reshape(B, nodes)
    err = np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)
    print("heavy-tailed cotangents", nodes, err, flush=True)
    assert err.max() <= 1e-4


def _dropin_inputs(rng, B, nx, ny, k, dtype):
    xb = np.sort(rng.uniform(0.001, 0.999, (B, nx, k)), axis=2)
    qb = np.sort(rng.uniform(1.0, 100.0, (B, ny, k)), axis=2)
    xs = rng.uniform(0.1, 1.0, (B, ny, nx, k))
    qs = rng.uniform(0.1, 1.0, (B, nx, ny, k))
    w = rng.uniform(0.5, 1.5, (B, nx, ny))
    w /= w.sum()
    acc = rng.uniform(0, 1, (B, nx, ny)) < 0.8
    return [torch.from_numpy(a).to(dtype).cuda() for a in (xb, xs, qb, qs, w)] + \
        [torch.from_numpy(acc).cuda()]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_bounded_dropin_matches_reference_path_and_replays_in_a_graph(hip_device, dtype):
    """The sync-free reference forward/backward (device-side nmax / active,
    capacity buffers) against the host-sized binding, eagerly and as a HIP graph
    replayed on new weights with eager work between replays; capacities that
    are too small report an error and sample nothing."""
    from quantom_loits_amd import load_extension
    from quantom_loits_amd.backend import BoundedLOITS

    ext = load_extension()
    rng = np.random.default_rng(40)
    xb, xs, qb, qs, w, acc = _dropin_inputs(rng, 2, 7, 5, 10, dtype)
    E, seed, seq = 20000, 99, 3
    b = BoundedLOITS(xb, qb, dtype, nmax_cap=2000, active_cap=2 * E)

    def reference(weights, grad_rows):
        st = ext.forward(xb, xs, qb, qs, weights, acc, E, seed, seq, False)
        g = grad_rows[: st[0].shape[0]].contiguous()
        gx, gq = ext.backward(g, xb, xs, qb, qs, acc, *st[1:], False)
        return st, gx, gq

    def check(weights):
        st, gx, gq = reference(weights, b.grad_events)
        got = b.state()
        for a, r in zip(got, st):
            assert a.shape == r.shape and torch.equal(a, r)
        assert torch.equal(b.grad_x, gx) and torch.equal(b.grad_q, gq)

    print("bounded eager", dtype, flush=True)
    b.grad_events.copy_(torch.randn(b.grad_events.shape, dtype=dtype, device="cuda"))
    b.forward(xb, xs, qb, qs, w, acc, E, seed, seq)
    b.backward(xb, xs, qb, qs, acc)
    check(w)

    print("bounded graph", dtype, flush=True)
    w_in = w.clone()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        b.forward(xb, xs, qb, qs, w_in, acc, E, seed, seq)
        b.backward(xb, xs, qb, qs, acc)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        b.forward(xb, xs, qb, qs, w_in, acc, E, seed, seq)
        b.backward(xb, xs, qb, qs, acc)
    for r in range(3):
        nw = torch.from_numpy(rng.uniform(0.2, 2.0, tuple(w.shape))).to(dtype).cuda()
        w_in.copy_(nw / nw.sum())
        graph.replay()
        ext.forward(xb, xs, qb, qs, w, acc, 1000, 5, 1, False)  # unrelated eager work
        torch.cuda.synchronize()
        check(w_in)

    print("bounded overflow", dtype, flush=True)
    small = BoundedLOITS(xb, qb, dtype, nmax_cap=10, active_cap=2 * E)
    small.forward(xb, xs, qb, qs, w, acc, E, seed, seq)
    nmax, active, kept, err = small.sizes()
    assert nmax > 10 and err & 2 and kept == 0
    with pytest.raises(RuntimeError, match="capacities"):
        small.state()
