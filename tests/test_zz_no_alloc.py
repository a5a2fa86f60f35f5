"""Allocation-free forward (loits_hip_forward_noalloc) through the C ABI.

It must give exactly what loits_hip_forward gives with a caller workspace, on
two calls in a row (the second smaller, so stale scratch contents from the
first call would show), and reject what does not fit its fixed scratch.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from quantom_loits_amd._native import library_path


class Shape(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("batch", "nx", "ny", "kx", "kq", "cells")]


class ForwardArgs(C.Structure):
    _fields_ = [("shape", Shape), ("dtype", C.c_int), ("nmax", C.c_int64),
                ("active", C.c_int64), ("seed", C.c_uint64), ("sequence", C.c_uint64)] + [
        (n, C.c_void_p) for n in (
            "x_bins", "xsec_x", "q_bins", "xsec_q", "acceptance", "counts", "norm_x", "norm_q",
            "cdf_x", "cdf_q", "u_x", "u_q", "interval_x", "interval_q", "events", "packed",
            "row_offsets", "workspace", "valid")]


SCRATCH_BYTES = 16 * 1024 * 1024


@pytest.fixture(scope="module")
def lib():
    path = library_path()
    if not path.is_file():
        pytest.skip("HIP library not built")
    lib = C.CDLL(str(path))
    lib.loits_hip_last_error.restype = C.c_char_p
    lib.loits_hip_forward_workspace_size.restype = C.c_int64
    return lib


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _theory(rng, B, nx, ny, kx, kq, dev):
    """Reference-layout inputs (cpp/loits_core.cpp:62-94) with uneven weights."""
    f64 = dict(dtype=torch.float64, device=dev)
    xb = torch.from_numpy(np.sort(rng.uniform(0.001, 0.999, (B, nx, kx)), axis=-1)).to(**f64)
    qb = torch.from_numpy(np.sort(rng.uniform(1.0, 100.0, (B, ny, kq)), axis=-1)).to(**f64)
    xs = torch.from_numpy(rng.uniform(0.1, 2.0, (B, ny, nx, kx))).to(**f64)
    qs = torch.from_numpy(rng.uniform(0.1, 2.0, (B, nx, ny, kq))).to(**f64)
    w = torch.from_numpy(rng.dirichlet(np.ones(nx * ny), B).reshape(B, nx, ny)).to(**f64)
    acc = torch.from_numpy(rng.uniform(size=(B, nx, ny)) > 0.1).to(torch.uint8).to(dev)
    return xb, xs, qb, qs, w, acc


def _forward(lib, inputs, n_events, seed, sequence, noalloc):
    xb, xs, qb, qs, w, acc = inputs
    dev = xb.device
    B, nx, kx = xb.shape
    ny, kq = qb.shape[1], qb.shape[2]
    cells = nx * ny
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    counts = torch.empty(B * cells, dtype=torch.int64, device=dev)
    na = torch.empty(2, dtype=torch.int64, device=dev)
    assert lib.loits_hip_allocate(_ptr(w), 0, C.c_int64(B * cells), C.c_int64(n_events),
                                  _ptr(counts), _ptr(na), stream) == 0
    nmax, active = (int(v) for v in na.cpu())
    slots = B * cells * nmax
    f64 = dict(dtype=torch.float64, device=dev)
    out = dict(
        norm_x=torch.full((B, ny, nx), -1.0, **f64), norm_q=torch.full((B, nx, ny), -1.0, **f64),
        cdf_x=torch.full(xs.shape, -1.0, **f64), cdf_q=torch.full(qs.shape, -1.0, **f64),
        u_x=torch.full((max(slots, 1),), -1.0, dtype=torch.float32, device=dev),
        u_q=torch.full((max(slots, 1),), -1.0, dtype=torch.float32, device=dev),
        interval_x=torch.full((max(slots, 1),), -1, dtype=torch.int16, device=dev),
        interval_q=torch.full((max(slots, 1),), -1, dtype=torch.int16, device=dev),
        events=torch.full((max(active, 1), 2), -1.0, **f64),
        packed=torch.full((max(active, 1),), -1, dtype=torch.int64, device=dev),
        row_offsets=torch.full((B * cells + 1,), -1, dtype=torch.int64, device=dev),
        valid=torch.full((2,), -1, dtype=torch.int64, device=dev))
    shape = Shape(B, nx, ny, kx, kq, cells)
    ws = None
    if not noalloc:
        nbytes = lib.loits_hip_forward_workspace_size(C.byref(shape), C.c_int64(nmax))
        assert nbytes > 0
        ws = torch.full((nbytes,), 0xAB, dtype=torch.uint8, device=dev)
    a = ForwardArgs(shape, 0, nmax, active, seed, sequence,
                    _ptr(xb), _ptr(xs), _ptr(qb), _ptr(qs), _ptr(acc), _ptr(counts),
                    _ptr(out["norm_x"]), _ptr(out["norm_q"]), _ptr(out["cdf_x"]),
                    _ptr(out["cdf_q"]), _ptr(out["u_x"]), _ptr(out["u_q"]),
                    _ptr(out["interval_x"]), _ptr(out["interval_q"]), _ptr(out["events"]),
                    _ptr(out["packed"]), _ptr(out["row_offsets"]), _ptr(ws), _ptr(out["valid"]))
    fn = lib.loits_hip_forward_noalloc if noalloc else lib.loits_hip_forward
    rc = fn(C.byref(a), stream)
    assert rc == 0, lib.loits_hip_last_error().decode()
    torch.cuda.synchronize()
    return {k: v.cpu() for k, v in out.items()}, (nmax, active)


@pytest.mark.gpu
def test_noalloc_forward_equals_workspace_forward(lib, hip_device):
    rng = np.random.default_rng(7)
    # (B, nx, ny, kx, kq, n_events): the second call is smaller than the first
    for B, nx, ny, kx, kq, ne in ((2, 9, 7, 10, 12, 200_000), (1, 3, 2, 5, 4, 3_000)):
        inputs = _theory(rng, B, nx, ny, kx, kq, hip_device)
        print(f"noalloc case B={B} nx={nx} ny={ny} kx={kx} kq={kq} events={ne}", flush=True)
        ref, sizes = _forward(lib, inputs, ne, 1234, 3, noalloc=False)
        got, sizes2 = _forward(lib, inputs, ne, 1234, 3, noalloc=True)
        assert sizes == sizes2
        assert int(got["valid"][1]) == 0 and int(ref["valid"][1]) == 0
        nvalid = int(ref["valid"][0])
        assert int(got["valid"][0]) == nvalid and nvalid > 0
        for k in ref:
            r, g = ref[k], got[k]
            if k in ("events", "packed"):
                r, g = r[:nvalid], g[:nvalid]
            assert torch.equal(r, g), k


def test_noalloc_rejects_a_workspace_and_oversized_shapes(lib):
    shape = Shape(1, 1, 1, 4, 4, 1)
    a = ForwardArgs(shape, 0, 8, 8, 0, 0)
    a.workspace = C.c_void_p(16)
    assert lib.loits_hip_forward_noalloc(C.byref(a), None) == 1
    assert "workspace must be NULL" in lib.loits_hip_last_error().decode()
    # the look-back status words alone outgrow the scratch: 2^40 slots
    big = ForwardArgs(Shape(1, 1, 1, 4, 4, 1), 0, 1 << 40, 8, 0, 0)
    need = lib.loits_hip_forward_workspace_size(C.byref(big.shape), C.c_int64(1 << 40))
    assert need > SCRATCH_BYTES
    assert lib.loits_hip_forward_noalloc(C.byref(big), None) == 1
    assert "scratch" in lib.loits_hip_last_error().decode()
