"""The ``hip`` LOITS backend: drop-in for the reference's native backends.

``HipLOITS`` has the constructor, seed/sequence policy and autograd contract
of the reference's ``CppLOITS`` / ``SYCLLOITS`` (cpp/backend.py:34-101,
sycl/backend.py:231-304); ``load_extension().forward/backward`` have the
reference binding signatures (cpp/bindings.cpp:24-33, 132-147) plus a trailing
optional HIP stream, and return the openmp-style state tuple
``[events, norm_x, norm_q, cdf_x, cdf_q, u_x, u_q, interval_x, interval_q,
packed, row_offsets]`` (openmp/bindings.cpp:131).  All work runs on torch's
current HIP stream.
"""
from __future__ import annotations

from contextlib import nullcontext

import torch
from torch import nn
from torch.profiler import record_function

from ._native import BackendUnavailableError, load_extension

EPSILON = 1e-5
SUPPORTED_ARCH = "gfx950"


def availability(device="cuda"):
    """(ok, reason): True only for a visible gfx950 GPU with the library built."""
    device = torch.device(device)
    if device.type != "cuda":
        return False, "HIP backend needs a ROCm GPU device (torch device type 'cuda')"
    if not torch.cuda.is_available():
        return False, "PyTorch ROCm reports no GPU"
    try:
        ext = load_extension()
    except BackendUnavailableError as exc:
        return False, str(exc)
    index = device.index if device.index is not None else torch.cuda.current_device()
    arch, _ = ext.device_info(index)
    if not arch.startswith(SUPPORTED_ARCH):
        return False, f"HIP backend is built for {SUPPORTED_ARCH}; device {index} is {arch}"
    return True, ""


def prepare_extension(device):
    """Return the extension after checking the device (sycl/backend.py:222-228)."""
    ok, reason = availability(device)
    if not ok:
        raise BackendUnavailableError(reason)
    return load_extension()


def _region(enabled, name):
    return record_function(name) if enabled else nullcontext()


class _HipLOITSFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x_bins, xsec_x, q2_bins, xsec_q2, weights, acceptance, n_events, seed,
                sequence, profile_regions):
        profile_regions = bool(profile_regions)
        ext = load_extension()
        with _region(profile_regions, "loits::binding::forward"):
            state = ext.forward(x_bins.contiguous(), xsec_x.contiguous(), q2_bins.contiguous(),
                                xsec_q2.contiguous(), weights.contiguous(),
                                acceptance.contiguous(), int(n_events), int(seed), int(sequence),
                                profile_regions)
        events, *saved = state
        ctx.profile_regions = profile_regions
        ctx.save_for_backward(x_bins, xsec_x, q2_bins, xsec_q2, acceptance, *saved)
        return events

    @staticmethod
    def backward(ctx, grad_events):
        with _region(ctx.profile_regions, "loits::autograd::backward"):
            x_bins, xsec_x, q2_bins, xsec_q2, acceptance, *saved = ctx.saved_tensors
            ext = load_extension()
            with _region(ctx.profile_regions, "loits::binding::backward"):
                grad_x, grad_q = ext.backward(grad_events.contiguous(), x_bins.contiguous(),
                                              xsec_x.contiguous(), q2_bins.contiguous(),
                                              xsec_q2.contiguous(), acceptance.contiguous(),
                                              *saved, ctx.profile_regions)
            return None, grad_x, None, grad_q, None, None, None, None, None, None


class HipLOITS(nn.Module):
    """Drop-in ``hip`` backend: ``HipLOITS(device)(theory_outputs, n_events)``."""

    def __init__(self, device="cuda", compile=False, profile_regions=False, epsilon=EPSILON):
        super().__init__()
        self.device = torch.device(device)
        if epsilon != EPSILON:
            raise ValueError("The HIP backend currently uses the LOITS epsilon 1e-5")
        ok, reason = availability(self.device)
        if not ok:
            raise BackendUnavailableError(reason)
        self.profile_regions = bool(profile_regions)
        self.seed = torch.initial_seed()
        self.sequence = 0

    def forward(self, theory_outputs, n_events):
        x_bins, xsec_x, q2_bins, xsec_q2, weights, acceptance = theory_outputs[:6]
        sequence = self.sequence
        self.sequence += 1
        with _region(self.profile_regions, "loits::autograd::forward"):
            return _HipLOITSFunction.apply(x_bins, xsec_x, q2_bins, xsec_q2, weights, acceptance,
                                           n_events, self.seed, sequence, self.profile_regions)


class BoundedLOITS:
    """The reference forward/backward with the sizes kept on the device.

    The reference reads nmax and active on the host between allocation and
    sampling (cpp/bindings.cpp:80-92).  Here every buffer is allocated once for
    capacities (``nmax_cap`` slots per cell, ``active_cap`` events) and
    :meth:`forward` / :meth:`backward` only enqueue kernels, so a caller can
    capture them in a HIP graph.  :meth:`state` reads the sizes once and returns
    the reference's 11-tensor state as views of the buffers (each array of the
    reference is a prefix of its capacity buffer).  If a call needs more than
    the capacities it samples nothing and :meth:`sizes` reports error bit 2.
    Cotangents go into ``grad_events`` ([active_cap, 2]) before :meth:`backward`.
    """

    def __init__(self, x_bins, q_bins, dtype, nmax_cap, active_cap):
        ext = prepare_extension(x_bins.device)
        self.ext = ext
        B, nx, kx = x_bins.shape
        ny, kq = q_bins.shape[1], q_bins.shape[2]
        self.shape = (B, nx, ny, kx, kq)
        g = B * nx * ny
        dev = x_bins.device
        i64 = dict(dtype=torch.int64, device=dev)
        f64 = dict(dtype=torch.float64, device=dev)
        self.nmax_cap, self.active_cap = int(nmax_cap), int(active_cap)
        self.counts = torch.zeros(g, **i64)
        self.na = torch.zeros(2, **i64)
        self.valid = torch.zeros(2, **i64)
        self.norm_x = torch.empty(B, ny, nx, **f64)
        self.norm_q = torch.empty(B, nx, ny, **f64)
        self.cdf_x = torch.empty(B, ny, nx, kx, **f64)
        self.cdf_q = torch.empty(B, nx, ny, kq, **f64)
        cap = max(g * self.nmax_cap, 1)
        self.u_x = torch.empty(cap, dtype=torch.float32, device=dev)
        self.u_q = torch.empty(cap, dtype=torch.float32, device=dev)
        self.interval_x = torch.empty(cap, dtype=torch.int16, device=dev)
        self.interval_q = torch.empty(cap, dtype=torch.int16, device=dev)
        acap = max(self.active_cap, 1)
        self.events = torch.empty(acap, 2, dtype=dtype, device=dev)
        self.packed = torch.empty(acap, **i64)
        self.row_offsets = torch.empty(g + 1, **i64)
        self.fwd_ws = torch.empty(max(int(ext.forward_workspace_size(x_bins, q_bins,
                                                                     self.nmax_cap)), 1),
                                  dtype=torch.uint8, device=dev)
        self.bwd_ws = torch.empty(max(int(ext.backward_workspace_size(x_bins, q_bins)), 1),
                                  dtype=torch.uint8, device=dev)
        self.grad_events = torch.zeros(acap, 2, dtype=dtype, device=dev)
        self.grad_x = torch.empty(B, ny, nx, kx, dtype=dtype, device=dev)
        self.grad_q = torch.empty(B, nx, ny, kq, dtype=dtype, device=dev)

    def forward(self, x_bins, xsec_x, q_bins, xsec_q, weights, acceptance, n_events, seed,
                sequence):
        self.ext.forward_bounded(x_bins, xsec_x, q_bins, xsec_q, weights, acceptance,
                                 int(n_events), int(seed), int(sequence), self.counts, self.na,
                                 self.norm_x, self.norm_q, self.cdf_x, self.cdf_q, self.u_x,
                                 self.u_q, self.interval_x, self.interval_q, self.events,
                                 self.packed, self.row_offsets, self.fwd_ws, self.valid, None)

    def backward(self, x_bins, xsec_x, q_bins, xsec_q, acceptance):
        self.ext.backward_bounded(self.grad_events, x_bins, xsec_x, q_bins, xsec_q, acceptance,
                                  self.norm_x, self.norm_q, self.cdf_x, self.cdf_q, self.u_x,
                                  self.u_q, self.interval_x, self.interval_q, self.packed,
                                  self.na, self.valid, self.grad_x, self.grad_q, self.bwd_ws,
                                  None)

    def sizes(self):
        """(nmax, active, kept, error bits) with one host read."""
        nmax, active = (int(v) for v in self.na.cpu())
        kept, err = (int(v) for v in self.valid.cpu())
        return nmax, active, kept, err

    def state(self):
        nmax, _, kept, err = self.sizes()
        if err & 2:
            raise RuntimeError(f"bounded forward failed (error bits {err}); raise the capacities")
        if err:
            raise RuntimeError(f"bounded forward failed (error bits {err}): the compaction "
                               "did not complete")
        B, nx, ny, _, _ = self.shape
        g = B * nx * ny
        return [self.events[:kept], self.norm_x, self.norm_q, self.cdf_x, self.cdf_q,
                self.u_x[:g * nmax].view(B, nx * ny, nmax),
                self.u_q[:g * nmax].view(B, nx * ny, nmax),
                self.interval_x[:g * nmax], self.interval_q[:g * nmax], self.packed[:kept],
                self.row_offsets]
