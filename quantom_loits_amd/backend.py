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
