"""Backend registry and the public ``LOITS`` module.

Mirrors the reference's backend API (loits.py:5-145): ``register_backend``,
``registered_backends``, ``backend_status``, ``available_backends``,
``BackendUnavailableError`` and ``LOITS(backend=..., **kwargs)``, called as
``sampler(theory_outputs, n_events)``.  This package registers one backend,
``"hip"``; the reference registers its own torch/cpp/openmp/sycl/triton
backends in its copy of this module and adds ``"hip"`` with the two lines
shown in INTEGRATION.md.
"""
from __future__ import annotations

import torch
from torch import nn

from ._native import BackendUnavailableError

_BACKENDS = {}
_BACKEND_PROBES = {}


def register_backend(name, factory, probe=None):
    _BACKENDS[name] = factory
    _BACKEND_PROBES[name] = probe


def registered_backends():
    return tuple(sorted(_BACKENDS))


def backend_status(name, device="cpu"):
    """(available, reason) for backend `name` on `device` (loits.py:34-46)."""
    if name not in _BACKENDS:
        return False, f"backend {name!r} is not registered"
    probe = _BACKEND_PROBES.get(name)
    if probe is None:
        return True, ""
    try:
        result = probe(torch.device(device))
    except Exception as exc:  # a failing probe means unavailable, with its message
        return False, str(exc)
    if isinstance(result, tuple):
        return bool(result[0]), str(result[1])
    return bool(result), "" if result else "backend probe returned unavailable"


def available_backends(device="cpu"):
    return tuple(name for name in registered_backends() if backend_status(name, device)[0])


def _load_hip(**kwargs):
    from .backend import HipLOITS

    return HipLOITS(**kwargs)


def _probe_hip(device):
    from .backend import availability

    return availability(device)


register_backend("hip", _load_hip, _probe_hip)


class LOITS(nn.Module):
    """``LOITS(backend="hip", device="cuda")(theory_outputs, n_events)`` (loits.py:129-145)."""

    def __init__(self, backend="hip", **kwargs):
        super().__init__()
        if backend not in _BACKENDS:
            available = ", ".join(registered_backends())
            raise ValueError(f"backend={backend!r} is not registered; registered: {available}")
        device = kwargs.get("device", "cpu")
        ok, reason = backend_status(backend, device)
        if not ok:
            raise BackendUnavailableError(
                f"backend={backend!r} is unavailable for device={device!s}: {reason}")
        self.backend = backend
        self.impl = _BACKENDS[backend](**kwargs)

    def forward(self, theory_outputs, n_events):
        return self.impl(theory_outputs, n_events)
