"""MI355X-native LOITS sampler (HIP kernels for gfx950).

    from quantom_loits_amd import LOITS
    sampler = LOITS(backend="hip", device="cuda")
    events = sampler(theory_outputs, n_events)

and the batched 1-D path ``quantom_loits_amd.batched.BatchedLOITS``.
"""
from ._native import BackendUnavailableError, load_extension
from .loits import LOITS, available_backends, backend_status, register_backend, registered_backends

__all__ = ["LOITS", "BackendUnavailableError", "available_backends", "backend_status",
           "load_extension", "register_backend", "registered_backends"]
__version__ = "0.1.0"
