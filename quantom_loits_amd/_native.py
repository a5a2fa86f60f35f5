"""Loader for the in-tree native extension (``_C``) and its HIP library.

The product path has no CPU fallback: when the extension or the HIP library
is missing, ``load_extension`` raises ``BackendUnavailableError`` with the
build command to run.
"""
from __future__ import annotations

import importlib
from pathlib import Path

import torch  # noqa: F401  (loads PyTorch's HIP runtime before our library)

PKG = Path(__file__).resolve().parent
LIBRARY = PKG / "lib" / "libquantom_loits_hip.so"

_EXT = None


class BackendUnavailableError(RuntimeError):
    """The requested backend cannot run here (loits.py:9)."""


def library_path() -> Path:
    return LIBRARY


def load_extension():
    """Import quantom_loits_amd._C (built by ``python -m quantom_loits_amd.build``)."""
    global _EXT
    if _EXT is not None:
        return _EXT
    if not LIBRARY.is_file():
        raise BackendUnavailableError(
            f"HIP library {LIBRARY} is not built; run `python -m quantom_loits_amd.build`")
    try:
        _EXT = importlib.import_module("quantom_loits_amd._C")
    except ImportError as exc:
        raise BackendUnavailableError(
            f"native extension quantom_loits_amd._C is not importable ({exc}); run "
            "`python -m quantom_loits_amd.build`") from exc
    return _EXT
