"""ctypes front end for the CPU checker libraries.  TEST INFRASTRUCTURE ONLY.

* ``Oracle("port")`` wraps ``oracle/libloits_oracle.so``, the repo's own C
  restatement of the reference algorithm (every function cites the reference
  file:line it follows; see ``loits_oracle.h``).
* ``Oracle("cpp")`` / ``Oracle("openmp")`` wrap ``oracle/_ref/libref_*.so``,
  the reference's own cpp/openmp cores compiled from /root/reference by
  ``oracle/Makefile``.  They exist only where that build ran.

``forward``/``backward`` run the same stage sequence as the reference binding
(cpp/bindings.cpp:24-130, 132-208) and return numpy arrays named like the
reference state tuple.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
EPSILON = 1e-5

_P = C.c_void_p
_I64 = C.c_int64
_U64 = C.c_uint64


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def library_path(kind: str, portable: bool = False) -> Path:
    if kind == "port":
        return HERE / "libloits_oracle.so"
    suffix = "_portable" if portable else ""
    return HERE / "_ref" / f"libref_{kind}{suffix}.so"


def available(kind: str, portable: bool = False) -> bool:
    return library_path(kind, portable).is_file()


def build(quiet: bool = True) -> None:
    """Compile the oracle (and, where /root/reference exists, oracle/_ref)."""
    import subprocess

    out = subprocess.run(["make", "-C", str(HERE), "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{out.stdout}\n{out.stderr}")
    if not quiet:
        print(out.stdout)


class Oracle:
    """One CPU implementation of the reference algorithm (fp64)."""

    def __init__(self, kind: str = "port", portable: bool = False):
        path = library_path(kind, portable)
        if not path.is_file():
            raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
        self.kind = kind
        self.lib = C.CDLL(str(path))
        self._prefix = "loits_oracle_" if kind == "port" else "ref_"
        self._setup()

    def _fn(self, name, restype, *argtypes):
        f = getattr(self.lib, self._prefix + name)
        f.restype = restype
        f.argtypes = list(argtypes)
        return f

    def _setup(self):
        port = self.kind == "port"
        self.f_philox = (self._fn("philox4x32_10", None, _U64, _U64, _U64, _P) if port
                         else self._fn("philox", None, _U64, _U64, _U64, _P))
        self.f_fill = self._fn("fill_uniform", None, _P, _I64, _U64, _U64)
        if port:
            self.f_alloc = self._fn("allocate_counts", _I64, _P, _P, _I64, _P, _P)
        else:
            self.f_alloc = self._fn("allocate_counts", None, _P, _P, _I64, _P, _P)
        for axis in ("x", "q"):
            setattr(self, f"f_density_{axis}", self._fn(f"density_{axis}", None, _P, _P, _P, _P, _P))
            setattr(self, f"f_cdf_{axis}", self._fn(f"cdf_{axis}", None, _P, _P, _P, _P, _P))
            setattr(self, f"f_interp_{axis}",
                    self._fn(f"interpolate_{axis}", None, _P, _P, _P, _P, _I64, _P, _P, _P))
            if port:
                setattr(self, f"f_ivjp_{axis}",
                        self._fn(f"interpolation_vjp_{axis}", None, _P, _P, _I64, _P, _P, _P, _P,
                                 _I64, _P, _P))
            else:
                setattr(self, f"f_ivjp_{axis}",
                        self._fn(f"interpolation_vjp_{axis}", None, _P, _P, _I64, _P, _P, _P, _P,
                                 _P, _I64, _P, _P))
            setattr(self, f"f_cvjp_{axis}", self._fn(f"cdf_vjp_{axis}", None, _P, _P, _P, _P, _P))
            setattr(self, f"f_dvjp_{axis}",
                    self._fn(f"density_vjp_{axis}", None, _P, _P, _P, _P, _P, _P))
        self.f_compact = self._fn("compact", _I64, _P, _P, _P, _I64, _P, _P, _P, _P)
        if port:
            self.f_bfwd = self._fn("batched_forward", _I64, _P, _P, _P, _P, _P, _P, _P, _P, _I64,
                                   _U64, _P, _P, _P, _P, _P, _P)
            self.f_bbwd = self._fn("batched_backward", None, _P, _P, _P, _P, _P, _P, _P, _P, _I64,
                                   _U64, _P, _P, _P)
            self.f_words = self._fn("philox_words", None, _P, _I64, _I64, _U64, _U64)
            self.f_threads = self._fn("threads", C.c_int)

    # -- RNG ---------------------------------------------------------------
    def philox(self, block: int, stream: int, seed: int) -> np.ndarray:
        out = np.zeros(4, dtype=np.uint32)
        self.f_philox(block & (2**64 - 1), stream & (2**64 - 1), seed & (2**64 - 1), _ptr(out))
        return out

    def philox_words(self, first_block: int, blocks: int, stream: int, seed: int) -> np.ndarray:
        out = np.zeros(4 * blocks, dtype=np.uint32)
        if self.kind == "port":
            self.f_words(_ptr(out), first_block, blocks, stream, seed)
        else:
            for i in range(blocks):
                out[4 * i:4 * i + 4] = self.philox(first_block + i, stream, seed)
        return out

    def fill_uniform(self, count: int, seed: int, stream: int) -> np.ndarray:
        out = np.empty(count, dtype=np.float32)
        self.f_fill(_ptr(out), count, seed, stream)
        return out

    # -- full reference pipeline -------------------------------------------
    @staticmethod
    def shape_of(x_bins, q_bins):
        b, nx, kx = x_bins.shape
        _, ny, kq = q_bins.shape
        return np.array([b, nx, ny, kx, kq, nx * ny], dtype=np.int64)

    def forward(self, x_bins, xsec_x, q_bins, xsec_q, weights, acceptance, n_events, seed,
                sequence):
        """Reference forward (cpp/bindings.cpp:24-130); returns a dict."""
        x_bins, xsec_x, q_bins, xsec_q, weights = (
            np.ascontiguousarray(a, dtype=np.float64)
            for a in (x_bins, xsec_x, q_bins, xsec_q, weights))
        acc = np.ascontiguousarray(acceptance, dtype=np.uint8)
        s = self.shape_of(x_bins, q_bins)
        B, nx, ny, kx, kq, cells = (int(v) for v in s)
        counts = np.empty((B, cells), dtype=np.int64)
        if self.kind == "port":
            active = np.zeros(1, dtype=np.int64)
            nmax = int(self.f_alloc(_ptr(weights), _ptr(counts), n_events, _ptr(s), _ptr(active)))
            active = int(active[0])
        else:
            out = np.zeros(2, dtype=np.int64)
            self.f_alloc(_ptr(weights), _ptr(counts), n_events, _ptr(s), _ptr(out))
            nmax, active = int(out[0]), int(out[1])
        norm_x = np.empty((B, ny, nx))
        rho_x = np.empty_like(xsec_x)
        self.f_density_x(_ptr(x_bins), _ptr(xsec_x), _ptr(norm_x), _ptr(rho_x), _ptr(s))
        norm_q = np.empty((B, nx, ny))
        rho_q = np.empty_like(xsec_q)
        self.f_density_q(_ptr(q_bins), _ptr(xsec_q), _ptr(norm_q), _ptr(rho_q), _ptr(s))
        cdf_x = np.empty_like(xsec_x)
        self.f_cdf_x(_ptr(x_bins), _ptr(rho_x), _ptr(acc), _ptr(cdf_x), _ptr(s))
        cdf_q = np.empty_like(xsec_q)
        self.f_cdf_q(_ptr(q_bins), _ptr(rho_q), _ptr(acc), _ptr(cdf_q), _ptr(s))
        total = B * cells * nmax
        u_x = np.empty((B, cells, nmax), dtype=np.float32)
        self.f_fill(_ptr(u_x), total, seed, 2 * sequence)
        u_q = np.empty((B, cells, nmax), dtype=np.float32)
        self.f_fill(_ptr(u_q), total, seed, 2 * sequence + 1)
        dense_x = np.zeros(total)
        interval_x = np.zeros(total, dtype=np.int16)
        self.f_interp_x(_ptr(x_bins), _ptr(cdf_x), _ptr(u_x), _ptr(counts), nmax, _ptr(dense_x),
                        _ptr(interval_x), _ptr(s))
        dense_q = np.zeros(total)
        interval_q = np.zeros(total, dtype=np.int16)
        self.f_interp_q(_ptr(q_bins), _ptr(cdf_q), _ptr(u_q), _ptr(counts), nmax, _ptr(dense_q),
                        _ptr(interval_q), _ptr(s))
        events = np.empty((max(active, 1), 2))
        packed = np.empty(max(active, 1), dtype=np.int64)
        row_offsets = np.zeros(B * cells + 1, dtype=np.int64)
        valid = int(self.f_compact(_ptr(dense_x), _ptr(dense_q), _ptr(counts), nmax, _ptr(events),
                                   _ptr(packed), _ptr(row_offsets), _ptr(s)))
        if self.kind == "cpp":  # the serial core does not fill row_offsets
            row_offsets = None
        return dict(events=events[:valid].copy(), norm_x=norm_x, norm_q=norm_q, cdf_x=cdf_x,
                    cdf_q=cdf_q, u_x=u_x, u_q=u_q, interval_x=interval_x, interval_q=interval_q,
                    packed=packed[:valid].copy(), row_offsets=row_offsets, counts=counts,
                    nmax=nmax, active=active, rho_x=rho_x, rho_q=rho_q, dense_x=dense_x,
                    dense_q=dense_q)

    def backward(self, grad_events, x_bins, xsec_x, q_bins, xsec_q, acceptance, state):
        """Reference backward (cpp/bindings.cpp:132-208); returns (gx, gq)."""
        x_bins, xsec_x, q_bins, xsec_q = (
            np.ascontiguousarray(a, dtype=np.float64) for a in (x_bins, xsec_x, q_bins, xsec_q))
        acc = np.ascontiguousarray(acceptance, dtype=np.uint8)
        g = np.ascontiguousarray(grad_events, dtype=np.float64)
        s = self.shape_of(x_bins, q_bins)
        packed = np.ascontiguousarray(state["packed"], dtype=np.int64)
        nvalid = packed.shape[0]
        nmax = int(state["u_x"].shape[2])
        out = []
        for axis, bins, xsec, cdf, u, idx, norm in (
                ("x", x_bins, xsec_x, state["cdf_x"], state["u_x"], state["interval_x"],
                 state["norm_x"]),
                ("q", q_bins, xsec_q, state["cdf_q"], state["u_q"], state["interval_q"],
                 state["norm_q"])):
            cdf = np.ascontiguousarray(cdf, dtype=np.float64)
            gcdf = np.empty_like(cdf)
            if self.kind == "port":
                getattr(self, f"f_ivjp_{axis}")(_ptr(g), _ptr(packed), nvalid, _ptr(bins),
                                                _ptr(cdf), _ptr(u), _ptr(idx), nmax, _ptr(gcdf),
                                                _ptr(s))
            else:
                ro = state.get("row_offsets")
                if self.kind == "openmp" and ro is None:
                    raise ValueError("openmp backward needs row_offsets")
                getattr(self, f"f_ivjp_{axis}")(_ptr(g), _ptr(packed), nvalid, _ptr(ro),
                                                _ptr(bins), _ptr(cdf), _ptr(u), _ptr(idx), nmax,
                                                _ptr(gcdf), _ptr(s))
            grho = np.empty_like(cdf)
            getattr(self, f"f_cvjp_{axis}")(_ptr(bins), _ptr(acc), _ptr(gcdf), _ptr(grho), _ptr(s))
            gxs = np.empty_like(cdf)
            getattr(self, f"f_dvjp_{axis}")(_ptr(bins), _ptr(xsec), _ptr(np.ascontiguousarray(norm)),
                                            _ptr(grho), _ptr(gxs), _ptr(s))
            out.append(gxs)
        return tuple(out)

    # -- 1-D batched restatement (port only) -------------------------------
    def batched_forward(self, x, density, node_off, n_nodes, events, counter_base, streams, seed,
                        accept=None, window=None):
        """Returns (values, bins, kept, norm, cdf) for the 1-D batched path."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        density = np.ascontiguousarray(density, dtype=np.float64)
        node_off = np.ascontiguousarray(node_off, dtype=np.int64)
        n_nodes = np.ascontiguousarray(n_nodes, dtype=np.int64)
        events = np.ascontiguousarray(events, dtype=np.int64)
        counter_base = np.ascontiguousarray(counter_base, dtype=np.uint64)
        streams = np.ascontiguousarray(streams, dtype=np.uint64)
        n = node_off.shape[0]
        acc = None if accept is None else np.ascontiguousarray(accept, dtype=np.uint8)
        win = None if window is None else np.asarray(window, dtype=np.float64)
        cap = int(events.sum())
        values = np.empty(max(cap, 1))
        bins = np.empty(max(cap, 1), dtype=np.int32)
        kept = np.zeros(n, dtype=np.int64)
        norm = np.empty(n)
        cdf = np.empty(x.shape[0])
        total = int(self.f_bfwd(_ptr(x), _ptr(density), _ptr(node_off), _ptr(n_nodes),
                                _ptr(events), _ptr(counter_base), _ptr(streams), _ptr(acc), n,
                                seed, _ptr(win), _ptr(values), _ptr(bins), _ptr(kept),
                                _ptr(norm), _ptr(cdf)))
        return values[:total].copy(), bins[:total].copy(), kept, norm, cdf

    def batched_backward(self, x, density, node_off, n_nodes, events, counter_base, streams, seed,
                         grad_values, accept=None, window=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        density = np.ascontiguousarray(density, dtype=np.float64)
        node_off = np.ascontiguousarray(node_off, dtype=np.int64)
        n_nodes = np.ascontiguousarray(n_nodes, dtype=np.int64)
        events = np.ascontiguousarray(events, dtype=np.int64)
        counter_base = np.ascontiguousarray(counter_base, dtype=np.uint64)
        streams = np.ascontiguousarray(streams, dtype=np.uint64)
        g = np.ascontiguousarray(grad_values, dtype=np.float64)
        acc = None if accept is None else np.ascontiguousarray(accept, dtype=np.uint8)
        win = None if window is None else np.asarray(window, dtype=np.float64)
        out = np.zeros(x.shape[0])
        self.f_bbwd(_ptr(x), _ptr(density), _ptr(node_off), _ptr(n_nodes), _ptr(events),
                    _ptr(counter_base), _ptr(streams), _ptr(acc), node_off.shape[0], seed,
                    _ptr(win), _ptr(g), _ptr(out))
        return out

    def threads(self) -> int:
        return int(self.f_threads()) if self.kind == "port" else int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def uniform_batch(B, N, E, sequence=0):
    """Instance descriptors of the 1-D mapping of a reference call with B curves
    of N nodes and E events each (counter p = b*E + slot, stream 2*sequence)."""
    node_off = np.arange(B, dtype=np.int64) * N
    n_nodes = np.full(B, N, dtype=np.int64)
    events = np.full(B, E, dtype=np.int64)
    counter_base = (np.arange(B, dtype=np.uint64) * np.uint64(E)).astype(np.uint64)
    streams = np.full(B, 2 * sequence, dtype=np.uint64)
    return node_off, n_nodes, events, counter_base, streams
