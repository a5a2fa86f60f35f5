"""Batched 1-D LOITS: many x-curves of uneven sizes in single launches.

``BatchedLOITS`` samples ``events[i]`` values from each instance's density
(the reference's x-axis sampler under the 1-D mapping nx = ny = 1, see
cpp/loits_core.cpp:55-338) and returns them packed, instance after instance,
in slot order (the compaction order of cpp/loits_core.cpp:239-266).  Its
autograd backward is the hand-written HIP VJP w.r.t. the densities, which
recomputes the Philox stream instead of storing uniforms or bin indices.

Counter mapping: instance i, slot s uses Philox counter ``counter_base[i] + s``
on stream ``streams[i]``.  ``BatchedLOITS.uniform(B, N, E, sequence)`` uses
``counter_base = i * E`` and stream ``2 * sequence``: the x column of a
reference call on the same B curves with weights 1 and ``n_events = E``.
"""
from __future__ import annotations

from contextlib import nullcontext
from dataclasses import dataclass, field

import numpy as np
import torch

from ._native import load_extension

DTYPE_CODE = {torch.float64: 0, torch.float32: 1}


@dataclass
class Tuning:
    """Launch shape of a batched plan.

    ``persistent=1`` runs the forward as one launch and the VJP as one launch
    of resident work-groups taking the plan's items from a device queue
    (largest first); 0 launches each kind of work separately.  ``buckets``
    maps a grid-size bucket b = ceil(log2(n_nodes)) to ``(chunk_events,
    guide_shift)`` for the curves of that bucket (either may be None to keep
    the plan-wide value)."""

    threads: int = 1024
    guide_shift: int = 2
    chunk_events: int = 1 << 18
    scan_mode: int = 0        # 0 parallel CDF scan, 1 the reference's sequential chain
    persistent: int = 0
    table_format: int = 0     # 0 cell tables where they apply, 1 guide tables only
    buckets: dict = field(default_factory=dict)

    def bucket_arrays(self):
        """(chunk_log2[16], guide_shift + 1 [16]) for the native plan; 0 = plan-wide."""
        chunk = [0] * 16
        shift = [0] * 16
        for b, (ce, gs) in self.buckets.items():
            b = int(b)
            if not 1 <= b <= 15:
                raise ValueError(f"bucket {b} outside 1..15")
            if ce is not None:
                ce = int(ce)
                if ce <= 0 or ce & (ce - 1):
                    raise ValueError("bucket chunk_events must be a power of two")
                chunk[b] = ce.bit_length() - 1
            if gs is not None:
                shift[b] = int(gs) + 1
        return chunk, shift


class BatchedLOITS:
    """Plan and launch the batched forward / pack / VJP for fixed instance sizes."""

    def __init__(self, n_nodes, events, counter_base=None, streams=None, accept=None,
                 node_off=None, *, dtype=torch.float32, device="cuda", tuning=None, window=None,
                 profile_regions=False):
        n_nodes = np.asarray(n_nodes, dtype=np.int64).reshape(-1)
        events = np.asarray(events, dtype=np.int64).reshape(-1)
        n = n_nodes.shape[0]
        if events.shape[0] != n:
            raise ValueError("n_nodes and events must have one entry per instance")
        if node_off is None:
            node_off = np.concatenate([[0], np.cumsum(n_nodes)[:-1]]).astype(np.int64)
        counter_base = (np.zeros(n, np.uint64) if counter_base is None
                        else np.asarray(counter_base, dtype=np.uint64).reshape(-1))
        streams = (np.zeros(n, np.uint64) if streams is None
                   else np.asarray(streams, dtype=np.uint64).reshape(-1))
        accept = np.ones(n, np.int64) if accept is None else np.asarray(accept).astype(np.int64)
        out_off = np.concatenate([[0], np.cumsum(events)[:-1]]).astype(np.int64)
        table = np.zeros((n, 8), dtype=np.int64)
        table[:, 0] = node_off
        table[:, 1] = n_nodes
        table[:, 2] = events
        table[:, 3] = out_off
        table[:, 4] = counter_base.view(np.int64)
        table[:, 5] = streams.view(np.int64)
        table[:, 6] = accept
        if dtype not in DTYPE_CODE:
            raise TypeError("dtype must be torch.float32 or torch.float64")
        self.dtype = dtype
        self.device = torch.device(device)
        if isinstance(tuning, str):
            if tuning != "auto":
                raise ValueError("tuning must be a Tuning, None or 'auto'")
            from .autotune import auto_tuning
            tuning = auto_tuning(n_nodes, events, dtype, window=window)
        self.tuning = tuning or Tuning()
        self.window = window
        ext = load_extension()
        t = self.tuning
        bchunk, bshift = t.bucket_arrays()
        self.plan = ext.BatchedPlan(torch.from_numpy(table), DTYPE_CODE[dtype], t.threads,
                                    t.guide_shift, t.chunk_events, t.scan_mode, t.persistent,
                                    t.table_format, bchunk, bshift)
        self.layout = self.plan.layout()
        self.desc = self.plan.descriptors().to(self.device)
        self.n_instances = n
        self.capacity = int(self.layout["capacity"])
        self.total_nodes = int(self.layout["total_nodes"])
        self.events = events
        self.n_nodes = n_nodes
        self.node_off = node_off
        self.profile_regions = bool(profile_regions)
        self.event_log = None  # list of (kind, start, end) CUDA events when set

    @classmethod
    def uniform(cls, B, N, E, sequence=0, **kw):
        """B curves of N nodes, E events each, as one reference call would draw them."""
        counter_base = np.arange(B, dtype=np.uint64) * np.uint64(E)
        streams = np.full(B, 2 * sequence, dtype=np.uint64)
        return cls(np.full(B, N), np.full(B, E), counter_base, streams, **kw)

    # -- buffers ----------------------------------------------------------
    def new_state(self):
        """Device buffers for one forward/backward pair."""
        d = self.device
        L = self.layout
        return dict(
            tables=torch.empty(int(L["table_bytes"]), dtype=torch.uint8, device=d),
            scratch=torch.zeros(int(L["scratch_bytes"]), dtype=torch.uint8, device=d),
            events=torch.empty(max(self.capacity, 1), dtype=self.dtype, device=d),
            kept=torch.empty(self.n_instances, dtype=torch.int64, device=d),
            total=torch.empty(1, dtype=torch.int64, device=d),
        )

    def _window(self):
        if self.window is None:
            return False, 0.0, 0.0
        return True, float(self.window[0]), float(self.window[1])

    # -- raw launches (no host synchronisation) ---------------------------
    def forward_raw(self, x, density, seed, st, norm=None, cdf=None, bins=None, stream=None):
        use, lo, hi = self._window()
        self.plan.forward(self.desc, x, density, int(seed) & (2 ** 64 - 1), use, lo, hi,
                          st["tables"], st["scratch"], st["events"], st["kept"], st["total"],
                          norm, cdf, bins, stream)

    def pack_raw(self, x, density, st, packed, bins=None, packed_bins=None, stream=None):
        self.plan.pack(self.desc, x, density, st["tables"], st["scratch"], st["events"],
                       st["kept"], st["total"], packed, bins, packed_bins, stream)

    def backward_raw(self, x, density, seed, st, grad, grad_density, recompute_tables=False,
                     stream=None):
        use, lo, hi = self._window()
        self.plan.backward(self.desc, x, density, int(seed) & (2 ** 64 - 1), use, lo, hi,
                           st["tables"], st["scratch"], st["events"], st["kept"], st["total"],
                           grad, grad_density, bool(recompute_tables), stream)

    def dense_flag(self, st) -> bool:
        return bool(st["scratch"][:4].view(torch.int32).item())

    # -- convenience -------------------------------------------------------
    def run_forward(self, x, density, seed, want_bins=False, want_tables=False, state=None):
        """Forward with one host sync; returns (values, kept, state, extras).
        ``state`` reuses the buffers of an earlier call (new ones otherwise)."""
        self._check_inputs(x, density)
        with self.region("loits::forward::allocation"):
            st = self.new_state() if state is None else state
        extras = {}
        bins = torch.empty(max(self.capacity, 1), dtype=torch.int32, device=self.device) \
            if want_bins else None
        norm = cdf = None
        if want_tables:
            norm = torch.empty(self.n_instances, dtype=torch.float64, device=self.device)
            cdf = torch.empty(max(self.total_nodes, 1), dtype=torch.float64, device=self.device)
            extras.update(norm=norm, cdf=cdf)
        # one fused launch group: density, CDF, Philox, search and interpolation
        with self.region("loits::forward::interpolation_x"):
            self.forward_raw(x, density, seed, st, norm, cdf, bins)
        with self.region("loits::forward::stream_compaction"):
            return self._compact(x, density, st, bins, extras)

    def _compact(self, x, density, st, bins, extras):
        total = int(st["total"].item())
        if self.dense_flag(st):
            values = st["events"][:total]
            if bins is not None:
                extras["bins"] = bins[:total]
        else:
            packed = torch.empty(max(total, 1), dtype=self.dtype, device=self.device)
            pb = torch.empty(max(total, 1), dtype=torch.int32, device=self.device) \
                if bins is not None else None
            self.pack_raw(x, density, st, packed, bins, pb)
            values = packed[:total]
            if pb is not None:
                extras["bins"] = pb[:total]
        return values, st["kept"], st, extras

    def _check_inputs(self, x, density):
        for name, t in (("x", x), ("density", density)):
            if not t.is_cuda or t.dtype != self.dtype or not t.is_contiguous():
                raise ValueError(f"{name} must be a contiguous {self.dtype} GPU tensor")
            if t.numel() < self.total_nodes:
                raise ValueError(f"{name} has {t.numel()} nodes; the plan needs {self.total_nodes}")

    def __call__(self, x, density, seed):
        """Differentiable w.r.t. density: returns the packed kept values."""
        return self.sample(x, density, seed)[0]

    def sample(self, x, density, seed):
        """Differentiable w.r.t. density: (packed kept values, kept per instance)."""
        return _BatchedFunction.apply(density, x, self, int(seed))

    def set_streams(self, streams):
        """Move every instance to new Philox streams (one host encode and copy).

        The reference post-increments its sequence per call and draws stream
        2*sequence for x (cpp/backend.py:34-101); a training loop does the
        same here with ``set_streams(np.full(B, 2 * sequence))``.
        """
        streams = torch.as_tensor(np.asarray(streams, dtype=np.uint64).reshape(-1).view(np.int64))
        self.desc.copy_(self.plan.descriptors_for_streams(streams))

    # -- observability ------------------------------------------------------
    def region(self, name):
        """A ``loits::*`` profiler range, synchronised at both ends when profiling.

        Kernels run asynchronously, so a range that did not wait for the device
        would only time the launches; like the reference's region hooks
        (pytorch/profiler.py:9-61) the device is synchronised on entry and exit.
        """
        if not self.profile_regions:
            return nullcontext()
        return _SyncedRegion(name, self.device)

    def _record(self, kind):
        """CUDA-event pair around one launch group when ``event_log`` is a list."""
        if self.event_log is None:
            return nullcontext()
        return _EventPair(self.event_log, kind)

    def capture(self, x, density, seed, recompute_tables=False):
        """Capture forward, pack and VJP on fixed buffers into one HIP graph."""
        return BatchedGraph(self, x, density, seed, recompute_tables)


class BatchedGraph:
    """One training step of the batched sampler as a replayable HIP graph.

    The graph holds, in launch order: the forward (tables, sampling, row
    offsets), the pack (a no-op inside the kernel when nothing was dropped) and
    the VJP.  Every buffer is fixed at capture; between replays the caller
    writes new nodes, densities and cotangents in place (``x``, ``density``,
    ``grad``: one cotangent per row of the packed output, capacity rows) and
    may move the Philox streams with ``set_streams``.  The seed is a launch
    argument and stays fixed.  No kernel in the graph needs memory cleared
    beforehand (the build resets the completion counters and the VJP clears its
    own partials), so the graph holds no memset nodes.
    """

    def __init__(self, sampler, x, density, seed, recompute_tables=False):
        sampler._check_inputs(x, density)
        self.sampler = sampler
        self.seed = int(seed)
        d = sampler.device
        cap = max(sampler.capacity, 1)
        self.x = x
        self.density = density
        self.grad = torch.zeros(cap, dtype=sampler.dtype, device=d)
        self.grad_density = torch.zeros_like(density)
        self.packed = torch.zeros(cap, dtype=sampler.dtype, device=d)
        self.desc = sampler.desc.clone()
        self.st = sampler.new_state()
        self.recompute_tables = bool(recompute_tables)
        side = torch.cuda.Stream(device=d)
        side.wait_stream(torch.cuda.current_stream(d))
        with torch.cuda.stream(side):
            self._launch()  # eager run: validates the arguments before capture
        torch.cuda.current_stream(d).wait_stream(side)
        torch.cuda.synchronize(d)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._launch()

    def _launch(self):
        s = self.sampler
        use, lo, hi = s._window()
        seed = self.seed & (2 ** 64 - 1)
        st = self.st
        s.plan.forward(self.desc, self.x, self.density, seed, use, lo, hi, st["tables"],
                       st["scratch"], st["events"], st["kept"], st["total"], None, None, None,
                       None)
        s.plan.pack(self.desc, self.x, self.density, st["tables"], st["scratch"], st["events"],
                    st["kept"], st["total"], self.packed, None, None, None)
        s.plan.backward(self.desc, self.x, self.density, seed, use, lo, hi, st["tables"],
                        st["scratch"], st["events"], st["kept"], st["total"], self.grad,
                        self.grad_density, self.recompute_tables, None)

    def set_streams(self, streams):
        """Point every instance at new Philox streams (host encode + one copy)."""
        streams = torch.as_tensor(np.asarray(streams, dtype=np.uint64).view(np.int64))
        blob = self.sampler.plan.descriptors_for_streams(streams)
        self.desc.copy_(blob)

    def replay(self):
        self.graph.replay()

    def values(self):
        """Packed kept values of the last replay (reads two scalars on the host)."""
        st = self.st
        total = int(st["total"].item())
        src = st["events"] if self.sampler.dense_flag(st) else self.packed
        return src[:total]

    @property
    def kept(self):
        return self.st["kept"]


class _SyncedRegion:
    def __init__(self, name, device):
        self.name = name
        self.device = device
        self.ctx = None

    def __enter__(self):
        torch.cuda.synchronize(self.device)
        self.ctx = torch.profiler.record_function(self.name)
        self.ctx.__enter__()
        return self

    def __exit__(self, *exc):
        torch.cuda.synchronize(self.device)
        self.ctx.__exit__(*exc)
        return False


class _EventPair:
    def __init__(self, log, kind):
        self.log = log
        self.kind = kind

    def __enter__(self):
        self.start = torch.cuda.Event(enable_timing=True)
        self.end = torch.cuda.Event(enable_timing=True)
        self.start.record()
        return self

    def __exit__(self, *exc):
        self.end.record()
        self.log.append((self.kind, self.start, self.end))
        return False


class _BatchedFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, density, x, sampler, seed):
        x = x.contiguous()
        density = density.contiguous()
        with sampler.region("loits::forward"), sampler._record("forward"):
            values, kept, st, _ = sampler.run_forward(x, density, seed)
        ctx.sampler = sampler
        ctx.seed = seed
        ctx.st = st
        ctx.save_for_backward(x, density)
        ctx.mark_non_differentiable(kept)
        return values, kept

    @staticmethod
    def backward(ctx, grad, _grad_kept):
        x, density = ctx.saved_tensors
        sampler = ctx.sampler
        with sampler.region("loits::backward"), sampler._record("backward"):
            with sampler.region("loits::backward::interpolation_x"):
                if grad is None:
                    grad = torch.zeros(1, dtype=sampler.dtype, device=sampler.device)
                grad = grad.contiguous().to(sampler.dtype)
                if grad.numel() == 0:
                    grad = torch.zeros(1, dtype=sampler.dtype, device=sampler.device)
                gd = torch.empty_like(density)
                sampler.backward_raw(x, density, ctx.seed, ctx.st, grad, gd)
        return gd, None, None, None
