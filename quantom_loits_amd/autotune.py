"""Launch-shape autotuner for the batched sampler.

The launch shape of a batched plan (threads per work-group, events per
work-group chunk, guide resolution) changes speed, not results: the forward
values, bins, kept counts and compaction order are identical for every shape,
and the gradients agree to float32 rounding (the VJP's fixed-point scale is
chosen per wave segment).  The scan mode changes bits and is left to the
caller.

Batches whose curves all fall in one grid-size bucket (b = ceil(log2 n_nodes))
are keyed by (dtype, largest curve rounded up to a power of two, mean events
per curve rounded up to a power of two, number of curves rounded up to a power
of two).  For a key seen for the first time the tuner times forward + VJP for
every candidate shape the planner accepts, on a synthetic batch of the key's
shape.

Ragged batches (curves in several grid-size buckets) are tuned per bucket
inside the one plan, on the whole batch: first per-kind against persistent
launches, then, bucket by bucket (most keys first), every chunk size (and
guide shift, where the bucket's curves take guide tables) with the other
buckets' current choices, keeping a change that saves at least 1%
(``Tuning.buckets``).  A bucket timed as a batch of its own would favour
small chunks that only pay off on an otherwise idle GPU.  Each choice is
stored in a JSON cache (``QUANTOM_LOITS_TUNE_CACHE``, default
``~/.cache/quantom_loits_amd/tuning.json``).
"""
from __future__ import annotations

import json
import os
from dataclasses import asdict
from pathlib import Path

import numpy as np
import torch

from .batched import DTYPE_CODE, Tuning
from ._native import load_extension

THREADS = (256, 512, 1024)
CHUNKS = (1 << 16, 1 << 17, 1 << 18)
GUIDE_SHIFTS = (1, 2, 3)
BUCKET_CHUNKS = (1 << 15, 1 << 16, 1 << 17, 1 << 18)
FAST_MAX_NODES = 4096  # float32 curves up to this size take the cell tables
MAX_NODES = 32767


def grid_bucket(n: int) -> int:
    """The planner's grid-size bucket of a curve: ceil(log2 n_nodes), at most 15."""
    return min(15, max(1, (int(n) - 1).bit_length()))


def _pow2(v: int) -> int:
    v = max(int(v), 1)
    return 1 << (v - 1).bit_length()


def bucket_key(n_nodes, events, dtype) -> str:
    n_nodes = np.asarray(n_nodes, dtype=np.int64).reshape(-1)
    events = np.asarray(events, dtype=np.int64).reshape(-1)
    name = "f32" if dtype == torch.float32 else "f64"
    mean_e = int(np.ceil(events.mean())) if events.size else 1
    return (f"{name}/N{_pow2(n_nodes.max() if n_nodes.size else 2)}"
            f"/E{_pow2(mean_e)}/B{_pow2(n_nodes.size)}")


def bucket_shape(key: str):
    """Representative (B, N, E) of a bucket: its upper corner (N at most 32767)."""
    _, n, e, b = key.split("/")
    return int(b[1:]), min(int(n[1:]), MAX_NODES), int(e[1:])


def candidates(scan_mode: int = 0):
    return [Tuning(threads=t, guide_shift=g, chunk_events=c, scan_mode=scan_mode)
            for t in THREADS for c in CHUNKS for g in GUIDE_SHIFTS]


def plannable(n_nodes, events, dtype, tuning: Tuning) -> bool:
    """Whether the planner accepts this shape for these curves (host only)."""
    ext = load_extension()
    n = len(n_nodes)
    table = np.zeros((n, 8), dtype=np.int64)
    table[:, 0] = np.concatenate([[0], np.cumsum(n_nodes)[:-1]])
    table[:, 1] = n_nodes
    table[:, 2] = events
    table[:, 3] = np.concatenate([[0], np.cumsum(events)[:-1]])
    table[:, 4] = np.arange(n, dtype=np.int64) * np.asarray(events, np.int64)
    table[:, 6] = 1
    try:
        bchunk, bshift = tuning.bucket_arrays()
        ext.BatchedPlan(torch.from_numpy(table), DTYPE_CODE[dtype], tuning.threads,
                        tuning.guide_shift, tuning.chunk_events, tuning.scan_mode,
                        tuning.persistent, tuning.table_format, bchunk, bshift)
    except (ValueError, RuntimeError):
        return False
    return True


_CURVES = {}  # the synthetic curves of the last batch timed


def gpu_step_ms(n_nodes, events, dtype, tuning: Tuning, warmup: int = 2, reps: int = 5,
                device="cuda", window=None) -> float:
    """Mean forward (+ pack with a window) + VJP time (ms) of one synthetic
    batch, HIP events on the current stream."""
    from .batched import BatchedLOITS
    from .workloads import density_params, eval_density

    n_nodes = np.asarray(n_nodes, dtype=np.int64)
    events = np.asarray(events, dtype=np.int64)
    s = BatchedLOITS(n_nodes, events, np.concatenate([[0], np.cumsum(events)[:-1]]),
                     dtype=dtype, device=device, tuning=tuning, window=window)
    key = (n_nodes.tobytes(), dtype, str(device))
    if _CURVES.get("key") != key:  # the coordinate descent times one batch many times
        params = density_params(len(n_nodes), 0)
        xs, ds = [], []
        for i, n in enumerate(n_nodes):
            grid = np.linspace(0.001, 0.999, int(n))
            xs.append(grid)
            ds.append(eval_density(grid, params[i]))
        _CURVES.clear()
        _CURVES.update(key=key, x=torch.from_numpy(np.concatenate(xs)).to(dtype).to(device),
                       d=torch.from_numpy(np.concatenate(ds)).to(dtype).to(device))
    x, d = _CURVES["x"], _CURVES["d"]
    st = s.new_state()
    grad = torch.randn(max(s.capacity, 1), dtype=dtype, device=device)
    gd = torch.empty_like(d)

    packed = torch.empty(max(s.capacity, 1), dtype=dtype, device=device) if window else None

    def step():
        s.forward_raw(x, d, 1, st)
        if packed is not None:
            s.pack_raw(x, d, st, packed)
        s.backward_raw(x, d, 1, st, grad, gd)

    for _ in range(warmup):
        step()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        step()
    t1.record()
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / reps


def cache_path() -> Path:
    env = os.environ.get("QUANTOM_LOITS_TUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "quantom_loits_amd" / "tuning.json"


class Autotuner:
    """Pick and cache the fastest launch shape per bucket.

    ``timer(n_nodes, events, dtype, tuning) -> ms`` measures one candidate; the
    default runs it on the GPU.  Shapes the planner rejects are skipped.
    """

    def __init__(self, path=None, timer=None, candidates_fn=candidates):
        self.path = Path(path) if path is not None else cache_path()
        self.timer = timer or gpu_step_ms
        self.candidates_fn = candidates_fn
        self.cache = {}
        if self.path.is_file():
            with open(self.path) as f:
                self.cache = json.load(f)

    def _save(self):
        self.path.parent.mkdir(parents=True, exist_ok=True)
        tmp = self.path.with_suffix(".tmp")
        with open(tmp, "w") as f:
            json.dump(self.cache, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    def tune_bucket(self, key: str, scan_mode: int = 0):
        dtype = torch.float32 if key.startswith("f32") else torch.float64
        B, N, E = bucket_shape(key)
        n_nodes = np.full(B, N, np.int64)
        events = np.full(B, E, np.int64)
        timings = []
        for cand in self.candidates_fn(scan_mode):
            if not plannable(n_nodes, events, dtype, cand):
                continue
            timings.append((float(self.timer(n_nodes, events, dtype, cand)), cand))
        if not timings:
            raise ValueError(f"no launch shape can plan bucket {key}")
        ms, best = min(timings, key=lambda t: t[0])
        entry = {k: v for k, v in asdict(best).items() if k in ("threads", "guide_shift",
                                                                 "chunk_events")}
        entry["ms"] = ms
        entry["candidates"] = len(timings)
        self.cache[key] = entry
        self._save()
        return best

    def _time(self, n_nodes, events, dtype, tuning, window):
        if window is None:
            return float(self.timer(n_nodes, events, dtype, tuning))
        return float(self.timer(n_nodes, events, dtype, tuning, window=window))

    def bucket_candidates(self, n_nodes, dtype):
        """(chunk_events, guide_shift) pairs worth timing for one grid-size bucket:
        every chunk size, and every guide shift where its curves take guide tables."""
        guide = dtype != torch.float32 or int(np.max(n_nodes)) > FAST_MAX_NODES
        shifts = GUIDE_SHIFTS if guide else (2,)
        return [(ce, gs) for ce in BUCKET_CHUNKS for gs in shifts]

    def tune_ragged(self, n_nodes, events, dtype, scan_mode=0, window=None) -> Tuning:
        """Per-kind or persistent launches, then each grid-size bucket's chunk size
        and guide shift by coordinate descent on the whole batch: a bucket's
        choices are timed with the other buckets' current ones, and a change is
        kept when it saves at least 1%."""
        n_nodes = np.asarray(n_nodes, np.int64).reshape(-1)
        events = np.asarray(events, np.int64).reshape(-1)
        b_of = np.array([grid_bucket(n) for n in n_nodes])
        key = f"ragged/{bucket_key(n_nodes, events, dtype)}/{'w' if window else 'd'}"
        entry = self.cache.get(key)
        if entry is None:
            def run(p, buckets):
                t = Tuning(threads=1024, scan_mode=scan_mode, persistent=p, buckets=buckets)
                if not plannable(n_nodes, events, dtype, t):
                    return float("inf")
                return self._time(n_nodes, events, dtype, t, window)

            modes = [(run(p, {}), p) for p in (0, 1)]
            best_ms, p = min(modes)
            buckets, tried = {}, 2
            # the buckets with the most keys first
            order = sorted(set(b_of.tolist()), key=lambda b: -int(events[b_of == b].sum()))
            for b in order:
                sel = b_of == b
                for ce, gs in self.bucket_candidates(n_nodes[sel], dtype):
                    if (ce, gs) == (Tuning().chunk_events, Tuning().guide_shift):
                        continue  # the plan-wide default, already timed
                    trial = dict(buckets)
                    trial[int(b)] = (ce, gs)
                    ms = run(p, trial)
                    tried += 1
                    if ms < 0.99 * best_ms:
                        best_ms, buckets = ms, trial
            entry = dict(persistent=p, ms=best_ms, per_kind_ms=modes[0][0],
                         persistent_ms=modes[1][0], candidates=tried,
                         buckets={str(b): list(v) for b, v in buckets.items()})
            self.cache[key] = entry
            self._save()
        buckets = {int(b): tuple(v) for b, v in entry["buckets"].items()}
        return Tuning(threads=1024, scan_mode=scan_mode, persistent=int(entry["persistent"]),
                      buckets=buckets)

    def tuning_for(self, n_nodes, events, dtype, scan_mode: int = 0, window=None) -> Tuning:
        n_flat = np.asarray(n_nodes, np.int64).reshape(-1)
        if len({grid_bucket(n) for n in n_flat}) > 1:
            best = self.tune_ragged(n_flat, events, dtype, scan_mode, window)
            if plannable(n_flat, np.asarray(events).reshape(-1), dtype, best):
                return best
            return Tuning(scan_mode=scan_mode)
        key = bucket_key(n_nodes, events, dtype)
        entry = self.cache.get(key)
        if entry is None:
            best = self.tune_bucket(key, scan_mode)
        else:
            best = Tuning(threads=int(entry["threads"]), guide_shift=int(entry["guide_shift"]),
                          chunk_events=int(entry["chunk_events"]), scan_mode=scan_mode)
        if not plannable(np.asarray(n_nodes).reshape(-1), np.asarray(events).reshape(-1), dtype,
                         best):  # the bucket's corner planned; this batch may not
            best = Tuning(scan_mode=scan_mode)
        return best


_DEFAULT = None


def auto_tuning(n_nodes, events, dtype, scan_mode: int = 0, window=None) -> Tuning:
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Autotuner()
    return _DEFAULT.tuning_for(n_nodes, events, dtype, scan_mode, window)
