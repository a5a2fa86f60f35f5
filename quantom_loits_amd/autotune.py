"""Launch-shape autotuner for the batched sampler.

The launch shape of a batched plan (threads per work-group, events per
work-group chunk, guide resolution) changes speed, not results: the forward
values, bins, kept counts and compaction order are identical for every shape,
and the gradients agree to float32 rounding (the VJP's fixed-point scale is
chosen per wave segment).  The scan mode changes bits and is left to the
caller.

Batches are grouped into buckets by (dtype, largest curve rounded up to a
power of two, mean events per curve rounded up to a power of two, number of
curves rounded up to a power of two).  For a bucket seen for the first time
the tuner times forward + VJP for every candidate shape the planner accepts,
on a synthetic batch of the bucket's shape, and stores the fastest in a JSON
cache (``QUANTOM_LOITS_TUNE_CACHE``, default
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
MAX_NODES = 32767


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


def gpu_step_ms(n_nodes, events, dtype, tuning: Tuning, warmup: int = 2, reps: int = 5,
                device="cuda") -> float:
    """Mean forward + VJP time (ms) of one synthetic batch, HIP events on the current stream."""
    from .batched import BatchedLOITS
    from .workloads import density_params, eval_density

    n_nodes = np.asarray(n_nodes, dtype=np.int64)
    events = np.asarray(events, dtype=np.int64)
    s = BatchedLOITS(n_nodes, events, np.concatenate([[0], np.cumsum(events)[:-1]]),
                     dtype=dtype, device=device, tuning=tuning)
    params = density_params(len(n_nodes), 0)
    xs, ds = [], []
    for i, n in enumerate(n_nodes):
        grid = np.linspace(0.001, 0.999, int(n))
        xs.append(grid)
        ds.append(eval_density(grid, params[i]))
    x = torch.from_numpy(np.concatenate(xs)).to(dtype).to(device)
    d = torch.from_numpy(np.concatenate(ds)).to(dtype).to(device)
    st = s.new_state()
    grad = torch.randn(max(s.capacity, 1), dtype=dtype, device=device)
    gd = torch.empty_like(d)

    def step():
        s.forward_raw(x, d, 1, st)
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

    def tuning_for(self, n_nodes, events, dtype, scan_mode: int = 0) -> Tuning:
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


def auto_tuning(n_nodes, events, dtype, scan_mode: int = 0) -> Tuning:
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Autotuner()
    return _DEFAULT.tuning_for(n_nodes, events, dtype, scan_mode)
