"""Training-step timing and ``loits::*`` region rows in the reference's CSV format.

The reference's benchmark driver writes two kinds of CSV files that its
plotting reads (benchmark_training.py:34-98, plotting/common.py:37-91):

* ``training_<implementation>_<device>_e<events>[_t<threads>].csv``: one row
  per timed training iteration, region ``gan::training_iteration``, column
  ``wall_ms`` (pytorch/profiler.py:126-131);
* ``regions_...csv``: one row per ``gan::*`` / ``loits::*`` profiler range,
  with its parent range and CPU / device times (pytorch/profiler.py:134-175).

Both start with the same metadata columns (benchmark_training.py:52-65).  The
column names and order, the file-name stem and the synchronised wall-clock
idiom are kept so that the reference's plotting reads these files unchanged.
"""
from __future__ import annotations

import csv
import re
import time
from pathlib import Path

import torch

METADATA_COLUMNS = ("site", "experiment", "backend", "implementation", "device", "events",
                    "threads", "grid_size", "warmup", "iterations", "seed")
TIMING_COLUMNS = METADATA_COLUMNS + ("region", "occurrence", "wall_ms")
REGION_COLUMNS = METADATA_COLUMNS + ("region_timing", "region", "occurrence", "event_id",
                                     "parent_event_id", "parent_region", "thread_id",
                                     "sequence_nr", "start_us", "end_us", "cpu_ms", "device_ms",
                                     "self_cpu_ms", "self_device_ms")
ITERATION_REGION = "gan::training_iteration"


def safe_component(value) -> str:
    """File-name-safe form of a metadata value (benchmark_training.py:11-12)."""
    return re.sub(r"[^A-Za-z0-9_.-]+", "-", str(value)).strip("-") or "run"


def result_stem(kind, implementation, device, events, threads="") -> str:
    parts = [kind, safe_component(implementation), safe_component(str(device).replace(":", "-")),
             f"e{events}", f"t{threads}" if threads else ""]
    return "_".join(p for p in parts if p)


def row_metadata(site, experiment, backend, implementation, device, events, threads, grid_size,
                 warmup, iterations, seed) -> dict:
    values = (site, experiment, backend, implementation, str(device), events, threads, grid_size,
              warmup, iterations, seed)
    return dict(zip(METADATA_COLUMNS, values))


def synchronize(device):
    device = torch.device(device)
    if device.type == "cuda":
        torch.cuda.synchronize(device)


def measure(trainer, warmup=5, iterations=10):
    """Synchronised wall-clock ms per training step (pytorch/profiler.py:85-99).

    The generator parameters are restored before every step so that each
    sample times the same step.
    """
    params = trainer.snapshot_generator_params()
    for _ in range(warmup):
        trainer.restore_generator_params(params)
        synchronize(trainer.device)
        trainer.step()
    synchronize(trainer.device)
    samples = []
    for _ in range(iterations):
        trainer.restore_generator_params(params)
        synchronize(trainer.device)
        t0 = time.perf_counter()
        trainer.step()
        synchronize(trainer.device)
        samples.append((time.perf_counter() - t0) * 1e3)
    return samples


def profile(trainer, warmup=5, iterations=10, trace_path=None):
    """Run the steps under torch.profiler, each inside ``gan::training_iteration``."""
    params = trainer.snapshot_generator_params()
    for _ in range(warmup):
        trainer.restore_generator_params(params)
        trainer.step()
    synchronize(trainer.device)
    acts = [torch.profiler.ProfilerActivity.CPU]
    if torch.device(trainer.device).type == "cuda":
        acts.append(torch.profiler.ProfilerActivity.CUDA)
    with torch.profiler.profile(activities=acts, record_shapes=False, profile_memory=False,
                                with_stack=False, acc_events=True) as prof:
        for _ in range(iterations):
            trainer.restore_generator_params(params)
            synchronize(trainer.device)
            with torch.profiler.record_function(ITERATION_REGION):
                trainer.step()
                synchronize(trainer.device)
    if trace_path:
        path = Path(trace_path)
        path.parent.mkdir(parents=True, exist_ok=True)
        prof.export_chrome_trace(str(path))
    return prof


def timing_rows(samples, metadata):
    return [dict(metadata, region=ITERATION_REGION, occurrence=i, wall_ms=ms)
            for i, ms in enumerate(samples)]


def _tracked(name: str) -> bool:
    return name.startswith("gan::") or name.startswith("loits::")


def region_rows(events, metadata):
    """Rows for every ``gan::*`` / ``loits::*`` range of a profiler event list.

    ``events`` is ``prof.events()``; the parent is the nearest enclosing range
    that is itself tracked.  Ranges with no CPU time are skipped.
    """
    meta = dict(metadata, region_timing="synchronized_wall")
    rows, seen = [], {}
    for ev in events:
        if not _tracked(ev.name) or float(ev.cpu_time_total or 0.0) <= 0.0:
            continue
        parent = ev.cpu_parent
        while parent is not None and not _tracked(parent.name):
            parent = parent.cpu_parent
        occ = seen.get(ev.name, 0)
        seen[ev.name] = occ + 1
        rows.append(dict(meta, region=ev.name, occurrence=occ, event_id=ev.id,
                         parent_event_id=parent.id if parent is not None else "",
                         parent_region=parent.name if parent is not None else "",
                         thread_id=ev.thread, sequence_nr=ev.sequence_nr,
                         start_us=ev.time_range.start, end_us=ev.time_range.end,
                         cpu_ms=ev.cpu_time_total / 1e3, device_ms=ev.device_time_total / 1e3,
                         self_cpu_ms=ev.self_cpu_time_total / 1e3,
                         self_device_ms=ev.self_device_time_total / 1e3))
    return rows


def write_csv(path, rows):
    if not rows:
        return
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with path.open("w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0].keys()))
        w.writeheader()
        w.writerows(rows)
