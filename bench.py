"""Benchmark: fused batched LOITS forward + VJP on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-plain] [--no-cpu]
                    [--no-extras] [--no-parity]

Workload (one "step"): B = 256 float32 densities of N = 2048 nodes on
x in [0.001, 0.999] (the density family x^a (1-x)^b + Gaussian bump, seeded),
2^18 events per density (67,108,864 events), Philox seed 1234, sequence 0:
the batched forward (curve tables, sampling, compaction) followed by the
reverse VJP w.r.t. the densities for an N(0,1) float32 cotangent.  Inputs are
resident in HBM before the timed region; nothing is skipped inside it.

Printed JSON line: events/s for the whole step (K steps after 0.25 s of
untimed steps that bring the GPU to its working clocks and W warm-up steps, no
events inside the timed loop; the parity check below runs first); the roofline of the dominant kernel (algorithmic bytes / its
average duration from HIP events on the launching stream, in a second loop);
the same-size parity of the float32 path against the float64 path on the
reference's sequential CDF chain (bin flips, largest value and gradient
errors); the per-instance-launch comparator ("plain"); the other workloads
(B=1024 N=4096 forward, stored vs recomputed tables, float64 at the headline
shape, the 4096-curve ragged batch, the GAN training step, the two-axis
drop-in path); and the CPU baseline (the reference's own OpenMP core, compiled
from its sources into oracle/_ref, in a fresh pinned process).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0
B, N, E = 256, 2048, 1 << 18
SEED, SEQUENCE = 1234, 0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=100)
    p.add_argument("--no-plain", action="store_true")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-parity", action="store_true",
                   help="skip the float32 vs float64 comparison at the headline shape")
    p.add_argument("--threads", type=int, default=1024)
    p.add_argument("--chunk", type=int, default=1 << 18)
    p.add_argument("--guide-shift", type=int, default=2)
    p.add_argument("--scan-mode", type=int, default=0)
    p.add_argument("--table-format", type=int, default=0)
    p.add_argument("--persistent", type=int, default=0, choices=(0, 1),
                   help="1: the batched forward and VJP as one persistent launch each")
    p.add_argument("--ragged-mode", type=int, default=None, choices=(0, 1),
                   help="with --ragged-only: time only per-kind (0) or persistent (1) launches "
                        "instead of both and the autotuned plan")
    p.add_argument("--autotune", action="store_true",
                   help="pick threads / chunk / guide shift with the launch-shape autotuner")
    p.add_argument("--graph", action="store_true",
                   help="also time the step (forward, pack, VJP) replayed as one HIP graph")
    p.add_argument("--no-extras", action="store_true",
                   help="skip the other workloads (forward at B=1024 N=4096, recomputed "
                        "tables, float64, the 4096-curve ragged batch, GAN step, drop-in)")
    p.add_argument("--dropin-only", action="store_true",
                   help="only time the reference-API (two-axis) path at the GAN shapes and "
                        "print those timings")
    p.add_argument("--dropin-shape", default=None, metavar="GRID,EVENTS",
                   help="with --dropin-only: time only this grid and event count")
    p.add_argument("--ragged-only", action="store_true",
                   help="only time the 4096-curve ragged batch (fused graph) and print it")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="only copy a MIB-MiB buffer a few times (the plain-copy kernel that "
                        "scales rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    p.add_argument("--cpu-child", action="store_true", help=argparse.SUPPRESS)
    return p.parse_args()


def algorithmic_bytes(kept, nodes, s=4):
    """Minimum traffic the API forces: forward reads x and the density and
    writes the kept events; the VJP reads the cotangents, x and the density and
    writes the density gradient (s = bytes per value)."""
    fwd = nodes * (s + s) + kept * s
    bwd = kept * s + nodes * (s + s + s)
    return fwd, bwd


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from profiles/traffic.json (rocprofv3
    FETCH_SIZE and WRITE_SIZE runs scaled by a plain copy), or None when the
    kernel sources changed since those counters were taken."""
    sys.path.insert(0, str(ROOT))
    from quantom_loits_amd.rocprof_summary import sources_sha256

    tfile = ROOT / "profiles" / "traffic.json"
    try:
        t = json.loads(tfile.read_text())
    except (OSError, ValueError):
        return None, "profiles/traffic.json missing"
    if t.get("sources_sha256") != sources_sha256():
        return None, "profiles/traffic.json is from other kernel sources"
    k = t.get("kernels", {}).get(kernel)
    if k is None:
        return None, f"no counters for {kernel} in profiles/traffic.json"
    return k["traffic_bytes"], ("profiles/traffic.json: rocprofv3 --pmc FETCH_SIZE and "
                                "WRITE_SIZE runs, scaled by a plain copy; sources match")


def time_ms(fn, reps, warmup=2):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps


def parity_fields(dev, x, d, grad, tun):
    """The float32 path at the bench tuning against the float64 path on the
    reference's sequential CDF chain (bit-exact with the oracle in the parity
    tests), on the same 67,108,864 events: bins that differ, the largest value
    error and the largest density-gradient error."""
    import torch
    from quantom_loits_amd.batched import BatchedLOITS, Tuning

    s32 = BatchedLOITS.uniform(B, N, E, SEQUENCE, dtype=torch.float32, device=dev, tuning=tun)
    st32 = s32.new_state()
    bins32 = torch.empty(B * E, dtype=torch.int32, device=dev)
    gd32 = torch.empty_like(d)
    s32.forward_raw(x, d, SEED, st32, bins=bins32)
    s32.backward_raw(x, d, SEED, st32, grad, gd32)
    s64 = BatchedLOITS.uniform(B, N, E, SEQUENCE, dtype=torch.float64, device=dev,
                               tuning=Tuning(threads=1024, chunk_events=1 << 16, scan_mode=1))
    st64 = s64.new_state()
    bins64 = torch.empty(B * E, dtype=torch.int32, device=dev)
    x64, d64 = x.double(), d.double()
    gd64 = torch.empty_like(d64)
    s64.forward_raw(x64, d64, SEED, st64, bins=bins64)
    s64.backward_raw(x64, d64, SEED, st64, grad.double(), gd64)
    torch.cuda.synchronize()
    if int(st32["total"].item()) != B * E or int(st64["total"].item()) != B * E:
        raise RuntimeError("parity: the headline workload keeps every event")
    v32, v64 = st32["events"][:B * E].double(), st64["events"][:B * E]
    verr = (v32 - v64).abs()
    gerr = (gd32.double() - gd64).abs()
    out = dict(
        compared="float32 (bench tuning, parallel CDF scan) vs float64 (sequential CDF "
                 "chain, the reference's order) on the same events",
        events=B * E,
        bin_flips=int((bins32 != bins64).sum().item()),
        max_abs_value_err=float(verr.max().item()),
        max_rel_value_err=float((verr / v64.abs().clamp_min(1e-30)).max().item()),
        max_abs_grad_err=float(gerr.max().item()),
        max_rel_grad_err=float((gerr.max() / gd64.abs().max()).item()),
        kept_equal=bool(torch.equal(st32["kept"], st64["kept"])),
    )
    g64 = grad.double()
    ms_seq = time_ms(lambda: (s64.forward_raw(x64, d64, SEED, st64),
                              s64.backward_raw(x64, d64, SEED, st64, g64, gd64)), 5)
    del s32, st32, s64, st64, bins32, bins64, v32, v64
    sd = BatchedLOITS.uniform(B, N, E, SEQUENCE, dtype=torch.float64, device=dev, tuning=tun)
    std = sd.new_state()
    ms_par = time_ms(lambda: (sd.forward_raw(x64, d64, SEED, std),
                              sd.backward_raw(x64, d64, SEED, std, g64, gd64)), 5)
    del sd, std
    return out, dict(
        bench_tuning=dict(ms_per_step=ms_par, events_per_s=B * E / (ms_par * 1e-3),
                          tuning="the bench's (threads 1024, chunk 2^18, parallel CDF scan)"),
        sequential_chain=dict(ms_per_step=ms_seq, events_per_s=B * E / (ms_seq * 1e-3),
                              tuning="threads 1024, chunk 2^16, the reference's CDF chain"))


def extra_workloads(dev, steps, fp64_row):
    """Timings of the other workloads (not the headline value)."""
    import torch
    from quantom_loits_amd.batched import BatchedLOITS, Tuning
    from quantom_loits_amd.workloads import uniform_family

    out = {}
    # forward only, B=1024 curves of N=4096 nodes, 2^16 events each
    b1, n1, e1 = 1024, 4096, 1 << 16
    x1, d1 = uniform_family(b1, n1, seed=1, dtype=np.float32)
    x1 = torch.from_numpy(x1.reshape(-1)).to(dev)
    d1 = torch.from_numpy(d1.reshape(-1)).to(dev)
    s1 = BatchedLOITS.uniform(b1, n1, e1, SEQUENCE, dtype=torch.float32, device=dev,
                              tuning=Tuning(chunk_events=1 << 16))
    st1 = s1.new_state()
    ms = time_ms(lambda: s1.forward_raw(x1, d1, SEED, st1), steps)
    fb, _ = algorithmic_bytes(b1 * e1, b1 * n1)
    out["forward_B1024_N4096_E2p16"] = dict(
        ms=ms, events_per_s=b1 * e1 / (ms * 1e-3), algorithmic_gbs=fb / (ms * 1e-3) / 1e9,
        cell_table_curves=s1.layout["cell_instances"],
        global_table_curves=s1.layout["global_instances"])
    del s1, st1, x1, d1
    # VJP that rebuilds the curve tables instead of reading the forward's
    xb, db = uniform_family(B, N, seed=0, dtype=np.float32)
    xb = torch.from_numpy(xb.reshape(-1)).to(dev)
    db = torch.from_numpy(db.reshape(-1)).to(dev)
    sb = BatchedLOITS.uniform(B, N, E, SEQUENCE, dtype=torch.float32, device=dev)
    stb = sb.new_state()
    gb = torch.randn(B * E, device=dev, dtype=torch.float32)
    gdb = torch.empty_like(db)
    sb.forward_raw(xb, db, SEED, stb)
    stored = time_ms(lambda: sb.backward_raw(xb, db, SEED, stb, gb, gdb), steps)
    rebuilt = time_ms(lambda: sb.backward_raw(xb, db, SEED, stb, gb, gdb, True), steps)
    out["vjp_B256_N2048_E2p18"] = dict(stored_tables_ms=stored, recomputed_tables_ms=rebuilt)
    del sb, stb, gb, gdb
    if fp64_row is not None:
        out["fp64_B256_N2048_E2p18"] = fp64_row
    out["ragged_4096_window"] = ragged_workload(dev, steps)
    out["gan_B512_N2048_E2p16"] = gan_workload(dev)
    out["dropin_two_axis"] = dropin_workloads(dev, steps)
    return out


def ragged_workload(dev, steps, loop=True, persistent=None):
    """4096 ragged curves (N log-uniform in [128, 16384], E in [1024, 2^20],
    window cut [0.05, 0.95]) as one captured graph (forward, pack, VJP; the
    planner picks each curve's table format and chunks), timed with per-kind
    launches, with persistent launches and with the autotuner's per-bucket
    choice (reported as the fused time), against a per-curve launch loop.
    ``persistent`` (0 or 1) times that mode alone instead."""
    import torch
    from quantom_loits_amd.autotune import Autotuner
    from quantom_loits_amd.batched import BatchedLOITS, Tuning
    from quantom_loits_amd.workloads import ragged_family

    w = ragged_family(4096, n_range=(128, 16384), e_range=(1024, 1 << 20), seed=5)
    nn = len(w["n_nodes"])
    xr = torch.from_numpy(w["x"]).to(dev)
    dr = torch.from_numpy(w["density"]).to(dev)
    base = np.concatenate([[0], np.cumsum(w["events"])[:-1]]).astype(np.uint64)
    window = (0.05, 0.95)
    if persistent is None:
        t0 = time.perf_counter()
        auto = Autotuner().tuning_for(w["n_nodes"], w["events"], torch.float32, window=window)
        tune_s = time.perf_counter() - t0
        runs = (("per_kind", Tuning()), ("persistent", Tuning(persistent=1)), ("autotuned", auto))
    else:
        tune_s = None
        runs = ((("persistent" if persistent else "per_kind"), Tuning(persistent=persistent)),)
    modes = {}
    for label, tun in runs:
        sr = BatchedLOITS(w["n_nodes"], w["events"], base, None, None, w["node_off"],
                          dtype=torch.float32, device=dev, window=window, tuning=tun)
        g = sr.capture(xr, dr, SEED)
        g.grad.normal_()
        modes[label] = time_ms(g.replay, max(3, steps // 4))
        kept = int(g.kept.sum())
        layout = sr.layout
        del g, sr
    label, tun = runs[-1]
    fused = modes[label]
    nodes = int(np.sum(w["n_nodes"]))
    fb, bb = algorithmic_bytes(kept, nodes)
    ev = int(w["events"].sum())
    out = dict(curves=nn, events=ev, kept=kept, fused_graph_ms=fused, mode=label,
               modes_ms=modes, fused_events_per_s=ev / (fused * 1e-3),
               algorithmic_gbs=(fb + bb) / (fused * 1e-3) / 1e9,
               tuning=dict(threads=tun.threads, chunk_events=tun.chunk_events,
                           guide_shift=tun.guide_shift, persistent=tun.persistent,
                           buckets={int(b): list(v) for b, v in tun.buckets.items()}),
               autotune_s=tune_s,
               cell_table_curves=layout["cell_instances"],
               global_table_curves=layout["global_instances"],
               range_vjp_curves=layout["range_instances"])
    if not loop:
        return out
    singles = []
    for i in range(nn):
        si = BatchedLOITS(w["n_nodes"][i:i + 1], w["events"][i:i + 1], base[i:i + 1], None, None,
                          w["node_off"][i:i + 1], dtype=torch.float32, device=dev,
                          window=(0.05, 0.95), tuning=Tuning(chunk_events=1 << 16))
        singles.append((si, si.new_state(),
                        torch.empty(max(si.capacity, 1), device=dev, dtype=torch.float32)))
    gr = torch.randn(max(int(w["events"].max()), 1), device=dev, dtype=torch.float32)
    gdr = torch.empty_like(dr)

    def run_loop():
        for si, sti, packed in singles:
            si.forward_raw(xr, dr, SEED, sti)
            si.pack_raw(xr, dr, sti, packed)
            si.backward_raw(xr, dr, SEED, sti, gr, gdr)

    looped = time_ms(run_loop, 2, warmup=1)
    out.update(per_curve_loop_ms=looped, speedup=looped / fused)
    del singles
    return out


def gan_workload(dev, warmup=2, iterations=5):
    """The GAN-style training step (B=512 generated densities, N=2048, 2^16
    events each, window [0.05, 0.95], soft 64-bin histograms): step time and
    the share of LOITS forward + backward, for the HIP sampler and for the
    torch-ops comparator on the same GPU."""
    import torch
    from quantom_loits_amd import profiling
    from quantom_loits_amd.training import BatchedGanTrainer, loits_share

    out = {}
    for backend in ("hip", "torch"):
        tr = BatchedGanTrainer(backend, dev, 512, 2048, 1 << 16, seed=0)
        wall = profiling.measure(tr, warmup, iterations)
        step_ms, loits_ms = loits_share(tr, 1, iterations)
        out[backend] = dict(step_ms=float(np.mean(wall)), event_step_ms=step_ms,
                            loits_ms=loits_ms, loits_share=loits_ms / step_ms)
        del tr
        torch.cuda.empty_cache()
    out["step_speedup_vs_torch"] = out["torch"]["step_ms"] / out["hip"]["step_ms"]
    out["loits_speedup_vs_torch"] = out["torch"]["loits_ms"] / out["hip"]["loits_ms"]
    return out


def dropin_workloads(dev, steps, only=None):
    """The reference-API path (LOITS(backend="hip") bindings, both axes) at the
    reference's own GAN shapes: grid g x g cells of K = 10 nodes per axis, one
    theory instance; forward + backward, the forward's two host reads included.
    The cotangent (all ones) is a view of a buffer filled once per shape: at
    most n_events rows are kept, and refilling 160 MB per step would time
    the harness, not the path."""
    import torch
    from quantom_loits_amd import load_extension

    ext = load_extension()
    out = {}
    rng = np.random.default_rng(3)
    for grid, n_events in ((10, 10000), (99, 10000), (99, 1000000), (99, 10000000)):
        if only is not None and (grid, n_events) != only:
            continue
        k = 10
        xb = np.sort(rng.uniform(0.001, 0.999, (1, grid, k)), axis=2)
        qb = np.sort(rng.uniform(1.0, 100.0, (1, grid, k)), axis=2)
        xs = rng.uniform(0.1, 1.0, (1, grid, grid, k))
        qs = rng.uniform(0.1, 1.0, (1, grid, grid, k))
        w = rng.uniform(0.5, 1.5, (1, grid, grid))
        w /= w.sum()
        acc = rng.uniform(0, 1, (1, grid, grid)) < 0.95
        args = [torch.from_numpy(a).to(dev) for a in (xb, xs, qb, qs, w)]
        acc_t = torch.from_numpy(acc).to(dev)
        ones = torch.ones(n_events, 2, dtype=torch.float64, device=dev)

        def step():
            st = ext.forward(args[0], args[1], args[2], args[3], args[4], acc_t, n_events, SEED,
                             SEQUENCE, False)
            kept = st[0].shape[0]
            g = ones[:kept] if kept <= n_events else torch.ones_like(st[0])
            ext.backward(g, args[0], args[1], args[2], args[3], acc_t, *st[1:], False)
            return st[0].shape[0]

        kept = step()
        ms = time_ms(step, max(3, steps // 2))
        out[f"grid{grid}_e{n_events}"] = dict(cells=grid * grid, nodes_per_axis=k,
                                              events=kept, ms=ms,
                                              events_per_s=kept / (ms * 1e-3))
    return out


def cpu_description():
    """(CPU model, physical cores among this process's CPUs or None when
    /proc/cpuinfo does not say, logical CPUs allowed)."""
    model = None
    cores = set()
    allowed = os.sched_getaffinity(0)
    known = True
    try:
        cpu = phys = core = None
        for line in Path("/proc/cpuinfo").read_text().splitlines() + [""]:
            key, _, val = line.partition(":")
            key, val = key.strip(), val.strip()
            if key == "processor":
                cpu = int(val)
            elif key == "model name" and model is None:
                model = val
            elif key == "physical id":
                phys = val
            elif key == "core id":
                core = val
            elif not line.strip() and cpu is not None:
                if cpu in allowed:
                    known = known and phys is not None and core is not None
                    cores.add((phys, core))
                cpu = phys = core = None
    except OSError:
        known = False
    return model, (len(cores) if known and cores else None), len(allowed)


def cpu_child():
    """Runs in a fresh CPU-only process (OMP_PLACES=cores OMP_PROC_BIND=close):
    the reference's OpenMP core on one curve per thread, N = 2048, E = 2^18
    events per curve; forward and forward + VJP, median of 5 each."""
    sys.path.insert(0, str(ROOT))
    from oracle.oracle import Oracle, available
    from quantom_loits_amd.workloads import reference_1d_inputs, uniform_family

    # the parent describes the CPUs: in this process numpy's OpenMP runtime has
    # already bound the thread to one place (OMP_PROC_BIND), so this view is narrow
    model, phys, logical = cpu_description()
    if available("openmp", portable=True):
        lib, kind = Oracle("openmp", portable=True), "reference"
    else:
        lib, kind = Oracle("port"), "port"
    threads = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    b = threads if kind == "reference" else 1
    x, d = uniform_family(b, N, seed=0, dtype=np.float32)
    args = reference_1d_inputs(x, d)
    fwd, both = [], []
    for rep in range(6):  # the first is a warm-up
        t0 = time.perf_counter()
        st = lib.forward(*args, E, SEED, SEQUENCE)
        t1 = time.perf_counter()
        g = np.zeros_like(st["events"])
        g[:, 0] = 1.0
        lib.backward(g, args[0], args[1], args[2], args[3], args[5], st)
        t2 = time.perf_counter()
        if rep:
            fwd.append(t1 - t0)
            both.append(t2 - t0)
    n_ev = b * E
    print(json.dumps(dict(
        value=n_ev / statistics.median(both), unit="events/s", cores=threads if kind ==
        "reference" else 1, kind=kind,
        forward_events_per_s=n_ev / statistics.median(fwd),
        fwd_vjp_events_per_s=n_ev / statistics.median(both),
        cpu_model=model, physical_cores=phys, logical_cpus=logical,
        omp=dict(OMP_PLACES=os.environ.get("OMP_PLACES"),
                 OMP_PROC_BIND=os.environ.get("OMP_PROC_BIND"), OMP_NUM_THREADS=threads),
        sample=f"{b} curves x N={N} x 2^18 events ({n_ev} events), forward and forward + VJP "
               f"through the reference {'openmp core' if kind == 'reference' else 'oracle port'}"
               f" (1-D mapping: two-axis call with a dummy q axis), median of 5 after a "
               f"warm-up; value = forward + VJP")))


def cpu_baseline():
    env = dict(os.environ, OMP_PLACES="cores", OMP_PROC_BIND="close", HIP_VISIBLE_DEVICES="",
               ROCR_VISIBLE_DEVICES="", CUDA_VISIBLE_DEVICES="")
    r = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--cpu-child"], env=env,
                       capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError("cpu baseline child failed:\n" + r.stderr[-2000:])
    out = json.loads(r.stdout.strip().splitlines()[-1])
    model, phys, logical = cpu_description()  # this (unpinned) process's CPUs
    out.update(cpu_model=model, physical_cores=phys, logical_cpus=logical)
    return out


def main():
    args = parse()
    if args.cpu_child:
        cpu_child()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus must be 1 "
                                   "(across GPUs the batched path would shard by instances, "
                                   "whole densities per GPU)"}))
        raise SystemExit(2)
    import torch

    if args.copy_calibration:
        n = args.copy_calibration << 18  # float32 elements
        src = torch.ones(n, device="cuda", dtype=torch.float32)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "copies": 5}))
        return

    sys.path.insert(0, str(ROOT))
    from quantom_loits_amd.batched import BatchedLOITS, Tuning
    from quantom_loits_amd.workloads import uniform_family

    dev = torch.device("cuda", 0)
    if args.dropin_only:
        only = tuple(int(v) for v in args.dropin_shape.split(",")) if args.dropin_shape else None
        print(json.dumps({"dropin_two_axis": dropin_workloads(dev, args.steps, only)}))
        return
    if args.ragged_only:
        print(json.dumps({"ragged_4096_window": ragged_workload(
            dev, args.steps, loop=False, persistent=args.ragged_mode)}))
        return
    x_np, d_np = uniform_family(B, N, seed=0, dtype=np.float32)
    x = torch.from_numpy(x_np.reshape(-1)).to(dev)
    d = torch.from_numpy(d_np.reshape(-1)).to(dev)
    tun = Tuning(threads=args.threads, guide_shift=args.guide_shift, chunk_events=args.chunk,
                 scan_mode=args.scan_mode, table_format=args.table_format,
                 persistent=args.persistent)
    if args.autotune:
        from quantom_loits_amd.autotune import Autotuner
        tun = Autotuner().tuning_for(np.full(B, N), np.full(B, E), torch.float32,
                                     args.scan_mode)
    s = BatchedLOITS.uniform(B, N, E, SEQUENCE, dtype=torch.float32, device=dev, tuning=tun)
    st = s.new_state()
    gen = torch.Generator(device=dev).manual_seed(1)
    grad = torch.randn(B * E, generator=gen, device=dev, dtype=torch.float32)
    gd = torch.empty_like(d)
    stream = torch.cuda.current_stream()
    # parity first (it also brings the GPU to its working clocks before timing)
    parity, fp64_row = (None, None) if args.no_parity else parity_fields(dev, x, d, grad, tun)
    torch.cuda.empty_cache()

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        s.forward_raw(x, d, SEED, st)
        if ev is not None:
            ev[1].record(stream)
        s.backward_raw(x, d, SEED, st, grad, gd)
        if ev is not None:
            ev[2].record(stream)

    # untimed: run the step for 0.25 s so that the GPU reaches its working clocks
    # (from idle the same step reads ~15% slower over its first ~100 runs), then
    # the W warm-up steps
    ramp_end = time.perf_counter() + 0.25
    while time.perf_counter() < ramp_end:
        for _ in range(16):
            step()
        torch.cuda.synchronize()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    total = int(st["total"].item())
    dense = s.dense_flag(st)
    # the timed steps: nothing but the two launches per step
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    # per-kernel durations: HIP events on the launching stream around each launch
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    for k in range(args.steps):
        step(events[k])
    torch.cuda.synchronize()
    fwd_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in events]))
    bwd_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in events]))
    fwd_b, bwd_b = algorithmic_bytes(total, B * N)
    step_gbs = (fwd_b + bwd_b) / (ms_per_step * 1e-3) / 1e9

    cell = s.layout["cell_instances"] == B
    names = ("sample_cell_kernel", "vjp_cell_kernel") if cell else ("sample_kernel", "vjp_kernel")
    if bwd_ms >= fwd_ms:
        dom = dict(name=names[1], what="batched VJP", ms=bwd_ms, bytes=bwd_b)
    else:
        dom = dict(name=names[0], what="batched forward", ms=fwd_ms, bytes=fwd_b)
    achieved = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9
    traffic, traffic_source = profiled_traffic(dom["name"])

    graph = None
    if args.graph:
        g = s.capture(x, d, SEED)
        g.grad.copy_(grad)
        for _ in range(args.warmup):
            g.replay()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        for _ in range(args.steps):
            g.replay()
        torch.cuda.synchronize()
        graph_ms = (time.perf_counter() - t2) * 1e3 / args.steps
        graph = dict(ms_per_step=graph_ms, events_per_s=B * E / (graph_ms * 1e-3))
        del g

    plain = None
    if not args.no_plain:
        singles = [BatchedLOITS(np.array([N]), np.array([E]), np.array([i * E], np.uint64),
                                np.array([2 * SEQUENCE], np.uint64), None, np.array([i * N]),
                                dtype=torch.float32, device=dev,
                                tuning=Tuning(threads=1024, chunk_events=1 << 14))
                   for i in range(B)]
        sts = [p.new_state() for p in singles]
        gds = torch.empty_like(d)

        def plain_step():
            for i, (p, sti) in enumerate(zip(singles, sts)):
                p.forward_raw(x, d, SEED, sti)
                p.backward_raw(x, d, SEED, sti, grad[i * E:(i + 1) * E], gds)

        plain_step()
        torch.cuda.synchronize()
        ps = max(2, args.steps // 5)
        t1 = time.perf_counter()
        for _ in range(ps):
            plain_step()
        torch.cuda.synchronize()
        plain_ms = (time.perf_counter() - t1) * 1e3 / ps
        plain = dict(ms_per_step=plain_ms, events_per_s=B * E / (plain_ms * 1e-3),
                     speedup=plain_ms / ms_per_step)
        del singles, sts

    extras = None if args.no_extras else extra_workloads(dev, args.steps, fp64_row)
    cpu = None if args.no_cpu else cpu_baseline()
    out = {
        "metric": "events/s fwd+VJP, B=256 N=2048, fused batched vs plain; % HBM peak",
        "value": B * E / (ms_per_step * 1e-3),
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded density family, N(0,1) cotangent)",
        "config": {"workload": "batched LOITS forward + VJP, B=256 densities, N=2048 nodes, "
                               "2^18 events each (67,108,864 events), float32",
                   "instances": B, "nodes": N, "events_per_instance": E,
                   "kept_events": total, "all_kept": dense,
                   "tuning": {"threads": tun.threads, "chunk_events": tun.chunk_events,
                              "guide_shift": tun.guide_shift, "scan_mode": tun.scan_mode,
                              "table_format": tun.table_format}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_source,
                     "kernel": f"{dom['name']} ({dom['what']})", "kernel_ms": dom["ms"],
                     "algorithmic_bytes": dom["bytes"]},
        "step": {"forward_ms": fwd_ms, "backward_ms": bwd_ms,
                 "algorithmic_bytes": fwd_b + bwd_b, "achieved_gbs": step_gbs,
                 "frac": step_gbs / HBM_PEAK_GBS},
        "parity": parity,
        "plain": plain,
        "graph": graph,
        "other_workloads": extras,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
