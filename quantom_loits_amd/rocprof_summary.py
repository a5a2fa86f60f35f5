"""Text summaries of rocprofv3 runs: kernel traces, HBM traffic and SQ counters.

    python -m quantom_loits_amd.rocprof_summary trace DIR --title T --command C
    python -m quantom_loits_amd.rocprof_summary traffic --fetch D --write D \
        --cfetch D --cwrite D --copy-bytes BYTES --trace D --title T --command C \
        [--bytes KERNEL=ALGORITHMIC_BYTES ...] [--json traffic.json]
    python -m quantom_loits_amd.rocprof_summary sq DIR [DIR ...] --events N \
        --title T --command C [--label KERNEL=LABEL ...]

Each prints a markdown table; `traffic --json` also writes the HBM bytes per
launch of every kernel with the SHA-256 of the kernel sources they were taken
from (bench.py reports them only while the sources still match).

FETCH_SIZE and WRITE_SIZE are KiB per dispatch; they are scaled by what the
same counters read for a plain device copy of known size (__amd_rocclr_copyBuffer,
bench.py --copy-calibration).  SQ_INSTS_* count wave instructions, so
instructions per 64 events is the per-lane count when a lane takes one event;
SQ_LDS_* and SQ_WAIT_* are summed over the device, reported per CU.
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import hashlib
import json
import re
import sys
from pathlib import Path

PEAK_GBS = 8000.0
CUS = 256
ROOT = Path(__file__).resolve().parent.parent


# The sources behind the batched kernels' traffic (kernels, device helpers and
# the planner that sizes their chunks and tables).
BATCHED_SOURCES = ("batched.hip", "batched.hpp", "fast_tables.hpp", "loits_device.hpp",
                   "capi.hip", "capi_internal.hpp")


def sources_sha256(root=ROOT):
    """SHA-256 over the batched path's sources (bench.py compares it with the
    one stored beside the counters in profiles/traffic.json)."""
    h = hashlib.sha256()
    for name in BATCHED_SOURCES:
        p = Path(root) / "quantom_loits_amd" / "csrc" / name
        h.update(name.encode())
        h.update(p.read_bytes())
    return h.hexdigest()


def short_name(kernel: str) -> str:
    """'void qloits::vjp_cell_kernel<2>(...)' -> 'vjp_cell_kernel'."""
    k = kernel.strip().replace("(anonymous namespace)::", "")
    k = re.sub(r"^void\s+", "", k)
    k = k.split("(")[0]
    k = re.sub(r"<.*>$", "", k)
    return k.split("::")[-1]


def display_name(kernel: str) -> str:
    """'void qloits::sample_kernel<double, true>(...)' -> 'sample_kernel<double, true>'."""
    k = re.sub(r"^void\s+", "", kernel.strip().replace("(anonymous namespace)::", ""))
    depth, cut = 0, len(k)
    for i, ch in enumerate(k):  # the parameter list opens at the first '(' outside <>
        depth += ch == "<"
        depth -= ch == ">"
        if ch == "(" and depth == 0:
            cut = i
            break
    k = k[:cut]
    return k[k.rfind("::", 0, k.find("<") if "<" in k else len(k)) + 2:] if "::" in k else k


def _csvs(d, pattern):
    files = glob.glob(str(Path(d) / "**" / pattern), recursive=True)
    if not files:
        raise FileNotFoundError(f"no {pattern} under {d}")
    return files


def counters(d):
    """{kernel: {counter: [value per dispatch]}} from a --pmc run."""
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    for f in _csvs(d, "*counter_collection.csv"):
        acc = collections.defaultdict(float)
        names = {}
        for r in csv.DictReader(open(f)):
            acc[(r["Dispatch_Id"], r["Counter_Name"])] += float(r["Counter_Value"])
            names[r["Dispatch_Id"]] = r["Kernel_Name"]
        for (disp, cname), v in acc.items():
            per[names[disp]][cname].append(v)
    return per


def trace(d):
    """{kernel: [duration us per dispatch]} from a --kernel-trace run."""
    by = collections.defaultdict(list)
    for f in _csvs(d, "*kernel_trace.csv"):
        for r in csv.DictReader(open(f)):
            by[r["Kernel_Name"]].append(
                (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    return by


def _mean(v):
    return sum(v) / len(v) if v else float("nan")


def _median(v):
    if not v:
        return float("nan")
    w = sorted(v)
    m = len(w) // 2
    return w[m] if len(w) % 2 else 0.5 * (w[m - 1] + w[m])


def cmd_trace(a):
    by = trace(a.dir)
    print(f"# Kernel trace: {a.title}\n")
    print("Command (from the repository root on the GPU host, `cd /tmp && export "
          "TMPDIR=/tmp` first):\n\n```\n" + a.command + "\n```\n")
    print("| kernel | calls | average (us) | median (us) | min (us) | max (us) |")
    print("| --- | --- | --- | --- | --- | --- |")
    for k, v in sorted(by.items(), key=lambda kv: -sum(kv[1])):
        print(f"| `{k[:90]}` | {len(v)} | {_mean(v):.1f} | {_median(v):.1f} | {min(v):.1f} | "
              f"{max(v):.1f} |")


def _kv(pairs, conv=str):
    out = {}
    for p in pairs or []:
        k, _, v = p.partition("=")
        out[k] = conv(v)
    return out


def cmd_traffic(a):
    fetch, write = counters(a.fetch), counters(a.write)
    cf, cw = counters(a.cfetch), counters(a.cwrite)
    copy = [k for k in cf if "copyBuffer" in k]
    if not copy:
        raise SystemExit("no __amd_rocclr_copyBuffer dispatch in the calibration runs")
    big = lambda vs: max(vs)  # the calibration copy is the largest copy dispatch
    c_fetch = big(cf[copy[0]]["FETCH_SIZE"]) * 1024.0
    c_write = big(cw[copy[0]]["WRITE_SIZE"]) * 1024.0
    f_fetch, f_write = a.copy_bytes / c_fetch, a.copy_bytes / c_write
    durations = trace(a.trace) if a.trace else {}
    alg = _kv(a.bytes, float)
    print(f"# HBM traffic: {a.title}\n")
    print("Commands (each counter in its own run, `cd /tmp && export TMPDIR=/tmp` first):\n")
    print("```\n" + a.command + "\n```\n")
    print(f"Calibration: a plain device copy of {a.copy_bytes / 2**20:.0f} MiB "
          f"(`__amd_rocclr_copyBuffer`) reads FETCH_SIZE = {c_fetch / 1024:.1f} KiB and "
          f"WRITE_SIZE = {c_write / 1024:.1f} KiB per dispatch, so bytes = FETCH_SIZE x "
          f"{f_fetch:.4f} and WRITE_SIZE x {f_write:.4f}.\n")
    print("Achieved = scaled traffic / the kernel's median duration in the kernel trace; "
          "peak 8 TB/s.\n")
    print("| kernel | FETCH_SIZE (KiB) | WRITE_SIZE (KiB) | traffic (MB) | algorithmic (MB) | "
          "median (us) | traffic GB/s | algorithmic GB/s | algorithmic of peak |")
    print("| --- | --- | --- | --- | --- | --- | --- | --- | --- |")
    kernels = {}
    for k in sorted(set(fetch) | set(write)):
        if "copyBuffer" in k or "at::native" in k:
            continue
        fk = _mean(fetch.get(k, {}).get("FETCH_SIZE", []))
        wk = _mean(write.get(k, {}).get("WRITE_SIZE", []))
        t = fk * 1024 * f_fetch + wk * 1024 * f_write
        name = short_name(k)
        dur = _median(durations.get(k, []))
        ab = alg.get(name)
        kernels[name] = dict(traffic_bytes=t, fetch_kib=fk, write_kib=wk,
                             median_us=dur if dur == dur else None)
        row = [f"`{k[:80]}`", f"{fk:.1f}", f"{wk:.1f}", f"{t / 1e6:.1f}",
               f"{ab / 1e6:.1f}" if ab else "", f"{dur:.1f}" if dur == dur else "",
               f"{t / (dur * 1e-6) / 1e9:.0f}" if dur == dur else "",
               f"{ab / (dur * 1e-6) / 1e9:.0f}" if ab and dur == dur else "",
               f"{ab / (dur * 1e-6) / 1e9 / PEAK_GBS:.1%}" if ab and dur == dur else ""]
        print("| " + " | ".join(row) + " |")
    if a.json:
        Path(a.json).write_text(json.dumps(dict(
            sources_sha256=sources_sha256(), workload=a.title, command=a.command,
            calibration=dict(copy_bytes=a.copy_bytes, fetch_factor=f_fetch,
                             write_factor=f_write),
            kernels=kernels), indent=1) + "\n")


SQ_ORDER = ("SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_VMEM_RD",
            "SQ_INSTS_VMEM_WR", "SQ_INSTS_VALU_INT32", "SQ_INSTS_VALU_INT64")
SQ_PER_CU = ("SQ_LDS_IDX_ACTIVE", "SQ_LDS_BANK_CONFLICT", "SQ_ACTIVE_INST_VALU",
             "SQ_ACTIVE_INST_LDS", "SQ_WAIT_INST_LDS", "SQ_WAIT_INST_ANY")


def cmd_sq(a):
    merged = collections.defaultdict(dict)
    for d in a.dirs:
        for k, cs in counters(d).items():
            for c, v in cs.items():
                merged[k][c] = _mean(v)
    labels = _kv(a.label)
    print(f"# SQ counters: {a.title}\n")
    print("Commands (at most 8 SQ counters a run, `cd /tmp && export TMPDIR=/tmp` first):\n")
    print("```\n" + a.command + "\n```\n")
    print(f"Per dispatch, averaged over dispatches; {a.events} events per dispatch.  "
          "`/64ev` = wave instructions per 64 events (the per-lane count when a lane "
          "takes one event); `/CU` = device total / 256 CUs.\n")
    keep = {k: v for k, v in merged.items()
            if not labels or short_name(k) in labels}
    cols = [c for c in SQ_ORDER if any(c in v for v in keep.values())]
    cols_cu = [c for c in SQ_PER_CU if any(c in v for v in keep.values())]
    head = ["kernel"] + [c.replace("SQ_INSTS_", "") + " /64ev" for c in cols] + \
           [c.replace("SQ_", "").lower() + " /CU" for c in cols_cu]
    print("| " + " | ".join(head) + " |")
    print("| " + " | ".join(["---"] * len(head)) + " |")
    for k, v in sorted(keep.items()):
        name = labels.get(short_name(k), display_name(k))
        row = [name] + [f"{v[c] / (a.events / 64):.2f}" if c in v else "" for c in cols] + \
              [f"{v[c] / CUS:.0f}" if c in v else "" for c in cols_cu]
        print("| " + " | ".join(row) + " |")


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = p.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("dir")
    t.add_argument("--title", required=True)
    t.add_argument("--command", required=True)
    f = sub.add_parser("traffic")
    for n in ("fetch", "write", "cfetch", "cwrite"):
        f.add_argument("--" + n, required=True)
    f.add_argument("--copy-bytes", type=float, required=True)
    f.add_argument("--trace")
    f.add_argument("--bytes", action="append", help="KERNEL=algorithmic bytes per launch")
    f.add_argument("--json")
    f.add_argument("--title", required=True)
    f.add_argument("--command", required=True)
    q = sub.add_parser("sq")
    q.add_argument("dirs", nargs="+")
    q.add_argument("--events", type=float, required=True)
    q.add_argument("--label", action="append", help="KERNEL=label (only these kernels)")
    q.add_argument("--title", required=True)
    q.add_argument("--command", required=True)
    a = p.parse_args(argv)
    {"trace": cmd_trace, "traffic": cmd_traffic, "sq": cmd_sq}[a.cmd](a)


if __name__ == "__main__":
    sys.exit(main())
