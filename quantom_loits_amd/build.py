"""In-tree build of the gfx950 LOITS library and its pybind front end.

    python -m quantom_loits_amd.build          # or __graft_entry__.build()

1. hipcc compiles every csrc/*.hip for gfx950 and links
   quantom_loits_amd/lib/libquantom_loits_hip.so (the C ABI of include/loits_hip.h).
2. g++ compiles csrc/bindings.cpp against the PyTorch headers into
   quantom_loits_amd/_C*.so, linked to the library above.

All HIP code is built with -ffp-contract=off (every fused multiply-add is
explicit in the source) and without packed-FP32 instructions.  Both outputs
resolve libamdhip64 from PyTorch's lib directory, so one HIP runtime serves
the process.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
LIBDIR = PKG / "lib"
BUILD = ROOT / "_build"
LIBNAME = "libquantom_loits_hip.so"
ARCH = os.environ.get("QUANTOM_LOITS_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-command-line-argument",
]


def _run(cmd, quiet=True):
    out = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature flag; that warning is harmless
    noise = "-packed-fp32-ops' is not a recognized feature"
    err = "\n".join(l for l in out.stderr.splitlines() if noise not in l)
    if out.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(map(str, cmd))}\n{out.stdout}\n{err}")
    if err.strip() and not quiet:
        print(err, file=sys.stderr)
    return out


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (Path(cand).exists() or cand == "hipcc"):
            return cand
    raise RuntimeError("hipcc not found")


def torch_paths():
    import torch
    import torch.utils.cpp_extension as ce

    tdir = Path(torch.__file__).resolve().parent
    return tdir, [Path(p) for p in ce.include_paths()], tdir / "lib"


def build_library(quiet=True) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    objs = [BUILD / (s.stem + ".o") for s in sources]
    cmds = [[hipcc(), *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(s), "-o", str(o)]
            for s, o in zip(sources, objs)]
    with ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        list(ex.map(lambda c: _run(c, quiet), cmds))
    _, _, tlib = torch_paths()
    out = LIBDIR / LIBNAME
    tmp = out.with_suffix(".so.tmp")
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          f"-Wl,-rpath,{tlib}", f"-Wl,-soname,{LIBNAME}"], quiet)
    os.replace(tmp, out)
    return out


def extension_suffix() -> str:
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def build_bindings(quiet=True) -> Path:
    import torch
    import torch.utils.cpp_extension as ce

    _, incs, tlib = torch_paths()
    py_inc = sysconfig.get_paths()["include"]
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    flags = ["-O2", "-std=c++17", "-fPIC", "-shared", "-w",
             "-DTORCH_EXTENSION_NAME=_C", "-DTORCH_API_INCLUDE_EXTENSION_H",
             f"-D_GLIBCXX_USE_CXX11_ABI={abi}"]
    getter = getattr(ce, "_get_pybind11_abi_build_flags", None)
    if getter is not None:
        flags += list(getter())
    out = PKG / ("_C" + extension_suffix())
    tmp = out.with_suffix(".tmp.so")
    cmd = ["g++", *flags, f"-I{INCLUDE}", f"-I{py_inc}", *[f"-I{p}" for p in incs],
           str(CSRC / "bindings.cpp"), "-o", str(tmp),
           f"-L{LIBDIR}", "-lquantom_loits_hip", f"-L{tlib}", "-lc10", "-ltorch", "-ltorch_cpu",
           "-ltorch_python", "-Wl,-rpath,$ORIGIN/lib", f"-Wl,-rpath,{tlib}"]
    _run(cmd, quiet)
    os.replace(tmp, out)
    return out


def build(quiet=True) -> None:
    build_library(quiet)
    build_bindings(quiet)


if __name__ == "__main__":
    build(quiet=False)
    print("built", LIBDIR / LIBNAME, "and", PKG / ("_C" + extension_suffix()))
