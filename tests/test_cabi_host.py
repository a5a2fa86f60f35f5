"""C ABI and host logic without a GPU: exports, layouts, argument errors, registry."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest
import torch  # noqa: F401  (PyTorch's HIP runtime first, as in the package)

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "loits_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(loits_hip_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from quantom_loits_amd._native import library_path

    path = library_path()
    if not path.is_file():
        pytest.skip("HIP library not built")
    return C.CDLL(str(path))


def test_library_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 14
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


class Instance(C.Structure):
    _fields_ = [("node_off", C.c_int64), ("n_nodes", C.c_int64), ("events", C.c_int64),
                ("out_off", C.c_int64), ("counter_base", C.c_uint64), ("stream", C.c_uint64),
                ("accept", C.c_int32), ("reserved", C.c_int32)]


class Tuning(C.Structure):
    _fields_ = [("threads", C.c_int32), ("guide_shift", C.c_int32), ("chunk_events", C.c_int64),
                ("scan_mode", C.c_int32), ("persistent", C.c_int32), ("table_format", C.c_int32),
                ("reserved", C.c_int32), ("bucket_chunk_log2", C.c_int32 * 16),
                ("bucket_guide_shift", C.c_int32 * 16)]


class Layout(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "n_instances", "n_chunks", "capacity", "total_nodes", "max_nodes", "desc_bytes",
        "table_bytes", "scratch_bytes", "lds_forward", "lds_backward", "lds_tables", "n_segments",
        "mask_words", "partial_doubles")] + [(n, C.c_int32) for n in (
            "dtype", "threads", "scan_mode", "persistent", "grid_forward", "grid_backward",
            "global_instances", "range_instances", "cell_instances", "reserved")] + \
        [(n, C.c_int64) for n in ("backward_plain", "cell_chunks", "backward_cell",
                                  "forward_items", "lds_persistent_forward")]


def make_instances(nodes, events):
    arr = (Instance * len(nodes))()
    off = out = 0
    for i, (n, e) in enumerate(zip(nodes, events)):
        arr[i] = Instance(off, n, e, out, 1000 * i, 0, 1, 0)
        off += n
        out += e
    return arr


def last_error(lib):
    lib.loits_hip_last_error.restype = C.c_char_p
    return lib.loits_hip_last_error().decode()


def test_batched_layout_sizes_and_tiles(lib):
    inst = make_instances([2048] * 3, [10000, 0, 4096])
    tun = Tuning(256, 1, 4096, 0, 0)
    lay = Layout()
    assert lib.loits_hip_batched_layout(inst, 3, 1, C.byref(tun), C.byref(lay)) == 0
    assert lay.capacity == 14096 and lay.total_nodes == 3 * 2048 and lay.max_nodes == 2048
    assert lay.n_chunks == 3 + 1 + 1  # ceil(10000/4096), empty instance keeps one tile, 1
    assert lay.n_segments == lay.n_chunks * 4 * 4  # 256 threads = 4 waves, 4 segments each
    assert lay.lds_forward > 0 and lay.lds_backward > lay.lds_forward - 32 * 2048
    assert lay.partial_doubles == lay.n_chunks * 2 * 2048  # 2 pad4(n) per chunk
    blob = (C.c_uint8 * lay.desc_bytes)()
    assert lib.loits_hip_batched_encode(inst, 3, C.byref(tun), C.byref(lay), blob) == 0
    # a layout for different tuning is rejected by encode
    tun2 = Tuning(512, 1, 4096, 0, 0)
    assert lib.loits_hip_batched_encode(inst, 3, C.byref(tun2), C.byref(lay), blob) == 1
    assert "does not match" in last_error(lib)


def test_batched_persistent_queue_and_buckets(lib):
    # 12000-node float32 curves keep global tables (built by a launch of their own)
    inst = make_instances([12000, 512, 2048], [70000, 5000, 300000])
    tun = Tuning(1024, 2, 1 << 16, 0, 1)
    lay = Layout()
    assert lib.loits_hip_batched_layout(inst, 3, 1, C.byref(tun), C.byref(lay)) == 0, last_error(lib)
    assert lay.persistent == 1 and lay.global_instances == 1
    assert lay.forward_items == lay.n_chunks
    assert lay.lds_persistent_forward == lay.lds_forward
    blob = (C.c_uint8 * lay.desc_bytes)()
    assert lib.loits_hip_batched_encode(inst, 3, C.byref(tun), C.byref(lay), blob) == 0
    # per-bucket chunk sizes: the 2048-node curve (bucket 11) in chunks of 2^17, the
    # 512-node curve (bucket 9) in one chunk of 2^13
    base_chunks = lay.n_chunks
    tun.bucket_chunk_log2[11] = 17
    tun.bucket_chunk_log2[9] = 13
    lay2 = Layout()
    assert lib.loits_hip_batched_layout(inst, 3, 1, C.byref(tun), C.byref(lay2)) == 0
    # 12000 nodes: ceil(70000 / 2^16) = 2; 512: 1; 2048: ceil(300000 / 2^17) = 3 (was 5)
    assert base_chunks == 2 + 1 + 5 and lay2.n_chunks == 2 + 1 + 3
    for field, val, msg in (("bucket_chunk_log2", 9, "bucket_chunk_log2"),
                            ("bucket_guide_shift", 5, "bucket_guide_shift")):
        t = Tuning(1024, 2, 1 << 16, 0, 0)
        getattr(t, field)[3] = val
        assert lib.loits_hip_batched_layout(inst, 3, 1, C.byref(t), C.byref(Layout())) == 1
        assert msg in last_error(lib)
    t = Tuning(1024, 2, 1 << 16, 0, 2)
    assert lib.loits_hip_batched_layout(inst, 3, 1, C.byref(t), C.byref(Layout())) == 1
    assert "persistent must be 0 or 1" in last_error(lib)


@pytest.mark.parametrize("bad,msg", [
    (dict(threads=300), "threads must be"),
    (dict(nodes=[1]), "n_nodes"),
    (dict(out_shift=5), "out_off"),
    (dict(nodes=[40000]), "n_nodes <= 32767"),
    (dict(nodes=[20000], threads=256), "n_nodes / threads"),
])
def test_batched_layout_rejects_bad_arguments(lib, bad, msg):
    nodes = bad.get("nodes", [64])
    inst = make_instances(nodes, [100] * len(nodes))
    if "out_shift" in bad:
        inst[0].out_off = bad["out_shift"]
    tun = Tuning(bad.get("threads", 256), 1, 4096, 0, 0)
    if "nodes" in bad and bad["nodes"] == [20000]:
        tun = Tuning(256, 1, 4096, 0, 0)
    lay = Layout()
    rc = lib.loits_hip_batched_layout(inst, len(nodes), 1, C.byref(tun), C.byref(lay))
    assert rc in (1, 3) and msg in last_error(lib)


def test_workspace_size_queries(lib):
    class Shape(C.Structure):
        _fields_ = [(n, C.c_int64) for n in ("batch", "nx", "ny", "kx", "kq", "cells")]

    lib.loits_hip_forward_workspace_size.restype = C.c_int64
    lib.loits_hip_backward_workspace_size.restype = C.c_int64
    s = Shape(1, 3, 3, 5, 5, 9)
    assert lib.loits_hip_forward_workspace_size(C.byref(s), 80) > 0
    assert lib.loits_hip_backward_workspace_size(C.byref(s), 80) >= 2 * 8 * 45 * 2
    bad = Shape(1, 3, 3, 1, 5, 9)  # kx < 2
    assert lib.loits_hip_forward_workspace_size(C.byref(bad), 80) == -1


def test_version_string(lib):
    lib.loits_hip_version.restype = C.c_char_p
    assert b"gfx950" in lib.loits_hip_version()


def test_pybind_plan_layout_on_host():
    from quantom_loits_amd.batched import BatchedLOITS, Tuning as T

    s = BatchedLOITS.uniform(4, 1024, 65536, 0, dtype=torch.float32, device="cpu",
                             tuning=T(threads=1024, chunk_events=1 << 16))
    assert s.capacity == 4 * 65536 and s.layout["n_chunks"] == 4
    with pytest.raises(RuntimeError, match="n_nodes"):
        BatchedLOITS(np.array([1]), np.array([5]), device="cpu")


# --- registry (loits.py semantics) -------------------------------------------
def test_registry_reports_hip_unavailable_on_cpu():
    from quantom_loits_amd import BackendUnavailableError, LOITS, backend_status
    from quantom_loits_amd.loits import registered_backends

    assert "hip" in registered_backends()
    ok, reason = backend_status("hip", "cpu")
    assert not ok and "GPU" in reason
    with pytest.raises(BackendUnavailableError, match="unavailable for device=cpu"):
        LOITS(backend="hip", device="cpu")
    with pytest.raises(ValueError, match="not registered"):
        LOITS(backend="nope")


def test_registry_probe_semantics():
    from quantom_loits_amd.loits import _BACKENDS, _BACKEND_PROBES, backend_status, register_backend

    def boom(device):
        raise RuntimeError("probe exploded")

    try:
        register_backend("t_plain", lambda **kw: None)
        register_backend("t_bool", lambda **kw: None, lambda d: False)
        register_backend("t_raise", lambda **kw: None, boom)
        assert backend_status("t_plain") == (True, "")
        assert backend_status("t_bool") == (False, "backend probe returned unavailable")
        assert backend_status("t_raise") == (False, "probe exploded")
    finally:
        for n in ("t_plain", "t_bool", "t_raise"):
            _BACKENDS.pop(n, None)
            _BACKEND_PROBES.pop(n, None)


def test_hip_backend_rejects_other_epsilon():
    from quantom_loits_amd.backend import HipLOITS

    with pytest.raises(ValueError, match="epsilon 1e-5"):
        HipLOITS(device="cuda", epsilon=1e-4)


def test_large_curves_use_global_tables(lib):
    """Curves whose tables do not fit in LDS keep them in global memory."""
    inst = make_instances([2048, 6000, 16384], [1000, 1000, 1000])
    tun = Tuning(1024, 2, 4096, 0, 0)
    lay = Layout()
    assert lib.loits_hip_batched_layout(inst, 3, 1, C.byref(tun), C.byref(lay)) == 0
    assert lay.global_instances == 2
    assert lay.lds_forward <= 160 * 1024 and lay.lds_backward <= 160 * 1024
    # float32 VJP bin ranges of at most 7168 bins: 6000 nodes -> 1, 16384 -> 3 work items
    assert lay.range_instances == 2 and lay.grid_backward == 1 + 1 + 3
    assert lay.backward_plain == 1


def test_vjp_bin_ranges_without_global_tables(lib):
    """Guide tables (table_format 1): at 4096 nodes the forward builds its tables
    in LDS and the VJP takes one bin range; 2500 nodes keep both in LDS.  With
    cell tables (the default) both curves keep everything in LDS."""
    inst = make_instances([4096, 2500], [300000, 10])
    lay = Layout()
    cell = Tuning(1024, 2, 1 << 17, 0, 0)
    assert lib.loits_hip_batched_layout(inst, 2, 1, C.byref(cell), C.byref(lay)) == 0
    assert lay.cell_instances == 2 and lay.global_instances == 0 and lay.range_instances == 0
    assert lay.n_chunks == 3 + 1 and lay.grid_backward == 4 and lay.backward_plain == 4
    assert lay.cell_chunks == 4 and lay.backward_cell == 4
    assert max(lay.lds_forward, lay.lds_backward) <= 160 * 1024 - 2048
    # float64 curves and chunks above 2^18 events keep the guide tables
    assert lib.loits_hip_batched_layout(inst, 2, 0, C.byref(cell), C.byref(lay)) == 0
    assert lay.cell_instances == 0
    big = Tuning(1024, 2, 1 << 19, 0, 0)
    assert lib.loits_hip_batched_layout(inst, 2, 1, C.byref(big), C.byref(lay)) == 0
    assert lay.cell_instances == 0
    tun = Tuning(1024, 2, 1 << 17, 0, 0, 1)
    assert lib.loits_hip_batched_layout(inst, 2, 1, C.byref(tun), C.byref(lay)) == 0
    assert lay.cell_instances == 0 and lay.cell_chunks == 0 and lay.backward_cell == 0
    assert lay.global_instances == 0 and lay.range_instances == 1
    assert lay.n_chunks == 3 + 1 and lay.grid_backward == 4 and lay.backward_plain == 1
    assert lay.lds_backward <= 160 * 1024 - 2048
