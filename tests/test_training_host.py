"""Host-side checks of the training harness: CSV rows in the reference's
format, the soft-binning layer and the torch-ops comparator (CPU only)."""
import csv
import math
from types import SimpleNamespace

import numpy as np
import torch

from quantom_loits_amd import profiling
from quantom_loits_amd.training import TorchSampler, soft_histogram


def test_metadata_and_file_stems_follow_the_reference():
    # benchmark_training.py:34-65: stem parts and metadata column order
    assert profiling.result_stem("training", "hip-batched", "cuda:0", 65536) == \
        "training_hip-batched_cuda-0_e65536"
    assert profiling.result_stem("regions", "openmp", "cpu", 10, "16") == \
        "regions_openmp_cpu_e10_t16"
    assert profiling.safe_component("a b/c") == "a-b-c"
    meta = profiling.row_metadata("s", "fixed", "hip", "hip-batched", "cuda", 8, "", 256, 1, 2, 0)
    assert tuple(meta) == ("site", "experiment", "backend", "implementation", "device",
                           "events", "threads", "grid_size", "warmup", "iterations", "seed")
    rows = profiling.timing_rows([1.5, 2.5], meta)
    assert tuple(rows[0]) == profiling.TIMING_COLUMNS
    assert [r["occurrence"] for r in rows] == [0, 1]
    assert rows[1]["region"] == "gan::training_iteration" and rows[1]["wall_ms"] == 2.5


def _ev(id_, name, parent, cpu_us):
    return SimpleNamespace(id=id_, name=name, cpu_parent=parent, cpu_time_total=cpu_us,
                           thread=1, sequence_nr=-1,
                           time_range=SimpleNamespace(start=0.0, end=cpu_us),
                           device_time_total=0.0, self_cpu_time_total=cpu_us,
                           self_device_time_total=0.0)


def test_region_rows_skip_untracked_parents(tmp_path):
    it = _ev(1, "gan::training_iteration", None, 100.0)
    step = _ev(2, "gan::generator_step", it, 80.0)
    other = _ev(3, "aten::mm", step, 5.0)
    fwd = _ev(4, "loits::forward", other, 20.0)
    empty = _ev(5, "loits::forward::allocation", fwd, 0.0)
    meta = profiling.row_metadata("s", "fixed", "hip", "hip", "cuda", 8, "", 256, 1, 1, 0)
    rows = profiling.region_rows([it, step, other, fwd, empty], meta)
    assert [r["region"] for r in rows] == ["gan::training_iteration", "gan::generator_step",
                                           "loits::forward"]
    assert rows[2]["parent_region"] == "gan::generator_step" and rows[2]["parent_event_id"] == 2
    assert rows[0]["parent_region"] == "" and rows[0]["region_timing"] == "synchronized_wall"
    assert tuple(rows[0]) == profiling.REGION_COLUMNS
    path = tmp_path / "regions_x.csv"
    profiling.write_csv(path, rows)
    with path.open() as f:
        got = list(csv.DictReader(f))
    assert len(got) == 3 and float(got[2]["cpu_ms"]) == 0.02


def test_soft_histogram_weights_and_gradient():
    lo, hi, bins = 0.05, 0.95, 64
    bw = (hi - lo) / bins
    centre = lo + 10.5 * bw
    values = torch.tensor([centre, centre + 0.2 * bw, 0.93], requires_grad=True)
    kept = torch.tensor([2, 1])
    h = soft_histogram(values, kept)
    assert h.shape == (2, bins)
    # a value at a bin centre: weights exp(-k^2 / (2 * 0.5^2)) on bins 10 +- k
    expect = np.exp(-np.arange(-3, 4) ** 2 / 0.5)
    single = soft_histogram(torch.tensor([centre]), torch.tensor([1]))[0]
    np.testing.assert_allclose(single[7:14].numpy(), expect, rtol=1e-5)
    assert float(single.sum()) == float(single[7:14].sum())
    # the value near the top edge loses the taps beyond bin 63
    assert float(h[1].sum()) < expect.sum()
    # d/dv of row 0, bin 11: w(d) = exp(-d^2/(2 s^2)), d = t - 11.5, t = (v - lo) / bw
    h[0, 11].backward()
    t = (centre + 0.2 * bw - lo) / bw
    d = t - 11.5
    analytic = math.exp(-d * d / 0.5) * (-d / 0.25) / bw / 2.0
    assert abs(float(values.grad[1]) - analytic) <= 1e-3 * abs(analytic)


def test_torch_comparator_keeps_window_and_counts():
    torch.manual_seed(0)
    B, N, E = 3, 64, 2000
    grid = torch.linspace(0.001, 0.999, N).repeat(B, 1)
    dens = (torch.rand(B, N) + 0.5).requires_grad_(True)
    s = TorchSampler(grid, E)
    values, kept = s.sample(dens)
    assert int(kept.sum()) == values.numel() > 0
    assert float(values.min()) >= 0.05 and float(values.max()) <= 0.95
    values.sum().backward()
    assert dens.grad is not None and torch.isfinite(dens.grad).all()
