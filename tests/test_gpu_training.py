"""GAN-style training step through the HIP sampler: gradients reach the
generator, the ``gan::*`` / ``loits::*`` regions nest as in the reference's
profiler rows, and the CLI writes training_* / regions_* CSV files."""
import csv
import json

import pytest
import torch

pytestmark = pytest.mark.gpu


def test_training_step_updates_generator_through_hip_vjp(hip_device):
    from quantom_loits_amd.training import BatchedGanTrainer

    print("gan step hip", flush=True)
    t = BatchedGanTrainer("hip", "cuda", batch=8, nodes=256, events=4096, seed=3)
    before = [p.detach().clone() for p in t.generator.parameters()]
    d_loss, g_loss = t.step()
    assert torch.isfinite(d_loss) and torch.isfinite(g_loss)
    last = list(t.generator.parameters())[-1]
    assert last.grad is not None and torch.isfinite(last.grad).all()
    assert float(last.grad.abs().max()) > 0
    moved = [float((p.detach() - b).abs().max()) for p, b in zip(t.generator.parameters(), before)]
    assert all(m > 0 for m in moved)
    assert t.sequence == 3  # real sample at construction + two calls per step


def test_regions_nest_like_the_reference(hip_device):
    from quantom_loits_amd import profiling
    from quantom_loits_amd.training import BatchedGanTrainer

    print("gan regions", flush=True)
    t = BatchedGanTrainer("hip", "cuda", batch=4, nodes=128, events=2048, profile_regions=True)
    prof = profiling.profile(t, warmup=1, iterations=2)
    meta = profiling.row_metadata("s", "fixed", "hip", "hip-batched", "cuda", 2048, "", 128, 1,
                                  2, 0)
    rows = profiling.region_rows(prof.events(), meta)
    names = {r["region"] for r in rows}
    for want in ("gan::training_iteration", "gan::discriminator_step", "gan::generator_step",
                 "loits::forward", "loits::forward::allocation",
                 "loits::forward::interpolation_x", "loits::forward::stream_compaction",
                 "loits::backward", "loits::backward::interpolation_x"):
        assert want in names, want
    parents = {(r["region"], r["parent_region"]) for r in rows}
    assert ("loits::forward", "gan::discriminator_step") in parents
    assert ("loits::forward", "gan::generator_step") in parents
    assert ("loits::forward::interpolation_x", "loits::forward") in parents
    assert ("loits::backward::interpolation_x", "loits::backward") in parents


def test_cli_writes_reference_csv_files(hip_device, tmp_path, capsys):
    from quantom_loits_amd import profiling
    from quantom_loits_amd.training import main

    print("gan cli", flush=True)
    summary = main(["--batch", "4", "--grid-size", "128", "--events", "2048", "--warmup", "1",
                    "--iterations", "2", "--output", str(tmp_path), "--site", "box",
                    "--regions"])
    assert 0 < summary["loits_ms"] < summary["event_step_ms"]
    d = tmp_path / "box" / "fixed"
    with (d / "training_hip-batched_cuda_e2048.csv").open() as f:
        rows = list(csv.DictReader(f))
    assert len(rows) == 2 and tuple(rows[0]) == profiling.TIMING_COLUMNS
    assert rows[0]["backend"] == "hip" and rows[0]["grid_size"] == "128"
    with (d / "regions_hip-batched_cuda_e2048.csv").open() as f:
        rrows = list(csv.DictReader(f))
    assert tuple(rrows[0]) == profiling.REGION_COLUMNS
    assert any(r["region"] == "loits::backward" for r in rrows)
    out = capsys.readouterr().out.strip().splitlines()[-1]
    assert json.loads(out)["backend"] == "hip"


def test_torch_comparator_step_runs(hip_device):
    from quantom_loits_amd.training import BatchedGanTrainer, loits_share

    print("gan step torch", flush=True)
    t = BatchedGanTrainer("torch", "cuda", batch=4, nodes=128, events=2048)
    step_ms, loits_ms = loits_share(t, warmup=1, iterations=2)
    assert 0 < loits_ms < step_ms


@pytest.mark.parametrize("width", [0.5, 0.1])
def test_hip_soft_histogram_matches_torch_layer(hip_device, width):
    """The HIP soft-binning kernels against the torch formulation: values,
    gradients, an empty instance, instances spanning work-groups, values
    outside the window; width 0.5 bins takes the factored taps (three
    exponentials per row), width 0.1 the seven exponentials."""
    from quantom_loits_amd.training import BINS, WINDOW, soft_histogram, soft_histogram_hip

    print("soft histogram width", width, flush=True)
    g = torch.Generator(device="cuda").manual_seed(5)
    kept = torch.tensor([30000, 0, 5, 20000, 1, 17000, 9000], device="cuda")
    v = torch.rand(int(kept.sum()), device="cuda", generator=g) * 1.1 - 0.05
    v[-9000:] = 0.5 + 0.01 * torch.randn(9000, device="cuda", generator=g)  # one crowded bin
    v1 = v.clone().requires_grad_(True)
    v2 = v.clone().requires_grad_(True)
    h1 = soft_histogram_hip(v1, kept, WINDOW[0], WINDOW[1], BINS, width)
    h2 = soft_histogram(v2, kept, WINDOW[0], WINDOW[1], BINS, width)
    assert torch.allclose(h1, h2, rtol=1e-4, atol=1e-6)
    w = torch.randn(h1.shape, device="cuda", generator=g)
    (h1 * w).sum().backward()
    (h2 * w).sum().backward()
    assert torch.allclose(v1.grad, v2.grad, rtol=1e-3, atol=1e-3 * float(v2.grad.abs().max()))
