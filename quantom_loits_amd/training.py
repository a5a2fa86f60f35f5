"""GAN-style training step around the batched HIP sampler.

The reference's training benchmark (pytorch/gan.py:26-122, driven by
benchmark_training.py:66-98) alternates a discriminator step and a generator
step, with the generator's gradient flowing back through LOITS.  This harness
keeps that step structure and its ``gan::*`` regions at a much larger shape:

* generator: per-instance latents z [B, 32] -> Linear 256 -> GELU -> Linear
  256 -> GELU -> Linear N -> softplus, giving B float32 densities of N nodes
  on x = linspace(0.001, 0.999, N).  (Latents are per instance; one latent
  mapped to all B*N values would put a 256 x B*N layer in the generator.)
* sampler: ``BatchedLOITS`` (HIP forward + VJP), E events per density, kept
  only for x in [0.05, 0.95]; stream 2*sequence with the sequence advanced per
  call as the reference's backends do (cpp/backend.py:34-101).
* soft binning: 64 bins over [0.05, 0.95], Gaussian kernel of width 0.5 bin,
  seven taps per event, normalised by the instance's kept count;
* discriminator: 64 -> 128 -> GELU -> 128 -> GELU -> 1 MLP on the histograms,
  BCE-with-logits losses, Adam (lr 1e-3) for both players.

``backend="torch"`` swaps the sampler for a torch-ops comparator with the same
arithmetic (torch.rand uniforms, a sorted search, the same interpolation and
window) so that the step time and LOITS share can be compared on one GPU.  It
is a comparison leg only and is never used by the HIP backend.
"""
from __future__ import annotations

import argparse
import json
import math
import os

import numpy as np
import torch
from torch import nn

from . import profiling
from .batched import BatchedLOITS, Tuning
from .workloads import uniform_family

WINDOW = (0.05, 0.95)
BINS = 64
EPSILON = 1e-5


class DensityGenerator(nn.Module):
    def __init__(self, n_nodes, latent=32, width=256):
        super().__init__()
        self.net = nn.Sequential(nn.Linear(latent, width), nn.GELU(), nn.Linear(width, width),
                                 nn.GELU(), nn.Linear(width, n_nodes))

    def forward(self, z):
        return nn.functional.softplus(self.net(z))


class HistogramDiscriminator(nn.Module):
    def __init__(self, bins=BINS, width=128):
        super().__init__()
        self.net = nn.Sequential(nn.Linear(bins, width), nn.GELU(), nn.Linear(width, width),
                                 nn.GELU(), nn.Linear(width, 1))

    def forward(self, hist):
        return self.net(hist)


def soft_histogram(values, kept, lo=WINDOW[0], hi=WINDOW[1], bins=BINS, width=0.5):
    """Differentiable per-instance histogram of packed values.

    Event v adds exp(-(t - c - 1/2)^2 / (2 width^2)) to bin c for the seven
    bins nearest t = (v - lo) / bin_width (width in bins; the kernel is below
    1e-7 beyond three bins); each instance's row is divided by its kept count.
    """
    B = kept.numel()
    bw = (hi - lo) / bins
    inst = torch.repeat_interleave(torch.arange(B, device=values.device), kept,
                                   output_size=values.numel())
    t = (values.float() - lo) / bw
    c = torch.floor(t).long().unsqueeze(1) + torch.arange(-3, 4, device=values.device)
    d = t.unsqueeze(1) - (c.float() + 0.5)
    w = torch.exp(-(d * d) / (2.0 * width * width))
    ok = (c >= 0) & (c < bins)
    idx = (inst.unsqueeze(1) * bins + c.clamp(0, bins - 1))
    hist = torch.zeros(B * bins, device=values.device, dtype=torch.float32)
    hist = hist.index_add(0, idx.reshape(-1), (w * ok).reshape(-1))
    return hist.view(B, bins) / kept.clamp(min=1).unsqueeze(1).float()


class _SoftHistogramHip(torch.autograd.Function):
    """soft_histogram on the GPU: the HIP kernels of loits_hip_soft_histogram."""

    @staticmethod
    def forward(ctx, values, kept, lo, hi, bins, width):
        from ._native import load_extension

        ext = load_extension()
        v = values.detach().float().contiguous()
        row_start = torch.zeros(kept.numel() + 1, dtype=torch.int64, device=values.device)
        torch.cumsum(kept, 0, out=row_start[1:])
        hist = ext.soft_histogram(v, row_start, lo, hi, bins, width)
        ctx.save_for_backward(v, row_start)
        ctx.params = (lo, hi, bins, width)
        ctx.dtype = values.dtype
        return hist

    @staticmethod
    def backward(ctx, grad_hist):
        from ._native import load_extension

        v, row_start = ctx.saved_tensors
        lo, hi, bins, width = ctx.params
        g = load_extension().soft_histogram_backward(v, row_start, lo, hi, bins, width,
                                                     grad_hist.float().contiguous())
        return g.to(ctx.dtype), None, None, None, None, None


def soft_histogram_hip(values, kept, lo=WINDOW[0], hi=WINDOW[1], bins=BINS, width=0.5):
    """soft_histogram with the forward and backward in HIP kernels (GPU tensors)."""
    hist = _SoftHistogramHip.apply(values, kept, lo, hi, bins, width)
    return hist / kept.clamp(min=1).unsqueeze(1).float()


class TorchSampler:
    """Torch-ops comparator for the 1-D sampler (comparison leg only).

    Same arithmetic as the reference's torch backend (pytorch/loits.py:16-71):
    trapezoid normalisation and CDF, interval j = clamp(#{t: u >= C_t} - 1, 0,
    N-2) (a sorted search on the non-decreasing CDF), value x_j + (x_{j+1} -
    x_j) (u - C_j) / (C_{j+1} - C_j + eps), then the window cut; uniforms come
    from torch.rand.  Gradients come from torch autograd.
    """

    def __init__(self, x, events, window=WINDOW):
        self.x = x  # [B, N]
        self.events = events
        self.window = window
        self.event_log = None

    def sample(self, density):
        fn = _TimedTorchSampler.apply
        return fn(density, self)

    def _forward(self, density):
        x = self.x
        dx = x[:, 1:] - x[:, :-1]
        area = 0.5 * dx * (density[:, 1:] + density[:, :-1])
        norm = area.sum(1, keepdim=True)
        rho = density / norm
        seg = 0.5 * dx * (rho[:, 1:] + rho[:, :-1])
        cdf = torch.cat([torch.zeros_like(rho[:, :1]), torch.cumsum(seg, 1)], 1)
        B, N = x.shape
        u = torch.rand(B, self.events, device=x.device, dtype=x.dtype)
        j = torch.searchsorted(cdf.detach().contiguous(), u, right=True) - 1
        j = j.clamp(0, N - 2)
        c0 = cdf.gather(1, j)
        c1 = cdf.gather(1, j + 1)
        x0 = x.gather(1, j)
        x1 = x.gather(1, j + 1)
        v = x0 + (x1 - x0) * (u - c0) / (c1 - c0 + EPSILON)
        keep = (v >= self.window[0]) & (v <= self.window[1])
        return v[keep], keep.sum(1)


class _TimedTorchSampler(torch.autograd.Function):
    """Runs the comparator's torch graph inside one autograd node so that its
    forward and backward can each be bracketed by a CUDA-event pair."""

    @staticmethod
    def forward(ctx, density, sampler):
        log = sampler.event_log
        start = torch.cuda.Event(enable_timing=True) if log is not None else None
        if start is not None:
            start.record()
        with torch.enable_grad():
            d = density.detach().requires_grad_(density.requires_grad)
            values, kept = sampler._forward(d)
        if start is not None:
            end = torch.cuda.Event(enable_timing=True)
            end.record()
            log.append(("forward", start, end))
        ctx.graph = (d, values)
        ctx.sampler = sampler
        ctx.mark_non_differentiable(kept)
        return values.detach(), kept

    @staticmethod
    def backward(ctx, grad, _):
        d, values = ctx.graph
        log = ctx.sampler.event_log
        start = torch.cuda.Event(enable_timing=True) if log is not None else None
        if start is not None:
            start.record()
        (gd,) = torch.autograd.grad(values, d, grad)
        if start is not None:
            end = torch.cuda.Event(enable_timing=True)
            end.record()
            log.append(("backward", start, end))
        return gd, None


class BatchedGanTrainer:
    """Discriminator step then generator step (pytorch/gan.py:107-122)."""

    def __init__(self, backend="hip", device="cuda", batch=512, nodes=2048, events=1 << 16,
                 seed=0, profile_regions=False, tuning=None):
        if backend not in ("hip", "torch"):
            raise ValueError("backend must be 'hip' or 'torch'")
        self.seed = int(seed)
        torch.manual_seed(self.seed)
        self.backend = backend
        self.device = torch.device(device)
        self.batch, self.nodes, self.events = batch, nodes, events
        self.profile_regions = profile_regions
        grid = np.linspace(0.001, 0.999, nodes).astype(np.float32)
        self.x = torch.from_numpy(np.tile(grid, batch)).to(self.device)
        self.z = torch.randn(batch, 32, device=self.device)
        self.generator = DensityGenerator(nodes).to(self.device)
        self.discriminator = HistogramDiscriminator().to(self.device)
        self.g_optimizer = torch.optim.Adam(self.generator.parameters(), lr=1e-3)
        self.d_optimizer = torch.optim.Adam(self.discriminator.parameters(), lr=1e-3)
        self.loss = nn.BCEWithLogitsLoss()
        self.sequence = 0
        self.loits_seed = 1234
        if backend == "hip":
            self.sampler = BatchedLOITS.uniform(batch, nodes, events, 0, dtype=torch.float32,
                                                device=self.device, window=WINDOW,
                                                tuning=tuning or Tuning(),
                                                profile_regions=profile_regions)
        else:
            self.sampler = TorchSampler(self.x.view(batch, nodes), events)
        _, target = uniform_family(batch, nodes, seed=self.seed, dtype=np.float32)
        target = torch.from_numpy(target.reshape(-1)).to(self.device)
        self.histogram = soft_histogram_hip if self.device.type == "cuda" else soft_histogram
        with torch.no_grad():
            values, kept = self._sample(target)
            self.real_hist = self.histogram(values, kept)

    # -- the sampler call, one new sequence per call ---------------------------
    def _sample(self, density):
        if self.backend == "hip":
            self.sampler.set_streams(np.full(self.batch, 2 * self.sequence, np.uint64))
            out = self.sampler.sample(self.x, density.reshape(-1).contiguous(), self.loits_seed)
        else:
            out = self.sampler.sample(density.reshape(self.batch, self.nodes))
        self.sequence += 1
        return out

    def _region(self, name):
        if not self.profile_regions:
            return _Null()
        return torch.profiler.record_function(name)

    def snapshot_generator_params(self):
        return [p.detach().clone() for p in self.generator.parameters()]

    def restore_generator_params(self, params):
        with torch.no_grad():
            for p, v in zip(self.generator.parameters(), params):
                p.copy_(v)

    def step(self):
        with self._region("gan::discriminator_step"):
            self.d_optimizer.zero_grad(set_to_none=True)
            with torch.no_grad():
                values, kept = self._sample(self.generator(self.z))
                fake_hist = self.histogram(values, kept)
            real = self.discriminator(self.real_hist)
            fake = self.discriminator(fake_hist)
            d_loss = (self.loss(real, torch.ones_like(real))
                      + self.loss(fake, torch.zeros_like(fake)))
            d_loss.backward()
            self.d_optimizer.step()
        with self._region("gan::generator_step"):
            self.g_optimizer.zero_grad(set_to_none=True)
            values, kept = self._sample(self.generator(self.z))
            logits = self.discriminator(self.histogram(values, kept))
            g_loss = self.loss(logits, torch.ones_like(logits))
            g_loss.backward()
            self.g_optimizer.step()
        return d_loss.detach(), g_loss.detach()


class _Null:
    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


def loits_share(trainer, warmup=2, iterations=5):
    """(step ms, LOITS forward+backward ms per step) from CUDA events."""
    params = trainer.snapshot_generator_params()
    for _ in range(warmup):
        trainer.restore_generator_params(params)
        trainer.step()
    profiling.synchronize(trainer.device)
    steps, loits = [], []
    for _ in range(iterations):
        trainer.restore_generator_params(params)
        log = []
        trainer.sampler.event_log = log
        profiling.synchronize(trainer.device)
        s0 = torch.cuda.Event(enable_timing=True)
        s1 = torch.cuda.Event(enable_timing=True)
        s0.record()
        trainer.step()
        s1.record()
        profiling.synchronize(trainer.device)
        trainer.sampler.event_log = None
        steps.append(s0.elapsed_time(s1))
        loits.append(sum(a.elapsed_time(b) for _, a, b in log))
    return float(np.mean(steps)), float(np.mean(loits))


def main(argv=None):
    p = argparse.ArgumentParser(description="GAN-style training step with the HIP sampler")
    p.add_argument("--backend", default="hip", choices=("hip", "torch"))
    p.add_argument("--device", default="cuda")
    p.add_argument("--batch", type=int, default=512)
    p.add_argument("--grid-size", type=int, default=2048, help="nodes per density")
    p.add_argument("--events", type=int, default=1 << 16)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--iterations", type=int, default=10)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--output", default="results/training")
    p.add_argument("--site", default=os.environ.get("QUANTOM_SITE", "mi355x"))
    p.add_argument("--experiment", default="fixed", choices=("fixed", "strong", "weak"))
    p.add_argument("--regions", action="store_true")
    p.add_argument("--trace", action="store_true")
    args = p.parse_args(argv)

    impl = "hip-batched" if args.backend == "hip" else "torch-ops"
    meta = profiling.row_metadata(args.site, args.experiment, args.backend, impl, args.device,
                                  args.events, "", args.grid_size, args.warmup, args.iterations,
                                  args.seed)
    out = os.path.join(args.output, profiling.safe_component(args.site),
                       profiling.safe_component(args.experiment))
    trainer = BatchedGanTrainer(args.backend, args.device, args.batch, args.grid_size,
                                args.events, args.seed)
    samples = profiling.measure(trainer, args.warmup, args.iterations)
    stem = profiling.result_stem("training", impl, args.device, args.events)
    profiling.write_csv(os.path.join(out, stem + ".csv"), profiling.timing_rows(samples, meta))
    step_ms, loits_ms = loits_share(trainer, 1, max(2, args.iterations // 2))
    summary = {"backend": args.backend, "batch": args.batch, "nodes": args.grid_size,
               "events": args.events, "step_ms": float(np.mean(samples)),
               "step_ms_std": float(np.std(samples)), "event_step_ms": step_ms,
               "loits_ms": loits_ms, "loits_share": loits_ms / step_ms if step_ms else math.nan}
    if args.regions:
        rtrainer = BatchedGanTrainer(args.backend, args.device, args.batch, args.grid_size,
                                     args.events, args.seed, profile_regions=True)
        trace = (os.path.join(out, profiling.result_stem("trace", impl, args.device,
                                                         args.events) + ".json")
                 if args.trace else None)
        prof = profiling.profile(rtrainer, args.warmup, args.iterations, trace)
        rstem = profiling.result_stem("regions", impl, args.device, args.events)
        profiling.write_csv(os.path.join(out, rstem + ".csv"),
                            profiling.region_rows(prof.events(), meta))
    print(json.dumps(summary))
    return summary


if __name__ == "__main__":
    main()
