"""Synthetic LOITS workloads: the density family and instance layouts.

Density family (one curve per instance, nodes x in [0.001, 0.999]):

    f(x) = x^a (1-x)^b + A exp(-(x-mu)^2 / (2 sigma^2))
    a ~ U(-0.5, 1.5), b ~ U(1, 5), A ~ U(0, 0.5), mu ~ U(0.2, 0.8),
    sigma ~ U(0.02, 0.1)

evaluated in float64 from a seeded numpy PCG64 stream and stored in the
requested dtype.  Ragged batches draw N_i and E_i log-uniformly.

The 1-D mapping of the reference call (one x-curve per instance, nx = ny = 1,
dummy q axis with two nodes and a constant density, weights 1, acceptance
True) is produced by ``reference_1d_inputs`` so that a reference-style call and
the batched path can be compared on the same data.
"""
from __future__ import annotations

import numpy as np


def density_params(B: int, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    a = rng.uniform(-0.5, 1.5, B)
    b = rng.uniform(1.0, 5.0, B)
    amp = rng.uniform(0.0, 0.5, B)
    mu = rng.uniform(0.2, 0.8, B)
    sigma = rng.uniform(0.02, 0.1, B)
    return np.stack([a, b, amp, mu, sigma], axis=1)


def eval_density(x: np.ndarray, p: np.ndarray) -> np.ndarray:
    a, b, amp, mu, sigma = p
    return x ** a * (1.0 - x) ** b + amp * np.exp(-((x - mu) ** 2) / (2.0 * sigma ** 2))


def uniform_family(B: int, N: int, seed: int = 0, dtype=np.float32):
    """B curves of N nodes: returns (x [B, N], density [B, N]) in ``dtype``."""
    grid = np.linspace(0.001, 0.999, N).astype(dtype)
    x = np.broadcast_to(grid, (B, N)).copy()
    params = density_params(B, seed)
    xd = x.astype(np.float64)
    dens = np.stack([eval_density(xd[i], params[i]) for i in range(B)]).astype(dtype)
    return x, dens


def ragged_family(n_instances: int, n_range=(128, 16384), e_range=(1024, 2 ** 20), seed: int = 0,
                  dtype=np.float32):
    """Ragged batch: N_i, E_i log-uniform.  Returns dict with flat x/density."""
    rng = np.random.default_rng(seed + 1)
    n_nodes = np.round(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), n_instances)))
    n_nodes = n_nodes.astype(np.int64)
    events = np.round(np.exp(rng.uniform(np.log(e_range[0]), np.log(e_range[1]), n_instances)))
    events = events.astype(np.int64)
    params = density_params(n_instances, seed)
    node_off = np.zeros(n_instances, dtype=np.int64)
    node_off[1:] = np.cumsum(n_nodes)[:-1]
    xs, ds = [], []
    for i in range(n_instances):
        grid = np.linspace(0.001, 0.999, int(n_nodes[i])).astype(dtype)
        xs.append(grid)
        ds.append(eval_density(grid.astype(np.float64), params[i]).astype(dtype))
    return dict(x=np.concatenate(xs), density=np.concatenate(ds), node_off=node_off,
                n_nodes=n_nodes, events=events)


def reference_1d_inputs(x: np.ndarray, density: np.ndarray):
    """Reference-layout theory outputs for the 1-D mapping of (x, density).

    Returns float64 arrays (x_bins [B,1,N], xsec_x [B,1,1,N], q_bins [B,1,2],
    xsec_q [B,1,1,2], weights [B,1,1], acceptance [B,1,1]) with weights 1 so that
    count = n_events exactly.
    """
    B, N = x.shape
    x_bins = x.astype(np.float64).reshape(B, 1, N)
    xsec_x = density.astype(np.float64).reshape(B, 1, 1, N)
    q_bins = np.broadcast_to(np.array([0.001, 0.999]), (B, 1, 2)).copy()
    xsec_q = np.ones((B, 1, 1, 2))
    weights = np.ones((B, 1, 1))
    acceptance = np.ones((B, 1, 1), dtype=bool)
    return x_bins, xsec_x, q_bins, xsec_q, weights, acceptance
