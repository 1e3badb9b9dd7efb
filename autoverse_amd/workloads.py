"""Synthetic Flow Lenia workloads at the benchmark sizes.

Every random draw uses ``numpy.random.default_rng(seed)`` (PCG64, stable across numpy
releases) in the order written below, so a (workload, seed) pair always yields the
same configuration and initial state.  Where the workload description leaves a
value open the choice is stated next to it.
"""
from __future__ import annotations

import numpy as np

from .schema import BlobSpec, FlowConfig, KernelConfig, Pattern, RingConfig, Seed, SimulationConfig


def plain_256() -> tuple[SimulationConfig, Seed]:
    """2D 256x256, 1 channel, dt 0.2, kernel_radius 13, one kernel (radius 1.0, one ring
    amplitude 1 / distance 0.5 / width 0.15, weight 1, mu 0.15, sigma 0.015), flow
    beta_a 1, n 2, distribution_size 1; GaussianBlob at (0.5, 0.5), radius 0.1, amplitude 1."""
    cfg = SimulationConfig(width=256, height=256, depth=1, channels=1, dt=0.2, kernel_radius=13,
                           kernels=[KernelConfig(1.0, [RingConfig(1.0, 0.5, 0.15)], 1.0, 0.15, 0.015, 0, 0)],
                           flow=FlowConfig(1.0, 2.0, 1.0))
    seed = Seed(Pattern("GaussianBlob", center=(0.5, 0.5), radius=0.1, amplitude=1.0, channel=0))
    return cfg, seed


def _random_kernels(rng, channels: int, pairs=None):
    """Kernels over source->target pairs (default: every pair, k = src*C + tgt), each with
    1-3 rings: amplitude U(0.2,1), distance U(0.2,0.8), width U(0.05,0.2); weight U(0.3,1),
    mu U(0.1,0.3), sigma U(0.01,0.05).  KernelConfig.radius is 1.0 (not drawn)."""
    if pairs is None:
        pairs = [(s, t) for s in range(channels) for t in range(channels)]
    ks = []
    for (s, t) in pairs:
        nr = int(rng.integers(1, 4))
        rings = [RingConfig(float(rng.uniform(0.2, 1.0)), float(rng.uniform(0.2, 0.8)),
                            float(rng.uniform(0.05, 0.2))) for _ in range(nr)]
        ks.append(KernelConfig(1.0, rings, float(rng.uniform(0.3, 1.0)), float(rng.uniform(0.1, 0.3)),
                               float(rng.uniform(0.01, 0.05)), s, t))
    return ks


def _random_blobs(rng, channels: int, per_channel: int = 8):
    """Per channel: centres U(0,1)^2, radius U(0.03,0.08), amplitude U(0.5,1)."""
    blobs = []
    for c in range(channels):
        for _ in range(per_channel):
            cx, cy = float(rng.uniform(0, 1)), float(rng.uniform(0, 1))
            blobs.append(BlobSpec((cx, cy), float(rng.uniform(0.03, 0.08)), float(rng.uniform(0.5, 1.0)), c))
    return blobs


def multispecies(size: int = 1024, channels: int = 3, seed: int = 0x5EED1) -> tuple[SimulationConfig, Seed]:
    """2D multi-species: size^2, `channels` channels, dt 0.2, kernel_radius 15, one kernel per
    source->target pair (9 at 3 channels), flow beta_a 1, n 2, distribution_size 1; 8 Gaussian
    blobs per channel."""
    rng = np.random.default_rng(seed)
    kernels = _random_kernels(rng, channels)
    blobs = _random_blobs(rng, channels)
    cfg = SimulationConfig(width=size, height=size, depth=1, channels=channels, dt=0.2, kernel_radius=15,
                           kernels=kernels, flow=FlowConfig(1.0, 2.0, 1.0))
    return cfg, Seed(Pattern("MultiBlob", blobs=blobs))


def sphere_3d(size: int = 128) -> tuple[SimulationConfig, Seed]:
    """3D size^3, 1 channel, dt 0.1, kernel_radius 10, one shell kernel (one ring amplitude 1 /
    distance 0.5 / width 0.15, weight 1.0, mu 0.15, sigma 0.017), distribution_size 1 with
    beta_a 1 and n 2 (the 2D workloads' flow values; not stated for 3D); GaussianSphere at
    the centre, radius 0.12, amplitude 1."""
    cfg = SimulationConfig(width=size, height=size, depth=size, channels=1, dt=0.1, kernel_radius=10,
                           kernels=[KernelConfig(1.0, [RingConfig(1.0, 0.5, 0.15)], 1.0, 0.15, 0.017, 0, 0)],
                           flow=FlowConfig(1.0, 2.0, 1.0))
    seed = Seed(Pattern("GaussianSphere", center=(0.5, 0.5, 0.5), radius=0.12, amplitude=1.0, channel=0))
    return cfg, seed


def population(members: int = 256, seed: int = 0x5EED3):
    """Population of independent members: side U{64,128,192,256}, channels U{1,2,3}, kernels
    for every source->target pair drawn as in `multispecies`, 8 blobs per channel."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(members):
        side = int(rng.choice([64, 128, 192, 256]))
        c = int(rng.integers(1, 4))
        kernels = _random_kernels(rng, c)
        blobs = _random_blobs(rng, c)
        cfg = SimulationConfig(width=side, height=side, depth=1, channels=c, dt=0.2, kernel_radius=15,
                               kernels=kernels, flow=FlowConfig(1.0, 2.0, 1.0))
        out.append((cfg, Seed(Pattern("MultiBlob", blobs=blobs))))
    return out


def differentiation(seed: int = 0x5EED4):
    """512^2, 2 channels, 4 kernels (every source->target pair), drawn as in `multispecies`."""
    return multispecies(size=512, channels=2, seed=seed)
