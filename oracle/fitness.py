"""Fitness evaluation of the reference, restated in float32 (TEST INFRASTRUCTURE ONLY).

Only the tests import this module, as the checker of the device fitness evaluation; the
product package never does.  Every function follows src/compute/evolution/fitness.rs of the
reference statement by statement in f32: sums run left to right over the cells in raster order
(np.add.accumulate in float32 reproduces Rust's sequential iterator sums), products and
quotients round in f32, f32::max ignores NaN.  The trajectory is recorded from the f32 oracle
propagator (oracle.Propagator), as FitnessEvaluator::evaluate records it from CpuPropagator.

    EvaluationTrajectory         fitness.rs:93-191
    compute_center_of_mass       fitness.rs:196-219
    compute_radius               fitness.rs:222-244
    compute_metric and metrics   fitness.rs:247-520
    FitnessEvaluator::evaluate   fitness.rs:26-86
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

F = np.float32


def seqsum(a) -> np.float32:
    """Rust's f32 iterator sum: left to right from zero."""
    a = np.asarray(a, dtype=F).reshape(-1)
    if a.size == 0:
        return F(0.0)
    return F(np.add.accumulate(a, dtype=F)[-1])


def fmax_fold(a) -> np.float32:
    """fold(0.0, f32::max): NaN operands are ignored."""
    a = np.asarray(a, dtype=F).reshape(-1)
    m = F(0.0)
    if a.size:
        v = np.nanmax(a) if not np.all(np.isnan(a)) else F(0.0)
        m = F(max(m, v))
    return m


def channel_sum(channels) -> np.ndarray:
    """SimulationState::channel_sum (propagator.rs:109-120): 0 + A_0 + A_1 + ... per cell."""
    s = np.zeros(np.asarray(channels[0]).size, dtype=F)
    for c in channels:
        s = (s + np.asarray(c, dtype=F).reshape(-1)).astype(F)
    return s


def total_mass(channels) -> np.float32:
    """total_mass_all_channels (reintegration.rs:221-228)."""
    return seqsum([seqsum(c) for c in channels])


def center_of_mass(grid, w: int, h: int) -> Tuple[np.float32, np.float32]:
    """fitness.rs:196-219: mass-weighted mean over cells with m > 0; the grid centre below 1e-6."""
    g = np.asarray(grid, dtype=F).reshape(-1)
    idx = np.nonzero(g > F(0.0))[0]
    m = g[idx]
    x = (idx % w).astype(F)
    y = (idx // w).astype(F)
    total = seqsum(m)
    cx = seqsum(x * m)
    cy = seqsum(y * m)
    if total > F(1e-6):
        return F(cx / total), F(cy / total)
    return F(F(w) / F(2.0)), F(F(h) / F(2.0))


def radius(grid, cx, cy, w: int, h: int) -> np.float32:
    """fitness.rs:222-244: sqrt of the second moment about (cx, cy) over cells with m > 0."""
    g = np.asarray(grid, dtype=F).reshape(-1)
    idx = np.nonzero(g > F(0.0))[0]
    m = g[idx]
    dx = (idx % w).astype(F) - F(cx)
    dy = (idx // w).astype(F) - F(cy)
    total = seqsum(m)
    moment = seqsum(m * (dx * dx + dy * dy))
    if total > F(1e-6):
        return F(math.sqrt(F(moment / total)))
    return F(0.0)


@dataclass
class Trajectory:
    """EvaluationTrajectory (fitness.rs:93-191)."""
    width: int
    height: int
    initial_mass: np.float32 = F(0.0)
    initial_center: Tuple[np.float32, np.float32] = (F(0.0), F(0.0))
    initial_radius: np.float32 = F(0.0)
    initial_snapshot: Optional[np.ndarray] = None
    center_samples: List[Tuple[np.float32, np.float32]] = field(default_factory=list)
    radius_samples: List[np.float32] = field(default_factory=list)
    mass_samples: List[np.float32] = field(default_factory=list)
    max_samples: List[np.float32] = field(default_factory=list)
    active_cell_samples: List[int] = field(default_factory=list)
    state_snapshots: List[Tuple[int, np.ndarray]] = field(default_factory=list)

    @classmethod
    def new(cls, channels, w: int, h: int) -> "Trajectory":
        snap = channel_sum(channels)
        cx, cy = center_of_mass(snap, w, h)
        return cls(w, h, total_mass(channels), (cx, cy), radius(snap, cx, cy, w, h), snap)

    def record_sample(self, channels, step: int) -> None:
        s = channel_sum(channels)
        cx, cy = center_of_mass(s, self.width, self.height)
        self.center_samples.append((cx, cy))
        self.radius_samples.append(radius(s, cx, cy, self.width, self.height))
        self.mass_samples.append(seqsum(s))
        self.max_samples.append(fmax_fold(s))
        self.active_cell_samples.append(int((s > F(1e-6)).sum()))
        self.state_snapshots.append((step, s))

    def displacement(self) -> np.float32:
        if len(self.center_samples) < 2:
            return F(0.0)
        (sx, sy), (ex, ey) = self.center_samples[0], self.center_samples[-1]
        ddx, ddy = F(ex - sx), F(ey - sy)
        return F(math.sqrt(F(F(ddx * ddx) + F(ddy * ddy))))

    def behavior(self) -> dict:
        """to_behavior_stats (fitness.rs:163-191)."""
        return {
            "final_mass": self.mass_samples[-1] if self.mass_samples else F(0.0),
            "initial_mass": self.initial_mass,
            "center_of_mass_trajectory": list(self.center_samples),
            "total_displacement": self.displacement(),
            "radius_over_time": list(self.radius_samples),
            "final_radius": self.radius_samples[-1] if self.radius_samples else F(0.0),
            "active_cells": self.active_cell_samples[-1] if self.active_cell_samples else 0,
            "max_activation": fmax_fold(self.max_samples) if self.max_samples else F(0.0),
        }


def similarity(a, b) -> np.float32:
    """compute_state_similarity (fitness.rs:353-366)."""
    a = np.asarray(a, dtype=F)
    b = np.asarray(b, dtype=F)
    if a.size != b.size:
        return F(0.0)
    dot = seqsum(a * b)
    na = F(math.sqrt(seqsum(a * a)))
    nb = F(math.sqrt(seqsum(b * b)))
    if na > F(1e-6) and nb > F(1e-6):
        return F(dot / F(na * nb))
    return F(0.0)


def clamp01(v) -> np.float32:
    v = F(v)
    if v != v:
        return v
    return F(min(max(v, F(0.0)), F(1.0)))


def persistence(t: Trajectory) -> np.float32:
    """fitness.rs:271-291."""
    if t.initial_mass < F(1e-6):
        return F(0.0)
    final_mass = t.mass_samples[-1] if t.mass_samples else F(0.0)
    mass_ratio = F(final_mass / t.initial_mass)
    final_max = t.max_samples[-1] if t.max_samples else F(0.0)
    if final_mass > F(1e-6):
        active = F(t.active_cell_samples[-1] if t.active_cell_samples else 1)
        with np.errstate(divide="ignore", invalid="ignore"):
            concentration = F(final_max / F(final_mass / active))
    else:
        concentration = F(0.0)
    return clamp01(F(F(mass_ratio * F(min(concentration, F(10.0)))) / F(10.0)))


def compactness(t: Trajectory, w: int, h: int) -> np.float32:
    """fitness.rs:294-306."""
    final_radius = t.radius_samples[-1] if t.radius_samples else F(np.finfo(np.float32).max)
    max_radius = F(F(min(w, h)) / F(2.0))
    return clamp01(F(1.0) - F(final_radius / max_radius))


def locomotion(t: Trajectory) -> np.float32:
    """fitness.rs:309-326."""
    if len(t.center_samples) < 2:
        return F(0.0)
    maxd = F(math.sqrt(F(t.width ** 2 + t.height ** 2)))
    return clamp01(F(t.displacement() / maxd))


def periodicity(t: Trajectory, period: int, tolerance: float) -> np.float32:
    """fitness.rs:329-350."""
    for step, snap in t.state_snapshots:
        if step == period:
            sim = similarity(t.initial_snapshot, snap)
            if sim > F(F(1.0) - F(tolerance)):
                return sim
    best = F(0.0)
    for step, snap in t.state_snapshots:
        if step > 0:
            best = F(max(best, similarity(t.initial_snapshot, snap)))
    return best


def complexity(channels, w: int, h: int) -> np.float32:
    """fitness.rs:369-398."""
    s = channel_sum(channels)
    n = s.size
    mean = F(seqsum(s) / F(n))
    if mean < F(1e-6):
        return F(0.0)
    d = (s - mean).astype(F)
    variance = F(seqsum(d * d) / F(n))
    g = s.reshape(h, w)
    terms = np.zeros((h, w, 2), dtype=F)  # per cell: the x term, then the y term (loop order)
    present = np.zeros((h, w, 2), dtype=bool)
    terms[:, 1:, 0] = np.abs(g[:, 1:] - g[:, :-1])
    present[:, 1:, 0] = True
    terms[1:, :, 1] = np.abs(g[1:, :] - g[:-1, :])
    present[1:, :, 1] = True
    gsum = seqsum(terms[present])
    gmean = F(gsum / F(2 * n))
    return F(min(F(F(F(math.sqrt(variance)) + gmean) / F(2.0)), F(1.0)))


def mass_concentration(channels) -> np.float32:
    """fitness.rs:401-413."""
    s = channel_sum(channels)
    mean = F(seqsum(s) / F(s.size))
    mx = fmax_fold(s)
    if mean < F(1e-6):
        return F(0.0)
    return F(min(F(F(mx / mean) / F(100.0)), F(1.0)))


def glider_score(t: Trajectory, w: int, h: int, min_displacement: float) -> np.float32:
    """fitness.rs:416-453."""
    loco = locomotion(t)
    comp = compactness(t, w, h)
    if t.displacement() < F(min_displacement):
        return F(0.0)
    if len(t.radius_samples) >= 2:
        r = np.asarray(t.radius_samples, dtype=F)
        mean_r = F(seqsum(r) / F(r.size))
        dr = (r - mean_r).astype(F)
        var = F(seqsum(dr * dr) / F(r.size))
        cv = F(F(math.sqrt(var)) / F(max(mean_r, F(1e-6))))
        consistency = clamp01(F(1.0) - cv)
    else:
        consistency = F(0.5)
    return F(F(F(loco * comp) * consistency) * persistence(t))


def oscillator_score(t: Trajectory, max_period: int, threshold: float) -> np.float32:
    """fitness.rs:456-483."""
    best = F(0.0)
    snaps = t.state_snapshots
    for period in range(1, max_period + 1):
        for i, (si, ai) in enumerate(snaps):
            for sj, aj in snaps[i + 1:]:
                if sj - si == period:
                    sim = similarity(ai, aj)
                    if sim > F(threshold):
                        ps = F(sim * F(F(1.0) - F(F(F(period) / F(max_period)) * F(0.5))))
                        best = F(max(best, ps))
    return best


def stability(t: Trajectory, channels) -> np.float32:
    """fitness.rs:486-511."""
    s = channel_sum(channels)
    neg = F(min(F(F(int((s < F(0.0)).sum())) / F(s.size)), F(1.0)))
    mx = fmax_fold(s)
    extreme = F(F(mx - F(10.0)) / mx) if mx > F(10.0) else F(0.0)
    final_mass = t.mass_samples[-1] if t.mass_samples else F(0.0)
    if t.initial_mass > F(1e-6):
        merr = F(abs(F(F(final_mass - t.initial_mass) / t.initial_mass)))
    else:
        merr = F(0.0)
    return clamp01(F(F(F(F(1.0) - neg) - extreme) - merr))


def compute_metric(metric: dict, t: Trajectory, channels, w: int, h: int) -> np.float32:
    """compute_metric (fitness.rs:247-268); `metric` is the serde form: {"type": ..., params}."""
    k = metric["type"]
    if k == "Persistence":
        return persistence(t)
    if k == "Compactness":
        return compactness(t, w, h)
    if k == "Locomotion":
        return locomotion(t)
    if k == "Periodicity":
        return periodicity(t, int(metric["period"]), metric["tolerance"])
    if k == "Complexity":
        return complexity(channels, w, h)
    if k == "MassConcentration":
        return mass_concentration(channels)
    if k == "GliderScore":
        return glider_score(t, w, h, metric["min_displacement"])
    if k == "OscillatorScore":
        return oscillator_score(t, int(metric["max_period"]), metric["threshold"])
    if k == "Stability":
        return stability(t, channels)
    return F(0.0)  # Custom: scored outside the evaluator


def evaluate(cfg, seed, metrics, steps: int, warmup: int = 0, sample_interval: int = 10,
             normalize: bool = True, oracle=None):
    """FitnessEvaluator::evaluate (fitness.rs:26-86) over the f32 oracle propagator.  `metrics`
    is a list of (metric dict, weight).  Returns (combined, [scores], behavior dict, trajectory)."""
    if oracle is None:
        from oracle import oracle as oracle  # noqa: PLW0127
    w, h = cfg.width, cfg.height
    st = oracle.seed_generate(seed, w, h, 1, cfg.channels)
    p = oracle.Propagator(cfg, np.float32)
    t = Trajectory.new(st, w, h)
    if warmup:
        st, _, _ = p.run(st, warmup)
    t.record_sample(st, 0)
    k = max(1, int(sample_interval))
    done = 0
    for step in range(k, steps + 1, k):
        st, _, _ = p.run(st, step - done)
        done = step
        t.record_sample(st, step)
    if steps > done:
        st, _, _ = p.run(st, steps - done)
    p.close()
    scores = [compute_metric(m, t, st, w, h) for m, _ in metrics]
    used = [clamp01(s) for s in scores] if normalize else scores
    total_w = seqsum([wt for _, wt in metrics])
    combined = F(seqsum([F(s * F(wt)) for s, (_, wt) in zip(used, metrics)]) / F(max(total_w, F(1e-6))))
    return combined, scores, t.behavior(), t
