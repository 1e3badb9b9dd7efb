"""Pins of the fitness restatement (oracle/fitness.py) -- CPU only.

Two kinds of check: the reference's own fitness tests (src/compute/evolution/fitness.rs:534-946)
re-asserted with the same inputs and bars, and float64 numpy closed forms of every statistic
the metrics are built from (centre of mass, radius, similarity, variance and gradient means,
peak ratio, the metric formulas), so that a dropped term or a wrong axis fails one of them.
"""
import math

import numpy as np
import pytest

from autoverse_amd.schema import FlowConfig, KernelConfig, Pattern, RingConfig, Seed, SimulationConfig
from oracle import fitness as FZ

F = np.float32


def _traj(w, h, centers, radii, masses):
    """create_test_trajectory (fitness.rs:574-606)."""
    t = FZ.Trajectory(w, h)
    t.initial_mass = F(masses[0] if masses else 1.0)
    t.initial_center = centers[0] if centers else (F(w / 2), F(h / 2))
    t.initial_radius = F(radii[0] if radii else 5.0)
    t.initial_snapshot = np.full(w * h, F(t.initial_mass / F(w * h)), F)
    t.center_samples = [(F(x), F(y)) for x, y in centers]
    t.radius_samples = [F(r) for r in radii]
    t.mass_samples = [F(m) for m in masses]
    t.max_samples = [F(1.0)] * 10
    t.active_cell_samples = [100] * 10
    return t


def _default_64():
    """test_config (fitness.rs:567-572): SimulationConfig::default() at 64 x 64."""
    return SimulationConfig(width=64, height=64)


def _blob(r, amp=1.0):
    return Seed(Pattern("GaussianBlob", center=(0.5, 0.5), radius=r, amplitude=amp, channel=0))


def _state(oracle, seed, cfg):
    return oracle.seed_generate(seed, cfg.width, cfg.height, 1, cfg.channels)


# ------------------------------------------------------------------ reference tests, restated
def test_fitness_evaluator(oracle):
    """fitness.rs:563-578: default metrics (Persistence 1.0, Compactness 0.5), 10 steps, sample
    interval 5 on the 64 x 64 default configuration."""
    combined, scores, beh, t = FZ.evaluate(_default_64(), _blob(0.1), [({"type": "Persistence"}, 1.0),
                                                                       ({"type": "Compactness"}, 0.5)],
                                           steps=10, sample_interval=5, oracle=oracle)
    assert combined >= 0.0 and len(scores) == 2 and beh["initial_mass"] > 0
    assert [s for s, _ in t.state_snapshots] == [0, 5, 10]


def test_center_of_mass_known_answer():
    cx, cy = FZ.center_of_mass(np.array([0.0, 0.0, 1.0, 0.0], F), 2, 2)
    assert abs(cx) < 1e-6 and abs(cy - 1.0) < 1e-6


def test_state_similarity_known_answers():
    a, b, c = np.array([1, 0, 0, 1], F), np.array([1, 0, 0, 1], F), np.array([0, 1, 1, 0], F)
    assert abs(FZ.similarity(a, b) - 1.0) < 1e-6 and abs(FZ.similarity(a, c)) < 1e-6


def test_locomotion_moving_vs_static():
    st = _traj(64, 64, [(32, 32)] * 3, [5.0] * 3, [1.0] * 3)
    mv = _traj(64, 64, [(32, 32), (40, 40), (50, 50)], [5.0] * 3, [1.0] * 3)
    assert FZ.locomotion(mv) > FZ.locomotion(st) and FZ.locomotion(st) < 0.01


def test_compactness_small_vs_diffuse(oracle):
    cfg = _default_64()
    out = []
    for r in (0.05, 0.3):
        s = _state(oracle, _blob(r), cfg)
        t = FZ.Trajectory.new(s, 64, 64)
        t.record_sample(s, 0)
        out.append(FZ.compactness(t, 64, 64))
    assert out[0] > out[1]


def test_persistence_surviving_vs_dissipating():
    a = _traj(64, 64, [(32, 32)] * 5, [5.0] * 5, [1.0, 0.98, 0.96, 0.95, 0.95])
    b = _traj(64, 64, [(32, 32)] * 5, [5, 10, 20, 40, 60], [1.0, 0.8, 0.5, 0.2, 0.05])
    assert FZ.persistence(a) > FZ.persistence(b)


def test_complexity_structured_vs_uniform(oracle):
    cfg = _default_64()
    s = FZ.complexity(_state(oracle, _blob(0.1), cfg), 64, 64)
    u = FZ.complexity(_state(oracle, _blob(0.5, 0.1), cfg), 64, 64)
    assert s > u or s > 0.01


def test_mass_concentration_peaked_vs_flat(oracle):
    cfg = _default_64()
    assert FZ.mass_concentration(_state(oracle, _blob(0.03), cfg)) > FZ.mass_concentration(
        _state(oracle, _blob(0.4), cfg))


def test_glider_score_requires_movement(oracle):
    cfg = _default_64()
    g = _traj(64, 64, [(20, 20), (25, 25), (30, 30), (35, 35), (40, 40)], [5.0] * 5,
              [1.0, 0.98, 0.96, 0.95, 0.94])
    s = _traj(64, 64, [(32, 32)] * 5, [5.0] * 5, [1.0, 0.98, 0.96, 0.95, 0.94])
    gs, ss = FZ.glider_score(g, 64, 64, 10.0), FZ.glider_score(s, 64, 64, 10.0)
    assert gs > ss and ss < 0.01


def test_periodicity_with_repeating_state():
    t = _traj(8, 8, [(4, 4)] * 10, [2.0] * 10, [1.0] * 10)
    a = np.array([1.0 if i % 2 == 0 else 0.0 for i in range(64)], F)
    b = np.array([1.0 if i % 2 == 1 else 0.0 for i in range(64)], F)
    t.initial_snapshot = a.copy()
    t.state_snapshots = [(0, a), (5, b), (10, a), (15, b), (20, a)]
    assert FZ.periodicity(t, 10, 0.1) > 0.5


def test_oscillator_score_finds_period():
    t = _traj(8, 8, [(4, 4)] * 10, [2.0] * 10, [1.0] * 10)
    a = np.array([1.0 if i < 32 else 0.0 for i in range(64)], F)
    b = np.array([1.0 if i >= 32 else 0.0 for i in range(64)], F)
    t.state_snapshots = [(0, a), (1, b), (2, a), (3, b), (4, a)]
    assert FZ.oscillator_score(t, 10, 0.8) > 0.3


def test_stability_of_a_fresh_state(oracle):
    cfg = _default_64()
    s = _state(oracle, _blob(0.1), cfg)
    t = FZ.Trajectory.new(s, 64, 64)
    t.record_sample(s, 0)
    assert FZ.stability(t, s) >= 0.0


def test_radius_known_answer():
    g = np.zeros(100, F)
    for (x, y) in ((3, 3), (7, 3), (3, 7), (7, 7)):
        g[y * 10 + x] = 1.0
    cx, cy = FZ.center_of_mass(g, 10, 10)
    assert abs(cx - 5) < 0.1 and abs(cy - 5) < 0.1
    assert abs(FZ.radius(g, cx, cy, 10, 10) - 2.83) < 0.1


# ------------------------------------------------------------------ float64 closed forms
def _grid(seed, w=37, h=29, neg=True):
    rng = np.random.default_rng(seed)
    g = rng.uniform(-0.2 if neg else 0.0, 1.0, (h, w)).astype(F)
    g[rng.uniform(size=(h, w)) < 0.3] = 0.0
    return g


def test_statistics_match_float64_closed_forms():
    for seed in range(4):
        g = _grid(seed)
        h, w = g.shape
        d = g.astype(np.float64)
        pos = np.where(d > 0, d, 0.0)
        ys, xs = np.mgrid[0:h, 0:w]
        m = pos.sum()
        cx, cy = (xs * pos).sum() / m, (ys * pos).sum() / m
        gx, gy = FZ.center_of_mass(g.reshape(-1), w, h)
        assert gx == pytest.approx(cx, rel=1e-5) and gy == pytest.approx(cy, rel=1e-5)
        r = math.sqrt((pos * ((xs - float(gx)) ** 2 + (ys - float(gy)) ** 2)).sum() / m)
        assert FZ.radius(g.reshape(-1), gx, gy, w, h) == pytest.approx(r, rel=1e-5)
        # similarity against a shifted copy, complexity and concentration
        g2 = np.roll(g, 3, 1)
        e = d.ravel()
        f = g2.astype(np.float64).ravel()
        assert FZ.similarity(g.ravel(), g2.ravel()) == pytest.approx(
            e @ f / math.sqrt((e @ e) * (f @ f)), rel=1e-5)
        mean = d.mean()
        grad = np.abs(np.diff(d, axis=1)).sum() + np.abs(np.diff(d, axis=0)).sum()
        want = min((d.std() + grad / (2 * d.size)) / 2, 1.0)
        assert FZ.complexity([g.reshape(-1)], w, h) == pytest.approx(want, rel=1e-5)
        assert FZ.mass_concentration([g.reshape(-1)]) == pytest.approx(min(d.max() / mean / 100, 1.0), rel=1e-5)
        # stability: negative fraction and mass error, no extreme values here
        t = FZ.Trajectory.new([g.reshape(-1)], w, h)
        t.record_sample([(g * F(1.01)).reshape(-1)], 0)
        want = 1.0 - (d < 0).mean() - 0.01
        assert FZ.stability(t, [g.reshape(-1)]) == pytest.approx(max(0.0, want), abs=2e-5)


def test_channel_sum_and_masses_over_channels():
    a, b = _grid(10, neg=False), _grid(11, neg=False)
    s = FZ.channel_sum([a.ravel(), b.ravel()])
    np.testing.assert_array_equal(s, (a + b).ravel())
    assert FZ.total_mass([a.ravel(), b.ravel()]) == pytest.approx(
        a.astype(np.float64).sum() + b.astype(np.float64).sum(), rel=1e-6)


def test_metric_formulas_closed_form():
    """Persistence, locomotion, compactness and glider from their definitions in float64 on a
    hand-built trajectory (a dropped factor or a swapped ratio fails)."""
    t = _traj(64, 48, [(10, 12), (20, 18), (31, 30)], [4.0, 5.0, 6.0], [2.0, 1.9, 1.8])
    t.max_samples = [F(0.5), F(0.7), F(0.9)]
    t.active_cell_samples = [40, 50, 60]
    conc = 0.9 / (1.8 / 60)
    assert FZ.persistence(t) == pytest.approx(min(max(1.8 / 2.0 * min(conc, 10) / 10, 0), 1), rel=1e-6)
    disp = math.hypot(21, 18)
    assert FZ.locomotion(t) == pytest.approx(disp / math.hypot(64, 48), rel=1e-6)
    assert FZ.compactness(t, 64, 48) == pytest.approx(1 - 6.0 / 24.0, rel=1e-6)
    r = np.array([4.0, 5.0, 6.0])
    cons = 1 - r.std() / r.mean()
    want = (disp / math.hypot(64, 48)) * (1 - 6.0 / 24.0) * cons * FZ.persistence(t)
    assert FZ.glider_score(t, 64, 48, 5.0) == pytest.approx(want, rel=1e-5)
    assert FZ.glider_score(t, 64, 48, 100.0) == 0.0


def test_evaluate_sampling_schedule(oracle):
    """Samples at 0 (after the warmup) and every sample_interval steps; the state after the last
    step is the one the state metrics see (steps not a multiple of the interval)."""
    cfg = SimulationConfig(width=32, height=32, channels=1, dt=0.2, kernel_radius=6,
                           kernels=[KernelConfig(1.0, [RingConfig(1.0, 0.5, 0.15)], 1.0, 0.15, 0.015, 0, 0)],
                           flow=FlowConfig(1.0, 2.0, 1.0))
    metrics = [({"type": "Stability"}, 1.0), ({"type": "Complexity"}, 1.0)]
    _, scores, beh, t = FZ.evaluate(cfg, _blob(0.2), metrics, steps=7, warmup=3, sample_interval=3,
                                    oracle=oracle)
    assert [s for s, _ in t.state_snapshots] == [0, 3, 6]
    st = oracle.seed_generate(_blob(0.2), 32, 32, 1, 1)
    fin, _, _ = oracle.Propagator(cfg, np.float32).run(st, 10)
    assert scores[1] == FZ.complexity(fin, 32, 32)
    assert beh["initial_mass"] == FZ.total_mass(st)
