"""Population fitness evaluation on the device (hp_batch_evaluate) against the f32 restatement of
the reference's FitnessEvaluator (oracle/fitness.py, fitness.rs:26-520).

Two bars:
  * statistics: the device's evaluation against oracle/fitness.py applied to the device's own
    states at the same sample steps (a second population with the same plan follows the same
    trajectory bit for bit).  Counts and peaks are exact (the channel sum is formed as the
    reference forms it); sums are FP64 on the device and sequential f32 in the restatement, so
    they agree to f32 rounding (relative 1e-4);
  * end to end: against oracle/fitness.py over the oracle propagator's own trajectory, within
    the trajectory tolerance (the FFT's rounding differs), 2e-3 on every score.
"""
import numpy as np
import pytest

from autoverse_amd import workloads
from autoverse_amd.fitness import EvaluationConfig, FitnessConfig, FitnessMetric, WeightedMetric
from autoverse_amd.population import HipPopulation
from autoverse_amd.schema import FlowConfig, Pattern, Seed, SimulationConfig
from oracle import fitness as FZ

pytestmark = pytest.mark.gpu

ALL_METRICS = [
    (FitnessMetric("Persistence"), 1.0),
    (FitnessMetric("Compactness"), 0.5),
    (FitnessMetric("Locomotion"), 0.7),
    (FitnessMetric("Periodicity", period=8, tolerance=0.2), 0.3),
    (FitnessMetric("Complexity"), 0.4),
    (FitnessMetric("MassConcentration"), 0.2),
    (FitnessMetric("GliderScore", min_displacement=0.05), 0.6),
    (FitnessMetric("OscillatorScore", max_period=8, threshold=0.5), 0.3),
    (FitnessMetric("Stability"), 0.9),
    (FitnessMetric("Custom", name="external"), 0.1),
]


def _say(*a):
    print(*a, flush=True)


def _members():
    rng = np.random.default_rng(77)
    out = []
    for side, c in ((64, 1), (96, 2), (128, 3), (64, 2)):
        ks = workloads._random_kernels(rng, c)
        blobs = workloads._random_blobs(rng, c)
        cfg = SimulationConfig(width=side, height=side, depth=1, channels=c, dt=0.2, kernel_radius=15,
                               kernels=ks, flow=FlowConfig(1.0, 2.0, 1.0))
        out.append((cfg, Seed(Pattern("MultiBlob", blobs=blobs))))
    return out


def _fc():
    return FitnessConfig([WeightedMetric(m, w) for m, w in ALL_METRICS], normalize=True)


def _dicts():
    return [(m.to_dict(), w) for m, w in ALL_METRICS]


def _restated_on_device_states(mem, ev):
    """oracle/fitness.py's trajectory and metrics over the device's own states."""
    out = []
    with HipPopulation([c for c, _ in mem]) as pop:
        pop.set_seeds([s for _, s in mem])
        states = lambda: [pop.get_state(i).channels for i in range(len(mem))]  # noqa: E731
        trajs = [FZ.Trajectory.new(ch, c.width, c.height) for ch, (c, _) in zip(states(), mem)]
        pop.advance(ev.warmup_steps)
        for t, ch in zip(trajs, states()):
            t.record_sample(ch, 0)
        done = 0
        for step in range(ev.sample_interval, ev.steps + 1, ev.sample_interval):
            pop.advance(step - done)
            done = step
            for t, ch in zip(trajs, states()):
                t.record_sample(ch, step)
        pop.advance(ev.steps - done)
        finals = states()
    for (cfg, _), t, fin in zip(mem, trajs, finals):
        scores = [FZ.compute_metric(m, t, fin, cfg.width, cfg.height) for m, _ in _dicts()]
        out.append((scores, t.behavior(), t))
    return out


def test_evaluate_matches_restated_fitness_on_device_states():
    mem = _members()
    ev = EvaluationConfig(steps=24, warmup_steps=5, sample_interval=4)
    _say("[fitness statistics] 4 members, 24 steps after 5, sample every 4, all metrics")
    with HipPopulation([c for c, _ in mem]) as pop:
        pop.set_seeds([s for _, s in mem])
        got = pop.evaluate(_fc(), ev)
    want = _restated_on_device_states(mem, ev)
    for i, ((fit, res, bs), (ws, wb, wt)) in enumerate(zip(got, want)):
        gs = np.array([r.score for r in res], np.float64)
        ws = np.array(ws, np.float64)
        _say(f"  member {i}: max |score - restated| = {np.abs(gs - ws).max():.3e}")
        np.testing.assert_allclose(gs, ws, rtol=1e-4, atol=2e-6, err_msg=f"member {i}")
        assert bs.active_cells == wb["active_cells"]
        assert bs.max_activation == wb["max_activation"]
        assert bs.active_over_time == wt.active_cell_samples
        assert bs.max_over_time == [float(x) for x in wt.max_samples]
        for k in ("final_mass", "initial_mass", "total_displacement", "final_radius"):
            assert bs.__dict__[k] == pytest.approx(float(wb[k]), rel=1e-4, abs=1e-5), (i, k)
        np.testing.assert_allclose(np.array(bs.center_of_mass_trajectory), np.array(wt.center_samples, np.float64),
                                   rtol=1e-5, atol=1e-4)
        np.testing.assert_allclose(bs.radius_over_time, np.array(wt.radius_samples, np.float64), rtol=1e-4, atol=1e-5)
        # the combination (normalize clamps, weights over their f32 sum)
        wsum = sum(np.float32(w) for _, w in ALL_METRICS)
        want_fit = sum(min(max(s, 0.0), 1.0) * w for s, (_, w) in zip(ws, ALL_METRICS)) / wsum
        assert fit == pytest.approx(want_fit, rel=1e-4, abs=1e-6)


def test_evaluate_matches_oracle_trajectory(oracle):
    mem = _members()[:3]
    ev = EvaluationConfig(steps=20, warmup_steps=3, sample_interval=5)
    _say("[fitness end to end] 3 members vs the oracle propagator")
    with HipPopulation([c for c, _ in mem]) as pop:
        pop.set_seeds([s for _, s in mem])
        got = pop.evaluate(_fc(), ev)
    for i, ((cfg, sd), (fit, res, bs)) in enumerate(zip(mem, got)):
        wc, wsc, wb, _ = FZ.evaluate(cfg, sd, _dicts(), steps=ev.steps, warmup=ev.warmup_steps,
                                     sample_interval=ev.sample_interval, oracle=oracle)
        gs = np.array([r.score for r in res], np.float64)
        err = np.abs(gs - np.array(wsc, np.float64)).max()
        _say(f"  member {i}: max |score - oracle| = {err:.3e}, fitness {fit:.6f} vs {float(wc):.6f}")
        assert err <= 2e-3
        assert abs(fit - float(wc)) <= 2e-3
        assert bs.initial_mass == pytest.approx(float(wb["initial_mass"]), rel=1e-5)


def test_evaluate_leaves_final_state_and_counts_samples():
    """The members end in the state of warmup + steps; a sample interval of 0 counts as 1."""
    mem = _members()[:2]
    with HipPopulation([c for c, _ in mem]) as pop, HipPopulation([c for c, _ in mem]) as ref:
        pop.set_seeds([s for _, s in mem])
        ref.set_seeds([s for _, s in mem])
        out = pop.evaluate(FitnessConfig(), EvaluationConfig(steps=7, warmup_steps=2, sample_interval=0))
        ref.advance(9)
        for i in range(2):
            a, b = pop.get_state(i), ref.get_state(i)
            for c in range(len(a.channels)):
                np.testing.assert_array_equal(a.channels[c], b.channels[c])
            assert a.step == 9
        assert len(out[0][2].radius_over_time) == 8
