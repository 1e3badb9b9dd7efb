"""Populations (HipPopulation, the batched step) against lone propagators and the oracle.

A member of a population must follow exactly the trajectory a lone HipPropagator computes
for it with the same FFT plan (row-major LDS FFTs, one spectral item per source): the batched
kernels run the same per-item code.  Against the oracle the FFT tolerance of the plain path
applies.  Centroids are the FP64 centre of mass of the channel sum over cells with m > 0
(src/compute/evolution/fitness.rs:196-219).
"""
import numpy as np
import pytest

import autoverse_amd as av
from autoverse_amd import _abi, workloads
from autoverse_amd.population import HipPopulation
from autoverse_amd.schema import FlowConfig, Pattern, Seed, SimulationConfig

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _members(sides, chans, seed):
    rng = np.random.default_rng(seed)
    out = []
    for side, c in zip(sides, chans):
        ks = workloads._random_kernels(rng, c)
        blobs = workloads._random_blobs(rng, c)
        cfg = SimulationConfig(width=side, height=side, depth=1, channels=c, dt=0.2, kernel_radius=15,
                               kernels=ks, flow=FlowConfig(1.0, 2.0, 1.0))
        out.append((cfg, Seed(Pattern("MultiBlob", blobs=blobs))))
    return out


def _lone(cfg, seed, steps, monkeypatch, plan, stats=False):
    """The lone trajectory under the population's plan (the process-wide plan cache may hold
    a tuned plan for this shape from an earlier test, with another FFT rounding)."""
    monkeypatch.setenv("HP_FFT_PATH", "generic")
    monkeypatch.setenv("HP_SPECTRAL_GROUP", "1")
    with av.HipPropagator(cfg) as hp:
        hp.set_plan(plan)
        st = av.SimulationState.from_seed(seed, cfg)
        hp.run(st, steps)
        _, mass = hp.mass_f64()
        gs = hp.stats()
    return (st, mass, gs) if stats else (st, mass)


def _com64(state):
    m = state.channel_sum().astype(np.float64).reshape(state.height, state.width)
    m = np.where(m > 0, m, 0.0)
    tot = m.sum()
    ys, xs = np.mgrid[0:state.height, 0:state.width]
    if tot <= 1e-6:
        return state.width / 2.0, state.height / 2.0, tot
    return float((xs * m).sum() / tot), float((ys * m).sum() / tot), float(tot)


def test_population_members_match_lone_propagators_bitwise(monkeypatch):
    sides = [64, 96, 128, 64, 192, 256]
    chans = [1, 2, 3, 2, 1, 3]
    mem = _members(sides, chans, seed=21)
    _say(f"[population_bitwise] sides={sides} channels={chans} steps=15")
    with HipPopulation([c for c, _ in mem]) as pop:
        pop.set_seeds([s for _, s in mem])
        pop.advance(9)
        pop.advance(6)
        got = [pop.get_state(i) for i in range(len(mem))]
        masses = pop.mass_f64()
        r, g = pop.stats()
        plan = pop.plan()
    assert g == 0
    for i, (cfg, sd) in enumerate(mem):
        want, wmass = _lone(cfg, sd, 15, monkeypatch, plan)
        assert got[i].step == 15 and got[i].time == want.time
        for c in range(cfg.channels):
            np.testing.assert_array_equal(got[i].channels[c], want.channels[c], err_msg=f"member {i} ch {c}")
        assert masses[i] == wmass


def test_population_vs_oracle(oracle):
    mem = _members([64, 128, 96], [3, 1, 2], seed=5)
    steps = 40
    _say(f"[population_oracle] steps={steps}")
    with HipPopulation([c for c, _ in mem]) as pop:
        pop.set_seeds([s for _, s in mem])
        pop.advance(steps)
        got = [pop.get_state(i) for i in range(len(mem))]
    for i, (cfg, sd) in enumerate(mem):
        st = oracle.seed_generate(sd, cfg.width, cfg.height, 1, cfg.channels)
        p = oracle.Propagator(cfg, np.float32)
        want, _, _ = p.run(st, steps)
        err = max(float(np.abs(got[i].channels[c] - want[c].reshape(-1)).max()) for c in range(cfg.channels))
        _say(f"  member {i}: max|gpu-oracle| = {err:.3e}")
        assert err <= 1e-4


def test_population_centroids_and_fp64_mass():
    mem = _members([64, 128, 192, 96], [2, 3, 1, 1], seed=9)
    _say("[population_centroids]")
    with HipPopulation([c for c, _ in mem]) as pop:
        pop.set_seeds([s for _, s in mem])
        pop.advance(12)
        xy, mass = pop.centroids()
        tot = pop.mass_f64()
        states = [pop.get_state(i) for i in range(len(mem))]
    for i, st in enumerate(states):
        cx, cy, m = _com64(st)
        assert abs(xy[i, 0] - cx) <= 1e-9 * max(1.0, cx)
        assert abs(xy[i, 1] - cy) <= 1e-9 * max(1.0, cy)
        assert abs(mass[i] - m) <= 1e-10 * m
        allv = np.concatenate(st.channels).astype(np.float64)
        assert abs(tot[i] - allv.sum()) <= 1e-10 * allv.sum()


def test_population_empty_member_centroid_is_grid_centre():
    mem = _members([64, 96], [1, 1], seed=2)
    with HipPopulation([c for c, _ in mem]) as pop:
        pop.set_state(0, av.SimulationState.from_seed(mem[0][1], mem[0][0]))
        pop.set_state(1, av.SimulationState([np.zeros(96 * 96, np.float32)], 96, 96))
        pop.advance(2)
        xy, mass = pop.centroids()
    assert mass[1] == 0.0
    assert tuple(xy[1]) == (48.0, 48.0)


def test_population_general_path_replay(monkeypatch):
    """One member with a large dt leaves the fused radius: the batch replays with the general
    gather and still reproduces every member's lone trajectory bit for bit."""
    mem = _members([64, 96, 64], [2, 1, 1], seed=3)
    mem[1][0].dt = 12.0
    _say("[population_general] member 1 dt=12.0, 4 steps")
    with HipPopulation([c for c, _ in mem]) as pop:
        pop.set_seeds([s for _, s in mem])
        pop.advance(4)
        got = [pop.get_state(i) for i in range(len(mem))]
        r, g = pop.stats()
        plan = pop.plan()
    _, _, (lr, lg) = _lone(mem[1][0], mem[1][1], 4, monkeypatch, plan, stats=True)
    _say(f"  batch: radius {r}, general steps {g}; lone member 1: radius {lr}, general steps {lg}")
    assert lg >= 1, "the case must leave the fused radius"
    assert g >= 1 and r == lr
    for i, (cfg, sd) in enumerate(mem):
        want, _ = _lone(cfg, sd, 4, monkeypatch, plan)
        for c in range(cfg.channels):
            np.testing.assert_array_equal(got[i].channels[c], want.channels[c], err_msg=f"member {i} ch {c}")


def test_population_rejects_3d_member():
    cfg3, _ = workloads.sphere_3d(16)
    cfg2, _ = workloads.plain_256()
    with pytest.raises(_abi.HipError):
        HipPopulation([cfg2, cfg3])


def test_population_256_uneven_members_conserve_mass():
    pop_def = workloads.population(256)
    _say("[population_256] 256 members, 20 steps")
    with HipPopulation([c for c, _ in pop_def]) as pop:
        pop.set_seeds([s for _, s in pop_def])
        m0 = pop.mass_f64()
        disp = pop.evaluate_displacement(None, 0, 20)
        m1 = pop.mass_f64()
        r, g = pop.stats()
    assert np.all(np.isfinite(disp))
    drift = np.abs(m1 - m0) / np.maximum(m0, 1e-30)
    _say(f"  max relative mass drift {drift.max():.3e}, general steps {g}, displacement mean {disp.mean():.3f}")
    assert drift.max() <= 1e-6
