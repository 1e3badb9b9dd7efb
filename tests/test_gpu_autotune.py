"""Execution plans and their autotuner: every plan computes the same step.

Tile shapes and spectral grouping are bitwise neutral (each destination still adds its
sources in the scatter's order; each kernel's spectral product is the same arithmetic), so
the fused stencil under every tile must reproduce the oracle's next state bit for bit.  FFT
variants (wave or LDS path, radix order, column split) change only the FFT's rounding: the
affinity stays within the plain path's tolerance of the oracle.
"""
import numpy as np
import pytest

import autoverse_amd as av
from autoverse_amd import workloads
from autoverse_amd.population import HipPopulation

pytestmark = pytest.mark.gpu

AFF_TOL = 2e-5


def _say(*a):
    print(*a, flush=True)


def _evolve(O, cfg, seed, steps):
    st = O.seed_generate(seed, cfg.width, cfg.height, cfg.depth, cfg.channels)
    p = O.Propagator(cfg, np.float32)
    ch, _, _ = p.run(st, steps)
    return p, ch


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4, 5])
def test_every_stencil_tile_is_bitwise_exact(oracle, tile):
    _say(f"[tile] {tile}")
    cfg, sd = workloads.multispecies(size=96, channels=3, seed=12)
    p, ch = _evolve(oracle, cfg, sd, 6)
    aff = p.affinity(ch)
    want = p.stencil(ch, aff)
    with av.HipPropagator(cfg) as hp:
        hp.set_plan({"stencil_tile": tile})
        assert hp.plan()["stencil_tile"] == tile
        got = hp.stencil(ch, aff)
        _, g = hp.stats()
    assert g == 0
    for c in range(cfg.channels):
        np.testing.assert_array_equal(got[c], want[c])


FFT_PLANS = [dict(fft_path=1, spectral_group=gr, radix_order=rp, column_split=cs)
             for gr in (0, 1) for rp in (0, 1) for cs in (1, 2)] + [dict(fft_path=0, spectral_group=1),
                                                                    dict(fft_path=0, spectral_group=0)]


@pytest.mark.parametrize("plan", FFT_PLANS, ids=lambda p: "-".join(f"{k[0]}{v}" for k, v in p.items()))
@pytest.mark.parametrize("size", [192, 256])
def test_every_fft_variant_matches_the_oracle(oracle, plan, size):
    cfg, sd = workloads.multispecies(size=size, channels=2, seed=size)
    p, ch = _evolve(oracle, cfg, sd, 3)
    want = p.affinity(ch)
    with av.HipPropagator(cfg) as hp:
        hp.set_plan(plan)
        got_plan = hp.plan()
        if plan["fft_path"] == 0 and size != 256:
            assert got_plan["fft_path"] == 1  # the wave path needs sides 256..2048
        got = hp.affinity(ch)
    err = max(float(np.abs(g - w).max()) for g, w in zip(got, want))
    assert err <= AFF_TOL, (plan, got_plan, err)


@pytest.mark.parametrize("size", [128, 256])
def test_autotune_keeps_state_and_picks_a_measured_plan(oracle, size):
    cfg, sd = workloads.multispecies(size=size, channels=3, seed=5)
    with av.HipPropagator(cfg) as hp:
        st = av.SimulationState.from_seed(sd, cfg)
        hp.set_state(st)
        hp.advance(3)
        before = hp.get_state()
        best, table = hp.autotune(steps=8)
        after = hp.get_state()
        _say(f"[autotune {size}] best {best}")
        for plan, us in table:
            _say(f"   {plan} -> {us}")
        assert hp.plan() == best
        assert after.step == before.step and after.time == before.time
        for c in range(cfg.channels):
            np.testing.assert_array_equal(after.channels[c], before.channels[c])
        timed = [us for _, us in table if us is not None]
        assert len(timed) >= (4 if size == 256 else 3) and min(timed) > 0
        assert any(p == best for p, _ in table)
        hp.run(st, 10)  # from the seed state (run uploads it), under the tuned plan
    # a later propagator of the same shape starts from the tuned plan
    with av.HipPropagator(cfg) as hp2:
        assert hp2.plan() == best
    ref = oracle.Propagator(cfg, np.float32)
    want, _, _ = ref.run(oracle.seed_generate(sd, size, size, 1, 3), 10)
    assert max(float(np.abs(st.channels[c] - want[c]).max()) for c in range(3)) <= 1e-4


def test_population_autotune_keeps_states_and_trajectories():
    mem = workloads.population(24, seed=77)
    cfgs = [c for c, _ in mem]
    with HipPopulation(cfgs) as pop:
        pop.set_seeds([s for _, s in mem])
        pop.advance(2)
        before = [pop.get_state(i) for i in range(len(cfgs))]
        best, table = pop.autotune(steps=6)
        _say(f"[population autotune] best {best}; {len(table)} plans timed")
        after = [pop.get_state(i) for i in range(len(cfgs))]
        for a, b in zip(before, after):
            assert a.step == b.step
            for c in range(len(a.channels)):
                np.testing.assert_array_equal(a.channels[c], b.channels[c])
        assert pop.plan() == best
        pop.advance(5)
        tuned = [pop.get_state(i) for i in range(len(cfgs))]
    with HipPopulation(cfgs) as ref:
        ref.set_seeds([s for _, s in mem])
        ref.advance(7)
        want = [ref.get_state(i) for i in range(len(cfgs))]
    for a, b in zip(tuned, want):
        for c in range(len(a.channels)):
            assert float(np.abs(a.channels[c] - b.channels[c]).max()) <= 1e-4
