"""GPU parity on the reference's own example configurations, the 3D workload at its full length,
and the 2D fused stencil with a distribution box whose area (2s)^2 is not a power of two.

Bars: max-abs 1e-4 against the f32 oracle after 100 steps (the FFT's rounding is not rustfft's,
so whole trajectories are compared with a tolerance); the error against the FP64 oracle is
printed next to it.  Stencil results given the oracle's affinity are compared bit for bit.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import autoverse_amd as av
from autoverse_amd import workloads
from autoverse_amd.population import HipPopulation
from autoverse_amd.schema import FlowConfig, Pattern, Seed, SimulationConfig

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(*a):
    print(*a, flush=True)


def _load_example(name):
    with open(os.path.join(GOLDEN, f"{name}.json")) as f:
        cfg = SimulationConfig.from_json(f.read())
    with open(os.path.join(GOLDEN, f"{name}.seed.json")) as f:
        seed = Seed.from_json(f.read())
    return cfg, seed


def _oracle_runs(oracle, cfg, st0, steps):
    """f32 and FP64 oracle runs of `steps` steps from st0, side by side on host threads (the
    oracle's C code runs outside the interpreter lock)."""
    def f32():
        return oracle.Propagator(cfg, np.float32, threads=8, line_threads=8).run([c.copy() for c in st0], steps)[0]

    def f64():
        return oracle.Propagator(cfg, np.float64, threads=8, line_threads=8).run(
            [np.asarray(c, np.float64) for c in st0], steps)[0]

    with ThreadPoolExecutor(2) as ex:
        a, b = ex.submit(f32), ex.submit(f64)
        return a.result(), b.result()


def _compare(oracle, cfg, seed, steps, tag):
    st0 = oracle.seed_generate(seed, cfg.width, cfg.height, cfg.depth, cfg.channels)
    state = av.SimulationState([c.copy() for c in st0], cfg.width, cfg.height, cfg.depth)
    with av.HipPropagator(cfg) as hp:
        hp.set_state(state)
        hp.advance(steps)
        got = hp.get_state()
        _, mass = hp.mass_f64()
        r, g = hp.stats()
    want32, want64 = _oracle_runs(oracle, cfg, st0, steps)
    e32 = max(float(np.abs(got.channels[c] - want32[c]).max()) for c in range(cfg.channels))
    e64 = max(float(np.abs(got.channels[c] - want64[c]).max()) for c in range(cfg.channels))
    o64 = max(float(np.abs(np.asarray(want32[c], np.float64) - want64[c]).max()) for c in range(cfg.channels))
    m0 = sum(float(np.asarray(c, np.float64).sum()) for c in st0)
    _say(f"  {tag}: max|gpu-f32 oracle|={e32:.3e} max|gpu-f64|={e64:.3e} (f32 oracle vs FP64 {o64:.3e}); "
         f"FP64 mass drift {(mass - m0) / m0:.3e}; gather radius {r}, general-path steps {g}")
    assert got.step == steps
    return e32


@pytest.mark.parametrize("name", ["glider", "sphere3d", "torus3d"])
def test_reference_examples_100_steps(oracle, name):
    """The reference's examples/{glider,sphere3d,torus3d}{,.seed}.json (128^2 with R 13; 64^3 with
    R 10 and R 12), parsed by the schema as the reference's CLI reads them, 100 steps."""
    _say(f"[example {name}, 100 steps]")
    cfg, seed = _load_example(name)
    assert _compare(oracle, cfg, seed, 100, name) <= 1e-4


def test_3d_sphere_128_100_steps(oracle):
    """The 3D workload at its full size and the north-star length: 128^3, R 10, 100 steps."""
    _say("[3d sphere 128^3, 100 steps]")
    cfg, seed = workloads.sphere_3d(128)
    assert _compare(oracle, cfg, seed, 100, "sphere 128^3") <= 1e-4


@pytest.mark.parametrize("s", [0.75, 0.6])
@pytest.mark.parametrize("size", [(128, 64), (96, 48)])
def test_stencil2d_non_pow2_area_bitwise(oracle, s, size):
    """(2s)^2 = 2.25 or 1.44: the fused 2D stencil's division path m * (area / (2s)^2) on a grid
    of whole tiles (128 x 64) and one whose sides are not tile multiples (96 x 48, where the
    edge-rotated gather runs), bit for bit against the oracle given its affinity."""
    W, H = size
    _say(f"[stencil2d non-pow2] s={s} {W}x{H}")
    cfg, sd = workloads.multispecies(size=W, channels=3, seed=W + H)
    cfg.width, cfg.height = W, H
    cfg.flow = FlowConfig(1.0, 2.0, s)
    st = oracle.seed_generate(sd, W, H, 1, 3)
    p = oracle.Propagator(cfg, np.float32)
    ch, _, _ = p.run(st, 6)
    aff = p.affinity(ch)
    want = p.stencil(ch, aff)
    with av.HipPropagator(cfg) as hp:
        got = hp.stencil(ch, aff)
        r, g = hp.stats()
    for c in range(3):
        np.testing.assert_array_equal(got[c], want[c])
    assert g == 0, "the fused path should have covered this flow"


def test_population_non_pow2_members_match_lone_bitwise(monkeypatch):
    """Population members with s = 0.75 and 0.6 beside s = 1 members (their fused stencils run
    as separate launch groups) follow the lone propagators' trajectories bit for bit."""
    rng = np.random.default_rng(41)
    mem = []
    for side, c, s in ((64, 2, 0.75), (96, 1, 0.6), (64, 3, 1.0), (128, 2, 0.6)):
        ks = workloads._random_kernels(rng, c)
        blobs = workloads._random_blobs(rng, c)
        cfg = SimulationConfig(width=side, height=side, depth=1, channels=c, dt=0.2, kernel_radius=15,
                               kernels=ks, flow=FlowConfig(1.0, 2.0, s))
        mem.append((cfg, Seed(Pattern("MultiBlob", blobs=blobs))))
    _say("[population non-pow2] 4 members, 12 steps")
    with HipPopulation([c for c, _ in mem]) as pop:
        pop.set_seeds([s for _, s in mem])
        pop.advance(12)
        got = [pop.get_state(i) for i in range(len(mem))]
        plan = pop.plan()
        r, g = pop.stats()
    assert g == 0
    monkeypatch.setenv("HP_FFT_PATH", "generic")
    monkeypatch.setenv("HP_SPECTRAL_GROUP", "1")
    for i, (cfg, sd) in enumerate(mem):
        with av.HipPropagator(cfg) as hp:
            hp.set_plan(plan)
            st = av.SimulationState.from_seed(sd, cfg)
            hp.run(st, 12)
        for c in range(cfg.channels):
            np.testing.assert_array_equal(got[i].channels[c], st.channels[c], err_msg=f"member {i} ch {c}")


def test_tape_backward_refusal_keeps_final_state():
    """A backward call refused for a NULL dL/dA_T channel leaves the population holding A_T, and
    a valid backward call still runs afterwards."""
    import ctypes

    from autoverse_amd import _abi
    rng = np.random.default_rng(3)
    ks = workloads._random_kernels(rng, 2)
    cfg = SimulationConfig(width=64, height=64, depth=1, channels=2, dt=0.2, kernel_radius=15, kernels=ks,
                           flow=FlowConfig(1.0, 2.0, 1.0))
    seed = Seed(Pattern("MultiBlob", blobs=workloads._random_blobs(rng, 2)))
    _say("[tape_backward refusal]")
    with HipPopulation([cfg]) as pop:
        pop.set_seeds([seed])
        pop.tape_forward(6, 2)
        final = pop.get_state(0)
        ok = np.ones(64 * 64, np.float32)
        seeds = _abi.float_ptr_array([ok, ok])
        seeds[1] = None
        out_bufs = [np.empty(64 * 64, np.float32) for _ in range(2)]  # kept alive across the call
        outs = _abi.float_ptr_array(out_bufs)
        th = np.zeros(pop.theta_count(0), np.float64)
        st = pop._lib.hp_batch_tape_backward(pop._ptr, seeds, outs, th.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert st == _abi.HP_ERR_NULL
        after = pop.get_state(0)
        for c in range(2):
            np.testing.assert_array_equal(after.channels[c], final.channels[c])
        g0, dth = pop.tape_backward([[ok, ok]])
        assert np.isfinite(dth[0]).all() and all(np.isfinite(g).all() for g in g0[0])
        again = pop.get_state(0)
        for c in range(2):
            np.testing.assert_array_equal(again.channels[c], final.channels[c])
