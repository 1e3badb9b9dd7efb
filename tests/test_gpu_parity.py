"""GPU parity of the HIP Flow Lenia step against the CPU oracle (oracle/), through the C ABI.

Bars: bit-exact where the arithmetic is the reference's own f32 sequence (Sobel, alpha,
flow, reintegration, seeds); within a stated tolerance where the FFT is involved
(the FFT's rounding is not the reference's rustfft's; see oracle/flowlenia_oracle.h).
"""
import numpy as np
import pytest

import autoverse_amd as av
from autoverse_amd import workloads

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _evolve(O, cfg, seed, steps):
    st = O.seed_generate(seed, cfg.width, cfg.height, cfg.depth, cfg.channels)
    p = O.Propagator(cfg, np.float32)
    ch, _, _ = p.run(st, steps)
    return p, ch


def _small_multispecies(size=128, channels=3, seed=7):
    cfg, sd = workloads.multispecies(size=size, channels=channels, seed=seed)
    return cfg, sd


# ---------------------------------------------------------------------------------------------
# stencil (stages 2-5) given the oracle's affinity: bit-exact
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("case", ["plain256", "multi128x3", "multi96x2"])
def test_stencil_bitwise(oracle, case):
    _say(f"[stencil_bitwise] {case}")
    if case == "plain256":
        cfg, sd = workloads.plain_256()
        warm = 20
    elif case == "multi128x3":
        cfg, sd = _small_multispecies(128, 3)
        warm = 10
    else:
        cfg, sd = _small_multispecies(96, 2, seed=11)
        warm = 5
    p, ch = _evolve(oracle, cfg, sd, warm)
    aff = p.affinity(ch)
    want = p.stencil(ch, aff)
    with av.HipPropagator(cfg) as hp:
        got = hp.stencil(ch, aff)
        r, g = hp.stats()
    for c in range(cfg.channels):
        np.testing.assert_array_equal(got[c], want[c])
    assert g == 0, "the fused path should have covered this flow"


def test_stencil_general_path_bitwise(oracle):
    """Large dt pushes sources beyond the fused radius: the general path must still be bit-exact."""
    _say("[stencil_general] dt=4")
    cfg, sd = _small_multispecies(64, 2, seed=3)
    p, ch = _evolve(oracle, cfg, sd, 5)
    cfg.dt = 4.0
    p2 = oracle.Propagator(cfg, np.float32)
    aff = p2.affinity(ch)
    want = p2.stencil(ch, aff)
    with av.HipPropagator(cfg) as hp:
        got = hp.stencil(ch, aff)
        r, g = hp.stats()
    for c in range(cfg.channels):
        np.testing.assert_array_equal(got[c], want[c])
    assert g >= 1 and r > 2


# ---------------------------------------------------------------------------------------------
# affinity (stage 1): FFT convolution + growth, tolerance
# ---------------------------------------------------------------------------------------------
AFF_TOL = 2e-5  # max-abs on U; G amplifies conv error by up to ~1.2/sigma


@pytest.mark.parametrize("size,channels", [(256, 1), (128, 3), (192, 2), (96, 1), (48, 2), (1024, 3), (512, 2),
                                           ((256, 512), 1), ((512, 256), 2), (2048, 1)])
def test_affinity_vs_oracle(oracle, size, channels):
    """Sides 256..2048 run the column-major wave-FFT path, the others the generic passes."""
    _say(f"[affinity] {size}x{channels}")
    if channels == 1 and size == 256:
        cfg, sd = workloads.plain_256()
    else:
        w, h = (size, size) if isinstance(size, int) else size
        cfg, sd = _small_multispecies(w, channels, seed=w + h + channels)
        cfg.width, cfg.height = w, h
    st = oracle.seed_generate(sd, cfg.width, cfg.height, 1, cfg.channels)
    p32 = oracle.Propagator(cfg, np.float32)
    p64 = oracle.Propagator(cfg, np.float64)
    want32 = p32.affinity(st)
    want64 = p64.affinity([s.astype(np.float64) for s in st])
    with av.HipPropagator(cfg) as hp:
        got = hp.affinity(st)
    e32 = max(float(np.abs(got[c] - want32[c]).max()) for c in range(cfg.channels))
    e64 = max(float(np.abs(got[c] - want64[c]).max()) for c in range(cfg.channels))
    o64 = max(float(np.abs(want32[c] - want64[c]).max()) for c in range(cfg.channels))
    _say(f"  max|gpu-f32 oracle|={e32:.3e} max|gpu-f64|={e64:.3e} max|f32 oracle-f64|={o64:.3e}")
    assert e32 <= AFF_TOL
    # against FP64 the bar is the f32 oracle's own error (f32 arithmetic, growth-amplified)
    assert e64 <= max(AFF_TOL, 1.5 * o64)


def test_spectral_grouping_is_bitwise_neutral(oracle, monkeypatch):
    """Grouping the spectral pass per source or per kernel runs the same transforms."""
    _say("[spectral grouping]")
    cfg, sd = _small_multispecies(512, 3, seed=9)
    st = oracle.seed_generate(sd, 512, 512, 1, 3)
    outs = []
    for grp in ("0", "1"):
        monkeypatch.setenv("HP_SPECTRAL_GROUP", grp)
        with av.HipPropagator(cfg) as hp:
            outs.append(hp.affinity(st))
    for c in range(3):
        np.testing.assert_array_equal(outs[0][c], outs[1][c])


# ---------------------------------------------------------------------------------------------
# full steps
# ---------------------------------------------------------------------------------------------
def test_plain256_100_steps(oracle):
    """The plain configuration, 100 steps: max-abs <= 1e-4 against the f32 oracle, and the error
    against the FP64 oracle is stated."""
    _say("[plain256_100]")
    cfg, sd = workloads.plain_256()
    st0 = oracle.seed_generate(sd, 256, 256, 1, 1)
    want32, t32, s32 = oracle.Propagator(cfg, np.float32).run(st0, 100)
    want64, _, _ = oracle.Propagator(cfg, np.float64).run([st0[0].astype(np.float64)], 100)
    state = av.SimulationState([st0[0].copy()], 256, 256)
    with av.HipPropagator(cfg) as hp:
        hp.run(state, 100)
        _, mass = hp.mass_f64()
    e32 = float(np.abs(state.channels[0] - want32[0]).max())
    e64 = float(np.abs(state.channels[0] - want64[0]).max())
    m0 = float(st0[0].astype(np.float64).sum())
    _say(f"  max|gpu-f32|={e32:.3e} max|gpu-f64|={e64:.3e} mass drift={(mass - m0) / m0:.3e}")
    assert e32 <= 1e-4
    assert state.step == s32 == 100
    assert state.time == pytest.approx(t32, abs=0)
    assert abs(mass - float(want64[0].sum())) / m0 < 1e-6


@pytest.mark.parametrize("kind", ["empty", "one_cell"])
def test_edge_states_vs_oracle(oracle, kind):
    """Edge states on the 128^2 x 3 multispecies configuration, 10 steps: an all-zero grid (every
    tile and channel skipped) stays exactly zero, with the oracle's step counter and time; a
    single cell of mass 1 in channel 1 keeps its mass and matches the f32 oracle to 1e-5."""
    _say(f"[edge_states {kind}]")
    cfg, _ = _small_multispecies(128, 3)
    n = 128 * 128
    chans = [np.zeros(n, np.float32) for _ in range(3)]
    if kind == "one_cell":
        chans[1][37 * 128 + 90] = 1.0
    want, t32, s32 = oracle.Propagator(cfg, np.float32).run([c.copy() for c in chans], 10)
    state = av.SimulationState([c.copy() for c in chans], 128, 128)
    with av.HipPropagator(cfg) as hp:
        hp.run(state, 10)
        per, mass = hp.mass_f64()
    assert state.step == s32 == 10
    assert state.time == pytest.approx(t32, abs=0)
    err = max(float(np.abs(state.channels[c] - want[c]).max()) for c in range(3))
    _say(f"  max|gpu-f32|={err:.3e} mass={mass:.9f}")
    if kind == "empty":
        assert all(not np.any(c) for c in state.channels)
        assert mass == 0.0
    else:
        assert err <= 1e-5
        assert abs(mass - 1.0) < 1e-6
        assert not np.any(state.channels[0]) and not np.any(state.channels[2])


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_nonfinite_input_is_rejected_and_context_stays_usable(oracle, bad):
    """A state holding NaN or Inf is refused on the host with HP_ERR_NONFINITE (the reference
    would carry it through the step); the context then steps a valid state as before."""
    from autoverse_amd import _abi
    _say(f"[nonfinite_input {bad}]")
    cfg, sd = workloads.plain_256()
    st0 = oracle.seed_generate(sd, 256, 256, 1, 1)
    want, _, _ = oracle.Propagator(cfg, np.float32).run([st0[0].copy()], 3)
    poisoned = st0[0].copy()
    poisoned[1000] = bad
    with av.HipPropagator(cfg) as hp:
        with pytest.raises(_abi.HipError) as ei:
            hp.set_state(av.SimulationState([poisoned], 256, 256))
        assert ei.value.status == _abi.HP_ERR_NONFINITE
        with pytest.raises(_abi.HipError) as ei:
            hp.step(av.SimulationState([poisoned.copy()], 256, 256))
        assert ei.value.status == _abi.HP_ERR_NONFINITE
        state = av.SimulationState([st0[0].copy()], 256, 256)
        hp.run(state, 3)
    assert state.step == 3
    assert float(np.abs(state.channels[0] - want[0]).max()) <= 1e-5


def test_refused_state_leaves_the_device_state_unchanged(oracle):
    """Three channels with NaN only in the last one: every channel is checked before any is
    copied, so the context and a population member both keep their previous state bit for bit."""
    from autoverse_amd import _abi
    from autoverse_amd.population import HipPopulation
    _say("[refused_state_unchanged]")
    cfg, sd = _small_multispecies(128, 3)
    good = oracle.seed_generate(sd, 128, 128, 1, 3)
    bad = [np.full(128 * 128, 0.25, np.float32) for _ in range(3)]
    bad[2][77] = np.nan
    with av.HipPropagator(cfg) as hp:
        hp.set_state(av.SimulationState([c.copy() for c in good], 128, 128))
        with pytest.raises(_abi.HipError) as ei:
            hp.set_state(av.SimulationState([c.copy() for c in bad], 128, 128))
        assert ei.value.status == _abi.HP_ERR_NONFINITE
        got = hp.get_state()
    for c in range(3):
        assert np.array_equal(got.channels[c], good[c])
    pop = HipPopulation([cfg])
    try:
        pop.set_state(0, av.SimulationState([c.copy() for c in good], 128, 128))
        with pytest.raises(_abi.HipError) as ei:
            pop.set_state(0, av.SimulationState([c.copy() for c in bad], 128, 128))
        assert ei.value.status == _abi.HP_ERR_NONFINITE
        got = pop.get_state(0)
    finally:
        pop.close()
    for c in range(3):
        assert np.array_equal(got.channels[c], good[c])


def test_plain256_mass_drift_1000_steps(oracle):
    """Mass conservation over 1000 steps of the plain configuration.  Reintegration in f32 (the
    reference's arithmetic, which the stencil reproduces bit for bit) is not exactly
    conservative: the f32 oracle itself drifts ~2e-6 per 1000 steps here, the FP64 oracle ~2e-8.
    The device's FP64-summed drift is stated next to both and must stay within 3x the f32
    oracle's."""
    _say("[plain256_drift_1000]")
    cfg, sd = workloads.plain_256()
    st0 = oracle.seed_generate(sd, 256, 256, 1, 1)
    m0 = float(np.asarray(st0[0], np.float64).sum())
    want32, _, _ = oracle.Propagator(cfg, np.float32).run(st0, 1000)
    want64, _, _ = oracle.Propagator(cfg, np.float64).run([np.asarray(st0[0], np.float64)], 1000)
    d32 = (float(np.asarray(want32[0], np.float64).sum()) - m0) / m0
    d64 = (float(want64[0].sum()) - m0) / m0
    state = av.SimulationState([st0[0].copy()], 256, 256)
    with av.HipPropagator(cfg) as hp:
        hp.set_state(state)
        hp.advance(1000)
        _, mass = hp.mass_f64()
    dg = (mass - m0) / m0
    _say(f"  relative mass drift over 1000 steps: device {dg:.3e}, f32 oracle {d32:.3e}, FP64 oracle {d64:.3e}")
    assert abs(dg) <= 3.0 * abs(d32) + 2e-7


@pytest.mark.parametrize("size,channels", [(128, 2), (512, 3)])
def test_step_equals_run_and_is_deterministic(oracle, size, channels):
    """hp_step (one checked step at a time, launched kernel by kernel) against hp_run (hipGraph
    replay with one flag check per run; 512^2 takes the wave-level FFT path): the same
    trajectory bit for bit, and repeatable."""
    _say(f"[step_vs_run] {size}^2 x {channels}")
    cfg, sd = _small_multispecies(size, channels, seed=5)
    st0 = oracle.seed_generate(sd, size, size, 1, channels)
    a = av.SimulationState([c.copy() for c in st0], size, size)
    b = av.SimulationState([c.copy() for c in st0], size, size)
    c = av.SimulationState([x.copy() for x in st0], size, size)
    with av.HipPropagator(cfg) as hp:
        for _ in range(12):
            hp.step(a)
        hp.run(b, 12)
        hp.run(c, 12)
    for i in range(channels):
        np.testing.assert_array_equal(a.channels[i], b.channels[i])
        np.testing.assert_array_equal(b.channels[i], c.channels[i])
    assert a.step == b.step == 12 and a.time == b.time


def test_multispecies_1024_steps(oracle):
    """The 1024^2 x 3 workload for 100 steps: max-abs <= 1e-4 against the f32 oracle; the error
    against the FP64 oracle is stated (the f32 oracle's own FP64 error is ~1.2e-4 here)."""
    _say("[multispecies_1024x3 100 steps]")
    cfg, sd = workloads.multispecies()
    st0 = oracle.seed_generate(sd, 1024, 1024, 1, 3)
    want, _, _ = oracle.Propagator(cfg, np.float32, threads=16).run(st0, 100)
    want64, _, _ = oracle.Propagator(cfg, np.float64, threads=16).run([np.asarray(c, np.float64) for c in st0], 100)
    state = av.SimulationState([c.copy() for c in st0], 1024, 1024)
    with av.HipPropagator(cfg) as hp:
        hp.run(state, 100)
    e = max(float(np.abs(state.channels[c] - want[c]).max()) for c in range(3))
    e64 = max(float(np.abs(state.channels[c] - want64[c]).max()) for c in range(3))
    o64 = max(float(np.abs(np.asarray(want[c], np.float64) - want64[c]).max()) for c in range(3))
    _say(f"  max|gpu-f32|={e:.3e} max|gpu-f64|={e64:.3e} (f32 oracle vs FP64: {o64:.3e})")
    assert e <= 1e-4


# ---------------------------------------------------------------------------------------------
# stage entry points
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("fx,fy,s,dt", [(0.0, 0.0, 1.0, 0.2), (5.0, 2.0, 1.0, 0.2), (8.0, 8.0, 1.0, 0.2),
                                        (100.0, 50.0, 1.0, 0.2), (1.0, 0.0, 0.01, 0.2), (2.0, 1.0, 0.5, 0.2),
                                        (0.3, -0.7, 1e-6, 1.0)])
def test_advect_stage_bitwise(oracle, fx, fy, s, dt):
    """advect_mass on the reference's own cases (reintegration.rs:235-470) and random fields."""
    _say(f"[advect] flow=({fx},{fy}) s={s}")
    rng = np.random.default_rng(1)
    for (w, h) in [(16, 16), (32, 32), (40, 24)]:
        cur = rng.uniform(0, 1, (h, w)).astype(np.float32)
        cur[rng.uniform(size=(h, w)) < 0.2] = 0.0
        flx = (np.full((h, w), fx) + rng.normal(0, 0.3, (h, w))).astype(np.float32)
        fly = (np.full((h, w), fy) + rng.normal(0, 0.3, (h, w))).astype(np.float32)
        want = oracle.advect2d(cur, flx, fly, dt, s)
        got = av.advect_mass(cur, flx, fly, w, h, dt, s).reshape(h, w)
        np.testing.assert_array_equal(got, want)


def test_sobel_flow_growth_stages(oracle):
    _say("[stages]")
    rng = np.random.default_rng(2)
    g = rng.normal(size=(37, 53)).astype(np.float32)
    gx, gy = av.sobel_gradient(g, 53, 37)
    wx, wy = oracle.sobel2d(g)
    np.testing.assert_array_equal(gx.reshape(37, 53), wx)
    np.testing.assert_array_equal(gy.reshape(37, 53), wy)
    n = 4096
    a = [rng.normal(size=n).astype(np.float32) for _ in range(4)]
    m = rng.uniform(0, 3, n).astype(np.float32)
    for (beta, nexp) in [(1.0, 2.0), (2.0, 4.0), (1.5, 1.0), (1.0, 3.0)]:
        fx, fy = av.compute_flow_field(a[0], a[1], a[2], a[3], m, av.FlowConfig(beta, nexp, 1.0))
        wx, wy = oracle.flow2d(a[0], a[1], a[2], a[3], m, beta, nexp)
        np.testing.assert_array_equal(fx, wx)
        np.testing.assert_array_equal(fy, wy)
    conv = rng.uniform(0, 0.4, n).astype(np.float32)
    t = rng.normal(size=n).astype(np.float32)
    want = t.copy()
    oracle.growth_accumulate(conv, want, 0.7, 0.15, 0.015)
    av.growth_accumulate(conv, t, 0.7, 0.15, 0.015)
    assert float(np.abs(t - want).max()) <= 4e-7  # device expf is not glibc expf (<= 2 ulp)


# ---------------------------------------------------------------------------------------------
# 3D
# ---------------------------------------------------------------------------------------------
def test_3d_small_vs_oracle(oracle):
    _say("[3d 32^3]")
    cfg, sd = workloads.sphere_3d(32)
    cfg.kernel_radius = 6
    st0 = oracle.seed_generate(sd, 32, 32, 32, 1)
    p = oracle.Propagator(cfg, np.float32)
    want, _, _ = p.run(st0, 10)
    aff = p.affinity(want)
    nxt = p.stencil(want, aff)
    state = av.SimulationState([st0[0].copy()], 32, 32, 32)
    with av.HipPropagator(cfg) as hp:
        got_aff = hp.affinity(want)
        got_nxt = hp.stencil(want, aff)
        hp.run(state, 10)
    ea = float(np.abs(got_aff[0] - aff[0]).max())
    e = float(np.abs(state.channels[0] - want[0]).max())
    _say(f"  affinity max err {ea:.3e}; 10-step max err {e:.3e}")
    np.testing.assert_array_equal(got_nxt[0], nxt[0])
    assert ea <= AFF_TOL
    assert e <= 1e-4


def test_3d_sphere_128_steps(oracle):
    """The 3D configuration at its full size, 128^3 with R = 10: 20 steps against the f32 oracle
    (max-abs <= 1e-4), the FP64 error stated, and the FP64 mass drift."""
    _say("[3d sphere 128^3, 20 steps]")
    cfg, sd = workloads.sphere_3d(128)
    st0 = oracle.seed_generate(sd, 128, 128, 128, 1)
    want, _, _ = oracle.Propagator(cfg, np.float32, threads=16).run(st0, 20)
    want64, _, _ = oracle.Propagator(cfg, np.float64, threads=16).run([np.asarray(st0[0], np.float64)], 20)
    state = av.SimulationState([st0[0].copy()], 128, 128, 128)
    with av.HipPropagator(cfg) as hp:
        hp.set_state(state)
        hp.advance(20)
        got = hp.get_state()
        _, mass = hp.mass_f64()
    e = float(np.abs(got.channels[0] - want[0]).max())
    e64 = float(np.abs(got.channels[0] - want64[0]).max())
    m0 = float(np.asarray(st0[0], np.float64).sum())
    _say(f"  max|gpu-f32|={e:.3e} max|gpu-f64|={e64:.3e} mass drift={(mass - m0) / m0:.3e}")
    assert e <= 1e-4


@pytest.mark.parametrize("shape,channels,dt,s", [((24, 16, 8), 2, 0.1, 1.0), ((8, 8, 8), 1, 0.1, 1.0),
                                                 ((32, 24, 16), 3, 0.2, 1.0), ((16, 16, 16), 1, 40.0, 1.0),
                                                 ((48, 24, 16), 2, 0.1, 0.75), ((64, 48, 32), 1, 0.1, 0.5),
                                                 ((48, 32, 24), 3, 0.15, 0.6), ((24, 96, 8), 1, 0.2, 1.0)])
def test_3d_stencil_bitwise(oracle, shape, channels, dt, s):
    """The fused 3D stencil (64-source gather) on random states: the oracle's next state bit for
    bit, with (2s)^3 a power of two and not, on grids that are and are not tile multiples (so
    that both the unrotated and the edge-rotated gathers run); a large dt leaves the fused window
    and must come back through the general path."""
    from autoverse_amd.schema import FlowConfig, SimulationConfig
    W, H, D = shape
    _say(f"[3d_stencil] {shape} C={channels} dt={dt} s={s}")
    rng = np.random.default_rng(W * H + D + channels)
    ks = workloads._random_kernels(rng, channels)
    cfg = SimulationConfig(width=W, height=H, depth=D, channels=channels, dt=dt, kernel_radius=3, kernels=ks,
                           flow=FlowConfig(1.0, 2.0, s))
    ch = [rng.random(W * H * D).astype(np.float32) for _ in range(channels)]
    p = oracle.Propagator(cfg, np.float32)
    aff = p.affinity(ch)
    want = p.stencil(ch, aff)
    with av.HipPropagator(cfg) as hp:
        got = hp.stencil(ch, aff)
        r, g = hp.stats()
    for c in range(channels):
        np.testing.assert_array_equal(got[c], want[c])
    if dt >= 3.0:
        assert g >= 1  # the fused window was left: the general path produced the result
    else:
        assert g == 0



@pytest.mark.parametrize("shape", [(64, 64, 1), (32, 32, 24)])
def test_stencil_huge_mass_leaves_fused_path(oracle, shape):
    """With (2s)^d a power of two the fused stencils pre-scale each mass by 1/(2s)^d; for
    s = 0.25 and a mass of 1e38 that product overflows while the reference's m * (V_o / V)
    does not.  The source must be sent to the general path and the result stay bit-exact."""
    from autoverse_amd.schema import FlowConfig, SimulationConfig
    W, H, D = shape
    _say(f"[huge_mass] {shape}")
    rng = np.random.default_rng(17)
    ks = workloads._random_kernels(rng, 1)
    cfg = SimulationConfig(width=W, height=H, depth=D, channels=1, dt=0.1, kernel_radius=3, kernels=ks,
                           flow=FlowConfig(1.0, 2.0, 0.25))
    a = np.zeros((D, H, W), np.float32)
    a[: max(1, D // 4), : H // 4, :] = rng.random((max(1, D // 4), H // 4, W)).astype(np.float32)
    a[D // 2, H // 2, W // 2] = 1e38
    ch = [a.reshape(-1)]
    aff = [(0.1 * rng.random(W * H * D)).astype(np.float32)]
    want = oracle.Propagator(cfg, np.float32).stencil(ch, aff)
    with av.HipPropagator(cfg) as hp:
        got = hp.stencil(ch, aff)
        r, g = hp.stats()
    assert np.isfinite(want[0]).all()
    np.testing.assert_array_equal(got[0], want[0])
    assert g >= 1


@pytest.mark.parametrize("shape", [(128, 128, 1), (48, 32, 24)])
def test_stencil_sparse_sources_bitwise(oracle, shape):
    """Isolated masses on and around tile boundaries (so that some tiles hold their only active
    sources in the halo), a second channel that is zero everywhere, and a third whose one mass
    sits at the far corner: tiles with no active source store +0 and the rest match the oracle
    bit for bit."""
    from autoverse_amd.schema import FlowConfig, SimulationConfig
    W, H, D = shape
    _say(f"[sparse_sources] {shape}")
    rng = np.random.default_rng(23)
    ks = workloads._random_kernels(rng, 3)
    cfg = SimulationConfig(width=W, height=H, depth=D, channels=3, dt=0.2, kernel_radius=3, kernels=ks,
                           flow=FlowConfig(1.0, 2.0, 1.0))
    a = np.zeros((3, D, H, W), np.float32)
    xs = [x for x in (0, 1, 15, 16, 17, 31, 32, 33, 63, 64, 65, W - 2, W - 1) if x < W]
    ys = [y for y in (0, 7, 8, 9, 15, 16, 17, H - 2, H - 1) if y < H]
    zs = [z for z in (0, 7, 8, 9, D - 1) if z < D] if D > 1 else [0]
    for z in zs[::2]:
        for y in ys[::2]:
            for x in xs[::3]:
                a[0, z, y, x] = 0.25 + 0.5 * rng.random()
    a[2, D - 1, H - 1, W - 1] = 0.7
    ch = [a[c].reshape(-1) for c in range(3)]
    aff = [(0.2 * rng.random(W * H * D) - 0.1).astype(np.float32) for _ in range(3)]
    want = oracle.Propagator(cfg, np.float32).stencil(ch, aff)
    with av.HipPropagator(cfg) as hp:
        got = hp.stencil(ch, aff)
        r, g = hp.stats()
    for c in range(3):
        np.testing.assert_array_equal(got[c], want[c])
    assert not got[1].any() and got[0].any() and got[2].any()
    assert g == 0


@pytest.mark.parametrize("shape", [(96, 64, 1), (32, 24, 16)])
def test_stencil_five_channels_bitwise(oracle, shape):
    """Five channels: the 2D stencil's body for a channel count known only at run time (masses
    re-read per channel, every channel treated as active) and the 3D kernel's per-channel mass
    prefetch; one channel is zero everywhere."""
    from autoverse_amd.schema import FlowConfig, SimulationConfig
    W, H, D = shape
    C = 5
    _say(f"[five_channels] {shape}")
    rng = np.random.default_rng(31)
    ks = workloads._random_kernels(rng, C)
    cfg = SimulationConfig(width=W, height=H, depth=D, channels=C, dt=0.2, kernel_radius=3, kernels=ks,
                           flow=FlowConfig(1.0, 2.0, 1.0))
    ch = [rng.random(W * H * D).astype(np.float32) for _ in range(C)]
    ch[3][:] = 0.0
    aff = [(0.2 * rng.random(W * H * D) - 0.1).astype(np.float32) for _ in range(C)]
    want = oracle.Propagator(cfg, np.float32).stencil(ch, aff)
    with av.HipPropagator(cfg) as hp:
        got = hp.stencil(ch, aff)
        r, g = hp.stats()
    for c in range(C):
        np.testing.assert_array_equal(got[c], want[c])
    assert not got[3].any()
    assert g == 0
