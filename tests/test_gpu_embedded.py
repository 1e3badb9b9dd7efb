"""The embedded-parameter propagator (hp_emb_*) on the GPU against the oracle's restatement of
EmbeddedPropagator (oracle/oracle_embedded.inc).

Bars:
  * affinity: the device convolves by FFT, the reference sums kernel taps directly; the two agree
    to the FFT's rounding, amplified by the growth's slope (max-abs 1e-4 on the growth sums);
  * stages 2-5 given the oracle's affinity: the mass and, with linear mixing, the parameters bit
    for bit (integer n, so alpha takes the reference's exact fast paths); with softmax mixing the
    parameters within 1e-6 relative (exp is FP64-rounded on the device, libm's expf in the
    reference); what next_params keeps, bit for bit; the search radius equal;
  * the reference's own tests (embedded_propagator.rs:398-452, param_advection.rs:364-457) and
    the duplicate-visit closed form through the stage entry point;
  * whole trajectories: 20 steps within 1e-4 of the f32 oracle; the 100-step distance is printed.
"""
import numpy as np
import pytest

from autoverse_amd import _abi
from autoverse_amd.embedded import (CellParams, EmbeddedState, HipEmbeddedPropagator, ParameterGrid,
                                    advect_mass_and_params, default_params)
from autoverse_amd.schema import EmbeddingConfig, FlowConfig, KernelConfig, Pattern, RingConfig, Seed, SimulationConfig
from autoverse_amd import workloads

pytestmark = pytest.mark.gpu

MODES = {"softmax": EmbeddingConfig(True, 0.7, False), "linear": EmbeddingConfig(True, 1.0, True),
         "disabled": EmbeddingConfig(False, 1.0, False)}


def _say(*a):
    print(*a, flush=True)


def _cfg(size=64, channels=2, kernel_radius=6, mode="softmax", seed=3, dt=0.2, s=1.0):
    rng = np.random.default_rng(seed)
    ks = workloads._random_kernels(rng, channels)
    return SimulationConfig(width=size, height=size, depth=1, channels=channels, dt=dt, kernel_radius=kernel_radius,
                            kernels=ks, flow=FlowConfig(1.0, 2.0, s), embedding=MODES[mode])


def _state(cfg, seed=5, integer_n=True):
    rng = np.random.default_rng(seed)
    n = cfg.width * cfg.height
    sd = Seed(Pattern("MultiBlob", blobs=workloads._random_blobs(rng, cfg.channels, 4)))
    from oracle import oracle as O
    ch = O.seed_generate(sd, cfg.width, cfg.height, 1, cfg.channels)
    nvals = [1.0, 2.0, 3.0, 4.0] if integer_n else [1.5, 2.0, 2.5, 3.0]
    ps = [np.stack([rng.uniform(0.1, 0.3, n), rng.uniform(0.01, 0.05, n), rng.uniform(0.3, 1.0, n),
                    rng.uniform(0.5, 2.0, n), rng.choice(nvals, n)], 1).astype(np.float32) for _ in range(cfg.channels)]
    return ch, ps


@pytest.mark.parametrize("mode", ["softmax", "linear", "disabled"])
def test_affinity_against_direct_convolution(oracle, mode):
    _say(f"[emb affinity] {mode}")
    cfg = _cfg(mode=mode)
    ch, ps = _state(cfg)
    # evolve a few oracle steps so the fields are not just blobs
    o = oracle.EmbeddedPropagator(cfg, cfg.embedding.__dict__)
    ch, ps, _, _ = o.run(ch, ps, 5)
    want = o.affinity(ch, ps)
    with HipEmbeddedPropagator(cfg) as hp:
        got = hp.affinity(ch, ps)
    err = max(float(np.abs(g - w).max()) for g, w in zip(got, want))
    _say(f"  max|gpu - oracle| affinity = {err:.3e}")
    assert err <= 1e-4


@pytest.mark.parametrize("mode,size,s", [("softmax", 64, 1.0), ("linear", 64, 1.0), ("disabled", 64, 1.0),
                                         ("linear", 48, 0.6), ("softmax", 96, 0.75)])
def test_stencil_given_affinity(oracle, mode, size, s):
    _say(f"[emb stencil] {mode} {size}^2 s={s}")
    cfg = _cfg(size=size, mode=mode, s=s, seed=size)
    ch, ps = _state(cfg, seed=size + 1)
    # a few steps with embedding disabled evolve the mass and leave the parameters (integer n) as
    # they are
    ch, _, _, _ = oracle.EmbeddedPropagator(cfg, MODES["disabled"].__dict__).run(ch, ps, 3)
    o = oracle.EmbeddedPropagator(cfg, cfg.embedding.__dict__)
    aff = o.affinity(ch, ps)
    sentinel = [np.full((cfg.width * cfg.height, 5), -3.0, np.float32) for _ in range(cfg.channels)]
    want_m, want_p = o.stencil(ch, ps, aff, sentinel)
    with HipEmbeddedPropagator(cfg) as hp:
        got_m, got_p = hp.stencil(ch, ps, aff, sentinel)
        radius = hp.last_radius()
    for c in range(cfg.channels):
        np.testing.assert_array_equal(got_m[c], want_m[c])
        kept_w = (want_p[c] == -3.0).all(1)
        np.testing.assert_array_equal((got_p[c] == -3.0).all(1), kept_w)
        if mode == "softmax":
            np.testing.assert_allclose(got_p[c], want_p[c], rtol=1e-6, atol=0)
        else:
            np.testing.assert_array_equal(got_p[c], want_p[c])
        assert radius[c] == o.last_radius(c)
    _say(f"  radius {radius}; cells keeping next_params: {[int((p == -3.0).all(1).sum()) for p in got_p]}")


def test_stencil_non_integer_exponents(oracle):
    """Mixed parameters give non-integer n, where alpha is a powf (one-ulp differences from
    libm's): the masses within 1e-6."""
    _say("[emb stencil] non-integer n")
    cfg = _cfg(mode="linear", seed=9)
    ch, ps = _state(cfg, seed=10, integer_n=False)
    o = oracle.EmbeddedPropagator(cfg, cfg.embedding.__dict__)
    aff = o.affinity(ch, ps)
    nxt = [p.copy() for p in ps]
    want_m, _ = o.stencil(ch, ps, aff, nxt)
    with HipEmbeddedPropagator(cfg) as hp:
        got_m, _ = hp.stencil(ch, ps, aff, nxt)
    err = max(float(np.abs(g - w).max()) for g, w in zip(got_m, want_m))
    _say(f"  max|gpu - oracle| mass = {err:.3e}")
    assert err <= 1e-6


@pytest.mark.parametrize("mode", ["softmax", "linear", "disabled"])
def test_trajectory_against_oracle(oracle, mode):
    _say(f"[emb trajectory] {mode}")
    cfg = _cfg(mode=mode, seed=21)
    ch, ps = _state(cfg, seed=22)
    n = cfg.width * cfg.height
    st = EmbeddedState(ch, [ParameterGrid(cfg.width, cfg.height, data=p) for p in ps], cfg.width, cfg.height)
    o = oracle.EmbeddedPropagator(cfg, cfg.embedding.__dict__)
    errs = {}
    with HipEmbeddedPropagator(cfg) as hp:
        hp.set_state(st)
        done = 0
        och, ops = ch, ps
        for k in (20, 100):
            hp.advance(k - done)
            och, ops, _, _ = o.run(och, ops, k - done)
            done = k
            g = hp.get_state()
            errs[k] = (max(float(np.abs(g.channels[c] - och[c]).max()) for c in range(cfg.channels)),
                       max(float(np.abs(g.params[c].data - ops[c]).max()) for c in range(cfg.channels)))
        _, mass = hp.mass_f64()
    m0 = sum(float(np.asarray(c, np.float64).sum()) for c in ch)
    _say(f"  max|gpu - oracle| (mass, params): 20 steps {errs[20][0]:.3e}, {errs[20][1]:.3e}; "
         f"100 steps {errs[100][0]:.3e}, {errs[100][1]:.3e}; FP64 mass {m0:.6f} -> {mass:.6f}")
    assert g.step == 100
    assert errs[20][0] <= 1e-4


def test_step_and_run_agree_and_resident_state(oracle):
    """hp_emb_step on host buffers twice equals hp_emb_run of 2 steps and a resident advance."""
    _say("[emb step/run]")
    cfg = _cfg(mode="softmax", seed=31)
    ch, ps = _state(cfg, seed=32)
    mk = lambda: EmbeddedState([c.copy() for c in ch], [ParameterGrid(cfg.width, cfg.height, data=p.copy())
                                                        for p in ps], cfg.width, cfg.height)
    a, b = mk(), mk()
    with HipEmbeddedPropagator(cfg) as hp:
        hp.step(a)
        hp.step(a)
    with HipEmbeddedPropagator(cfg) as hp:
        hp.run(b, 2)
    with HipEmbeddedPropagator(cfg) as hp:
        hp.set_state(mk())
        hp.advance(2)
        c = hp.get_state()
    assert a.step == b.step == c.step == 2 and abs(a.time - 0.4) < 1e-6
    for i in range(cfg.channels):
        np.testing.assert_array_equal(a.channels[i], b.channels[i])
        np.testing.assert_array_equal(a.channels[i], c.channels[i])
        np.testing.assert_array_equal(a.params[i].data, c.params[i].data)


def test_reference_embedded_propagator_tests():
    """embedded_propagator.rs:398-452 on the device: mass within 5% over 10 steps; custom
    parameters at the blob keep a positive mu; from_seed's defaults come from the config."""
    _say("[emb reference tests]")
    cfg = SimulationConfig(width=32, height=32, depth=1, channels=1, dt=0.2, kernel_radius=5,
                           kernels=[KernelConfig(1.0, [RingConfig(1.0, 0.5, 0.15)], 1.0, 0.15, 0.015, 0, 0)],
                           flow=FlowConfig(1.0, 2.0, 1.0), embedding=EmbeddingConfig.enabled_softmax())
    seed = Seed(Pattern("GaussianBlob", center=(0.5, 0.5), radius=0.15, amplitude=1.0, channel=0))
    st = EmbeddedState.from_seed(seed, cfg)
    assert st.params[0].get(3, 3) == default_params(cfg) == CellParams(0.15, 0.015, 1.0, 1.0, 2.0)
    m0 = st.total_mass()
    with HipEmbeddedPropagator(cfg) as hp:
        hp.run(st, 10)
    assert abs(st.total_mass() - m0) / m0 < 0.05
    seed = Seed(Pattern("GaussianBlob", center=(0.25, 0.5), radius=0.1, amplitude=1.0, channel=0))
    st = EmbeddedState.from_seed(seed, cfg)
    st.params[0].set(8, 16, CellParams(0.3, 0.03, 2.0, 1.5, 3.0))
    with HipEmbeddedPropagator(cfg) as hp:
        hp.run(st, 5)
    assert st.get_params(8, 16, 0).mu > 0.0


def test_stage_entry_reference_tests_and_closed_form(oracle):
    """param_advection.rs:364-457 through hp_advect_mass_and_params, the duplicate-visit closed
    form, and random inputs bit for bit against the oracle (linear mixing)."""
    _say("[emb stage entry]")
    lin = EmbeddingConfig.enabled_linear()
    w = h = 16
    m = np.zeros(w * h, np.float32)
    m[8 * w + 8] = 1.0
    pg = ParameterGrid.from_defaults(w, h)
    pg.set(8, 8, CellParams(0.25, 0.025, 2.0, 1.5, 3.0))
    z = np.zeros(w * h, np.float32)
    nm, npg = advect_mass_and_params(m, pg, z, z, EmbeddingConfig.enabled_softmax(), 0.2, 1.0, w, h)
    assert abs(float(nm.sum()) - 1.0) < 0.01 and abs(npg.get(8, 8).mu - 0.25) < 0.1
    m = np.zeros(w * h, np.float32)
    m[8 * w + 4] = m[8 * w + 10] = 1.0
    pg = ParameterGrid.from_defaults(w, h)
    pg.set(4, 8, CellParams(0.1, 0.01, 1.0, 1.0, 2.0))
    pg.set(10, 8, CellParams(0.3, 0.03, 1.0, 1.0, 2.0))
    fx = np.zeros(w * h, np.float32)
    fx[8 * w + 4], fx[8 * w + 10] = 15.0, -15.0
    nm, npg = advect_mass_and_params(m, pg, fx, z, lin, 0.2, 1.0, w, h)
    assert abs(float(nm.sum()) - 2.0) < 0.01
    i = int(np.argmax(nm))
    if nm[i] > 0.5:
        assert 0.1 < npg.get_idx(i).mu < 0.3
    # duplicate visits on 3 x 3
    m = np.zeros(9, np.float32)
    m[4] = 1.0
    nm, _ = advect_mass_and_params(m, ParameterGrid.from_defaults(3, 3), np.zeros(9, np.float32),
                                   np.zeros(9, np.float32), lin, 0.2, 1.0, 3, 3)
    nv, ov = np.array([2, 1, 2]), np.array([1, 1, 0])
    np.testing.assert_array_equal(nm, (np.outer(nv * ov, nv * ov) / 4.0).astype(np.float32).reshape(-1))
    # random, bit for bit
    rng = np.random.default_rng(4)
    w, h = 33, 20
    n = w * h
    m = rng.random(n).astype(np.float32) * (rng.random(n) < 0.5)
    pr = rng.uniform(0.01, 3.0, (n, 5)).astype(np.float32)
    fx = (rng.standard_normal(n) * 3).astype(np.float32)
    fy = (rng.standard_normal(n) * 3).astype(np.float32)
    sent = np.full((n, 5), -1.0, np.float32)
    want_m, want_p, _ = oracle.advect_params(m, pr, fx, fy, w, h, 0.4, 0.9, linear=True, next_params=sent)
    got_m, got_p = advect_mass_and_params(m, ParameterGrid(w, h, data=pr), fx, fy, lin, 0.4, 0.9, w, h,
                                          next_params=ParameterGrid(w, h, data=sent))
    np.testing.assert_array_equal(got_m, want_m)
    np.testing.assert_array_equal(got_p.data, want_p)


@pytest.mark.parametrize("mode", ["softmax", "linear"])
def test_flow_beyond_work_bound_is_refused(mode):
    """A time step so large that the mixing gather's search radius (which grows with max|F| dt)
    passes its work bound: the run is refused with HP_ERR_UNSUPPORTED_SIZE and the resident
    state is the one before it.  (With embedding disabled the reintegration clamps its boxes to
    [-W, 2W] as the reference does, so any time step runs.)"""
    _say(f"[emb refused] {mode}")
    cfg = _cfg(mode=mode, seed=41, dt=1e7)
    ch, ps = _state(cfg, seed=42)
    st = EmbeddedState(ch, [ParameterGrid(cfg.width, cfg.height, data=p) for p in ps], cfg.width, cfg.height)
    with HipEmbeddedPropagator(cfg) as hp:
        hp.set_state(st)
        with pytest.raises(_abi.HipError) as e:
            hp.advance(3)
        assert e.value.status == _abi.HP_ERR_UNSUPPORTED_SIZE
        g = hp.get_state()
    assert g.step == 0
    for c in range(cfg.channels):
        np.testing.assert_array_equal(g.channels[c], ch[c])
        np.testing.assert_array_equal(g.params[c].data, ps[c])


@pytest.mark.parametrize("linear", [True, False])
def test_stage_entry_many_contributors(oracle, linear):
    """s = 2: a destination receives from up to 25 sources, more than the tiled gather keeps in
    LDS, so it recomputes them; dense masses on 40 x 32 (radius 4: the tiled path)."""
    _say(f"[emb many contributors] linear={linear}")
    rng = np.random.default_rng(51 + linear)
    w, h = 40, 32
    n = w * h
    m = rng.random(n).astype(np.float32)
    pr = rng.uniform(0.01, 3.0, (n, 5)).astype(np.float32)
    fx = rng.uniform(-1, 1, n).astype(np.float32)
    fy = rng.uniform(-1, 1, n).astype(np.float32)
    sent = np.full((n, 5), -1.0, np.float32)
    emb = EmbeddingConfig(True, 0.5, linear)
    want_m, want_p, R = oracle.advect_params(m, pr, fx, fy, w, h, 0.2, 2.0, linear=linear, temperature=0.5,
                                             next_params=sent)
    assert R == 4
    got_m, got_p = advect_mass_and_params(m, ParameterGrid(w, h, data=pr), fx, fy, emb, 0.2, 2.0, w, h,
                                          next_params=ParameterGrid(w, h, data=sent))
    np.testing.assert_array_equal(got_m, want_m)
    if linear:
        np.testing.assert_array_equal(got_p.data, want_p)
    else:
        np.testing.assert_allclose(got_p.data, want_p, rtol=1e-6, atol=0)
