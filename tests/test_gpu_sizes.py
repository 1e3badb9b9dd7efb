"""Grid sides outside the FFT passes' lengths (2^a * 3^b, 8 .. 4096): the direct convolution path
(csrc/direct_conv.hip), and the two largest FFT lengths, against the f32 oracle (whose FFT
plans every length, as rustfft does in the reference).

Bars: the affinity within 2e-5 of the oracle's (the direct sum and the FFT agree to rounding);
whole trajectories within 1e-4 after 100 steps (2D) and 50 steps (3D), the reference's own bar
for its FFT tests; the error against the FP64 oracle is printed beside it."""
import numpy as np
import pytest

import autoverse_amd as av
from autoverse_amd import _abi, workloads
from autoverse_amd.population import HipPopulation
from autoverse_amd.schema import FlowConfig, KernelConfig, Pattern, RingConfig, Seed, SimulationConfig

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _cfg2d(W, H, C=2, R=13, seed=1):
    rng = np.random.default_rng(seed)
    ks = workloads._random_kernels(rng, C)
    return SimulationConfig(width=W, height=H, depth=1, channels=C, dt=0.2, kernel_radius=R, kernels=ks,
                            flow=FlowConfig(1.0, 2.0, 1.0)), \
        Seed(Pattern("MultiBlob", blobs=workloads._random_blobs(rng, C, 4)))


def _cfg3d(W, H, D, R=10):
    cfg = SimulationConfig(width=W, height=H, depth=D, channels=1, dt=0.1, kernel_radius=R,
                           kernels=[KernelConfig(1.0, [RingConfig(1.0, 0.5, 0.15)], 1.0, 0.15, 0.017, 0, 0)],
                           flow=FlowConfig(1.0, 2.0, 1.0))
    return cfg, Seed(Pattern("GaussianSphere", center=(0.5, 0.5, 0.5), radius=0.2, amplitude=1.0, channel=0))


def _check(oracle, cfg, seed, steps, tag):
    st0 = oracle.seed_generate(seed, cfg.width, cfg.height, cfg.depth, cfg.channels)
    p32 = oracle.Propagator(cfg, np.float32, threads=8, line_threads=8)
    warm, _, _ = p32.run(st0, 5)
    want_u = p32.affinity(warm)
    with av.HipPropagator(cfg) as hp:
        got_u = hp.affinity(warm)
        state = av.SimulationState([c.copy() for c in st0], cfg.width, cfg.height, cfg.depth)
        hp.set_state(state)
        hp.advance(steps)
        got = hp.get_state()
    eu = max(float(np.abs(g - w).max()) for g, w in zip(got_u, want_u))
    want, _, _ = p32.run(st0, steps)
    w64, _, _ = oracle.Propagator(cfg, np.float64, threads=8, line_threads=8).run(
        [np.asarray(c, np.float64) for c in st0], steps)
    e32 = max(float(np.abs(got.channels[c] - want[c]).max()) for c in range(cfg.channels))
    e64 = max(float(np.abs(got.channels[c] - w64[c]).max()) for c in range(cfg.channels))
    _say(f"  {tag}: affinity {eu:.3e}; after {steps} steps max|gpu - f32 oracle| {e32:.3e}, "
         f"max|gpu - FP64| {e64:.3e}")
    assert eu <= 2e-5
    assert e32 <= 1e-4
    assert got.step == steps


@pytest.mark.parametrize("W,H", [(100, 100), (320, 320), (97, 61)])
def test_direct_path_2d(oracle, W, H):
    _say(f"[sizes 2D direct] {W}x{H}")
    cfg, seed = _cfg2d(W, H, seed=W + H)
    _check(oracle, cfg, seed, 100, f"{W}x{H}")


@pytest.mark.parametrize("W,H", [(3072, 64), (4096, 48)])
def test_long_fft_sides(oracle, W, H):
    """The FFT lengths above 2048, on a noise state whose convolution sits near the growth
    centre (so the growth is not saturated anywhere)."""
    _say(f"[sizes FFT] {W}x{H}")
    cfg = SimulationConfig(width=W, height=H, depth=1, channels=1, dt=0.2, kernel_radius=10,
                           kernels=[KernelConfig(1.0, [RingConfig(1.0, 0.5, 0.15)], 1.0, 0.15, 0.05, 0, 0)],
                           flow=FlowConfig(1.0, 2.0, 1.0))
    seed = Seed(Pattern("Noise", amplitude=0.3, channel=0, seed=W))
    _check(oracle, cfg, seed, 20, f"{W}x{H}")


@pytest.mark.parametrize("W,H,D", [(80, 80, 80), (40, 36, 30)])
def test_direct_path_3d(oracle, W, H, D):
    _say(f"[sizes 3D direct] {W}x{H}x{D}")
    cfg, seed = _cfg3d(W, H, D)
    _check(oracle, cfg, seed, 50, f"{W}x{H}x{D}")


def test_direct_path_embedded(oracle):
    """The embedded-parameter step on a 100 x 100 grid (convolutions from the direct path)."""
    from autoverse_amd.embedded import EmbeddedState, HipEmbeddedPropagator, ParameterGrid
    from autoverse_amd.schema import EmbeddingConfig
    _say("[sizes embedded 100^2]")
    cfg, seed = _cfg2d(100, 100, seed=7)
    cfg.embedding = EmbeddingConfig(True, 1.0, True)
    ch = oracle.seed_generate(seed, 100, 100, 1, 2)
    ps = [np.tile(np.array([0.2, 0.03, 0.8, 1.0, 2.0], np.float32), (10000, 1)) for _ in range(2)]
    o = oracle.EmbeddedPropagator(cfg, cfg.embedding.__dict__, threads=8)
    want_u = o.affinity(ch, ps)
    want, _, _, _ = o.run(ch, ps, 20)
    st = EmbeddedState(ch, [ParameterGrid(100, 100, data=p) for p in ps], 100, 100)
    with HipEmbeddedPropagator(cfg) as hp:
        got_u = hp.affinity(ch, ps)
        hp.run(st, 20)
    eu = max(float(np.abs(g - w).max()) for g, w in zip(got_u, want_u))
    e = max(float(np.abs(st.channels[c] - want[c]).max()) for c in range(2))
    _say(f"  affinity {eu:.3e}; 20 steps {e:.3e}")
    assert eu <= 2e-5 and e <= 1e-4


def test_population_refuses_direct_path_members():
    cfg, _ = _cfg2d(100, 100)
    ok, _ = _cfg2d(64, 64)
    with pytest.raises(_abi.HipError) as e:
        HipPopulation([ok, cfg])
    assert e.value.status == _abi.HP_ERR_UNSUPPORTED_SIZE
