"""Pins of the CPU oracle (oracle/) -- CPU only.

Two kinds of check, chosen so that a plausible mistake in the oracle (a dropped term,
a wrong sign or index, a transposed operand) fails at least one:
  * the reference crate's own unit tests, re-asserted on the oracle with the same
    inputs and tolerances (file:line of each is given);
  * what the mathematics fixes, computed independently of the oracle: numpy.fft in
    float64, brute-force periodic convolution, an independent Python scatter on tiny
    grids, closed forms and invariants.
"""
import math

import numpy as np
import pytest

from autoverse_amd.schema import BlobSpec, FlowConfig, KernelConfig, Pattern, RingConfig, Seed, SimulationConfig


def default_kernel():
    return KernelConfig(1.0, [RingConfig(1.0, 0.5, 0.15)], 1.0, 0.15, 0.015, 0, 0)


# ------------------------------------------------------------------ kernel (kernel.rs:143-182)
def test_kernel_normalization(oracle):
    k = oracle.kernel2d(default_kernel(), 10)
    assert abs(float(k.sum(dtype=np.float32)) - 1.0) < 1e-5


def test_kernel_symmetry(oracle):
    k = oracle.kernel2d(KernelConfig(), 10)
    c = k.shape[0] // 2
    for d in range(1, c):
        v1, v2, v3, v4 = k[c, c + d], k[c, c - d], k[c + d, c], k[c - d, c]
        assert abs(v1 - v2) < 1e-6 and abs(v1 - v3) < 1e-6 and abs(v1 - v4) < 1e-6


def test_kernel_matches_closed_form(oracle):
    """Ring sum a*exp(-(r/R - d)^2 / (2 w^2)) inside r/R <= 1, normalised -- float64 numpy."""
    kc = KernelConfig(0.8, [RingConfig(0.7, 0.3, 0.1), RingConfig(0.4, 0.8, 0.05)], 1.0, 0.2, 0.02, 0, 0)
    R = round(0.8 * 15)
    k = oracle.kernel2d(kc, 15)
    assert k.shape == (2 * R + 1, 2 * R + 1)
    y, x = np.mgrid[-R:R + 1, -R:R + 1]
    nd = np.sqrt(x * x + y * y) / R
    v = sum(r.amplitude * np.exp(-(nd - r.distance) ** 2 / (2 * r.width ** 2)) for r in kc.rings)
    v[nd > 1] = 0
    v /= v.sum()
    np.testing.assert_allclose(k, v, rtol=2e-5, atol=1e-9)
    k3 = oracle.kernel3d(kc, 6)
    R3 = round(0.8 * 6)
    z, y, x = np.mgrid[-R3:R3 + 1, -R3:R3 + 1, -R3:R3 + 1]
    nd = np.sqrt(x * x + y * y + z * z) / R3
    v = sum(r.amplitude * np.exp(-(nd - r.distance) ** 2 / (2 * r.width ** 2)) for r in kc.rings)
    v[nd > 1] = 0
    v /= v.sum()
    np.testing.assert_allclose(k3, v, rtol=2e-5, atol=1e-9)


def test_pad_places_centre_at_origin(oracle):
    k = oracle.kernel2d(default_kernel(), 5)
    p = oracle.pad2d(k, 16, 12)
    for dy in range(-5, 6):
        for dx in range(-5, 6):
            assert p[dy % 12, dx % 16] == k[dy + 5, dx + 5]
    assert p.sum(dtype=np.float64) == pytest.approx(k.sum(dtype=np.float64), rel=1e-12)


# ------------------------------------------------------------------ growth (growth.rs:66-199)
def test_growth_known_answers(oracle):
    assert abs(oracle.growth(0.15, 0.15, 0.015) - 1.0) < 1e-6
    assert abs(oracle.growth(1.0, 0.15, 0.015) + 1.0) < 0.01
    assert abs(oracle.growth(0.4, 0.5, 0.1) - oracle.growth(0.6, 0.5, 0.1)) < 1e-6
    for i in range(1000):
        g = oracle.growth(i * 0.01 - 5.0, 0.5, 0.1)
        assert -1.0 - 1e-6 <= g <= 1.0 + 1e-6
    assert oracle.growth(0.6, 0.5, 0.2) > oracle.growth(0.6, 0.5, 0.05)
    conv = np.array([0.1, 0.15, 0.2, 0.3], dtype=np.float32)
    t = np.ones(4, dtype=np.float32)
    oracle.growth_accumulate(conv, t, 0.5, 0.15, 0.015)
    for i in range(4):
        assert abs(t[i] - (1.0 + 0.5 * oracle.growth(float(conv[i]), 0.15, 0.015))) < 1e-6
    # closed form in float64
    u = np.linspace(-1, 1, 101)
    g = [oracle.growth(float(x), 0.15, 0.015, dtype=np.float64) for x in u]
    np.testing.assert_allclose(g, 2 * np.exp(-(u - 0.15) ** 2 / (2 * 0.015 ** 2)) - 1, rtol=0, atol=1e-14)


# ------------------------------------------------------------------ Sobel (gradient.rs:115-339)
def numpy_sobel(g):
    tl = np.roll(np.roll(g, 1, 0), 1, 1)
    tc = np.roll(g, 1, 0)
    tr = np.roll(np.roll(g, 1, 0), -1, 1)
    ml = np.roll(g, 1, 1)
    mr = np.roll(g, -1, 1)
    bl = np.roll(np.roll(g, -1, 0), 1, 1)
    bc = np.roll(g, -1, 0)
    br = np.roll(np.roll(g, -1, 0), -1, 1)
    two, e = g.dtype.type(2), g.dtype.type(0.125)
    gx = (-tl + tr - two * ml + two * mr - bl + br) * e
    gy = (-tl - two * tc - tr + bl + two * bc + br) * e
    return gx, gy


def test_sobel_known_answers(oracle):
    gx, gy = oracle.sobel2d(np.ones((16, 16), np.float32))
    assert np.abs(gx).max() < 1e-6 and np.abs(gy).max() < 1e-6
    ramp = np.tile(np.arange(16, dtype=np.float32), (16, 1))
    gx, _ = oracle.sobel2d(ramp)
    assert (gx[1:-1, 1:-1] > 0).all()
    _, gy = oracle.sobel2d(ramp.T.copy())
    assert (gy[1:-1, :] > 0).all()


def test_sobel_matches_numpy_bitwise(oracle):
    rng = np.random.default_rng(0)
    g = rng.normal(size=(23, 31)).astype(np.float32)
    gx, gy = oracle.sobel2d(g)
    nx, ny = numpy_sobel(g)
    np.testing.assert_array_equal(gx, nx)
    np.testing.assert_array_equal(gy, ny)
    # a linear ramp in float64 has the exact slope away from the wrap
    xx = np.tile(np.arange(20, dtype=np.float64) * 0.5, (9, 1))
    gx, gy = oracle.sobel2d(xx)
    np.testing.assert_allclose(gx[:, 1:-1], 0.5, rtol=0, atol=1e-15)
    np.testing.assert_allclose(gy, 0.0, atol=0)


def test_sobel3d_matches_numpy(oracle):
    rng = np.random.default_rng(1)
    g = rng.normal(size=(6, 7, 9))
    gx, gy, gz = oracle.sobel3d(g)
    s = np.array([1.0, 2.0, 1.0])
    d = np.array([-1.0, 0.0, 1.0])

    def sep(wz, wy, wx):
        out = np.zeros_like(g)
        for iz in range(3):
            for iy in range(3):
                for ix in range(3):
                    out += wz[iz] * wy[iy] * wx[ix] * np.roll(g, (1 - iz, 1 - iy, 1 - ix), (0, 1, 2))
        return out / 32

    np.testing.assert_allclose(gx, sep(s, s, d), atol=1e-12)
    np.testing.assert_allclose(gy, sep(s, d, s), atol=1e-12)
    np.testing.assert_allclose(gz, sep(d, s, s), atol=1e-12)


# ------------------------------------------------------------------ alpha / flow (flow.rs:159-379)
def test_alpha_and_flow_known_answers(oracle):
    d = FlowConfig()
    assert oracle.alpha(0.0, d.beta_a, d.n) == 0.0
    assert oracle.alpha(d.beta_a, d.beta_a, d.n) == 1.0
    assert oracle.alpha(2 * d.beta_a, d.beta_a, d.n) == 1.0
    a = [oracle.alpha(f, 1.0, 2.0) for f in (0.0, 0.25, 0.5, 0.75, 1.0)]
    assert a == sorted(a) and len(set(a)) == 5
    assert abs(a[2] - 0.25) < 1e-6
    assert abs(oracle.alpha(0.5, 1.0, 1.0) - 0.5) < 1e-6
    assert oracle.alpha(0.5, 1.0, 4.0) < oracle.alpha(0.5, 1.0, 2.0)
    assert oracle.alpha(0.5, 1.0, 2.5, dtype=np.float64) == pytest.approx(0.5 ** 2.5, rel=1e-15)
    gux = np.array([1, 2, 3], np.float32)
    guy = np.array([0.5, 1, 1.5], np.float32)
    fx, fy = oracle.flow2d(gux, guy, -gux, -guy, np.zeros(3, np.float32), 2.0, 4.0)
    np.testing.assert_allclose(fx, gux, atol=1e-6)
    np.testing.assert_allclose(fy, guy, atol=1e-6)
    fx, fy = oracle.flow2d(gux, guy, -gux, -guy, np.full(3, 20.0, np.float32), 2.0, 4.0)
    np.testing.assert_allclose(fx, gux, atol=1e-6)  # -grad_A with grad_A = -grad_U
    np.testing.assert_allclose(fy, guy, atol=1e-6)
    m = np.array([0.3, 0.9, 1.4], np.float32)
    fx, _ = oracle.flow2d(gux, guy, 2 * gux, guy, m, 1.0, 2.0)
    al = np.clip((m.astype(np.float64) / 1.0) ** 2, 0, 1)
    np.testing.assert_allclose(fx, (1 - al) * gux - al * 2 * gux, rtol=1e-6)


# ------------------------------------------------------------------ reintegration
def test_wrap_coord(oracle):
    for c, s, want in [(0, 10, 0), (5, 10, 5), (10, 10, 0), (-1, 10, 9), (-10, 10, 0), (15, 10, 5)]:
        assert oracle.wrap_coord(c, s) == want


def _mass(a):
    return float(np.add.accumulate(a.ravel(), dtype=np.float32)[-1])


@pytest.mark.parametrize("case", ["no_flow", "with_flow", "boundary", "corner", "large", "small_s"])
def test_reintegration_reference_cases(oracle, case):
    """reintegration.rs:235-470, same inputs and bars."""
    if case == "no_flow":
        w = h = 16; cur = np.zeros((h, w), np.float32); cur[8, 8] = 1; cur[7, 8] = 0.5
        fx, fy, dt, s, tol = 0.0, 0.0, 0.2, 1.0, 1e-5
    elif case == "with_flow":
        w = h = 32; cur = np.zeros((h, w), np.float32); cur[10:20, 10:20] = 1
        fx, fy, dt, s, tol = 5.0, 2.0, 0.2, 1.0, 1e-4
    elif case == "boundary":
        w = h = 16; cur = np.zeros((h, w), np.float32); cur[8, 14] = 1
        fx, fy, dt, s, tol = 5.0, 0.0, 0.5, 1.0, 1e-5
    elif case == "corner":
        w = h = 16; cur = np.zeros((h, w), np.float32); cur[15, 15] = 1
        fx, fy, dt, s, tol = 8.0, 8.0, 0.5, 1.0, 1e-5
    elif case == "large":
        w = h = 16; cur = np.zeros((h, w), np.float32); cur[8, 8] = 1
        fx, fy, dt, s, tol = 100.0, 50.0, 0.2, 1.0, 1e-5
    else:
        w = h = 16; cur = np.zeros((h, w), np.float32); cur[8, 8] = 1
        fx, fy, dt, s, tol = 1.0, 0.0, 0.2, 0.01, 1e-4
    nxt = oracle.advect2d(cur, np.full((h, w), fx, np.float32), np.full((h, w), fy, np.float32), dt, s)
    assert abs(_mass(cur) - _mass(nxt)) < tol
    if case == "boundary":
        assert nxt[8, 0:3].sum() > 0.1 and nxt[8, 14] < 1.0


def python_scatter(cur, fx, fy, dt, s):
    """Independent restatement of distribute_mass in numpy float32 scalars (tiny grids only)."""
    f = np.float32
    h, w = cur.shape
    nxt = np.zeros_like(cur)
    ta = (f(2) * f(s)) * (f(2) * f(s))
    for y in range(h):
        for x in range(w):
            m = cur[y, x]
            if abs(m) < f(1e-10):
                continue
            dx = f(x) + f(dt) * fx[y, x]
            dy = f(y) + f(dt) * fy[y, x]
            x0, x1, y0, y1 = dx - f(s), dx + f(s), dy - f(s), dy + f(s)
            for iy in range(max(math.floor(y0), -h), min(math.ceil(y1), 2 * h) + 1):
                for ix in range(max(math.floor(x0), -w), min(math.ceil(x1), 2 * w) + 1):
                    ow = max(min(f(ix + 1), x1) - max(f(ix), x0), f(0))
                    oh = max(min(f(iy + 1), y1) - max(f(iy), y0), f(0))
                    a = f(ow * oh)
                    if a > 0:
                        nxt[iy % h, ix % w] = f(nxt[iy % h, ix % w] + f(m * f(a / ta)))
    return nxt


@pytest.mark.parametrize("dt,s,scale", [(0.2, 1.0, 1.0), (0.7, 0.5, 3.0), (1.0, 1.3, 6.0), (0.2, 0.01, 1.0)])
def test_reintegration_matches_independent_scatter(oracle, dt, s, scale):
    rng = np.random.default_rng(int(dt * 10 + s * 100))
    h, w = 9, 11
    cur = rng.uniform(0, 1, (h, w)).astype(np.float32)
    cur[rng.uniform(size=(h, w)) < 0.3] = 0
    fx = rng.normal(0, scale, (h, w)).astype(np.float32)
    fy = rng.normal(0, scale, (h, w)).astype(np.float32)
    np.testing.assert_array_equal(oracle.advect2d(cur, fx, fy, dt, s), python_scatter(cur, fx, fy, dt, s))


def test_reintegration_conserves_and_translates(oracle):
    """FP64 scatter conserves mass to rounding; shifting input and flow by whole cells shifts
    the output (up to the absolute-frame rounding)."""
    rng = np.random.default_rng(5)
    cur = rng.uniform(0, 1, (24, 20))
    fx, fy = rng.normal(0, 2, (24, 20)), rng.normal(0, 2, (24, 20))
    nxt = oracle.advect2d(cur, fx, fy, 0.4, 1.0)
    assert abs(nxt.sum() - cur.sum()) < 1e-12 * cur.sum() * 50
    sh = (3, 5)
    nxt2 = oracle.advect2d(np.roll(cur, sh, (0, 1)), np.roll(fx, sh, (0, 1)), np.roll(fy, sh, (0, 1)), 0.4, 1.0)
    np.testing.assert_allclose(nxt2, np.roll(nxt, sh, (0, 1)), atol=1e-12)
    # point fallback: every source lands whole on round(dest)
    c = np.zeros((8, 8)); c[2, 3] = 1.5
    n = oracle.advect2d(c, np.full((8, 8), 1.2), np.full((8, 8), -3.6), 1.0, 1e-6)
    assert n[(2 - 4) % 8, (3 + 1) % 8] == 1.5 and n.sum() == 1.5


def test_reintegration3d_matches_2d_when_flat(oracle):
    """A 3D grid with zero z-flow and mass on plane 1, s = 0.5: each source's cube spans
    z in [0.5, 1.5], so planes 0 and 1 each receive half of the 2D reintegration (the
    square has area (2s)^2 = 1, the cube volume (2s)^3 = 1), and planes 2, 3 nothing."""
    rng = np.random.default_rng(3)
    d, h, w = 4, 10, 12
    cur = np.zeros((d, h, w)); cur[1] = rng.uniform(0, 1, (h, w))
    fx, fy = rng.normal(0, 1, (d, h, w)), rng.normal(0, 1, (d, h, w))
    fz = np.zeros((d, h, w))
    n3 = oracle.advect3d(cur, fx, fy, fz, 0.3, 0.5)
    n2 = oracle.advect2d(cur[1].copy(), fx[1].copy(), fy[1].copy(), 0.3, 0.5)
    np.testing.assert_allclose(n3[1], 0.5 * n2, atol=1e-14)
    np.testing.assert_allclose(n3[0], 0.5 * n2, atol=1e-14)
    assert np.abs(n3[[2, 3]]).max() == 0


# ------------------------------------------------------------------ FFT (fft.rs:493-787)
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 12, 16, 30, 49, 64, 96, 192, 243, 1000, 1024])
def test_fft_matches_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    np.testing.assert_allclose(oracle.fft(x, dtype=np.float64), np.fft.fft(x), atol=1e-11 * max(1, n))
    np.testing.assert_allclose(oracle.fft(x, True, dtype=np.float64), np.fft.ifft(x) * n, atol=1e-11 * max(1, n))
    y = oracle.fft(x, dtype=np.float32)
    assert np.abs(y - np.fft.fft(x)).max() <= 2e-6 * math.sqrt(n) * (1 + math.log2(max(n, 2)))


def test_fft2d_3d_match_numpy(oracle):
    rng = np.random.default_rng(9)
    x = rng.normal(size=(12, 20)) + 1j * rng.normal(size=(12, 20))
    np.testing.assert_allclose(oracle.fftn(x, dtype=np.float64), np.fft.fft2(x), atol=1e-11)
    np.testing.assert_allclose(oracle.fftn(x, True, dtype=np.float64), np.fft.ifft2(x) * x.size, atol=1e-11)
    x = rng.normal(size=(6, 8, 10)) + 1j * rng.normal(size=(6, 8, 10))
    np.testing.assert_allclose(oracle.fftn(x, dtype=np.float64), np.fft.fftn(x), atol=1e-11)


def _conv(oracle, inp, ker, dtype=np.float32):
    kf = oracle.fftn(ker.astype(np.complex128), dtype=dtype)
    buf = np.empty(2 * ker.size, dtype=dtype)
    buf[0::2] = kf.real.ravel()
    buf[1::2] = kf.imag.ravel()
    out = np.zeros(inp.size, dtype=dtype)
    shp = inp.shape if inp.ndim == 3 else (1,) + inp.shape
    getattr(oracle.lib(), f"orc_convolve_{'f32' if dtype == np.float32 else 'f64'}")(
        oracle._p(np.ascontiguousarray(inp, dtype=dtype).ravel()), oracle._p(buf), shp[2], shp[1], shp[0],
        oracle._p(out))
    return out.reshape(inp.shape)


def test_fft_convolution_identities(oracle):
    w = h = 16
    x = (np.arange(w * h) % 10).astype(np.float32).reshape(h, w)
    delta = np.zeros((h, w), np.float32); delta[0, 0] = 1
    np.testing.assert_allclose(_conv(oracle, x, delta), x, atol=1e-4)
    inp = np.zeros((h, w), np.float32); inp[5, 5] = 1
    k = np.zeros((h, w), np.float32); k[2, 3] = 1
    r = _conv(oracle, inp, k)
    assert abs(r[7, 8] - 1) < 1e-4 and abs(r.sum() - 1) < 1e-4
    inp = np.zeros((h, w), np.float32); inp[8, 8] = 1
    k = np.zeros((h, w), np.float32)
    for ky in range(3):
        for kx in range(3):
            k[0 if ky == 0 else h - (3 - ky), 0 if kx == 0 else w - (3 - kx)] = 1 / 9
    r = _conv(oracle, inp, k)
    assert abs(r.sum() - 1) < 1e-4 and 0.05 < r[8, 8] < 0.5
    a = ((np.arange(w * h) * 17) % 100 / 100).astype(np.float32).reshape(h, w)
    b = ((np.arange(w * h) * 31 + 7) % 100 / 100).astype(np.float32).reshape(h, w)
    np.testing.assert_allclose(_conv(oracle, a, b), _conv(oracle, b, a), atol=1e-4)
    for (ww, hh) in [(8, 8), (16, 16), (32, 32), (16, 32), (32, 16)]:
        x = (np.arange(ww * hh) % 10).astype(np.float32).reshape(hh, ww)
        d = np.zeros((hh, ww), np.float32); d[0, 0] = 1
        np.testing.assert_allclose(_conv(oracle, x, d), x, atol=1e-4)


def direct_periodic(inp, k):
    """sum_d inp[p - d] k[d] over the wrap-padded kernel, float64 brute force."""
    out = np.zeros_like(inp, dtype=np.float64)
    R = k.shape[0] // 2
    for dy in range(-R, R + 1):
        for dx in range(-R, R + 1):
            out += k[dy + R, dx + R] * np.roll(inp, (dy, dx), (0, 1))
    return out


def test_fft_vs_direct_convolution(oracle):
    """fft.rs:685-787: 64^2, kernel radius 7, Gaussian input, max diff < 1e-4 (f32); 1e-12 in f64."""
    kc = default_kernel()
    y, x = np.mgrid[0:64, 0:64].astype(np.float32)
    inp = np.exp(-((x - 32) ** 2 + (y - 32) ** 2) / np.float32(50.0)).astype(np.float32)
    k32 = oracle.kernel2d(kc, 7)
    got = _conv(oracle, inp, oracle.pad2d(k32, 64, 64))
    want = direct_periodic(inp.astype(np.float64), k32.astype(np.float64))
    assert np.abs(got - want).max() < 1e-4
    k64 = oracle.kernel2d(kc, 7, dtype=np.float64)
    got64 = _conv(oracle, inp.astype(np.float64), oracle.pad2d(k64, 64, 64), dtype=np.float64)
    assert np.abs(got64 - direct_periodic(inp.astype(np.float64), k64)).max() < 1e-12


def test_propagator_kernel_spectrum_is_real_even(oracle):
    cfg = SimulationConfig(width=48, height=32, kernel_radius=9, kernels=[default_kernel()],
                           flow=FlowConfig(1.0, 2.0, 1.0))
    p = oracle.Propagator(cfg, np.float64)
    kf = p.kernel_freq(0).reshape(32, 48)
    assert np.abs(kf.imag).max() < 1e-12
    np.testing.assert_allclose(kf.real, kf.real[::-1, :][np.r_[-1, 0:31], :], atol=1e-12)
    np.testing.assert_allclose(kf.real, kf.real[:, ::-1][:, np.r_[-1, 0:47]], atol=1e-12)
    k = oracle.kernel2d(default_kernel(), 9, dtype=np.float64)
    np.testing.assert_allclose(kf, np.fft.fft2(oracle.pad2d(k, 48, 32)), atol=1e-12)


# ------------------------------------------------------------------ propagator (propagator.rs:518-676)
def plain64_config():
    return SimulationConfig(width=64, height=64, depth=1, channels=1, dt=0.2, kernel_radius=10,
                            kernels=[default_kernel()], flow=FlowConfig(1.0, 2.0, 1.0))


def test_propagator_mass_conservation(oracle):
    cfg = plain64_config()
    seed = Seed(Pattern("GaussianBlob", center=(0.5, 0.5), radius=0.15, amplitude=1.0, channel=0))
    st = oracle.seed_generate(seed, 64, 64, 1, 1)
    m0 = oracle.total_mass_f32(st)
    assert m0 > 0
    p = oracle.Propagator(cfg)
    ch, t, s = p.run(st, 10)
    assert abs(oracle.total_mass_f32(ch) - m0) / m0 < 0.01
    ch, t, s = p.run(st, 100)
    assert abs(oracle.total_mass_f32(ch) - m0) / m0 < 0.001
    assert s == 100 and t == pytest.approx(20.0, rel=1e-6)


def test_propagator_multichannel_mass(oracle):
    cfg = plain64_config()
    cfg.channels = 2
    cfg.kernels = [default_kernel(), KernelConfig(1.0, [RingConfig(1.0, 0.5, 0.15)], 1.0, 0.15, 0.015, 1, 1)]
    seed = Seed(Pattern("MultiBlob", blobs=[BlobSpec((0.3, 0.5), 0.1, 1.0, 0), BlobSpec((0.7, 0.5), 0.1, 2.0, 1)]))
    st = oracle.seed_generate(seed, 64, 64, 1, 2)
    m = [float(np.add.accumulate(c, dtype=np.float32)[-1]) for c in st]
    ch, _, _ = oracle.Propagator(cfg).run(st, 20)
    for i in range(2):
        assert abs(float(np.add.accumulate(ch[i], dtype=np.float32)[-1]) - m[i]) / m[i] < 0.01


def test_propagator_f32_tracks_f64(oracle):
    """The f32 oracle stays within 1e-4 of the same algorithm in float64 after 100 steps on the
    plain 256^2 configuration (so a 1e-4 GPU-vs-oracle bar is meaningful)."""
    from autoverse_amd import workloads
    cfg, seed = workloads.plain_256()
    st = oracle.seed_generate(seed, 256, 256, 1, 1)
    a, _, _ = oracle.Propagator(cfg, np.float32).run(st, 100)
    b, _, _ = oracle.Propagator(cfg, np.float64).run([st[0].astype(np.float64)], 100)
    assert np.abs(a[0] - b[0]).max() < 1e-4


def test_threads_do_not_change_results(oracle):
    from autoverse_amd import workloads
    cfg, seed = workloads.multispecies(size=64, channels=3, seed=2)
    st = oracle.seed_generate(seed, 64, 64, 1, 3)
    a, _, _ = oracle.Propagator(cfg, np.float32, threads=1).run(st, 3)
    b, _, _ = oracle.Propagator(cfg, np.float32, threads=8).run(st, 3)
    for i in range(3):
        np.testing.assert_array_equal(a[i], b[i])


# ------------------------------------------------------------------ seeds (seed.rs:160-524)
def test_seed_gaussian_closed_form(oracle):
    seed = Seed(Pattern("GaussianBlob", center=(0.5, 0.25), radius=0.1, amplitude=0.8, channel=0))
    g = oracle.seed_generate(seed, 40, 60, 1, 1)[0].reshape(60, 40)
    y, x = np.mgrid[0:60, 0:40]
    r = 0.1 * 40
    want = 0.8 * np.exp(-((x - 20.0) ** 2 + (y - 15.0) ** 2) / (2 * (r / 2) ** 2))
    np.testing.assert_allclose(g, want, rtol=1e-6, atol=1e-30)
    assert g[15, 20] == np.float32(0.8)  # integer cell coordinates, no +0.5 offset


def test_seed_noise_lcg(oracle):
    seed = Seed(Pattern("Noise", amplitude=0.5, channel=None, seed=42))
    out = oracle.seed_generate(seed, 4, 3, 1, 2)
    s = 42
    for c in range(2):
        for i in range(12):
            s = (s * 6364136223846793005 + 1) % (1 << 64)
            assert out[c][i] == np.float32(0.5) * (np.float32(s >> 33) / np.float32(2 ** 31))


def test_center_of_mass(oracle):
    g = np.zeros((10, 12), np.float32)
    g[2, 3] = 1.0
    g[6, 9] = 3.0
    cx, cy = oracle.center_of_mass(g)
    assert cx == pytest.approx((3 + 27) / 4) and cy == pytest.approx((2 + 18) / 4)
    assert oracle.center_of_mass(np.zeros((10, 12), np.float32)) == (6.0, 5.0)
    assert oracle.locomotion(0, 0, 3, 4, 30, 40) == pytest.approx(5 / 50)
