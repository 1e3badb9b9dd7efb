"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Only the tests, the smoke check and the benchmark's
CPU-baseline leg import this module, always as the checker; the product package
``autoverse_amd`` never imports it.  The C sources (flowlenia_oracle.c /
oracle_impl.inc) restate the reference crate's CPU path; see
flowlenia_oracle.h for the file:line map.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_int32, c_uint32, c_uint64, c_void_p

import numpy as np

from autoverse_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc)."""
    if force and os.path.exists(LIB):
        os.remove(LIB)
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        FPf = POINTER(c_float)
        FPd = POINTER(c_double)
        for S, FPx, T in (("f32", FPf, c_float), ("f64", FPd, c_double)):
            PP = POINTER(FPx)
            sig = {
                f"orc_kernel2d_{S}": (c_uint32, [POINTER(_abi.HpKernel), POINTER(_abi.HpRing), c_uint32, FPx]),
                f"orc_kernel3d_{S}": (c_uint32, [POINTER(_abi.HpKernel), POINTER(_abi.HpRing), c_uint32, FPx]),
                f"orc_pad2d_{S}": (None, [FPx, c_uint32, c_uint32, c_uint32, FPx]),
                f"orc_pad3d_{S}": (None, [FPx, c_uint32, c_uint32, c_uint32, c_uint32, FPx]),
                f"orc_fft_{S}": (None, [FPx, c_uint32, c_int]),
                f"orc_fft2d_{S}": (None, [FPx, c_uint32, c_uint32, c_int]),
                f"orc_fft3d_{S}": (None, [FPx, c_uint32, c_uint32, c_uint32, c_int]),
                f"orc_growth_{S}": (T, [T, T, T]),
                f"orc_growth_accumulate_{S}": (None, [FPx, FPx, c_uint64, T, T, T]),
                f"orc_sobel2d_{S}": (None, [FPx, FPx, FPx, c_uint32, c_uint32]),
                f"orc_sobel3d_{S}": (None, [FPx, FPx, FPx, FPx, c_uint32, c_uint32, c_uint32]),
                f"orc_alpha_{S}": (T, [T, T, T]),
                f"orc_flow2d_{S}": (None, [FPx, FPx, FPx, FPx, FPx, c_uint64, T, T, FPx, FPx]),
                f"orc_flow3d_{S}": (None, [FPx] * 7 + [c_uint64, T, T, FPx, FPx, FPx]),
                f"orc_advect2d_{S}": (None, [FPx, FPx, FPx, FPx, c_uint32, c_uint32, T, T]),
                f"orc_advect3d_{S}": (None, [FPx] * 5 + [c_uint32, c_uint32, c_uint32, T, T]),
                f"orc_convolve_{S}": (None, [FPx, FPx, c_uint32, c_uint32, c_uint32, FPx]),
                f"orc_prop_create_{S}": (c_void_p, [POINTER(_abi.HpConfig), c_int]),
                f"orc_prop_destroy_{S}": (None, [c_void_p]),
                f"orc_prop_set_threads_{S}": (None, [c_void_p, c_int]),
                f"orc_prop_set_line_threads_{S}": (None, [c_void_p, c_int]),
                f"orc_prop_set_conserve_{S}": (None, [c_void_p, c_int]),
                f"orc_advect2d_mode_{S}": (None, [FPx, FPx, FPx, FPx, c_uint32, c_uint32, T, T, c_int]),
                f"orc_advect3d_mode_{S}": (None, [FPx] * 5 + [c_uint32, c_uint32, c_uint32, T, T, c_int]),
                f"orc_prop_kernel_freq_{S}": (FPx, [c_void_p, c_uint32]),
                f"orc_prop_affinity_{S}": (None, [c_void_p, PP, PP]),
                f"orc_prop_stencil_{S}": (None, [c_void_p, PP, PP, PP]),
                f"orc_prop_step_{S}": (None, [c_void_p, PP, POINTER(c_float), POINTER(c_uint64)]),
                f"orc_prop_run_{S}": (None, [c_void_p, PP, POINTER(c_float), POINTER(c_uint64), c_uint64]),
                f"orc_prop_create_theta_{S}": (c_void_p, [POINTER(_abi.HpConfig), c_int, FPd]),
                f"orc_theta_count_{S}": (c_uint32, [c_void_p]),
                f"orc_grad_{S}": (c_int, [c_void_p, PP, c_uint64, PP, PP, PP, FPd]),
                f"orc_conv_direct2d_{S}": (None, [FPx, FPx, c_uint32, c_uint32, c_uint32, FPx]),
                f"orc_conv_growth_emb_{S}": (None, [FPx, FPx, c_uint32, FPx, c_uint32, c_uint32, FPx, c_int,
                                                    c_int]),
                f"orc_flow_emb_{S}": (None, [FPx] * 6 + [c_uint64, FPx, FPx]),
                f"orc_overlap_emb_{S}": (T, [T, T, T, T, T]),
                f"orc_max_flow_mag_{S}": (T, [FPx, FPx, c_uint64]),
                f"orc_mix_softmax_{S}": (None, [FPx, FPx, c_uint32, T, FPx]),
                f"orc_mix_linear_{S}": (None, [FPx, FPx, c_uint32, FPx]),
                f"orc_advect_params_{S}": (c_int32, [FPx, FPx, FPx, FPx, c_int, T, T, T, c_uint32, c_uint32, FPx,
                                                     FPx]),
                f"orc_emb_create_{S}": (c_void_p, [POINTER(_abi.HpConfig), POINTER(_abi.HpEmbedding), c_int]),
                f"orc_emb_destroy_{S}": (None, [c_void_p]),
                f"orc_emb_next_params_{S}": (FPx, [c_void_p, c_uint32]),
                f"orc_emb_last_radius_{S}": (c_int32, [c_void_p, c_uint32]),
                f"orc_emb_affinity_{S}": (None, [c_void_p, PP, PP, PP]),
                f"orc_emb_stencil_{S}": (None, [c_void_p, PP, PP, PP, PP, PP]),
                f"orc_emb_step_{S}": (None, [c_void_p, PP, PP, POINTER(c_float), POINTER(c_uint64)]),
                f"orc_emb_run_{S}": (None, [c_void_p, PP, PP, POINTER(c_float), POINTER(c_uint64), c_uint64]),
            }
            for name, (res, args) in sig.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
        L.orc_wrap_coord.restype = c_int
        L.orc_wrap_coord.argtypes = [c_int32, c_uint32]
        L.orc_validate.restype = c_int
        L.orc_validate.argtypes = [POINTER(_abi.HpConfig), POINTER(c_uint32), POINTER(c_uint32)]
        L.orc_seed_generate.restype = c_int
        L.orc_seed_generate.argtypes = [POINTER(_abi.HpSeed), c_uint32, c_uint32, c_uint32, c_uint32,
                                        POINTER(POINTER(c_float))]
        L.orc_total_mass_f32.restype = c_float
        L.orc_total_mass_f32.argtypes = [POINTER(POINTER(c_float)), c_uint32, c_uint64]
        L.orc_center_of_mass.restype = None
        L.orc_center_of_mass.argtypes = [POINTER(c_float), c_uint32, c_uint32, POINTER(c_float), POINTER(c_float)]
        L.orc_locomotion.restype = c_float
        L.orc_locomotion.argtypes = [c_float] * 4 + [c_uint32, c_uint32]
        _lib = L
    return _lib


def _sfx(dtype) -> str:
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def _p(a: np.ndarray):
    ct = c_float if a.dtype == np.float32 else c_double
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(POINTER(ct))


def _pp(arrs):
    ct = c_float if arrs[0].dtype == np.float32 else c_double
    arr = (POINTER(ct) * len(arrs))()
    for i, a in enumerate(arrs):
        assert a.flags["C_CONTIGUOUS"] and a.dtype == arrs[0].dtype
        arr[i] = a.ctypes.data_as(POINTER(ct))
    return arr


# ---------------------------------------------------------------------------------------
def kernel2d(kc, kernel_radius: int, dtype=np.float32) -> np.ndarray:
    """Kernel::from_config (kernel.rs:32-86) -> (size, size) array."""
    from autoverse_amd.schema import SimulationConfig
    cfg = SimulationConfig(kernels=[kc]).to_c()
    f = getattr(lib(), f"orc_kernel2d_{_sfx(dtype)}")
    size = f(cfg.kernels, cfg.rings, kernel_radius, None)
    out = np.zeros(size * size, dtype=dtype)
    f(cfg.kernels, cfg.rings, kernel_radius, _p(out))
    return out.reshape(size, size)


def kernel3d(kc, kernel_radius: int, dtype=np.float32) -> np.ndarray:
    from autoverse_amd.schema import SimulationConfig
    cfg = SimulationConfig(kernels=[kc]).to_c()
    f = getattr(lib(), f"orc_kernel3d_{_sfx(dtype)}")
    size = f(cfg.kernels, cfg.rings, kernel_radius, None)
    out = np.zeros(size ** 3, dtype=dtype)
    f(cfg.kernels, cfg.rings, kernel_radius, _p(out))
    return out.reshape(size, size, size)


def pad2d(k: np.ndarray, w: int, h: int) -> np.ndarray:
    out = np.zeros(w * h, dtype=k.dtype)
    getattr(lib(), f"orc_pad2d_{_sfx(k.dtype)}")(_p(np.ascontiguousarray(k)), k.shape[0], w, h, _p(out))
    return out.reshape(h, w)


def pad3d(k: np.ndarray, w: int, h: int, d: int) -> np.ndarray:
    out = np.zeros(w * h * d, dtype=k.dtype)
    getattr(lib(), f"orc_pad3d_{_sfx(k.dtype)}")(_p(np.ascontiguousarray(k)), k.shape[0], w, h, d, _p(out))
    return out.reshape(d, h, w)


def fft(x: np.ndarray, inverse: bool = False, dtype=np.float32) -> np.ndarray:
    """Own 1D complex FFT (unnormalised both ways)."""
    buf = np.empty(2 * x.size, dtype=dtype)
    buf[0::2] = x.real
    buf[1::2] = x.imag
    getattr(lib(), f"orc_fft_{_sfx(dtype)}")(_p(buf), x.size, int(inverse))
    return buf[0::2] + 1j * buf[1::2]


def fftn(x: np.ndarray, inverse: bool = False, dtype=np.float32) -> np.ndarray:
    """Own 2D (h, w) or 3D (d, h, w) complex FFT."""
    buf = np.empty(2 * x.size, dtype=dtype)
    buf[0::2] = x.real.ravel()
    buf[1::2] = x.imag.ravel()
    if x.ndim == 2:
        getattr(lib(), f"orc_fft2d_{_sfx(dtype)}")(_p(buf), x.shape[1], x.shape[0], int(inverse))
    else:
        getattr(lib(), f"orc_fft3d_{_sfx(dtype)}")(_p(buf), x.shape[2], x.shape[1], x.shape[0], int(inverse))
    return (buf[0::2] + 1j * buf[1::2]).reshape(x.shape)


def growth(u, mu, sigma, dtype=np.float32):
    return getattr(lib(), f"orc_growth_{_sfx(dtype)}")(u, mu, sigma)


def growth_accumulate(conv: np.ndarray, target: np.ndarray, weight, mu, sigma) -> None:
    getattr(lib(), f"orc_growth_accumulate_{_sfx(conv.dtype)}")(_p(conv), _p(target), conv.size, weight, mu, sigma)


def sobel2d(g: np.ndarray):
    h, w = g.shape
    g = np.ascontiguousarray(g)
    gx, gy = np.zeros_like(g), np.zeros_like(g)
    getattr(lib(), f"orc_sobel2d_{_sfx(g.dtype)}")(_p(g), _p(gx), _p(gy), w, h)
    return gx, gy


def sobel3d(g: np.ndarray):
    d, h, w = g.shape
    g = np.ascontiguousarray(g)
    gx, gy, gz = np.zeros_like(g), np.zeros_like(g), np.zeros_like(g)
    getattr(lib(), f"orc_sobel3d_{_sfx(g.dtype)}")(_p(g), _p(gx), _p(gy), _p(gz), w, h, d)
    return gx, gy, gz


def alpha(mass, beta_a, n, dtype=np.float32):
    return getattr(lib(), f"orc_alpha_{_sfx(dtype)}")(mass, beta_a, n)


def flow2d(gux, guy, gax, gay, mass, beta_a, n):
    fx, fy = np.zeros_like(gux), np.zeros_like(gux)
    getattr(lib(), f"orc_flow2d_{_sfx(gux.dtype)}")(_p(gux), _p(guy), _p(gax), _p(gay), _p(mass), gux.size,
                                                     beta_a, n, _p(fx), _p(fy))
    return fx, fy


def advect2d(current: np.ndarray, fx: np.ndarray, fy: np.ndarray, dt: float, s: float,
             conserve: bool = False) -> np.ndarray:
    """advect_mass (reintegration.rs:21-43): scatter into a zeroed grid.  conserve: the
    exact-conservation variant (each source's mass rescaled by (2s)^2 / its overlap-width product)."""
    h, w = current.shape
    nxt = np.zeros_like(current)
    getattr(lib(), f"orc_advect2d_mode_{_sfx(current.dtype)}")(_p(np.ascontiguousarray(current)), _p(fx), _p(fy),
                                                                _p(nxt), w, h, dt, s, int(conserve))
    return nxt


def advect3d(current, fx, fy, fz, dt, s, conserve: bool = False):
    d, h, w = current.shape
    nxt = np.zeros_like(current)
    getattr(lib(), f"orc_advect3d_mode_{_sfx(current.dtype)}")(_p(np.ascontiguousarray(current)), _p(fx), _p(fy),
                                                                _p(fz), _p(nxt), w, h, d, dt, s, int(conserve))
    return nxt


def wrap_coord(c: int, size: int) -> int:
    return lib().orc_wrap_coord(c, size)


def validate(cfg) -> tuple:
    c = cfg.to_c()
    bk, bc = c_uint32(0), c_uint32(0)
    st = lib().orc_validate(ctypes.byref(c.cfg), ctypes.byref(bk), ctypes.byref(bc))
    return st, bk.value, bc.value


def seed_generate(seed, w, h, d, c):
    out = [np.zeros(w * h * d, dtype=np.float32) for _ in range(c)]
    cs = seed.to_c()
    st = lib().orc_seed_generate(ctypes.byref(cs.seed), w, h, d, c, _pp(out))
    assert st == 0, st
    return out


def total_mass_f32(channels) -> float:
    return lib().orc_total_mass_f32(_pp(channels), len(channels), channels[0].size)


def center_of_mass(grid: np.ndarray):
    h, w = grid.shape
    cx, cy = c_float(0), c_float(0)
    lib().orc_center_of_mass(_p(np.ascontiguousarray(grid, dtype=np.float32)), w, h, ctypes.byref(cx),
                             ctypes.byref(cy))
    return cx.value, cy.value


def locomotion(x0, y0, x1, y1, w, h) -> float:
    return lib().orc_locomotion(x0, y0, x1, y1, w, h)


def theta_of(cfg) -> np.ndarray:
    """Differentiable parameters of a config, float64: per kernel (weight, mu, sigma), then
    per ring in config order (amplitude, distance, width)."""
    out = []
    for k in cfg.kernels:
        out += [k.weight, k.mu, k.sigma]
    for k in cfg.kernels:
        for r in k.rings:
            out += [r.amplitude, r.distance, r.width]
    return np.array(out, dtype=np.float64)


class Propagator:
    """CpuPropagator / CpuPropagator3D restated (propagator.rs:124-427), f32 or f64.
    `theta` (float64, layout of theta_of) overrides the config's kernel and ring values."""

    def __init__(self, cfg, dtype=np.float32, threads: int = 1, theta=None, line_threads: int = 1,
                 conserve: bool = False):
        self.cfg = cfg
        self.dtype = np.dtype(dtype)
        self.S = _sfx(dtype)
        self._c = cfg.to_c()
        if theta is None:
            self.ptr = getattr(lib(), f"orc_prop_create_{self.S}")(ctypes.byref(self._c.cfg), threads)
        else:
            self._theta = np.ascontiguousarray(theta, dtype=np.float64)
            self.ptr = getattr(lib(), f"orc_prop_create_theta_{self.S}")(
                ctypes.byref(self._c.cfg), threads, self._theta.ctypes.data_as(POINTER(c_double)))
        if not self.ptr:
            raise ValueError("invalid configuration")
        self.n = cfg.width * cfg.height * cfg.depth
        if line_threads > 1:
            self.set_line_threads(line_threads)
        if conserve:
            getattr(lib(), f"orc_prop_set_conserve_{self.S}")(self.ptr, 1)

    def close(self):
        if self.ptr:
            getattr(lib(), f"orc_prop_destroy_{self.S}")(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_threads(self, t: int):
        getattr(lib(), f"orc_prop_set_threads_{self.S}")(self.ptr, t)

    def set_line_threads(self, t: int):
        """Workers inside each FFT axis pass and 3D Sobel (results do not depend on it)."""
        getattr(lib(), f"orc_prop_set_line_threads_{self.S}")(self.ptr, t)

    def kernel_freq(self, k: int) -> np.ndarray:
        p = getattr(lib(), f"orc_prop_kernel_freq_{self.S}")(self.ptr, k)
        a = np.ctypeslib.as_array(p, shape=(2 * self.n,)).copy()
        return a[0::2] + 1j * a[1::2]

    def _chans(self, channels):
        return [np.ascontiguousarray(c, dtype=self.dtype).reshape(-1) for c in channels]

    def affinity(self, channels):
        ch = self._chans(channels)
        aff = [np.zeros(self.n, dtype=self.dtype) for _ in ch]
        getattr(lib(), f"orc_prop_affinity_{self.S}")(self.ptr, _pp(ch), _pp(aff))
        return aff

    def stencil(self, channels, affinity):
        ch = self._chans(channels)
        af = self._chans(affinity)
        nxt = [np.zeros(self.n, dtype=self.dtype) for _ in ch]
        getattr(lib(), f"orc_prop_stencil_{self.S}")(self.ptr, _pp(ch), _pp(af), _pp(nxt))
        return nxt

    def run(self, channels, steps: int, time: float = 0.0, step: int = 0):
        """Returns (channels, time, step) after `steps` steps; inputs are not modified."""
        ch = [c.copy() for c in self._chans(channels)]
        t, s = c_float(time), c_uint64(step)
        getattr(lib(), f"orc_prop_run_{self.S}")(self.ptr, _pp(ch), ctypes.byref(t), ctypes.byref(s), steps)
        return ch, t.value, s.value

    def grad(self, channels, steps: int, seed):
        """Reverse-mode derivative of L = sum_c <seed_c, A_T,c> over `steps` steps (2D).
        Returns (A_T, dL/dA_0, dL/dtheta) with theta laid out as theta_of()."""
        ch = self._chans(channels)
        sd = self._chans(seed)
        at = [np.zeros(self.n, dtype=self.dtype) for _ in ch]
        g0 = [np.zeros(self.n, dtype=self.dtype) for _ in ch]
        nth = getattr(lib(), f"orc_theta_count_{self.S}")(self.ptr)
        dth = np.zeros(max(1, nth), dtype=np.float64)
        st = getattr(lib(), f"orc_grad_{self.S}")(self.ptr, _pp(ch), steps, _pp(sd), _pp(at), _pp(g0),
                                                 dth.ctypes.data_as(POINTER(c_double)))
        if st != 0:
            raise ValueError(f"orc_grad failed ({st}): 3D, or a kernel wider than the grid")
        return at, g0, dth[:nth]


def loss(cfg, channels, steps: int, seed, theta=None, dtype=np.float64) -> float:
    """L = sum_c <seed_c, A_T,c> after `steps` oracle steps with parameters theta."""
    p = Propagator(cfg, dtype, theta=theta)
    at, _, _ = p.run(channels, steps)
    p.close()
    return float(sum(np.dot(np.asarray(a, np.float64), np.asarray(s, np.float64).reshape(-1))
                     for a, s in zip(at, seed)))


# ---- embedded-parameter mode (oracle_embedded.inc) ------------------------------------------
# Parameter grids are float arrays of shape (N, 5): mu, sigma, weight, beta_a, n per cell, the
# layout of the reference's Vec<CellParams>.
CELL_DEFAULT = (0.15, 0.015, 1.0, 1.0, 2.0)  # CellParams::default (embedding.rs:35-45)


def _params(a, n, dtype):
    a = np.ascontiguousarray(a, dtype=dtype).reshape(n, 5)
    return a


def conv_direct2d(inp: np.ndarray, k: np.ndarray, w: int, h: int) -> np.ndarray:
    """convolve_direct (direct_convolution.rs:20-60)."""
    dt = inp.dtype
    inp = np.ascontiguousarray(inp).reshape(-1)
    k = np.ascontiguousarray(k, dtype=dt)
    out = np.zeros(w * h, dtype=dt)
    getattr(lib(), f"orc_conv_direct2d_{_sfx(dt)}")(_p(inp), _p(k.reshape(-1)), k.shape[0], w, h, _p(out))
    return out


def conv_growth_emb(inp, k, params, w, h, target=None, threads: int = 1):
    """convolve_growth_embedded (target None) or convolve_growth_accumulate_embedded
    (direct_convolution.rs:66-174)."""
    dt = inp.dtype
    inp = np.ascontiguousarray(inp).reshape(-1)
    k = np.ascontiguousarray(k, dtype=dt)
    pr = _params(params, w * h, dt)
    acc = target is not None
    out = target if acc else np.zeros(w * h, dtype=dt)
    getattr(lib(), f"orc_conv_growth_emb_{_sfx(dt)}")(_p(inp), _p(k.reshape(-1)), k.shape[0], _p(pr.reshape(-1)),
                                                    w, h, _p(out), int(acc), threads)
    return out


def flow_emb(gux, guy, gax, gay, mass, params):
    dt = gux.dtype
    n = gux.size
    fx = np.zeros(n, dtype=dt)
    fy = np.zeros(n, dtype=dt)
    pr = _params(params, n, dt)
    getattr(lib(), f"orc_flow_emb_{_sfx(dt)}")(_p(gux), _p(guy), _p(gax), _p(gay), _p(mass), _p(pr.reshape(-1)), n,
                                             _p(fx), _p(fy))
    return fx, fy


def overlap_emb(dest_x, dest_y, s, cell_x, cell_y, dtype=np.float32) -> float:
    return getattr(lib(), f"orc_overlap_emb_{_sfx(dtype)}")(dest_x, dest_y, s, cell_x, cell_y)


def mix_softmax(params, masses, temperature, dtype=np.float32):
    pr = np.ascontiguousarray(params, dtype=dtype).reshape(-1, 5)
    m = np.ascontiguousarray(masses, dtype=dtype).reshape(-1)
    out = np.zeros(5, dtype=dtype)
    getattr(lib(), f"orc_mix_softmax_{_sfx(dtype)}")(_p(pr.reshape(-1)), _p(m), m.size, temperature, _p(out))
    return out


def mix_linear(params, masses, dtype=np.float32):
    pr = np.ascontiguousarray(params, dtype=dtype).reshape(-1, 5)
    m = np.ascontiguousarray(masses, dtype=dtype).reshape(-1)
    out = np.zeros(5, dtype=dtype)
    getattr(lib(), f"orc_mix_linear_{_sfx(dtype)}")(_p(pr.reshape(-1)), _p(m), m.size, _p(out))
    return out


def advect_params(mass, params, fx, fy, w, h, dt, s, linear=False, temperature=1.0, next_params=None):
    """advect_mass_and_params_into (param_advection.rs:83-166).  next_params (N x 5) is the
    grid written into (cells that receive no mix keep its values; default CellParams::default).
    Returns (next_mass, next_params, R)."""
    d = mass.dtype
    n = w * h
    pr = _params(params, n, d)
    nm = np.zeros(n, dtype=d)
    if next_params is None:
        nps = np.tile(np.array(CELL_DEFAULT, dtype=d), (n, 1))
    else:
        nps = _params(next_params, n, d).copy()
    R = getattr(lib(), f"orc_advect_params_{_sfx(d)}")(_p(np.ascontiguousarray(mass)), _p(pr.reshape(-1)),
                                                     _p(np.ascontiguousarray(fx)), _p(np.ascontiguousarray(fy)),
                                                     int(linear), temperature, dt, s, w, h, _p(nm),
                                                     _p(nps.reshape(-1)))
    return nm, nps, R


class EmbeddedPropagator:
    """EmbeddedPropagator restated (embedded_propagator.rs:144-340), f32 or f64.  `embedding`
    is a dict with enabled / mixing_temperature / linear_mixing (EmbeddingConfig)."""

    def __init__(self, cfg, embedding=None, dtype=np.float32, threads: int = 1):
        self.cfg = cfg
        self.dtype = np.dtype(dtype)
        self.S = _sfx(dtype)
        self._c = cfg.to_c()
        e = embedding or {}
        self._e = _abi.HpEmbedding(int(bool(e.get("enabled", False))), float(e.get("mixing_temperature", 1.0)),
                                   int(bool(e.get("linear_mixing", False))))
        self.ptr = getattr(lib(), f"orc_emb_create_{self.S}")(ctypes.byref(self._c.cfg), ctypes.byref(self._e),
                                                              threads)
        if not self.ptr:
            raise ValueError("invalid configuration")
        self.n = cfg.width * cfg.height

    def close(self):
        if self.ptr:
            getattr(lib(), f"orc_emb_destroy_{self.S}")(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chans(self, channels):
        return [np.ascontiguousarray(c, dtype=self.dtype).reshape(-1) for c in channels]

    def _pars(self, params):
        return [np.ascontiguousarray(p, dtype=self.dtype).reshape(-1) for p in params]

    def next_params(self, c: int) -> np.ndarray:
        p = getattr(lib(), f"orc_emb_next_params_{self.S}")(self.ptr, c)
        return np.ctypeslib.as_array(p, shape=(self.n * 5,)).copy().reshape(self.n, 5)

    def set_next_params(self, c: int, values) -> None:
        p = getattr(lib(), f"orc_emb_next_params_{self.S}")(self.ptr, c)
        np.ctypeslib.as_array(p, shape=(self.n * 5,))[:] = np.asarray(values, dtype=self.dtype).reshape(-1)

    def last_radius(self, c: int) -> int:
        return getattr(lib(), f"orc_emb_last_radius_{self.S}")(self.ptr, c)

    def affinity(self, channels, params):
        ch, pr = self._chans(channels), self._pars(params)
        aff = [np.zeros(self.n, dtype=self.dtype) for _ in ch]
        getattr(lib(), f"orc_emb_affinity_{self.S}")(self.ptr, _pp(ch), _pp(pr), _pp(aff))
        return aff

    def stencil(self, channels, params, affinity, next_params):
        """Stages 2-5 into fresh mass grids and copies of next_params (N x 5 per channel)."""
        ch, pr, af = self._chans(channels), self._pars(params), self._chans(affinity)
        nxt = [np.zeros(self.n, dtype=self.dtype) for _ in ch]
        nps = [np.array(p, dtype=self.dtype).reshape(-1) for p in next_params]
        getattr(lib(), f"orc_emb_stencil_{self.S}")(self.ptr, _pp(ch), _pp(pr), _pp(af), _pp(nxt), _pp(nps))
        return nxt, [p.reshape(self.n, 5) for p in nps]

    def run(self, channels, params, steps: int, time: float = 0.0, step: int = 0):
        """Returns (channels, params, time, step); inputs are not modified (the propagator's
        own next grids are)."""
        ch = [c.copy() for c in self._chans(channels)]
        pr = [p.copy() for p in self._pars(params)]
        t, s = c_float(time), c_uint64(step)
        getattr(lib(), f"orc_emb_run_{self.S}")(self.ptr, _pp(ch), _pp(pr), ctypes.byref(t), ctypes.byref(s), steps)
        return ch, [p.reshape(self.n, 5) for p in pr], t.value, s.value
