"""hp_advect_mass_2d_noalloc (device buffers, static device scratch, no allocation) against
hp_advect_mass_2d (host buffers, allocating), through the C ABI: bit for bit on two calls in a
row, the second on a smaller grid, and the size limit refused on the host."""
import ctypes

import numpy as np
import pytest

from autoverse_amd import _abi

pytestmark = pytest.mark.gpu


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _fields(w, h, seed, scale):
    rng = np.random.default_rng(seed)
    cur = rng.uniform(0, 1, (h, w)).astype(np.float32)
    cur[rng.uniform(size=(h, w)) < 0.25] = 0.0
    fx = rng.normal(0, scale, (h, w)).astype(np.float32)
    fy = rng.normal(0, scale, (h, w)).astype(np.float32)
    return cur, fx, fy


def _allocating(lib, cur, fx, fy, w, h, dt, s):
    out = np.empty(w * h, np.float32)
    st = lib.hp_advect_mass_2d(0, _fp(cur), _fp(fx), _fp(fy), _fp(out), w, h, dt, s)
    assert st == _abi.HP_OK, lib.hp_last_global_error().decode()
    return out


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime the library itself links (loaded by
    file, so the test and the library share one runtime instance)."""

    def __init__(self, lib):
        self.h = ctypes.CDLL("libamdhip64.so.7")  # the soname the library links
        V = ctypes.c_void_p
        self.h.hipMalloc.argtypes = [ctypes.POINTER(V), ctypes.c_size_t]
        self.h.hipFree.argtypes = [V]
        self.h.hipMemcpy.argtypes = [V, V, ctypes.c_size_t, ctypes.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def upload(self, a):
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, a):
        assert self.h.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), p, a.nbytes, 2) == 0  # device to host
        return a


def _noalloc(lib, hip, cur, fx, fy, w, h, dt, s):
    bufs = [hip.upload(np.ascontiguousarray(a.reshape(-1))) for a in (cur, fx, fy)]
    bufs.append(hip.upload(np.full(w * h, np.nan, np.float32)))  # every cell must be written
    try:
        st = lib.hp_advect_mass_2d_noalloc(0, bufs[0], bufs[1], bufs[2], bufs[3], w, h, dt, s)
        assert st == _abi.HP_OK, lib.hp_last_global_error().decode()
        return hip.download(bufs[3], np.empty(w * h, np.float32))
    finally:
        for b in bufs:
            hip.h.hipFree(b)


def test_noalloc_advect_matches_allocating_entry_point():
    lib = _abi.load_library()
    hip = _Hip(lib)
    print("[noalloc] 1024x1024 then 40x24", flush=True)
    for (w, h, seed, scale, dt, s) in ((1024, 1024, 1, 3.0, 0.2, 1.0), (40, 24, 2, 1.0, 0.7, 0.75)):
        cur, fx, fy = _fields(w, h, seed, scale)
        want = _allocating(lib, cur, fx, fy, w, h, dt, s)
        got = _noalloc(lib, hip, cur, fx, fy, w, h, dt, s)
        np.testing.assert_array_equal(got, want)
        assert abs(float(got.astype(np.float64).sum()) - float(cur.astype(np.float64).sum())) < 1e-3 * max(
            1.0, float(cur.sum()))
    # beyond the scratch: refused before any device work
    big = 1025 * 1024
    p = hip.alloc(4)
    try:
        assert lib.hp_advect_mass_2d_noalloc(0, p, p, p, p, 1025, 1024, 0.2, 1.0) == _abi.HP_ERR_UNSUPPORTED_SIZE
    finally:
        hip.h.hipFree(p)
    assert b"1048576" in lib.hp_last_global_error()
    assert big > (1 << 20)
