"""HipPropagator: the reference CpuPropagator API over the native HIP library.

Mirrors src/compute/propagator.rs (SimulationState :19-121, CpuPropagator::new /
step / run / config :157-432, SimulationStats :436-474) and its 3D twin
(src/compute/propagator3d.rs).  All numerics run in libflowlenia_hip.so on the GPU;
this module only moves host buffers across the C ABI.  There is no CPU path: if
the library or the device is missing, construction raises.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List

import numpy as np

from . import _abi
from .schema import ConfigError, Seed, SimulationConfig


class SimulationState:
    """SimulationState (propagator.rs:19-121): per-channel flat f32 grids
    indexed z*H*W + y*W + x, plus time (f32) and step (u64)."""

    def __init__(self, channels: List[np.ndarray], width: int, height: int, depth: int = 1,
                 time: float = 0.0, step: int = 0):
        self.channels = [np.ascontiguousarray(c, dtype=np.float32).reshape(-1) for c in channels]
        self.width, self.height, self.depth = width, height, depth
        self.time = float(np.float32(time))
        self.step = int(step)
        n = width * height * depth
        for c in self.channels:
            if c.size != n:
                raise ValueError(f"channel has {c.size} cells, expected {n}")

    @classmethod
    def from_seed(cls, seed: Seed, config: SimulationConfig) -> "SimulationState":
        """SimulationState::from_seed (propagator.rs:36-59)."""
        chans = seed.generate(config.width, config.height, config.depth, config.channels)
        return cls(chans, config.width, config.height, config.depth)

    def is_3d(self) -> bool:
        return self.depth > 1

    def grid_size(self) -> int:
        return self.width * self.height * self.depth

    def idx(self, x: int, y: int, z: int = 0) -> int:
        return z * self.height * self.width + y * self.width + x

    def get(self, x: int, y: int, channel: int) -> float:
        return float(self.channels[channel][y * self.width + x])

    def get_3d(self, x: int, y: int, z: int, channel: int) -> float:
        return float(self.channels[channel][self.idx(x, y, z)])

    def total_mass(self) -> float:
        """total_mass_all_channels (reintegration.rs:221-228): f32 sums in order."""
        total = np.float32(0.0)
        for c in self.channels:
            # ufunc.accumulate adds strictly left to right in f32, as the reference's iterator sum
            s = np.add.accumulate(c, dtype=np.float32)[-1] if c.size else np.float32(0.0)
            total = np.float32(total + s)
        return float(total)

    def channel_sum(self) -> np.ndarray:
        out = np.zeros(self.grid_size(), dtype=np.float32)
        for c in self.channels:
            out += c
        return out

    def copy(self) -> "SimulationState":
        return SimulationState([c.copy() for c in self.channels], self.width, self.height, self.depth,
                               self.time, self.step)


@dataclass
class SimulationStats:
    """SimulationStats (propagator.rs:436-474)."""
    total_mass: float
    max_value: float
    min_value: float
    mean_value: float
    active_cells: int

    @classmethod
    def from_state(cls, state: SimulationState) -> "SimulationStats":
        allv = np.concatenate(state.channels)
        total = float(np.sum(allv, dtype=np.float64))
        return cls(total, float(allv.max()), float(allv.min()), total / allv.size, int((allv > 1e-6).sum()))


class HipPropagator:
    """Drop-in for CpuPropagator / CpuPropagator3D (propagator.rs:124-432) on one GPU."""

    def __init__(self, config: SimulationConfig, device: int = 0):
        self._lib = _abi.load_library()
        self._config = config
        self._c = config.to_c()
        ptr = ctypes.c_void_p()
        st = self._lib.hp_create(ctypes.byref(self._c.cfg), device, ctypes.byref(ptr))
        if st in (_abi.HP_ERR_INVALID_DIMENSIONS, _abi.HP_ERR_INVALID_CHANNELS, _abi.HP_ERR_INVALID_TIME_STEP,
                  _abi.HP_ERR_INVALID_KERNEL_RADIUS, _abi.HP_ERR_INVALID_CHANNEL_INDEX):
            config.validate()  # raises ConfigError with the offending kernel/channel
            raise ConfigError(st)
        if st != _abi.HP_OK:
            raise _abi.HipError(st, self._lib.hp_last_global_error().decode())
        self._ptr = ptr
        self.n = config.width * config.height * config.depth

    # -- lifecycle ---------------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_ptr", None):
            self._lib.hp_destroy(self._ptr)
            self._ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st: int, what: str) -> None:
        if st != _abi.HP_OK:
            raise _abi.HipError(st, f"{what}: {self._lib.hp_last_error(self._ptr).decode()}")

    def config(self) -> SimulationConfig:
        """CpuPropagator::config (propagator.rs:430-432)."""
        return self._config

    # -- execution plan (no reference counterpart) ---------------------------------------------
    def plan(self) -> dict:
        """The resolved plan: fft_path, spectral_group, radix_order, column_split, stencil_tile."""
        p = _abi.HpPlan()
        self._check(self._lib.hp_get_plan(self._ptr, ctypes.byref(p)), "hp_get_plan")
        return p.as_dict()

    def set_plan(self, plan: dict) -> None:
        """Rebuild the device plan (the state, time and step are kept)."""
        p = _abi.HpPlan.from_dict({**self.plan(), **plan})
        self._check(self._lib.hp_set_plan(self._ptr, ctypes.byref(p)), "hp_set_plan")

    def autotune(self, steps: int = 20):
        """Time every candidate plan from the current state (restored afterwards), keep the
        fastest, and remember it for later propagators of the same shape.  Returns (best plan,
        [(plan, us per step or None if it disagreed with the first plan)])."""
        st, best, table = _abi.autotune_call(self._lib.hp_autotune, self._ptr, steps)
        self._check(st, "hp_autotune")
        return best, table

    def _ptrs(self, state: SimulationState):
        if len(state.channels) != self._config.channels:
            raise ValueError("channel count mismatch")
        for c in state.channels:
            if c.size != self.n or c.dtype != np.float32 or not c.flags["C_CONTIGUOUS"]:
                raise ValueError("state channels must be contiguous float32 of W*H*D cells")
        return _abi.float_ptr_array(state.channels)

    # -- CpuPropagator API -------------------------------------------------------------------
    def step(self, state: SimulationState) -> None:
        """CpuPropagator::step (propagator.rs:239-420): in place, time += dt, step += 1."""
        p = self._ptrs(state)
        t, s = ctypes.c_float(state.time), ctypes.c_uint64(state.step)
        self._check(self._lib.hp_step(self._ptr, p, ctypes.byref(t), ctypes.byref(s)), "hp_step")
        state.time, state.step = t.value, s.value

    def run(self, state: SimulationState, steps: int) -> None:
        """CpuPropagator::run (propagator.rs:423-427): one upload, device-resident steps, one download."""
        p = self._ptrs(state)
        t, s = ctypes.c_float(state.time), ctypes.c_uint64(state.step)
        self._check(self._lib.hp_run(self._ptr, p, ctypes.byref(t), ctypes.byref(s), steps), "hp_run")
        state.time, state.step = t.value, s.value

    # -- device-resident use ---------------------------------------------------------------------
    def set_state(self, state: SimulationState) -> None:
        self._check(self._lib.hp_set_state(self._ptr, self._ptrs(state), state.time, state.step), "hp_set_state")

    def get_state(self) -> SimulationState:
        cfg = self._config
        chans = [np.empty(self.n, dtype=np.float32) for _ in range(cfg.channels)]
        t, s = ctypes.c_float(), ctypes.c_uint64()
        self._check(self._lib.hp_get_state(self._ptr, _abi.float_ptr_array(chans), ctypes.byref(t),
                                           ctypes.byref(s)), "hp_get_state")
        return SimulationState(chans, cfg.width, cfg.height, cfg.depth, t.value, s.value)

    def advance(self, steps: int) -> None:
        self._check(self._lib.hp_advance(self._ptr, steps), "hp_advance")

    def mass_f64(self):
        """Per-channel and total FP64 mass of the device-resident state (fixed reduction order)."""
        per = (ctypes.c_double * self._config.channels)()
        tot = ctypes.c_double()
        self._check(self._lib.hp_mass_f64(self._ptr, per, ctypes.byref(tot)), "hp_mass_f64")
        return list(per), tot.value

    def set_conservation(self, mode: str) -> None:
        """'reference' (default: the reference's f32 reintegration, bit for bit) or 'exact' (FP64
        contributions and sums per destination; mass conserved to FP64-reference level)."""
        m = {"reference": 0, "exact": 1}[mode]
        self._check(self._lib.hp_set_conservation(self._ptr, m), "hp_set_conservation")

    def conservation(self) -> str:
        m = ctypes.c_int(0)
        self._check(self._lib.hp_get_conservation(self._ptr, ctypes.byref(m)), "hp_get_conservation")
        return "exact" if m.value else "reference"

    def stats(self):
        r, g = ctypes.c_uint32(), ctypes.c_uint64()
        self._check(self._lib.hp_stats(self._ptr, ctypes.byref(r), ctypes.byref(g)), "hp_stats")
        return r.value, g.value

    def profile_enable(self, on: bool, repeats: int = 1) -> None:
        """Per-pass event timing (graph replay off while on).  `repeats` > 1 launches each
        out-of-place pass that many times back to back inside its event pair."""
        self._check(self._lib.hp_profile_repeats(self._ptr, int(repeats)), "hp_profile_repeats")
        self._check(self._lib.hp_profile_enable(self._ptr, int(on)), "hp_profile_enable")

    def profile_read(self):
        ms = (ctypes.c_double * _abi.HP_K_NUM_CLASSES)()
        n = (ctypes.c_uint64 * _abi.HP_K_NUM_CLASSES)()
        self._check(self._lib.hp_profile_read(self._ptr, ms, n), "hp_profile_read")
        return {k: (ms[i], n[i]) for i, k in enumerate(_abi.KERNEL_CLASSES)}

    def kernel_bytes(self):
        b = (ctypes.c_double * _abi.HP_K_NUM_CLASSES)()
        self._check(self._lib.hp_kernel_bytes(self._ptr, b), "hp_kernel_bytes")
        return {k: b[i] for i, k in enumerate(_abi.KERNEL_CLASSES)}

    # -- stage access (parity tests) -------------------------------------------------------------
    def affinity(self, channels) -> List[np.ndarray]:
        """Stage 1 of the step (propagator.rs:244-280) for the given state."""
        ch = [np.ascontiguousarray(c, dtype=np.float32).reshape(-1) for c in channels]
        out = [np.empty(self.n, dtype=np.float32) for _ in ch]
        self._check(self._lib.hp_affinity(self._ptr, _abi.float_ptr_array(ch), _abi.float_ptr_array(out)),
                    "hp_affinity")
        return out

    def stencil(self, channels, affinity) -> List[np.ndarray]:
        """Stages 2-5 of the step (propagator.rs:303-414) with the affinity supplied."""
        ch = [np.ascontiguousarray(c, dtype=np.float32).reshape(-1) for c in channels]
        af = [np.ascontiguousarray(c, dtype=np.float32).reshape(-1) for c in affinity]
        out = [np.empty(self.n, dtype=np.float32) for _ in ch]
        self._check(self._lib.hp_stencil_with_affinity(self._ptr, _abi.float_ptr_array(ch), _abi.float_ptr_array(af),
                                                       _abi.float_ptr_array(out)), "hp_stencil_with_affinity")
        return out


# ---- stage functions (the crate's re-exported L1 API, src/compute/mod.rs:5-41) -----------------
def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def advect_mass(current, flow_x, flow_y, width, height, dt, distribution_size, device=0) -> np.ndarray:
    """advect_mass (reintegration.rs:21-43) on the GPU."""
    lib = _abi.load_library()
    cur, fx, fy = _f32(current).reshape(-1), _f32(flow_x).reshape(-1), _f32(flow_y).reshape(-1)
    out = np.empty(width * height, dtype=np.float32)
    st = lib.hp_advect_mass_2d(device, _fp(cur), _fp(fx), _fp(fy), _fp(out), width, height, dt, distribution_size)
    if st != _abi.HP_OK:
        raise _abi.HipError(st, lib.hp_last_global_error().decode())
    return out


def advect_mass_3d(current, flow_x, flow_y, flow_z, width, height, depth, dt, distribution_size, device=0):
    """advect_mass_3d (reintegration3d.rs:20-54) on the GPU."""
    lib = _abi.load_library()
    a = [_f32(v).reshape(-1) for v in (current, flow_x, flow_y, flow_z)]
    out = np.empty(width * height * depth, dtype=np.float32)
    st = lib.hp_advect_mass_3d(device, *[_fp(v) for v in a], _fp(out), width, height, depth, dt, distribution_size)
    if st != _abi.HP_OK:
        raise _abi.HipError(st, lib.hp_last_global_error().decode())
    return out


def sobel_gradient(grid, width, height, device=0):
    """sobel_gradient (gradient.rs:15-62) on the GPU."""
    lib = _abi.load_library()
    g = _f32(grid).reshape(-1)
    gx, gy = np.empty_like(g), np.empty_like(g)
    st = lib.hp_sobel_gradient_2d(device, _fp(g), _fp(gx), _fp(gy), width, height)
    if st != _abi.HP_OK:
        raise _abi.HipError(st, lib.hp_last_global_error().decode())
    return gx, gy


def compute_flow_field(gux, guy, gax, gay, mass_sum, flow, device=0):
    """compute_flow_field (flow.rs:39-63) on the GPU."""
    lib = _abi.load_library()
    a = [_f32(v).reshape(-1) for v in (gux, guy, gax, gay, mass_sum)]
    fx, fy = np.empty_like(a[0]), np.empty_like(a[0])
    st = lib.hp_compute_flow_2d(device, *[_fp(v) for v in a], a[0].size,
                                _abi.HpFlow(flow.beta_a, flow.n, flow.distribution_size), _fp(fx), _fp(fy))
    if st != _abi.HP_OK:
        raise _abi.HipError(st, lib.hp_last_global_error().decode())
    return fx, fy


def growth_accumulate(conv, target, weight, mu, sigma, device=0) -> None:
    """growth_accumulate (growth.rs:44-59) on the GPU; `target` (float32) is updated in place."""
    lib = _abi.load_library()
    c = _f32(conv).reshape(-1)
    if target.dtype != np.float32 or not target.flags["C_CONTIGUOUS"]:
        raise ValueError("target must be contiguous float32")
    st = lib.hp_growth_accumulate(device, _fp(c), _fp(target), c.size, weight, mu, sigma)
    if st != _abi.HP_OK:
        raise _abi.HipError(st, lib.hp_last_global_error().decode())
