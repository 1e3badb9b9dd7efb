"""Embedded-parameter Flow Lenia over the native HIP library.

Mirrors the reference's parameter embedding: the per-cell parameter types of
src/schema/embedding.rs (CellParams :20-45, ParameterGrid :172-320, SpeciesConfig :370-395),
the state and propagator of src/compute/embedded_propagator.rs (EmbeddedState :24-137,
EmbeddedPropagator :144-340) and the stage function advect_mass_and_params of
src/compute/param_advection.rs:47-79.  Every step runs in libflowlenia_hip.so on the GPU;
this module moves host buffers across the C ABI.  There is no CPU path.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from . import _abi
from .schema import ConfigError, EmbeddingConfig, Seed, SimulationConfig

FIELDS = ("mu", "sigma", "weight", "beta_a", "n")


@dataclass(frozen=True)
class CellParams:
    """CellParams (embedding.rs:20-58): growth centre and width, growth weight, and the alpha
    threshold and exponent of one cell; the default is embedding.rs:35-45's."""
    mu: float = 0.15
    sigma: float = 0.015
    weight: float = 1.0
    beta_a: float = 1.0
    n: float = 2.0

    def __post_init__(self):
        # the reference's fields are f32
        for f in FIELDS:
            object.__setattr__(self, f, float(np.float32(getattr(self, f))))

    @classmethod
    def new(cls, mu: float, sigma: float, weight: float, beta_a: float, n: float) -> "CellParams":
        return cls(mu, sigma, weight, beta_a, n)

    def as_array(self) -> np.ndarray:
        return np.array([self.mu, self.sigma, self.weight, self.beta_a, self.n], dtype=np.float32)

    @classmethod
    def from_array(cls, a) -> "CellParams":
        a = np.asarray(a, dtype=np.float32)
        return cls(*(float(v) for v in a[:5]))


class ParameterGrid:
    """ParameterGrid (embedding.rs:172-320): W*H cells of CellParams in row-major order, held as
    a (W*H, 5) float32 array, the memory layout of the reference's Vec<CellParams>."""

    def __init__(self, width: int, height: int, initial: CellParams = CellParams(), data=None):
        self.width, self.height = int(width), int(height)
        if data is not None:
            self.data = np.ascontiguousarray(data, dtype=np.float32).reshape(self.width * self.height, 5)
        else:
            self.data = np.tile(initial.as_array(), (self.width * self.height, 1))

    @classmethod
    def new(cls, width: int, height: int, initial: CellParams) -> "ParameterGrid":
        return cls(width, height, initial)

    @classmethod
    def from_defaults(cls, width: int, height: int) -> "ParameterGrid":
        return cls(width, height, CellParams())

    @classmethod
    def from_species(cls, width: int, height: int,
                     species: Sequence[Tuple[CellParams, Callable[[int, int], bool]]],
                     default: CellParams) -> "ParameterGrid":
        """The first species whose predicate holds at (x, y) sets the cell (embedding.rs:203-227)."""
        g = cls(width, height, default)
        for y in range(height):
            for x in range(width):
                for params, pred in species:
                    if pred(x, y):
                        g.set(x, y, params)
                        break
        return g

    def dimensions(self) -> Tuple[int, int]:
        return self.width, self.height

    def get(self, x: int, y: int) -> CellParams:
        return CellParams.from_array(self.data[y * self.width + x])

    def get_idx(self, idx: int) -> CellParams:
        return CellParams.from_array(self.data[idx])

    def set(self, x: int, y: int, params: CellParams) -> None:
        self.data[y * self.width + x] = params.as_array()

    def set_idx(self, idx: int, params: CellParams) -> None:
        self.data[idx] = params.as_array()

    def fill(self, params: CellParams) -> None:
        self.data[:] = params.as_array()

    def get_wrapped(self, x: int, y: int) -> CellParams:
        """Periodic lookup (embedding.rs:284-289)."""
        return self.get(((x % self.width) + self.width) % self.width, ((y % self.height) + self.height) % self.height)

    def extract(self, name: str) -> np.ndarray:
        """One parameter as a flat grid (extract_mu / _sigma / _weight / _beta_a / _n)."""
        return self.data[:, FIELDS.index(name)].copy()

    def extract_mu(self) -> np.ndarray:
        return self.extract("mu")

    def extract_sigma(self) -> np.ndarray:
        return self.extract("sigma")

    def extract_weight(self) -> np.ndarray:
        return self.extract("weight")

    def extract_beta_a(self) -> np.ndarray:
        return self.extract("beta_a")

    def extract_n(self) -> np.ndarray:
        return self.extract("n")

    def copy(self) -> "ParameterGrid":
        return ParameterGrid(self.width, self.height, data=self.data.copy())


@dataclass
class SpeciesConfig:
    """SpeciesConfig (embedding.rs:370-395)."""
    name: str
    params: CellParams
    initial_region: Optional[Tuple[float, float, float]] = None

    def with_region(self, center_x: float, center_y: float, radius: float) -> "SpeciesConfig":
        return SpeciesConfig(self.name, self.params, (center_x, center_y, radius))


def default_params(config: SimulationConfig) -> CellParams:
    """The parameters EmbeddedState::from_seed gives every cell (embedded_propagator.rs:50-57)."""
    lib = _abi.load_library()
    c = config.to_c()
    out = _abi.HpCellParams()
    st = lib.hp_emb_default_params(ctypes.byref(c.cfg), ctypes.byref(out))
    if st != _abi.HP_OK:
        raise _abi.HipError(st, "hp_emb_default_params")
    return CellParams(out.mu, out.sigma, out.weight, out.beta_a, out.n)


class EmbeddedState:
    """EmbeddedState (embedded_propagator.rs:24-137): 2D mass channels and one parameter grid
    per channel."""

    def __init__(self, channels: List[np.ndarray], params: List[ParameterGrid], width: int, height: int,
                 time: float = 0.0, step: int = 0):
        self.channels = [np.ascontiguousarray(c, dtype=np.float32).reshape(-1) for c in channels]
        self.params = list(params)
        self.width, self.height = int(width), int(height)
        self.time = float(np.float32(time))
        self.step = int(step)
        n = self.width * self.height
        if len(self.params) != len(self.channels):
            raise ValueError("one parameter grid per channel")
        for c, p in zip(self.channels, self.params):
            if c.size != n or p.data.shape != (n, 5):
                raise ValueError("channel or parameter grid does not match the grid size")

    @classmethod
    def from_seed(cls, seed: Seed, config: SimulationConfig) -> "EmbeddedState":
        """EmbeddedState::from_seed (embedded_propagator.rs:40-72): 2D seed, every cell of every
        channel at default_params(config)."""
        chans = seed.generate(config.width, config.height, 1, config.channels)
        d = default_params(config)
        return cls(chans, [ParameterGrid(config.width, config.height, d) for _ in range(config.channels)],
                   config.width, config.height)

    @classmethod
    def from_seed_with_params(cls, seed: Seed, config: SimulationConfig,
                              initial_params: List[ParameterGrid]) -> "EmbeddedState":
        """EmbeddedState::from_seed_with_params (embedded_propagator.rs:75-104)."""
        if len(initial_params) != config.channels:
            raise ValueError("Must provide one parameter grid per channel")
        chans = seed.generate(config.width, config.height, 1, config.channels)
        return cls(chans, list(initial_params), config.width, config.height)

    def total_mass(self) -> float:
        """total_mass_all_channels (reintegration.rs:221-228): f32 sums in order."""
        total = np.float32(0.0)
        for c in self.channels:
            s = np.add.accumulate(c, dtype=np.float32)[-1] if c.size else np.float32(0.0)
            total = np.float32(total + s)
        return float(total)

    def get(self, x: int, y: int, channel: int) -> float:
        return float(self.channels[channel][y * self.width + x])

    def get_params(self, x: int, y: int, channel: int) -> CellParams:
        return self.params[channel].get(x, y)

    def sum_at(self, x: int, y: int) -> float:
        i = y * self.width + x
        s = np.float32(0.0)
        for c in self.channels:
            s = np.float32(s + c[i])
        return float(s)

    def channel_sum(self) -> np.ndarray:
        out = np.zeros(self.width * self.height, dtype=np.float32)
        for c in self.channels:
            out += c
        return out

    def copy(self) -> "EmbeddedState":
        return EmbeddedState([c.copy() for c in self.channels], [p.copy() for p in self.params], self.width,
                             self.height, self.time, self.step)


def _pp(arrays) -> ctypes.Array:
    """float* per array; converted copies stay alive as long as the returned array."""
    conv = [np.ascontiguousarray(a, dtype=np.float32) for a in arrays]
    arr = _abi.float_ptr_array(conv)
    arr._keep = conv
    return arr


class HipEmbeddedPropagator:
    """Drop-in for EmbeddedPropagator (embedded_propagator.rs:144-340) on one GPU.  The embedding
    settings are `config.embedding` unless `embedding` is given."""

    def __init__(self, config: SimulationConfig, embedding: Optional[EmbeddingConfig] = None, device: int = 0):
        self._lib = _abi.load_library()
        self._config = config
        self._embedding = embedding if embedding is not None else config.embedding
        self._c = config.to_c()
        self._e = self._embedding.to_c()
        ptr = ctypes.c_void_p()
        st = self._lib.hp_emb_create(ctypes.byref(self._c.cfg), ctypes.byref(self._e), device, ctypes.byref(ptr))
        if st in (_abi.HP_ERR_INVALID_DIMENSIONS, _abi.HP_ERR_INVALID_CHANNELS, _abi.HP_ERR_INVALID_TIME_STEP,
                  _abi.HP_ERR_INVALID_KERNEL_RADIUS, _abi.HP_ERR_INVALID_CHANNEL_INDEX):
            config.validate()
            raise ConfigError(st)
        if st != _abi.HP_OK:
            raise _abi.HipError(st, self._lib.hp_emb_last_error(None).decode())
        self._ptr = ptr
        self.width, self.height, self.C = config.width, config.height, config.channels
        self.n = self.width * self.height

    def close(self) -> None:
        if getattr(self, "_ptr", None):
            self._lib.hp_emb_destroy(self._ptr)
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
            raise _abi.HipError(st, f"{what}: {self._lib.hp_emb_last_error(self._ptr).decode()}")

    def config(self) -> SimulationConfig:
        return self._config

    def embedding(self) -> EmbeddingConfig:
        return self._embedding

    # -- the reference's interface -----------------------------------------------------------
    def step(self, state: EmbeddedState) -> None:
        """EmbeddedPropagator::step: the state's channels, parameters, time and step in place."""
        self.run(state, 1)

    def run(self, state: EmbeddedState, steps: int) -> None:
        t, s = ctypes.c_float(state.time), ctypes.c_uint64(state.step)
        params = [p.data for p in state.params]
        self._check(self._lib.hp_emb_run(self._ptr, _abi.float_ptr_array(state.channels), _pp(params),
                                         ctypes.byref(t), ctypes.byref(s), steps), "hp_emb_run")
        state.time, state.step = t.value, s.value

    # -- device-resident state -----------------------------------------------------------------
    def set_state(self, state: EmbeddedState) -> None:
        self._check(self._lib.hp_emb_set_state(self._ptr, _abi.float_ptr_array(state.channels),
                                               _pp([p.data for p in state.params]), state.time, state.step),
                    "hp_emb_set_state")

    def get_state(self) -> EmbeddedState:
        ch = [np.empty(self.n, dtype=np.float32) for _ in range(self.C)]
        ps = [np.empty((self.n, 5), dtype=np.float32) for _ in range(self.C)]
        t, s = ctypes.c_float(), ctypes.c_uint64()
        self._check(self._lib.hp_emb_get_state(self._ptr, _abi.float_ptr_array(ch), _pp(ps), ctypes.byref(t),
                                               ctypes.byref(s)), "hp_emb_get_state")
        return EmbeddedState(ch, [ParameterGrid(self.width, self.height, data=p) for p in ps], self.width,
                             self.height, t.value, s.value)

    def advance(self, steps: int) -> None:
        self._check(self._lib.hp_emb_advance(self._ptr, steps), "hp_emb_advance")

    def mass_f64(self):
        per = (ctypes.c_double * self.C)()
        tot = ctypes.c_double()
        self._check(self._lib.hp_emb_mass_f64(self._ptr, per, ctypes.byref(tot)), "hp_emb_mass_f64")
        return list(per), tot.value

    def next_params(self) -> List[ParameterGrid]:
        """The propagator's next parameter grids (what destinations without a mix keep)."""
        ps = [np.empty((self.n, 5), dtype=np.float32) for _ in range(self.C)]
        self._check(self._lib.hp_emb_get_next_params(self._ptr, _pp(ps)), "hp_emb_get_next_params")
        return [ParameterGrid(self.width, self.height, data=p) for p in ps]

    def set_next_params(self, grids: List[ParameterGrid]) -> None:
        self._check(self._lib.hp_emb_set_next_params(self._ptr, _pp([g.data for g in grids])),
                    "hp_emb_set_next_params")

    def last_radius(self) -> List[int]:
        r = (ctypes.c_int32 * self.C)()
        self._check(self._lib.hp_emb_last_radius(self._ptr, r), "hp_emb_last_radius")
        return list(r)

    # -- stage access (parity tests) -------------------------------------------------------------
    def affinity(self, channels, params) -> List[np.ndarray]:
        out = [np.empty(self.n, dtype=np.float32) for _ in range(self.C)]
        self._check(self._lib.hp_emb_affinity(self._ptr, _pp(channels), _pp(params), _abi.float_ptr_array(out)),
                    "hp_emb_affinity")
        return out

    def stencil(self, channels, params, affinity, next_params):
        """Stages 2-5 with the affinity given; returns (next mass, next parameter arrays)."""
        nxt = [np.empty(self.n, dtype=np.float32) for _ in range(self.C)]
        nps = [np.array(p, dtype=np.float32).reshape(self.n, 5) for p in next_params]
        self._check(self._lib.hp_emb_stencil(self._ptr, _pp(channels), _pp(params), _pp(affinity),
                                             _abi.float_ptr_array(nxt), _pp(nps)), "hp_emb_stencil")
        return nxt, nps


def advect_mass_and_params(current_mass, current_params: ParameterGrid, flow_x, flow_y, config: EmbeddingConfig,
                           dt: float, distribution_size: float, width: int, height: int,
                           next_params: Optional[ParameterGrid] = None, device: int = 0):
    """advect_mass_and_params (param_advection.rs:47-79) on the GPU: returns (next mass, next
    parameter grid); the grid written into starts as `next_params` (default: from_defaults)."""
    lib = _abi.load_library()
    n = width * height
    m = np.ascontiguousarray(current_mass, dtype=np.float32).reshape(-1)
    fx = np.ascontiguousarray(flow_x, dtype=np.float32).reshape(-1)
    fy = np.ascontiguousarray(flow_y, dtype=np.float32).reshape(-1)
    pr = np.ascontiguousarray(current_params.data, dtype=np.float32)
    out_p = (next_params.copy() if next_params is not None else ParameterGrid.from_defaults(width, height))
    out_m = np.empty(n, dtype=np.float32)
    e = config.to_c()
    FP = ctypes.POINTER(ctypes.c_float)
    st = lib.hp_advect_mass_and_params(m.ctypes.data_as(FP), pr.ctypes.data_as(FP), fx.ctypes.data_as(FP),
                                       fy.ctypes.data_as(FP), ctypes.byref(e), dt, distribution_size, width, height,
                                       out_m.ctypes.data_as(FP), out_p.data.ctypes.data_as(FP), device)
    if st != _abi.HP_OK:
        raise _abi.HipError(st, lib.hp_emb_last_error(None).decode())
    return out_m, out_p
