"""HipPopulation: many independent 2D simulations stepped together on one GPU.

The reference scores a population by building one CpuPropagator per candidate and running
FitnessEvaluator::evaluate on each (src/compute/evolution/fitness.rs:26-88), in parallel
over rayon (src/compute/evolution/search.rs:153-172).  Here the whole population shares one
device arena; a step is one launch per FFT pass per grid size plus one fused-stencil launch
per channel count, replayed from a hipGraph.  Members may differ in size, channel count,
kernels and flow parameters.  Each member's trajectory is bit-identical to the one a lone
HipPropagator computes for it.

Fitness here is the displacement of the mass centroid between the sample after the warmup
and the last one (EvaluationTrajectory::to_behavior_stats, fitness.rs:166-191, with the
centre of mass of fitness.rs:196-219 reduced in FP64 on the device).
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _abi
from .propagator import SimulationState
from .schema import Seed, SimulationConfig


class HipPopulation:
    def __init__(self, configs: Sequence[SimulationConfig], device: int = 0):
        if not configs:
            raise ValueError("a population needs at least one member")
        self._lib = _abi.load_library()
        self._configs = list(configs)
        self._cs = [c.to_c() for c in self._configs]  # keeps the kernel/ring arrays alive
        arr = (_abi.HpConfig * len(self._cs))()
        for i, c in enumerate(self._cs):
            arr[i] = c.cfg
        self._arr = arr
        ptr = ctypes.c_void_p()
        st = self._lib.hp_batch_create(arr, len(self._cs), device, ctypes.byref(ptr))
        if st != _abi.HP_OK:
            raise _abi.HipError(st, self._lib.hp_batch_last_error(None).decode())
        self._ptr = ptr

    def __len__(self) -> int:
        return len(self._configs)

    def close(self) -> None:
        if getattr(self, "_ptr", None):
            self._lib.hp_batch_destroy(self._ptr)
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
            raise _abi.HipError(st, f"{what}: {self._lib.hp_batch_last_error(self._ptr).decode()}")

    def config(self, member: int) -> SimulationConfig:
        return self._configs[member]

    def set_state(self, member: int, state: SimulationState) -> None:
        cfg = self._configs[member]
        if len(state.channels) != cfg.channels or state.grid_size() != cfg.width * cfg.height:
            raise ValueError("state does not match the member's configuration")
        chans = [np.ascontiguousarray(c, dtype=np.float32) for c in state.channels]
        self._check(self._lib.hp_batch_set_state(self._ptr, member, _abi.float_ptr_array(chans)),
                    "hp_batch_set_state")

    def set_seeds(self, seeds: Sequence[Seed]) -> None:
        for i, s in enumerate(seeds):
            self.set_state(i, SimulationState.from_seed(s, self._configs[i]))

    def get_state(self, member: int) -> SimulationState:
        cfg = self._configs[member]
        n = cfg.width * cfg.height
        out = [np.empty(n, dtype=np.float32) for _ in range(cfg.channels)]
        t, s = ctypes.c_float(0.0), ctypes.c_uint64(0)
        self._check(self._lib.hp_batch_get_state(self._ptr, member, _abi.float_ptr_array(out), ctypes.byref(t),
                                                 ctypes.byref(s)), "hp_batch_get_state")
        return SimulationState(out, cfg.width, cfg.height, 1, t.value, s.value)

    def advance(self, steps: int) -> None:
        """Step every member `steps` times on the device."""
        self._check(self._lib.hp_batch_advance(self._ptr, int(steps)), "hp_batch_advance")

    def centroids(self) -> Tuple[np.ndarray, np.ndarray]:
        """(M, 2) centres of mass of the channel sums and their (M,) masses, both FP64."""
        m = len(self)
        xy = np.empty(2 * m, dtype=np.float64)
        mass = np.empty(m, dtype=np.float64)
        self._check(self._lib.hp_batch_centroids(self._ptr, xy.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                 mass.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                    "hp_batch_centroids")
        return xy.reshape(m, 2), mass

    def mass_f64(self) -> np.ndarray:
        """Per-member total mass, FP64, in the same summation order as HipPropagator.mass_f64."""
        out = np.empty(len(self), dtype=np.float64)
        self._check(self._lib.hp_batch_mass_f64(self._ptr, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                    "hp_batch_mass_f64")
        return out

    def stats(self) -> Tuple[int, int]:
        r, g = ctypes.c_uint32(0), ctypes.c_uint64(0)
        self._check(self._lib.hp_batch_stats(self._ptr, ctypes.byref(r), ctypes.byref(g)), "hp_batch_stats")
        return r.value, g.value

    # -- execution plan ----------------------------------------------------------------------
    def plan(self) -> dict:
        p = _abi.HpPlan()
        self._check(self._lib.hp_batch_get_plan(self._ptr, ctypes.byref(p)), "hp_batch_get_plan")
        return p.as_dict()

    def set_plan(self, plan: dict) -> None:
        p = _abi.HpPlan.from_dict({**self.plan(), **plan})
        self._check(self._lib.hp_batch_set_plan(self._ptr, ctypes.byref(p)), "hp_batch_set_plan")

    STEP_MODES = {"multi_launch": 0, "persistent": 1}

    def set_step_mode(self, mode: str) -> None:
        """'multi_launch' (default: one launch per FFT pass and size class, one stencil launch per
        channel-count group) or 'persistent' (one launch per step with grid barriers between the
        passes; member sides 64, 128, 192 or 256 and radix order 0).  Trajectories are identical."""
        self._check(self._lib.hp_batch_set_step_mode(self._ptr, self.STEP_MODES[mode]), "hp_batch_set_step_mode")

    def step_mode(self) -> Tuple[str, int]:
        """(mode, grid-barrier timeouts so far: runs recomputed through the checked path)."""
        m, t = ctypes.c_int32(0), ctypes.c_uint64(0)
        self._check(self._lib.hp_batch_step_mode(self._ptr, ctypes.byref(m), ctypes.byref(t)), "hp_batch_step_mode")
        return {v: k for k, v in self.STEP_MODES.items()}[m.value], t.value

    def autotune(self, steps: int = 20):
        """Time the LDS FFT variants and stencil tiles on the population from its current states
        (restored afterwards); keeps the fastest.  Returns (best plan, [(plan, us per step)])."""
        st, best, table = _abi.autotune_call(self._lib.hp_batch_autotune, self._ptr, steps)
        self._check(st, "hp_batch_autotune")
        return best, table

    # -- gradient tape ---------------------------------------------------------------------
    def theta_count(self, member: int) -> int:
        n = ctypes.c_uint32(0)
        self._check(self._lib.hp_batch_theta_count(self._ptr, member, ctypes.byref(n)), "hp_batch_theta_count")
        return n.value

    def tape_forward(self, steps: int, checkpoint_every: int = 0) -> int:
        """Advance every member `steps` steps, keeping a state every `checkpoint_every` steps
        for tape_backward (0: chosen from free device memory).  Returns the interval used."""
        used = ctypes.c_uint32(0)
        self._check(self._lib.hp_batch_tape_forward(self._ptr, int(steps), int(checkpoint_every),
                                                    ctypes.byref(used)), "hp_batch_tape_forward")
        return used.value

    def tape_backward(self, dl_dfinal: Sequence[Sequence[np.ndarray]]):
        """For L = sum over members and channels of <dl_dfinal, A_T>: per member, the list of
        dL/dA_0 channels and dL/dtheta (per kernel weight, mu, sigma; per ring amplitude,
        distance, width), from the last tape_forward."""
        seeds, outs = [], []
        for i, cfg in enumerate(self._configs):
            n = cfg.width * cfg.height
            if len(dl_dfinal[i]) != cfg.channels:
                raise ValueError(f"member {i}: expected {cfg.channels} channels of dL/dA_T")
            for c in range(cfg.channels):
                a = np.ascontiguousarray(dl_dfinal[i][c], dtype=np.float32).reshape(-1)
                if a.size != n:
                    raise ValueError(f"member {i}: dL/dA_T channel has {a.size} cells, expected {n}")
                seeds.append(a)
                outs.append(np.empty(n, dtype=np.float32))
        counts = [self.theta_count(i) for i in range(len(self))]
        dth = np.zeros(max(1, sum(counts)), dtype=np.float64)
        self._check(self._lib.hp_batch_tape_backward(self._ptr, _abi.float_ptr_array(seeds),
                                                     _abi.float_ptr_array(outs),
                                                     dth.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                    "hp_batch_tape_backward")
        g0, th, ci, ti = [], [], 0, 0
        for i, cfg in enumerate(self._configs):
            g0.append(outs[ci:ci + cfg.channels])
            ci += cfg.channels
            th.append(dth[ti:ti + counts[i]].copy())
            ti += counts[i]
        return g0, th

    def evaluate(self, fitness_config, eval_config):
        """FitnessEvaluator::evaluate (fitness.rs:26-86) for every member from its current state:
        per member (combined fitness, [MetricResult], BehaviorStats); see autoverse_amd.fitness."""
        from .fitness import evaluate_population
        return evaluate_population(self, fitness_config, eval_config)

    def evaluate_displacement(self, seeds: Optional[Sequence[Seed]], warmup: int, steps: int) -> np.ndarray:
        """Centroid displacement of every member over `steps` steps after `warmup` steps
        (FitnessEvaluator::evaluate with the trajectory's total_displacement)."""
        if seeds is not None:
            self.set_seeds(seeds)
        if warmup:
            self.advance(warmup)
        c0, _ = self.centroids()
        if steps:
            self.advance(steps)
        c1, _ = self.centroids()
        return np.hypot(c1[:, 0] - c0[:, 0], c1[:, 1] - c0[:, 1])
