"""Fitness evaluation on the GPU: the reference's FitnessEvaluator over HipPopulation.

Mirrors src/compute/evolution/fitness.rs (FitnessEvaluator::evaluate :26-86, MetricResult
:90-95, the metrics :247-520) and the schema types of src/schema/evolution.rs (FitnessConfig
:273-305, WeightedMetric :309-314, FitnessMetric :319-375 with its serde forms, EvaluationConfig
:479-517, BehaviorStats :1015-1032).  A whole population is evaluated at once by
hp_batch_evaluate: the members step on the device, every sample's sums are FP64 reductions on
the device, and the metric formulas run in f32 on the host as the reference writes them.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _abi
from .schema import Seed, SimulationConfig


@dataclass
class FitnessMetric:
    """FitnessMetric (evolution.rs:319-375): ``type`` is the serde tag; parameters of the
    variants that take them."""
    type: str
    period: int = 0
    tolerance: float = 0.0
    min_displacement: float = 0.0
    max_period: int = 0
    threshold: float = 0.0
    name: str = ""

    @classmethod
    def from_json_value(cls, v) -> "FitnessMetric":
        """Either the bare string of a parameterless variant or the tagged object
        (FitnessMetric's Deserialize, evolution.rs:377-443)."""
        if isinstance(v, str):
            if v not in ("Persistence", "Compactness", "Locomotion", "Complexity", "MassConcentration",
                         "Stability"):
                raise ValueError(f"Unknown fitness metric string: {v}")
            return cls(v)
        if isinstance(v, dict):
            t = v.get("type")
            if t not in _abi.METRIC_KINDS:
                raise ValueError(f"Invalid fitness metric format: {v}")
            return cls(t, **{k: x for k, x in v.items() if k != "type"})
        raise ValueError(f"Invalid fitness metric format: {v}")

    def to_dict(self) -> dict:
        d = {"type": self.type}
        if self.type == "Periodicity":
            d.update(period=self.period, tolerance=self.tolerance)
        elif self.type == "GliderScore":
            d.update(min_displacement=self.min_displacement)
        elif self.type == "OscillatorScore":
            d.update(max_period=self.max_period, threshold=self.threshold)
        elif self.type == "Custom":
            d.update(name=self.name)
        return d


@dataclass
class WeightedMetric:
    """WeightedMetric (evolution.rs:309-314)."""
    metric: FitnessMetric
    weight: float


@dataclass
class FitnessConfig:
    """FitnessConfig (evolution.rs:273-305); default Persistence 1.0 + Compactness 0.5."""
    metrics: List[WeightedMetric] = field(default_factory=lambda: [
        WeightedMetric(FitnessMetric("Persistence"), 1.0), WeightedMetric(FitnessMetric("Compactness"), 0.5)])
    archive_threshold: Optional[float] = None
    normalize: bool = True


@dataclass
class EvaluationConfig:
    """EvaluationConfig (evolution.rs:479-517): 200 steps, no warmup, a sample every 10."""
    steps: int = 200
    warmup_steps: int = 0
    sample_interval: int = 10
    use_gpu: bool = False
    parallel_workers: int = 0


@dataclass
class MetricResult:
    """MetricResult (fitness.rs:90-95)."""
    metric: FitnessMetric
    score: float
    weight: float


@dataclass
class BehaviorStats:
    """BehaviorStats (evolution.rs:1015-1032)."""
    final_mass: float
    initial_mass: float
    center_of_mass_trajectory: List[Tuple[float, float]]
    total_displacement: float
    radius_over_time: List[float]
    final_radius: float
    active_cells: int
    max_activation: float
    # the per-sample masses, peaks and active-cell counts the trajectory recorded as well
    mass_over_time: List[float] = field(default_factory=list)
    max_over_time: List[float] = field(default_factory=list)
    active_over_time: List[int] = field(default_factory=list)


def _metric_structs(fc: FitnessConfig):
    arr = (_abi.HpMetric * max(1, len(fc.metrics)))()
    for i, wm in enumerate(fc.metrics):
        m = wm.metric
        arr[i] = _abi.HpMetric(_abi.METRIC_KINDS[m.type], wm.weight, int(m.period), m.tolerance,
                               m.min_displacement, int(m.max_period), m.threshold)
    return arr


def evaluate_population(pop, fitness: FitnessConfig, ev: EvaluationConfig):
    """hp_batch_evaluate over every member of a HipPopulation from its current state.  Returns
    per member (combined fitness, [MetricResult], BehaviorStats); the members end in their final
    states."""
    lib = pop._lib
    e = _abi.HpEvalConfig(int(ev.steps), int(ev.warmup_steps), int(ev.sample_interval))
    n = ctypes.c_uint32(0)
    st = lib.hp_eval_sample_count(ctypes.byref(e), ctypes.byref(n))
    if st != _abi.HP_OK:
        raise _abi.HipError(st, "hp_eval_sample_count")
    M, K, S = len(pop), len(fitness.metrics), n.value
    mets = _metric_structs(fitness)
    fit = np.zeros(M, np.float32)
    scores = np.zeros(max(1, M * K), np.float32)
    beh = (_abi.HpBehavior * M)()
    smp = (_abi.HpSample * (M * S))()
    fp = ctypes.POINTER(ctypes.c_float)
    pop._check(lib.hp_batch_evaluate(pop._ptr, ctypes.byref(e), mets, K, int(fitness.normalize),
                                     fit.ctypes.data_as(fp), scores.ctypes.data_as(fp), beh, smp),
               "hp_batch_evaluate")
    out = []
    for i in range(M):
        res = [MetricResult(wm.metric, float(scores[i * K + k]), wm.weight) for k, wm in enumerate(fitness.metrics)]
        b = beh[i]
        ss = [smp[i * S + j] for j in range(S)]
        bs = BehaviorStats(b.final_mass, b.initial_mass, [(s.center_x, s.center_y) for s in ss],
                           b.total_displacement, [s.radius for s in ss], b.final_radius, int(b.active_cells),
                           b.max_activation, [s.mass for s in ss], [s.max_value for s in ss],
                           [int(s.active_cells) for s in ss])
        out.append((float(fit[i]), res, bs))
    return out


class FitnessEvaluator:
    """FitnessEvaluator (fitness.rs:10-86) on the GPU: evaluate() builds a one-member population
    from (config, seed) as the reference builds a CpuPropagator; evaluate_many() scores a whole
    list of candidates in one population, as EvolutionEngine::evaluate_population
    (search.rs:153-172) scores a generation."""

    def __init__(self, config: FitnessConfig, eval_config: EvaluationConfig, device: int = 0):
        self.config = config
        self.eval_config = eval_config
        self.device = device

    def evaluate(self, sim_config: SimulationConfig, seed: Seed):
        return self.evaluate_many([(sim_config, seed)])[0]

    def evaluate_many(self, candidates: Sequence[Tuple[SimulationConfig, Seed]]):
        from .population import HipPopulation
        with HipPopulation([c for c, _ in candidates], device=self.device) as pop:
            pop.set_seeds([s for _, s in candidates])
            return evaluate_population(pop, self.config, self.eval_config)
