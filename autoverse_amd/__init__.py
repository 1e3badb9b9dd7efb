"""MI355X-native Flow Lenia time step: a drop-in for the reference crate's CpuPropagator.

The numerics live in libflowlenia_hip.so (hand-written HIP kernels for gfx950, C ABI in
include/flowlenia_hip.h); this package mirrors the reference's Rust API on top of it.
"""
from .schema import (BlobSpec, ConfigError, EmbeddingConfig, FlowConfig, KernelConfig, Pattern, RingConfig, Seed,
                     SimulationConfig, SphereSpec)
from .propagator import (HipPropagator, SimulationState, SimulationStats, advect_mass, advect_mass_3d,
                         compute_flow_field, growth_accumulate, sobel_gradient)

__all__ = ["BlobSpec", "ConfigError", "EmbeddingConfig", "FlowConfig", "KernelConfig", "Pattern", "RingConfig", "Seed",
           "SimulationConfig", "SphereSpec", "HipPropagator", "SimulationState", "SimulationStats",
           "advect_mass", "advect_mass_3d", "compute_flow_field", "growth_accumulate", "sobel_gradient"]
