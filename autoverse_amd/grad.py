"""Reverse-mode derivatives of the Flow Lenia step (no counterpart in the reference).

The loss is any L whose derivative with respect to the final state is supplied
(dL/dA_T); the tape gives dL/dA_0 and dL/dtheta, where theta is, per kernel in config
order, (weight, mu, sigma), then per ring in config order (amplitude, distance, width) --
the parameters of KernelConfig / RingConfig (src/schema/config.rs:55-101) that the step
depends on smoothly.  Work runs in libflowlenia_hip.so: a checkpointed forward tape and
hand-written adjoint kernels for every stage.

    FlowLeniaSteps.apply(a0, theta, cfg, steps, checkpoint_every) is a torch.autograd.Function
    over the same tape, for use inside a torch graph (tensors on the CPU; the function
    moves data across the C ABI itself).
"""
from __future__ import annotations

import copy
from typing import List, Sequence, Tuple

import numpy as np

from .population import HipPopulation
from .propagator import SimulationState
from .schema import SimulationConfig


def theta_of(cfg: SimulationConfig) -> np.ndarray:
    """The differentiable parameters of a config, float64, in the tape's layout."""
    out: List[float] = []
    for k in cfg.kernels:
        out += [k.weight, k.mu, k.sigma]
    for k in cfg.kernels:
        for r in k.rings:
            out += [r.amplitude, r.distance, r.width]
    return np.array(out, dtype=np.float64)


def with_theta(cfg: SimulationConfig, theta: Sequence[float]) -> SimulationConfig:
    """A copy of cfg with theta's values (rounded to f32 by the config) written back."""
    out = copy.deepcopy(cfg)
    th = [float(v) for v in theta]
    if len(th) != len(theta_of(cfg)):
        raise ValueError(f"theta has {len(th)} values, the config {len(theta_of(cfg))}")
    i = 0
    for k in out.kernels:
        k.weight, k.mu, k.sigma = th[i:i + 3]
        i += 3
    for k in out.kernels:
        for r in k.rings:
            r.amplitude, r.distance, r.width = th[i:i + 3]
            i += 3
    return out


def grad(cfg: SimulationConfig, state: SimulationState, steps: int, dl_dfinal: Sequence[np.ndarray],
         checkpoint_every: int = 8, device: int = 0) -> Tuple[SimulationState, List[np.ndarray], np.ndarray]:
    """Final state after `steps` steps from `state`, dL/dA_0 and dL/dtheta for the loss whose
    derivative with respect to the final state is dl_dfinal."""
    with HipPopulation([cfg], device) as pop:
        pop.set_state(0, state)
        pop.tape_forward(steps, checkpoint_every)
        final = pop.get_state(0)
        g0, th = pop.tape_backward([list(dl_dfinal)])
    return final, g0[0], th[0]


def _function():
    import torch

    class FlowLeniaSteps(torch.autograd.Function):
        """a_T = FlowLeniaSteps.apply(a0 [C, H, W] float32, theta [P] float64, cfg, steps, k)."""

        @staticmethod
        def forward(ctx, a0, theta, cfg, steps, checkpoint_every=8):
            c = with_theta(cfg, theta.detach().cpu().numpy())
            C, H, W = a0.shape
            pop = HipPopulation([c])
            chans = [a0[i].detach().cpu().numpy().astype(np.float32).reshape(-1) for i in range(C)]
            pop.set_state(0, SimulationState(chans, W, H))
            pop.tape_forward(int(steps), int(checkpoint_every))
            final = pop.get_state(0)
            ctx.pop = pop
            ctx.shape = (C, H, W)
            ctx.theta_dtype = theta.dtype
            return torch.from_numpy(np.stack(final.channels).reshape(C, H, W)).to(a0.device)

        @staticmethod
        def backward(ctx, grad_out):
            C, H, W = ctx.shape
            seed = [grad_out[i].detach().cpu().numpy().astype(np.float32).reshape(-1) for i in range(C)]
            g0, th = ctx.pop.tape_backward([seed])
            ctx.pop.close()
            ga = torch.from_numpy(np.stack(g0[0]).reshape(C, H, W)).to(grad_out.device)
            gt = torch.from_numpy(th[0]).to(dtype=ctx.theta_dtype)
            return ga, gt, None, None, None

    return FlowLeniaSteps


def __getattr__(name):
    if name == "FlowLeniaSteps":
        return _function()
    raise AttributeError(name)
