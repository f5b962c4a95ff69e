"""torch.autograd view of a differentiable PBF rollout.

    x_T, v_T = pbf_rollout(x0, v0, rollout, control, parameters)

x0 and v0 are (3, N) float32 CUDA tensors (SoA rows x, y, z); the backward
pass chains PBF::vjp_step in FP64 and returns the gradients of x0 and v0
cast to float32. The control and parameter adjoints of the last backward
pass are kept on the rollout object (`last_control_adjoint`,
`last_parameter_adjoint`)."""
from __future__ import annotations

import torch

from .rollout import PbfRollout


def _fill(field, rows: torch.Tensor) -> None:
    field.x.copy_(rows[0])
    field.y.copy_(rows[1])
    field.z.copy_(rows[2])


def _stack(field) -> torch.Tensor:
    return torch.stack([field.x, field.y, field.z])


class PbfRolloutFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0: torch.Tensor, v0: torch.Tensor, rollout: PbfRollout, control, parameters):
        solver = rollout.solver
        state = solver.allocate_state()
        _fill(state.particles.positions, x0.detach().to(torch.float32).contiguous())
        _fill(state.particles.velocities, v0.detach().to(torch.float32).contiguous())
        final = rollout.forward(state, control, parameters)
        ctx.rollout = rollout
        return _stack(final.particles.positions).clone(), _stack(final.particles.velocities).clone()

    @staticmethod
    def backward(ctx, grad_x: torch.Tensor, grad_v: torch.Tensor):
        rollout: PbfRollout = ctx.rollout
        solver = rollout.solver
        adjoint = solver.allocate_state_adjoint()
        if grad_x is not None:
            _fill(adjoint.particles.positions, grad_x.to(torch.float64))
        if grad_v is not None:
            _fill(adjoint.particles.velocities, grad_v.to(torch.float64))
        initial, control_adjoint, parameter_adjoint = rollout.backward(adjoint)
        rollout.last_control_adjoint = control_adjoint
        rollout.last_parameter_adjoint = parameter_adjoint
        gx = _stack(initial.particles.positions).to(torch.float32)
        gv = _stack(initial.particles.velocities).to(torch.float32)
        return gx, gv, None, None, None


def pbf_rollout(x0: torch.Tensor, v0: torch.Tensor, rollout: PbfRollout, control, parameters):
    """Differentiable rollout: returns (x_T, v_T) as (3, N) tensors."""
    return PbfRolloutFunction.apply(x0, v0, rollout, control, parameters)
