"""Differentiable PBF rollouts: forward steps recorded in step caches, then
PBF::vjp_step chained backwards (pbf/pbf.cpp:217-258).

`checkpoint_every=1` keeps every step cache resident (about 260 bytes per
particle per step at 4 constraint iterations, so a 262,144-particle,
120-step rollout holds roughly 8 GB of HBM and the backward pass recomputes
nothing). `checkpoint_every=k` keeps only every k-th state and recomputes each
k-step segment's caches before differentiating it.
"""
from __future__ import annotations

from typing import List, Optional

import torch

from .particle import PBF, Control, ParticleParameterAdjoint


class PbfRollout:
    """A fixed-length rollout of one PBF solver created in differentiable mode."""

    def __init__(self, solver: PBF, steps: int, checkpoint_every: int = 1, parameter_gradients: bool = True):
        if steps < 0 or checkpoint_every < 1:
            raise ValueError("steps must be >= 0 and checkpoint_every >= 1")
        self.solver = solver
        self.steps = steps
        self.checkpoint_every = checkpoint_every
        # without parameter gradients the VJPs receive null parameter adjoints
        # and skip the terms that only feed them
        self.parameter_gradients = parameter_gradients
        self._states: List[Optional[PBF.State]] = []
        self._caches: List[Optional[PBF.StepCache]] = []
        self._control: Optional[Control] = None
        self._parameters: Optional[PBF.Parameters] = None

    def _step(self, state: PBF.State, keep_cache: bool):
        nxt = self.solver.allocate_state()
        cache = self.solver.allocate_step_cache()
        self.solver.forward_step(state, self._control, self._parameters, nxt, cache)
        return nxt, cache if keep_cache else None

    def forward(self, state: PBF.State, control: Control, parameters: PBF.Parameters) -> PBF.State:
        """Runs the rollout from `state` (not modified) and returns the final state."""
        self._control, self._parameters = control, parameters
        start = self.solver.allocate_state()
        self.solver.copy_state(state, start)
        start.particles.step_index = state.particles.step_index
        self._states = [start]
        self._caches = []
        current = start
        for k in range(self.steps):
            keep = self.checkpoint_every == 1
            nxt, cache = self._step(current, keep)
            self._caches.append(cache)
            checkpoint = self.checkpoint_every == 1 or (k + 1) % self.checkpoint_every == 0 or k + 1 == self.steps
            self._states.append(nxt if checkpoint else None)
            current = nxt
        return current

    def backward(self, final_adjoint: PBF.StateAdjoint):
        """Chains the VJPs from the adjoint of the final state. Returns
        (adjoint of the initial state, control adjoint, parameter adjoint),
        all freshly allocated and accumulated over every step (the parameter
        adjoint is None without parameter gradients)."""
        solver = self.solver
        control_adjoint = solver.allocate_control_adjoint()
        if self.parameter_gradients:
            parameter_adjoint = solver.allocate_parameter_adjoint()
        else:
            parameter_adjoint = PBF.ParameterAdjoint(ParticleParameterAdjoint(None, None, None, None),
                                                     None, None, None, None, None, None)
        adjoint = solver.allocate_state_adjoint()
        solver.copy_state_adjoint(final_adjoint, adjoint)
        k = self.steps
        while k > 0:
            first = ((k - 1) // self.checkpoint_every) * self.checkpoint_every
            if self.checkpoint_every == 1:
                states, caches = [self._states[k - 1], self._states[k]], [self._caches[k - 1]]
            else:
                # recompute the segment [first, k) from its checkpoint
                states, caches = [self._states[first]], []
                for _ in range(first, k):
                    nxt, cache = self._step(states[-1], True)
                    states.append(nxt)
                    caches.append(cache)
            for local in reversed(range(len(caches))):
                previous = solver.allocate_state_adjoint()
                solver.vjp_step(states[local], self._control, self._parameters, states[local + 1], caches[local],
                                adjoint, previous, control_adjoint, parameter_adjoint)
                adjoint = previous
            k = first if self.checkpoint_every > 1 else k - 1
        return adjoint, control_adjoint, parameter_adjoint if self.parameter_gradients else None

    def release(self) -> None:
        """Frees the recorded states and caches."""
        self._states, self._caches = [], []
        torch.cuda.synchronize()
