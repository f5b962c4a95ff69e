"""Independent PBF scenes of uneven particle counts stepped together.

The reference steps one scene per PBF solver and stream
(domain/domain.ixx:121-124), each substep a fixed launch sequence
(pbf/pbf.cpp:167-191). PbfBatch runs that sequence for every scene at once:
the scenes' particles are concatenated in scene order (scene k holds
particles offsets[k] .. offsets[k + 1] - 1 of every array), each pass of a
substep is one launch over all scenes, and with static boxes a substep is
captured once in a HIP graph and replayed. Each scene's states equal its
own PBF solver's bit for bit. The scenes share the support radius, time
step and iteration count; boxes, gravity and particle counts are their own.
"""
from __future__ import annotations

import ctypes
import json
from pathlib import Path
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _capi
from .particle import PBF, Control, Domain, DomainConfiguration, _release_handle, _require_gpu, _SolverStream

TUNING = Path(__file__).resolve().parent / "tuning" / "pbf_batch.json"


def tuned_settings() -> dict:
    """The autotuner's choice for this GPU (physica_amd/tuning/pbf_batch.json),
    or the defaults (mask skin 0.3 h, graph replay)."""
    settings = {"mask_skin": 0.3, "use_graph": True}
    if TUNING.exists():
        try:
            data = json.loads(TUNING.read_text())
        except ValueError:
            return settings
        best = data.get("best", {})
        settings.update({k: best[k] for k in ("mask_skin", "use_graph") if k in best})
    return settings


class PbfBatch:
    """A batch of PBF scenes in forward mode (phx_pbf_batch_*)."""

    def __init__(self, domains: Sequence[DomainConfiguration], configuration: Optional[PBF.Configuration] = None,
                 mask_skin: Optional[float] = None, use_graph: Optional[bool] = None, device="cuda"):
        self.device = torch.device(device)
        _require_gpu(self.device)
        self.configuration = configuration or PBF.Configuration()
        tuned = tuned_settings()
        self.mask_skin = float(tuned["mask_skin"] if mask_skin is None else mask_skin)
        self.use_graph = bool(tuned["use_graph"] if use_graph is None else use_graph)
        self.scene_count = len(domains)
        self._domains_c = (_capi.DomainConfiguration * max(1, self.scene_count))()
        for k, d in enumerate(domains):
            self._domains_c[k] = d.to_c()
        self._stream = torch.cuda.current_stream(self.device)
        handle = ctypes.c_void_p()
        batch_cfg = _capi.PbfBatchConfiguration(self.mask_skin, 1 if self.use_graph else 0)
        _capi.call("phx_pbf_batch_create", self._domains_c, self.scene_count, ctypes.byref(self.configuration.to_c()),
                   ctypes.byref(batch_cfg), self._stream.cuda_stream, ctypes.byref(handle))
        self._handle = handle
        self._release = _capi.library().phx_pbf_batch_destroy
        offsets = (ctypes.c_uint64 * (self.scene_count + 1))()
        _capi.call("phx_pbf_batch_particle_offsets", self._handle, offsets)
        self.offsets: List[int] = [int(v) for v in offsets]
        self.particle_count = self.offsets[-1]
        first = domains[0]
        total = DomainConfiguration(self.particle_count, first.time_step, first.support_radius,
                                    first.particle_radius, first.gravity, first.boundary)
        self.domain = Domain(total, self.device)

    __del__ = _release_handle

    def allocate_state(self) -> PBF.State:
        return PBF.State(self.domain.allocate_particle_state())

    def allocate_control(self) -> Control:
        return self.domain.allocate_control()

    def allocate_parameters(self) -> PBF.Parameters:
        d = self.domain
        return PBF.Parameters(d.allocate_particle_parameters(), *(d.allocate_scalar_field() for _ in range(6)))

    def scene_slice(self, k: int) -> slice:
        return slice(self.offsets[k], self.offsets[k + 1])

    def upload(self, scenes, state: PBF.State, parameters: PBF.Parameters) -> None:
        """Concatenates the scenes' initial states and parameters (each a
        physica_amd.scenes.Scene) into the batch's arrays."""
        positions = np.concatenate([sc.positions for sc in scenes], axis=1)
        velocities = np.concatenate([sc.velocities for sc in scenes], axis=1)
        state.particles.positions.upload(positions)
        state.particles.velocities.upload(velocities)
        for name in ("masses", "rest_densities", "viscosities", "surface_tensions"):
            values = np.concatenate([sc.parameters[name] for sc in scenes]).astype(np.float32)
            getattr(parameters.particles, name).copy_(torch.from_numpy(values))
        for name in ("relaxation", "artificial_pressure_strength", "artificial_pressure_exponent",
                     "artificial_pressure_radius", "xsph_viscosity", "vorticity_confinement"):
            values = np.concatenate([sc.parameters[name] for sc in scenes]).astype(np.float32)
            getattr(parameters, name).copy_(torch.from_numpy(values))

    def run(self, state: PBF.State, control: Control, parameters: PBF.Parameters, steps: int) -> None:
        """`steps` forward substeps of every scene, in place on `state`."""
        c = state.to_c()
        with _SolverStream(self):
            _capi.call("phx_pbf_batch_run", self._handle, ctypes.byref(c), ctypes.byref(control.to_c()),
                       ctypes.byref(parameters.to_c()), int(steps))
        state.particles.step_index = int(c.particles.step_index)

    @property
    def kernels_per_step(self) -> int:
        """Kernel launches of one substep over all scenes."""
        v = ctypes.c_uint32()
        _capi.call("phx_pbf_batch_kernels_per_step", self._handle, ctypes.byref(v))
        return int(v.value)
