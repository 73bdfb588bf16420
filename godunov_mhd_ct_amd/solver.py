"""SolverGodunovMHD counterpart: owns U, U2, Bface, Bface2 (and U_RK, Bface_RK under RK2_SSP)
in HBM and drives the step.

One step follows SolverGodunovMHD::next_iteration_impl (SolverGodunovMHD.cpp:409-484):

- HANCOCK (godunov_unsplit, :494-549): compute_dt, fill the outside blocks, do_time_step
  (with U2 == U, B2 == B on entry), then the new state becomes current.
- RK2_SSP (:551-647): fill the outside blocks, U_RK = U, do_time_step(U -> U_RK), combine with
  the stage-1 coefficients, fill U_RK's outside blocks, U2 = U_RK, do_time_step(U_RK -> U2),
  U2 = 1/2 U + 1/2 U2 (stage-2 coefficients).

The reference copies U2 back into U; here the buffers swap roles and the copy moves to the
start of the next step (the same traffic, the same contract). The fused forms write every owned
value of their output from the input alone, so they skip these copies (and RK2_SSP's U_RK = U and
U2 = U_RK): the owned results are the same, and the outside blocks are refilled before use.

`do_amr_cycle` is the reference's AMR cycle (:263-319) on this mesh: refinement flags on the
device, adaptation and balance of a copy of the host octree, remap of U and face B on the
device (amr.py); `next_iteration` runs it every `cycle_frequency` steps before the step, as
next_iteration_impl does (:409-418).
"""
from __future__ import annotations

import numpy as np

from .amr import AMRSettings, Remap, adapted, refine_flags
from .implem import (IMPL_PLAIN, RK2_SSP_STAGE1_COEFS, RK2_SSP_STAGE2_COEFS, TIME_INTEGRATOR, GodunovImplemHIP,
                     hydro_settings)
from .mesh import Mesh


class SolverGodunovMHD:
    def __init__(self, mesh: Mesh, U: np.ndarray, B: np.ndarray, settings=None, impl_version: int = IMPL_PLAIN,
                 device: str = "cuda"):
        import torch

        self.torch = torch
        self.mesh = mesh if mesh.finalized else mesh.finalize()
        self.settings = settings if settings is not None else hydro_settings()
        self.implem = GodunovImplemHIP(self.settings, mesh.dim, mesh.b, impl_version)
        self.implem.resize_auxiliary_data(self.mesh)
        nt = mesh.num_total
        if U.shape != (nt, 5, mesh.cells_per_block) and U.size != nt * 5 * mesh.cells_per_block:
            raise ValueError("U has the wrong size for this mesh")
        if B.size != nt * mesh.faces_per_block:
            raise ValueError("B has the wrong size for this mesh")
        self.U = torch.from_numpy(np.ascontiguousarray(U, dtype=np.float64).reshape(-1)).to(device)
        self.B = torch.from_numpy(np.ascontiguousarray(B, dtype=np.float64).reshape(-1)).to(device)
        self.U2 = self.U.clone()
        self.B2 = self.B.clone()
        self.rk2 = self.settings.time_integrator == TIME_INTEGRATOR["RK2_SSP"]
        # the plain form accumulates into U2 (U2 == U on entry, the reference's contract); the
        # fused forms write every owned value of U2 and B2 from U and B, so they skip the copies
        self.copies = impl_version == IMPL_PLAIN
        if self.rk2:
            self.U_RK = self.U.clone()
            self.B_RK = self.B.clone()
        self.t = 0.0
        self.iteration = 0
        self.amr_cycles = 0
        self.dt = None
        self.total_num_cell_updates = 0

    def compute_dt(self) -> float:
        self.dt = self.implem.compute_dt(self.U, self.B)
        return self.dt

    def step(self, dt: float | None = None) -> float:
        if self.mesh.num_outside:
            self.implem.fill_outside(self.U, self.B)
        if dt is None:
            dt = self.compute_dt()
        if self.rk2:
            self._rk2_stages(dt)
        else:
            if self.copies:
                self.U2.copy_(self.U)
                self.B2.copy_(self.B)
            self.implem.do_time_step(self.U, self.U2, self.B, self.B2, dt)
        self.U, self.U2 = self.U2, self.U
        self.B, self.B2 = self.B2, self.B
        self.t += dt
        self.iteration += 1
        # SolverGodunovMHD.cpp:482
        self.total_num_cell_updates += self.mesh.num_leaves * self.mesh.cells_per_block
        return dt

    def _rk2_stages(self, dt: float):
        # SolverGodunovMHD.cpp:555-622, after the outside fill of U
        im = self.implem
        if self.copies:
            self.U_RK.copy_(self.U)
            self.B_RK.copy_(self.B)
        im.do_time_step(self.U, self.U_RK, self.B, self.B_RK, dt)
        im.linear_combination(self.U, self.U_RK, self.B, self.B_RK, RK2_SSP_STAGE1_COEFS)
        if self.mesh.num_outside:
            im.fill_outside(self.U_RK, self.B_RK)
        if self.copies:
            self.U2.copy_(self.U_RK)
            self.B2.copy_(self.B_RK)
        im.do_time_step(self.U_RK, self.U2, self.B_RK, self.B2, dt)
        im.linear_combination(self.U, self.U2, self.B, self.B2, RK2_SSP_STAGE2_COEFS)

    def do_amr_cycle(self, amr: AMRSettings):
        """One AMR cycle; returns the (copied, prolonged, restricted) leaf counts, or None when
        no leaf is flagged (the mesh and the state are then unchanged)."""
        torch = self.torch
        if self.mesh.num_outside:
            self.implem.fill_outside(self.U, self.B)
        flags, _ = refine_flags(self.implem, self.U, amr, B=self.B)
        self.amr_cycles += 1
        if not flags.any():
            return None
        new_mesh = adapted(self.mesh, flags)
        rm = Remap(self.mesh, new_mesh)
        nt = new_mesh.num_total
        Un = torch.zeros(nt * 5 * new_mesh.cells_per_block, dtype=torch.float64, device=self.U.device)
        Bn = torch.zeros(nt * new_mesh.faces_per_block, dtype=torch.float64, device=self.U.device)
        rm.apply(self.U, self.B, Un, Bn)
        self.implem.resize_auxiliary_data(new_mesh)
        self.mesh = new_mesh
        self.U, self.B = Un, Bn
        self.U2, self.B2 = Un.clone(), Bn.clone()
        if self.rk2:
            self.U_RK, self.B_RK = Un.clone(), Bn.clone()
        return rm.counts()

    def should_do_amr_cycle(self, amr: AMRSettings | None, disable_first: bool = False) -> bool:
        # no cycle when level_min == level_max (SolverGodunovMHD.cpp:413)
        if amr is None or amr.cycle_frequency <= 0 or amr.level_min >= amr.level_max:
            return False
        if self.iteration == 0 and disable_first:
            return False
        return self.iteration % amr.cycle_frequency == 0

    def next_iteration(self, amr: AMRSettings | None = None, disable_first: bool = False) -> float:
        """next_iteration_impl: the AMR cycle when due, then one step."""
        if self.should_do_amr_cycle(amr, disable_first):
            self.do_amr_cycle(amr)
        return self.step()

    def state(self):
        nt = self.mesh.num_total
        U = self.U.cpu().numpy().reshape(nt, 5, -1)
        B = self.B.cpu().numpy().reshape(nt, -1)
        return U, B
