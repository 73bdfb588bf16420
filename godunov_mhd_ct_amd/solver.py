"""SolverGodunovMHD counterpart: owns U, U2, Bface, Bface2 in HBM and drives the step.

One step follows SolverGodunovMHD::next_iteration_impl with the HANCOCK integrator
(SolverGodunovMHD.cpp:409-484, 490-549): compute_dt, fill the outside blocks, do_time_step
(with U2 == U, B2 == B on entry), then the new state becomes current.  The reference copies
U2 back into U; here the buffers swap roles and the copy moves to the start of the next step
(the same traffic, the same contract).
"""
from __future__ import annotations

import numpy as np

from .implem import IMPL_PLAIN, GodunovImplemHIP, hydro_settings
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
        self.t = 0.0
        self.iteration = 0
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

    def state(self):
        nt = self.mesh.num_total
        U = self.U.cpu().numpy().reshape(nt, 5, -1)
        B = self.B.cpu().numpy().reshape(nt, -1)
        return U, B
