"""Batches: many meshes advanced together through one device work list per stage.

The reference advances one mesh per process (SolverGodunovMHD::next_iteration_impl,
SolverGodunovMHD.cpp:409-484).  A `BatchGodunovMHD` holds any number of independent solves
("slots") of one dimension -- each with its own mesh, block size (8, 16 or 32), settings and
CFL dt -- and advances them together: per batch step the outside blocks are filled, each slot's
dt = cfl / inv_dt is reduced on the device, and the fused stages run once over the
concatenated work list of all slots.  Slots join and leave between steps; a slot's result is
the same alone and in any mix.  The state of each slot stays in HBM and ping-pongs between two
buffers.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import MhdctError, Settings, lib
from .implem import IMPL_FUSED, GodunovImplemHIP, _stream, hydro_settings
from .mesh import Mesh


class _Slot:
    def __init__(self, implem, U, U2, B, B2, mesh):
        self.implem = implem
        self.bufs = ((U, B), (U2, B2))
        self.mesh = mesh


class BatchGodunovMHD:
    def __init__(self, dim: int, impl_version: int = IMPL_FUSED):
        self.dim = int(dim)
        self.impl_version = int(impl_version)
        h = C.c_void_p()
        if lib().mhdct_batch_create(self.dim, self.impl_version, C.byref(h)) != 0:
            raise MhdctError("mhdct_batch_create failed (dim 2 or 3, impl 11 or 12)")
        self._h = h
        self.slots: dict[int, _Slot] = {}
        self.front = False

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().mhdct_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().mhdct_batch_last_error(self._h)
            raise MhdctError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    @staticmethod
    def footprint_bytes(mesh: Mesh, impl_version: int = IMPL_FUSED) -> int:
        """HBM a slot needs: its state (two buffer pairs) and the fused scratch, estimated
        from the block count (used for admission before the slot is built)."""
        b, dim, n = mesh.b, mesh.dim, mesh.num_total
        g2 = (b + 4) ** dim
        g1 = (b + 2) ** dim
        nb = b + 4
        bg = 3 * (nb + 1) * nb * nb if dim == 3 else 2 * (nb + 1) * nb + nb * nb
        per = 5 * g2 + bg + 8 * g1 + dim * 8 * g1 + dim * g1 + (6 if dim == 3 else 2) * g1 \
            + dim * 12 * b ** (dim - 1) + (3 if dim == 3 else 1) * ((b + 1) ** dim + (b + 3) ** dim)
        scratch = mesh.num_leaves * per * (4 if impl_version == 12 else 8)
        state = 2 * n * (5 * mesh.cells_per_block + mesh.faces_per_block) * 8
        return int(scratch + state)

    def add(self, mesh: Mesh, U: np.ndarray, B: np.ndarray, settings: Settings | None = None,
            device: str = "cuda") -> int:
        import torch

        mesh = mesh if mesh.finalized else mesh.finalize()
        if mesh.dim != self.dim:
            raise MhdctError("mesh dimension differs from the batch")
        st = settings if settings is not None else hydro_settings()
        implem = GodunovImplemHIP(st, mesh.dim, mesh.b, self.impl_version)
        implem.resize_auxiliary_data(mesh)
        Ut = torch.from_numpy(np.ascontiguousarray(U, dtype=np.float64).reshape(-1)).to(device)
        Bt = torch.from_numpy(np.ascontiguousarray(B, dtype=np.float64).reshape(-1)).to(device)
        if Ut.numel() != mesh.num_total * 5 * mesh.cells_per_block or Bt.numel() != mesh.num_total * mesh.faces_per_block:
            raise MhdctError("state arrays do not match the mesh")
        U2 = Ut.clone()
        B2 = Bt.clone()
        sid = C.c_int()
        self._check(lib().mhdct_batch_add(self._h, implem._h, Ut.data_ptr(), U2.data_ptr(), Bt.data_ptr(),
                                          B2.data_ptr(), C.byref(sid)), "batch_add")
        self.slots[sid.value] = _Slot(implem, Ut, U2, Bt, B2, mesh)
        return sid.value

    def remove(self, slot: int):
        self._check(lib().mhdct_batch_remove(self._h, int(slot)), "batch_remove")
        self.slots.pop(int(slot), None)

    def step(self, nsteps: int = 1, stream=None):
        self._check(lib().mhdct_batch_step(self._h, int(nsteps), C.c_void_p(_stream(stream))), "batch_step")

    def step_timed(self, nsteps: int = 1, stream=None) -> dict:
        """Steps with HIP events between the stage launches; returns the average milliseconds
        per step of each stage (synchronises)."""
        n = 10
        ms = (C.c_double * n)()
        self._check(lib().mhdct_batch_step_timed(self._h, int(nsteps), C.c_void_p(_stream(stream)), ms),
                    "batch_step_timed")
        names = [lib().mhdct_stage_name(k).decode() for k in range(n)]
        if self.front:
            # k_front runs the primitive, electric-field and cell stages in the first of them
            names[2:5] = ["k_front", "k_efield (in k_front)", "k_cell (in k_front)"]
        return {name: ms[k] for k, name in enumerate(names)}

    def slot_state(self, slot: int):
        cur, t, n = C.c_int(), C.c_double(), C.c_int64()
        self._check(lib().mhdct_batch_slot_state(self._h, int(slot), C.byref(cur), C.byref(t), C.byref(n)),
                    "batch_slot_state")
        return cur.value, t.value, n.value

    def current(self, slot: int):
        """Device tensors (U, B) holding the slot's current state."""
        cur, _, _ = self.slot_state(slot)
        return self.slots[int(slot)].bufs[cur]

    def state(self, slot: int):
        U, B = self.current(slot)
        m = self.slots[int(slot)].mesh
        nt = m.num_total
        return U.cpu().numpy().reshape(nt, 5, -1), B.cpu().numpy().reshape(nt, -1)

    def check(self):
        bad = C.c_int(-1)
        rc = lib().mhdct_batch_check(self._h, C.byref(bad))
        if rc != 0:
            msg = lib().mhdct_batch_last_error(self._h)
            raise MhdctError(f"slot {bad.value}: {msg.decode() if msg else ''} ({rc})")

    def set_graph(self, enable: bool = True):
        self._check(lib().mhdct_batch_set_graph(self._h, int(bool(enable))), "batch_set_graph")

    def set_overlap(self, enable: bool = True):
        self._check(lib().mhdct_batch_set_overlap(self._h, int(bool(enable))), "batch_set_overlap")

    def set_tune(self, grid_cap: int, tile_cap: int, tile_z: int = 8):
        self._check(lib().mhdct_batch_set_tune(self._h, int(grid_cap), int(tile_cap), int(tile_z)),
                    "batch_set_tune")

    def set_class_launch(self, enable: bool = True):
        """Slots of several block sizes: each stage launched once per block-size class with the
        kernels specialised to it (the default), or once over all slots; same results."""
        self._check(lib().mhdct_batch_set_class_launch(self._h, int(bool(enable))), "batch_set_class_launch")

    def set_front(self, enable: bool = True):
        """k_front (one LDS-tiled workgroup per sub-tile) for the primitive, electric-field and
        cell stages, or the three kernels through HBM scratch (the default); same results bit for
        bit, and on MI355X the three kernels are the faster."""
        self._check(lib().mhdct_batch_set_front(self._h, int(bool(enable))), "batch_set_front")
        self.front = bool(enable)

    def set_stage_threads(self, cell_threads: int, edge_threads: int):
        """Workgroup sizes of k_cell (256 or 512) and k_edge (256 or 512); same results."""
        self._check(lib().mhdct_batch_set_stage_threads(self._h, int(cell_threads), int(edge_threads)),
                    "batch_set_stage_threads")

    def autotune(self, reps: int = 2, stream=None):
        """Pick the launch caps and the flux-tile depth by timing; returns (grid_cap, tile_cap, tile_z)."""
        g, t, z = C.c_int(), C.c_int(), C.c_int()
        self._check(lib().mhdct_batch_autotune(self._h, int(reps), C.c_void_p(_stream(stream)), C.byref(g),
                                               C.byref(t), C.byref(z)), "batch_autotune")
        return g.value, t.value, z.value

    def num_slots(self) -> int:
        return int(lib().mhdct_batch_num_slots(self._h))

    def cells(self) -> int:
        return sum(s.mesh.num_leaves * s.mesh.cells_per_block for s in self.slots.values())
