"""The AMR cycle (SolverGodunovMHD::do_amr_cycle, SolverGodunovMHD.cpp:263-319) over the C ABI.

1. fill the outside blocks (the indicator reads one ghost layer);
2. refinement flags of every owned block on the device (`mhdct_refine_flags`, in place of
   compute_refine_flags, :1094-1185);
3. adapt and 2:1-balance a copy of the host octree (in place of p4est's refine/coarsen/balance);
4. remap U and face B to the new leaves on the device (`mhdct_remap_*`, in place of
   map_userdata_after_adapt, :1191-1269);
5. the implementation takes the new mesh (resize_auxiliary_data).

The indicator and the remap restate kalypsso::core's ComputeRefineFlags and UserDataRemapper,
which are not in the reference tree; include/mhdct.h and csrc/amr.hip state what they compute.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import AmrParams, MhdctError, lib
from .mesh import Mesh

# conserved variables of U by name ([amr] variable_used_to_compute_refine_criterion)
VARIABLES = {"rho": 0, "e_tot": 1, "rho_vx": 2, "rho_vy": 3, "rho_vz": 4}


@dataclass
class AMRSettings:
    """[amr] keys of the cycle (app/test_mhd_blast_3d.ini:9-69, SolverGodunovMHD.cpp:1107-1113)."""

    level_min: int
    level_max: int
    epsilon_refine: float = 0.01
    epsilon_coarsen: float = 0.00125
    epsilon_lohner: float = 0.02
    variable: str = "rho"
    cycle_frequency: int = 10

    def params(self) -> AmrParams:
        if self.variable not in VARIABLES:
            raise MhdctError(f"refine criterion variable {self.variable!r} is not supported "
                             f"(one of {sorted(VARIABLES)})")
        p = AmrParams()
        p.level_min = int(self.level_min)
        p.level_max = int(self.level_max)
        p.epsilon_refine = float(self.epsilon_refine)
        p.epsilon_coarsen = float(self.epsilon_coarsen)
        p.epsilon_lohner = float(self.epsilon_lohner)
        p.ivar = VARIABLES[self.variable]
        return p


def refine_flags(implem, U, settings: AMRSettings, stream=None):
    """(flags int8[num_leaves], indicator float64[num_leaves]) of the implementation's mesh."""
    from .implem import _ptr, _stream

    n = implem.mesh.num_leaves
    flags = np.zeros(max(n, 1), dtype=np.int8)
    ind = np.zeros(max(n, 1), dtype=np.float64)
    p = settings.params()
    implem._check(lib().mhdct_refine_flags(implem._h, _ptr(U), C.byref(p), flags.ctypes.data_as(C.POINTER(C.c_int8)),
                                           ind.ctypes.data_as(C.POINTER(C.c_double)), C.c_void_p(_stream(stream))),
                  "refine_flags")
    return flags[:n], ind[:n]


def adapted(mesh: Mesh, flags) -> Mesh:
    """A new, finalized mesh: a copy of `mesh` refined/coarsened by `flags`, then 2:1 balanced."""
    new = mesh.clone()
    new.adapt(flags)
    new.balance()
    return new.finalize()


class Remap:
    """Remap of the owned leaves' U and face B from `old_mesh` to `new_mesh`."""

    def __init__(self, old_mesh: Mesh, new_mesh: Mesh):
        h = C.c_void_p()
        rc = lib().mhdct_remap_create(old_mesh.handle, new_mesh.handle, C.byref(h))
        if rc != 0:
            msg = lib().mhdct_remap_last_error(None)
            raise MhdctError(f"remap_create failed ({rc}): {msg.decode() if msg else ''}")
        self._h = h
        self.old_mesh = old_mesh
        self.new_mesh = new_mesh

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().mhdct_remap_destroy(h)
            except Exception:
                pass
            self._h = None

    def counts(self):
        """(copied, prolonged, restricted) new leaves."""
        c = (C.c_int * 3)()
        lib().mhdct_remap_counts(self._h, c)
        return tuple(int(x) for x in c)

    def apply(self, U_old, B_old, U_new, B_new, stream=None):
        from .implem import _ptr, _stream

        rc = lib().mhdct_remap_apply(self._h, _ptr(U_old), _ptr(B_old), _ptr(U_new), _ptr(B_new),
                                     C.c_void_p(_stream(stream)))
        if rc != 0:
            msg = lib().mhdct_remap_last_error(self._h)
            raise MhdctError(f"remap_apply failed ({rc}): {msg.decode() if msg else ''}")
