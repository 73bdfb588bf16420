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

# refinement variables by name ([amr] variable_used_to_compute_refine_criterion,
# SolverGodunovMHD.cpp:1117-1167): the cell variables of U (MHD::m_id2names_all, models/MHD.cpp:20-32)
# and the derived quantities (DERIVED_QUANTITY, utils/ComputeDerivedQuantities.h:44-53, any case)
VARIABLES = {"rho": 0, "e_tot": 1, "rho_vx": 2, "rho_vy": 3, "rho_vz": 4}
DERIVED = {"thermal_pressure": 10, "speed_of_sound": 11, "specific_ekin": 12, "magnetic_pressure": 13,
           "local_mach_number": 14, "vx": 15, "vy": 16, "vz": 17}
CRITERIA = {"LOHNER_UNSPLIT": 0, "SIMPLE_GRADIENT": 1}


def variable_code(name: str) -> int:
    """A derived quantity's code (any case), else a variable of U, else rho (the reference's
    fallback, SolverGodunovMHD.cpp:1149-1160)."""
    if name.lower() in DERIVED:
        return DERIVED[name.lower()]
    return VARIABLES.get(name, 0)


@dataclass
class AMRSettings:
    """[amr] keys of the cycle (app/test_mhd_blast_3d.ini:9-69, SolverGodunovMHD.cpp:1107-1113).
    `variable` is one name or a list of up to four (refine where any asks, coarsen where all do);
    `criterion` is LOHNER_UNSPLIT or SIMPLE_GRADIENT."""

    level_min: int
    level_max: int
    epsilon_refine: float = 0.01
    epsilon_coarsen: float = 0.00125
    epsilon_lohner: float = 0.02
    variable: object = "rho"
    cycle_frequency: int = 10
    criterion: str = "LOHNER_UNSPLIT"

    def names(self):
        return [self.variable] if isinstance(self.variable, str) else list(self.variable)

    def params(self) -> AmrParams:
        names = self.names()
        if not 1 <= len(names) <= 4:
            raise MhdctError("one to four refinement variables")
        if self.criterion not in CRITERIA:
            raise MhdctError(f"refine criterion {self.criterion!r} is not supported (one of {sorted(CRITERIA)})")
        p = AmrParams()
        p.level_min = int(self.level_min)
        p.level_max = int(self.level_max)
        p.epsilon_refine = float(self.epsilon_refine)
        p.epsilon_coarsen = float(self.epsilon_coarsen)
        p.epsilon_lohner = float(self.epsilon_lohner)
        codes = [variable_code(n) for n in names]
        p.ivar = codes[0] if codes[0] < 5 else 0
        p.criterion = CRITERIA[self.criterion]
        p.nvars = len(codes)
        for i, c in enumerate(codes):
            p.vars[i] = c
        return p


def refine_flags(implem, U, settings: AMRSettings, stream=None, B=None):
    """(flags int8[num_leaves], indicator float64[num_leaves]) of the implementation's mesh; B (the
    face field) is needed for derived refinement variables."""
    from .implem import _ptr, _stream

    n = implem.mesh.num_leaves
    flags = np.zeros(max(n, 1), dtype=np.int8)
    ind = np.zeros(max(n, 1), dtype=np.float64)
    p = settings.params()
    implem._check(lib().mhdct_refine_flags(implem._h, _ptr(U), _ptr(B) if B is not None else None, C.byref(p),
                                           flags.ctypes.data_as(C.POINTER(C.c_int8)),
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
