"""Test-side binding of the CPU oracle (oracle/mhd_oracle.cpp, built into oracle/_build).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
# MHDCT_ORACLE_LIB names another build of the oracle's sources (e.g. one with a slip planted on
# purpose, to show that the pin tests fail on it); by default the one oracle/Makefile builds
ORACLE_LIB = Path(os.environ.get("MHDCT_ORACLE_LIB", ORACLE_DIR / "_build" / "libmhd_oracle.so"))

# the -O3 build of the same source, timed as the CPU baseline (not used for parity)
ORACLE_LIB_O3 = ORACLE_DIR / "_build" / "libmhd_oracle_o3.so"

_lib = None


def use_library(path):
    """Load the oracle from `path` from now on (the -O3 timing build for the CPU baseline)."""
    global ORACLE_LIB, _lib
    ORACLE_LIB = Path(path)
    _lib = None


def build_oracle():
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not ORACLE_LIB.exists():
            build_oracle()
        L = C.CDLL(str(ORACLE_LIB))
        P = C.POINTER
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.c_int, C.c_int, P(C.c_double), C.c_int, C.c_int, P(C.c_int), P(C.c_int),
                                    P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int), C.c_int]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.oracle_do_time_step.restype = C.c_int
        L.oracle_do_time_step.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_double),
                                          C.c_double]
        L.oracle_compute_inv_dt.restype = C.c_double
        L.oracle_compute_inv_dt.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double)]
        L.oracle_fill_outside.restype = C.c_int
        L.oracle_fill_outside.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double)]
        L.oracle_linear_combination.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), P(C.c_double),
                                                P(C.c_double), C.c_double, C.c_double]
        L.oracle_refine_indicator.restype = C.c_int
        L.oracle_refine_indicator.argtypes = [C.c_void_p, P(C.c_double), C.c_int, C.c_double, P(C.c_double),
                                              P(C.c_double)]
        L.oracle_refine_indicator_vars.restype = C.c_int
        L.oracle_refine_indicator_vars.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), C.c_int, P(C.c_int),
                                                   C.c_int, C.c_double, P(C.c_double), P(C.c_double)]
        L.oracle_refine_flags.argtypes = [C.c_int, P(C.c_int), P(C.c_double), C.c_int, C.c_int, C.c_double,
                                          C.c_double, P(C.c_int8)]
        L.oracle_remap.restype = C.c_int
        L.oracle_remap.argtypes = [C.c_int, C.c_int, P(C.c_int), P(C.c_int), C.c_int, P(C.c_int), P(C.c_int),
                                   C.c_int, P(C.c_int), P(C.c_int), P(C.c_double), P(C.c_double), P(C.c_double),
                                   P(C.c_double)]
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_last_error.argtypes = [C.c_void_p]
        L.oracle_riemann_hlld.argtypes = [P(C.c_double)] * 4
        L.oracle_compute_emf.restype = C.c_double
        L.oracle_compute_emf.argtypes = [P(C.c_double), P(C.c_double), C.c_int]
        L.oracle_compute_primitives.argtypes = [P(C.c_double)] * 3
        L.oracle_slope.restype = C.c_double
        L.oracle_slope.argtypes = [C.c_double] * 4
        L.oracle_speed_info.argtypes = [P(C.c_double), P(C.c_double), C.c_int, P(C.c_double)]
        L.oracle_scratch.restype = C.c_long
        L.oracle_scratch.argtypes = [C.c_void_p, C.c_int, P(C.c_double)]
        _lib = L
    return _lib


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a.ctypes.data_as(C.POINTER(C.c_int)), a


TIME_INTEGRATOR = {"HANCOCK": 0, "RK2_SSP": 1}
# SSP-RK2 stage coefficients (c0, c1) of y = c0 x + c1 y (Shu & Osher 1988; call sites
# SolverGodunovMHD.cpp:583-588, 617-622)
RK2_STAGE1 = (0.0, 1.0)
RK2_STAGE2 = (0.5, 0.5)


PROLONGATION = {"SIMPLE_COPY": 0, "EXTRAPOLATE_LINEAR_MINMOD": 1}


def settings_array(gamma0=1.4, cfl=0.5, slope_type=1.0, smallr=1e-10, smallc=1e-10, scaling=1.0,
                   abort_eint=0, time_integrator="HANCOCK", gravity=None, mu=None, viscous_predictor=False,
                   prolongation="SIMPLE_COPY"):
    """The oracle's settings vector; mu (not None) enables viscosity."""
    pr = PROLONGATION[prolongation] if isinstance(prolongation, str) else int(prolongation)
    ti = TIME_INTEGRATOR[time_integrator] if isinstance(time_integrator, str) else int(time_integrator)
    g = (0.0, 0.0, 0.0) if gravity is None else tuple(float(x) for x in gravity) + (0.0,) * (3 - len(gravity))
    return np.array([gamma0, cfl, slope_type, smallr, smallc, scaling, abort_eint, ti,
                     0.0 if gravity is None else 1.0, *g, 0.0 if mu is None else 1.0,
                     0.0 if mu is None else float(mu), 1.0 if viscous_predictor else 0.0, float(pr)],
                    dtype=np.float64)


class OracleSolver:
    """The oracle's GodunovImplemV0 restatement on a given mesh (leaves from the product's
    host octree; the oracle does its own neighbour search on them)."""

    def __init__(self, mesh, settings=None, nthreads=1, bc=None):
        self.dim = mesh.dim
        self.b = mesh.b
        self.settings = settings if settings is not None else settings_array()
        lv, co = mesh.leaves()
        src, dirs = mesh.outside()
        self.nowned = mesh.num_leaves
        self.noutside = mesh.num_outside
        keep = []
        lvp, a1 = iptr(lv)
        cop, a2 = iptr(co.reshape(-1))
        nbp, a3 = iptr(mesh.nbrick)
        pep, a4 = iptr(mesh.periodic)
        bcv = list(bc) if bc is not None else list(mesh.bc)
        bcp, a5 = iptr(bcv)
        srcp, a6 = iptr(src if len(src) else np.zeros(1))
        dirp, a7 = iptr(dirs.reshape(-1) if len(dirs) else np.zeros(3))
        keep += [a1, a2, a3, a4, a5, a6, a7]
        self._h = lib().oracle_create(self.dim, self.b, dptr(self.settings), self.nowned, self.noutside, lvp, cop,
                                      nbp, pep, bcp, srcp, dirp, int(nthreads))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_destroy(self._h)
            self._h = None

    def set_threads(self, n):
        lib().oracle_set_threads(self._h, int(n))

    def do_time_step(self, U, U2, B, B2, dt):
        rc = lib().oracle_do_time_step(self._h, dptr(U), dptr(U2), dptr(B), dptr(B2), float(dt))
        if rc != 0:
            raise RuntimeError(lib().oracle_last_error(self._h).decode())

    def inv_dt(self, U, B):
        return lib().oracle_compute_inv_dt(self._h, dptr(U), dptr(B))

    def scratch(self, which):
        """A scratch array of the last step: 0 Bg, 1 Q, 2 Q2, 3-5 slopes, 6 E, 7 sF."""
        n = lib().oracle_scratch(self._h, int(which), None)
        out = np.empty(n)
        lib().oracle_scratch(self._h, int(which), dptr(out))
        return out

    def compute_dt(self, U, B):
        return self.settings[1] / self.inv_dt(U, B)

    def fill_outside(self, U, B):
        lib().oracle_fill_outside(self._h, dptr(U), dptr(B))

    def linear_combination(self, U, U_RK, B, B_RK, coefs):
        lib().oracle_linear_combination(self._h, dptr(U), dptr(U_RK), dptr(B), dptr(B_RK), float(coefs[0]),
                                        float(coefs[1]))

    def refine_indicator(self, U, ivar=0, eps=0.02, per_cell=False):
        """Per owned block the maximum of Loehner's estimate (and the per-cell values)."""
        emax = np.zeros(max(self.nowned, 1))
        ecell = np.zeros((max(self.nowned, 1), self.b ** self.dim)) if per_cell else None
        rc = lib().oracle_refine_indicator(self._h, dptr(U), int(ivar), float(eps), dptr(emax),
                                           dptr(ecell) if per_cell else None)
        if rc != 0:
            raise RuntimeError(lib().oracle_last_error(self._h).decode())
        return (emax[:self.nowned], ecell[:self.nowned]) if per_cell else emax[:self.nowned]

    def refine_indicator_vars(self, U, B, variables, criterion=0, eps=0.02, per_cell=False):
        """Per owned block the maximum over its cells and the refinement variables (codes 0-4 of
        U, 10-17 derived) of the criterion (0 LOHNER_UNSPLIT, 1 SIMPLE_GRADIENT)."""
        emax = np.zeros(max(self.nowned, 1))
        ecell = np.zeros((max(self.nowned, 1), self.b ** self.dim)) if per_cell else None
        vp, keep = iptr(list(variables))
        rc = lib().oracle_refine_indicator_vars(self._h, dptr(U), dptr(B), len(variables), vp, int(criterion),
                                                float(eps), dptr(emax), dptr(ecell) if per_cell else None)
        if rc != 0:
            raise RuntimeError(lib().oracle_last_error(self._h).decode())
        return (emax[:self.nowned], ecell[:self.nowned]) if per_cell else emax[:self.nowned]

    @property
    def rk2(self):
        return int(self.settings[7]) == TIME_INTEGRATOR["RK2_SSP"]

    def step(self, U, B, dt=None):
        """One solver step (SolverGodunovMHD::next_iteration_impl): dt, outside fill, then
        HANCOCK (godunov_unsplit, SolverGodunovMHD.cpp:494-549): U2 = U, B2 = B, do_time_step;
        or RK2_SSP (:551-647): U_RK = U, do_time_step(U -> U_RK), stage-1 combination, outside
        fill of U_RK, U2 = U_RK, do_time_step(U_RK -> U2), stage-2 combination into U2.
        Returns (U2, B2, dt); U and B are left as the outside fill made them."""
        if self.noutside:
            self.fill_outside(U, B)
        if dt is None:
            dt = self.compute_dt(U, B)
        if not self.rk2:
            U2 = U.copy()
            B2 = B.copy()
            self.do_time_step(U, U2, B, B2, dt)
            return U2, B2, dt
        U_RK = U.copy()
        B_RK = B.copy()
        self.do_time_step(U, U_RK, B, B_RK, dt)
        self.linear_combination(U, U_RK, B, B_RK, RK2_STAGE1)
        if self.noutside:
            self.fill_outside(U_RK, B_RK)
        U2 = U_RK.copy()
        B2 = B_RK.copy()
        self.do_time_step(U_RK, U2, B_RK, B2, dt)
        self.linear_combination(U, U2, B, B2, RK2_STAGE2)
        return U2, B2, dt


def refine_flags(levels, emax, level_min, level_max, eps_refine, eps_coarsen):
    n = len(emax)
    lv = np.ascontiguousarray(levels[:n], dtype=np.int32)
    e = np.ascontiguousarray(emax, dtype=np.float64)
    out = np.zeros(max(n, 1), dtype=np.int8)
    lib().oracle_refine_flags(n, lv.ctypes.data_as(C.POINTER(C.c_int)), dptr(e), int(level_min), int(level_max),
                              float(eps_refine), float(eps_coarsen), out.ctypes.data_as(C.POINTER(C.c_int8)))
    return out[:n]


def remap(old_mesh, new_mesh, U_old, B_old):
    """The oracle's remap of the owned leaves: (U_new, B_new) in the new mesh's layout (outside
    blocks zero)."""
    lo, co = old_mesh.leaves()
    ln, cn = new_mesh.leaves()
    no, nn = old_mesh.num_leaves, new_mesh.num_leaves
    ncell = new_mesh.cells_per_block
    fper = new_mesh.faces_per_block
    Uo = np.ascontiguousarray(U_old, dtype=np.float64).reshape(-1)
    Bo = np.ascontiguousarray(B_old, dtype=np.float64).reshape(-1)
    Un = np.zeros(new_mesh.num_total * 5 * ncell)
    Bn = np.zeros(new_mesh.num_total * fper)
    lop, a1 = iptr(lo[:no])
    cop, a2 = iptr(co[:no].reshape(-1))
    lnp, a3 = iptr(ln[:nn])
    cnp, a4 = iptr(cn[:nn].reshape(-1))
    nbp, a5 = iptr(new_mesh.nbrick)
    pep, a6 = iptr(new_mesh.periodic)
    rc = lib().oracle_remap(new_mesh.dim, new_mesh.b, nbp, pep, no, lop, cop, nn, lnp, cnp, dptr(Uo), dptr(Bo),
                            dptr(Un), dptr(Bn))
    if rc != 0:
        raise RuntimeError("oracle remap: a new leaf has no source one level away")
    return Un.reshape(new_mesh.num_total, 5, ncell), Bn.reshape(new_mesh.num_total, fper)
