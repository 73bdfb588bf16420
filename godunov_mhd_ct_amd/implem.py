"""GodunovImplemBase counterpart over the C ABI (scheme/GodunovImplemBase.h:45-170).

`GodunovImplemHIP` keeps the reference's public surface -- resize_auxiliary_data(),
total_mem_size_in_bytes(), do_time_step(U, U2, B, B2, dt) -- on HBM-resident FP64 arrays in
the reference layout (torch tensors are only the allocator and stream plumbing).
`create_godunov_implem` mirrors SolverGodunovMHD::create_godunov_implem
(SolverGodunovMHD.h:326-340), keyed on [hydro] implementation_version.
"""
from __future__ import annotations

import ctypes as C

from ._lib import MhdctError, Settings, lib
from .mesh import Mesh

IMPL_PLAIN = 10
IMPL_FUSED = 11
IMPL_FUSED_FP32 = 12

RIEMANN = {"UNKNOWN": 0, "APPROX": 1, "LLF": 2, "HLL": 3, "HLLC": 4, "HLLD": 5}
# [run] time_integrator (SolverGodunovMHD.cpp:494, 551); RK3_SSP is not accepted, as in the
# reference (SolverGodunovMHD.cpp:174-177)
TIME_INTEGRATOR = {"HANCOCK": 0, "RK2_SSP": 1}
# stage coefficients (c0, c1) of y = c0 * x + c1 * y after each forward-Euler stage of the
# two-stage SSP Runge-Kutta scheme (Shu & Osher 1988): u1 = u + dt L(u),
# u_new = 1/2 u + 1/2 (u1 + dt L(u1)) (SolverGodunovMHD.cpp:583-588, 617-622)
RK2_SSP_STAGE1_COEFS = (0.0, 1.0)
RK2_SSP_STAGE2_COEFS = (0.5, 0.5)
PROLONGATION = {"SIMPLE_COPY": 0, "EXTRAPOLATE_LINEAR_MINMOD": 1}


def hydro_settings(gamma0=1.4, cfl=0.5, slope_type=1.0, smallr=1e-10, smallc=1e-10, scaling_factor=1.0,
                   riemann="HLLD", prolongation="SIMPLE_COPY", abort_when_negative_eint=False,
                   time_integrator="HANCOCK", gravity=None, viscosity=None, viscous_predictor=False) -> Settings:
    """Defaults of HydroSettings (models/HydroSettings.cpp:19-30), plus [run] time_integrator,
    [gravity] (gravity=(gx, gy, gz) enables a uniform field; None disables it) and [viscosity]
    (viscosity=mu enables viscous fluxes, FP64 forms only; viscous_predictor adds the viscous
    term of the half-step predictor)."""
    s = Settings()
    s.gamma0 = float(gamma0)
    s.cfl = float(cfl)
    s.slope_type = float(slope_type)
    s.smallr = float(smallr)
    s.smallc = float(smallc)
    s.scaling_factor = float(scaling_factor)
    s.riemann = RIEMANN[riemann] if isinstance(riemann, str) else int(riemann)
    s.prolongation = PROLONGATION[prolongation] if isinstance(prolongation, str) else int(prolongation)
    s.abort_when_negative_eint = int(bool(abort_when_negative_eint))
    if isinstance(time_integrator, str):
        if time_integrator not in TIME_INTEGRATOR:
            raise MhdctError(f"time_integrator {time_integrator!r} is not supported (HANCOCK or RK2_SSP)")
        s.time_integrator = TIME_INTEGRATOR[time_integrator]
    else:
        s.time_integrator = int(time_integrator)
    if gravity is not None:
        g = tuple(float(x) for x in gravity) + (0.0,) * (3 - len(gravity))
        s.gravity_enabled = 1
        for d in range(3):
            s.gravity[d] = g[d]
    if viscosity is not None:
        s.viscosity_enabled = 1
        s.viscosity_mu = float(viscosity)
        s.viscosity_hancock_predictor = int(bool(viscous_predictor))
    return s


def _ptr(t) -> int:
    if t is None:
        return 0
    if hasattr(t, "data_ptr"):
        if not t.is_cuda:
            raise MhdctError("state arrays must be device tensors")
        if not t.is_contiguous():
            raise MhdctError("state arrays must be contiguous")
        return t.data_ptr()
    return int(t)


def _stream(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if hasattr(stream, "cuda_stream"):
        return stream.cuda_stream
    return int(stream)


class GodunovImplemHIP:
    def __init__(self, settings: Settings, dim: int, block_size: int, impl_version: int = IMPL_PLAIN):
        self.settings = settings
        self.dim = int(dim)
        self.b = int(block_size)
        self.impl_version = int(impl_version)
        h = C.c_void_p()
        bs = (C.c_int * 3)(self.b, self.b, self.b)
        rc = lib().mhdct_create(C.byref(settings), self.dim, bs, self.impl_version, C.byref(h))
        if rc == -5:
            raise MhdctError("mhdct_create failed (-5): settings not run by this build (HLLD and SIMPLE_COPY only; "
                             "HANCOCK or RK2_SSP; viscosity in the FP64 forms only, mu >= 0)")
        if rc != 0:
            raise MhdctError(f"mhdct_create failed ({rc})")
        self._h = h
        self.mesh = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().mhdct_destroy(h)
            except Exception:
                pass
            self._h = None

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().mhdct_last_error(self._h)
            raise MhdctError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    # GodunovImplemBase interface ------------------------------------------------------
    def resize_auxiliary_data(self, mesh: Mesh):
        if not mesh.finalized:
            mesh.finalize()
        self._check(lib().mhdct_set_mesh(self._h, mesh.handle), "set_mesh")
        self.mesh = mesh

    def total_mem_size_in_bytes(self) -> int:
        return int(lib().mhdct_mem_bytes(self._h))

    def do_time_step(self, U, U2, B, B2, dt: float, stream=None):
        self._check(lib().mhdct_do_time_step(self._h, _ptr(U), _ptr(U2), _ptr(B), _ptr(B2), float(dt),
                                             C.c_void_p(_stream(stream))), "do_time_step")

    # solver-side kernels next to the boundary -------------------------------------------
    def compute_dt(self, U, B, stream=None) -> float:
        out = C.c_double()
        self._check(lib().mhdct_compute_dt(self._h, _ptr(U), _ptr(B), C.byref(out), C.c_void_p(_stream(stream))),
                    "compute_dt")
        return float(out.value)

    def fill_outside(self, U, B, stream=None):
        self._check(lib().mhdct_fill_outside(self._h, _ptr(U), _ptr(B), C.c_void_p(_stream(stream))),
                    "fill_outside")

    def linear_combination(self, U, U_RK, B, B_RK, coefs, stream=None):
        """U_RK = c0 U + c1 U_RK, B_RK = c0 B + c1 B_RK (LinearCombination of the RK2-SSP driver)."""
        c0, c1 = coefs
        self._check(lib().mhdct_linear_combination(self._h, _ptr(U), _ptr(U_RK), _ptr(B), _ptr(B_RK), float(c0),
                                                   float(c1), C.c_void_p(_stream(stream))), "linear_combination")


def create_godunov_implem(settings: Settings, dim: int, block_size: int, implementation_version: int):
    if implementation_version not in (IMPL_PLAIN, IMPL_FUSED, IMPL_FUSED_FP32):
        raise MhdctError(
            f"Wrong value for input parameter hydro/implementation_version={implementation_version} "
            f"(allowed: {IMPL_PLAIN} plain, {IMPL_FUSED} fused, {IMPL_FUSED_FP32} fused FP32 bulk)")
    return GodunovImplemHIP(settings, dim, block_size, implementation_version)
