"""GPU parity of the viscous terms (ComputeViscousFluxesAndStoreFunctor.cpp:96-352 after the
hydro fluxes, GodunovImplemV0.cpp:179-192; the predictor term,
ComputeUpdatedPrimitiveVariablesFunctor.cpp:178-272) in the plain and the fused FP64 forms.
The kernels are compared against the oracle at the 1e-12 relative L-inf bar with equal dt on:

* 2D and 3D periodic meshes with AMR (coarse/fine viscous flux sums), predictor on and off;
* a walled 2D mesh under RK2_SSP;
* the Kelvin-Helmholtz setup with the reference file's mu = 1e-4.

The FP32-bulk form and the batch engine refuse viscosity with a clear error.
"""
import pytest

from godunov_mhd_ct_amd import MhdctError, problems
from godunov_mhd_ct_amd.implem import IMPL_FUSED, IMPL_FUSED_FP32, IMPL_PLAIN, GodunovImplemHIP, hydro_settings
from godunov_mhd_ct_amd.mesh import Mesh, refine_where, uniform_mesh
from test_gpu_parity import run_pair
from test_oracle_viscosity import random_flow, state


def amr2d():
    return refine_where(Mesh(2, 8, 2), lambda L, x0, x1, y0, y1, z0, z1: x0 < 0.6 and x1 > 0.4 and y0 < 0.6 and y1 > 0.4, 3)


def amr3d():
    return refine_where(Mesh(3, 8, 1), lambda L, x0, x1, y0, y1, z0, z1: x0 < 0.5 and y0 < 0.5 and z0 < 0.5, 2)


FP64_FORMS = [IMPL_PLAIN, IMPL_FUSED]


@pytest.mark.gpu
@pytest.mark.parametrize("impl", FP64_FORMS)
@pytest.mark.parametrize("predictor", [False, True])
@pytest.mark.parametrize("mk", [amr2d, amr3d])
def test_viscous_flow_matches_oracle(mk, predictor, impl):
    m = mk()
    U, B = state(m, random_flow(5))
    run_pair(m, U, B, 1.4, 0.6, 2.0, 4, impl, viscosity=0.02, viscous_predictor=predictor)


@pytest.mark.gpu
@pytest.mark.parametrize("impl", FP64_FORMS)
def test_viscous_rk2_with_walls(impl):
    bc = [3, 3, 3, 3, 0, 0]
    m = Mesh(2, 8, 2, periodic=(0, 0, 1))
    m.set_bc(bc)
    m.finalize()
    U, B = problems.blast(m, radius=0.2, center=(0.4, 0.5, 0.0), bfield=(0.4, 0.2, 0.1))
    run_pair(m, U, B, 5.0 / 3.0, 0.6, 1.0, 3, impl, bc=bc, time_integrator="RK2_SSP", viscosity=0.01)


@pytest.mark.gpu
@pytest.mark.parametrize("impl", FP64_FORMS)
def test_kelvin_helmholtz_reference_viscosity(impl):
    m = uniform_mesh(3, 8, 1)
    U, B = problems.kelvin_helmholtz_3d(m, gamma=1.666)
    run_pair(m, U, B, 1.666, 0.8, 2.0, 3, impl, viscosity=1e-4, viscous_predictor=True)


@pytest.mark.gpu
def test_viscous_16cube_blocks_fused():
    # the 16^3 kernels of the fused form (the benchmark block size) with the predictor term
    m = refine_where(Mesh(3, 16, 1), lambda L, x0, x1, y0, y1, z0, z1: x0 < 0.5 and y0 < 0.5, 2)
    U, B = state(m, random_flow(7))
    run_pair(m, U, B, 1.4, 0.6, 2.0, 2, IMPL_FUSED, viscosity=0.01, viscous_predictor=True)


@pytest.mark.gpu
def test_viscosity_refused_where_not_run():
    from godunov_mhd_ct_amd.batch import BatchGodunovMHD

    with pytest.raises(MhdctError, match="viscosity"):
        GodunovImplemHIP(hydro_settings(viscosity=0.01), 3, 8, IMPL_FUSED_FP32)
    with pytest.raises(MhdctError):
        GodunovImplemHIP(hydro_settings(viscosity=-1.0), 3, 8, IMPL_PLAIN)
    # the batch engine keeps dt on the device; the viscous stage takes it from the host
    m = uniform_mesh(3, 8, 1)
    U, B = state(m, random_flow(3))
    bt = BatchGodunovMHD(3, IMPL_FUSED)
    with pytest.raises(MhdctError, match="viscosity"):
        bt.add(m, U, B, hydro_settings(viscosity=0.01))
