"""GPU parity of the ghost-cell prolongation EXTRAPOLATE_LINEAR_MINMOD
(ConvertToPrimitivesVariablesFunctor.cpp:471-635, chosen at :697-710) against the oracle: the
plain and fused FP64 forms on 2D and 3D AMR meshes (ghost cells beside coarser blocks, slope
stencils reaching into finer and same-level blocks), on meshes refined into a corner between
walls (slope stencils of coarse outside blocks), and inside a batch with device dt.  Bar: relative
L-inf <= 1e-12 per field after every step (per-mesh runs) or after the last step (batch)."""
import numpy as np
import pytest

from godunov_mhd_ct_amd import problems
from godunov_mhd_ct_amd.implem import IMPL_FUSED, IMPL_PLAIN, hydro_settings
from godunov_mhd_ct_amd.mesh import Mesh, refine_where
from parity import assert_parity
from oracle_ref import OracleSolver, settings_array
from test_gpu_parity import amr_mesh_2d, amr_mesh_3d, run_pair, walled_refined

LIN = "EXTRAPOLATE_LINEAR_MINMOD"


@pytest.mark.gpu
@pytest.mark.parametrize("impl", [IMPL_PLAIN, IMPL_FUSED])
def test_linear_prolongation_2d_amr(impl):
    m = amr_mesh_2d()
    U, B = problems.blast(m, radius=0.15, center=(0.5, 0.5, 0.0), bfield=(0.4, 0.3, 0.2))
    run_pair(m, U, B, 5.0 / 3.0, 0.8, 1.0, 5, impl, prolongation=LIN)


@pytest.mark.gpu
@pytest.mark.parametrize("impl", [IMPL_PLAIN, IMPL_FUSED])
def test_linear_prolongation_3d_amr(impl):
    m = amr_mesh_3d()
    U, B = problems.random_mhd_3d(m, seed=12, nmodes=6, device="cpu")
    run_pair(m, U, B, 5.0 / 3.0, 0.8, 2.0, 3, impl, prolongation=LIN)


@pytest.mark.gpu
@pytest.mark.parametrize("impl", [IMPL_PLAIN, IMPL_FUSED])
@pytest.mark.parametrize("dim", [2, 3])
def test_linear_prolongation_at_walls(impl, dim):
    bc = [3, 3, 2, 3, 3, 2]
    m = walled_refined(dim, bc)
    U, B = problems.blast(m, radius=0.25, center=(0.2, 0.25, 0.2), rho_in=2.0, bfield=(0.5, 0.3, 0.2))
    run_pair(m, U, B, 5.0 / 3.0, 0.8, 1.0, 4, impl, bc=bc, prolongation=LIN)


@pytest.mark.gpu
def test_linear_prolongation_in_a_batch():
    from godunov_mhd_ct_amd.batch import BatchGodunovMHD

    print("[prolongation] batch of 8^3 and 16^3 AMR meshes vs oracle", flush=True)
    cases = []
    m = amr_mesh_3d()
    U, B = problems.blast(m, radius=0.3, center=(0.3, 0.3, 0.3))
    cases.append((m, U, B))
    m = refine_where(Mesh(3, 16, 1), lambda L, x0, x1, y0, y1, z0, z1: x1 > 0.5 and y0 < 0.5, 2)
    U, B = problems.random_mhd_3d(m, seed=3, nmodes=6, device="cpu")
    cases.append((m, U, B))
    bt = BatchGodunovMHD(3, IMPL_FUSED)
    ids = [bt.add(m, U, B, hydro_settings(gamma0=5.0 / 3.0, cfl=0.8, prolongation=LIN)) for (m, U, B) in cases]
    nsteps = 3
    bt.step(nsteps)
    bt.check()
    for (m, U, B), sid in zip(cases, ids):
        ora = OracleSolver(m, settings_array(gamma0=5.0 / 3.0, cfl=0.8, prolongation=LIN), nthreads=4)
        Uo = U.reshape(m.num_total, 5, -1).copy()
        Bo = B.reshape(m.num_total, -1).copy()
        for _ in range(nsteps):
            Uo, Bo, _ = ora.step(Uo, Bo)
        Ug, Bg = bt.state(sid)
        assert_parity(Ug, Bg, Uo, Bo, m.num_leaves, f"b={m.b}")
