"""GPU parity with uniform gravity (AddGravitySourceTerm.cpp, the predictor's half-step kick,
ComputeDtMHDFunctor::compute_cfl_with_gravity): the plain and fused FP64 forms on the magnetic
Rayleigh-Taylor configurations (app/test_mhd_rayleigh_taylor_{2d,3d}.ini: walls all around, a
1x3 (2D) or 1x1x2 (3D) brick), HANCOCK and RK2_SSP, and a batch holding a slot with gravity next
to one without, against the oracle at the 1e-12 relative L-inf bar with equal dt.
"""
import numpy as np
import pytest

from godunov_mhd_ct_amd import problems
from godunov_mhd_ct_amd.implem import IMPL_FUSED, IMPL_PLAIN, hydro_settings
from godunov_mhd_ct_amd.mesh import Mesh, uniform_mesh
from parity import assert_parity
from oracle_ref import OracleSolver, settings_array
from parity import TOL, rel_linf
from test_gpu_parity import run_pair

WALLS = [3, 3, 3, 3, 3, 3]


def rt_case(dim, b=8):
    nb = (1, 3, 1) if dim == 2 else (1, 1, 2)
    m = Mesh(dim, b, 2 if dim == 2 else 1, nbrick=nb, periodic=(0, 0, 0))
    m.set_bc(WALLS)
    m.finalize()
    g = (0.0, -0.1) if dim == 2 else (0.0, 0.0, -0.1)
    U, B = problems.rayleigh_taylor(m, g, gamma=1.4, bfield=(0.0, 0.5, 0.0) if dim == 2 else (0.0, 0.0, 0.5),
                                    xmin=(-0.5, -1.5, 0.0) if dim == 2 else (-0.5, -0.5, -1.0))
    return m, U, B, g


@pytest.mark.gpu
@pytest.mark.parametrize("impl", [IMPL_PLAIN, IMPL_FUSED])
@pytest.mark.parametrize("dim", [2, 3])
def test_rayleigh_taylor_gravity(impl, dim):
    m, U, B, g = rt_case(dim)
    run_pair(m, U, B, 1.4, 0.7, 1.0, 4, impl, bc=WALLS, gravity=g)


@pytest.mark.gpu
@pytest.mark.parametrize("impl", [IMPL_PLAIN, IMPL_FUSED])
def test_rayleigh_taylor_gravity_rk2(impl):
    m, U, B, g = rt_case(2)
    run_pair(m, U, B, 1.4, 0.7, 1.0, 3, impl, bc=WALLS, gravity=g, time_integrator="RK2_SSP")


@pytest.mark.gpu
def test_batch_with_and_without_gravity():
    import torch
    from godunov_mhd_ct_amd.batch import BatchGodunovMHD

    print("[gravity] batch: a periodic 3D slot with gravity next to one without", flush=True)
    cases = []
    for gz in (-0.2, None):
        m = uniform_mesh(3, 8, 1)
        U, B = problems.rayleigh_taylor(m, (0.0, 0.0, gz or 0.0), gamma=1.4, bfield=(0.1, 0.0, 0.4))
        cases.append((m, U, B, None if gz is None else (0.0, 0.0, gz)))
    bt = BatchGodunovMHD(3, IMPL_FUSED)
    ids = [bt.add(m, U, B, hydro_settings(gamma0=1.4, cfl=0.7, gravity=g)) for (m, U, B, g) in cases]
    nsteps = 3
    bt.step(nsteps)
    bt.check()
    torch.cuda.synchronize()
    for sid, (m, U, B, g) in zip(ids, cases):
        ora = OracleSolver(m, settings_array(gamma0=1.4, cfl=0.7, gravity=g), nthreads=4)
        Uo = U.reshape(m.num_total, 5, -1).copy()
        Bo = B.reshape(m.num_total, -1).copy()
        t = 0.0
        for _ in range(nsteps):
            Uo, Bo, dt = ora.step(Uo, Bo)
            t += dt
        (Ug, Bg) = bt.state(sid)
        _, tg, steps = bt.slot_state(sid)
        assert steps == nsteps and abs(tg - t) <= 1e-14 * t, (tg, t)
        assert_parity(Ug, Bg, Uo, Bo, m.num_leaves, f"slot {sid}")
