"""GPU restart: a run saved to a checkpoint and restarted (godunov_mhd_ct_amd/io.py;
init_restart, init/init.cpp:107-193) continues bit for bit as the uninterrupted run does:
same leaves, same U and face B, same time and iteration.  The cases:

* the plain and fused FP64 forms with AMR cycles every two steps (2D blast), saved between
  cycles, so the restarted run adapts its rebuilt mesh on the same schedule;
* RK2_SSP with gravity on the walled 3D Rayleigh-Taylor configuration.
"""
import numpy as np
import pytest

from godunov_mhd_ct_amd import problems
from godunov_mhd_ct_amd.amr import AMRSettings
from godunov_mhd_ct_amd.implem import IMPL_FUSED, IMPL_PLAIN, hydro_settings
from godunov_mhd_ct_amd.io import restart_solver, save_solver
from godunov_mhd_ct_amd.mesh import Mesh

GAMMA = 5.0 / 3.0


def run_and_restart(tmp_path, mk, settings, impl, nsteps, nsave, amr=None):
    import torch
    from godunov_mhd_ct_amd.solver import SolverGodunovMHD

    m, U, B = mk()
    ref = SolverGodunovMHD(m, U, B, settings, impl_version=impl)
    for _ in range(nsteps):
        ref.next_iteration(amr)
    m, U, B = mk()
    sol = SolverGodunovMHD(m, U, B, settings, impl_version=impl)
    for _ in range(nsave):
        sol.next_iteration(amr)
    save_solver(tmp_path / "chk", sol)
    del sol
    rst = restart_solver(tmp_path / "chk", impl_version=impl)
    assert rst.iteration == nsave
    for _ in range(nsteps - nsave):
        rst.next_iteration(amr)
    torch.cuda.synchronize()
    assert rst.t == ref.t and rst.iteration == ref.iteration
    assert rst.amr_cycles == ref.amr_cycles
    lr, cr = ref.mesh.leaves()
    l2, c2 = rst.mesh.leaves()
    assert np.array_equal(lr, l2) and np.array_equal(cr, c2)
    n = ref.mesh.num_leaves
    Ur, Br = ref.state()
    Ug, Bg = rst.state()
    assert np.array_equal(Ug[:n], Ur[:n]), np.abs(Ug[:n] - Ur[:n]).max()
    assert np.array_equal(Bg[:n], Br[:n]), np.abs(Bg[:n] - Br[:n]).max()
    return ref


@pytest.mark.gpu
@pytest.mark.parametrize("impl", [IMPL_PLAIN, IMPL_FUSED])
def test_restart_with_amr_cycles(tmp_path, impl):
    print(f"[restart] impl={impl} 2D blast with AMR cycles", flush=True)

    def mk():
        m = Mesh(2, 8, 2).finalize()
        U, B = problems.blast(m, radius=0.15, center=(0.5, 0.5, 0.0), rho_in=2.0, bfield=(0.5, 0.3, 0.2))
        return m, U, B

    amr = AMRSettings(level_min=1, level_max=3, epsilon_refine=0.05, epsilon_coarsen=0.005, cycle_frequency=2)
    st = hydro_settings(gamma0=GAMMA, cfl=0.8)
    ref = run_and_restart(tmp_path, mk, st, impl, 7, 3, amr)
    assert ref.mesh.num_leaves != 16, "the cycles did not change the mesh"


@pytest.mark.gpu
def test_restart_rk2_gravity(tmp_path):
    print("[restart] RK2_SSP with gravity, 3D Rayleigh-Taylor with walls", flush=True)
    walls = [3] * 6

    def mk():
        m = Mesh(3, 8, 1, nbrick=(1, 1, 2), periodic=(0, 0, 0))
        m.set_bc(walls)
        m.finalize()
        U, B = problems.rayleigh_taylor(m, (0.0, 0.0, -0.1), gamma=1.4, bfield=(0.0, 0.0, 0.5),
                                        xmin=(-0.5, -0.5, -1.0))
        return m, U, B

    st = hydro_settings(gamma0=1.4, cfl=0.7, time_integrator="RK2_SSP", gravity=(0.0, 0.0, -0.1))
    run_and_restart(tmp_path, mk, st, IMPL_FUSED, 5, 2)
