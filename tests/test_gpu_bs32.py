"""GPU parity of the 32-cell block kernels against the CPU oracle.

Every kernel of the fused form is compiled once per block size (8, 16, 32), and a batch of
mixed block sizes runs the size-dispatching kernels; 32^3 blocks hold most of the cells of the
mixed-block benchmark mesh.  These tests run 32-cell blocks on two-level meshes:

* alone, through the per-mesh solver, in the plain and fused FP64 forms, 2D and 3D;
* inside batches of 8/16/32 blocks (3D) and 8/32 blocks (2D), per slot, with device dt.

Bar (tests/parity.py): relative L-inf <= 1e-12 on rho, E, momentum and face B after every step
of the per-mesh runs and after the last step of the batch runs.
"""
import numpy as np
import pytest

from godunov_mhd_ct_amd import problems
from godunov_mhd_ct_amd.implem import IMPL_FUSED, IMPL_PLAIN, hydro_settings
from godunov_mhd_ct_amd.mesh import Mesh, refine_where, uniform_mesh
from oracle_ref import OracleSolver, settings_array
from parity import assert_parity
from test_gpu_parity import run_pair

TOL = 1e-12


def rel_linf(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def mesh32_3d():
    # 8 leaves at level 1, the low corner refined: 15 leaves of 32^3 cells on two levels
    m = Mesh(3, 32, 1)
    return refine_where(m, lambda L, x0, x1, y0, y1, z0, z1: x0 < 0.5 and y0 < 0.5 and z0 < 0.5, 2)


def mesh32_2d():
    m = Mesh(2, 32, 1)
    return refine_where(m, lambda L, x0, x1, y0, y1, z0, z1: x0 < 0.6 and x1 > 0.4 and y0 < 0.6 and y1 > 0.4, 3)


@pytest.mark.gpu
@pytest.mark.parametrize("impl", [IMPL_PLAIN, IMPL_FUSED])
def test_blocks32_3d_two_levels(impl):
    m = mesh32_3d()
    assert set(m.leaves()[0].tolist()) == {1, 2}
    U, B = problems.random_mhd_3d(m, seed=9, nmodes=6, device="cpu")
    run_pair(m, U, B, 5.0 / 3.0, 0.8, 1.0, 3, impl)


@pytest.mark.gpu
@pytest.mark.parametrize("impl", [IMPL_PLAIN, IMPL_FUSED])
def test_blocks32_3d_blast_mc(impl):
    m = mesh32_3d()
    U, B = problems.blast(m, radius=0.2, center=(0.3, 0.3, 0.3), bfield=(0.4, 0.3, 0.5))
    run_pair(m, U, B, 5.0 / 3.0, 0.8, 2.0, 3, impl)


@pytest.mark.gpu
@pytest.mark.parametrize("impl", [IMPL_PLAIN, IMPL_FUSED])
def test_blocks32_2d_two_levels(impl):
    m = mesh32_2d()
    assert len(set(m.leaves()[0].tolist())) >= 2
    U, B = problems.blast(m, radius=0.15, center=(0.5, 0.5, 0.0), bfield=(0.6, 0.3, 0.2))
    run_pair(m, U, B, 5.0 / 3.0, 0.8, 1.0, 5, impl)


def _oracle_run(m, U, B, slope, nsteps):
    ora = OracleSolver(m, settings_array(gamma0=5.0 / 3.0, cfl=0.8, slope_type=slope), nthreads=4)
    Uo = U.reshape(m.num_total, 5, -1).copy()
    Bo = B.reshape(m.num_total, -1).copy()
    t = 0.0
    for _ in range(nsteps):
        Uo, Bo, dt = ora.step(Uo, Bo)
        t += dt
    return Uo, Bo, t


def _batch_vs_oracle(dim, cases, nsteps):
    from godunov_mhd_ct_amd.batch import BatchGodunovMHD

    bt = BatchGodunovMHD(dim, IMPL_FUSED)
    ids = [bt.add(m, U, B, hydro_settings(gamma0=5.0 / 3.0, cfl=0.8, slope_type=slope))
           for (m, U, B, slope) in cases]
    bt.step(nsteps)
    bt.check()
    for n, ((m, U, B, slope), sid) in enumerate(zip(cases, ids)):
        Uo, Bo, t = _oracle_run(m, U, B, slope, nsteps)
        Ug, Bg = bt.state(sid)
        _, tg, steps = bt.slot_state(sid)
        assert steps == nsteps and abs(tg - t) <= 1e-14 * t, f"slot {n} (b={m.b}): time {tg!r} vs {t!r}"
        assert_parity(Ug, Bg, Uo, Bo, m.num_leaves, f"slot {n} (b={m.b})")


@pytest.mark.gpu
def test_batch_3d_blocks_8_16_32_vs_oracle():
    """The size-dispatching kernels of a mixed batch, each slot against the oracle."""
    print("[bs32] 3D batch of 8^3, 16^3 and 32^3 AMR meshes vs oracle", flush=True)
    cases = []
    m = Mesh(3, 8, 1)
    m = refine_where(m, lambda L, x0, x1, y0, y1, z0, z1: x0 < 0.5 and y0 < 0.5 and z0 < 0.5, 2)
    U, B = problems.random_mhd_3d(m, seed=5, nmodes=6, device="cpu")
    cases.append((m, U, B, 1.0))
    m = Mesh(3, 16, 1)
    m = refine_where(m, lambda L, x0, x1, y0, y1, z0, z1: x1 > 0.5 and y0 < 0.5, 2)
    U, B = problems.kelvin_helmholtz_3d(m, seed=2)
    cases.append((m, U, B, 2.0))
    m = mesh32_3d()
    U, B = problems.random_mhd_3d(m, seed=6, nmodes=6, device="cpu")
    cases.append((m, U, B, 1.0))
    _batch_vs_oracle(3, cases, 3)


@pytest.mark.gpu
def test_batch_2d_blocks_8_32_vs_oracle():
    print("[bs32] 2D batch of 8^2 and 32^2 meshes vs oracle", flush=True)
    cases = []
    m = uniform_mesh(2, 8, 2)
    U, B = problems.orszag_tang_2d(m)
    cases.append((m, U, B, 1.0))
    m = mesh32_2d()
    U, B = problems.orszag_tang_2d(m)
    cases.append((m, U, B, 2.0))
    _batch_vs_oracle(2, cases, 5)


@pytest.mark.gpu
def test_batch_3d_only_32_blocks_vs_oracle():
    """A batch of one block size runs the kernels specialised to 32^3."""
    print("[bs32] 3D batch of one 32^3 mesh vs oracle", flush=True)
    m = mesh32_3d()
    U, B = problems.blast(m, radius=0.25, center=(0.35, 0.4, 0.3), bfield=(0.3, 0.5, 0.4))
    _batch_vs_oracle(3, [(m, U, B, 1.0)], 3)
