"""k_front (the primitive, electric-field and cell stages as one LDS-tiled workgroup per sub-tile
of 8 x 8 x 4 cells, 8 x 8 x 2 in 8^3 blocks, 8 x 8 in 2D) against the three kernels that write
and read the same values through HBM scratch.

Both paths run the same arithmetic on the same inputs, so each batch below is stepped twice, once
per path, and the states, the simulated times and the step counts must be equal bit for bit:

* 3D batches of mixed block sizes (8, 16, 32: launched per block-size class, and once with the
  size-dispatching kernels) and of one block size, on two-level meshes;
* 2D batches of mixed sizes, and bounded meshes with outside blocks (WALL, ZERO_GRADIENT);
* the FP32-bulk form, the RK2_SSP integrator, EXTRAPOLATE_LINEAR_MINMOD prolongation, gravity.

Both paths also meet the oracle at the parity bar: the rest of the GPU suite checks the
three-kernel path (the default) against the oracle, and the first test below checks the k_front
path against it directly.
"""
import numpy as np
import pytest

from godunov_mhd_ct_amd import problems
from godunov_mhd_ct_amd.implem import IMPL_FUSED, IMPL_FUSED_FP32, hydro_settings
from godunov_mhd_ct_amd.mesh import Mesh, refine_where, uniform_mesh
from oracle_ref import OracleSolver, settings_array
from parity import assert_parity


def amr3d(b, seed, corner=True):
    m = Mesh(3, b, 1)
    if corner:
        m = refine_where(m, lambda L, x0, x1, y0, y1, z0, z1: x0 < 0.5 and y0 < 0.5 and z0 < 0.5, 2)
    else:
        m = refine_where(m, lambda L, x0, x1, y0, y1, z0, z1: x1 > 0.5 and y0 < 0.5, 2)
    U, B = problems.random_mhd_3d(m, seed=seed, nmodes=6, device="cpu")
    return m, U, B


def amr2d(b):
    m = Mesh(2, b, 1)
    m = refine_where(m, lambda L, x0, x1, y0, y1, z0, z1: x0 < 0.6 and x1 > 0.4 and y0 < 0.6 and y1 > 0.4, 3)
    U, B = problems.blast(m, radius=0.15, center=(0.5, 0.5, 0.0), bfield=(0.6, 0.3, 0.2))
    return m, U, B


def run(dim, cases, nsteps, front, impl=IMPL_FUSED, class_launch=True, **kw):
    from godunov_mhd_ct_amd.batch import BatchGodunovMHD

    bt = BatchGodunovMHD(dim, impl)
    bt.set_front(front)
    bt.set_class_launch(class_launch)
    ids = [bt.add(m, U, B, hydro_settings(gamma0=5.0 / 3.0, cfl=0.8, slope_type=slope, **kw))
           for (m, U, B, slope) in cases]
    bt.step(nsteps)
    bt.check()
    return [(bt.state(s), bt.slot_state(s)[1:]) for s in ids]


def assert_same(dim, cases, nsteps, what, impl=IMPL_FUSED, **kw):
    print(f"[front] {what}: {len(cases)} slots, {nsteps} steps", flush=True)
    a = run(dim, cases, nsteps, True, impl, **kw)
    b = run(dim, cases, nsteps, False, impl, **kw)
    for n, (((Ua, Ba), ca), ((Ub, Bb), cb)) in enumerate(zip(a, b)):
        assert ca == cb, f"{what} slot {n}: clock {ca} vs {cb}"
        assert np.array_equal(Ua, Ub), f"{what} slot {n}: U differs, max {np.abs(Ua - Ub).max():.3e}"
        assert np.array_equal(Ba, Bb), f"{what} slot {n}: B differs, max {np.abs(Ba - Bb).max():.3e}"
    return a


@pytest.mark.gpu
def test_front_3d_mixed_block_sizes():
    """Also each slot of the k_front run against the oracle, relative L-inf <= 1e-12."""
    cases = [(*amr3d(8, 5), 1.0), (*amr3d(16, 2, corner=False), 2.0), (*amr3d(32, 6), 1.0)]
    nsteps = 3
    res = assert_same(3, cases, nsteps, "3D 8/16/32")
    # the kernels that dispatch on the block size (one launch over all slots)
    one = run(3, cases, nsteps, True, class_launch=False)
    for n, (((Ua, Ba), _), ((Ub, Bb), _)) in enumerate(zip(one, res)):
        assert np.array_equal(Ua, Ub) and np.array_equal(Ba, Bb), f"slot {n}: one launch vs per class"
    for n, ((m, U, B, slope), ((Ug, Bg), _)) in enumerate(zip(cases, res)):
        ora = OracleSolver(m, settings_array(gamma0=5.0 / 3.0, cfl=0.8, slope_type=slope), nthreads=4)
        Uo = U.reshape(m.num_total, 5, -1).copy()
        Bo = B.reshape(m.num_total, -1).copy()
        for _ in range(nsteps):
            Uo, Bo, _ = ora.step(Uo, Bo)
        assert_parity(Ug, Bg, Uo, Bo, m.num_leaves, f"k_front slot {n} (b={m.b})")


@pytest.mark.gpu
@pytest.mark.parametrize("b", [8, 16, 32])
def test_front_3d_one_block_size(b):
    assert_same(3, [(*amr3d(b, 7 + b), 2.0)], 3, f"3D {b}^3")


@pytest.mark.gpu
def test_front_2d_mixed_and_bounded():
    cases = [(*amr2d(8), 1.0), (*amr2d(32), 2.0)]
    for n, bc in enumerate([[2, 2, 3, 3, 0, 0], [3, 3, 2, 2, 0, 0]]):
        m = Mesh(2, 16, 2, periodic=(0, 0, 1))
        m.set_bc(bc)
        m.finalize()
        U, B = problems.blast(m, radius=0.2, center=(0.3 + 0.2 * n, 0.6, 0.0), bfield=(0.5, 0.2, 0.1))
        cases.append((m, U, B, 1.0))
    assert_same(2, cases, 5, "2D mixed and bounded")


@pytest.mark.gpu
def test_front_3d_bounded():
    m = Mesh(3, 16, 1, periodic=(0, 0, 0))
    m.set_bc([3, 3, 2, 2, 3, 2])
    m.finalize()
    U, B = problems.blast(m, radius=0.25, center=(0.3, 0.4, 0.6), bfield=(0.3, 0.5, 0.4))
    assert_same(3, [(m, U, B, 1.0)], 3, "3D walls")


@pytest.mark.gpu
def test_front_fp32_bulk():
    cases = [(*amr3d(8, 11), 1.0), (*amr3d(16, 12, corner=False), 1.0)]
    assert_same(3, cases, 3, "FP32 3D 8/16", impl=IMPL_FUSED_FP32)
    assert_same(2, [(*amr2d(16), 1.0)], 3, "FP32 2D 16", impl=IMPL_FUSED_FP32)


@pytest.mark.gpu
def test_front_rk2_prolongation_gravity():
    cases = [(*amr3d(8, 13), 1.0), (*amr3d(16, 14), 2.0)]
    assert_same(3, cases, 2, "RK2_SSP", time_integrator="RK2_SSP")
    assert_same(3, cases, 3, "EXTRAPOLATE_LINEAR_MINMOD", prolongation="EXTRAPOLATE_LINEAR_MINMOD")
    assert_same(2, [(*amr2d(8), 1.0)], 3, "2D linear prolongation", prolongation="EXTRAPOLATE_LINEAR_MINMOD")
    assert_same(3, cases, 3, "gravity", gravity=(0.1, -0.2, 0.3))
