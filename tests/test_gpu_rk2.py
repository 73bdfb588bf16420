"""GPU parity of the RK2_SSP integrator (SolverGodunovMHD.cpp:551-647): the plain and fused FP64
forms driven through two forward-Euler stages and the stage combinations, against the oracle's
RK2_SSP step at the 1e-12 relative L-inf bar; the stage combination entry point against numpy
bit for bit; the FP32-bulk form against the FP64 form; RK2_SSP batches (device dt, three
slots including a walled one) against the oracle per slot, alone and graph-replayed equal to
the mix, and one integrator per batch.
"""
import numpy as np
import pytest

from godunov_mhd_ct_amd import problems
from godunov_mhd_ct_amd.implem import IMPL_FUSED, IMPL_FUSED_FP32, IMPL_PLAIN, GodunovImplemHIP, hydro_settings
from godunov_mhd_ct_amd.mesh import Mesh, refine_where, uniform_mesh
from parity import assert_parity
from oracle_ref import OracleSolver, settings_array
from parity import TOL, rel_linf
from test_gpu_parity import amr_mesh_2d, amr_mesh_3d, run_pair

IMPLS = [IMPL_PLAIN, IMPL_FUSED]
RK2 = "RK2_SSP"


@pytest.mark.gpu
@pytest.mark.parametrize("impl", IMPLS)
def test_rk2_orszag_tang_2d(impl):
    m = uniform_mesh(2, 8, 2)
    U, B = problems.orszag_tang_2d(m)
    run_pair(m, U, B, 5.0 / 3.0, 0.8, 1.0, 5, impl, time_integrator=RK2)


@pytest.mark.gpu
@pytest.mark.parametrize("impl", IMPLS)
def test_rk2_kelvin_helmholtz_3d_mc(impl):
    m = uniform_mesh(3, 8, 1)
    U, B = problems.kelvin_helmholtz_3d(m, seed=3)
    run_pair(m, U, B, 1.4, 0.8, 2.0, 3, impl, time_integrator=RK2)


@pytest.mark.gpu
@pytest.mark.parametrize("impl", IMPLS)
def test_rk2_blast_amr(impl):
    m = amr_mesh_2d()
    U, B = problems.blast(m, radius=0.15, center=(0.5, 0.5, 0.0))
    run_pair(m, U, B, 5.0 / 3.0, 0.8, 1.0, 3, impl, time_integrator=RK2)
    m = amr_mesh_3d()
    U, B = problems.blast(m, radius=0.3, center=(0.3, 0.3, 0.3))
    run_pair(m, U, B, 5.0 / 3.0, 0.8, 1.0, 2, impl, time_integrator=RK2)


@pytest.mark.gpu
@pytest.mark.parametrize("impl", IMPLS)
def test_rk2_blast_2d_bounded(impl):
    bc = [2, 2, 3, 3, 0, 0]  # ZERO_GRADIENT in x, WALL in y: the outside fill runs before each stage
    m = Mesh(2, 8, 2, periodic=(0, 0, 1))
    m.set_bc(bc)
    m.finalize()
    U, B = problems.blast(m, radius=0.2, center=(0.3, 0.6, 0.0), bfield=(0.5, 0.2, 0.1))
    run_pair(m, U, B, 5.0 / 3.0, 0.8, 1.0, 4, impl, bc=bc, time_integrator=RK2)


@pytest.mark.gpu
def test_linear_combination_entry_point():
    import torch

    print("[rk2] mhdct_linear_combination 3D b=8 level 1", flush=True)
    m = uniform_mesh(3, 8, 1)
    im = GodunovImplemHIP(hydro_settings(time_integrator=RK2), 3, 8, IMPL_PLAIN)
    im.resize_auxiliary_data(m)
    rng = np.random.default_rng(11)
    n = m.num_total
    x = rng.uniform(-3.0, 3.0, (n, 5, m.cells_per_block))
    y = rng.uniform(-3.0, 3.0, (n, 5, m.cells_per_block))
    bx = rng.uniform(-1.0, 1.0, (n, m.faces_per_block))
    by = rng.uniform(-1.0, 1.0, (n, m.faces_per_block))
    dev = lambda a: torch.from_numpy(a.reshape(-1).copy()).cuda()
    X, Y, BX, BY = dev(x), dev(y), dev(bx), dev(by)
    for c0, c1 in ((0.0, 1.0), (0.5, 0.5), (0.25, -1.5)):
        im.linear_combination(X, Y, BX, BY, (c0, c1))
        torch.cuda.synchronize()
        y = c0 * x + c1 * y
        by = c0 * bx + c1 * by
        assert np.array_equal(Y.cpu().numpy(), y.reshape(-1)), (c0, c1)
        assert np.array_equal(BY.cpu().numpy(), by.reshape(-1)), (c0, c1)


@pytest.mark.gpu
def test_rk2_fp32_bulk_close_to_fp64():
    import torch
    from godunov_mhd_ct_amd.solver import SolverGodunovMHD

    print("[rk2] FP32 bulk vs FP64, KH3D 8^3 blocks level 1, 10 steps", flush=True)
    m = uniform_mesh(3, 8, 1)
    U, B = problems.kelvin_helmholtz_3d(m, seed=1)
    st = hydro_settings(gamma0=1.4, cfl=0.8, slope_type=1.0, time_integrator=RK2)
    s64 = SolverGodunovMHD(m, U, B, st, impl_version=IMPL_FUSED)
    s32 = SolverGodunovMHD(m, U, B, st, impl_version=IMPL_FUSED_FP32)
    for _ in range(10):
        dt = s64.compute_dt()
        s64.step(dt=dt)
        s32.step(dt=dt)
    torch.cuda.synchronize()
    (U64, B64), (U32, B32) = s64.state(), s32.state()
    nl = m.num_leaves
    # rho, E, then momentum as one vector (its z component starts at zero)
    for name, sl in (("rho", slice(0, 1)), ("E", slice(1, 2)), ("momentum", slice(2, 5))):
        ref = U64[:nl, sl]
        err = np.abs(U32[:nl, sl] - ref).sum() / max(np.abs(ref).sum(), 1e-300)
        assert err <= 1e-5, (name, err)
    err = np.abs(B32[:nl] - B64[:nl]).sum() / np.abs(B64[:nl]).sum()
    assert err <= 1e-5, err


def rk2_batch_cases():
    """Three 3D slots: a uniform KH (MC limiter), an AMR blast, a walled blast (the outside
    fill of the stage buffer runs between the stages)."""
    cases = []
    m = uniform_mesh(3, 8, 1)
    U, B = problems.kelvin_helmholtz_3d(m, seed=5)
    cases.append((m, U, B, 2.0, None))
    m = refine_where(Mesh(3, 8, 1), lambda L, x0, x1, y0, y1, z0, z1: x0 < 0.5 and y0 < 0.5 and z0 < 0.5, 2)
    U, B = problems.blast(m, radius=0.3, center=(0.3, 0.3, 0.3))
    cases.append((m, U, B, 1.0, None))
    bc = [3, 3, 2, 2, 3, 3]
    m = Mesh(3, 8, 1, periodic=(0, 0, 0))
    m.set_bc(bc)
    m.finalize()
    U, B = problems.blast(m, radius=0.3, center=(0.4, 0.5, 0.5), bfield=(0.3, 0.2, 0.4))
    cases.append((m, U, B, 1.0, bc))
    return cases


def run_rk2_batch(cases, nsteps, graph=False, only=None):
    import torch
    from godunov_mhd_ct_amd.batch import BatchGodunovMHD

    bt = BatchGodunovMHD(3, IMPL_FUSED)
    if graph:
        bt.set_graph(True)
    ids = {}
    for n, (m, U, B, slope, bc) in enumerate(cases):
        if only is None or n == only:
            ids[n] = bt.add(m, U, B, hydro_settings(gamma0=5.0 / 3.0, cfl=0.7, slope_type=slope,
                                                    time_integrator=RK2))
    bt.step(nsteps)
    bt.check()
    torch.cuda.synchronize()
    out = {n: (bt.state(sid), bt.slot_state(sid)) for n, sid in ids.items()}
    return out


@pytest.mark.gpu
def test_rk2_batch_matches_oracle_per_slot():
    cases = rk2_batch_cases()
    nsteps = 3
    print(f"[rk2] batch of {len(cases)} RK2_SSP slots, {nsteps} steps, device dt", flush=True)
    res = run_rk2_batch(cases, nsteps)
    for n, (m, U, B, slope, bc) in enumerate(cases):
        ora = OracleSolver(m, settings_array(gamma0=5.0 / 3.0, cfl=0.7, slope_type=slope, time_integrator=RK2),
                           nthreads=4)
        Uo = U.reshape(m.num_total, 5, -1).copy()
        Bo = B.reshape(m.num_total, -1).copy()
        t = 0.0
        for _ in range(nsteps):
            Uo, Bo, dt = ora.step(Uo, Bo)
            t += dt
        (Ug, Bg), (cur, tg, steps) = res[n]
        assert steps == nsteps and abs(tg - t) <= 1e-14 * t, (n, tg, t)
        assert_parity(Ug, Bg, Uo, Bo, m.num_leaves, f"slot {n}")


@pytest.mark.gpu
def test_rk2_batch_alone_graph_and_refusals():
    import torch
    from godunov_mhd_ct_amd._lib import MhdctError
    from godunov_mhd_ct_amd.batch import BatchGodunovMHD

    cases = rk2_batch_cases()
    mix = run_rk2_batch(cases, 4)
    graphed = run_rk2_batch(cases, 4, graph=True)
    for n in range(len(cases)):
        alone = run_rk2_batch(cases, 4, only=n)
        no = cases[n][0].num_leaves
        for other in (alone, graphed):
            (Ua, Ba), _ = other[n]
            (Um, Bm), _ = mix[n]
            assert np.array_equal(Ua[:no], Um[:no]) and np.array_equal(Ba[:no], Bm[:no]), n
    # one integrator per batch
    m, U, B, slope, _ = cases[0]
    bt = BatchGodunovMHD(3, IMPL_FUSED)
    bt.add(m, U, B, hydro_settings(time_integrator=RK2))
    with pytest.raises(MhdctError, match="one time integrator"):
        bt.add(m, U.copy(), B.copy(), hydro_settings())
    torch.cuda.synchronize()
