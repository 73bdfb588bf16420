"""GPU checks at the full sizes of the benchmark configurations (godunov_mhd_ct_amd/workloads.py),
where the oracle is too slow to follow every step: size-independent properties plus agreement
between the forms.

* OT2D (128^2 cells, 16 x 16 blocks of 8^2): small enough for the oracle; 25 steps of the fused
  form against it at the 1e-12 relative L-inf bar.
* BL3D (2808 leaves of 16^3, 11.5 M cells):
  - 10 fused steps conserve mass, momentum and energy to 1e-12 of their scale (periodic box);
  - div B stays at roundoff;
  - the plain and fused forms agree after 3 steps at the 1e-12 bar (equal dt);
  - a second run of the fused form is bitwise equal to the first (determinism).
* MIX (three meshes, 8^3, 16^3 and 32^3 blocks, 29 M cells, one batch): 5 batch steps conserve
  every slot's totals, and every slot equals the same mesh run in a batch of its own.
"""
import numpy as np
import pytest

from parity import assert_parity
from godunov_mhd_ct_amd import workloads
from godunov_mhd_ct_amd.implem import IMPL_FUSED, IMPL_PLAIN
from parity import TOL, rel_linf
from test_gpu_parity import run_pair
from test_oracle_invariants import cell_volumes, div_b


def totals(mesh, U):
    n = mesh.num_leaves
    vol = cell_volumes(mesh)
    X = U.reshape(mesh.num_total, 5, -1)[:n]
    return (X * vol[:, None, None]).sum(axis=(0, 2)), (np.abs(X) * vol[:, None, None]).sum(axis=(0, 2))


@pytest.mark.gpu
def test_ot2d_full_config_against_oracle():
    m, U, B, st = workloads.ot2d()
    assert m.num_leaves == 256 and m.cells_per_block == 64
    run_pair(m, U, B, st.gamma0, st.cfl, st.slope_type, 25, IMPL_FUSED)


@pytest.mark.gpu
def test_bl3d_full_size_invariants_and_forms():
    import torch
    from godunov_mhd_ct_amd.solver import SolverGodunovMHD

    m, U, B, st = workloads.bl3d()
    print(f"[full] BL3D {m.num_leaves} leaves, {m.owned_cells()} cells", flush=True)
    assert m.num_leaves == 2808
    t0, s0 = totals(m, U)
    d0 = np.abs(div_b(m, B.reshape(m.num_total, -1))).max()
    # fused vs plain at equal dt, 3 steps
    fz = SolverGodunovMHD(m, U, B, st, impl_version=IMPL_FUSED)
    pl = SolverGodunovMHD(m, U, B, st, impl_version=IMPL_PLAIN)
    for _ in range(3):
        dt = fz.compute_dt()
        dtp = pl.compute_dt()
        assert abs(dt - dtp) <= 1e-14 * dt
        fz.step(dt)
        pl.step(dt)
    torch.cuda.synchronize()
    Uf, Bf = fz.state()
    Up, Bp = pl.state()
    n = m.num_leaves
    assert_parity(Uf, Bf, Up, Bp, n, "fused vs plain")
    del pl
    torch.cuda.empty_cache()
    # 7 more fused steps: totals and div B
    for _ in range(7):
        fz.step()
    torch.cuda.synchronize()
    Uf, Bf = fz.state()
    t1, _ = totals(m, Uf)
    assert np.all(np.abs(t1 - t0) <= 1e-12 * s0.max()), (t1 - t0) / s0.max()
    d1 = np.abs(div_b(m, Bf)).max()
    assert d1 <= max(1e-13 * np.abs(Bf[:n]).max(), 4 * d0), (d1, d0)
    # determinism: the same 10 steps again give the same bits
    again = SolverGodunovMHD(m, U, B, st, impl_version=IMPL_FUSED)
    for _ in range(10):
        again.step()
    torch.cuda.synchronize()
    Ua, Ba = again.state()
    assert np.array_equal(Ua[:n], Uf[:n]) and np.array_equal(Ba[:n], Bf[:n])


@pytest.mark.gpu
def test_mix_full_size_batch():
    import torch
    from godunov_mhd_ct_amd.batch import BatchGodunovMHD

    cases = workloads.mix3d(seed=0)
    print(f"[full] MIX {[m.num_leaves for (m, _, _, _) in cases]} leaves", flush=True)
    bt = BatchGodunovMHD(3, IMPL_FUSED)
    ids = [bt.add(m, U, B, st) for (m, U, B, st) in cases]
    bt.step(5)
    bt.check()
    torch.cuda.synchronize()
    for sid, (m, U, B, st) in zip(ids, cases):
        t0, s0 = totals(m, U)
        Ug, Bg = bt.state(sid)
        t1, _ = totals(m, Ug)
        assert np.all(np.abs(t1 - t0) <= 1e-12 * s0.max()), (sid, (t1 - t0) / s0.max())
    # each slot equals its mesh advanced alone
    for sid, (m, U, B, st) in zip(ids, cases):
        solo = BatchGodunovMHD(3, IMPL_FUSED)
        k = solo.add(m, U, B, st)
        solo.step(5)
        torch.cuda.synchronize()
        Us, Bs = solo.state(k)
        Ug, Bg = bt.state(sid)
        n = m.num_leaves
        assert np.array_equal(Us[:n], Ug[:n]) and np.array_equal(Bs[:n], Bg[:n]), sid
        del solo
        torch.cuda.empty_cache()
