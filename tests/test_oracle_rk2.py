"""The oracle's RK2_SSP integrator (SolverGodunovMHD.cpp:551-647) against what the mathematics
fixes, on the CPU.

* Amplification factor: with slope_type 0 (no slopes) a density wave carried by a uniform
  supersonic flow (B = 0, uniform p) is advanced by first-order upwinding, a linear scheme whose
  forward-Euler stage multiplies the Fourier mode k by 1 + z, z = -c (1 - exp(-i k dx)),
  c = u dt / dx.  One RK2_SSP step must multiply it by 1 + z + z^2 / 2 and one HANCOCK step by
  1 + z.  A wrong stage coefficient, a missing stage or a missing outside fill breaks this.
* One RK2_SSP stage in closed form: with minmod slopes and a smooth tangential discontinuity
  (rho and B_y varying along x, uniform total pressure, uniform supersonic u) every flux is the
  upwind one, so the stage is the textbook forward-Euler MUSCL update of rho and of the y-face
  field from the unpredicted reconstruction rho_i + slope_i / 2, B_y,i + slope_i / 2.  The
  HANCOCK step (half-step predictor, face-B source at the edges) must differ from it.
* Invariants under RK2_SSP: a uniform state is unchanged bit for bit on an AMR mesh; mass,
  momentum and energy are conserved and div B stays at roundoff on periodic AMR meshes.
* The stage combination y = c0 x + c1 y is exact for the two stage coefficient pairs:
  (1/2, 1/2) gives the correctly rounded mean and (0, 1) leaves y unchanged.
"""
import numpy as np
import pytest

from godunov_mhd_ct_amd import problems
from godunov_mhd_ct_amd.mesh import Mesh, refine_where, uniform_mesh
from oracle_ref import RK2_STAGE1, RK2_STAGE2, OracleSolver, settings_array

GAMMA = 5.0 / 3.0


def columns_x(mesh):
    """x index of every owned cell in a uniform 2D mesh of one level (global column)."""
    lv, co = mesh.leaves()
    b = mesh.b
    n = mesh.num_leaves
    i = np.tile(np.arange(b), b)[None, :] + (co[:n, 0:1] * b)
    return i  # [block, cell]


def xface_columns(mesh):
    """global column of every y-face (component 1) of the owned blocks (2D: (b+1) rows of b)."""
    lv, co = mesh.leaves()
    b = mesh.b
    n = mesh.num_leaves
    i = np.tile(np.arange(b), b + 1)[None, :] + (co[:n, 0:1] * b)
    return i


def state_1d(mesh, rho_of_col, by_of_col, p_of_col, u0):
    """2D state that varies along x only: rho, B_y (on y-faces), p per global column; Bx = Bz = 0."""
    n = mesh.num_total
    b = mesh.b
    assert mesh.num_outside == 0
    ic = columns_x(mesh)
    rho = rho_of_col(ic)
    mom = [rho * u0, np.zeros_like(rho), np.zeros_like(rho)]
    p = p_of_col(ic)
    nxf = (b + 1) * b
    bx = np.zeros((n, nxf))
    by = by_of_col(xface_columns(mesh))
    bz = np.zeros((n, b * b))
    return problems._assemble(mesh, rho, mom, p, [bx, by, bz], GAMMA)


def one_step(mesh, U, B, dt, time_integrator, slope_type):
    ora = OracleSolver(mesh, settings_array(gamma0=GAMMA, cfl=0.8, slope_type=slope_type,
                                            time_integrator=time_integrator), nthreads=2)
    Uo = U.reshape(mesh.num_total, 5, -1).copy()
    Bo = B.reshape(mesh.num_total, -1).copy()
    U2, B2, _ = ora.step(Uo, Bo, dt=dt)
    return U2, B2, ora


@pytest.mark.parametrize("kmode", [1, 3])
def test_rk2_amplification_factor(kmode):
    m = uniform_mesh(2, 4, 2)  # 16 x 16 cells, periodic
    ncol = 16
    dx = 1.0 / ncol
    u0, a = 3.0, 0.2
    k = 2 * np.pi * kmode
    xc = lambda ic: (ic + 0.5) * dx
    U, B = state_1d(m, lambda ic: 1.0 + a * np.cos(k * xc(ic)), lambda ic: np.zeros(ic.shape),
                    lambda ic: np.full(ic.shape, 0.6), u0)
    dt = 0.4 * dx / u0
    c = u0 * dt / dx
    z = -c * (1.0 - np.exp(-1j * k * dx))
    ic = columns_x(m)
    for ti, G in (("RK2_SSP", 1 + z + z * z / 2), ("HANCOCK", 1 + z)):
        U2, _, _ = one_step(m, U, B, dt, ti, slope_type=0.0)
        want = 1.0 + a * np.real(G * np.exp(1j * k * xc(ic)))
        err = np.abs(U2[:m.num_leaves, 0] - want).max()
        assert err < 1e-13, f"{ti}: density differs from the amplification-factor closed form by {err}"
    # the two integrators really differ here (the check above discriminates them)
    Urk, _, _ = one_step(m, U, B, dt, "RK2_SSP", 0.0)
    Uh, _, _ = one_step(m, U, B, dt, "HANCOCK", 0.0)
    assert np.abs(Urk[:, 0] - Uh[:, 0]).max() > 1e-4


def minmod_slope(qc, qp, qm):
    dl, dr = qc - qm, qp - qc
    dcen = 0.5 * (qp - qm)
    lim = np.where(dl * dr <= 0.0, 0.0, np.minimum(np.abs(dl), np.abs(dr)))
    return np.where(dcen >= 0.0, 1.0, -1.0) * np.minimum(lim, np.abs(dcen))


def test_rk2_stage_closed_form():
    m = uniform_mesh(2, 4, 2)  # 16 x 16 cells, periodic
    ncol = 16
    dx = 1.0 / ncol
    u0 = 4.0
    x = lambda ic: (ic % ncol + 0.5) * dx
    rho_f = lambda ic: 1.0 + 0.3 * np.sin(2 * np.pi * x(ic))
    by_f = lambda ic: 0.6 + 0.3 * np.cos(2 * np.pi * x(ic))
    ptot = 1.0
    p_f = lambda ic: ptot - 0.5 * by_f(ic) ** 2
    U, B = state_1d(m, rho_f, by_f, p_f, u0)
    dt = 0.3 * dx / u0
    dtdx = dt / dx
    # one forward-Euler stage: do_time_step with the RK2_SSP functor variants
    ora = OracleSolver(m, settings_array(gamma0=GAMMA, cfl=0.8, slope_type=1.0, time_integrator="RK2_SSP"),
                       nthreads=2)
    Uo = U.reshape(m.num_total, 5, -1).copy()
    Bo = B.reshape(m.num_total, -1).copy()
    U1, B1 = Uo.copy(), Bo.copy()
    ora.do_time_step(Uo, U1, Bo, B1, dt)
    cols = np.arange(ncol)
    r = rho_f(cols)
    by = by_f(cols)
    rl = r + 0.5 * minmod_slope(r, np.roll(r, -1), np.roll(r, 1))      # left state at i + 1/2
    bl = by + 0.5 * minmod_slope(by, np.roll(by, -1), np.roll(by, 1))
    r_new = r - dtdx * u0 * (rl - np.roll(rl, 1))
    by_new = by - dtdx * u0 * (bl - np.roll(bl, 1))
    ic = columns_x(m)
    err_r = np.abs(U1[:m.num_leaves, 0] - r_new[ic]).max()
    nxf = (m.b + 1) * m.b
    fc = xface_columns(m)
    err_b = np.abs(B1[:m.num_leaves, nxf:2 * nxf] - by_new[fc]).max()
    assert err_r < 1e-13, f"rho differs from the forward-Euler MUSCL stage by {err_r}"
    assert err_b < 1e-13, f"B_y differs from the forward-Euler MUSCL stage by {err_b}"
    # B_x stays zero, and the HANCOCK step (predictor, face-B source at the edges) is another scheme
    assert np.abs(B1[:m.num_leaves, :nxf]).max() == 0.0
    oh = OracleSolver(m, settings_array(gamma0=GAMMA, cfl=0.8, slope_type=1.0), nthreads=2)
    Uh, Bh = Uo.copy(), Bo.copy()
    oh.do_time_step(Uo, Uh, Bo, Bh, dt)
    assert np.abs(Uh[:m.num_leaves, 0] - r_new[ic]).max() > 1e-5
    assert np.abs(Bh[:m.num_leaves, nxf:2 * nxf] - by_new[fc]).max() > 1e-5


def amr2d(b=4):
    m = Mesh(2, b, 2)
    return refine_where(m, lambda L, x0, x1, y0, y1, z0, z1: x0 < 0.6 and x1 > 0.4 and y0 < 0.6 and y1 > 0.4, 3)


def amr3d(b=4):
    m = Mesh(3, b, 1)
    return refine_where(m, lambda L, x0, x1, y0, y1, z0, z1: x0 < 0.5 and y0 < 0.5 and z0 < 0.5, 2)


def rk2_settings():
    return settings_array(gamma0=GAMMA, cfl=0.8, slope_type=1.0, time_integrator="RK2_SSP")


@pytest.mark.parametrize("mk", [amr2d, amr3d])
def test_rk2_uniform_state_unchanged(mk):
    from test_oracle_invariants import uniform_state

    m = mk()
    U, B = uniform_state(m)
    ora = OracleSolver(m, rk2_settings(), nthreads=2)
    Uo = U.reshape(m.num_total, 5, -1).copy()
    Bo = B.reshape(m.num_total, -1).copy()
    for _ in range(2):
        Uo, Bo, _ = ora.step(Uo, Bo)
    n = m.num_leaves
    assert np.array_equal(Uo[:n], U.reshape(m.num_total, 5, -1)[:n])
    assert np.array_equal(Bo[:n], B.reshape(m.num_total, -1)[:n])


@pytest.mark.parametrize("dim", [2, 3])
def test_rk2_conservation_and_div_b(dim):
    from test_oracle_invariants import cell_volumes, div_b

    m = amr2d() if dim == 2 else amr3d()
    center = (0.3, 0.3, 0.3) if dim == 3 else (0.5, 0.5, 0.0)
    U, B = problems.blast(m, radius=0.3, center=center, bfield=(0.6, 0.3, 0.2))
    ora = OracleSolver(m, rk2_settings(), nthreads=2)
    no = m.num_leaves
    Uo = U.reshape(m.num_total, 5, -1).copy()
    Bo = B.reshape(m.num_total, -1).copy()
    vol = cell_volumes(m)[:, None, None]
    tot0 = (Uo[:no] * vol).sum(axis=(0, 2))
    scale = np.abs(Uo[:no] * vol).sum(axis=(0, 2)).max()
    d0 = np.abs(div_b(m, Bo)).max()
    for _ in range(3):
        Uo, Bo, _ = ora.step(Uo, Bo)
    tot = (Uo[:no] * vol).sum(axis=(0, 2))
    assert np.all(np.abs(tot - tot0) <= 1e-13 * scale), (tot - tot0, scale)
    d = np.abs(div_b(m, Bo)).max()
    assert d <= max(1e-13 * np.abs(Bo[:no]).max(), 4 * d0), (d, d0)


def test_linear_combination_exact():
    rng = np.random.default_rng(5)
    m = uniform_mesh(3, 4, 1)
    ora = OracleSolver(m, rk2_settings())
    n = m.num_total
    U = rng.uniform(0.1, 10.0, (n, 5, m.cells_per_block))
    Y = rng.uniform(-10.0, 10.0, (n, 5, m.cells_per_block))
    B = rng.uniform(-1.0, 1.0, (n, m.faces_per_block))
    YB = rng.uniform(-1.0, 1.0, (n, m.faces_per_block))
    y, yb = Y.copy(), YB.copy()
    ora.linear_combination(U, y, B, yb, RK2_STAGE1)
    assert np.array_equal(y, Y) and np.array_equal(yb, YB)
    ora.linear_combination(U, y, B, yb, RK2_STAGE2)
    # 0.5 x and 0.5 y are exact, so the sum is the correctly rounded mean
    assert np.array_equal(y, (U + Y) / 2) and np.array_equal(yb, (B + YB) / 2)
