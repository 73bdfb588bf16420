"""The oracle's uniform gravity (AddGravitySourceTerm.cpp, the predictor's half-step kick in
ComputeUpdatedPrimitiveVariablesFunctor.cpp:372-377 and :505-511, and
ComputeDtMHDFunctor::compute_cfl_with_gravity, .cpp:108-249) against what physics and algebra
fix, on the CPU:

* a uniform state under gravity keeps uniform fluxes, so one step changes the momentum by exactly
  rho g dt and leaves rho and B unchanged; the total energy changes by the kinetic-energy change
  (HANCOCK) or by the mean of the two stages' changes (RK2_SSP);
* the time step with gravity solves the CFL condition it is built from:
  g dt'^2 / (2 dx) + |v| dt' / dx = 1 for dt' = (dt / cfl) sum_d v_d / |v|, with v_d the signal
  speeds (the oracle's speed function, pinned by its own known-answer test) and g = sum |g_d|
  (the code's k carries no cfl factor, so the cfl = 1 step is scaled by cfl);
* a magnetic Rayleigh-Taylor state in a periodic box keeps its mass and div B, and its total
  momentum changes by dt g (M_old + M_new) / 2 per step.
"""
import numpy as np
import pytest

from godunov_mhd_ct_amd import problems
from godunov_mhd_ct_amd.mesh import Mesh, uniform_mesh
from oracle_ref import OracleSolver, dptr, lib, settings_array
from test_oracle_invariants import div_b, uniform_state

GAMMA = 5.0 / 3.0


@pytest.mark.parametrize("ti", ["HANCOCK", "RK2_SSP"])
@pytest.mark.parametrize("dim", [2, 3])
def test_uniform_state_gains_rho_g_dt(dim, ti):
    m = uniform_mesh(dim, 4, 1)
    g = (0.3, -0.7, 0.45)[:dim]
    rho, vel = 1.3, (0.3, -0.2, 0.1)
    U, B = uniform_state(m, rho=rho, p=0.7, v=vel, bvec=(0.4, -0.25, 0.6), gamma=GAMMA)
    ora = OracleSolver(m, settings_array(gamma0=GAMMA, cfl=0.5, time_integrator=ti, gravity=g), nthreads=2)
    Uo = U.reshape(m.num_total, 5, -1).copy()
    Bo = B.reshape(m.num_total, -1).copy()
    dt = 0.01
    U2, B2, _ = ora.step(Uo, Bo, dt=dt)
    n = m.num_leaves
    assert np.array_equal(U2[:n, 0], Uo[:n, 0])
    assert np.array_equal(B2[:n], Bo[:n])
    mom0 = np.array([rho * x for x in vel])
    a = np.zeros(3)
    a[:dim] = rho * np.array(g) * dt
    mom1 = mom0 + a
    for d in range(3):
        assert np.abs(U2[:n, 2 + d] - mom1[d]).max() <= 1e-15 * 8, d
    if ti == "HANCOCK":
        de = 0.5 * (mom1 @ mom1 - mom0 @ mom0) / rho
    else:
        # the mean of two forward-Euler stages: E0 + (|m0 + 2a|^2 - |m0|^2) / (4 rho), which
        # exceeds the kinetic-energy change by |a|^2 / (2 rho)
        m2 = mom0 + 2 * a
        de = 0.25 * (m2 @ m2 - mom0 @ mom0) / rho
    assert np.abs(U2[:n, 1] - (Uo[:n, 1] + de)).max() <= 1e-14


@pytest.mark.parametrize("dim", [2, 3])
def test_dt_with_gravity_solves_its_cfl_condition(dim):
    import ctypes as C

    m = uniform_mesh(dim, 4, 1)
    gvec = (0.4, -2.5, 1.2)[:dim]
    rho, p, vel, bvec = 1.1, 0.9, (0.5, -0.3, 0.2), (0.3, 0.5, -0.4)
    U, B = uniform_state(m, rho=rho, p=p, v=vel, bvec=bvec, gamma=GAMMA)
    cfl = 0.8
    ora = OracleSolver(m, settings_array(gamma0=GAMMA, cfl=cfl, gravity=gvec), nthreads=2)
    dt = ora.compute_dt(U.reshape(m.num_total, 5, -1).copy(), B.reshape(m.num_total, -1).copy())
    # signal speeds of the uniform state from the oracle's speed function
    q = np.array([rho, p, *vel, *bvec])
    v = np.zeros(3)
    st = settings_array(gamma0=GAMMA)
    lib().oracle_speed_info(dptr(st), dptr(q), dim, dptr(v))
    v = v[:dim]
    dx = 1.0 / (4 * 2)
    G = sum(abs(x) for x in gvec)
    vnorm = np.sqrt((v * v).sum())
    # the code scales the cfl = 1 step by cfl (its k holds no cfl factor)
    dtp = dt / cfl * v.sum() / vnorm
    lhs = G * dtp * dtp / (2 * dx) + vnorm * dtp / dx
    assert abs(lhs - 1.0) <= 1e-13, lhs
    # and gravity shortens the step
    ora0 = OracleSolver(m, settings_array(gamma0=GAMMA, cfl=cfl), nthreads=2)
    assert dt < ora0.compute_dt(U.reshape(m.num_total, 5, -1).copy(), B.reshape(m.num_total, -1).copy())


@pytest.mark.parametrize("dim", [2, 3])
def test_rayleigh_taylor_momentum_budget(dim):
    # periodic box: the fluxes conserve mass and momentum, so each step adds exactly
    # dt g (M_old + M_new) / 2 to the total momentum (M the total mass, equal cells), and
    # div B stays at roundoff.  (With reflecting walls the predictor's gravity kick is applied
    # to the mirrored ghost cells too, so mass crosses the walls: a property of the scheme.)
    nb = (1, 3, 1) if dim == 2 else (1, 1, 2)
    m = Mesh(dim, 4, 2 if dim == 2 else 1, nbrick=nb).finalize()
    g = (0.0, -0.1) if dim == 2 else (0.0, 0.0, -0.1)
    U, B = problems.rayleigh_taylor(m, g, bfield=(0.0, 0.5, 0.0) if dim == 2 else (0.0, 0.0, 0.5),
                                    xmin=(-0.5, -1.5, 0.0) if dim == 2 else (-0.5, -0.5, -1.0))
    ora = OracleSolver(m, settings_array(gamma0=1.4, cfl=0.7, gravity=g), nthreads=2)
    Uo = U.reshape(m.num_total, 5, -1).copy()
    Bo = B.reshape(m.num_total, -1).copy()
    n = m.num_leaves
    d0 = np.abs(div_b(m, Bo)).max()
    for _ in range(4):
        M0, P0 = Uo[:n, 0].sum(), Uo[:n, 2:5].sum(axis=(0, 2))
        Uo, Bo, dt = ora.step(Uo, Bo)
        M1, P1 = Uo[:n, 0].sum(), Uo[:n, 2:5].sum(axis=(0, 2))
        assert abs(M1 - M0) <= 1e-13 * M0
        want = np.zeros(3)
        want[:dim] = dt * np.array(g) * (M0 + M1) / 2
        assert np.abs((P1 - P0) - want).max() <= 1e-12 * max(np.abs(Uo[:n, 2:5]).sum(), 1.0)
    assert np.abs(div_b(m, Bo)).max() <= max(1e-13, 4 * d0)


def test_predictor_kick_shifts_upwind_advection():
    # First order (slope_type 0), supersonic flow along x, B = 0, gravity along x: the Hancock
    # predictor's only change is the half-step kick u + g dt / 2, so every mass flux is the
    # upwind rho_{i-1} (u0 + g dt / 2) and rho_i' = rho_i - dt/dx (u0 + g dt/2) (rho_i - rho_{i-1}).
    from test_oracle_rk2 import columns_x, state_1d

    m = uniform_mesh(2, 4, 2)
    dx = 1.0 / 16
    u0, gx = 3.0, 0.8
    rho_f = lambda ic: 1.0 + 0.2 * np.cos(2 * np.pi * (ic + 0.5) * dx)  # noqa: E731
    U, B = state_1d(m, rho_f, lambda ic: np.zeros(ic.shape), lambda ic: np.full(ic.shape, 0.6), u0)
    dt = 0.3 * dx / u0
    ora = OracleSolver(m, settings_array(gamma0=GAMMA, cfl=0.8, slope_type=0.0, gravity=(gx, 0.0)), nthreads=2)
    Uo = U.reshape(m.num_total, 5, -1).copy()
    Bo = B.reshape(m.num_total, -1).copy()
    U1, B1 = Uo.copy(), Bo.copy()
    ora.do_time_step(Uo, U1, Bo, B1, dt)
    cols = np.arange(16)
    r = rho_f(cols)
    want = r - dt / dx * (u0 + gx * dt / 2) * (r - np.roll(r, 1))
    ic = columns_x(m)
    assert np.abs(U1[:m.num_leaves, 0] - want[ic]).max() <= 1e-14
    # without the kick the result would differ by dt/dx (g dt / 2) (rho_i - rho_{i-1})
    nokick = r - dt / dx * u0 * (r - np.roll(r, 1))
    assert np.abs(want - nokick).max() > 1e-6
