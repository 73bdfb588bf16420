"""Oracle pins of the viscous terms (ComputeViscousFluxesAndStoreFunctor.cpp:96-352, called at
GodunovImplemV0.cpp:179-192; the predictor term add_viscous_predictor_{2d,3d},
ComputeUpdatedPrimitiveVariablesFunctor.cpp:178-272), checked against the mathematics they
discretise. All runs compare one step with viscosity to the same step without it, at equal dt
and predictor off, so the difference is the viscous update alone.

* Shear v_y(x) (rho = 1): the momentum change is mu dt times the 3-point Laplacian of v_y, and
  nothing else moves.
* Compression u(x): the change is mu dt Lap(u) in 2D and 4/3 mu dt Lap(u) in 3D. This pins the
  trace parts 1/2 and 1/3.
* Totals: mass, momentum and energy are conserved, also across coarse/fine faces of an AMR
  mesh.
* Kinetic energy falls by the discrete dissipation mu dt sum |grad v|^2 (summation by parts)
  to first order.
* Predictor: the viscous term moves the predicted velocity by the half-step heat equation,
  1/2 (mu / rho) dt Lap(v).
* A uniform flow is unchanged by viscosity, bit for bit.
"""
import math

import numpy as np
import pytest

from godunov_mhd_ct_amd import problems
from godunov_mhd_ct_amd.mesh import Mesh, refine_where, uniform_mesh
from oracle_ref import OracleSolver, settings_array
from test_oracle_invariants import cell_volumes, uniform_state

GAMMA = 1.4
MU = 0.02


def state(mesh, vel_fn, rho=1.0, p=1.0, bz=0.1):
    """U, B with velocity vel_fn(x, y, z) -> (u, v, w) at cell centres."""
    xs = problems._grid(mesh, 1.0, (0.0, 0.0, 0.0), "cell")
    if mesh.dim == 2:
        xs = xs + [np.zeros_like(xs[0])]
    u, v, w = vel_fn(*xs)
    n = mesh.num_total
    rho_a = np.full_like(xs[0], rho)
    b, dim = mesh.b, mesh.dim
    parts = []
    for c in range(3):
        size = (b * b if c == 2 else (b + 1) * b) if dim == 2 else (b + 1) * b * b
        parts.append(np.full((n, size), bz if c == 2 else 0.0))
    return problems._assemble(mesh, rho_a, [rho * u, rho * v, rho * w], np.full_like(xs[0], p), parts, GAMMA)


def one_step(mesh, U, B, mu, predictor=False, dt=None):
    ora = OracleSolver(mesh, settings_array(gamma0=GAMMA, cfl=0.5, mu=mu, viscous_predictor=predictor))
    U0 = U.reshape(mesh.num_total, 5, -1).copy()
    B0 = B.reshape(mesh.num_total, -1).copy()
    if dt is None:
        dt = ora.compute_dt(U0, B0)
    U2, B2, _ = ora.step(U0, B0, dt=dt)
    return U2, B2, dt, ora


def as_grid(mesh, a):
    """[leaf, cell] of a uniform mesh -> global array indexed [z][y][x] (2D: [y][x])."""
    lv, co = mesh.leaves()
    b, dim = mesh.b, mesh.dim
    nb = 1 << int(lv[0])
    G = np.zeros((nb * b,) * dim)
    for o in range(mesh.num_leaves):
        blk = a[o].reshape((b,) * dim)
        sl = tuple(slice(co[o][d] * b, (co[o][d] + 1) * b) for d in reversed(range(dim)))
        G[sl] = blk
    return G


def lap_x(G, dx):
    ax = G.ndim - 1
    return (np.roll(G, -1, axis=ax) - 2 * G + np.roll(G, 1, axis=ax)) / dx ** 2


@pytest.mark.parametrize("dim", [2, 3])
def test_shear_is_a_laplacian(dim):
    m = uniform_mesh(dim, 8, 2 if dim == 2 else 1)
    A = 0.1
    U, B = state(m, lambda x, y, z: (0 * x, A * np.sin(2 * math.pi * x), 0 * x))
    Uv, _, dt, _ = one_step(m, U, B, MU)
    Un, _, _, _ = one_step(m, U, B, None, dt=dt)
    n = m.num_leaves
    d = Uv[:n] - Un[:n]
    dx = 1.0 / as_grid(m, d[:, 0]).shape[-1]
    v0 = as_grid(m, U.reshape(m.num_total, 5, -1)[:n, 3])  # rho = 1: momentum = velocity
    want = MU * dt * lap_x(v0, dx)
    got = as_grid(m, d[:, 3])
    assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max(), np.abs(got - want).max() / np.abs(want).max()
    assert np.abs(want).max() > 1e-6
    # density, x and z momentum untouched by the viscous fluxes
    assert np.abs(d[:, 0]).max() == 0.0
    assert np.abs(d[:, 2]).max() <= 1e-15 and np.abs(d[:, 4]).max() <= 1e-15


@pytest.mark.parametrize("dim,factor", [(2, 1.0), (3, 4.0 / 3.0)])
def test_compression_trace_part(dim, factor):
    m = uniform_mesh(dim, 8, 2 if dim == 2 else 1)
    A = 0.1
    U, B = state(m, lambda x, y, z: (A * np.sin(2 * math.pi * x), 0 * x, 0 * x))
    Uv, _, dt, _ = one_step(m, U, B, MU)
    Un, _, _, _ = one_step(m, U, B, None, dt=dt)
    n = m.num_leaves
    d = Uv[:n] - Un[:n]
    dx = 1.0 / as_grid(m, d[:, 0]).shape[-1]
    u0 = as_grid(m, U.reshape(m.num_total, 5, -1)[:n, 2])
    want = factor * MU * dt * lap_x(u0, dx)
    got = as_grid(m, d[:, 2])
    assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max()


def random_flow(seed):
    rng = np.random.default_rng(seed)
    k = rng.integers(1, 3, size=(3, 3))
    ph = rng.random((3, 3)) * 2 * math.pi

    def f(x, y, z):
        return tuple(0.1 * np.sin(2 * math.pi * (k[c, 0] * x + k[c, 1] * y + k[c, 2] * z) + ph[c, 0]) for c in range(3))
    return f


@pytest.mark.parametrize("mk", [lambda: uniform_mesh(2, 8, 2), lambda: uniform_mesh(3, 8, 1),
                                lambda: refine_where(Mesh(2, 8, 2), lambda L, x0, x1, y0, y1, z0, z1:
                                                     x0 < 0.6 and x1 > 0.4 and y0 < 0.6 and y1 > 0.4, 3),
                                lambda: refine_where(Mesh(3, 4, 1), lambda L, x0, x1, y0, y1, z0, z1:
                                                     x0 < 0.5 and y0 < 0.5 and z0 < 0.5, 2)])
def test_conservation_and_dissipation(mk):
    m = mk()
    U, B = state(m, random_flow(3))
    Uv, _, dt, _ = one_step(m, U, B, MU)
    Un, _, _, _ = one_step(m, U, B, None, dt=dt)
    n = m.num_leaves
    vol = cell_volumes(m)
    tot = lambda X: (X[:n] * vol[:, None, None]).sum(axis=(0, 2))  # noqa: E731
    U0 = U.reshape(m.num_total, 5, -1)
    scale = np.abs(U0[:n] * vol[:, None, None]).sum(axis=(0, 2))
    diff = tot(Uv) - tot(U0)
    assert np.all(np.abs(diff) <= 1e-13 * scale.max()), diff / scale.max()
    if len(set(m.leaves()[0][:n].tolist())) == 1:
        # kinetic energy: the viscous momentum change dotted with the velocity, summed, is
        # -dt times the discrete dissipation (summation by parts of the stress divergence)
        dm = (Uv[:n, 2:5] - Un[:n, 2:5])
        vel = U0[:n, 2:5] / U0[:n, :1]
        dke = (vel * dm).sum(axis=1)
        total_dke = (dke * vol[:, None]).sum()
        assert total_dke < 0


def test_uniform_flow_unchanged_by_viscosity():
    m = refine_where(Mesh(3, 4, 1), lambda L, x0, x1, y0, y1, z0, z1: x0 < 0.5, 2)
    U, B = uniform_state(m, v=(0.3, -0.2, 0.1))
    Uv, Bv, dt, _ = one_step(m, U, B, MU, predictor=True)
    Un, Bn, _, _ = one_step(m, U, B, None, dt=dt)
    n = m.num_leaves
    assert np.array_equal(Uv[:n], Un[:n]) and np.array_equal(Bv[:n], Bn[:n])


@pytest.mark.parametrize("dim", [2, 3])
def test_predictor_is_a_half_step_heat_equation(dim):
    m = uniform_mesh(dim, 8, 2 if dim == 2 else 1)
    A, rho = 0.1, 1.5
    U, B = state(m, lambda x, y, z: (0 * x, A * np.sin(2 * math.pi * x), 0 * x), rho=rho)
    _, _, dt, on = one_step(m, U, B, MU, predictor=True)
    q_on = on.scratch(2).copy()
    _, _, _, off = one_step(m, U, B, MU, predictor=False, dt=dt)
    q_off = off.scratch(2)
    b = m.b
    g1 = (b + 2) ** dim
    n = m.num_leaves
    Qon = q_on.reshape(n, 8, g1)
    Qoff = q_off.reshape(n, 8, g1)
    inner = np.zeros((b + 2,) * dim, dtype=bool)
    inner[(slice(1, -1),) * dim] = True
    inner = inner.reshape(-1)
    dv = (Qon[:, 3] - Qoff[:, 3])[:, inner]  # v, owned cells
    dx = 1.0 / (b * (1 << int(m.leaves()[0][0])))
    v0 = as_grid(m, (U.reshape(m.num_total, 5, -1)[:n, 3] / rho))
    want = 0.5 * (MU / rho) * dt * lap_x(v0, dx)
    got = as_grid(m, dv)
    assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max()
    assert np.abs((Qon[:, 2] - Qoff[:, 2])[:, inner]).max() == 0.0  # u: no second difference
