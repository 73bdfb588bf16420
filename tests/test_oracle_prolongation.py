"""Pins of the oracle's ghost-cell prolongation (EXTRAPOLATE_LINEAR_MINMOD,
scheme/ConvertToPrimitivesVariablesFunctor.cpp:471-635, chosen at :697-710) by closed forms.

A ghost cell beside a coarser block takes the conserved variables of the coarse cell that
contains it, extrapolated to its own centre with the minmod slopes of the coarse cell (a
quarter of a coarse slope per axis, towards the half of the coarse cell the ghost lies in):

* a field linear in x, y (and z) is prolonged exactly: every ghost cell of every fine block
  (same-level copies, fine-neighbour means and coarse-neighbour extrapolations alike) holds the
  field's value at its own centre, in 2D and 3D; SIMPLE_COPY, by contrast, leaves the coarse
  cell's value there;
* a field with an extremum at a coarse cell has a zero slope there, so the ghost cells inside
  that coarse cell take its value unchanged;
* the prolongation changes nothing but the ghost cells beside coarser blocks: on a uniform mesh
  both settings give the same step bit for bit.

The slopes' stencil beyond the coarse block follows the ghost-cell rules of the same functor
(copy, mean of the finer cells, containing coarser cell); that part restates kalypsso::core's
StencilHelper (absent from the reference tree) and is checked here only through these cases.
"""
import numpy as np
import pytest

from godunov_mhd_ct_amd import problems
from godunov_mhd_ct_amd.mesh import Mesh, refine_where, uniform_mesh
from oracle_ref import OracleSolver, settings_array

GAMMA = 1.4


def central_mesh(dim):
    # 4^dim blocks at level 2, those covering [0.25, 0.75]^dim refined once: every ghost cell and
    # every slope stencil of the fine blocks lies inside [0.15, 0.85]^dim, away from the wrap
    return refine_where(Mesh(dim, 8, 2), lambda L, x0, x1, y0, y1, z0, z1: 0.25 <= x0 < 0.75 and 0.25 <= y0 < 0.75
                        and (dim == 2 or 0.25 <= z0 < 0.75), 3)


def ghost_centres(mesh, o):
    """Centres (3, n) and (i, j, k) of the 2-ghost-layer cells of block o, ghost layer only."""
    b, dim = mesh.b, mesh.dim
    lv, co = mesh.leaves()
    h = 1.0 / (b * (1 << int(lv[o])))
    rng = np.arange(-2, b + 2)
    grids = np.meshgrid(*([rng] * dim), indexing="ij")
    idx = [g.reshape(-1) for g in grids]
    ghost = np.zeros(idx[0].size, dtype=bool)
    for d in range(dim):
        ghost |= (idx[d] < 0) | (idx[d] >= b)
    idx = [a[ghost] for a in idx]
    cen = [(co[o, d] * b + idx[d] + 0.5) * h for d in range(dim)]
    return cen, idx


def q_at(Q, mesh, o, v, idx):
    b, dim = mesh.b, mesh.dim
    n = b + 4
    g = (idx[0] + 2) + n * ((idx[1] + 2) + (n * (idx[2] + 2) if dim == 3 else 0))
    return Q.reshape(mesh.num_total, 11, -1)[o, v, g]


def linear_state(mesh, grad_rho, grad_e):
    """rho and E linear in the cell centres, no momentum, no field: p = (gamma - 1) E."""
    dim = mesh.dim
    xs = problems._grid(mesh, 1.0, (0.0, 0.0, 0.0), "cell")
    rho = 1.0 + sum(g * x for g, x in zip(grad_rho, xs))
    E = 2.0 + sum(g * x for g, x in zip(grad_e, xs))
    n = mesh.num_total
    U = np.zeros((n, 5, mesh.cells_per_block))
    U[:, 0] = rho
    U[:, 1] = E
    B = np.zeros((n, mesh.faces_per_block))
    return U, B


def primitives_after_step(mesh, U, B, prolongation):
    ora = OracleSolver(mesh, settings_array(gamma0=GAMMA, cfl=0.2, prolongation=prolongation), nthreads=4)
    ora.step(U.copy(), B.copy())
    return ora.scratch(1)


@pytest.mark.parametrize("dim", [2, 3])
def test_linear_field_prolonged_exactly(dim):
    mesh = central_mesh(dim)
    lv, _ = mesh.leaves()
    gr, ge = (0.3, -0.2, 0.15)[:dim], (0.5, 0.25, -0.4)[:dim]
    U, B = linear_state(mesh, gr, ge)
    worst = {}
    for prol in ("EXTRAPOLATE_LINEAR_MINMOD", "SIMPLE_COPY"):
        Q = primitives_after_step(mesh, U, B, prol)
        err = 0.0
        for o in np.nonzero(lv[:mesh.num_leaves] == 3)[0]:
            cen, idx = ghost_centres(mesh, o)
            rho = 1.0 + sum(g * x for g, x in zip(gr, cen))
            p = (GAMMA - 1) * (2.0 + sum(g * x for g, x in zip(ge, cen)))
            err = max(err, np.abs(q_at(Q, mesh, o, 0, idx) - rho).max(), np.abs(q_at(Q, mesh, o, 1, idx) - p).max())
        worst[prol] = err
    assert worst["EXTRAPOLATE_LINEAR_MINMOD"] <= 1e-14, worst
    assert worst["SIMPLE_COPY"] > 1e-3, worst  # the coarse-neighbour ghosts hold the coarse value


def test_extremum_gives_zero_slope():
    """rho = 1 + 0.2 cos(2 pi (x - x0)) (1 + y): a maximum along x at the coarse cell centred on
    x0 beside the refined region; the ghost cells inside that coarse cell see no x-slope, and
    their y-slope is the linear one."""
    mesh = central_mesh(2)
    lv, co = mesh.leaves()
    hc = 1.0 / 32.0
    x0 = 0.25 - 0.5 * hc  # centre of the coarse cell left of the fine region
    xs = problems._grid(mesh, 1.0, (0.0, 0.0, 0.0), "cell")

    def rho_of(x, y):
        return 1.0 + 0.2 * np.cos(2 * np.pi * (x - x0)) * (1.0 + y)

    U = np.zeros((mesh.num_total, 5, mesh.cells_per_block))
    U[:, 0] = rho_of(xs[0], xs[1])
    U[:, 1] = 2.0
    B = np.zeros((mesh.num_total, mesh.faces_per_block))
    Q = primitives_after_step(mesh, U, B, "EXTRAPOLATE_LINEAR_MINMOD")
    checked = 0
    for o in np.nonzero(lv[:mesh.num_leaves] == 3)[0]:
        cen, idx = ghost_centres(mesh, o)
        sel = (np.abs(cen[0] - x0) < 0.5 * hc) & (cen[1] > 0.25) & (cen[1] < 0.75)
        if not sel.any():
            continue
        # the coarse cell's own value, plus its y-slope: rho(x0, y_c) + 1/4 s_y (rho(x0, y_c + h) - rho(x0, y_c))
        yc = (np.floor(cen[1][sel] / hc) + 0.5) * hc
        sy = np.where((np.floor(cen[1][sel] / (hc / 2)) % 2) == 1, 1.0, -1.0)
        expect = rho_of(x0, yc) + 0.25 * sy * (rho_of(x0, yc + hc) - rho_of(x0, yc))
        got = q_at(Q, mesh, o, 0, [a[sel] for a in idx])
        assert np.abs(got - expect).max() <= 1e-14
        checked += int(sel.sum())
    assert checked >= 8


@pytest.mark.parametrize("dim", [2, 3])
def test_uniform_mesh_unaffected(dim):
    mesh = uniform_mesh(dim, 8, 1)
    if dim == 2:
        U, B = problems.orszag_tang_2d(mesh)
    else:
        U, B = problems.blast(mesh, radius=0.3)
    U = U.reshape(mesh.num_total, 5, -1)
    B = B.reshape(mesh.num_total, -1)
    out = []
    for prol in ("SIMPLE_COPY", "EXTRAPOLATE_LINEAR_MINMOD"):
        ora = OracleSolver(mesh, settings_array(gamma0=5.0 / 3.0, cfl=0.8, prolongation=prol), nthreads=4)
        out.append(ora.step(U.copy(), B.copy())[:2])
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
