"""Pin the CPU oracle's point functions by closed forms, invariants and special cases.

The reference has no tests or golden vectors for this path, so each check is chosen so that a
plausible slip (dropped term, wrong sign, transposed operand) fails it.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_ref
from oracle_ref import dptr, settings_array

GAMMA = 5.0 / 3.0
ST = settings_array(gamma0=GAMMA)


def prim(u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    q = np.zeros(8)
    oracle_ref.lib().oracle_compute_primitives(dptr(ST), dptr(u), dptr(q))
    return q


def hlld(ql, qr):
    ql = np.ascontiguousarray(ql, dtype=np.float64)
    qr = np.ascontiguousarray(qr, dtype=np.float64)
    f = np.zeros(8)
    oracle_ref.lib().oracle_riemann_hlld(dptr(ST), dptr(ql), dptr(qr), dptr(f))
    return f


def emf(qe, d=2):
    qe = np.ascontiguousarray(qe, dtype=np.float64)
    return oracle_ref.lib().oracle_compute_emf(dptr(ST), dptr(qe), d)


def physical_flux(q, gamma=GAMMA):
    """Ideal-MHD x-flux of a primitive state, total pressure in the momentum flux (the
    find_mhd_flux formula, models/mhd_utils.h:182-245)."""
    d, p, u, v, w, a, b, c = q
    ek = 0.5 * (u * u + v * v + w * w) * d
    em = 0.5 * (a * a + b * b + c * c)
    et = p / (gamma - 1) + ek + em
    pt = p + em
    return np.array([d * u, (et + pt) * u - a * (a * u + b * v + c * w), d * u * u - a * a + pt,
                     d * u * v - a * b, d * u * w - a * c, 0.0, b * u - a * v, c * u - a * w])


def test_primitives_closed_form():
    rho, vx, vy, vz, p = 1.3, 0.4, -0.7, 0.2, 2.5
    bx, by, bz = 0.3, -0.8, 1.1
    E = p / (GAMMA - 1) + 0.5 * rho * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx * bx + by * by + bz * bz)
    q = prim([rho, E, rho * vx, rho * vy, rho * vz, bx, by, bz])
    assert q[0] == rho
    np.testing.assert_allclose(q[2:5], [vx, vy, vz], rtol=1e-15)
    np.testing.assert_allclose(q[1], p, rtol=1e-13)
    assert list(q[5:]) == [bx, by, bz]


def test_primitives_floors():
    # negative internal energy -> p = rho * smallp; tiny density -> smallr
    q = prim([1.0, 0.1, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0])
    smallp = 1e-10 * 1e-10 / GAMMA
    assert q[1] == 1.0 * smallp
    q = prim([1e-20, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0])
    assert q[0] == 1e-10


@pytest.mark.parametrize("st,expect", [(1.0, 1.0), (2.0, 1.5)])
def test_slope_limiters_monotone(st, expect):
    assert oracle_ref.lib().oracle_slope(2.0, 4.0, 1.0, st) == expect   # dl=st*1, dr=st*2, dc=1.5
    assert oracle_ref.lib().oracle_slope(2.0, 1.0, 4.0, st) == -expect  # mirrored


@pytest.mark.parametrize("st", [1.0, 2.0])
def test_slope_limiters_extremum_and_flat(st):
    assert oracle_ref.lib().oracle_slope(2.0, 1.0, 1.0, st) == 0.0
    assert oracle_ref.lib().oracle_slope(2.0, 3.0, 3.0, st) == 0.0
    assert oracle_ref.lib().oracle_slope(2.0, 2.0, 2.0, st) == 0.0


def test_slope_minmod_is_min_of_one_sided():
    rng = np.random.default_rng(1)
    for _ in range(200):
        qm, q, qp = np.sort(rng.normal(size=3))
        s = oracle_ref.lib().oracle_slope(q, qp, qm, 1.0)
        assert s == min(q - qm, qp - q, 0.5 * (qp - qm))


STATES = [
    [1.0, 1.0, 0.0, 0.0, 0.0, 0.75, 1.0, 0.0],
    [0.125, 0.1, 0.0, 0.0, 0.0, 0.75, -1.0, 0.0],
    [1.08, 0.95, 1.2, 0.01, 0.5, 2.0 / math.sqrt(4 * math.pi), 3.6 / math.sqrt(4 * math.pi), 2 / math.sqrt(4 * math.pi)],
    [0.9, 2.0, -0.3, 0.4, -0.2, -0.6, 0.2, 1.5],
]


@pytest.mark.parametrize("q", STATES)
def test_hlld_consistency(q):
    """F(q, q) is the physical flux."""
    f = hlld(q, q)
    ref = physical_flux(q)
    np.testing.assert_allclose(f, ref, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("i", range(len(STATES)))
def test_hlld_mirror_symmetry(i):
    """x -> -x: swap sides, negate u and Bx; the fluxes of the even quantities (mass, energy,
    transverse momentum and field) change sign, the normal-momentum flux does not."""
    rng = np.random.default_rng(i)
    ql = np.array(STATES[i])
    qr = ql * (1 + 0.3 * rng.random(8))
    qr[5] = ql[5]

    def mirror(q):
        m = q.copy()
        m[2] = -m[2]
        m[5] = -m[5]
        return m

    f = hlld(ql, qr)
    g = hlld(mirror(qr), mirror(ql))
    np.testing.assert_allclose(g[[0, 1, 3, 4]], -f[[0, 1, 3, 4]], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(g[2], f[2], rtol=1e-12)
    np.testing.assert_allclose(g[[6, 7]], -f[[6, 7]], rtol=1e-12, atol=1e-14)


def test_hlld_stationary_contact():
    """u = 0, Bn != 0, equal total pressure, density jump: no mass flux."""
    ql = [1.0, 1.0, 0.0, 0.0, 0.0, 0.5, 0.3, 0.1]
    qr = [0.3, 1.0, 0.0, 0.0, 0.0, 0.5, 0.3, 0.1]
    f = hlld(ql, qr)
    assert abs(f[0]) < 1e-15
    np.testing.assert_allclose(f[2], 1.0 + 0.5 * (0.25 + 0.09 + 0.01) - 0.25, rtol=1e-14)


def test_hlld_supersonic_upwind():
    """Flow faster than every wave: the flux is the left state's physical flux."""
    ql = [1.0, 0.1, 20.0, 0.1, 0.0, 0.2, 0.1, 0.0]
    qr = [0.5, 0.05, 20.0, -0.1, 0.0, 0.2, 0.3, 0.0]
    np.testing.assert_allclose(hlld(ql, qr), physical_flux(ql), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_emf_uniform_state(d):
    """Four equal corner states: E equals the ideal-MHD electric field of that state for the
    edge direction (u x B component)."""
    q = np.array([1.2, 0.9, 0.3, -0.4, 0.25, 0.6, -0.2, 0.45])
    qe = np.tile(q, (4, 1))
    u, v, w, a, b, c = q[2:]
    expect = {2: u * b - v * a, 1: w * a - u * c, 0: v * c - w * b}[d]
    np.testing.assert_allclose(emf(qe, d), expect, rtol=1e-13, atol=1e-16)


def test_emf_linear_in_field_sign():
    """Flipping every B component flips E (the 2D HLLD EMF is odd in B)."""
    rng = np.random.default_rng(5)
    qe = np.array([STATES[2]] * 4) * (1 + 0.1 * rng.random((4, 8)))
    qf = qe.copy()
    qf[:, 5:] *= -1
    np.testing.assert_allclose(emf(qf), -emf(qe), rtol=1e-13)


def test_speed_info_closed_form():
    q = np.array([1.3, 2.0, 0.4, -0.5, 0.1, 0.3, 0.9, -0.2])
    out = np.zeros(3)
    oracle_ref.lib().oracle_speed_info(dptr(ST), dptr(q), 3, dptr(out))
    d, p, u, v, w, a, b, c = q
    cs2 = GAMMA * p / d
    va2 = (a * a + b * b + c * c) / d
    for k, (vel, bn) in enumerate([(u, a), (v, b), (w, c)]):
        cf = math.sqrt(0.5 * (cs2 + va2 + math.sqrt((cs2 + va2) ** 2 - 4 * cs2 * bn * bn / d)))
        assert abs(out[k] - (cf + abs(vel))) < 1e-14
