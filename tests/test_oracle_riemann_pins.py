"""Pins of the oracle's HLLD solver and 2D-HLLD EMF (oracle/mhd_oracle.cpp restating
models/RiemannSolvers_MHD.h:311-551 and :629-869) by properties the mathematics fixes.

HLLD (Miyoshi & Kusano 2005):

* Rankine-Hugoniot across every wave of the fan, checked without knowing the fan: ideal MHD
  and the solver are Galilean invariant, so the solver applied to both states seen from a frame
  moving at speed s returns the flux through the ray x/t = s of the original fan, in that frame.
  Across a wave of speed S that flux is continuous at s = S (the wave is at rest there), for
  every conserved component, if and only if the states on both sides of the wave satisfy its
  jump conditions.  Each boost moves x/t = 0 through every branch of the solver (outer states,
  star states, double-star states), so a scan over s checks every branch of every state pair.
  A discontinuity is found by zooming into every kink of the flux as a function of s.
* Exact resolution of isolated discontinuities: a moving contact with Bn != 0, a tangential
  discontinuity (Bn = 0), rotational (Alfven) discontinuities travelling left and right (the
  Walen relation: dv_t = -+ sgn(Bn) dB_t / sqrt(rho)); the flux is the physical flux of the
  exact state at x/t = 0, including the double-star states.
* The outer wave speeds: seen from any frame moving slower than SL = min(u_L, u_R) -
  max(c_f,L, c_f,R), the flux is the left state's physical flux, and just faster it is not
  (likewise at SR with the right state).
* The degeneracy guard (|e*| < 1e-4 Bn^2): a state whose fast speed equals its Alfven speed
  exactly (B_t = 0, v_A > c_s) has e* == 0; the flux of that state against itself must be its
  physical flux, finite.

2D-HLLD EMF (states LL, RL, LR, RR: first letter the side in x, second in y; B_x shared
across the x-interface, B_y across the y-interface, as compute_emf builds them):

* supersonic corners: a flow faster than every wave in x and y returns the EMF of the upwind
  quadrant state;
* supersonic in y, any flow in x: Galilean boosts along x keep the EMF continuous (Faraday's
  jump condition across the x-waves), and likewise with x and y exchanged;
* a rotation of the plane by 90 degrees leaves E_z unchanged, and reflections x -> -x and
  y -> -y change its sign, on random subsonic states (every branch of the formula).
"""
import math

import numpy as np
import pytest

import oracle_ref
from oracle_ref import dptr, settings_array

GAMMA = 5.0 / 3.0
ST = settings_array(gamma0=GAMMA)


def hlld(ql, qr, st=ST):
    ql = np.ascontiguousarray(ql, dtype=np.float64)
    qr = np.ascontiguousarray(qr, dtype=np.float64)
    f = np.zeros(8)
    oracle_ref.lib().oracle_riemann_hlld(dptr(st), dptr(ql), dptr(qr), dptr(f))
    return f


def physical_flux(q, gamma=GAMMA):
    """Ideal-MHD x-flux of a primitive state (rho, p, u, v, w, Bx, By, Bz), total pressure in the
    momentum flux (the find_mhd_flux formula, models/mhd_utils.h:182-245)."""
    d, p, u, v, w, a, b, c = q
    ek = 0.5 * (u * u + v * v + w * w) * d
    em = 0.5 * (a * a + b * b + c * c)
    et = p / (gamma - 1) + ek + em
    pt = p + em
    return np.array([d * u, (et + pt) * u - a * (a * u + b * v + c * w), d * u * u - a * a + pt,
                     d * u * v - a * b, d * u * w - a * c, 0.0, b * u - a * v, c * u - a * w])


def fast_speed(q, n=0, gamma=GAMMA):
    """Textbook fast magnetosonic speed along axis n."""
    d, p = q[0], q[1]
    bb = q[5] ** 2 + q[6] ** 2 + q[7] ** 2
    cs2, va2, bn2 = gamma * p / d, bb / d, q[5 + n] ** 2 / d
    return math.sqrt(0.5 * (cs2 + va2 + math.sqrt((cs2 + va2) ** 2 - 4.0 * cs2 * bn2)))


def boost(q, s, axis=2):
    """A primitive state seen from a frame moving at speed s along the velocity component axis."""
    q = np.array(q, dtype=np.float64)
    q[axis] -= s
    return q


# components whose boosted flux is linear in s within a region of the fan (mass, transverse
# momentum, transverse field): their kinks locate the waves
LINEAR = [0, 3, 4, 6, 7]


def largest_jump(f, lo, hi, linear, n=2001, levels=4):
    """Largest discontinuity of a piecewise-smooth vector function f on [lo, hi], relative to
    each component's scale: every kink of the components listed in `linear` (zero second
    differences elsewhere) is zoomed into `levels` times; at the finest spacing the increment of
    every component across the kink is its jump (a continuous function's increment has shrunk
    below 1e-10 of its scale)."""
    s = np.linspace(lo, hi, n)
    g = np.array([f(x) for x in s])
    scale = np.abs(g).max(axis=0)
    keep = scale > 0
    scale = np.where(keep, scale, 1.0)
    worst = 0.0
    lin = [c for c in linear if keep[c]]

    def d2(gg):
        return (np.abs(np.diff(gg[:, lin], 2, axis=0)) / scale[lin]).max(axis=1)

    def zoom(a, b, level):
        nonlocal worst
        ss = np.linspace(a, b, n)
        gg = np.array([f(x) for x in ss])
        if level == levels:
            worst = max(worst, float((np.abs(np.diff(gg[:, keep], axis=0)) / scale[keep]).max()))
            return
        i = int(np.argmax(d2(gg)))
        zoom(ss[i], ss[i + 2], level + 1)

    cand = np.nonzero(d2(g) > 1e-11)[0]
    if len(cand) == 0:
        return 0.0, 0
    clusters = np.split(cand, np.nonzero(np.diff(cand) > 1)[0] + 1)
    for c in clusters:
        zoom(s[c[0]], s[c[-1] + 2], 1)
    return worst, len(clusters)


def random_pair(rng):
    def one():
        return np.array([rng.uniform(0.4, 2.5), rng.uniform(0.1, 2.0), rng.normal(0, 0.6), rng.normal(0, 0.6),
                         rng.normal(0, 0.6), 0.0, rng.normal(0, 0.8), rng.normal(0, 0.8)])
    ql, qr = one(), one()
    ql[5] = qr[5] = rng.choice([-1.0, 1.0]) * rng.uniform(0.2, 1.2)
    return ql, qr


def outer_speeds(ql, qr, axis=0):
    cf = max(fast_speed(ql, axis), fast_speed(qr, axis))
    return min(ql[2 + axis], qr[2 + axis]) - cf, max(ql[2 + axis], qr[2 + axis]) + cf


# ------------------------------------------------------------------------------------------
# HLLD
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("seed", range(12))
def test_hlld_jump_conditions_across_every_wave(seed):
    """The boosted flux is continuous in the boost speed for all seven fluxed components: the
    states of every region satisfy the jump conditions of the waves that bound it."""
    ql, qr = random_pair(np.random.default_rng(seed))
    sl, sr = outer_speeds(ql, qr)
    w = sr - sl
    jump, nkinks = largest_jump(lambda s: hlld(boost(ql, s), boost(qr, s)), sl - 0.2 * w, sr + 0.2 * w, LINEAR)
    # the fan has five waves (SL, SL*, SM, SR*, SR): each is a kink of some linear component
    assert nkinks == 5, f"{nkinks} kinks found in the fan"
    assert jump <= 1e-9, f"flux jumps by {jump:.3e} of its scale across a wave"


def test_hlld_outer_speeds():
    """Below SL = min(u) - max(c_f) the flux is the left state's physical flux; just above SL it
    is not (and symmetrically at SR)."""
    rng = np.random.default_rng(42)
    for _ in range(6):
        ql, qr = random_pair(rng)
        sl, sr = outer_speeds(ql, qr)
        eps = 1e-7 * (sr - sl)
        below = hlld(boost(ql, sl - eps), boost(qr, sl - eps))
        np.testing.assert_allclose(below, physical_flux(boost(ql, sl - eps)), rtol=1e-12, atol=1e-13)
        above = hlld(boost(ql, sl + eps), boost(qr, sl + eps))
        assert np.abs(above - physical_flux(boost(ql, sl + eps))).max() > 1e-9
        beyond = hlld(boost(ql, sr + eps), boost(qr, sr + eps))
        np.testing.assert_allclose(beyond, physical_flux(boost(qr, sr + eps)), rtol=1e-12, atol=1e-13)
        inside = hlld(boost(ql, sr - eps), boost(qr, sr - eps))
        assert np.abs(inside - physical_flux(boost(qr, sr - eps))).max() > 1e-9


def rotational(rho, p, u, a, bt, th_l, th_r, vt_l, going):
    """Left and right states of an isolated rotational discontinuity: rho, p, u, Bn = a and
    |B_t| = bt continuous, B_t rotated from angle th_l to th_r, and the transverse velocity
    jump of the Walen relation for a wave travelling left (-1, speed u - |a|/sqrt(rho)) or
    right (+1, speed u + |a|/sqrt(rho))."""
    bl = np.array([bt * math.cos(th_l), bt * math.sin(th_l)])
    br = np.array([bt * math.cos(th_r), bt * math.sin(th_r)])
    vl = np.array(vt_l, dtype=np.float64)
    vr = vl - going * math.copysign(1.0, a) * (br - bl) / math.sqrt(rho)
    ql = np.array([rho, p, u, vl[0], vl[1], a, bl[0], bl[1]])
    qr = np.array([rho, p, u, vr[0], vr[1], a, br[0], br[1]])
    return ql, qr, u + going * abs(a) / math.sqrt(rho)


@pytest.mark.parametrize("going", [-1, 1])
@pytest.mark.parametrize("u", [-0.9, -0.3, 0.3, 0.9])
@pytest.mark.parametrize("a", [0.7, -0.7])
def test_hlld_isolated_rotational_discontinuity_is_exact(going, u, a):
    """The flux is the physical flux of the exact state at x/t = 0: the left state if the wave
    moves right of the interface, else the right one.  With |a|/sqrt(rho) = 0.7 the wave speed
    u -+ 0.7 takes every sign relative to u, so x/t = 0 falls in each of the double-star
    regions (between the Alfven wave and the contact) and outside them."""
    ql, qr, speed = rotational(1.3, 0.8, u, a * math.sqrt(1.3), 0.9, 0.4, 2.3, (0.2, -0.1), going)
    f = hlld(ql, qr)
    exact = physical_flux(ql if speed > 0 else qr)
    np.testing.assert_allclose(f, exact, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("u", [-0.6, 0.6])
def test_hlld_moving_contact_is_exact(u):
    """Only the density jumps (Bn != 0, v, B, p continuous): the flux is the upwind state's."""
    ql = np.array([1.7, 0.9, u, 0.3, -0.2, 0.6, -0.4, 0.5])
    qr = ql.copy()
    qr[0] = 0.45
    np.testing.assert_allclose(hlld(ql, qr), physical_flux(ql if u > 0 else qr), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("u", [-0.4, 0.4])
def test_hlld_tangential_discontinuity_is_exact(u):
    """Bn = 0: density, pressure, transverse velocity and field jump at equal total pressure and
    move with the flow; the flux is the upwind state's."""
    ql = np.array([1.5, 1.2, u, 0.5, -0.3, 0.0, 0.8, -0.6])
    pt = ql[1] + 0.5 * (0.8 ** 2 + 0.6 ** 2)
    qr = np.array([0.6, 0.0, u, -0.7, 0.4, 0.0, -0.3, 1.1])
    qr[1] = pt - 0.5 * (0.3 ** 2 + 1.1 ** 2)
    np.testing.assert_allclose(hlld(ql, qr), physical_flux(ql if u > 0 else qr), rtol=1e-12, atol=1e-13)


def test_hlld_degeneracy_guard():
    """gamma = 2, rho = 1, p = 1/2, Bn = 2, B_t = 0: c_s = 1 < v_A = 2, the fast speed is 2
    exactly and e* = rho (S_L - u)(S_L - u*) - Bn^2 == 0.  The star states must fall back to the
    outer transverse state (0/0 otherwise), so the flux of the state against itself is its
    physical flux; a Riemann problem along the field keeps every transverse flux at zero."""
    st = settings_array(gamma0=2.0)
    q = np.array([1.0, 0.5, 0.0, 0.0, 0.0, 2.0, 0.0, 0.0])
    assert fast_speed(q, 0, gamma=2.0) == 2.0
    f = hlld(q, q, st)
    assert np.all(np.isfinite(f))
    np.testing.assert_allclose(f, physical_flux(q, gamma=2.0), rtol=1e-14, atol=0)
    qr = np.array([0.5, 0.2, 0.0, 0.0, 0.0, 2.0, 0.0, 0.0])
    f = hlld(q, qr, st)
    assert np.all(np.isfinite(f)) and f[3] == 0 and f[4] == 0 and f[6] == 0 and f[7] == 0


# ------------------------------------------------------------------------------------------
# 2D-HLLD EMF
# ------------------------------------------------------------------------------------------
IRT, IRB, ILT, ILB = 0, 1, 2, 3


def emf_llrr(LL, RL, LR, RR):
    """E_z of mag_riemann2d_hlld from the four quadrant states, through compute_emf<EMFZ> (which
    maps its corner states RT, LT, RB, LB to LL, RL, LR, RR and averages the shared field
    components, exact here since they are equal)."""
    qe = np.zeros((4, 8))
    qe[IRT], qe[ILT], qe[IRB], qe[ILB] = LL, RL, LR, RR
    return oracle_ref.lib().oracle_compute_emf(dptr(ST), dptr(np.ascontiguousarray(qe)), 2)


def quad_states(rng, vx=0.0, vy=0.0, spread=0.3):
    S = [np.array([rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0), vx + rng.normal(0, spread),
                   vy + rng.normal(0, spread), rng.normal(0, 0.3), 0.0, 0.0, rng.normal(0, 0.5)])
         for _ in range(4)]
    LL, RL, LR, RR = S
    LL[5] = RL[5] = rng.normal(0, 0.7)   # B_x continuous across the x-interface, bottom row
    LR[5] = RR[5] = rng.normal(0, 0.7)   # top row
    LL[6] = LR[6] = rng.normal(0, 0.7)   # B_y continuous across the y-interface, left column
    RL[6] = RR[6] = rng.normal(0, 0.7)   # right column
    return [LL, RL, LR, RR]


def ez(q):
    return q[2] * q[6] - q[3] * q[5]


@pytest.mark.parametrize("sx,sy,which", [(1, 1, 0), (-1, 1, 1), (1, -1, 2), (-1, -1, 3)])
def test_emf_supersonic_corners(sx, sy, which):
    """A flow faster than every wave in x and in y: the EMF of the upwind quadrant state
    (LL for +x,+y; RL for -x,+y; LR for +x,-y; RR for -x,-y)."""
    rng = np.random.default_rng(10 + which)
    for _ in range(5):
        st = quad_states(rng, vx=8.0 * sx, vy=8.0 * sy)
        assert abs(emf_llrr(*st) - ez(st[which])) <= 1e-13 * max(abs(ez(st[which])), 1.0)


def _boosted(states, vx, vy):
    out = []
    for q in states:
        q = q.copy()
        q[2] -= vx
        q[3] -= vy
        out.append(q)
    return out


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("case", ["x_up", "x_down", "y_right", "y_left"])
def test_emf_jump_conditions_of_the_one_dimensional_branches(case, seed):
    """Supersonic along one axis (SB > 0, ST < 0, SL > 0 or SR < 0), the EMF is a 1D problem
    along the other; boosting along that axis, E_z in the moving frame stays continuous across
    its waves (Faraday: the tangential electric field is continuous in a wave's rest frame)."""
    rng = np.random.default_rng(100 + seed)
    vy = {"x_up": 8.0, "x_down": -8.0}.get(case, 0.0)
    vx = {"y_right": 8.0, "y_left": -8.0}.get(case, 0.0)
    st = quad_states(rng, vx=vx, vy=vy)
    along_x = case.startswith("x")
    n = 0 if along_x else 1
    u = [q[2 + n] for q in st]
    cf = max(fast_speed(q, n) for q in st)
    lo, hi = min(u) - cf, max(u) + cf
    w = hi - lo
    f = (lambda s: np.array([emf_llrr(*_boosted(st, s, 0.0))])) if along_x else \
        (lambda s: np.array([emf_llrr(*_boosted(st, 0.0, s))]))
    jump, nk = largest_jump(f, lo - 0.2 * w, hi + 0.2 * w, [0])
    assert nk >= 2, f"only {nk} kinks: the outer waves were not crossed"
    assert jump <= 1e-9, f"EMF jumps by {jump:.3e} of its scale across a wave"


def rotate90(states):
    """(x, y) -> (y, -x): the quadrant states move (LL' = RL, RL' = RR, LR' = LL, RR' = LR) and
    (u, v) -> (v, -u), (Bx, By) -> (By, -Bx)."""
    LL, RL, LR, RR = states

    def r(q):
        q2 = q.copy()
        q2[2], q2[3] = q[3], -q[2]
        q2[5], q2[6] = q[6], -q[5]
        return q2

    return [r(RL), r(RR), r(LL), r(LR)]


def reflect(states, axis):
    """x -> -x (axis 0: L and R swap in x, u and Bx change sign) or y -> -y (axis 1)."""
    LL, RL, LR, RR = states

    def m(q):
        q2 = q.copy()
        q2[2 + axis] = -q2[2 + axis]
        q2[5 + axis] = -q2[5 + axis]
        return q2

    if axis == 0:
        return [m(RL), m(LL), m(RR), m(LR)]
    return [m(LR), m(RR), m(LL), m(RL)]


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("flow", [(0.0, 0.0), (8.0, 0.0), (0.0, -8.0), (1.5, 1.0)])
def test_emf_rotation_and_reflections(seed, flow):
    rng = np.random.default_rng(1000 + seed)
    st = quad_states(rng, vx=flow[0], vy=flow[1], spread=0.5)
    e = emf_llrr(*st)
    scale = max(max(abs(q[2]) + abs(q[3]) for q in st) * max(abs(q[5]) + abs(q[6]) for q in st), 1e-300)
    r = st
    for k in range(4):
        r = rotate90(r)
        assert abs(emf_llrr(*r) - e) <= 1e-13 * scale, f"rotation {90 * (k + 1)} deg"
    assert abs(emf_llrr(*reflect(st, 0)) + e) <= 1e-13 * scale
    assert abs(emf_llrr(*reflect(st, 1)) + e) <= 1e-13 * scale
