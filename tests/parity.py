"""The GPU-vs-oracle parity bar of the north star: relative L-inf <= 1e-12 on density, momentum,
energy and face B.  Momentum is measured as a vector: the largest change of any component over
the largest momentum component (a component that starts at zero, e.g. m_z in a 2D-like flow,
is not divided by its own tiny norm).  The per-component figures are reported beside it."""
import numpy as np

TOL = 1e-12


def rel_linf(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def parity_errors(Ug, Bg, Uo, Bo, n):
    Ug, Uo = np.asarray(Ug)[:n], np.asarray(Uo)[:n]
    Bg, Bo = np.asarray(Bg)[:n], np.asarray(Bo)[:n]
    errs = {"rho": rel_linf(Ug[:, 0], Uo[:, 0]), "E": rel_linf(Ug[:, 1], Uo[:, 1]),
            "momentum": rel_linf(Ug[:, 2:5], Uo[:, 2:5]), "B": rel_linf(Bg, Bo)}
    per = {f"m{c}": rel_linf(Ug[:, 2 + i], Uo[:, 2 + i]) for i, c in enumerate("xyz")}
    same = bool(np.array_equal(Ug, Uo) and np.array_equal(Bg, Bo))
    return errs, per, same


def assert_parity(Ug, Bg, Uo, Bo, n, tag=""):
    """Assert the bar; returns (errors, per-component momentum errors, bitwise equal)."""
    errs, per, same = parity_errors(Ug, Bg, Uo, Bo, n)
    bad = {k: e for k, e in errs.items() if not e <= TOL}
    assert not bad, f"{tag}: relative L-inf above {TOL}: {bad} (all: {errs}, per component: {per})"
    return errs, per, same
