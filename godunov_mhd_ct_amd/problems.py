"""Initial states of the benchmark problems, built on the host in the ABI layout.

Problem setup is not on the device hot path; these functions only produce the inputs
(U[block][var][cell], B[block][face]) the time step consumes.  Definitions follow
init/InitOrszagTang.cpp:84-399 and init/InitBlast.cpp:104-335 where the reference has them.
"""
from __future__ import annotations

import math

import numpy as np

from .mesh import Mesh


def _grid(mesh: Mesh, scaling: float, xmin, kind: str, comp: int = 0):
    """Physical coordinates of cell centres (kind 'cell'), of cells' lowest corners (kind
    'vertex') or of face centres of component `comp` (kind 'face') for every owned leaf and
    outside block, in ABI order."""
    lv, co = mesh.leaves()
    b, dim = mesh.b, mesh.dim
    n = len(lv)
    if kind in ("cell", "vertex"):
        shape = (b,) * dim
        off = [0.5 if kind == "cell" else 0.0] * 3
    else:
        if dim == 2 and comp == 2:
            shape = (b, b)
            off = [0.5, 0.5, 0.5]
        else:
            shape = [b] * dim
            shape[comp] = b + 1
            shape = tuple(shape)
            off = [0.5] * 3
            off[comp] = 0.0
    # index grids, x fastest
    idx = np.indices(shape[::-1])[::-1]  # idx[0] = i (fastest) ...
    out = []
    for d in range(dim):
        rel = idx[d].reshape(-1).astype(np.float64) + off[d]
        scale = (1 << lv).astype(np.float64)  # blocks per tree edge
        pos = (co[:, d:d + 1].astype(np.float64) * b + rel[None, :]) / (b * scale[:, None])
        out.append(xmin[d] + scaling * pos)
    return out  # list of [n, ncell]


def total_energy(rho, mom, p, bcc, gamma):
    ek = 0.5 * sum(m * m for m in mom) / rho
    em = 0.5 * sum(x * x for x in bcc)
    return p / (gamma - 1.0) + ek + em


def cell_centred_b(mesh: Mesh, B):
    """Cell-centred field from the face array (mean of the two faces; 2D Bz is cell centred)."""
    b, dim = mesh.b, mesh.dim
    n = B.shape[0]
    if dim == 2:
        nx = (b + 1) * b
        bx = B[:, :nx].reshape(n, b, b + 1)
        by = B[:, nx:2 * nx].reshape(n, b + 1, b)
        bz = B[:, 2 * nx:].reshape(n, b, b)
        a = 0.5 * (bx[:, :, :-1] + bx[:, :, 1:])
        bb = 0.5 * (by[:, :-1, :] + by[:, 1:, :])
        return [a.reshape(n, -1), bb.reshape(n, -1), bz.reshape(n, -1)]
    nf = (b + 1) * b * b
    bx = B[:, :nf].reshape(n, b, b, b + 1)
    by = B[:, nf:2 * nf].reshape(n, b, b + 1, b)
    bz = B[:, 2 * nf:].reshape(n, b + 1, b, b)
    a = 0.5 * (bx[..., :-1] + bx[..., 1:])
    bb = 0.5 * (by[:, :, :-1, :] + by[:, :, 1:, :])
    c = 0.5 * (bz[:, :-1] + bz[:, 1:])
    return [a.reshape(n, -1), bb.reshape(n, -1), c.reshape(n, -1)]


def _assemble(mesh, rho, mom, p, Bparts, gamma):
    n = mesh.num_total
    B = np.concatenate([x.reshape(n, -1) for x in Bparts], axis=1)
    bcc = cell_centred_b(mesh, B)
    E = total_energy(rho, mom, p, bcc, gamma)
    U = np.stack([rho, E, mom[0], mom[1], mom[2]], axis=1)
    return np.ascontiguousarray(U), np.ascontiguousarray(B)


def orszag_tang_2d(mesh: Mesh, gamma=5.0 / 3.0, scaling=1.0, xmin=(0.0, 0.0, 0.0)):
    """Orszag-Tang vortex (init/InitOrszagTang.cpp:102-105, 193, 220-225, 355-383):
    rho = gamma*p0, p0 = gamma/(4 pi), v = (-sin 2 pi y, sin 2 pi x), face B set pointwise:
    Bx = -B0 sin(2 pi y) on x-faces, By = B0 sin(4 pi x) on y-faces, B0 = 1/sqrt(4 pi)."""
    assert mesh.dim == 2
    two_pi = 2.0 * math.pi
    p0 = gamma / (2.0 * two_pi)
    d0 = gamma * p0
    B0 = 1.0 / math.sqrt(2.0 * two_pi)
    x, y = _grid(mesh, scaling, xmin, "cell")
    rho = np.full_like(x, d0)
    mom = [-d0 * np.sin(y * two_pi), d0 * np.sin(x * two_pi), np.zeros_like(x)]
    p = np.full_like(x, p0)
    fx, fy = _grid(mesh, scaling, xmin, "face", 0)
    gx, gy = _grid(mesh, scaling, xmin, "face", 1)
    Bx = -B0 * np.sin(fy * two_pi)
    By = B0 * np.sin(2 * gx * two_pi)
    Bz = np.zeros((mesh.num_total, mesh.b * mesh.b))
    return _assemble(mesh, rho, mom, p, [Bx, By, Bz], gamma)


def blast(mesh: Mesh, gamma=5.0 / 3.0, radius=0.1, center=(0.5, 0.5, 0.5), rho_in=1.0, rho_out=1.0,
          p_in=10.0, p_out=0.1, bfield=(1 / math.sqrt(3),) * 3, scaling=1.0, xmin=(0.0, 0.0, 0.0),
          test_point="centre"):
    """MHD blast (init/InitBlast.cpp:104-335): inside the disc/sphere when the cell's test point
    lies strictly within the radius; the test point is the cell centre, or with
    test_point="vertex" the cell's lowest corner, as the reference tests it
    (init/InitBlast.cpp:141-146, 162-168)."""
    if test_point not in ("centre", "vertex"):
        raise ValueError("test_point is 'centre' or 'vertex'")
    xs = _grid(mesh, scaling, xmin, "cell" if test_point == "centre" else "vertex")
    d2 = sum((xs[d] - center[d]) ** 2 for d in range(mesh.dim))
    inside = d2 < radius * radius
    rho = np.where(inside, rho_in, rho_out)
    p = np.where(inside, p_in, p_out)
    mom = [np.zeros_like(rho) for _ in range(3)]
    n = mesh.num_total
    b = mesh.b
    parts = []
    for c in range(3):
        if mesh.dim == 2 and c == 2:
            size = b * b
        elif mesh.dim == 2:
            size = (b + 1) * b
        else:
            size = (b + 1) * b * b
        parts.append(np.full((n, size), bfield[c]))
    return _assemble(mesh, rho, mom, p, parts, gamma)


def kelvin_helmholtz_3d(mesh: Mesh, gamma=1.4, seed=0, scaling=1.0, xmin=(0.0, 0.0, 0.0), bx=0.1):
    """3D Kelvin-Helmholtz: rho = 2 for |y-0.5| < 0.25 else 1, vx = +-0.5 across the layers,
    vy = 0.01 sin(4 pi x) + 1e-3 U(-1/2, 1/2) (seeded), p = 2.5, Bx = 0.1."""
    rng = np.random.default_rng(seed)
    x, y, z = _grid(mesh, scaling, xmin, "cell")
    inner = np.abs(y - 0.5) < 0.25
    rho = np.where(inner, 2.0, 1.0)
    vx = np.where(inner, 0.5, -0.5)
    vy = 0.01 * np.sin(4 * math.pi * x) + 1e-3 * (rng.random(x.shape) - 0.5)
    vz = np.zeros_like(x)
    p = np.full_like(x, 2.5)
    mom = [rho * vx, rho * vy, rho * vz]
    n = mesh.num_total
    b = mesh.b
    nf = (b + 1) * b * b
    parts = [np.full((n, nf), bx), np.zeros((n, nf)), np.zeros((n, nf))]
    return _assemble(mesh, rho, mom, p, parts, gamma)


def field_loop_2d(mesh: Mesh, gamma=5.0 / 3.0, scaling=1.0, xmin=(-1.0, -0.5, 0.0), amp=1e-3, r0=0.3,
                  v=(2.0, 1.0)):
    """Field-loop advection: A_z = amp (r0 - r) for r < r0, rho = p = 1, v = (2, 1); face B is
    the discrete curl of A_z sampled at the cell corners (divergence free to roundoff)."""
    assert mesh.dim == 2
    lv, co = mesh.leaves()
    b = mesh.b
    n = len(lv)
    # corner coordinates (b+1)^2 per block
    idx = np.indices((b + 1, b + 1))[::-1]
    h = scaling / (b * (1 << lv).astype(np.float64))
    cx = xmin[0] + (co[:, 0:1] * b + idx[0].reshape(1, -1)) * h[:, None]
    cy = xmin[1] + (co[:, 1:2] * b + idx[1].reshape(1, -1)) * h[:, None]
    r = np.sqrt(cx * cx + cy * cy)
    Az = np.where(r < r0, amp * (r0 - r), 0.0).reshape(n, b + 1, b + 1)
    Bx = (Az[:, 1:, :] - Az[:, :-1, :]) / h[:, None, None]      # (n, b, b+1): x-faces
    By = -(Az[:, :, 1:] - Az[:, :, :-1]) / h[:, None, None]     # (n, b+1, b): y-faces
    x, y = _grid(mesh, scaling, xmin, "cell")
    rho = np.ones_like(x)
    p = np.ones_like(x)
    mom = [rho * v[0], rho * v[1], np.zeros_like(x)]
    return _assemble(mesh, rho, mom, p, [Bx.reshape(n, -1), By.reshape(n, -1), np.zeros((n, b * b))], gamma)


def rayleigh_taylor(mesh: Mesh, gravity, gamma=1.4, rho_up=2.0, rho_down=1.0, P0=10.0, amplitude=0.01,
                    nmodes=(1, 1, 1), bfield=(0.0, 0.5, 0.0), scaling=1.0, xmin=(0.0, 0.0, 0.0)):
    """Magnetic Rayleigh-Taylor (init/InitRayleighTaylor.cpp, sine perturbation, not isothermal):
    heavy fluid above the mid-plane normal to the last axis, hydrostatic P = P0 + rho g.x, a
    velocity perturbation along that axis, uniform B."""
    dim = mesh.dim
    xs = _grid(mesh, scaling, xmin, "cell")
    L = [mesh.nbrick[d] * scaling for d in range(3)]
    up = xs[dim - 1] > xmin[dim - 1] + L[dim - 1] / 2
    rho = np.where(up, rho_up, rho_down)
    P = P0 + rho * sum(gravity[d] * xs[d] for d in range(dim))
    pert = amplitude * rho
    for d in range(dim):
        pert = pert * (1 + np.cos(2 * math.pi * xs[d] * nmodes[d] / L[d]))
    pert = pert / (4.0 if dim == 2 else 8.0)
    mom = [np.zeros_like(rho) for _ in range(3)]
    mom[dim - 1] = pert
    n = mesh.num_total
    b = mesh.b
    parts = []
    for c in range(3):
        size = (b * b if c == 2 else (b + 1) * b) if dim == 2 else (b + 1) * b * b
        parts.append(np.full((n, size), bfield[c]))
    return _assemble(mesh, rho, mom, P, parts, gamma)


# ---------------------------------------------------------------------------------------------
# random turbulent-like state (the batching workload); evaluated with torch so that large
# meshes are set up in seconds (on the GPU when there is one)
# ---------------------------------------------------------------------------------------------
def _modes(rng, kmax, n, slope):
    """n random integer wavevectors with 1 <= |k| <= kmax, amplitudes ~ |k|^slope (unit RMS sum),
    random phases."""
    ks = []
    while len(ks) < n:
        k = rng.integers(-kmax, kmax + 1, size=3)
        kk = math.sqrt(float((k * k).sum()))
        if 1.0 <= kk <= kmax:
            ks.append(k)
    ks = np.array(ks, dtype=np.float64)
    amp = np.linalg.norm(ks, axis=1) ** slope
    amp = amp / math.sqrt(0.5 * (amp ** 2).sum())
    phase = rng.uniform(0.0, 2.0 * math.pi, size=n)
    return ks, amp * rng.standard_normal(n), phase


def _series(torch, ks, amp, phase, x, y, z):
    out = torch.zeros_like(x)
    for k, a, ph in zip(ks, amp, phase):
        out += a * torch.cos(2.0 * math.pi * (k[0] * x + k[1] * y + k[2] * z) + ph)
    return out


def _line_integral(torch, ks, amp, phase, p0, axis, h):
    """Exact integral of sum_k a cos(2 pi k.x + ph) along `axis` from p0 to p0 + h e_axis."""
    x, y, z = p0
    out = torch.zeros_like(x)
    for k, a, ph in zip(ks, amp, phase):
        arg = 2.0 * math.pi * (k[0] * x + k[1] * y + k[2] * z) + ph
        ka = k[axis]
        if ka == 0:
            out += a * torch.cos(arg) * h
        else:
            w = 2.0 * math.pi * ka
            out += a * (torch.sin(arg + w * h) - torch.sin(arg)) / w
    return out


def random_mhd_3d(mesh: Mesh, seed=0, gamma=5.0 / 3.0, nmodes=16, kmax=8, device=None):
    """Smooth random MHD state on a periodic [0,1]^3 mesh: rho = exp(0.3 g1), p = max(1 + 0.1 g2,
    0.05), v = 0.5 (g3, g4, g5), with g_i unit-variance random Fourier fields (|k| <= 4, a
    Gaussian-smoothed white-noise spectrum), and face B = curl A from the exact edge line
    integrals of a random Fourier vector potential (|k| <= kmax, k^-2 spectrum, amplitude 0.5),
    so the discrete div B of every cell vanishes to roundoff."""
    import torch

    assert mesh.dim == 3
    dev = device or ("cuda" if torch.cuda.is_available() else "cpu")
    rng = np.random.default_rng(seed)
    lv, co = mesh.leaves()
    b, n = mesh.b, mesh.num_total
    h = 1.0 / (b * (1 << lv).astype(np.float64))   # cell size per block
    T = lambda a: torch.as_tensor(a, dtype=torch.float64, device=dev)  # noqa: E731
    hb = T(h)[:, None]
    ob = [T(co[:, d].astype(np.float64) * b * h)[:, None] for d in range(3)]  # block origins

    def pts(shape, off):
        idx = np.indices(shape[::-1])[::-1]
        return [ob[d] + (T(idx[d].reshape(1, -1).astype(np.float64)) + off[d]) * hb for d in range(3)]

    # scalar fields at cell centres
    xc = pts((b, b, b), (0.5, 0.5, 0.5))
    g = []
    for _ in range(5):
        ks, a, ph = _modes(rng, 4, nmodes, 0.0)
        a = a * np.exp(-0.5 * (np.linalg.norm(ks, axis=1) / 2.0) ** 2)
        a = a / math.sqrt(0.5 * (a ** 2).sum())
        g.append(_series(torch, ks, a, ph, *xc))
    rho = torch.exp(0.3 * g[0])
    p = torch.clamp(1.0 + 0.1 * g[1], min=0.05)
    v = [0.5 * g[2], 0.5 * g[3], 0.5 * g[4]]
    # vector potential modes
    Am = []
    for _ in range(3):
        ks, a, ph = _modes(rng, kmax, nmodes, -2.0)
        Am.append((ks, 0.5 * a, ph))
    # edge line integrals: I_c along axis c on the edge grid of each block
    def edge_int(c):
        shape = [b + 1] * 3
        shape[c] = b
        p0 = pts(tuple(shape), (0.0, 0.0, 0.0))
        I = _line_integral(torch, *Am[c], p0, c, hb)
        return I.reshape(n, *shape[::-1])  # [n, z, y, x]
    Ix, Iy, Iz = edge_int(0), edge_int(1), edge_int(2)
    area = (hb * hb).reshape(n, 1, 1, 1)
    # x-faces (n, b, b, b+1): Iz(j+1) - Iz(j) - Iy(k+1) + Iy(k)
    Bx = (Iz[:, :, 1:, :] - Iz[:, :, :-1, :] - Iy[:, 1:, :, :] + Iy[:, :-1, :, :]) / area
    # y-faces (n, b, b+1, b): Ix(k+1) - Ix(k) - Iz(i+1) + Iz(i)
    By = (Ix[:, 1:, :, :] - Ix[:, :-1, :, :] - Iz[:, :, :, 1:] + Iz[:, :, :, :-1]) / area
    # z-faces (n, b+1, b, b): Iy(i+1) - Iy(i) - Ix(j+1) + Ix(j)
    Bz = (Iy[:, :, :, 1:] - Iy[:, :, :, :-1] - Ix[:, :, 1:, :] + Ix[:, :, :-1, :]) / area
    np_ = lambda t: t.reshape(n, -1).cpu().numpy()  # noqa: E731
    rho_n = np_(rho)
    mom = [rho_n * np_(vc) for vc in v]
    return _assemble(mesh, rho_n, mom, np_(p), [np_(Bx), np_(By), np_(Bz)], gamma)
