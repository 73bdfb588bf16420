"""The benchmark workloads: meshes, initial states and settings.

  ot2d  2D Orszag-Tang, [0,1]^2 periodic, octree level 4 with 8x8 blocks (128^2 cells),
        gamma 5/3, minmod, HLLD, CFL 0.8
  bl3d  3D MHD blast, [0,1]^3 periodic, 16^3 blocks, levels 3-6 refined where the normalised
        density/pressure jump exceeds 0.1, p 10 / 0.1 inside / outside r < 0.1, B = (1,1,1)/sqrt 3
  kh3d  3D Kelvin-Helmholtz, 256^3 as 16^3 blocks at uniform level 4, gamma 1.4
"""
from __future__ import annotations

import math

import numpy as np

from . import problems
from .implem import hydro_settings
from .mesh import Mesh, uniform_mesh


def _block_jump(mesh: Mesh, field_fn, scaling=1.0, xmin=(0.0, 0.0, 0.0)):
    """Largest normalised jump |q_r - q_l| / max(|q_l|, |q_r|) between neighbouring cells
    inside each leaf, for a field given as a function of cell-centre coordinates."""
    xs = problems._grid(mesh, scaling, xmin, "cell")
    q = field_fn(*xs)
    n = q.shape[0]
    b, dim = mesh.b, mesh.dim
    q = q.reshape((n,) + (b,) * dim)
    out = np.zeros(n)
    for ax in range(1, dim + 1):
        a = np.take(q, range(0, b - 1), axis=ax)
        c = np.take(q, range(1, b), axis=ax)
        den = np.maximum(np.maximum(np.abs(a), np.abs(c)), 1e-300)
        j = np.abs(c - a) / den
        out = np.maximum(out, j.reshape(n, -1).max(axis=1))
    return out


def refine_on_jump(mesh: Mesh, fields, threshold: float, level_max: int) -> Mesh:
    while True:
        ind = np.zeros(mesh.num_total)
        for f in fields:
            ind = np.maximum(ind, _block_jump(mesh, f))
        lv, _ = mesh.leaves()
        n = mesh.num_leaves
        flags = ((ind[:n] > threshold) & (lv[:n] < level_max)).astype(np.int8)
        if not flags.any():
            break
        mesh.adapt(flags)
        mesh.balance()
    return mesh.finalize()


def ot2d(level=4, b=8):
    mesh = uniform_mesh(2, b, level)
    U, B = problems.orszag_tang_2d(mesh)
    st = hydro_settings(gamma0=5.0 / 3.0, cfl=0.8, slope_type=1.0)
    return mesh, U, B, st


def bl3d(b=16, level_min=3, level_max=6, threshold=0.1):
    mesh = Mesh(3, b, level_min)
    c = (0.5, 0.5, 0.5)
    r0 = 0.1

    def pressure(x, y, z):
        d2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
        return np.where(d2 < r0 * r0, 10.0, 0.1)

    mesh = refine_on_jump(mesh, [pressure], threshold, level_max)
    U, B = problems.blast(mesh, radius=r0, center=c, rho_in=1.0, rho_out=1.0, p_in=10.0, p_out=0.1,
                          bfield=(1 / math.sqrt(3),) * 3)
    st = hydro_settings(gamma0=5.0 / 3.0, cfl=0.8, slope_type=1.0)
    return mesh, U, B, st


def kh3d(level=4, b=16, seed=0):
    mesh = uniform_mesh(3, b, level)
    U, B = problems.kelvin_helmholtz_3d(mesh, gamma=1.4, seed=seed)
    st = hydro_settings(gamma0=1.4, cfl=0.8, slope_type=1.0)
    return mesh, U, B, st


def random_amr_mesh(b, target, seed, level_min=2, level_max=6) -> Mesh:
    """Periodic 2:1-balanced octree of about `target` leaves of b^3 cells over levels
    level_min..level_max: leaves are refined, highest score first, where a smooth random
    indicator (seeded) is large, until the leaf count reaches the target."""
    rng = np.random.default_rng(seed)
    ks = rng.integers(-3, 4, size=(6, 3)).astype(np.float64)
    ph = rng.uniform(0, 2 * math.pi, size=6)
    mesh = Mesh(3, b, level_min)
    while mesh.num_leaves < target:
        lv, co = mesh.leaves()
        n = mesh.num_leaves
        c = (co[:n].astype(np.float64) + 0.5) / (1 << lv[:n]).astype(np.float64)[:, None]
        score = sum(np.cos(2 * math.pi * (c @ k) + p) for k, p in zip(ks, ph))
        score = score + 1e-3 * rng.random(n)
        score[lv[:n] >= level_max] = -np.inf
        if not np.isfinite(score).any():
            break
        nref = max(1, min(int((target - n) / 7 / 2), int(np.isfinite(score).sum())))
        flags = np.zeros(n, dtype=np.int8)
        flags[np.argsort(-score)[:nref]] = 1
        mesh.adapt(flags)
        mesh.balance()
    return mesh.finalize()


def mix3d(seed=0, counts=((8, 2048), (16, 1434), (32, 614)), level_min=2, level_max=6):
    """The batching workload: three independent periodic meshes of 8^3, 16^3 and 32^3 blocks
    (about 2048, 1434 and 614 leaves, 27 M cells), levels 2-6, each with a smooth random MHD
    state (problems.random_mhd_3d), gamma 5/3, minmod, HLLD, CFL 0.8.  Returns a list of
    (mesh, U, B, settings)."""
    out = []
    for n, (b, target) in enumerate(counts):
        mesh = random_amr_mesh(b, target, seed + 101 * n, level_min, level_max)
        U, B = problems.random_mhd_3d(mesh, seed=seed + 7 * n)
        st = hydro_settings(gamma0=5.0 / 3.0, cfl=0.8, slope_type=1.0)
        out.append((mesh, U, B, st))
    return out


WORKLOADS = {"ot2d": ot2d, "bl3d": bl3d, "kh3d": kh3d}
