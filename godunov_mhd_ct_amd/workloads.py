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


def bl3d_amr_settings(level_min=3, level_max=6):
    """The blast configuration's AMR cycle: refine where the normalised jump of the density or of
    the thermal pressure exceeds 0.1 (SIMPLE_GRADIENT on rho and thermal_pressure), coarsen below
    0.01, a cycle every 10 steps after the first (cycle_frequency and disable_first_amr_cycle of
    app/test_mhd_blast_3d.ini)."""
    from .amr import AMRSettings

    return AMRSettings(level_min=level_min, level_max=level_max, epsilon_refine=0.1, epsilon_coarsen=0.01,
                       variable=["rho", "thermal_pressure"], criterion="SIMPLE_GRADIENT", cycle_frequency=10)


def kh3d(level=4, b=16, seed=0):
    mesh = uniform_mesh(3, b, level)
    U, B = problems.kelvin_helmholtz_3d(mesh, gamma=1.4, seed=seed)
    st = hydro_settings(gamma0=1.4, cfl=0.8, slope_type=1.0)
    return mesh, U, B, st


def random_amr_mesh(b, target, seed, level_min=2, level_max=6, nbrick=(1, 1, 1)) -> Mesh:
    """Periodic 2:1-balanced octree of exactly `target` leaves of b^3 cells over levels
    level_min..level_max on an nbrick brick of trees: leaves are refined, highest score first,
    where a smooth random indicator (seeded) is large.  Far from the target many leaves are
    refined per pass; near it one leaf at a time, skipping any whose refinement (with the 2:1
    balance it triggers) would pass the target.  Refining a leaf adds 7 leaves, so the target
    must equal the uniform level_min leaf count plus a multiple of 7."""
    ntrees = nbrick[0] * nbrick[1] * nbrick[2]
    base = ntrees * 8 ** level_min
    if target < base or (target - base) % 7:
        raise ValueError(f"{target} leaves are not reachable from {base} by refinement (steps of 7)")
    rng = np.random.default_rng(seed)
    ks = rng.integers(-3, 4, size=(6, 3)).astype(np.float64)
    ph = rng.uniform(0, 2 * math.pi, size=6)
    mesh = Mesh(3, b, level_min, nbrick=nbrick)
    while mesh.num_leaves < target:
        lv, co = mesh.leaves()
        n = mesh.num_leaves
        c = (co[:n].astype(np.float64) + 0.5) / (1 << lv[:n]).astype(np.float64)[:, None]
        score = sum(np.cos(2 * math.pi * (c @ k) + p) for k, p in zip(ks, ph))
        score = score + 1e-3 * rng.random(n)
        score[lv[:n] >= level_max] = -np.inf
        order = [int(i) for i in np.argsort(-score) if np.isfinite(score[i])]
        if not order:
            raise ValueError(f"no leaf below level {level_max} left to refine at {n} leaves")
        rem = (target - n) // 7
        nref = rem // 4
        if nref >= 2:
            flags = np.zeros(n, dtype=np.int8)
            flags[order[:nref]] = 1
            trial = mesh.clone()
            trial.adapt(flags)
            trial.balance()
            if trial.num_leaves <= target:
                mesh = trial
                continue
        # close to the target: the best leaf whose refinement stays within it
        for i in order:
            flags = np.zeros(n, dtype=np.int8)
            flags[i] = 1
            trial = mesh.clone()
            trial.adapt(flags)
            trial.balance()
            if trial.num_leaves <= target:
                mesh = trial
                break
        else:
            raise ValueError(f"no single refinement fits between {n} and {target} leaves")
    return mesh.finalize()


# The batching configuration: 4096 blocks, 2048 x 8^3 + 1434 x 16^3 + 614 x 32^3 = 27,041,792
# cells, levels 2-6.  A single tree's leaf count stays 1 modulo 7 under refinement, so each
# mesh lies on a periodic brick of trees whose count matches its target modulo 7 (4, 6 and 5
# trees); the random fields are periodic on the unit cube, hence on every brick of whole units.
MIX_MESHES = ((8, 2048, (2, 2, 1)), (16, 1434, (3, 2, 1)), (32, 614, (5, 1, 1)))


def mix3d(seed=0, meshes=MIX_MESHES, level_min=2, level_max=6):
    """The batching workload: three independent periodic meshes of 8^3, 16^3 and 32^3 blocks
    (2048, 1434 and 614 leaves, 27,041,792 cells), levels 2-6, each with a smooth random MHD
    state (problems.random_mhd_3d), gamma 5/3, minmod, HLLD, CFL 0.8.  Returns a list of
    (mesh, U, B, settings)."""
    out = []
    for n, (b, target, nbrick) in enumerate(meshes):
        mesh = random_amr_mesh(b, target, seed + 101 * n, level_min, level_max, nbrick)
        U, B = problems.random_mhd_3d(mesh, seed=seed + 7 * n)
        st = hydro_settings(gamma0=5.0 / 3.0, cfl=0.8, slope_type=1.0)
        out.append((mesh, U, B, st))
    return out


WORKLOADS = {"ot2d": ot2d, "bl3d": bl3d, "kh3d": kh3d}
