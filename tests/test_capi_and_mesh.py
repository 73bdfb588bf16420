"""CPU tests of the C ABI library and of the host block octree (no GPU needed).

* libmhdct.so loads and exports every function include/mhdct.h declares, and the Python
  binding declares a signature for each of them;
* the octree: leaves tile the domain exactly, refinement keeps 2:1 balance across faces,
  edges and corners, outside blocks appear only on non-periodic borders, and the error paths
  return codes with a message instead of aborting.
"""
import ctypes as C
import itertools
import re
from pathlib import Path

import numpy as np
import pytest

from godunov_mhd_ct_amd import MhdctError, lib
from godunov_mhd_ct_amd._lib import SIGNATURES
from godunov_mhd_ct_amd.mesh import Mesh, refine_where, uniform_mesh

HEADER = Path(__file__).resolve().parents[1] / "include" / "mhdct.h"


def declared():
    txt = HEADER.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mhdct_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_abi():
    names = declared()
    for must in ["mhdct_create", "mhdct_set_mesh", "mhdct_do_time_step", "mhdct_compute_dt", "mhdct_fill_outside",
                 "mhdct_mem_bytes", "mhdct_last_error", "mhdct_destroy", "mhdct_batch_step"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    L = lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert sorted(SIGNATURES) == declared()


def leaf_boxes(m):
    lv, co = m.leaves()
    n = m.num_leaves
    h = 1.0 / (1 << lv[:n]).astype(np.float64)
    lo = co[:n, :m.dim] * h[:, None]
    return lo, lo + h[:, None], lv[:n]


@pytest.mark.parametrize("dim", [2, 3])
def test_leaves_tile_the_domain(dim):
    m = Mesh(dim, 4, 1)
    m = refine_where(m, lambda L, x0, x1, y0, y1, z0, z1: x0 < 0.3 and y0 < 0.4, 3)
    lo, hi, lv = leaf_boxes(m)
    vol = np.prod(hi - lo, axis=1)
    assert abs(vol.sum() - 1.0) < 1e-12
    # no two leaves overlap: sample points at the centres of the finest cells
    f = 1 << int(lv.max())
    pts = (np.stack(np.meshgrid(*[np.arange(f)] * dim, indexing="ij"), -1).reshape(-1, dim) + 0.5) / f
    inside = np.zeros(len(pts), dtype=int)
    for a, b in zip(lo, hi):
        inside += np.all((pts >= a) & (pts < b), axis=1)
    assert np.all(inside == 1)


@pytest.mark.parametrize("dim", [2, 3])
def test_two_to_one_balance(dim):
    m = Mesh(dim, 4, 1)
    # refine towards a point repeatedly: without balancing this would break 2:1
    m = refine_where(m, lambda L, x0, x1, y0, y1, z0, z1: x0 <= 0.01 and y0 <= 0.01 and (dim == 2 or z0 <= 0.01), 5)
    lo, hi, lv = leaf_boxes(m)
    n = len(lv)
    eps = 1e-12
    for a, b in itertools.combinations(range(n), 2):
        # leaves touching across a face, an edge or a corner (periodic images included)
        touch = True
        for d in range(dim):
            gap = min(abs(lo[a, d] - hi[b, d]), abs(lo[b, d] - hi[a, d]),
                      abs(lo[a, d] - hi[b, d] + 1.0), abs(lo[b, d] - hi[a, d] + 1.0))
            overlap = lo[a, d] < hi[b, d] - eps and lo[b, d] < hi[a, d] - eps
            if not (overlap or gap < eps):
                touch = False
                break
        if touch:
            assert abs(int(lv[a]) - int(lv[b])) <= 1, (a, b, lv[a], lv[b])


def test_outside_blocks_only_on_bounded_borders():
    m = uniform_mesh(2, 4, 2)
    assert m.num_outside == 0
    m = Mesh(2, 4, 2, periodic=(0, 1, 1))
    m.set_bc(["ZERO_GRADIENT", "WALL", "ZERO_GRADIENT", "ZERO_GRADIENT", "ZERO_GRADIENT", "ZERO_GRADIENT"])
    m.finalize()
    # one column of outside blocks on each x border, 4 blocks high
    assert m.num_outside == 2 * 4
    src, dirs = m.outside()
    assert set(np.abs(dirs[:, 0]).tolist()) == {1} and not dirs[:, 1].any()


def test_error_paths_return_codes():
    h = C.c_void_p()
    nb = (C.c_int * 3)(1, 1, 1)
    pe = (C.c_int * 3)(1, 1, 1)
    assert lib().mhdct_mesh_create(3, 5, nb, pe, 1, C.byref(h)) != 0  # odd block size
    with pytest.raises(MhdctError):
        Mesh(3, 5, 1)
    m = Mesh(2, 4, 1)
    with pytest.raises(MhdctError):
        m.adapt(np.zeros(m.num_leaves + 3, dtype=np.int8))
    assert lib().mhdct_last_error(None) is not None


def test_settings_struct_and_time_integrator():
    # the ctypes mirror of mhdct_settings has the header's fields in the header's order
    from godunov_mhd_ct_amd._lib import Settings
    from godunov_mhd_ct_amd.implem import hydro_settings

    txt = HEADER.read_text()
    body = re.search(r"typedef struct mhdct_settings\s*\{(.*?)\}\s*mhdct_settings;", txt, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"\b(?:double|int)\s+([a-z0-9_]+)\s*(?:\[\d+\])?\s*;", body)
    assert fields == [f[0] for f in Settings._fields_]
    assert hydro_settings(time_integrator="RK2_SSP").time_integrator == 1
    assert hydro_settings().time_integrator == 0
    with pytest.raises(MhdctError, match="RK3_SSP"):
        hydro_settings(time_integrator="RK3_SSP")  # the reference aborts on it (SolverGodunovMHD.cpp:174-177)


def test_level0_origin_leaf_is_found():
    # the key of the level-0 leaf at the origin is 0: a refined neighbour's tables must still
    # find it (the coarser neighbour of every child beside it)
    for dim in (2, 3):
        m = Mesh(dim, 8, 0, nbrick=(2, 2, 2 if dim == 3 else 1), periodic=(1, 1, 1))
        lv, co = m.leaves()
        n = m.num_leaves
        flags = np.zeros(n, dtype=np.int8)
        flags[[i for i in range(n) if tuple(co[i][:dim]) == (1,) + (0,) * (dim - 1)]] = 1
        m.adapt(flags)
        m.balance()
        m.finalize()
        assert m.num_leaves == n - 1 + (1 << dim)


def test_balance_large_mesh_is_two_to_one():
    # the threaded balance on a mesh big enough to split over threads
    m = Mesh(3, 8, 2)
    def shell(L, x0, x1, y0, y1, z0, z1):
        # blocks crossed by the sphere of radius 0.3 about (0.5, 0.45, 0.55)
        c = ((x0 + x1) / 2 - 0.5, (y0 + y1) / 2 - 0.45, (z0 + z1) / 2 - 0.55)
        return abs(np.sqrt(sum(v * v for v in c)) - 0.3) < 0.9 * (x1 - x0)

    m = refine_where(m, shell, 4)
    assert m.num_leaves > 1000  # split over several threads (256 leaves or more each)
    lv, co = m.leaves()
    n = m.num_leaves
    h = 1.0 / (1 << lv[:n]).astype(np.float64)
    lo = co[:n] * h[:, None]
    hi = lo + h[:, None]
    # every pair of touching leaves (faces, edges, corners, periodic images) differs by <= 1
    for shift in itertools.product((-1.0, 0.0, 1.0), repeat=3):
        s = np.array(shift)
        lo2, hi2 = lo + s, hi + s
        touch = np.all((lo[:, None, :] <= hi2[None, :, :] + 1e-12) & (lo2[None, :, :] <= hi[:, None, :] + 1e-12),
                       axis=2)
        a, b = np.nonzero(touch)
        assert np.all(np.abs(lv[a].astype(int) - lv[b].astype(int)) <= 1)


def _order(nb, level, dim):
    m = Mesh(dim, 8, level, nbrick=nb).finalize()
    lv, co = m.leaves()
    return [tuple(int(x) for x in c[:dim]) for c in co[:m.num_leaves]]


def test_leaf_order_follows_p4est_brick_trees():
    """Owned leaves come in p4est's order: trees as p4est_connectivity_new_brick numbers them
    (bits of the tree position interleaved, the axis with fewer bits dropping out first, positions
    outside the brick skipped), then the leaves of each tree along the Morton curve (x bit
    fastest).  So U and B in the reference's DataArrayBlock order pass the ABI as they are."""
    # 3 x 2: 2 bits in x, 1 in y -> (0,0) (1,0) (0,1) (1,1) then the x >= 2 pair (2,0) (2,1)
    assert _order((3, 2, 1), 0, 2) == [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (2, 1)]
    # 3 x 3 inside the 4 x 4 Morton square
    assert _order((3, 3, 1), 0, 2) == [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (2, 1), (0, 2), (1, 2), (2, 2)]
    # a line of trees, and a power-of-two cube (pure Morton)
    assert _order((5, 1, 1), 0, 3) == [(0, 0, 0), (1, 0, 0), (2, 0, 0), (3, 0, 0), (4, 0, 0)]
    assert _order((2, 2, 2), 0, 3) == [(x, y, z) for z in range(2) for y in range(2) for x in range(2)]
    # 1 x 3 (the Rayleigh-Taylor brick): y only
    assert _order((1, 3, 1), 0, 2) == [(0, 0), (0, 1), (0, 2)]
    # 2 x 1 at level 1: tree 0's four leaves in Morton order, then tree 1's
    assert _order((2, 1, 1), 1, 2) == [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (3, 0), (2, 1), (3, 1)]
    # 3 x 2 x 2: 2 bits in x, 1 in y and z: one interleaved round of x, y, z, then x's last bit
    got = _order((3, 2, 2), 0, 3)
    assert got[:8] == [(x, y, z) for z in range(2) for y in range(2) for x in range(2)]
    assert got[8:] == [(2, y, z) for z in range(2) for y in range(2)]


def test_set_leaves_puts_a_shuffled_list_in_p4est_order():
    rng = np.random.default_rng(3)
    m = refine_where(Mesh(2, 8, 1, nbrick=(3, 2, 1)), lambda L, x0, x1, y0, y1, z0, z1: x0 < 1.5 and y0 < 0.5, 2)
    lv, co = m.leaves()
    n = m.num_leaves
    perm = rng.permutation(n)
    m2 = Mesh(2, 8, 1, nbrick=(3, 2, 1))
    m2.set_leaves(lv[:n][perm], co[:n][perm])
    m2.finalize()
    lv2, co2 = m2.leaves()
    assert np.array_equal(lv2[:n], lv[:n]) and np.array_equal(co2[:n], co[:n])
