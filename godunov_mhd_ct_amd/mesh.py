"""Host block octree (the p4est brick forest of the reference, kept on the host).

Wraps the mhdct_mesh_* entry points: a uniform forest, refine/coarsen flags, full 2:1
balance, then finalize() builds the outside blocks and the neighbour tables the kernels read.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import MhdctError, lib

# BC_MHD (border_conditions/FillOutside.h:55-61)
BC = {"NONE": 0, "PERIODIC": 1, "ZERO_GRADIENT": 2, "WALL": 3}


def _ivec(vals):
    arr = (C.c_int * len(vals))(*[int(v) for v in vals])
    return arr


class Mesh:
    """A forest of (quad|oct)trees on an nbrick brick, leaves of b^dim cells."""

    def __init__(self, dim: int, block_size: int, level: int, nbrick=(1, 1, 1), periodic=(1, 1, 1)):
        self.dim = int(dim)
        self.b = int(block_size)
        self.nbrick = tuple(int(x) for x in nbrick) + (1,) * (3 - len(nbrick))
        self.periodic = tuple(int(bool(x)) for x in periodic) + (1,) * (3 - len(periodic))
        h = C.c_void_p()
        rc = lib().mhdct_mesh_create(self.dim, self.b, _ivec(self.nbrick), _ivec(self.periodic), int(level),
                                     C.byref(h))
        if rc != 0:
            raise MhdctError(f"mhdct_mesh_create failed ({rc})")
        self._h = h
        self.bc = (0,) * 6
        self.finalized = False

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().mhdct_mesh_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    @classmethod
    def from_handle(cls, h, dim, b, nbrick, periodic, bc) -> "Mesh":
        """Wrap a mesh handle the library created (it is destroyed with the wrapper)."""
        m = cls.__new__(cls)
        m.dim, m.b, m.nbrick, m.periodic, m.bc = int(dim), int(b), tuple(nbrick), tuple(periodic), tuple(bc)
        m._h = h
        m.finalized = False
        return m

    def clone(self) -> "Mesh":
        """An independent copy (leaves, border conditions, tables)."""
        h = C.c_void_p()
        self._check(lib().mhdct_mesh_clone(self._h, C.byref(h)), "clone")
        m = Mesh.__new__(Mesh)
        m.dim, m.b, m.nbrick, m.periodic, m.bc = self.dim, self.b, self.nbrick, self.periodic, self.bc
        m._h = h
        m.finalized = self.finalized
        return m

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().mhdct_mesh_last_error(self._h)
            raise MhdctError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def set_bc(self, bc):
        """bc: 6 BC_MHD values or names (XMIN, XMAX, YMIN, YMAX, ZMIN, ZMAX)."""
        vals = [BC[v] if isinstance(v, str) else int(v) for v in bc] + [0] * (6 - len(bc))
        self._check(lib().mhdct_mesh_set_bc(self._h, _ivec(vals)), "set_bc")
        self.bc = tuple(vals[:6])
        self.finalized = False

    def adapt(self, flags):
        f = np.ascontiguousarray(np.asarray(flags, dtype=np.int8))
        if f.size != self.num_leaves:
            raise MhdctError("adapt: one flag per leaf")
        self._check(lib().mhdct_mesh_adapt(self._h, f.ctypes.data_as(C.POINTER(C.c_int8)), int(f.size)), "adapt")
        self.finalized = False

    def set_leaves(self, levels, coords):
        """Replace the leaves by a given tiling (levels[n], coords[n, 3])."""
        lv = np.ascontiguousarray(levels, dtype=np.int32)
        co = np.ascontiguousarray(coords, dtype=np.int32).reshape(-1, 3)
        if co.shape[0] != lv.size:
            raise MhdctError("set_leaves: one coordinate triple per level")
        self._check(lib().mhdct_mesh_set_leaves(self._h, int(lv.size), lv.ctypes.data_as(C.POINTER(C.c_int)),
                                                co.ctypes.data_as(C.POINTER(C.c_int))), "set_leaves")
        self.finalized = False

    def balance(self):
        self._check(lib().mhdct_mesh_balance(self._h), "balance")
        self.finalized = False

    def finalize(self):
        self._check(lib().mhdct_mesh_finalize(self._h), "finalize")
        self.finalized = True
        return self

    @property
    def num_leaves(self) -> int:
        return int(lib().mhdct_mesh_num_leaves(self._h))

    @property
    def num_outside(self) -> int:
        return int(lib().mhdct_mesh_num_outside(self._h))

    @property
    def num_total(self) -> int:
        return self.num_leaves + self.num_outside

    def leaves(self):
        """(levels[n], coords[n,3]) of the owned leaves then the outside blocks."""
        n = self.num_total
        lv = np.zeros(n, dtype=np.int32)
        co = np.zeros((n, 3), dtype=np.int32)
        self._check(lib().mhdct_mesh_leaves(self._h, lv.ctypes.data_as(C.POINTER(C.c_int)),
                                            co.ctypes.data_as(C.POINTER(C.c_int))), "leaves")
        return lv, co

    def outside(self):
        n = self.num_outside
        src = np.zeros(max(n, 1), dtype=np.int32)
        dirs = np.zeros((max(n, 1), 3), dtype=np.int32)
        self._check(lib().mhdct_mesh_outside(self._h, src.ctypes.data_as(C.POINTER(C.c_int)),
                                             dirs.ctypes.data_as(C.POINTER(C.c_int))), "outside")
        return src[:n], dirs[:n]

    # sizes of the ABI arrays -------------------------------------------------------
    @property
    def cells_per_block(self) -> int:
        return self.b ** self.dim

    @property
    def faces_per_block(self) -> int:
        b = self.b
        return 2 * (b + 1) * b + b * b if self.dim == 2 else 3 * (b + 1) * b * b

    def owned_cells(self) -> int:
        return self.num_leaves * self.cells_per_block


def uniform_mesh(dim, b, level, nbrick=(1, 1, 1), periodic=(1, 1, 1), bc=None) -> Mesh:
    m = Mesh(dim, b, level, nbrick, periodic)
    if bc is not None:
        m.set_bc(bc)
    return m.finalize()


def refine_where(mesh: Mesh, predicate, level_max: int, iterations: int | None = None) -> Mesh:
    """Refine leaves where predicate(level, x0, x1, y0, y1, z0, z1) is true (block bounds in
    brick units), up to level_max, balancing after each pass."""
    its = iterations if iterations is not None else level_max
    for _ in range(its):
        lv, co = mesh.leaves()
        n = mesh.num_leaves
        flags = np.zeros(n, dtype=np.int8)
        for i in range(n):
            L = int(lv[i])
            if L >= level_max:
                continue
            h = 1.0 / (1 << L)
            x0, y0, z0 = co[i] * h
            if predicate(L, x0, x0 + h, y0, y0 + h, z0, z0 + h):
                flags[i] = 1
        if not flags.any():
            break
        mesh.adapt(flags)
        mesh.balance()
    return mesh.finalize()
