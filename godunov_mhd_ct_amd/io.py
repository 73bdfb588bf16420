"""Checkpoint, restart and visualisation output (SolverGodunovMHD::save_solution_hdf5,
SolverGodunovMHD.cpp:856-1042; init_restart, init/init.cpp:107-193).

The reference writes HDF5 (through kalypsso::core) with an XDMF index, plus the p4est mesh
file. This image has no HDF5 library, so the same content is kept in a self-describing form:

* ``<prefix>.json``: the header. It holds the run time and iteration (the reference's
  ``run/time`` and ``run/iteration`` attributes), the mesh (dimension, block size, brick,
  periodicity, border conditions) and the byte layout of every array in ``<prefix>.bin``.
* ``<prefix>.bin``: raw little-endian arrays:
  - the leaf levels and block coordinates (what the p4est file provides on restart);
  - the conservative variables ``rho, e_tot, rho_vx, rho_vy, rho_vz`` of every owned leaf;
  - the face-centred ``magnetic_field`` (the reference's checkpoint name).
* ``<prefix>.xmf`` with ``<prefix>_cells.bin``: XDMF 2 with one co-rectilinear grid per leaf.
  Each grid carries the ``write_variables`` the reference knows:
  - the five conservative variables;
  - ``Bx``, ``By``, ``Bz`` (cell-centred, the mean of the two faces; 2D ``Bz`` is cell centred);
  - ``divB`` (per cell, divided by dx);
  - ``thermal_pressure``, ``magnetic_pressure`` and ``specific_ekin``.

  ParaView and VisIt read it as is.

Restart is exact: the state is read back bit for bit, and the mesh is rebuilt from the leaf
list (``Mesh.set_leaves``).
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from ._lib import MhdctError, Settings
from .mesh import Mesh

FORMAT = "godunov_mhd_ct_amd checkpoint"
VERSION = 1
# MHD::m_id2names_all (models/MHD.cpp:20-32), the conservative cell variables in U order
CELL_VARS = ["rho", "e_tot", "rho_vx", "rho_vy", "rho_vz"]
DERIVED = ["Bx", "By", "Bz", "divB", "thermal_pressure", "magnetic_pressure", "specific_ekin"]


def _align(n: int) -> int:
    return (n + 7) // 8 * 8


def settings_dict(settings) -> dict:
    """The fields of an mhdct_settings struct as plain values."""
    out = {}
    for name, _ in Settings._fields_:
        v = getattr(settings, name)
        out[name] = list(v) if not isinstance(v, (int, float)) else v
    return out


def settings_from_dict(d: dict) -> Settings:
    st = Settings()
    for name, _ in Settings._fields_:
        if name not in d:
            raise MhdctError(f"checkpoint settings lack {name}")
        v = d[name]
        if isinstance(v, list):
            getattr(st, name)[:] = v
        else:
            setattr(st, name, v)
    return st


def _named(prefix: Path, suffix: str) -> Path:
    """``<prefix><suffix>``: the suffix is appended, so a prefix with dots in its name (e.g.
    ``kh.mu1e-4_iter0000010``) keeps every character of it."""
    return prefix.parent / (prefix.name + suffix)


def save_checkpoint(prefix, mesh: Mesh, U, B, time: float, iteration: int, settings=None, extra=None) -> Path:
    """Write ``<prefix>.json`` and ``<prefix>.bin`` (owned leaves only); returns the header path."""
    prefix = Path(prefix)
    n = mesh.num_leaves
    ncell = mesh.cells_per_block
    fper = mesh.faces_per_block
    U = np.ascontiguousarray(np.asarray(U, dtype=np.float64).reshape(mesh.num_total, 5, ncell)[:n])
    B = np.ascontiguousarray(np.asarray(B, dtype=np.float64).reshape(mesh.num_total, fper)[:n])
    lv, co = mesh.leaves()
    arrays = [("levels", np.ascontiguousarray(lv[:n], dtype="<i4")),
              ("coords", np.ascontiguousarray(co[:n], dtype="<i4")),
              ("U", U.astype("<f8", copy=False)),
              ("magnetic_field", B.astype("<f8", copy=False))]
    layout, off = [], 0
    for name, a in arrays:
        layout.append({"name": name, "dtype": a.dtype.str, "shape": list(a.shape), "offset": off})
        off = _align(off + a.nbytes)
    head = {"format": FORMAT, "version": VERSION, "dim": mesh.dim, "block_size": mesh.b,
            "nbrick": list(mesh.nbrick[:3]), "periodic": list(mesh.periodic[:3]),
            "bc": list(mesh.bc), "time": float(time), "iteration": int(iteration),
            "num_leaves": n, "cell_variables": CELL_VARS, "arrays": layout,
            "settings": settings_dict(settings) if settings is not None else None, "extra": extra or {}}
    with open(_named(prefix, ".bin"), "wb") as f:
        pos = 0
        for (name, a), ent in zip(arrays, layout):
            f.write(b"\0" * (ent["offset"] - pos))
            f.write(a.tobytes())
            pos = ent["offset"] + a.nbytes
    path = _named(prefix, ".json")
    path.write_text(json.dumps(head, indent=1))
    return path


def load_checkpoint(prefix):
    """(mesh, U, B, header) from a checkpoint; U and B in the ABI layout of the rebuilt mesh,
    outside blocks zero (the step's outside fill sets them)."""
    prefix = Path(prefix)
    head = json.loads(_named(prefix, ".json").read_text())
    if head.get("format") != FORMAT or head.get("version") != VERSION:
        raise MhdctError(f"{prefix}.json is not a version {VERSION} checkpoint of this package")
    raw = np.fromfile(_named(prefix, ".bin"), dtype=np.uint8)
    arrs = {}
    for ent in head["arrays"]:
        dt = np.dtype(ent["dtype"])
        cnt = int(np.prod(ent["shape"])) if ent["shape"] else 1
        end = ent["offset"] + cnt * dt.itemsize
        if end > raw.size:
            raise MhdctError(f"{prefix}.bin is truncated ({ent['name']})")
        arrs[ent["name"]] = raw[ent["offset"]:end].view(dt).reshape(ent["shape"])
    dim, b = head["dim"], head["block_size"]
    mesh = Mesh(dim, b, 0, nbrick=tuple(head["nbrick"]), periodic=tuple(head["periodic"]))
    if any(head["bc"]):
        mesh.set_bc(head["bc"])
    mesh.set_leaves(arrs["levels"], arrs["coords"])
    mesh.finalize()
    n = mesh.num_leaves
    lv, co = mesh.leaves()
    # the rebuilt mesh orders its leaves itself: map the file's leaves onto it
    key = {(int(l), *map(int, c)): i for i, (l, c) in enumerate(zip(arrs["levels"], arrs["coords"]))}
    perm = np.array([key[(int(lv[i]), *map(int, co[i]))] for i in range(n)], dtype=np.int64).reshape(n)
    U = np.zeros((mesh.num_total, 5, mesh.cells_per_block))
    B = np.zeros((mesh.num_total, mesh.faces_per_block))
    U[:n] = arrs["U"][perm]
    B[:n] = arrs["magnetic_field"][perm]
    return mesh, U, B, head


def cell_fields(mesh: Mesh, U, B, gamma: float, scaling: float = 1.0) -> dict:
    """The output variables of every owned leaf, each [leaf, cell] (x fastest)."""
    from .problems import cell_centred_b

    n = mesh.num_leaves
    U = np.asarray(U, dtype=np.float64).reshape(mesh.num_total, 5, -1)[:n]
    Bf = np.asarray(B, dtype=np.float64).reshape(mesh.num_total, -1)[:n]
    out = {name: U[:, v] for v, name in enumerate(CELL_VARS)}
    bx, by, bz = cell_centred_b(mesh, Bf)
    out["Bx"], out["By"], out["Bz"] = bx, by, bz
    b, dim = mesh.b, mesh.dim
    lv, _ = mesh.leaves()
    dx = scaling / (b * (1 << lv[:n]).astype(np.float64))
    if dim == 2:
        nx = (b + 1) * b
        fx = Bf[:, :nx].reshape(n, b, b + 1)
        fy = Bf[:, nx:2 * nx].reshape(n, b + 1, b)
        div = (fx[:, :, 1:] - fx[:, :, :-1]) + (fy[:, 1:, :] - fy[:, :-1, :])
    else:
        nf = (b + 1) * b * b
        fx = Bf[:, :nf].reshape(n, b, b, b + 1)
        fy = Bf[:, nf:2 * nf].reshape(n, b, b + 1, b)
        fz = Bf[:, 2 * nf:].reshape(n, b + 1, b, b)
        div = (fx[..., 1:] - fx[..., :-1]) + (fy[:, :, 1:, :] - fy[:, :, :-1, :]) + (fz[:, 1:] - fz[:, :-1])
    out["divB"] = div.reshape(n, -1) / dx[:, None]
    rho = out["rho"]
    ekin = 0.5 * (out["rho_vx"] ** 2 + out["rho_vy"] ** 2 + out["rho_vz"] ** 2) / rho
    emag = 0.5 * (bx * bx + by * by + bz * bz)
    out["thermal_pressure"] = (gamma - 1.0) * (out["e_tot"] - ekin - emag)
    out["magnetic_pressure"] = emag
    out["specific_ekin"] = ekin / rho
    return out


def write_xdmf(prefix, mesh: Mesh, U, B, time: float, gamma: float, variables=None, scaling: float = 1.0,
               xmin=(0.0, 0.0, 0.0)) -> Path:
    """``<prefix>.xmf`` + ``<prefix>_cells.bin``: one co-rectilinear grid per owned leaf with the
    requested cell variables (default: all); returns the .xmf path."""
    prefix = Path(prefix)
    names = list(variables) if variables is not None else CELL_VARS + DERIVED
    unknown = [v for v in names if v not in CELL_VARS + DERIVED]
    if unknown:
        raise MhdctError(f"unknown output variables {unknown}")
    fields = cell_fields(mesh, U, B, gamma, scaling)
    n, b, dim = mesh.num_leaves, mesh.b, mesh.dim
    ncell = mesh.cells_per_block
    data = np.stack([fields[v] for v in names], axis=1).astype("<f8")  # [leaf, var, cell]
    binp = prefix.parent / (prefix.name + "_cells.bin")
    data.tofile(binp)
    lv, co = mesh.leaves()
    cdims = " ".join([str(b)] * dim)
    pdims = " ".join([str(b + 1)] * dim)
    topo = "3DCoRectMesh" if dim == 3 else "2DCoRectMesh"
    geo = "ORIGIN_DXDYDZ" if dim == 3 else "ORIGIN_DXDY"
    lines = ['<?xml version="1.0" ?>', '<Xdmf Version="2.0">', " <Domain>",
             f'  <Grid Name="leaves" GridType="Collection" CollectionType="Spatial">',
             f'   <Time Value="{float(time)!r}"/>']
    for o in range(n):
        h = scaling / (b * (1 << int(lv[o])))
        origin = [float(xmin[d] + co[o][d] * b * h) for d in range(dim)][::-1]  # slowest axis first
        lines.append(f'   <Grid Name="leaf{o}" GridType="Uniform">')
        lines.append(f'    <Topology TopologyType="{topo}" Dimensions="{pdims}"/>')
        lines.append(f'    <Geometry GeometryType="{geo}">')
        lines.append(f'     <DataItem Format="XML" Dimensions="{dim}">{" ".join(repr(x) for x in origin)}</DataItem>')
        lines.append(f'     <DataItem Format="XML" Dimensions="{dim}">{" ".join([repr(float(h))] * dim)}</DataItem>')
        lines.append("    </Geometry>")
        for v, name in enumerate(names):
            seek = ((o * len(names) + v) * ncell) * 8
            lines.append(f'    <Attribute Name="{name}" AttributeType="Scalar" Center="Cell">')
            lines.append(f'     <DataItem Format="Binary" NumberType="Float" Precision="8" Endian="Little" '
                         f'Seek="{seek}" Dimensions="{cdims}">{binp.name}</DataItem>')
            lines.append("    </Attribute>")
        lines.append("   </Grid>")
    lines += ["  </Grid>", " </Domain>", "</Xdmf>"]
    path = _named(prefix, ".xmf")
    path.write_text("\n".join(lines) + "\n")
    return path


def save_solver(prefix, solver, write_vis: bool = True, variables=None, xmin=(0.0, 0.0, 0.0)) -> Path:
    """save_solution (SolverGodunovMHD.cpp:856-1042) for a SolverGodunovMHD: the checkpoint and,
    with write_vis, the XDMF output of the same state; returns the checkpoint header path."""
    U, B = solver.state()
    st = solver.settings
    path = save_checkpoint(prefix, solver.mesh, U, B, solver.t, solver.iteration, settings=st,
                           extra={"amr_cycles": solver.amr_cycles,
                                  "total_num_cell_updates": solver.total_num_cell_updates})
    if write_vis:
        write_xdmf(prefix, solver.mesh, U, B, solver.t, st.gamma0, variables, st.scaling_factor, xmin)
    return path


def restart_solver(prefix, settings=None, impl_version=None, device: str = "cuda"):
    """init_restart (init/init.cpp:107-193): a SolverGodunovMHD holding a checkpoint's mesh, state,
    time and iteration. settings default to the checkpoint's own."""
    from .implem import IMPL_PLAIN
    from .solver import SolverGodunovMHD

    mesh, U, B, head = load_checkpoint(prefix)
    if settings is None:
        if head.get("settings") is None:
            raise MhdctError("the checkpoint holds no settings: pass them")
        settings = settings_from_dict(head["settings"])
    sol = SolverGodunovMHD(mesh, U, B, settings, impl_version=IMPL_PLAIN if impl_version is None else impl_version,
                           device=device)
    sol.t = float(head["time"])
    sol.iteration = int(head["iteration"])
    ex = head.get("extra") or {}
    sol.amr_cycles = int(ex.get("amr_cycles", 0))
    sol.total_num_cell_updates = int(ex.get("total_num_cell_updates", 0))
    return sol
