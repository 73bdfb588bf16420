"""CPU tests of checkpoint, restart and visualisation output (godunov_mhd_ct_amd/io.py; the
reference's save_solution_hdf5, SolverGodunovMHD.cpp:856-1042, and init_restart,
init/init.cpp:107-193):

* Mesh.set_leaves rebuilds a refined mesh exactly (leaves, outside blocks, neighbour-derived
  counts) and rejects a leaf set that is not a tiling (duplicate, nested, gap, outside);
* a checkpoint reads back bit for bit, also when the file lists its leaves in another order;
* the XDMF index points at the right bytes: every grid's origin and spacing are the leaf's,
  every attribute's Seek reads that leaf's values; divB is at roundoff for a field made from a
  vector potential and the thermal pressure is the initial one.
"""
import json
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from godunov_mhd_ct_amd import MhdctError, problems
from godunov_mhd_ct_amd.implem import hydro_settings
from godunov_mhd_ct_amd.io import (CELL_VARS, DERIVED, load_checkpoint, save_checkpoint, settings_dict,
                                   settings_from_dict, write_xdmf)
from godunov_mhd_ct_amd.mesh import Mesh, refine_where


def refined(dim, b=8, nbrick=(1, 1, 1), periodic=(1, 1, 1), bc=None):
    m = Mesh(dim, b, 1, nbrick=nbrick, periodic=periodic)
    if bc is not None:
        m.set_bc(bc)
    # a central box: refinement stays clear of the non-periodic borders
    pred = (lambda L, x0, x1, y0, y1, z0, z1: x1 > 0.3 * nbrick[0] and x0 < 0.7 * nbrick[0]
            and y1 > 0.3 * nbrick[1] and y0 < 0.7 * nbrick[1])
    return refine_where(m, pred, 3)


def random_state(m, seed=0):
    rng = np.random.default_rng(seed)
    U = rng.standard_normal((m.num_total, 5, m.cells_per_block))
    B = rng.standard_normal((m.num_total, m.faces_per_block))
    return U, B


@pytest.mark.parametrize("dim", [2, 3])
def test_set_leaves_rebuilds_the_mesh(dim):
    m = refined(dim, nbrick=(2, 1, 1), periodic=(0, 1, 1), bc=[3, 3, 0, 0, 0, 0])
    lv, co = m.leaves()
    n = m.num_leaves
    r = Mesh(dim, m.b, 0, nbrick=m.nbrick, periodic=m.periodic)
    r.set_bc(m.bc)
    perm = np.random.default_rng(1).permutation(n)
    r.set_leaves(lv[:n][perm], co[:n][perm])
    r.finalize()
    l2, c2 = r.leaves()
    assert r.num_leaves == n and r.num_outside == m.num_outside
    assert np.array_equal(l2, lv) and np.array_equal(c2, co)
    assert [np.array_equal(a, b) for a, b in zip(r.outside(), m.outside())] == [True] * len(m.outside())


def test_set_leaves_rejects_non_tilings():
    m = refined(2)
    lv, co = m.leaves()
    n = m.num_leaves
    lv, co = lv[:n].copy(), co[:n].copy()
    bad = {
        "duplicate": (np.append(lv, lv[0]), np.vstack([co, co[:1]])),
        "gap": (lv[1:], co[1:]),
        "nested": (np.append(lv, 0), np.vstack([co, [[0, 0, 0]]])),
        "outside": (np.append(lv[1:], lv[0]), np.vstack([co[1:], [[co[0, 0] + (1 << 12), co[0, 1], 0]]])),
        "level": (np.append(lv[1:], 30), np.vstack([co[1:], co[:1]])),
    }
    for name, (l, c) in bad.items():
        r = Mesh(2, 8, 0)
        with pytest.raises(MhdctError):
            r.set_leaves(l, c)
    # the mesh keeps working after a rejected set
    r = Mesh(2, 8, 0)
    with pytest.raises(MhdctError):
        r.set_leaves(lv[1:], co[1:])
    r.set_leaves(lv, co)
    r.finalize()
    assert r.num_leaves == n


@pytest.mark.parametrize("dim", [2, 3])
def test_checkpoint_round_trip(tmp_path, dim):
    m = refined(dim, nbrick=(1, 2, 1), periodic=(1, 0, 1), bc=[0, 0, 3, 2, 0, 0])
    U, B = random_state(m)
    st = hydro_settings(gamma0=1.4, cfl=0.6, time_integrator="RK2_SSP", gravity=(0.0, -0.5, 0.0)[:dim])
    save_checkpoint(tmp_path / "chk", m, U, B, 0.125, 37, settings=st, extra={"note": "x"})
    m2, U2, B2, head = load_checkpoint(tmp_path / "chk")
    n = m.num_leaves
    assert head["time"] == 0.125 and head["iteration"] == 37 and head["extra"] == {"note": "x"}
    assert m2.bc == m.bc and m2.nbrick == m.nbrick and m2.periodic == m.periodic
    assert m2.num_leaves == n and m2.num_outside == m.num_outside
    assert np.array_equal(m2.leaves()[0], m.leaves()[0]) and np.array_equal(m2.leaves()[1], m.leaves()[1])
    assert np.array_equal(U2[:n], U[:n]) and np.array_equal(B2[:n], B[:n])
    assert not U2[n:].any() and not B2[n:].any()
    st2 = settings_from_dict(head["settings"])
    assert settings_dict(st2) == settings_dict(st)


def test_dotted_prefix_keeps_every_file_apart(tmp_path):
    """A prefix with dots in its name ('kh.mu1e-4_iterNNN') names its files by appending: two
    checkpoints of one run do not overwrite each other, and the XDMF data file sits beside its
    index under the same name."""
    m = refined(2, b=4)
    U, B = random_state(m, seed=1)
    U2, B2 = random_state(m, seed=2)
    save_checkpoint(tmp_path / "kh.mu1e-4_iter0000010", m, U, B, 0.1, 10)
    save_checkpoint(tmp_path / "kh.mu1e-4_iter0000020", m, U2, B2, 0.2, 20)
    for it in (10, 20):
        assert (tmp_path / f"kh.mu1e-4_iter00000{it}.json").exists()
        assert (tmp_path / f"kh.mu1e-4_iter00000{it}.bin").exists()
    assert not (tmp_path / "kh.json").exists() and not (tmp_path / "kh.bin").exists()
    n = m.num_leaves
    _, Ua, _, ha = load_checkpoint(tmp_path / "kh.mu1e-4_iter0000010")
    _, Ub, _, hb = load_checkpoint(tmp_path / "kh.mu1e-4_iter0000020")
    assert ha["iteration"] == 10 and hb["iteration"] == 20
    assert np.array_equal(Ua[:n], U[:n]) and np.array_equal(Ub[:n], U2[:n])
    path = write_xdmf(tmp_path / "kh.mu1e-4_iter0000020", m, U2, B2, 0.2, 1.4)
    assert path == tmp_path / "kh.mu1e-4_iter0000020.xmf"
    assert (tmp_path / "kh.mu1e-4_iter0000020_cells.bin").exists()
    assert "kh.mu1e-4_iter0000020_cells.bin" in path.read_text()


def test_checkpoint_with_leaves_in_another_order(tmp_path):
    m = refined(3)
    U, B = random_state(m, 4)
    save_checkpoint(tmp_path / "chk", m, U, B, 1.0, 2)
    head = json.loads((tmp_path / "chk.json").read_text())
    raw = np.fromfile(tmp_path / "chk.bin", dtype=np.uint8)
    n = m.num_leaves
    perm = np.random.default_rng(3).permutation(n)
    for ent in head["arrays"]:
        dt = np.dtype(ent["dtype"])
        cnt = int(np.prod(ent["shape"]))
        view = raw[ent["offset"]:ent["offset"] + cnt * dt.itemsize].view(dt).reshape(ent["shape"])
        view[...] = view[perm].copy()
    raw.tofile(tmp_path / "chk.bin")
    _, U2, B2, _ = load_checkpoint(tmp_path / "chk")
    assert np.array_equal(U2[:n], U[:n]) and np.array_equal(B2[:n], B[:n])


def test_checkpoint_errors(tmp_path):
    m = refined(2)
    U, B = random_state(m)
    save_checkpoint(tmp_path / "chk", m, U, B, 0.0, 0)
    raw = (tmp_path / "chk.bin").read_bytes()
    (tmp_path / "chk.bin").write_bytes(raw[:-8])
    with pytest.raises(MhdctError, match="truncated"):
        load_checkpoint(tmp_path / "chk")
    head = json.loads((tmp_path / "chk.json").read_text())
    head["version"] = 99
    (tmp_path / "chk.json").write_text(json.dumps(head))
    with pytest.raises(MhdctError):
        load_checkpoint(tmp_path / "chk")


def read_xdmf(path):
    root = ET.parse(path).getroot()
    coll = root.find("Domain/Grid")
    return coll, coll.findall("Grid")


@pytest.mark.parametrize("dim", [2, 3])
def test_xdmf_points_at_the_right_bytes(tmp_path, dim):
    m = refined(dim, b=4)
    gamma, scaling = 1.4, 2.0
    if dim == 2:
        U, B = problems.orszag_tang_2d(m, gamma=gamma, scaling=scaling)
        p0 = gamma / (4.0 * np.pi)  # the Orszag-Tang initial pressure
    else:
        U, B = problems.random_mhd_3d(m, seed=2, gamma=gamma, nmodes=4, device="cpu")
        p0 = None
    U = np.asarray(U).reshape(m.num_total, 5, -1)
    B = np.asarray(B).reshape(m.num_total, -1)
    xmin = (-1.0, 0.5, 0.25)
    path = write_xdmf(tmp_path / "out", m, U, B, 0.75, gamma, scaling=scaling, xmin=xmin)
    coll, grids = read_xdmf(path)
    n, b = m.num_leaves, m.b
    assert len(grids) == n and float(coll.find("Time").get("Value")) == 0.75
    names = CELL_VARS + DERIVED
    data = np.fromfile(tmp_path / "out_cells.bin", dtype="<f8")
    lv, co = m.leaves()
    divmax = 0.0
    for o, g in enumerate(grids):
        geo = g.find("Geometry").findall("DataItem")
        origin = np.array(geo[0].text.split(), dtype=float)[::-1]
        h = np.array(geo[1].text.split(), dtype=float)
        hh = scaling / (b * (1 << int(lv[o])))
        assert np.allclose(h, hh, rtol=0, atol=1e-15)
        assert np.allclose(origin, [xmin[d] + co[o][d] * b * hh for d in range(dim)], rtol=0, atol=1e-14)
        assert g.find("Topology").get("Dimensions") == " ".join([str(b + 1)] * dim)
        attrs = {a.get("Name"): a.find("DataItem") for a in g.findall("Attribute")}
        assert list(attrs) == names
        for name, item in attrs.items():
            seek = int(item.get("Seek")) // 8
            vals = data[seek:seek + b ** dim]
            if name in CELL_VARS:
                assert np.array_equal(vals, U[o, CELL_VARS.index(name)])
            elif name == "divB":
                divmax = max(divmax, np.abs(vals).max())
            elif name == "thermal_pressure" and p0 is not None:
                assert np.allclose(vals, p0, rtol=1e-12, atol=0)
            elif name == "Bx":
                want = 0.5 * (B[o, :(b + 1) * b ** (dim - 1)].reshape(-1, b + 1)[:, :-1]
                              + B[o, :(b + 1) * b ** (dim - 1)].reshape(-1, b + 1)[:, 1:]).reshape(-1)
                assert np.array_equal(vals, want)
    bscale = np.abs(B[:n]).max() / (scaling / (b * (1 << int(lv[:n].max()))))
    assert divmax <= 1e-12 * bscale, divmax


def test_xdmf_rejects_unknown_variables(tmp_path):
    m = refined(2)
    U, B = random_state(m)
    with pytest.raises(MhdctError):
        write_xdmf(tmp_path / "o", m, U, B, 0.0, 1.4, variables=["rho", "velocity"])
    path = write_xdmf(tmp_path / "o", m, U, B, 0.0, 1.4, variables=["rho", "divB"])
    _, grids = read_xdmf(path)
    assert [a.get("Name") for a in grids[0].findall("Attribute")] == ["rho", "divB"]
