"""GPU parity of the AMR cycle (SolverGodunovMHD::do_amr_cycle, SolverGodunovMHD.cpp:263-319)
against the oracle's restatement on the same inputs:

* the refinement indicator of every owned block (bit for bit) and the flags, for 8^d and 16^3
  blocks (the latter through the wavefront-reduction path);
* the remap of U and face B to an adapted mesh (bit for bit), with leaves refined and families
  coarsened in one cycle;
* the solver's step loop with an AMR cycle every two steps, plain and fused forms, against the
  oracle driven through the same cycle (1e-12 relative L-inf per field, equal dt);
* mass, momentum and energy conserved across a cycle on the device.
"""
import numpy as np
import pytest

from godunov_mhd_ct_amd import problems
from godunov_mhd_ct_amd.amr import AMRSettings, Remap, adapted, refine_flags
from godunov_mhd_ct_amd.implem import IMPL_FUSED, IMPL_PLAIN, GodunovImplemHIP, hydro_settings
from godunov_mhd_ct_amd.mesh import Mesh, refine_where, uniform_mesh
from oracle_ref import OracleSolver, remap, settings_array
from oracle_ref import refine_flags as oracle_flags
from test_gpu_parity import compare

GAMMA = 5.0 / 3.0


def amr2d(b=8):
    m = Mesh(2, b, 2)
    return refine_where(m, lambda L, x0, x1, y0, y1, z0, z1: x0 < 0.6 and x1 > 0.4 and y0 < 0.6 and y1 > 0.4, 3)


def amr3d(b=8):
    m = Mesh(3, b, 1)
    return refine_where(m, lambda L, x0, x1, y0, y1, z0, z1: x0 < 0.5 and y0 < 0.5 and z0 < 0.5, 2)


def state(m, seed=0):
    if m.dim == 3:
        U, B = problems.random_mhd_3d(m, seed=seed, nmodes=6, device="cpu")
    else:
        U, B = problems.blast(m, radius=0.2, center=(0.45, 0.5, 0.0), rho_in=2.0, bfield=(0.5, 0.3, 0.2))
    return U.reshape(m.num_total, 5, -1).copy(), B.reshape(m.num_total, -1).copy()


def dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a).reshape(-1)).cuda()


CASES = [("amr2d_b8", amr2d), ("amr3d_b8", amr3d), ("uniform3d_b16", lambda: uniform_mesh(3, 16, 1))]


@pytest.mark.gpu
@pytest.mark.parametrize("name,mk", CASES)
def test_refine_flags_match_oracle(name, mk):
    print(f"[amr] refine flags {name}", flush=True)
    m = mk()
    U, B = state(m, 1)
    amr = AMRSettings(level_min=1, level_max=3, epsilon_refine=0.05, epsilon_coarsen=0.01, epsilon_lohner=0.02)
    im = GodunovImplemHIP(hydro_settings(gamma0=GAMMA), m.dim, m.b, IMPL_PLAIN)
    im.resize_auxiliary_data(m)
    flags, ind = refine_flags(im, dev(U), amr)
    ora = OracleSolver(m, settings_array(gamma0=GAMMA))
    emax = ora.refine_indicator(U, ivar=0, eps=amr.epsilon_lohner)
    assert np.array_equal(ind, emax), np.abs(ind - emax).max()
    lv, _ = m.leaves()
    want = oracle_flags(lv, emax, amr.level_min, amr.level_max, amr.epsilon_refine, amr.epsilon_coarsen)
    assert np.array_equal(flags, want)
    assert ind.max() > ind.min()


def mixed_flags(m):
    lv, co = m.leaves()
    n = m.num_leaves
    x0 = co[:n, 0] / (1 << lv[:n])
    lmax, lmin = lv[:n].max(), lv[:n].min()
    f = np.where((lv[:n] == lmax) & (x0 < 0.5), -1, np.where((lv[:n] == lmin) & (x0 >= 0.5), 1, 0))
    return f.astype(np.int8)


@pytest.mark.gpu
@pytest.mark.parametrize("name,mk", CASES[:2])
def test_remap_matches_oracle(name, mk):
    import torch

    print(f"[amr] remap {name}", flush=True)
    m = mk()
    U, B = state(m, 2)
    new = adapted(m, mixed_flags(m))
    rm = Remap(m, new)
    assert rm.counts()[1] > 0 and rm.counts()[2] > 0
    Un = torch.zeros(new.num_total * 5 * new.cells_per_block, dtype=torch.float64, device="cuda")
    Bn = torch.zeros(new.num_total * new.faces_per_block, dtype=torch.float64, device="cuda")
    rm.apply(dev(U), dev(B), Un, Bn)
    torch.cuda.synchronize()
    Uo, Bo = remap(m, new, U, B)
    n = new.num_leaves
    Ug = Un.cpu().numpy().reshape(new.num_total, 5, -1)
    Bg = Bn.cpu().numpy().reshape(new.num_total, -1)
    assert np.array_equal(Ug[:n], Uo[:n])
    assert np.array_equal(Bg[:n], Bo[:n]), np.abs(Bg[:n] - Bo[:n]).max()


def oracle_cycle(m, Uo, Bo, amr, st, ora):
    if m.num_outside:
        ora.fill_outside(Uo, Bo)
    p = amr.params()
    emax = ora.refine_indicator_vars(Uo, Bo, list(p.vars)[:p.nvars], criterion=p.criterion, eps=amr.epsilon_lohner)
    lv, _ = m.leaves()
    flags = oracle_flags(lv, emax, amr.level_min, amr.level_max, amr.epsilon_refine, amr.epsilon_coarsen)
    if not flags.any():
        return m, Uo, Bo, ora
    new = adapted(m, flags)
    Un, Bn = remap(m, new, Uo, Bo)
    return new, Un, Bn, OracleSolver(new, st, nthreads=4)


@pytest.mark.gpu
@pytest.mark.parametrize("impl", [IMPL_PLAIN, IMPL_FUSED])
@pytest.mark.parametrize("dim", [2, 3])
def test_solver_with_amr_cycles_matches_oracle(impl, dim):
    import torch
    from godunov_mhd_ct_amd.solver import SolverGodunovMHD

    m = Mesh(dim, 8, 2 if dim == 2 else 1).finalize()
    if dim == 2:
        U, B = problems.blast(m, radius=0.15, center=(0.5, 0.5, 0.0), rho_in=2.0, bfield=(0.5, 0.3, 0.2))
    else:
        U, B = problems.blast(m, radius=0.3, center=(0.5, 0.5, 0.5), rho_in=2.0, bfield=(0.4, 0.3, 0.2))
    amr = AMRSettings(level_min=1, level_max=3 if dim == 2 else 2, epsilon_refine=0.05, epsilon_coarsen=0.005,
                      cycle_frequency=2)
    print(f"[amr] solver impl={impl} dim={dim} with AMR cycles", flush=True)
    st = hydro_settings(gamma0=GAMMA, cfl=0.8, slope_type=1.0)
    sol = SolverGodunovMHD(m, U, B, st, impl_version=impl)
    sa = settings_array(gamma0=GAMMA, cfl=0.8, slope_type=1.0)
    ora = OracleSolver(m, sa, nthreads=4)
    om, Uo, Bo = m, U.reshape(m.num_total, 5, -1).copy(), B.reshape(m.num_total, -1).copy()
    changed = 0
    for it in range(6):
        if sol.should_do_amr_cycle(amr):
            counts = sol.do_amr_cycle(amr)
            om, Uo, Bo, ora = oracle_cycle(om, Uo, Bo, amr, sa, ora)
            assert sol.mesh.num_leaves == om.num_leaves
            assert np.array_equal(sol.mesh.leaves()[0], om.leaves()[0])
            changed += counts is not None
        dto = ora.compute_dt(Uo, Bo)
        dtg = sol.compute_dt()
        assert abs(dtg - dto) <= 1e-14 * dto, (dtg, dto)
        Uo, Bo, _ = ora.step(Uo, Bo, dt=dto)
        sol.step(dt=dto)
        torch.cuda.synchronize()
        Ug, Bg = sol.state()
        compare(Ug, Bg, Uo, Bo, om.num_leaves, f"iteration {it}")
    assert changed >= 2, "the cycles did not change the mesh"


@pytest.mark.gpu
def test_amr_cycle_conserves_on_device():
    import torch
    from godunov_mhd_ct_amd.solver import SolverGodunovMHD
    from test_oracle_invariants import cell_volumes

    m = amr3d()
    U, B = state(m, 3)
    print("[amr] conservation across a device cycle, 3D", flush=True)
    sol = SolverGodunovMHD(m, U, B, hydro_settings(gamma0=GAMMA), impl_version=IMPL_FUSED)
    amr = AMRSettings(level_min=1, level_max=3, epsilon_refine=0.1, epsilon_coarsen=0.05)
    U0, _ = sol.state()
    t0 = (U0[:m.num_leaves] * cell_volumes(m)[:, None, None]).sum(axis=(0, 2))
    counts = sol.do_amr_cycle(amr)
    torch.cuda.synchronize()
    assert counts is not None and counts[1] + counts[2] > 0
    U1, _ = sol.state()
    nm = sol.mesh
    t1 = (U1[:nm.num_leaves] * cell_volumes(nm)[:, None, None]).sum(axis=(0, 2))
    scale = np.abs(U0[:m.num_leaves] * cell_volumes(m)[:, None, None]).sum(axis=(0, 2)).max()
    assert np.all(np.abs(t1 - t0) <= 1e-13 * scale), t1 - t0


@pytest.mark.gpu
@pytest.mark.parametrize("impl", [IMPL_PLAIN, IMPL_FUSED])
def test_amr_cycles_with_walls_and_gravity(impl):
    # the Rayleigh-Taylor configuration (walls all around, a 1 x 3 brick, gravity): cycles
    # refine leaves that touch the walls, so leaves of several levels meet a wall
    import torch
    from godunov_mhd_ct_amd.solver import SolverGodunovMHD

    walls = [3] * 6
    m = Mesh(2, 8, 1, nbrick=(1, 3, 1), periodic=(0, 0, 0))
    m.set_bc(walls)
    m.finalize()
    g = (0.0, -0.1)
    U, B = problems.rayleigh_taylor(m, g, gamma=1.4, amplitude=0.05, bfield=(0.0, 0.5, 0.0), xmin=(-0.5, -1.5, 0.0))
    amr = AMRSettings(level_min=1, level_max=3, epsilon_refine=0.01, epsilon_coarsen=0.001, cycle_frequency=2)
    print(f"[amr] impl={impl} Rayleigh-Taylor with walls, gravity and AMR cycles", flush=True)
    st = hydro_settings(gamma0=1.4, cfl=0.7, slope_type=2.0, gravity=g)
    sol = SolverGodunovMHD(m, U, B, st, impl_version=impl)
    sa = settings_array(gamma0=1.4, cfl=0.7, slope_type=2.0, gravity=g)
    ora = OracleSolver(m, sa, nthreads=4)
    om, Uo, Bo = m, U.reshape(m.num_total, 5, -1).copy(), B.reshape(m.num_total, -1).copy()
    wall_levels = set()
    for it in range(6):
        if sol.should_do_amr_cycle(amr):
            sol.do_amr_cycle(amr)
            om, Uo, Bo, ora = oracle_cycle(om, Uo, Bo, amr, sa, ora)
            assert np.array_equal(sol.mesh.leaves()[0], om.leaves()[0])
            src, _ = om.outside()
            wall_levels |= set(om.leaves()[0][src].tolist())
        dto = ora.compute_dt(Uo, Bo)
        dtg = sol.compute_dt()
        assert abs(dtg - dto) <= 1e-14 * dto, (dtg, dto)
        Uo, Bo, _ = ora.step(Uo, Bo, dt=dto)
        sol.step(dt=dto)
        torch.cuda.synchronize()
        Ug, Bg = sol.state()
        compare(Ug, Bg, Uo, Bo, om.num_leaves, f"iteration {it}")
    assert len(wall_levels) >= 2, "no refinement reached a wall"


def varied_state(m):
    """A state in which every refinement variable varies: 3D the random one; 2D Orszag-Tang (flow
    and a varying field) with a density bump (momenta scaled with it, so velocities and pressure
    are kept)."""
    if m.dim == 3:
        return state(m, 4)
    U, B = problems.orszag_tang_2d(m)
    U, B = U.reshape(m.num_total, 5, -1).copy(), B.reshape(m.num_total, -1).copy()
    xs = problems._grid(m, 1.0, (0.0, 0.0, 0.0), "cell")
    f = 1.0 + 0.5 * np.exp(-((xs[0] - 0.45) ** 2 + (xs[1] - 0.5) ** 2) / 0.02)
    ek = 0.5 * (U[:, 2] ** 2 + U[:, 3] ** 2 + U[:, 4] ** 2) / U[:, 0]
    U[:, 0] *= f
    U[:, 2:5] *= f[:, None, :]
    U[:, 1] += (f - 1.0) * ek
    return U, B


CRITERIA_CASES = [
    ("SIMPLE_GRADIENT", ["rho"]),
    ("SIMPLE_GRADIENT", ["rho", "thermal_pressure"]),
    ("LOHNER_UNSPLIT", ["magnetic_pressure", "local_mach_number"]),
    ("SIMPLE_GRADIENT", ["speed_of_sound", "specific_ekin", "vx", "vz"]),
    ("LOHNER_UNSPLIT", ["e_tot", "vy"]),
]


@pytest.mark.gpu
@pytest.mark.parametrize("crit,variables", CRITERIA_CASES)
@pytest.mark.parametrize("name,mk", CASES)
def test_refine_criteria_and_derived_variables_match_oracle(name, mk, crit, variables):
    """SIMPLE_GRADIENT and LOHNER_UNSPLIT over conserved and derived variables (several at once):
    the device indicators equal the oracle's bit for bit, and so do the flags."""
    print(f"[amr] {crit} {variables} {name}", flush=True)
    m = mk()
    U, B = varied_state(m)
    amr = AMRSettings(level_min=1, level_max=3, epsilon_refine=0.1, epsilon_coarsen=0.01, epsilon_lohner=0.02,
                      variable=variables, criterion=crit)
    im = GodunovImplemHIP(hydro_settings(gamma0=GAMMA), m.dim, m.b, IMPL_PLAIN)
    im.resize_auxiliary_data(m)
    flags, ind = refine_flags(im, dev(U), amr, B=dev(B))
    p = amr.params()
    ora = OracleSolver(m, settings_array(gamma0=GAMMA))
    emax = ora.refine_indicator_vars(U, B, list(p.vars)[:p.nvars], criterion=p.criterion, eps=amr.epsilon_lohner)
    assert np.array_equal(ind, emax), np.abs(ind - emax).max()
    lv, _ = m.leaves()
    want = oracle_flags(lv, emax, amr.level_min, amr.level_max, amr.epsilon_refine, amr.epsilon_coarsen)
    assert np.array_equal(flags, want)
    assert ind.max() > ind.min()


@pytest.mark.gpu
def test_derived_variable_with_walls_matches_oracle():
    """Outside blocks read as zero for a derived quantity (the reference fills its derived array
    for the owned blocks only), as in the oracle."""
    bc = [2, 3, 2, 2, 0, 0]
    m = Mesh(2, 8, 2, periodic=(0, 0, 1))
    m.set_bc(bc)
    m = refine_where(m, lambda L, x0, x1, y0, y1, z0, z1: x0 < 0.3 and y0 < 0.3, 3)
    U, B = state(m, 5)
    amr = AMRSettings(level_min=1, level_max=4, epsilon_refine=0.2, variable=["thermal_pressure", "rho"],
                      criterion="SIMPLE_GRADIENT")
    im = GodunovImplemHIP(hydro_settings(gamma0=GAMMA), m.dim, m.b, IMPL_PLAIN)
    im.resize_auxiliary_data(m)
    Ud, Bd = dev(U), dev(B)
    im.fill_outside(Ud, Bd)
    flags, ind = refine_flags(im, Ud, amr, B=Bd)
    ora = OracleSolver(m, settings_array(gamma0=GAMMA), bc=bc)
    ora.fill_outside(U, B)
    emax = ora.refine_indicator_vars(U, B, [10, 0], criterion=1)
    assert np.array_equal(ind, emax)
    assert (ind == 1.0).any()


@pytest.mark.gpu
@pytest.mark.parametrize("impl", [IMPL_PLAIN, IMPL_FUSED])
def test_solver_cycles_on_density_and_pressure_gradient(impl):
    """The blast configuration's criterion: SIMPLE_GRADIENT on rho and the thermal pressure,
    normalised jump > 0.1, cycles every two steps, against the oracle driven through the same
    cycles (1e-12 per field, equal dt)."""
    import torch
    from godunov_mhd_ct_amd.solver import SolverGodunovMHD

    m = Mesh(3, 8, 1).finalize()
    U, B = problems.blast(m, radius=0.3, center=(0.5, 0.5, 0.5), p_in=10.0, p_out=0.1, bfield=(0.4, 0.3, 0.2))
    amr = AMRSettings(level_min=1, level_max=2, epsilon_refine=0.1, epsilon_coarsen=0.01, cycle_frequency=2,
                      variable=["rho", "thermal_pressure"], criterion="SIMPLE_GRADIENT")
    print(f"[amr] solver impl={impl} SIMPLE_GRADIENT on rho and p", flush=True)
    st = hydro_settings(gamma0=GAMMA, cfl=0.8, slope_type=1.0)
    sol = SolverGodunovMHD(m, U, B, st, impl_version=impl)
    sa = settings_array(gamma0=GAMMA, cfl=0.8, slope_type=1.0)
    ora = OracleSolver(m, sa, nthreads=4)
    om, Uo, Bo = m, U.reshape(m.num_total, 5, -1).copy(), B.reshape(m.num_total, -1).copy()
    changed = 0
    for it in range(4):
        if sol.should_do_amr_cycle(amr):
            counts = sol.do_amr_cycle(amr)
            om, Uo, Bo, ora = oracle_cycle(om, Uo, Bo, amr, sa, ora)
            assert np.array_equal(sol.mesh.leaves()[0], om.leaves()[0])
            changed += counts is not None
        dto = ora.compute_dt(Uo, Bo)
        dtg = sol.compute_dt()
        assert abs(dtg - dto) <= 1e-14 * dto, (dtg, dto)
        Uo, Bo, _ = ora.step(Uo, Bo, dt=dto)
        sol.step(dt=dto)
        torch.cuda.synchronize()
        Ug, Bg = sol.state()
        compare(Ug, Bg, Uo, Bo, om.num_leaves, f"iteration {it}")
    assert changed >= 1, "the cycles did not change the mesh"
