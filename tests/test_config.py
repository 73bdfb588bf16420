"""CPU tests of the .ini reader (mhdct_config_*, godunov_mhd_ct_amd/config.py) on the
reference's own .ini files (tests/golden/ini, copied from app/) and on small texts:

* the keys the reference reads and their defaults (models/HydroSettings.cpp:19-30,
  SolverGodunovMHD.cpp:76, 1107-1113; [gravity] as in AddGravitySourceTerm.cpp and the
  Rayleigh-Taylor files; [p4est_connectivity], [mesh] as in every app/*.ini);
* syntax: comments, blanks around '=', case-insensitive names, the last of repeated keys,
  and the line number of a malformed line;
* what is refused with a message: unknown prolongation or border condition, RK3_SSP, unequal
  block sizes, a derived refine variable, more than one refine variable, another criterion.
"""
from pathlib import Path

import pytest

from godunov_mhd_ct_amd import MhdctError
from godunov_mhd_ct_amd.config import Config
from godunov_mhd_ct_amd.implem import IMPL_PLAIN, RIEMANN, TIME_INTEGRATOR

INI = Path(__file__).resolve().parent / "golden" / "ini"


def test_every_reference_file_maps():
    files = sorted(INI.glob("*.ini"))
    assert len(files) >= 7
    for f in files:
        cfg = Config.load(f)
        st = cfg.settings()
        assert st.riemann == RIEMANN["HLLD"] and st.time_integrator == TIME_INTEGRATOR["HANCOCK"], f
        amr = cfg.amr()
        assert 0 <= amr.level_min <= amr.level_max, f
        m = cfg.mesh()
        assert m.num_leaves == m.nbrick[0] * m.nbrick[1] * m.nbrick[2] << (m.dim * amr.level_min), f


def test_blast_3d():
    cfg = Config.load(INI / "test_mhd_blast_3d.ini")
    st = cfg.settings()
    assert (st.gamma0, st.cfl, st.slope_type) == (1.666, 0.8, 2.0)
    assert (st.smallr, st.smallc, st.scaling_factor) == (1e-10, 1e-10, 1.0)  # defaults
    assert st.prolongation == 0 and st.abort_when_negative_eint == 0 and st.gravity_enabled == 0
    amr = cfg.amr()
    assert (amr.level_min, amr.level_max, amr.cycle_frequency) == (2, 5, 10)
    assert (amr.epsilon_refine, amr.epsilon_coarsen, amr.epsilon_lohner) == (0.01, 0.00125, 0.02)
    assert amr.variable == "rho"
    m = cfg.mesh().finalize()
    assert (m.dim, m.b, m.num_leaves, m.num_outside) == (3, 4, 64, 0)
    assert m.periodic == (1, 1, 1) and m.bc == (3, 3, 3, 3, 3, 3)
    assert cfg.get("hydro", "problem") == "blast" and cfg.get_real("blast", "radius", 0) == 0.25
    assert cfg.get_bool("output", "hdf5_enabled", False) is True
    assert cfg.get_string_vector("output", "write_variables", []) == ["rho", "rho_vx", "rho_vy", "e_tot", "Bx", "By",
                                                                      "Bz", "divB"]


def test_rayleigh_taylor_gravity_and_walls():
    cfg = Config.load(INI / "test_mhd_rayleigh_taylor_2d.ini")
    st = cfg.settings()
    assert st.gravity_enabled == 1 and list(st.gravity) == [0.0, -0.1, 0.0]
    m = cfg.mesh().finalize()
    assert m.nbrick[:2] == (1, 3) and m.periodic[:2] == (0, 0) and m.num_outside > 0
    # "amplitude = 0.01" (blanks around '='), "P0" read in any case
    assert cfg.get_real("rayleigh_taylor", "amplitude", 0) == 0.01
    assert cfg.get_real("Rayleigh_Taylor", "p0", 0) == 10.0
    st3 = Config.load(INI / "test_mhd_rayleigh_taylor_3d.ini").settings()
    assert list(st3.gravity) == [0.0, 0.0, -0.1]


def test_restart_file_and_cavity():
    cfg = Config.load(INI / "test_mhd_blast_2d_restart.ini")
    assert cfg.get_bool("restart", "enabled", False) and cfg.get("restart", "filePrefix") == "test_mhd_blast_2d_iter0001421"
    st = cfg.settings()
    assert st.scaling_factor == 0.5
    m = cfg.mesh()
    assert m.nbrick[:2] == (2, 3) and m.b == 8
    # the shipped cavity files write time_integrator in [amr]
    cav = Config.load(INI / "test_mhd_diamag_cavity_2d.ini")
    assert cav.get("amr", "time_integrator") == "HANCOCK"
    kh = Config.load(INI / "test_mhd_kelvin_helmholtz_3d.ini")
    assert kh.get_bool("viscosity", "enabled", True) is False
    st = kh.settings()
    assert st.viscosity_enabled == 0 and st.viscosity_mu == 1e-4 and st.viscosity_hancock_predictor == 0
    st = Config.parse(text(*BASE, "[viscosity]", "enabled=yes", "mu=0.5", "hancock_predictor_enabled=true")).settings()
    assert (st.viscosity_enabled, st.viscosity_mu, st.viscosity_hancock_predictor) == (1, 0.5, 1)


def text(*sections):
    return "\n".join(sections) + "\n"


BASE = ["[run]", "dimension=2", "[amr]", "level_min=1", "level_max=2", "bx=8", "by=8", "[hydro]", "riemann=HLLD"]


def test_defaults_and_syntax():
    cfg = Config.parse(text(*BASE, "  # a comment", "; another", "", "[HYDRO]", "GAMMA0 = 1.5", "gamma0=1.6"))
    st = cfg.settings()
    assert st.gamma0 == 1.6  # the last repeated key; names in any case
    assert (st.cfl, st.slope_type) == (0.5, 1.0)
    assert cfg.get("hydro", "missing") is None and cfg.get("nosection", "x", "d") == "d"
    cfg = Config.parse(text("[a]", "t=Yes", "f=off", "n=0x10", "r=1e-3", "bad=1.5x", "v= a , b,,c "))
    assert cfg.get_bool("a", "t", False) and not cfg.get_bool("a", "f", True)
    assert cfg.get_integer("a", "n", 0) == 16 and cfg.get_real("a", "r", 0) == 1e-3
    assert cfg.get_real("a", "bad", 7.0) == 7.0
    assert cfg.get_string_vector("a", "v", []) == ["a", "b", "c"]
    # no riemann: UNKNOWN, which the context refuses (the reference fails on it at the first flux)
    st = Config.parse(text("[hydro]", "gamma0=1.4")).settings()
    assert st.riemann == RIEMANN["UNKNOWN"]
    st = Config.parse(text(*BASE, "[run]", "time_integrator=RK2_SSP")).settings()
    assert st.time_integrator == TIME_INTEGRATOR["RK2_SSP"]


def test_malformed_lines_name_their_line():
    with pytest.raises(MhdctError, match="line 3"):
        Config.parse("[a]\nx=1\nnot a key value line\n")
    with pytest.raises(MhdctError, match="line 2"):
        Config.parse("# c\n[open\n")
    with pytest.raises(MhdctError, match="line 1"):
        Config.parse("= value\n")
    with pytest.raises(MhdctError, match="cannot be read"):
        Config.load(INI / "no_such_file.ini")


@pytest.mark.parametrize("extra,what,match", [
    (["[amr]", "cell_centered_prolongation_type=QUADRATIC"], "settings", "prolongation"),
    (["[run]", "time_integrator=RK3_SSP"], "settings", "time_integrator"),
    (["[amr]", "by=16"], "mesh", "equal"),
    (["[mesh]", "boundary_type_xmin=CUSTOM"], "mesh", "boundary_type_xmin"),
    (["[amr]", "variable_used_to_compute_refine_criterion=rho,e_tot,vx,vy,vz"], "amr", "one to four"),
    (["[amr]", "refine_criterion=QUADTREE_MAGIC"], "amr", "SIMPLE_GRADIENT"),
])
def test_refusals(extra, what, match):
    cfg = Config.parse(text(*BASE, *extra))
    with pytest.raises(MhdctError, match=match):
        getattr(cfg, what)()


def test_refine_criterion_and_derived_variables():
    """refine_criterion SIMPLE_GRADIENT; several variables, derived quantities matched in any case
    (SolverGodunovMHD.cpp:1117-1160), a name that is neither falls back to rho."""
    cfg = Config.parse(text(*BASE, "[amr]", "refine_criterion=SIMPLE_GRADIENT",
                            "variable_used_to_compute_refine_criterion=rho, Thermal_Pressure,magnetic_pressure, bogus"))
    amr = cfg.amr()
    assert amr.criterion == "SIMPLE_GRADIENT"
    assert amr.variable == ["rho", "thermal_pressure", "magnetic_pressure", "rho"]
    p = amr.params()
    assert p.criterion == 1 and p.nvars == 4 and list(p.vars) == [0, 10, 13, 0]
    assert Config.parse(text(*BASE)).amr().criterion == "LOHNER_UNSPLIT"


def test_refine_variable_names():
    cfg = Config.parse(text(*BASE, "[amr]", "variable_used_to_compute_refine_criterion=e_tot"))
    assert cfg.amr().variable == "e_tot"
    # another name falls back to rho, as SolverGodunovMHD.cpp:1149-1160 does
    cfg = Config.parse(text(*BASE, "[amr]", "variable_used_to_compute_refine_criterion=Bx"))
    assert cfg.amr().variable == "rho"


def test_missing_required_keys():
    with pytest.raises(MhdctError, match="dimension"):
        Config.parse(text("[amr]", "bx=8", "level_min=1")).mesh()
    with pytest.raises(MhdctError, match="level_min"):
        Config.parse(text("[run]", "dimension=3", "[amr]", "bx=8", "by=8", "bz=8")).mesh()
    with pytest.raises(MhdctError, match="level_min and level_max"):
        Config.parse(text("[amr]", "level_min=1")).amr()


def test_implementation_version():
    from godunov_mhd_ct_amd.app import impl_of

    assert impl_of(Config.parse(text(*BASE, "implementation_version=0"))) == IMPL_PLAIN
    assert impl_of(Config.parse(text(*BASE, "implementation_version=11"))) == 11
    assert impl_of(Config.parse(text(*BASE)), "fused") == 11
    with pytest.raises(MhdctError):
        impl_of(Config.parse(text(*BASE, "implementation_version=2")))


def test_viscosity_refused_for_the_fp32_form_before_any_device_work():
    from godunov_mhd_ct_amd.app import IMPLS, impl_of, run_simulation

    cfg = Config.parse(text(*BASE, "[viscosity]", "enabled=true", "mu=1e-4", "hancock_predictor_enabled=yes"))
    st = cfg.settings()
    assert st.viscosity_enabled == 1 and st.viscosity_mu == 1e-4 and st.viscosity_hancock_predictor == 1
    # both FP64 forms take viscosity; the FP32-bulk form is refused by the driver itself
    assert impl_of(cfg) == IMPL_PLAIN and impl_of(cfg, "fused") == IMPLS["fused"]
    with pytest.raises(MhdctError, match="FP64 forms"):
        run_simulation(cfg, impl="fp32", log=lambda *a: None, max_steps=1)
