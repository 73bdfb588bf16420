"""GPU runs of the .ini driver (godunov_mhd_ct_amd/app.py, the counterpart of
solver_godunov_mhd_ct --ini, app/solver_godunov_mhd_ct.cpp:55-192):

* an Orszag-Tang run with AMR cycles and outputs every 4 steps, then a run restarted from its
  step-4 output through ``[restart] enabled, filePrefix`` (init/init.cpp:107-193), ends bit for
  bit where the uninterrupted run ends, and the last step is clamped onto tEnd;
* the reference's own Rayleigh-Taylor 2D and blast 3D files (tests/golden/ini, with the output
  directory, level_max and step count overridden by keys appended to the text) run with
  gravity and walls, and with the periodic blast conserving mass, momentum and energy.
"""
from pathlib import Path

import numpy as np
import pytest

INI = Path(__file__).resolve().parent / "golden" / "ini"

OT = """
[run]
dimension=2
tEnd={tend}
nStepmax={nmax}
nOutput=4
nlog=1
[amr]
level_min=2
level_max=3
bx=8
by=8
epsilon_refine=0.05
epsilon_coarsen=0.005
cycle_frequency=2
disable_first_amr_cycle=yes
[p4est_connectivity]
nbrick_x=1
nbrick_y=1
periodic_x=1
periodic_y=1
[hydro]
gamma0=1.666
cfl=0.8
slope_type=2
problem=orszag_tang
riemann=HLLD
implementation_version=11
[output]
outputDir={out}
outputPrefix=ot
write_variables=rho,e_tot,Bx,By,divB
"""


def quiet(*a):
    print(*a, flush=True)


@pytest.mark.gpu
def test_ot_run_outputs_and_restart(tmp_path):
    from godunov_mhd_ct_amd.app import run_simulation
    from godunov_mhd_ct_amd.config import Config

    print("[app] Orszag-Tang 2D, 8 steps with outputs, then a restart from step 4", flush=True)
    full = run_simulation(Config.parse(OT.format(tend=1.0, nmax=8, out=tmp_path / "a")), log=quiet)
    assert full.iteration == 8 and full.amr_cycles >= 3
    names = sorted(p.name for p in (tmp_path / "a").glob("*.json"))
    assert names == ["ot_iter0000000.json", "ot_iter0000004.json", "ot_iter0000008.json"]
    assert (tmp_path / "a" / "ot_iter0000008.xmf").exists()
    # restart from the step-4 output; outputs go to a second directory, the checkpoint is read
    # from outputDir as the reference reads it
    rst_text = OT.format(tend=1.0, nmax=8, out=tmp_path / "a") + "[restart]\nenabled=yes\nfilePrefix=ot_iter0000004\n"
    rst = run_simulation(Config.parse(rst_text), log=quiet)
    assert rst.iteration == 8 and rst.t == full.t
    Ua, Ba = full.state()
    Ub, Bb = rst.state()
    n = full.mesh.num_leaves
    assert np.array_equal(full.mesh.leaves()[0], rst.mesh.leaves()[0])
    assert np.array_equal(Ua[:n], Ub[:n]) and np.array_equal(Ba[:n], Bb[:n])
    # tEnd inside the run: the last step lands on it
    tend = full.t * 0.6
    short = run_simulation(Config.parse(OT.format(tend=tend, nmax=100, out=tmp_path / "c")), log=quiet)
    assert short.t == tend and short.iteration < 8


@pytest.mark.gpu
def test_reference_rayleigh_taylor_2d_file(tmp_path):
    from godunov_mhd_ct_amd.app import run_simulation
    from godunov_mhd_ct_amd.config import Config

    print("[app] app/test_mhd_rayleigh_taylor_2d.ini, 6 steps", flush=True)
    text = (INI / "test_mhd_rayleigh_taylor_2d.ini").read_text()
    text += f"\n[amr]\nlevel_max=3\ncycle_frequency=3\n[output]\noutputDir={tmp_path}\n[run]\nnOutput=0\n"
    sol = run_simulation(Config.parse(text), impl="fused", log=quiet, max_steps=6)
    assert sol.iteration == 6 and sol.t > 0 and sol.settings.gravity_enabled == 1
    assert sol.mesh.num_outside > 0
    U, _ = sol.state()
    n = sol.mesh.num_leaves
    assert np.isfinite(U[:n]).all() and (U[:n, 0] > 0).all()


@pytest.mark.gpu
def test_reference_blast_3d_file_conserves(tmp_path):
    from godunov_mhd_ct_amd.app import run_simulation, volume_integrals
    from godunov_mhd_ct_amd.config import Config

    print("[app] app/test_mhd_blast_3d.ini at levels 2-3, 6 steps", flush=True)
    text = (INI / "test_mhd_blast_3d.ini").read_text()
    text += f"\n[amr]\nlevel_max=3\ncycle_frequency=2\n[output]\noutputDir={tmp_path}\n[run]\nnOutput=0\n"
    cfg = Config.parse(text)
    sol = run_simulation(cfg, impl="plain", log=quiet, max_steps=6)
    assert sol.iteration == 6 and sol.mesh.num_leaves > 64
    # the whole run, AMR cycles included, conserves the periodic box's totals
    from godunov_mhd_ct_amd.app import initial_mesh, initial_state
    from godunov_mhd_ct_amd.solver import SolverGodunovMHD

    st, amr = cfg.settings(), cfg.amr()
    m0 = initial_mesh(cfg, st, amr, 10, log=quiet)
    U0, B0 = initial_state(cfg, m0, st)
    i0 = volume_integrals(SolverGodunovMHD(m0, U0, B0, st))
    i1 = volume_integrals(sol)
    scale = np.abs(i0).max()
    assert np.all(np.abs(i1 - i0) <= 1e-12 * scale), (i1 - i0) / scale
