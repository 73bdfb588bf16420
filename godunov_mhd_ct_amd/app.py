"""Run driver for an .ini file: ``python -m godunov_mhd_ct_amd.app --ini file.ini``, as
``solver_godunov_mhd_ct --ini file.ini`` runs one (run_simulation,
app/solver_godunov_mhd_ct.cpp:55-192).

* Settings, AMR thresholds and the mesh come from the .ini file (config.py, the library's
  reader).
* The start is a restart when ``[restart] enabled`` (init/init.cpp:40-46, 107-193: the
  checkpoint ``<outputDir>/<filePrefix>``). Otherwise it is the problem's initial state on the
  mesh refined from ``[amr] level_min`` towards ``level_max``.
* The loop is the solver loop around next_iteration_impl (SolverGodunovMHD.cpp:409-484), while
  ``t < tEnd`` and ``iteration < nStepmax``. Each iteration runs:
  - the AMR cycle every ``cycle_frequency`` steps (not the first with
    ``disable_first_amr_cycle``, none with ``[amr] static``);
  - an output every ``nOutput`` steps (every step for -1);
  - the dt, clamped so that the last step ends at tEnd;
  - the step.
* After the loop come a last output (unless ``nOutput = 0``), the conservation report
  (``[diagnostic] conservativity_checks``) and the cell-update rate (SolverGodunovMHD.cpp:692-720).

Outputs are ``<outputDir>/<outputPrefix>_iter<7 digits>``. Each is a checkpoint (restartable)
plus an XDMF file with the ``[output] write_variables``. An output is written after its
iteration's AMR cycle, so a run restarted from it skips that iteration's cycle and output and
goes on exactly as the run that wrote it.

The run loop, the tEnd clamp and the cadence keys come from kalypsso::core's SolverBase, which
the reference tree does not contain. Initial conditions are this package's (problems.py) for
``problem`` = blast, orszag_tang (2D) and rayleigh_taylor, read from the reference's problem
sections. The initial refinement uses the run's own refinement criterion.
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

from . import problems
from ._lib import MhdctError
from .amr import adapted, refine_flags
from .config import Config
from .implem import IMPL_FUSED, IMPL_FUSED_FP32, IMPL_PLAIN, GodunovImplemHIP
from .io import CELL_VARS, DERIVED, restart_solver, save_checkpoint, write_xdmf

IMPLS = {"plain": IMPL_PLAIN, "fused": IMPL_FUSED, "fp32": IMPL_FUSED_FP32}


def impl_of(cfg: Config, override: str | None = None) -> int:
    """[hydro] implementation_version: 0 (GodunovImplemV0) is the plain form here; 10, 11, 12
    select a form directly."""
    if override is not None:
        return IMPLS[override]
    v = cfg.get_integer("hydro", "implementation_version", 0)
    if v == 0:
        return IMPL_PLAIN
    if v in (IMPL_PLAIN, IMPL_FUSED, IMPL_FUSED_FP32):
        return v
    raise MhdctError(f"[hydro] implementation_version {v} is not known (0, 10, 11 or 12)")


def initial_state(cfg: Config, mesh, st):
    """The problem's U and face B on `mesh` (host arrays in the ABI layout)."""
    name = cfg.get("hydro", "problem", "")
    dim = mesh.dim
    xmin = cfg.xmin()
    sc = st.scaling_factor
    if name == "blast":
        ext = [mesh.nbrick[d] * sc for d in range(3)]
        centre = tuple(cfg.get_real("blast", f"center_{a}", xmin[d] + ext[d] / 2) for d, a in enumerate("xyz"))
        return problems.blast(mesh, gamma=st.gamma0, radius=cfg.get_real("blast", "radius", 0.1), center=centre,
                              rho_in=cfg.get_real("blast", "density_in", 1.0),
                              rho_out=cfg.get_real("blast", "density_out", 1.2),
                              p_in=cfg.get_real("blast", "pressure_in", 10.0),
                              p_out=cfg.get_real("blast", "pressure_out", 0.1),
                              bfield=tuple(cfg.get_real("blast", f"b{a}", 0.0) for a in "xyz"), scaling=sc,
                              xmin=xmin, test_point="vertex")
    if name == "orszag_tang" and dim == 2:
        return problems.orszag_tang_2d(mesh, gamma=st.gamma0, scaling=sc, xmin=xmin)
    if name == "rayleigh_taylor":
        if cfg.get("rayleigh_taylor", "perturbation", "sine") != "sine":
            raise MhdctError("[rayleigh_taylor] perturbation: only sine is built")
        g = tuple(st.gravity[d] for d in range(dim))
        return problems.rayleigh_taylor(mesh, g, gamma=st.gamma0,
                                        rho_up=cfg.get_real("rayleigh_taylor", "rho_up", 2.0),
                                        rho_down=cfg.get_real("rayleigh_taylor", "rho_down", 1.0),
                                        P0=cfg.get_real("rayleigh_taylor", "P0", 10.0),
                                        amplitude=cfg.get_real("rayleigh_taylor", "amplitude", 0.01),
                                        bfield=tuple(cfg.get_real("rayleigh_taylor", f"b{a}", 0.0) for a in "xyz"),
                                        scaling=sc, xmin=xmin)
    raise MhdctError(f"problem {name!r} (dimension {dim}): no initial condition is built for it here; "
                     "build U and B and use SolverGodunovMHD directly")


def initial_mesh(cfg: Config, st, amr, impl: int, log=print):
    """The mesh at [amr] level_min refined towards level_max, one level per pass. Each pass
    refines where the run's criterion asks for it on the problem's state
    (refine_criterion_at_init ALWAYS_REFINE: everywhere)."""
    import torch

    mesh = cfg.mesh().finalize()
    always = cfg.get("amr", "refine_criterion_at_init", "GEOMETRIC") == "ALWAYS_REFINE"
    im = GodunovImplemHIP(st, mesh.dim, mesh.b, impl)
    for _ in range(amr.level_max - amr.level_min):
        U, B = initial_state(cfg, mesh, st)
        if always:
            flags = np.ones(mesh.num_leaves, dtype=np.int8)
        else:
            im.resize_auxiliary_data(mesh)
            Ud = torch.from_numpy(np.ascontiguousarray(U).reshape(-1)).cuda()
            Bd = torch.from_numpy(np.ascontiguousarray(B).reshape(-1)).cuda()
            if mesh.num_outside:
                im.fill_outside(Ud, Bd)
            flags, _ = refine_flags(im, Ud, amr, B=Bd)
            flags = np.where(flags > 0, 1, 0).astype(np.int8)
        if not flags.any():
            break
        mesh = adapted(mesh, flags)
        log(f"[init] refined to {mesh.num_leaves} leaves")
    return mesh


def volume_integrals(sol):
    """Volume integrals of rho, E, rho*u, rho*v, rho*w over the owned leaves."""
    U, _ = sol.state()
    m = sol.mesh
    n = m.num_leaves
    lv, _ = m.leaves()
    vol = (sol.settings.scaling_factor / (m.b * (1 << lv[:n]).astype(np.float64))) ** m.dim
    return (U[:n] * vol[:, None, None]).sum(axis=(0, 2))


class Output:
    def __init__(self, cfg: Config, gamma: float):
        self.dir = Path(cfg.get("output", "outputDir", "./"))
        self.prefix = cfg.get("output", "outputPrefix", "output")
        want = cfg.get_string_vector("output", "write_variables", [])
        self.variables = [v for v in want if v in CELL_VARS + DERIVED] or None
        self.skipped = [v for v in want if v not in CELL_VARS + DERIVED]
        self.gamma = gamma
        self.xmin = cfg.xmin()

    def save(self, sol):
        self.dir.mkdir(parents=True, exist_ok=True)
        base = self.dir / f"{self.prefix}_iter{sol.iteration:07d}"
        U, B = sol.state()
        save_checkpoint(base, sol.mesh, U, B, sol.t, sol.iteration, settings=sol.settings,
                        extra={"amr_cycles": sol.amr_cycles, "total_num_cell_updates": sol.total_num_cell_updates})
        write_xdmf(base, sol.mesh, U, B, sol.t, self.gamma, self.variables, sol.settings.scaling_factor, self.xmin)
        return base


def run_simulation(ini, impl: str | None = None, log=print, max_steps: int | None = None):
    """Run the .ini file's simulation on the GPU; returns the solver at the end."""
    import torch
    from .solver import SolverGodunovMHD

    cfg = ini if isinstance(ini, Config) else Config.load(ini)
    st = cfg.settings()
    amr = cfg.amr()
    iv = impl_of(cfg, impl)
    if st.viscosity_enabled and iv == IMPL_FUSED_FP32:
        raise MhdctError("[viscosity] enabled: viscous fluxes run in the FP64 forms (implementation_version 0, 10 "
                         "or 11)")
    t_end = cfg.get_real("run", "tEnd", 0.0)
    nstep_max = cfg.get_integer("run", "nStepmax", 0)
    n_output = cfg.get_integer("run", "nOutput", 0)
    nlog = max(1, cfg.get_integer("run", "nlog", 10))
    disable_first = cfg.get_bool("amr", "disable_first_amr_cycle", False)
    amr_on = not cfg.get_bool("amr", "static", False)
    out = Output(cfg, st.gamma0)
    if out.skipped:
        log(f"[output] not written (not built): {', '.join(out.skipped)}")
    # an output is written after its iteration's AMR cycle: a restart from it does not repeat
    # the cycle (nor the output) of the iteration it resumes
    resumed = cfg.get_bool("restart", "enabled", False)
    if resumed:
        prefix = Path(cfg.get("output", "outputDir", ".")) / cfg.get("restart", "filePrefix", "")
        sol = restart_solver(prefix, settings=st, impl_version=iv)
        log(f"[restart] {prefix}: t={sol.t} iteration={sol.iteration} leaves={sol.mesh.num_leaves}")
    else:
        mesh = initial_mesh(cfg, st, amr, iv, log)
        U, B = initial_state(cfg, mesh, st)
        sol = SolverGodunovMHD(mesh, U, B, st, impl_version=iv)
    if max_steps is not None:
        nstep_max = min(nstep_max, sol.iteration + max_steps)
    checks = cfg.get_bool("diagnostic", "conservativity_checks", True)
    i0 = volume_integrals(sol) if checks else None
    updates0 = sol.total_num_cell_updates
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    while sol.t < t_end and sol.iteration < nstep_max:
        if amr_on and not resumed and sol.should_do_amr_cycle(amr, disable_first):
            sol.do_amr_cycle(amr)
        if not resumed and (n_output == -1 or (n_output > 0 and sol.iteration % n_output == 0)):
            base = out.save(sol)
            log(f"[output] {base} t={sol.t}")
        resumed = False
        dt = sol.compute_dt()
        if sol.t + dt > t_end:
            dt = t_end - sol.t
        if sol.iteration % nlog == 0:
            log(f"time step={sol.iteration:07d} (dt={dt:10.8f} t={sol.t:10.8f}) leaves={sol.mesh.num_leaves}")
        sol.step(dt)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    if n_output != 0:
        base = out.save(sol)
        log(f"[output] {base} t={sol.t}")
    if checks:
        i1 = volume_integrals(sol)
        names = ["mass", "energy", "momentum x", "momentum y", "momentum z"]
        for k, name in enumerate(names):
            ref = max(abs(i0[k]), 1e-300)
            log(f"[conservation] {name}: {i0[k]:.15e} -> {i1[k]:.15e} (relative change {(i1[k] - i0[k]) / ref:.3e})")
    rate = (sol.total_num_cell_updates - updates0) / max(wall, 1e-12)
    log(f"final time is {sol.t:.10f} after {sol.iteration} steps, {sol.amr_cycles} AMR cycles; "
        f"{rate / 1e6:.1f} Mcell-updates/s")
    sol.wall_seconds = wall
    return sol


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--ini", required=True)
    ap.add_argument("--impl", choices=sorted(IMPLS), default=None,
                    help="override [hydro] implementation_version")
    ap.add_argument("--max-steps", type=int, default=None, help="stop after this many steps")
    a = ap.parse_args(argv)
    run_simulation(a.ini, a.impl, max_steps=a.max_steps)
    return 0


if __name__ == "__main__":
    sys.exit(main())
