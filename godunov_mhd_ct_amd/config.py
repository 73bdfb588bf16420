"""The run's .ini settings (mhdct_config_* in include/mhdct.h), read as the reference reads them
through kalypsso::core's ConfigMap: the same sections, keys and defaults.

    cfg = Config.load("test_mhd_blast_3d.ini")
    st = cfg.settings()          # mhdct_settings: [hydro], [mesh] scaling_factor, [gravity], ...
    amr = cfg.amr()              # AMRSettings: [amr] thresholds, cycle frequency
    mesh = cfg.mesh()            # uniform at [amr] level_min over the [p4est_connectivity] brick

The reader and the settings/mesh mapping are in the library (csrc/config.cpp, capi.hip); the
typed reads here follow the same rules for the keys the library does not map (problem
parameters, [run] and [output] keys).
"""
from __future__ import annotations

import ctypes as C

from ._lib import AmrParams, MhdctError, Settings, lib
from .amr import CRITERIA, DERIVED, VARIABLES, AMRSettings
from .mesh import BC, Mesh

_TRUE = {"true", "yes", "on", "1"}
_FALSE = {"false", "no", "off", "0"}


class Config:
    def __init__(self, handle: C.c_void_p, source: str = "<text>"):
        self._h = handle
        self.source = source

    @classmethod
    def parse(cls, text: str) -> "Config":
        h = C.c_void_p()
        rc = lib().mhdct_config_parse(text.encode(), C.byref(h))
        if rc > 0:
            raise MhdctError(f"ini text, line {rc}: expected [section], key = value, or a comment")
        if rc != 0:
            raise MhdctError(f"mhdct_config_parse failed ({rc})")
        return cls(h)

    @classmethod
    def load(cls, path) -> "Config":
        h = C.c_void_p()
        rc = lib().mhdct_config_load(str(path).encode(), C.byref(h))
        if rc == -2:
            raise MhdctError(f"{path}: cannot be read")
        if rc > 0:
            raise MhdctError(f"{path}, line {rc}: expected [section], key = value, or a comment")
        if rc != 0:
            raise MhdctError(f"mhdct_config_load failed ({rc})")
        return cls(h, str(path))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().mhdct_config_destroy(h)
            except Exception:
                pass
            self._h = None

    def _err(self, what, rc):
        msg = lib().mhdct_config_last_error(self._h)
        raise MhdctError(f"{self.source}: {what} failed ({rc}): {msg.decode() if msg else ''}")

    # ---- typed reads (ConfigMap getString / getReal / getInteger / getBool / getStringVector)
    def get(self, section: str, key: str, default=None):
        v = lib().mhdct_config_get(self._h, section.encode(), key.encode())
        return v.decode() if v is not None else default

    def get_real(self, section, key, default: float) -> float:
        v = self.get(section, key)
        try:
            return float(v) if v not in (None, "") else default
        except ValueError:
            return default

    def get_integer(self, section, key, default: int) -> int:
        v = self.get(section, key)
        try:
            return int(v, 0) if v not in (None, "") else default
        except ValueError:
            return default

    def get_bool(self, section, key, default: bool) -> bool:
        v = self.get(section, key)
        if v is None:
            return default
        v = v.lower()
        return True if v in _TRUE else (False if v in _FALSE else default)

    def get_string_vector(self, section, key, default):
        v = self.get(section, key)
        if v is None:
            return list(default)
        return [x.strip() for x in v.split(",") if x.strip()]

    # ---- what the run is built from
    @property
    def dim(self) -> int:
        return self.get_integer("run", "dimension", 0)

    def settings(self) -> Settings:
        st = Settings()
        rc = lib().mhdct_config_settings(self._h, C.byref(st))
        if rc != 0:
            self._err("settings", rc)
        return st

    def amr(self) -> AMRSettings:
        p = AmrParams()
        rc = lib().mhdct_config_amr_params(self._h, C.byref(p))
        if rc != 0:
            self._err("amr params", rc)
        names = {v: k for k, v in list(VARIABLES.items()) + list(DERIVED.items())}
        crit = {v: k for k, v in CRITERIA.items()}[p.criterion]
        variables = [names[p.vars[i]] for i in range(p.nvars)] if p.nvars > 0 else [names[p.ivar]]
        return AMRSettings(level_min=p.level_min, level_max=p.level_max, epsilon_refine=p.epsilon_refine,
                           epsilon_coarsen=p.epsilon_coarsen, epsilon_lohner=p.epsilon_lohner,
                           variable=variables[0] if len(variables) == 1 else variables, criterion=crit,
                           cycle_frequency=self.get_integer("amr", "cycle_frequency", 1))

    def mesh(self) -> Mesh:
        h = C.c_void_p()
        rc = lib().mhdct_config_mesh(self._h, C.byref(h))
        if rc != 0:
            self._err("mesh", rc)
        dim = self.dim
        b = self.get_integer("amr", "bx", 0)
        nbrick = tuple(self.get_integer("p4est_connectivity", f"nbrick_{a}", 1) if d < dim else 1
                       for d, a in enumerate("xyz"))
        periodic = tuple(int(self.get_integer("p4est_connectivity", f"periodic_{a}", 1) != 0) if d < dim else 1
                         for d, a in enumerate("xyz"))
        bc = tuple(BC[self.get("mesh", f"boundary_type_{a}{s}", "ZERO_GRADIENT")] if f < 2 * dim else 0
                   for f, (a, s) in enumerate((a, s) for a in "xyz" for s in ("min", "max")))
        return Mesh.from_handle(h, dim, b, nbrick, periodic, bc)

    def xmin(self):
        return tuple(self.get_real("mesh", f"{a}min", 0.0) for a in "xyz")
