"""MI355X-native per-block MHD-CT time step (MUSCL-Hancock, HLLD, constrained transport).

Host side: the block octree (`mesh`), the GodunovImplemBase counterpart (`implem`), the
solver driver (`solver`) and problem setups (`problems`); device side: libmhdct.so, built
from csrc/ by `python -m godunov_mhd_ct_amd.build`.
"""
from ._lib import MhdctError, lib, lib_path  # noqa: F401
from .implem import IMPL_FUSED, IMPL_PLAIN, GodunovImplemHIP, create_godunov_implem, hydro_settings  # noqa: F401
from .mesh import Mesh, refine_where, uniform_mesh  # noqa: F401
