"""GPU test of mhdct_refine_flags_noalloc, the refinement-flag entry point that allocates nothing
(its device buffers are a file-scope array of the library): through the C ABI, its flags and
indicators must equal mhdct_refine_flags' bit for bit, on two calls in a row, the second on a
smaller mesh whose indicators (a weak density step) are smaller than the first's (a strong one) (so a per-block maximum left over
from the first call would show), then once more on the first mesh."""
import ctypes as C

import numpy as np
import pytest

from godunov_mhd_ct_amd import problems
from godunov_mhd_ct_amd._lib import lib
from godunov_mhd_ct_amd.amr import AMRSettings
from godunov_mhd_ct_amd.implem import IMPL_PLAIN, GodunovImplemHIP, _ptr, hydro_settings
from godunov_mhd_ct_amd.mesh import Mesh, refine_where


def _flags(fn, implem, U, amr):
    n = implem.mesh.num_leaves
    flags = np.full(n, 7, dtype=np.int8)
    ind = np.full(n, np.nan)
    p = amr.params()
    rc = fn(implem._h, _ptr(U), None, C.byref(p), flags.ctypes.data_as(C.POINTER(C.c_int8)),
            ind.ctypes.data_as(C.POINTER(C.c_double)), C.c_void_p(0))
    assert rc == 0, lib().mhdct_last_error(implem._h)
    return flags, ind


def _case(b, level, pred, small):
    m = refine_where(Mesh(3, b, level), pred, level + 1)
    # a density step of 2 % (small): Loehner's noise filter keeps its indicator near 0.2,
    # against about 0.8 for the 50 % step of the large mesh
    rho_in = 1.02 if small else 1.5
    U, B = problems.blast(m, radius=0.3, center=(0.4, 0.45, 0.5), p_in=50.0, p_out=0.1, rho_in=rho_in)
    st = hydro_settings(gamma0=5.0 / 3.0)
    implem = GodunovImplemHIP(st, 3, b, IMPL_PLAIN)
    implem.resize_auxiliary_data(m)
    import torch
    Ud = torch.as_tensor(U, device="cuda")
    return implem, Ud


@pytest.mark.gpu
def test_refine_flags_noalloc_matches_refine_flags():
    amr = AMRSettings(level_min=1, level_max=3, epsilon_refine=0.05, epsilon_coarsen=0.01, epsilon_lohner=0.02)
    big, Ubig = _case(8, 2, lambda L, x0, x1, y0, y1, z0, z1: x0 < 0.5, False)
    small, Usmall = _case(8, 1, lambda L, x0, x1, y0, y1, z0, z1: x0 < 0.5 and y0 < 0.5, True)
    assert small.mesh.num_leaves < big.mesh.num_leaves
    print(f"[noalloc] meshes of {big.mesh.num_leaves} and {small.mesh.num_leaves} leaves", flush=True)
    ref_big = _flags(lib().mhdct_refine_flags, big, Ubig, amr)
    ref_small = _flags(lib().mhdct_refine_flags, small, Usmall, amr)
    # a maximum left from the first call in any of the small mesh's blocks would show
    assert ref_small[1].max() < ref_big[1][:small.mesh.num_leaves].max()
    assert (ref_big[0] == 1).any() and (ref_big[0] == -1).any()
    for implem, U, ref in ((big, Ubig, ref_big), (small, Usmall, ref_small), (big, Ubig, ref_big)):
        got = _flags(lib().mhdct_refine_flags_noalloc, implem, U, amr)
        assert np.array_equal(got[0], ref[0]), "flags differ"
        assert np.array_equal(got[1].view(np.int64), ref[1].view(np.int64)), "indicators differ"
