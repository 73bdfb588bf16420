"""One implementation context moved between meshes of different sizes, as the AMR cycle does
(resize_auxiliary_data, GodunovImplemV0.cpp:40-75).  The device mesh tables, the border-edge
gather list and the scratch arena grow only: a new mesh that fits them reuses them.  A step on
the reused context must therefore equal, bit for bit, the step of a fresh context on the same
mesh, whether the previous mesh was larger (stale entries past the new sizes) or smaller
(growth), and whether or not it had outside blocks.
"""
import numpy as np
import pytest

from godunov_mhd_ct_amd.implem import IMPL_FUSED, IMPL_PLAIN, GodunovImplemHIP, hydro_settings
from godunov_mhd_ct_amd.mesh import uniform_mesh
from test_gpu_amr import amr3d, dev, state

GAMMA = 5.0 / 3.0


def one_step(im, m, U, B):
    import torch

    Ud, Bd = dev(U), dev(B)
    if m.num_outside:
        im.fill_outside(Ud, Bd)
    U2, B2 = Ud.clone(), Bd.clone()
    dt = im.compute_dt(Ud, Bd)
    im.do_time_step(Ud, U2, Bd, B2, dt)
    torch.cuda.synchronize()
    return dt, U2.cpu().numpy(), B2.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("impl", [IMPL_PLAIN, IMPL_FUSED])
def test_context_reused_across_meshes_matches_fresh(impl):
    small = uniform_mesh(3, 8, 1)  # 8 leaves of one level: no border-edge corrections
    big = amr3d(8)  # 15 leaves over two levels: coarse-fine borders
    # x bounded: 8 outside blocks (the outside-block tables in use)
    walled = uniform_mesh(3, 8, 1, periodic=(0, 1, 1),
                          bc=["ZERO_GRADIENT", "WALL", "PERIODIC", "PERIODIC", "PERIODIC", "PERIODIC"])
    assert big.num_leaves > small.num_leaves and walled.num_outside > 0
    st = hydro_settings(gamma0=GAMMA)
    reused = GodunovImplemHIP(st, 3, 8, impl)
    for n, m in enumerate((small, big, walled, small, walled, big)):
        print(f"[reuse] impl={impl} mesh {n}: {m.num_leaves} leaves, {m.num_outside} outside", flush=True)
        U, B = state(m, 3 + n)
        reused.resize_auxiliary_data(m)
        fresh = GodunovImplemHIP(st, 3, 8, impl)
        fresh.resize_auxiliary_data(m)
        dta, Ua, Ba = one_step(reused, m, U, B)
        dtb, Ub, Bb = one_step(fresh, m, U, B)
        assert dta == dtb
        assert np.array_equal(Ua, Ub) and np.array_equal(Ba, Bb), n
        del fresh
