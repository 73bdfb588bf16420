"""CPU tests of the benchmark workload meshes (godunov_mhd_ct_amd/workloads.py)."""
import numpy as np
import pytest

from godunov_mhd_ct_amd import workloads


def test_mix_meshes_have_the_stated_block_counts():
    """The batching configuration: 2048 x 8^3 + 1434 x 16^3 + 614 x 32^3 = 4096 blocks and
    27,041,792 cells over levels 2-6, every mesh 2:1 balanced and periodic."""
    total_blocks = total_cells = 0
    for n, (b, target, nbrick) in enumerate(workloads.MIX_MESHES):
        m = workloads.random_amr_mesh(b, target, 101 * n, 2, 6, nbrick)
        lv, co = m.leaves()
        nl = m.num_leaves
        assert nl == target and m.num_outside == 0
        assert lv[:nl].min() >= 2 and lv[:nl].max() <= 6
        # the leaves tile the brick: their volumes (in tree units) add up to the tree count
        vol = (1.0 / 8.0 ** lv[:nl].astype(np.float64)).sum()
        assert vol == pytest.approx(nbrick[0] * nbrick[1] * nbrick[2], rel=1e-12)
        total_blocks += nl
        total_cells += nl * b ** 3
    assert total_blocks == 4096 and total_cells == 27_041_792


def test_random_amr_mesh_is_seeded_and_rejects_unreachable_targets():
    a = workloads.random_amr_mesh(8, 64 + 7 * 20, 3, 2, 5)
    b = workloads.random_amr_mesh(8, 64 + 7 * 20, 3, 2, 5)
    assert np.array_equal(a.leaves()[0], b.leaves()[0]) and np.array_equal(a.leaves()[1], b.leaves()[1])
    with pytest.raises(ValueError):
        workloads.random_amr_mesh(8, 100, 0, 2, 5)  # 100 - 64 is not a multiple of 7
