"""The depth-12 divergence scene's generator (ve_gen_caves: a heightfield
ground with value-noise caves under its surface, only the shell and the cave
walls as voxels) against the same rule restated over whole numpy arrays, and
against a property the rule must give: every solid voxel that faces air
(above the ground or in a cave) is a voxel of the scene, so no ray can look
into the solid ground."""
import numpy as np
import pytest

CASES = [(32, 10, 8, 0.55, 3), (64, 20, 8, 0.5, 3), (48, 63, 16, 0.45, 5), (80, 48, 32, 0.55, 1), (64, 1, 4, 0.3, 9)]


@pytest.mark.parametrize("n,band,cell,thr,seed", CASES)
def test_gen_caves_equals_array_rule(ve, scenes, n, band, cell, thr, seed):
    hi, lo = scenes.shell_columns(n, seed)
    a = ve.gen_caves(hi, lo, band, cell, thr, 77, seed, threads=3)
    b = scenes.caves_numpy(hi, lo, band, cell, thr, 77, seed)
    a = a[np.lexsort((a["y"], a["z"], a["x"]))]
    assert len(a) > 0 and np.array_equal(a, b)


@pytest.mark.parametrize("n,band,cell,thr,seed", CASES[:3])
def test_caves_leave_no_solid_voxel_facing_air_unlisted(ve, scenes, n, band, cell, thr, seed):
    hi, lo = scenes.shell_columns(n, seed)
    ops = ve.gen_caves(hi, lo, band, cell, thr, 77, seed)
    listed = np.zeros((n, n, n), bool)
    listed[ops["x"], ops["y"], ops["z"]] = True
    # air: above the ground, or a cave position (from the noise directly)
    X, Y, Z = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    HI = hi[:, None, :].repeat(n, 1)
    v = scenes.value_noise3((X + 0.5) / cell, (Y + 0.5) / cell, (Z + 0.5) / cell, 77)
    air = (Y >= HI) | ((Y >= HI - band) & (v > thr))
    assert not np.any(listed & air)
    solid = ~air
    pad = np.pad(air, 1, constant_values=False)
    faces_air = (pad[2:, 1:-1, 1:-1] | pad[:-2, 1:-1, 1:-1] | pad[1:-1, 2:, 1:-1] | pad[1:-1, :-2, 1:-1] |
                 pad[1:-1, 1:-1, 2:] | pad[1:-1, 1:-1, :-2])
    assert np.all(listed[solid & faces_air])
    # every voxel carries the seeded material rule
    assert np.array_equal(ops["material"], scenes.voxel_materials(ops["x"], ops["y"], ops["z"], seed))


def test_gen_caves_rejects_bad_input(ve, scenes):
    hi, lo = scenes.shell_columns(16, 1)
    for band, cell in ((0, 8), (64, 8), (8, 0)):
        with pytest.raises(ve.VeError):
            ve.gen_caves(hi, lo, band, cell, 0.5, 1, 1)
    bad = lo.copy()
    bad[3, 4] = hi[3, 4] + 1
    with pytest.raises(ve.VeError):
        ve.gen_caves(hi, bad, 8, 8, 0.5, 1, 1)
    tall = hi.copy()
    tall[0, 0] = 17
    with pytest.raises(ve.VeError):
        ve.gen_caves(tall, lo, 8, 8, 0.5, 1, 1)


def test_pillars_stand_on_the_ground(scenes):
    n = 512
    hi, _ = scenes.shell_columns(n, 2)
    ops = scenes.pillar_ops(n, 2, hi, count=64, hmin=16, hmax=40)
    cols = {}
    for x, y, z in zip(ops["x"].tolist(), ops["y"].tolist(), ops["z"].tolist()):
        cols.setdefault((x, z), []).append(y)
    assert len(cols) <= 64 and len(cols) > 48   # a few pillars may share a column
    for (x, z), ys in cols.items():
        ys = sorted(set(ys))
        assert ys[0] == hi[x, z] and ys == list(range(ys[0], ys[-1] + 1))
        assert 16 <= len(ys) <= 40 and 0.05 * n <= x < 0.95 * n and 0.05 * n <= z < 0.95 * n


def test_scene_caves_patch_builds(ve, scenes):
    depth, ops = scenes.scene_caves(4, depth=12, patch=96, pillars=32)
    assert ops["x"].max() >= 1024 or ops["z"].max() >= 1024   # pillars beyond the voxel-key range
    words, info = ve.octree_build(ops, depth)
    assert info.depth == 12 and info.num_voxels == len(np.unique(ops[["x", "y", "z"]]))
