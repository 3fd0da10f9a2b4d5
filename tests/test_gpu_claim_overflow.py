"""Claim rows holding more distinct visible voxels than the schedule pass's
per-row table (schedule.comp:92-120 keeps the first 64 claimants of each of
the 16384 rows): scenes whose voxels all hash to one or two claim rows, so a
row sees hundreds of distinct ids.  The row's candidates, claim and
pixel-count maps, the shade list and the lBuffer match the oracle bit for
bit over moving frames, and the inexact-order flag stays clear."""
import numpy as np
import pytest

from frame_common import DET_KNOBS, assert_out_close, frame_seed, hash_map

pytestmark = pytest.mark.gpu

M32 = np.uint64(0xFFFFFFFF)


def _mix(x):
    x = x ^ (x >> np.uint64(16))
    x = (x * np.uint64(0x7FEB352D)) & M32
    x = x ^ (x >> np.uint64(15))
    x = (x * np.uint64(0x846CA68B)) & M32
    return x ^ (x >> np.uint64(16))


def voxel_ids(x, y, z):
    """voxelKey + 1 (internal.glsl:263-291)."""
    k = (x.astype(np.uint64) | (y.astype(np.uint64) << np.uint64(10)) | (z.astype(np.uint64) << np.uint64(20))) & M32
    return (_mix(k) & np.uint64(0x7FFFFFFF)) + np.uint64(1)


def claim_rows(v):
    """claimHash(vid) % 16384 (schedule.comp:36-43)."""
    v = v ^ (v >> np.uint64(16))
    v = (v * np.uint64(0x7FEB352D)) & M32
    return ((v ^ (v >> np.uint64(15))) % np.uint64(16384)).astype(np.int64)


def _crowded_scene(ve, scenes, rows, terrain):
    g = np.arange(256)
    X, Y, Z = [a.ravel() for a in np.meshgrid(g, g, g, indexing="ij")]
    r = claim_rows(voxel_ids(X, Y, Z))
    sel = np.isin(r, rows) & (Y >= 70)
    parts = [scenes._ops(X[sel], Y[sel], Z[sel], scenes.voxel_materials(X[sel], Y[sel], Z[sel], 1))]
    if terrain:
        parts.insert(0, scenes.terrain_solid(128, 2))   # y < 64 in the first 128 x 128 columns
    ops = np.concatenate(parts)
    words, info = ve.octree_build(ops, 8)
    return words, info


@pytest.mark.parametrize("rows,terrain", [((777,), False), ((777, 9001), True)])
def test_crowded_claim_rows_match_oracle(ve, scenes, oracle, rows, terrain):
    words, info = _crowded_scene(ve, scenes, list(rows), terrain)
    W, H = 256, 256
    rend = ve.Renderer(W, H, max_bounces=2)
    of = oracle.Frame(W, H, 16384, 2)
    try:
        rend.set_materials(scenes.palette(1))
        of.set_materials(scenes.palette(1))
        rend.set_sky_constant((140, 180, 230))
        of.set_sky_rgb8((140, 180, 230))
        scene = rend.upload(words, 8, info.size)
        fg, fo = ve.FrameConfig(**DET_KNOBS), oracle.FrameConfig(**DET_KNOBS)
        for f in range(3):
            e, d = scenes.look_at((128.0 + 9.0 * f, 420.0, -260.0 + 5.0 * f), (128.0, 110.0, 128.0))
            ubo = ve.camera_ubo(e, d, 55.0, 1.0)
            rend.set_camera(ubo)
            print("case crowded rows", rows, "terrain", terrain, "frame", f, flush=True)
            rend.frame(scene, fg, ve.FrameInputs(W, H, f, frame_seed(f), f + 1))
            of.frame(words, 8, ubo, fo, f, frame_seed(f), f + 1, n_words=info.size)
            g, _ = rend.gbuffer(W, H)
            assert np.array_equal(g, of.read("GBUFFER").reshape(H, W, 4))
            vis = np.unique(g[..., 3][g[..., 3] != 0]).astype(np.uint64)
            crowd = np.bincount(claim_rows(vis), minlength=16384)[list(rows)]
            assert crowd.min() > 256                      # far past the per-row table
            cg, co = rend.counters(), of.counters()
            assert cg[ve.CNT["OVERFLOW"]] == 0
            assert [cg[0], cg[1], cg[7], cg[4]] == [co[0], co[1], co[7], co[4]]
            nc, ns = int(cg[0]), int(cg[1])
            assert np.array_equal(rend.read("CANDIDATES", np.uint32, nc * 16), of.read("CANDIDATES"))
            assert np.array_equal(rend.read("CLAIM_MAP"), of.read("CLAIM_MAP"))
            assert np.array_equal(rend.read("SAMPLE_COUNTS"), of.read("SAMPLE_COUNTS"))
            assert np.array_equal(rend.read("SHADE_LIST", np.uint32, ns * 8), of.read("SHADE_LIST"))
        gl, ol = hash_map(rend.read("LBUFFER"), 12, 32), hash_map(of.read("LBUFFER"), 12, 32)
        assert np.array_equal(gl[0], ol[0]) and np.array_equal(gl[1], ol[1])
        assert_out_close(rend.read("OUTPUT", np.float32, W * H * 16), of.read("OUTPUT", np.float32))
        rend.release(scene)
    finally:
        rend.close()
        of.close()
