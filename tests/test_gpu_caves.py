"""The depth-12 divergence configuration on the GPU against the oracle: a
4096^3 grid (voxel keys alias past coordinate 1024, node words beyond the
reference's texture-buffer size), fbm ground hollowed by value-noise caves,
1x1 pillars; 4 samples per visible voxel (bottom rank tier), MAX_BOUNCES 5.
GBuffer, counters, candidate and shade lists and every lBuffer integer sum
bit-exact; resolved image within 1e-4 relative.  The first test uses the
same rules on a 256 x 256 patch of ground with pillars across the whole grid
(seconds); the second builds the full scene (about 40 M voxels, 147 M node
words) and compares a 1280x720 primary pass and two small frames."""
import numpy as np
import pytest

from frame_common import DET_KNOBS, assert_out_close, frame_seed, hash_map

pytestmark = pytest.mark.gpu

SPP4 = dict(DET_KNOBS, top_fraction=0.0, mid_fraction=0.0, min_samples=4)


def _frames_vs_oracle(ve, scenes, oracle, words, info, depth, views, W, H, mb, frames, tag):
    rend = ve.Renderer(W, H, max_bounces=mb)
    of = oracle.Frame(W, H, 16384, mb)
    try:
        mats = scenes.palette(1)
        rend.set_materials(mats)
        of.set_materials(mats)
        rend.set_sky_constant((140, 180, 230))
        of.set_sky_rgb8((140, 180, 230))
        scene = rend.upload(words, depth, info.size)
        fg, fo = ve.FrameConfig(**SPP4), oracle.FrameConfig(**SPP4)
        f = 0
        for eye, tgt in views:
            e, d = scenes.look_at(eye, tgt)
            ubo = ve.camera_ubo(e, d, 60.0, W / H)
            rend.set_camera(ubo)
            for _ in range(frames):
                print("case", tag, "view", eye, "frame", f, flush=True)
                rend.frame(scene, fg, ve.FrameInputs(W, H, f, frame_seed(f), f + 1))
                of.frame(words, depth, ubo, fo, f, frame_seed(f), f + 1, n_words=info.size)
                f += 1
                g, _ = rend.gbuffer(W, H)
                assert np.array_equal(g, of.read("GBUFFER").reshape(H, W, 4))
                cg, co = rend.counters(), of.counters()
                assert [cg[0], cg[1], cg[7], cg[4]] == [co[0], co[1], co[7], co[4]]
            nc, ns = int(cg[0]), int(cg[1])
            assert ns > 0
            assert np.array_equal(rend.read("CANDIDATES", np.uint32, nc * 16), of.read("CANDIDATES"))
            assert np.array_equal(rend.read("SHADE_LIST", np.uint32, ns * 8), of.read("SHADE_LIST"))
            gl, ol = hash_map(rend.read("LBUFFER"), 12, 32), hash_map(of.read("LBUFFER"), 12, 32)
            assert np.array_equal(gl[0], ol[0]) and np.array_equal(gl[1], ol[1])
            assert_out_close(rend.read("OUTPUT", np.float32, W * H * 16), of.read("OUTPUT", np.float32))
        rend.release(scene)
    finally:
        rend.close()
        of.close()


def test_cave_patch_frames_vs_oracle(ve, scenes, oracle):
    depth, ops = scenes.scene_caves(4, depth=12, patch=256, pillars=2048)
    words, info = ve.octree_build(ops, depth)
    views = [((300.0, 240.0, 300.0), (128.0, 70.0, 128.0)),        # down into the caves
             ((2048.0, 380.0, 1200.0), (2300.0, 100.0, 2200.0)),    # in the pillar field (keys alias)
             ((-40.0, 130.0, 120.0), (300.0, 60.0, 140.0))]         # grazing across the patch
    _frames_vs_oracle(ve, scenes, oracle, words, info, depth, views, 128, 96, 5, 3, "cave patch")


def test_divergence_scene_vs_oracle(ve, scenes, oracle):
    depth, ops = scenes.scene_caves(1)
    words, info = ve.octree_build(ops, depth, max_words=1 << 28)
    del ops
    assert depth == 12 and info.size > (1 << 27)   # past the reference's 2^27-word texture buffer
    W, H = 1280, 720
    eye, d, fov = scenes.orbit_camera(1 << depth, 5, 32, 1)
    ubo = ve.camera_ubo(eye, d, fov, W / H)
    rend = ve.Renderer(W, H, max_bounces=5)
    try:
        scene = rend.upload(words, depth, info.size)
        rend.set_camera(ubo)
        print("case divergence primary", W, H, flush=True)
        rend.run_pass(scene, "PRIMARY", ve.FrameConfig(**SPP4), ve.FrameInputs(W, H, 0, 1, 1))
        g, p = rend.gbuffer(W, H)
        og, op, st = oracle.generate_primary(words, depth, ubo, W, H, 70, n_words=info.size)
        assert np.array_equal(g, og) and np.array_equal(p, op)
        assert (g[..., 3] != 0).mean() > 0.3 and ((g[..., 3] == 0) & (g[..., 0] >= 70)).mean() > 0.01
        rend.release(scene)
    finally:
        rend.close()
    views = [(tuple(float(v) for v in eye), (2048.0, 900.0, 2048.0)),
             ((1800.0, 1300.0, 1500.0), (2100.0, 900.0, 2300.0))]
    _frames_vs_oracle(ve, scenes, oracle, words, info, depth, views, 192, 108, 5, 2, "divergence")
