"""The fused bounce pass with a grid smaller than its ray list (tuning
bounce_blocks): lanes whose ray ends take the next ray from a queue.  Every
structure equals the full-grid pass's, bit for bit, frame by frame -- the
deposits are ordered by their keys afterwards, so no result depends on which
lane traced which ray or when."""
import numpy as np
import pytest

from frame_common import frame_seed, hash_map

pytestmark = pytest.mark.gpu

BUFS = ("GBUFFER", "GBUFFER_POS", "COUNTERS", "LBUFFER", "NBUFFER", "EMISSIVES", "OUTPUT")


def _run(ve, scenes, words, depth, info, mats, W, H, mb, frames, knobs, blocks):
    a = ve.Renderer(W, H, hash_width=16384, max_bounces=mb)
    b = ve.Renderer(W, H, hash_width=16384, max_bounces=mb)
    try:
        for r in (a, b):
            r.set_overlap(False)
            r.set_materials(mats)
            r.set_sky_constant((120, 160, 210))
        a.set_tuning(bounce_blocks=blocks)
        assert a.tuning()["bounce_blocks"] == blocks
        sa, sb = a.upload(words, depth, info.size), b.upload(words, depth, info.size)
        for f in range(frames):
            eye, d, fov = scenes.orbit_camera(1 << depth, f, 16, 5)
            ubo = ve.camera_ubo(eye, d, fov, W / H)
            fc, fi = ve.FrameConfig(**knobs), ve.FrameInputs(W, H, f, frame_seed(f), f + 1)
            for r, s in ((a, sa), (b, sb)):
                r.set_camera(ubo)
                r.frame(s, fc, fi)
            print("case bounce queue blocks", blocks, "frame", f, W, H, mb, flush=True)
            for name in BUFS:
                assert np.array_equal(a.read(name), b.read(name)), "frame %d %s" % (f, name)
            ra, rb = hash_map(a.read("RBUFFER"), 8, 16), hash_map(b.read("RBUFFER"), 8, 16)
            assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1], rb[1]), "frame %d RBUFFER" % f
            assert a.ray_totals() == b.ray_totals()
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("blocks", [1, 3, 40])
def test_bounce_queue_matches_full_grid(ve, scenes, small_scene, blocks):
    knobs = dict(primary_controlchecks=70, bounce_controlchecks=100, min_samples=2)
    _run(ve, scenes, small_scene["words"], small_scene["depth"], small_scene["info"], scenes.palette_spec(1), 160, 120,
         5, 3, knobs, blocks)


def test_bounce_queue_with_lights(ve, scenes):
    depth, ops = scenes.scene_lights(1)
    words, info = ve.octree_build(ops, depth)
    knobs = dict(restir_spatial=0, weight_pixels=0.0, weight_variance=0.0, top_fraction=0.0, mid_fraction=1.0)
    _run(ve, scenes, words, depth, info, scenes.palette_lights(1), 128, 96, 4, 3, knobs, 5)
