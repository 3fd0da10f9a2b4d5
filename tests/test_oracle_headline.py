"""Pin the C oracle on the measured workload itself against the reference's
own GLSL on Mesa llvmpipe (tests/golden/headline_llvmpipe.*): the depth-10
scene of 10,401,603 voxels, 1920x1080, MAX_BOUNCES 3, the benchmark's knobs,
orbit frames 0..3 accumulating (renderer.cpp:282-649).  At this size the
shading budget truncates the shade list, the ray ring runs four chunks of
49,152 rays, and the traversal runs from outside the grid at coordinates
near 1024, where the association llvmpipe's compiler gives the nudged
intersections (internal.glsl:245-261) decides 70 pixels' step counts and 7
voxel ids of frame 0 -- the fixture pins that association."""
import numpy as np
import pytest

from frame_common import (HEADLINE_FRAMES, HEADLINE_KNOBS, HEADLINE_SKY, compare_headline_frame, frame_seed,
                          hash_map, headline_digests, load_headline_gold)


@pytest.fixture(scope="module")
def extreme(ve, scenes):
    depth, ops = scenes.scene_extreme(1)
    words, info = ve.octree_build(ops, depth)
    return depth, words, info


def test_oracle_headline_frames_match_llvmpipe(ve, scenes, oracle, extreme):
    meta, arr = load_headline_gold()
    depth, words, info = extreme
    import hashlib
    assert hashlib.sha256(np.ascontiguousarray(words[: info.size]).tobytes()).hexdigest() == meta["words_sha256"]
    W, H = meta["W"], meta["H"]
    of = oracle.Frame(W, H, 16384, meta["max_bounces"])
    try:
        of.set_materials(scenes.palette(1))
        of.set_sky_rgb8(HEADLINE_SKY)
        fc = oracle.FrameConfig(**HEADLINE_KNOBS)
        for f in range(HEADLINE_FRAMES):
            eye, d, fov = scenes.orbit_camera(1 << depth, f, 64, 1)
            of.frame(words, depth, ve.camera_ubo(eye, d, fov, W / H), fc, f, frame_seed(f), f + 1,
                     n_words=info.size)
            c = of.counters()
            nc, ns = int(c[0]), int(c[1])
            lmap = hash_map(of.read("LBUFFER"), 12, 32)
            dg = headline_digests(of.read("GBUFFER"), of.read("GBUFFER_POS"), of.read("CANDIDATES")[: 4 * nc],
                                  of.read("SHADE_LIST")[: 2 * ns], of.read("RANK_STATS"),
                                  hash_map(of.read("NBUFFER"), 3, 32), lmap)
            compare_headline_frame(meta, arr, f, [nc, ns, int(c[7]), int(c[4])], dg, lmap,
                                   of.read("OUTPUT", np.float32))
    finally:
        of.close()


def test_headline_fixture_rejects_the_source_order_nudges(ve, scenes, oracle, extreme):
    """The frame-0 G-buffer with the nudges evaluated in GLSL source order
    ((b - o) -+ 0.001, an FP variant of the oracle) differs from llvmpipe's."""
    meta, _ = load_headline_gold()
    depth, words, info = extreme
    W, H = meta["W"], meta["H"]
    eye, d, fov = scenes.orbit_camera(1 << depth, 0, 64, 1)
    ubo = ve.camera_ubo(eye, d, fov, W / H)
    oracle.lib().so_set_fp_variant(5)
    try:
        g, p, _ = oracle.generate_primary(words, depth, ubo, W, H, HEADLINE_KNOBS["primary_controlchecks"],
                                          threads=8, n_words=info.size)
    finally:
        oracle.lib().so_set_fp_variant(0)
    from frame_common import sha
    assert sha(np.asarray(g, np.uint32).reshape(-1)) != meta["frames"][0]["digests"]["gbuf"]
    g0, p0, _ = oracle.generate_primary(words, depth, ubo, W, H, HEADLINE_KNOBS["primary_controlchecks"],
                                        threads=8, n_words=info.size)
    assert sha(np.asarray(g0, np.uint32).reshape(-1)) == meta["frames"][0]["digests"]["gbuf"]
