"""GPU parity on the measured workload itself: the depth-10 scene of
10,401,603 voxels (fbm terrain shell + 64 spheres), 1920x1080, MAX_BOUNCES 3
(two bounces), the benchmark's knobs and orbit camera, through the C ABI,
against the C oracle (Renderer::runWavefrontFrame, renderer.cpp:282-649).

At this size the 470-480 k candidates per frame exceed the 196,608-entry
shading budget, so the shade list is the budget-truncated prefix of the
rank pass's sequential order (rank.comp:120-129): a wrong order would show
in the shade list, the emitted rays and every lBuffer bucket.

Bit-exact: G-buffer and packed positions, counters, candidate list (order
and representative pixel), shade list, per-voxel pixel counts, claim map,
RankStats, l/nBuffer, emit.comp's ray records.  Resolved RGBA32F within
1e-4 relative."""
import numpy as np
import pytest

from frame_common import HEADLINE_KNOBS, HEADLINE_SKY, compare_frame, frame_seed

pytestmark = pytest.mark.gpu

W, H = 1920, 1080
MAX_BOUNCES = 3
ORBIT = 64


@pytest.fixture(scope="module")
def extreme(ve, scenes):
    depth, ops = scenes.scene_extreme(1)
    words, info = ve.octree_build(ops, depth)
    assert info.num_voxels == 10401603
    return depth, words, info


def _ubo(ve, scenes, depth, f):
    eye, d, fov = scenes.orbit_camera(1 << depth, f, ORBIT, 1)
    return ve.camera_ubo(eye, d, fov, W / H)


def _pair(ve, scenes, oracle):
    rend = ve.Renderer(W, H, hash_width=16384, max_bounces=MAX_BOUNCES)
    of = oracle.Frame(W, H, 16384, MAX_BOUNCES)
    mats = scenes.palette(1)
    rend.set_materials(mats)
    of.set_materials(mats)
    rend.set_sky_constant(HEADLINE_SKY)
    of.set_sky_rgb8(HEADLINE_SKY)
    return rend, of


def test_headline_frames_match_oracle(ve, scenes, oracle, extreme):
    """Three accumulating orbit frames through ve_frame (cross-frame overlap
    on, as the benchmark renders them)."""
    depth, words, info = extreme
    rend, of = _pair(ve, scenes, oracle)
    try:
        scene = rend.upload(words, depth, info.size)
        fc_g, fc_o = ve.FrameConfig(**HEADLINE_KNOBS), oracle.FrameConfig(**HEADLINE_KNOBS)
        for f in range(3):
            print("case headline frame", f, flush=True)
            ubo = _ubo(ve, scenes, depth, f)
            rend.set_camera(ubo)
            rend.frame(scene, fc_g, ve.FrameInputs(W, H, f, frame_seed(f), f + 1))
            of.frame(words, depth, ubo, fc_o, f, frame_seed(f), f + 1, n_words=info.size)
            c = compare_frame(ve, rend, of, W, H, "headline frame %d" % f)
            assert c[0] > 196608 and c[1] == 196608, "the budget truncates the shade list"
        rend.release(scene)
    finally:
        rend.close()
        of.close()


def test_headline_emit_records_match_oracle(ve, scenes, oracle, extreme):
    """emit.comp's ray records (VE_BUF_RAYS after the emit pass, dispatched
    pass by pass) equal the oracle's for the whole truncated shade list."""
    depth, words, info = extreme
    rend, of = _pair(ve, scenes, oracle)
    try:
        scene = rend.upload(words, depth, info.size)
        fc_g, fc_o = ve.FrameConfig(**HEADLINE_KNOBS), oracle.FrameConfig(**HEADLINE_KNOBS)
        f = 5
        ubo = _ubo(ve, scenes, depth, f)
        rend.set_camera(ubo)
        inputs = ve.FrameInputs(W, H, f, frame_seed(f), f + 1)
        print("case headline emit frame", f, flush=True)
        for name in ("RESET", "PRIMARY", "NORMAL", "SCHEDULE", "RANK", "THRESHOLD", "EMIT"):
            rend.run_pass(scene, name, fc_g, inputs)
        ns = int(rend.counters()[1])
        rays = rend.read("RAYS", np.uint32, ns * 96).reshape(ns, 24)
        of.frame(words, depth, ubo, fc_o, f, frame_seed(f), f + 1, n_words=info.size)
        assert ns == int(of.counters()[1]) == 196608
        ref = of.emit_records(0, ns)
        assert np.array_equal(rays, ref)
        rend.release(scene)
    finally:
        rend.close()
        of.close()


def test_headline_orbit_primaries_match_oracle(ve, scenes, oracle, extreme):
    """Primary passes of four more orbit frames (generate_primary.comp)."""
    depth, words, info = extreme
    rend = ve.Renderer(W, H)
    try:
        scene = rend.upload(words, depth, info.size)
        for f in (9, 23, 37, 51):
            print("case headline primary frame", f, flush=True)
            ubo = _ubo(ve, scenes, depth, f)
            rend.set_camera(ubo)
            rend.run_pass(scene, "PRIMARY", ve.FrameConfig(**HEADLINE_KNOBS), ve.FrameInputs(W, H, f, 1, 1))
            g, p = rend.gbuffer(W, H)
            go, po, _ = oracle.generate_primary(words, depth, ubo, W, H, 70, n_words=info.size)
            assert np.array_equal(g, go) and np.array_equal(p, po)
        rend.release(scene)
    finally:
        rend.close()

