"""GPU cube-map sky (texture(samplerCube, dir) under the reference skybox's
sampler state, skybox.cpp:87-122) through the C ABI: whole frames against
the reference GLSL on llvmpipe under a random 8x8 cube
(tests/golden/frame_sky.*) -- voxel-ID GBuffer, counters and every lBuffer
integer sum bit-exact (escaping bounce rays sample the sky, shade.comp:559),
resolved image within 1e-4 relative; and against the oracle over random
cameras and cube sizes, where every sky pixel (resolve.comp:218-288) is bit
for bit the oracle's sample along the pixel direction."""
import json
import os

import numpy as np
import pytest

from frame_common import DET_KNOBS, assert_out_close, frame_seed, hash_map

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("run", ["sky_det", "sky_timeout"])
def test_sky_frames_match_llvmpipe(ve, scenes, small_scene, run):
    meta = json.load(open(os.path.join(GOLD, "frame_sky.json")))
    arr = np.load(os.path.join(GOLD, "frame_sky.npz"))
    r = meta["runs"][run]
    W, H = r["W"], r["H"]
    rend = ve.Renderer(W, H, hash_width=16384, max_bounces=r["max_bounces"])
    try:
        rend.set_materials(scenes.palette(1))
        rend.set_sky_cube(arr["sky_faces"])
        rend.set_camera(arr[run + "_ubo"])
        scene = rend.upload(small_scene["words"], small_scene["depth"], small_scene["info"].size)
        fc = ve.FrameConfig(**r["knobs"])
        counters = []
        for f in range(r["frames"]):
            print("case sky frame", run, f, flush=True)
            rend.frame(scene, fc, ve.FrameInputs(W, H, f, frame_seed(f), f + 1))
            c = rend.counters()
            counters.append([c[0], c[1], c[7], c[4]])
        assert np.array_equal(np.array(counters, np.uint32), arr[run + "_counters"])
        g, _ = rend.gbuffer(W, H)
        assert np.array_equal(g, arr[run + "_gbuf"])
        lv, lval = hash_map(rend.read("LBUFFER"), 12, 32)
        assert np.array_equal(lv, arr[run + "_lbuf_vids"]) and np.array_equal(lval, arr[run + "_lbuf_vals"])
        assert_out_close(rend.read("OUTPUT", np.float32, W * H * 16), arr[run + "_out"])
        rend.release(scene)
    finally:
        rend.close()


def test_sky_pixels_equal_oracle(ve, scenes, oracle, small_scene):
    """Random cameras near the horizon under cubes of edge 1 (six different faces), 3, 16 and 64, then
    a uniform cube (kept as a constant): GPU == oracle on the voxel-ID image,
    every lBuffer integer sum and, bit for bit, on every sky pixel; the rest
    of the image within tolerance."""
    rng = np.random.default_rng(31)
    W, H = 96, 72
    words, depth = small_scene["words"], small_scene["depth"]
    rend = ve.Renderer(W, H, hash_width=4096, max_bounces=3)
    of = oracle.Frame(W, H, 4096, 3)
    try:
        rend.set_materials(scenes.palette(1))
        of.set_materials(scenes.palette(1))
        scene = rend.upload(words, depth, small_scene["info"].size)
        fc_kw = dict(DET_KNOBS)
        fc, ofc = ve.FrameConfig(**fc_kw), oracle.FrameConfig(**fc_kw)
        n = 0
        for edge in (1, 3, 16, 64, 0):
            if edge:
                faces = rng.integers(0, 256, (6, edge, edge, 3)).astype(np.uint8)
            else:
                faces = np.broadcast_to(np.array([200, 17, 96], np.uint8), (6, 5, 5, 3))
            rend.set_sky_cube(faces)
            of.set_sky_cube(faces)
            for k in range(2):
                # near the horizon: terrain below, sky above
                eye = rng.uniform([10, 60, 10], [118, 110, 118]).astype(np.float32)
                a = rng.uniform(0, 2 * np.pi)
                tgt = eye + np.array([50 * np.cos(a), rng.uniform(-25, 5), 50 * np.sin(a)], np.float32)
                e, d = scenes.look_at(eye, tgt)
                ubo = ve.camera_ubo(e, d, float(rng.uniform(50, 100)), W / H)
                rend.set_camera(ubo)
                print("case sky pixels edge", edge, "view", k, flush=True)
                rend.frame(scene, fc, ve.FrameInputs(W, H, n, frame_seed(n), n + 1))
                of.frame(words, depth, ubo, ofc, n, frame_seed(n), n + 1)
                n += 1
                g, _ = rend.gbuffer(W, H)
                og = of.read("GBUFFER").reshape(H, W, 4)
                assert np.array_equal(g, og)
                out = rend.read("OUTPUT", np.float32, W * H * 16).reshape(-1, 4)
                oout = of.read("OUTPUT", np.float32).reshape(-1, 4)
                sky = ((g[..., 3] == 0) & (g[..., 0] < fc_kw["primary_controlchecks"])).ravel()
                assert sky.sum() > 0
                assert np.array_equal(out[sky], oout[sky])
                assert_out_close(out, oout)
            lv, lval = hash_map(rend.read("LBUFFER"), 12, 32)
            olv, olval = hash_map(of.read("LBUFFER"), 12, 32)
            assert np.array_equal(lv, olv) and np.array_equal(lval, olval)
        rend.release(scene)
    finally:
        rend.close()
        of.close()
