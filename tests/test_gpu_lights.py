"""GPU next-event estimation against the reference GLSL on llvmpipe
(tests/golden/frame_lights.*) on an emissive scene.

Temporal ReSTIR (bounce 0) + single-pick RIS (bounce 1): the emissive
registry (its order included), every lBuffer integer sum and every reservoir
bit-exact, the resolved image within 1e-4.

Spatial reuse reads the reservoirs of neighbouring voxels while the same
dispatch writes them; the reference's result there depends on execution order
(llvmpipe's serial SIMD-8 order versus a GPU's concurrent waves), so the GPU
run is held to what does not depend on it: the registry, which voxels carry a
reservoir, the per-voxel sample counts, and the image mean within 2%."""
import json
import os

import numpy as np
import pytest

from frame_common import GOLD, assert_out_close, frame_seed, hash_map

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lights_gold():
    with open(os.path.join(GOLD, "frame_lights.json")) as fh:
        meta = json.load(fh)
    return meta, np.load(os.path.join(GOLD, "frame_lights.npz"))


@pytest.fixture(scope="module")
def lights_scene(ve, scenes):
    depth, ops = scenes.scene_lights(1)
    words, info = ve.octree_build(ops, depth)
    return dict(depth=depth, words=words, info=info)


def _gpu(ve, scenes, meta, arr, run, scene):
    r = meta["runs"][run]
    W, H = r["W"], r["H"]
    rend = ve.Renderer(W, H, hash_width=16384, max_bounces=r["max_bounces"])
    rend.set_materials(scenes.palette_lights(1))
    rend.set_sky_constant(meta["sky_rgb8"])
    rend.set_camera(arr[run + "_ubo"])
    sc = rend.upload(scene["words"], scene["depth"], scene["info"].size)
    fc = ve.FrameConfig(**r["knobs"])
    counters, shadow = [], 0
    for f in range(r["frames"]):
        print("case lights", run, "frame", f, flush=True)
        rend.frame(sc, fc, ve.FrameInputs(W, H, f, frame_seed(f), f + 1))
        c = rend.counters()
        counters.append([c[0], c[1], c[7], c[4], c[2]])
        shadow += int(c[ve.CNT["SHADOW_RAYS"]])
    return rend, sc, counters, shadow


def test_nee_temporal_matches_llvmpipe(ve, scenes, lights_gold, lights_scene):
    meta, arr = lights_gold
    run = "lights_temporal"
    rend, sc, counters, shadow = _gpu(ve, scenes, meta, arr, run, lights_scene)
    try:
        assert np.array_equal(np.array(counters, np.uint32), arr[run + "_counters"])
        assert shadow > 0
        ne = int(counters[-1][4])
        em = rend.read("EMISSIVES", np.uint32, ne * 32).reshape(-1, 8)
        assert np.array_equal(em, arr[run + "_emissives"])
        lv, lval = hash_map(rend.read("LBUFFER"), 12, 32)
        assert np.array_equal(lv, arr[run + "_lbuf_vids"])
        assert np.array_equal(lval, arr[run + "_lbuf_vals"])
        rv, rval = hash_map(rend.read("RBUFFER"), 8, 16)
        assert np.array_equal(rv, arr[run + "_rbuf_vids"])
        assert np.array_equal(rval, arr[run + "_rbuf_vals"])
        W, H = meta["runs"][run]["W"], meta["runs"][run]["H"]
        assert_out_close(rend.read("OUTPUT", np.float32, W * H * 16), arr[run + "_out"])
        rend.release(sc)
    finally:
        rend.close()


def test_nee_spatial_statistics(ve, scenes, lights_gold, lights_scene):
    meta, arr = lights_gold
    run = "lights_spatial"
    rend, sc, counters, shadow = _gpu(ve, scenes, meta, arr, run, lights_scene)
    try:
        ref_c = arr[run + "_counters"]
        got = np.array(counters, np.uint32)
        assert np.array_equal(got[:, [0, 1, 4]], ref_c[:, [0, 1, 4]])   # candidates, shade list, registry size
        ne = int(counters[-1][4])
        em = rend.read("EMISSIVES", np.uint32, ne * 32).reshape(-1, 8)
        assert np.array_equal(em, arr[run + "_emissives"])
        rv, _ = hash_map(rend.read("RBUFFER"), 8, 16)
        assert np.array_equal(rv, arr[run + "_rbuf_vids"])
        lv, lval = hash_map(rend.read("LBUFFER"), 12, 32)
        assert np.array_equal(lv, arr[run + "_lbuf_vids"])
        assert np.array_equal(lval[:, [0, 5]], arr[run + "_lbuf_vals"][:, [0, 5]])   # direct / indirect counts
        W, H = meta["runs"][run]["W"], meta["runs"][run]["H"]
        out = rend.read("OUTPUT", np.float32, W * H * 16).reshape(-1, 4)[:, :3]
        ref = arr[run + "_out"].reshape(-1, 4)[:, :3]
        np.testing.assert_allclose(out.mean(0), ref.mean(0), rtol=0.02)
        rend.release(sc)
    finally:
        rend.close()


def test_registry_persists_and_matches_oracle(ve, scenes, oracle, lights_scene):
    """A second view of the same scene: lights first seen there are appended
    after the ones already registered, in the oracle's order."""
    W, H = 96, 72
    knobs = dict(restir_spatial=0, weight_pixels=0.0, weight_variance=0.0, top_fraction=0.0, mid_fraction=1.0,
                 min_samples=1, sample_cap_direct=1024, sample_cap_indirect=4096, disable_stale_reset=1,
                 update_time=0xFFFFFFFF, firefly_floor=1e30, normal_precision=1)
    views = [scenes.look_at((20.0, 110.0, 20.0), (40.0, 70.0, 50.0)), scenes.look_at((120.0, 100.0, 110.0), (80.0, 70.0, 60.0))]
    rend = ve.Renderer(W, H, max_bounces=2)
    of = oracle.Frame(W, H, 16384, 2)
    try:
        for x in (rend, of):
            x.set_materials(scenes.palette_lights(1))
        rend.set_sky_constant((140, 180, 230))
        of.set_sky_rgb8((140, 180, 230))
        sc = rend.upload(lights_scene["words"], lights_scene["depth"], lights_scene["info"].size)
        fc_g, fc_o = ve.FrameConfig(**knobs), oracle.FrameConfig(**knobs)
        for f, (e, d) in enumerate(views * 2):
            ubo = ve.camera_ubo(e, d, 70.0, W / H)
            print("case registry frame", f, flush=True)
            rend.set_camera(ubo)
            rend.frame(sc, fc_g, ve.FrameInputs(W, H, f, frame_seed(f), f + 1))
            of.frame(lights_scene["words"], lights_scene["depth"], ubo, fc_o, f, frame_seed(f), f + 1)
            cg, co = rend.counters(), of.counters()
            assert cg[2] == co[2] and cg[4] == co[4] and cg[5] == co[5]   # registry size, traced, shadow rays
        ne = int(cg[2])
        assert np.array_equal(rend.read("EMISSIVES", np.uint32, ne * 32).reshape(-1, 8),
                              of.read("EMISSIVES").reshape(-1, 8)[:ne])
        gl, ol = hash_map(rend.read("LBUFFER"), 12, 32), hash_map(of.read("LBUFFER"), 12, 32)
        assert np.array_equal(gl[0], ol[0]) and np.array_equal(gl[1], ol[1])
        assert_out_close(rend.read("OUTPUT", np.float32, W * H * 16), of.read("OUTPUT", np.float32))
        rend.release(sc)
    finally:
        rend.close()
        of.close()
