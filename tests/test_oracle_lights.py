"""Pin the C oracle's next-event estimation (shade.comp:276-539 ReSTIR DI at
bounce 0, shade.comp:816-836 single-pick RIS at bounce 1, the emissive
registry of shade.comp:592-619) against the reference GLSL run on llvmpipe on
an emissive scene (tests/golden/frame_lights.*): the registry (order
included), every lBuffer integer sum, every reservoir of the rBuffer and the
per-frame counters bit-exact; the resolved image within 1e-4 relative.  Two
runs: temporal reuse only, and temporal + spatial reuse."""
import json
import os

import numpy as np
import pytest

from frame_common import GOLD, assert_out_close, frame_seed, hash_map


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


def run_oracle(oracle, scenes, meta, arr, run, scene):
    r = meta["runs"][run]
    W, H = r["W"], r["H"]
    fc = oracle.FrameConfig(**r["knobs"])
    of = oracle.Frame(W, H, 16384, r["max_bounces"])
    of.set_materials(scenes.palette_lights(1))
    of.set_sky_rgb8(meta["sky_rgb8"])
    counters = []
    for f in range(r["frames"]):
        of.frame(scene["words"], scene["depth"], arr[run + "_ubo"], fc, f, frame_seed(f), f + 1)
        c = of.counters()
        counters.append([c[0], c[1], c[7], c[4], c[2]])
    return of, counters


@pytest.mark.parametrize("run", ["lights_temporal", "lights_spatial"])
def test_oracle_nee_matches_llvmpipe(lights_gold, lights_scene, oracle, scenes, run):
    meta, arr = lights_gold
    of, counters = run_oracle(oracle, scenes, meta, arr, run, lights_scene)
    try:
        assert np.array_equal(np.array(counters, np.uint32), arr[run + "_counters"])
        ne = int(counters[-1][4])
        assert ne > 0
        em = of.read("EMISSIVES").reshape(-1, 8)[:ne]
        assert np.array_equal(em, arr[run + "_emissives"])
        lv, lval = hash_map(of.read("LBUFFER"), 12, 32)
        assert np.array_equal(lv, arr[run + "_lbuf_vids"])
        assert np.array_equal(lval, arr[run + "_lbuf_vals"])
        rv, rval = hash_map(of.read("RBUFFER"), 8, 16)
        assert np.array_equal(rv, arr[run + "_rbuf_vids"])
        assert np.array_equal(rval, arr[run + "_rbuf_vals"])
        assert_out_close(of.read("OUTPUT", np.float32), arr[run + "_out"])
    finally:
        of.close()


def test_nee_adds_direct_light(lights_gold):
    """Mutation check on the fixtures themselves: the scene's lights reach the
    direct channel (non-zero direct sums in the lBuffer), which only
    next-event estimation and bounce-0 emissive hits produce."""
    meta, arr = lights_gold
    vals = arr["lights_temporal_lbuf_vals"]
    assert (vals[:, 1] > 0).sum() > 0.5 * len(vals)   # direct red sums
