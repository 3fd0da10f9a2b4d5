"""GPU parity on the llvmpipe fixtures of the order-sensitive and material
paths (tests/golden/make_golden.py: overflow, material), through the C ABI:

  - 64-bucket hash images, every bucket full: the nBuffer's and lBuffer's raw
    images (slot order included) bit-exact after every run, under
    deterministic, multi-sample (halving, firefly) and default knobs with a
    moving camera -- the ordered per-bucket replay of normal.comp and
    depositSampleDual (ve_passes.hip);
  - GGX VNDF / Blinn-Phong materials, the demo's caps 16/8, and 72 frames
    reaching the trusted resolve path: lBuffer sums bit-exact, output within
    1e-4, and every debug render type of resolve.comp;
  - the same frames dispatched pass by pass (ve_pass) equal ve_frame."""
import json
import os

import numpy as np
import pytest

from frame_common import assert_out_close, frame_seed, hash_map

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLD, name + ".json")) as fh:
        meta = json.load(fh)
    return meta, np.load(os.path.join(GOLD, name + ".npz"))


def moving_ubo(ve, scenes, W, H, f):
    eye, d, fov = scenes.camera_small()
    e = np.asarray(eye, np.float32) + np.float32(2.0 * f) * np.array([0.5, 0.0, 1.0], np.float32)
    return ve.camera_ubo(e, d, fov, W / H)


PASSES = ("RESET", "PRIMARY", "NORMAL", "SCHEDULE", "RANK", "THRESHOLD", "EMIT")


@pytest.mark.parametrize("by_pass", [False, True])
@pytest.mark.parametrize("run", ["ovf_det", "ovf_spp", "ovf_default"])
def test_overflow_runs_match_llvmpipe(ve, scenes, small_scene, run, by_pass):
    meta, arr = _load("frame_overflow")
    r = meta["runs"][run]
    W, H, hw, mb = r["W"], r["H"], meta["hash_width"], r["max_bounces"]
    rend = ve.Renderer(W, H, hash_width=hw, max_bounces=mb)
    try:
        rend.set_materials(scenes.palette(1))
        rend.set_sky_constant(meta["sky_rgb8"])
        sc = rend.upload(small_scene["words"], small_scene["depth"], small_scene["info"].size)
        fc = ve.FrameConfig(**r["knobs"])
        counters = []
        for f in range(r["frames"]):
            print("case overflow", run, "by_pass", by_pass, "frame", f, flush=True)
            rend.set_camera(moving_ubo(ve, scenes, W, H, f))
            inp = ve.FrameInputs(W, H, f, frame_seed(f), f + 1)
            if by_pass:
                for name in PASSES:
                    rend.run_pass(sc, name, fc, inp)
                for _ in range(mb):
                    rend.run_pass(sc, "TRACE", fc, inp)
                    rend.run_pass(sc, "SHADE", fc, inp)
                rend.run_pass(sc, "RESOLVE", fc, inp)
            else:
                rend.frame(sc, fc, inp)
            c = rend.counters()
            counters.append([c[0], c[1], c[7], c[4]])
        assert np.array_equal(np.array(counters, np.uint32), arr[run + "_counters"])
        g, _ = rend.gbuffer(W, H)
        assert np.array_equal(g, arr[run + "_gbuf"])
        assert np.array_equal(rend.read("NBUFFER").reshape(hw, 96), arr[run + "_nbuf"])
        assert np.array_equal(rend.read("LBUFFER").reshape(hw, 384), arr[run + "_lbuf"])
        assert_out_close(rend.read("OUTPUT", np.float32, W * H * 16), arr[run + "_out"])
        rend.release(sc)
    finally:
        rend.close()


@pytest.mark.parametrize("run", ["spec_det", "caps_det", "long_det"])
def test_material_runs_match_llvmpipe(ve, scenes, small_scene, run):
    meta, arr = _load("frame_material")
    r = meta["runs"][run]
    W, H = r["W"], r["H"]
    rend = ve.Renderer(W, H, hash_width=16384, max_bounces=r["max_bounces"])
    try:
        rend.set_materials(arr[run + "_materials"].view(scenes.MATERIAL_DTYPE))
        rend.set_sky_constant(meta["sky_rgb8"])
        rend.set_camera(arr[run + "_ubo"])
        sc = rend.upload(small_scene["words"], small_scene["depth"], small_scene["info"].size)
        fc = ve.FrameConfig(**r["knobs"])
        counters = []
        for f in range(r["frames"]):
            if f % 10 == 0:
                print("case material", run, "frame", f, flush=True)
            rend.frame(sc, fc, ve.FrameInputs(W, H, f, frame_seed(f), f + 1))
            c = rend.counters()
            counters.append([c[0], c[1], c[7], c[4]])
        assert np.array_equal(np.array(counters, np.uint32), arr[run + "_counters"])
        lv, lval = hash_map(rend.read("LBUFFER"), 12, 32)
        assert np.array_equal(lv, arr[run + "_lbuf_vids"]) and np.array_equal(lval, arr[run + "_lbuf_vals"])
        assert_out_close(rend.read("OUTPUT", np.float32, W * H * 16), arr[run + "_out"])
        if run == "long_det":
            last = r["frames"] - 1
            for rt in meta["resolve_types"]:
                kn = ve.FrameConfig(**dict(r["knobs"], render_type=rt, primary_controlchecks=meta["resolve_checks"]))
                rend.run_pass(sc, "RESOLVE", kn, ve.FrameInputs(W, H, last, frame_seed(last), last + 1))
                assert_out_close(rend.read("OUTPUT", np.float32, W * H * 16), arr["%s_resolve%d" % (run, rt)])
        rend.release(sc)
    finally:
        rend.close()
