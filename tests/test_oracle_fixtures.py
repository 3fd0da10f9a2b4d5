"""The C oracle against the llvmpipe fixtures of the paths the other
fixtures leave out (tests/golden/make_golden.py: fp, overflow, material):

  - GLSL built-ins of the shading code: inversesqrt (the VNDF frame,
    internal.glsl:199) is 1/sqrt exactly, sqrt exact; log, log2, exp, exp2,
    pow, sin, cos within the bounds stated below (llvmpipe's polynomial cores
    against correctly rounded ones: transcendentals are pinned within
    tolerance, not bit for bit);
  - hash-bucket overflow (64-bucket hash images): LRU eviction in the nBuffer
    (normal.comp:160-176, one invocation after another) and the lBuffer
    (shade.comp:162-257, SIMD-8 lanes in lockstep), raw images bit-exact,
    slot order included, under deterministic, multi-sample and default knobs;
  - GGX VNDF + Blinn-Phong materials (shade.comp:847-873, internal.glsl:120-239),
    the demo's caps 16/8 with halving and the firefly clamp
    (shade.comp:188-234), and 72 accumulated frames that reach the trusted
    resolve path (resolve.comp:235) and every debug view (resolve.comp:266-382).
"""
import json
import os

import numpy as np
import pytest

from frame_common import assert_out_close, frame_seed, hash_map

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BUILTIN_OPS = {"inversesqrt": 6, "sqrt": 7, "pow": 8, "exp": 9, "log": 10, "exp2": 11, "log2": 12, "sin": 13, "cos": 14}
# max relative difference llvmpipe vs the oracle's cores (measured: pow
# 1.14e-5 with exponents up to 128 -- log2's absolute error near 1 times the
# exponent --, exp 5.5e-6, log/log2 1.8e-5 where the result is near 0, exp2
# 2.6e-7, sin/cos 1.2e-7), and for log/log2 an absolute bound of 5e-7
BUILTIN_REL = {"pow": 2e-5, "exp": 1e-5, "log": 3e-5, "log2": 3e-5, "exp2": 5e-7, "sin": 2.5e-7, "cos": 2.5e-7}


@pytest.fixture(scope="module")
def fp():
    return np.load(os.path.join(GOLD, "fp_probes.npz"))


def test_inversesqrt_and_sqrt_are_exact(fp, oracle):
    """inversesqrt(x) == 1.0 / sqrt(x) bit for bit on llvmpipe (the kernels and
    the oracle use exactly that), and sqrt is correctly rounded; the product
    of two roundings differs from a correctly rounded 1/sqrt in 29% of these
    inputs, so the probe tells them apart."""
    for name in ("inversesqrt", "sqrt"):
        out = oracle.fp_probe(BUILTIN_OPS[name], fp["inputs"])
        assert np.array_equal(out.view(np.uint32), fp[name].view(np.uint32)), name
    x = (np.abs(fp["inputs"][:, 0]) + np.float32(1e-6)).astype(np.float32)
    once = (1.0 / np.sqrt(x.astype(np.float64))).astype(np.float32)
    assert np.mean(once != fp["inversesqrt"]) > 0.2


@pytest.mark.parametrize("name", sorted(BUILTIN_REL))
def test_transcendentals_within_bounds(fp, oracle, name):
    out = oracle.fp_probe(BUILTIN_OPS[name], fp["inputs"]).astype(np.float64)
    ref = fp[name].astype(np.float64)
    fin = np.isfinite(ref)
    assert np.array_equal(fin, np.isfinite(out)) or name == "pow"   # pow: overflow at the top of the range
    ok = fin & np.isfinite(out) & (np.abs(ref) > 1e-30)
    rel = np.abs(out[ok] - ref[ok]) / np.abs(ref[ok])
    if name in ("log", "log2"):
        assert np.all((rel <= BUILTIN_REL[name]) | (np.abs(out[ok] - ref[ok]) <= 5e-7))
    else:
        assert rel.max() <= BUILTIN_REL[name], rel.max()
    assert np.mean(rel == 0.0) > 0.2   # not a loose match: a good share is bit-identical


def _load(name):
    with open(os.path.join(GOLD, name + ".json")) as fh:
        meta = json.load(fh)
    return meta, np.load(os.path.join(GOLD, name + ".npz"))


def moving_ubo(ve, scenes, W, H, f):
    eye, d, fov = scenes.camera_small()
    e = np.asarray(eye, np.float32) + np.float32(2.0 * f) * np.array([0.5, 0.0, 1.0], np.float32)
    return ve.camera_ubo(e, d, fov, W / H)


@pytest.mark.parametrize("run", ["ovf_det", "ovf_spp", "ovf_default"])
def test_overflow_runs_match_llvmpipe(ve, scenes, oracle, small_scene, run):
    meta, arr = _load("frame_overflow")
    r = meta["runs"][run]
    W, H, hw = r["W"], r["H"], meta["hash_width"]
    of = oracle.Frame(W, H, hw, r["max_bounces"])
    try:
        of.set_materials(scenes.palette(1))
        of.set_sky_rgb8(meta["sky_rgb8"])
        fc = oracle.FrameConfig(**r["knobs"])
        counters = []
        for f in range(r["frames"]):
            of.frame(small_scene["words"], small_scene["depth"], moving_ubo(ve, scenes, W, H, f), fc, f,
                     frame_seed(f), f + 1, n_words=small_scene["info"].size)
            c = of.counters()
            counters.append([c[0], c[1], c[7], c[4]])
        assert np.array_equal(np.array(counters, np.uint32), arr[run + "_counters"])
        assert np.array_equal(of.read("GBUFFER").reshape(H, W, 4), arr[run + "_gbuf"])
        nb = of.read("NBUFFER")[: hw * 96].reshape(hw, 96)
        lb = of.read("LBUFFER")[: hw * 384].reshape(hw, 384)
        assert np.array_equal(nb, arr[run + "_nbuf"])
        assert np.array_equal(lb, arr[run + "_lbuf"])
        assert_out_close(of.read("OUTPUT", np.float32), arr[run + "_out"])
    finally:
        of.close()
    # the runs overflow: every bucket is full, far more voxels than slots
    own = arr[run + "_lbuf"].reshape(hw, 32, 12)[:, :, 0]
    assert np.all((own[1:] != 0).sum(1) == 32)


def _material_run(ve, scenes, oracle, small_scene, meta, arr, run, keep=False):
    r = meta["runs"][run]
    W, H = r["W"], r["H"]
    of = oracle.Frame(W, H, 16384, r["max_bounces"])
    of.set_materials(arr[run + "_materials"].view(scenes.MATERIAL_DTYPE))
    of.set_sky_rgb8(meta["sky_rgb8"])
    fc = oracle.FrameConfig(**r["knobs"])
    counters = []
    for f in range(r["frames"]):
        of.frame(small_scene["words"], small_scene["depth"], arr[run + "_ubo"], fc, f, frame_seed(f), f + 1,
                 n_words=small_scene["info"].size)
        c = of.counters()
        counters.append([c[0], c[1], c[7], c[4]])
    return of, fc, np.array(counters, np.uint32)


@pytest.mark.parametrize("run", ["spec_det", "caps_det", "long_det"])
def test_material_runs_match_llvmpipe(ve, scenes, oracle, small_scene, run):
    meta, arr = _load("frame_material")
    of, fc, counters = _material_run(ve, scenes, oracle, small_scene, meta, arr, run)
    try:
        assert np.array_equal(counters, arr[run + "_counters"])
        lv, lval = hash_map(of.read("LBUFFER"), 12, 32)
        assert np.array_equal(lv, arr[run + "_lbuf_vids"]) and np.array_equal(lval, arr[run + "_lbuf_vals"])
        assert_out_close(of.read("OUTPUT", np.float32), arr[run + "_out"])
        if run == "long_det":
            W, H = meta["runs"][run]["W"], meta["runs"][run]["H"]
            words, depth = small_scene["words"], small_scene["depth"]
            for rt in meta["resolve_types"]:
                kn = oracle.FrameConfig(**dict(meta["runs"][run]["knobs"], render_type=rt,
                                               primary_controlchecks=meta["resolve_checks"]))
                of.resolve(words, depth, arr[run + "_ubo"], kn, n_words=small_scene["info"].size)
                assert_out_close(of.read("OUTPUT", np.float32), arr["%s_resolve%d" % (run, rt)])
    finally:
        of.close()


def test_material_fixtures_reach_their_paths(scenes):
    """What the runs were made for actually happens in them."""
    meta, arr = _load("frame_material")
    spec = arr["spec_det_materials"].view(scenes.MATERIAL_DTYPE)
    assert spec[3]["metallic"] * (1 - spec[3]["roughness"]) > 0.5           # GGX VNDF entry
    assert spec[5]["specular"] * spec[5]["metallic"] > 0 and spec[7]["specular"] * spec[7]["metallic"] > 0
    # caps 16 / 8: halving keeps every count below its cap plus one
    v = arr["caps_det_lbuf_vals"]
    assert v[:, 0].max() <= 16 and v[:, 5].max() <= 8 and (v[:, 5] == 8).any()
    # 72 frames: every visible voxel passes the trusted threshold (count >= 128,
    # variance < 0.02), so the whole image takes the trusted path; the
    # bilateral path below it is what every other frame fixture resolves
    lv = arr["long_det_lbuf_vals"]
    cnt = lv[:, 0].astype(np.int64) + lv[:, 5]
    ic = np.maximum(lv[:, 5], 1).astype(np.float64)
    mean_l = (lv[:, 6].astype(np.float64) + lv[:, 7] + lv[:, 8]) / ic / 3.0
    var = np.maximum(lv[:, 9] / ic - mean_l ** 2, 0.0) / (255.0 * 255.0)
    assert ((cnt >= 128) & (var < 0.02)).sum() > 1000
    small = np.load(os.path.join(GOLD, "frame_small.npz"))
    sv = small["small_det_lbuf_vals"]
    assert (sv[:, 0].astype(np.int64) + sv[:, 5]).max() < 128


def test_llvmpipe_loop_limit():
    """llvmpipe stops a shader's loops after 65535 iterations in total per
    SIMD vector, nested loops included (gallivm's loop limiter); GLSL has no
    such limit and neither do the oracle nor the kernels.  The parity runs stay
    below it (the overflow run with default knobs uses normal radius 1:
    radius 6 under heavy eviction exceeds it)."""
    lp = np.load(os.path.join(GOLD, "loop_limit.npz"))
    assert np.all(lp["single_60000"] == 60000)
    assert np.all(lp["single_100000"] == 65535)
    assert np.all(lp["nested_300x300"] < 90000) and np.all(lp["nested_40x40x40"] < 64000)
