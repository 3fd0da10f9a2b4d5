"""Pin the C oracle's generate_primary + RaycastCounted
(src/shd/generate_primary.comp:38-85, src/shd/internal.glsl:245-400) against
the reference GLSL run on Mesa llvmpipe (tests/golden/primary_small.*), and
against closed forms: the camera UBO and the voxel key hash.
"""
import hashlib
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def gold():
    with open(os.path.join(GOLD, "primary_small.json")) as fh:
        meta = json.load(fh)
    return meta, np.load(os.path.join(GOLD, "primary_small.npz"))


def test_scene_matches_fixture(gold, small_scene):
    meta, _ = gold
    info = small_scene["info"]
    assert (info.size, info.capacity, info.num_voxels) == (meta["size"], meta["capacity"], meta["num_voxels"])
    dig = hashlib.sha256(np.ascontiguousarray(small_scene["words"][: info.size]).tobytes()).hexdigest()
    assert dig == meta["words_sha256"]


def test_camera_ubo_matches_fixture(gold, ve, oracle, scenes):
    _, arr = gold
    eye, d, fov = scenes.camera_small()
    assert np.array_equal(ve.camera_ubo(eye, d, fov, 1.0), arr["small_cfg_ubo"])
    assert np.array_equal(oracle.camera_ubo(eye, d, fov, 1.0), arr["small_cfg_ubo"])


def test_camera_ubo_closed_form(ve):
    # looking down +z with 90 deg fov: right = (1,0,0)*tan(45)*aspect, up = (0,-1,0)*tan(45)
    u = ve.camera_ubo((1, 2, 3), (0, 0, 1), 90.0, 2.0)
    t = np.float32(np.tan(np.float32(45.0) * np.float32(0.017453292519943295)))
    assert np.array_equal(u[:3], [1, 2, 3]) and np.array_equal(u[4:7], [0, 0, 1])
    assert np.array_equal(u[8:11], np.float32([t * 2, 0, 0]))
    assert np.array_equal(u[12:15], np.float32([0, -t, 0]))


def test_voxel_key_closed_form(oracle):
    def h(x):
        x ^= x >> 16
        x = (x * 0x7FEB352D) & 0xFFFFFFFF
        x ^= x >> 15
        x = (x * 0x846CA68B) & 0xFFFFFFFF
        x ^= x >> 16
        return x & 0x7FFFFFFF
    for (x, y, z) in [(0, 0, 0), (1, 2, 3), (1023, 1023, 1023), (1024, 0, 0), (5, 700, 12)]:
        assert oracle.voxel_key(x, y, z) == h((x | (y << 10) | (z << 20)) & 0xFFFFFFFF)
    assert oracle.voxel_key(1024, 0, 0) == oracle.voxel_key(0, 1, 0)   # 10-bit aliasing (internal.glsl:284)


@pytest.mark.parametrize("view", ["small_cfg", "inside", "axis_z", "axis_y", "grazing"])
def test_oracle_primary_bit_exact(gold, oracle, small_scene, view):
    meta, arr = gold
    v = meta["views"][view]
    g, p, _ = oracle.generate_primary(small_scene["words"], small_scene["depth"], arr[view + "_ubo"], v["W"], v["H"],
                                      v["controlchecks"])
    assert np.array_equal(g, arr[view + "_gbuf"])
    assert np.array_equal(p, arr[view + "_pos"])


FP_OPS = {"dot": 0, "normalize": 1, "div": 2, "mad": 3, "min": 4, "max": 5}


@pytest.fixture(scope="module")
def fp():
    return np.load(os.path.join(GOLD, "fp_probes.npz"))


@pytest.mark.parametrize("name", sorted(FP_OPS))
def test_fp_primitives_match_llvmpipe(fp, oracle, name):
    """The oracle's FP primitives equal llvmpipe's bit for bit on 4096 random
    inputs plus NaN / signed-zero / infinity cases."""
    out = oracle.fp_probe(FP_OPS[name], fp["inputs"])
    ref = fp[name]
    if name in ("min", "max"):
        same = (out.view(np.uint32) == ref.view(np.uint32))
        assert same.all()
    else:
        ok = np.isfinite(ref) if name != "normalize" else np.isfinite(ref).all(axis=1)
        assert np.array_equal(out[ok].view(np.uint32), ref[ok].view(np.uint32))


@pytest.mark.parametrize("variant,op", [(1, "mad"), (2, "dot"), (3, "normalize"), (4, "div")])
def test_probes_reject_other_fp_lowerings(fp, oracle, variant, op):
    """Each plausible alternative lowering (fma contraction, left-to-right
    dot, v/len normalize, reciprocal division) disagrees with llvmpipe."""
    lib = oracle.lib()
    lib.so_set_fp_variant(variant)
    try:
        out = oracle.fp_probe(FP_OPS[op], fp["inputs"])
    finally:
        lib.so_set_fp_variant(0)
    ref = fp[op]
    ok = np.isfinite(ref) if op != "normalize" else np.isfinite(ref).all(axis=1)
    assert not np.array_equal(out[ok].view(np.uint32), ref[ok].view(np.uint32))


def test_primary_fixture_rejects_fma(gold, small_scene, oracle):
    """An FMA-contracted traversal (d*t + o fused) changes voxel IDs on the
    primary fixtures: the pass-level fixture pins the contraction policy."""
    meta, arr = gold
    lib = oracle.lib()
    lib.so_set_fp_variant(1)
    try:
        diff = 0
        for view, v in meta["views"].items():
            g, p, _ = oracle.generate_primary(small_scene["words"], small_scene["depth"], arr[view + "_ubo"], v["W"],
                                              v["H"], v["controlchecks"])
            diff += int((g != arr[view + "_gbuf"]).any(axis=-1).sum())
    finally:
        lib.so_set_fp_variant(0)
    assert diff > 0
