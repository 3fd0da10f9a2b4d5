"""GPU parity of RayPass's primary stage (generate_primary.comp + RaycastCounted)
through the C ABI: bit-exact against the llvmpipe golden fixtures and the C
oracle."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def gold():
    with open(os.path.join(GOLD, "primary_small.json")) as fh:
        meta = json.load(fh)
    return meta, np.load(os.path.join(GOLD, "primary_small.npz"))


def _primary(ve, renderer, scene, ubo, W, H, cc):
    renderer.set_camera(ubo)
    fc = ve.FrameConfig(primary_controlchecks=cc)
    renderer.run_pass(scene, "PRIMARY", fc, ve.FrameInputs(W, H, 0, 1, 1))
    return renderer.gbuffer(W, H)


@pytest.mark.parametrize("view", ["small_cfg", "inside", "axis_z", "axis_y", "grazing"])
def test_primary_matches_llvmpipe(ve, renderer, small_scene, gold, view):
    meta, arr = gold
    v = meta["views"][view]
    print("case primary", view, v, flush=True)
    scene = renderer.upload(small_scene["words"], small_scene["depth"], small_scene["info"].size)
    try:
        g, p = _primary(ve, renderer, scene, arr[view + "_ubo"], v["W"], v["H"], v["controlchecks"])
    finally:
        renderer.release(scene)
    assert np.array_equal(g, arr[view + "_gbuf"])
    assert np.array_equal(p, arr[view + "_pos"])


def test_primary_random_cameras_vs_oracle(ve, renderer, oracle, small_scene):
    rng = np.random.default_rng(7)
    scene = renderer.upload(small_scene["words"], small_scene["depth"], small_scene["info"].size)
    try:
        for k in range(6):
            eye = rng.uniform(-60, 190, 3).astype(np.float32)
            tgt = rng.uniform(20, 110, 3)
            d = (tgt - eye) / np.linalg.norm(tgt - eye)
            W, H = int(rng.integers(17, 200)), int(rng.integers(9, 150))
            ubo = ve.camera_ubo(eye, d.astype(np.float32), float(rng.uniform(30, 110)), W / H)
            cc = int(rng.integers(1, 130))
            print("case random camera", k, W, H, cc, flush=True)
            g, p = _primary(ve, renderer, scene, ubo, W, H, cc)
            go, po, _ = oracle.generate_primary(small_scene["words"], small_scene["depth"], ubo, W, H, cc)
            assert np.array_equal(g, go) and np.array_equal(p, po)
    finally:
        renderer.release(scene)


def test_node_read_stats_match_oracle(ve, renderer, oracle, small_scene, gold):
    meta, arr = gold
    v = meta["views"]["small_cfg"]
    scene = renderer.upload(small_scene["words"], small_scene["depth"], small_scene["info"].size)
    try:
        renderer.enable_node_stats(True)
        _primary(ve, renderer, scene, arr["small_cfg_ubo"], v["W"], v["H"], v["controlchecks"])
        st = renderer.node_stats()
        renderer.enable_node_stats(False)
    finally:
        renderer.release(scene)
    _, _, ost = oracle.generate_primary(small_scene["words"], small_scene["depth"], arr["small_cfg_ubo"], v["W"],
                                        v["H"], v["controlchecks"])
    assert st["primary_reads"] == ost["node_reads"] and st["primary_steps"] == ost["steps"]


def test_malformed_octree_rejected(ve, renderer):
    words = np.zeros(16, np.uint32)
    words[0] = 1 | (1 << 1) | (100 << 4)   # next block outside the array
    with pytest.raises(ve.VeError) as e:
        renderer.upload(words, 2)
    assert e.value.code == ve.VE_EFORMAT


def test_zero_viewport_rejected(ve, renderer, small_scene):
    scene = renderer.upload(small_scene["words"], small_scene["depth"], small_scene["info"].size)
    try:
        with pytest.raises(ve.VeError) as e:
            renderer.run_pass(scene, "PRIMARY", ve.FrameConfig(), ve.FrameInputs(0, 10, 0, 1, 1))
        assert e.value.code == ve.VE_EINVAL
    finally:
        renderer.release(scene)
