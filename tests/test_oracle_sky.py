"""Pin the oracle's cube-map sky (texture(samplerCube, dir) under the
reference skybox's sampler state, skybox.cpp:87-122) against the GL driver
the reference runs on (tests/golden/cube_probes.npz: six random RGB8 faces
per edge size, random directions plus axes, face ties, cube edges and
corners, unnormalised vectors and texel centres), and the whole frame with
a non-constant sky (tests/golden/frame_sky.*): escaping bounce rays sample
it into the lBuffer (shade.comp:559), sky pixels and timed-out pixels with
no usable neighbour resolve to it (resolve.comp:218,227)."""
import json
import os

import numpy as np
import pytest

from frame_common import assert_out_close, frame_seed, hash_map

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def probes():
    return np.load(os.path.join(GOLD, "cube_probes.npz"))


@pytest.mark.parametrize("edge", [1, 2, 4, 7, 16])
def test_cube_sampling_matches_driver(oracle, probes, edge):
    of = oracle.Frame(8, 8, hash_width=64, max_bounces=2)
    of.set_sky_cube(probes["faces_%d" % edge])
    got = of.sky_sample(probes["dirs_%d" % edge])
    assert np.array_equal(got, probes["rgb_%d" % edge])
    of.close()


def test_cube_sampling_closed_forms(oracle):
    """Cases the GL spec fixes without the driver: a uniform cube returns its
    texel for every direction; the centre of a face returns the byte-exact
    average of its four centre texels (weights 1/2); a texel centre returns
    that texel; the six axes select the six faces in +X -X +Y -Y +Z -Z order."""
    rng = np.random.default_rng(3)
    of = oracle.Frame(8, 8, hash_width=64, max_bounces=2)
    dirs = rng.normal(size=(512, 3)).astype(np.float32)
    for c in ((0, 0, 0), (255, 255, 255), (230, 7, 129)):
        of.set_sky_cube(np.broadcast_to(np.array(c, np.uint8), (6, 5, 5, 3)))
        assert np.all(of.sky_sample(dirs) == np.float32(np.array(c, np.float32) * np.float32(1.0 / 255.0)))
    faces = rng.integers(0, 256, (6, 4, 4, 3)).astype(np.uint8)
    of.set_sky_cube(faces)
    axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)
    got = np.rint(of.sky_sample(axes) * 255).astype(int)
    for f in range(6):
        q = faces[f, 1:3, 1:3].astype(int)
        lerp = lambda a, b: a + (((b - a) * 128 + 128) >> 8)
        want = lerp(lerp(q[0, 0], q[0, 1]), lerp(q[1, 0], q[1, 1]))
        assert np.array_equal(got[f], want), f
    # texel (i, j) of +Z: s = (i + 1/2) / edge, t = (j + 1/2) / edge, dir = (2s - 1, -(2t - 1), 1)
    for i, j in ((0, 0), (3, 1), (2, 3)):
        d = np.array([[(i + 0.5) / 2 - 1, -((j + 0.5) / 2 - 1), 1.0]], np.float32)
        assert np.array_equal(np.rint(of.sky_sample(d)[0] * 255).astype(int), faces[4, j, i].astype(int))
    of.close()


@pytest.mark.parametrize("run", ["sky_det", "sky_timeout"])
def test_oracle_sky_frames_match_llvmpipe(oracle, small_scene, scenes, run):
    meta = json.load(open(os.path.join(GOLD, "frame_sky.json")))
    arr = np.load(os.path.join(GOLD, "frame_sky.npz"))
    r = meta["runs"][run]
    W, H = r["W"], r["H"]
    fc = oracle.FrameConfig(**r["knobs"])
    of = oracle.Frame(W, H, 16384, r["max_bounces"])
    of.set_materials(scenes.palette(1))
    of.set_sky_cube(arr["sky_faces"])
    words, depth = small_scene["words"], small_scene["depth"]
    counters = []
    for f in range(r["frames"]):
        of.frame(words, depth, arr[run + "_ubo"], fc, f, frame_seed(f), f + 1)
        c = of.counters()
        counters.append([c[0], c[1], c[7], c[4]])
    assert np.array_equal(np.array(counters, np.uint32), arr[run + "_counters"])
    g = of.read("GBUFFER").reshape(H, W, 4)
    assert np.array_equal(g, arr[run + "_gbuf"])
    lv, lval = hash_map(of.read("LBUFFER"), 12, 32)
    assert np.array_equal(lv, arr[run + "_lbuf_vids"]) and np.array_equal(lval, arr[run + "_lbuf_vals"])
    out = of.read("OUTPUT", np.float32)
    assert_out_close(out, arr[run + "_out"])
    # the sky shows: the miss pixels are not one colour
    miss = (g[..., 3] == 0).ravel()
    assert miss.sum() > 0 and len(np.unique(out.reshape(-1, 4)[miss, :3], axis=0)) > 1
    of.close()
