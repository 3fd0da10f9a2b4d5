"""Pin the C oracle's whole wavefront frame (Renderer::runWavefrontFrame,
src/renderer/renderer.cpp:282-649: normal, schedule, rank, threshold, emit,
trace, shade, resolve) against the reference GLSL run on llvmpipe
(tests/golden/frame_small.*): per-frame counters, the candidate and shade
lists (order included), per-voxel pixel counts, RankStats, the nBuffer normals
and every lBuffer integer sum bit-exact; the resolved RGBA32F image within
1e-4 relative."""
import numpy as np
import pytest

from frame_common import assert_out_close, frame_seed, hash_map, load_frame_gold


@pytest.fixture(scope="module")
def gold():
    return load_frame_gold()


@pytest.mark.parametrize("run", ["small_det", "small_default"])
def test_oracle_frames_match_llvmpipe(gold, oracle, small_scene, scenes, run):
    meta, arr = gold
    r = meta["runs"][run]
    W, H = r["W"], r["H"]
    fc = oracle.FrameConfig(**r["knobs"])
    of = oracle.Frame(W, H, 16384, r["max_bounces"])
    of.set_materials(scenes.palette(1))
    of.set_sky_rgb8(meta["sky_rgb8"])
    words, depth = small_scene["words"], small_scene["depth"]
    counters = []
    for f in range(r["frames"]):
        of.frame(words, depth, arr[run + "_ubo"], fc, f, frame_seed(f), f + 1)
        c = of.counters()
        counters.append([c[0], c[1], c[7], c[4]])
        if f == 0:
            assert np.array_equal(of.read("GBUFFER").reshape(H, W, 4), arr[run + "_gbuf"])
            assert_out_close(of.read("OUTPUT", np.float32), arr[run + "_out0"])
    assert np.array_equal(np.array(counters, np.uint32), arr[run + "_counters"])
    lv, lval = hash_map(of.read("LBUFFER"), 12, 32)
    assert np.array_equal(lv, arr[run + "_lbuf_vids"]) and np.array_equal(lval, arr[run + "_lbuf_vals"])
    nv, nval = hash_map(of.read("NBUFFER"), 3, 32)
    assert np.array_equal(nv, arr[run + "_nbuf_vids"]) and np.array_equal(nval[:, 0], arr[run + "_nbuf_vals"][:, 0])
    assert np.array_equal(of.read("CANDIDATES").reshape(-1, 4), arr[run + "_candidates"])
    assert np.array_equal(of.read("SHADE_LIST").reshape(-1, 2), arr[run + "_shade_list"])
    sc = of.read("SAMPLE_COUNTS").reshape(16384, 64).T
    assert np.array_equal(sc, arr[run + "_sample_counts"].reshape(64, 16384))
    assert np.array_equal(of.read("RANK_STATS"), arr[run + "_rank_stats"])
    assert_out_close(of.read("OUTPUT", np.float32), arr[run + "_out"])
    of.close()


def test_sample_counts_are_vid_histogram(gold, oracle, small_scene, scenes):
    """Per-voxel hit counts (schedule.comp:92-120) equal the histogram of the
    voxel-ID image: a closed form independent of the claim-hash layout."""
    meta, arr = gold
    r = meta["runs"]["small_det"]
    g = arr["small_det_gbuf"][..., 3].ravel()
    vids, counts = np.unique(g[g != 0], return_counts=True)
    cand = arr["small_det_candidates"]
    sc = arr["small_det_sample_counts"].reshape(64, 16384)
    got = dict(zip(cand[:, 0].tolist(), sc[cand[:, 3], cand[:, 2]].tolist()))
    assert len(got) == len(vids)
    assert all(got[int(v)] == int(c) for v, c in zip(vids, counts))
    # representative pixel of each candidate carries its vid
    assert np.all(g[cand[:, 1]] == cand[:, 0])
    assert r["W"] * r["H"] == len(g)
