"""GPU parity of the whole wavefront frame (Renderer::runWavefrontFrame,
src/renderer/renderer.cpp:282-649) through the C ABI, against the reference
GLSL run on llvmpipe (tests/golden/frame_small.*) and against the C oracle.

Bit-exact: voxel-ID GBuffer, candidate list (order and representative pixel),
shade list, per-voxel pixel counts, RankStats, nBuffer normals and every
lBuffer integer sum (AccumPass).  Resolved RGBA32F (AvgPass) within 1e-4
relative."""
import numpy as np
import pytest

from frame_common import DET_KNOBS, assert_out_close, frame_seed, hash_map, load_frame_gold

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gold():
    return load_frame_gold()


def _fc(ve, knobs):
    return ve.FrameConfig(**knobs)


def _gpu_run(ve, scenes, small_scene, meta, arr, run, frames=None):
    r = meta["runs"][run]
    W, H = r["W"], r["H"]
    rend = ve.Renderer(W, H, hash_width=16384, max_bounces=r["max_bounces"])
    rend.set_materials(scenes.palette(1))
    rend.set_sky_constant(meta["sky_rgb8"])
    rend.set_camera(arr[run + "_ubo"])
    scene = rend.upload(small_scene["words"], small_scene["depth"], small_scene["info"].size)
    fc = _fc(ve, r["knobs"])
    counters, first = [], None
    for f in range(r["frames"] if frames is None else frames):
        print("case frame", run, f, flush=True)
        rend.frame(scene, fc, ve.FrameInputs(W, H, f, frame_seed(f), f + 1))
        c = rend.counters()
        counters.append([c[0], c[1], c[7], c[4]])
        assert c[ve.CNT["OVERFLOW"]] == 0 if "OVERFLOW" in ve.CNT else True
        if f == 0:
            g, _ = rend.gbuffer(W, H)
            first = (g, rend.read("OUTPUT", np.float32, W * H * 16))
    return rend, scene, counters, first


@pytest.mark.parametrize("run", ["small_det", "small_default"])
def test_frames_match_llvmpipe(ve, scenes, small_scene, gold, run):
    meta, arr = gold
    rend, scene, counters, first = _gpu_run(ve, scenes, small_scene, meta, arr, run)
    try:
        r = meta["runs"][run]
        W, H = r["W"], r["H"]
        assert np.array_equal(first[0], arr[run + "_gbuf"])
        assert_out_close(first[1], arr[run + "_out0"])
        assert np.array_equal(np.array(counters, np.uint32), arr[run + "_counters"])
        nc, ns = int(counters[-1][0]), int(counters[-1][1])
        cand = rend.read("CANDIDATES", np.uint32, nc * 16).reshape(-1, 4)
        assert np.array_equal(cand, arr[run + "_candidates"])
        sl = rend.read("SHADE_LIST", np.uint32, ns * 8).reshape(-1, 2)
        assert np.array_equal(sl, arr[run + "_shade_list"])
        sc = rend.read("SAMPLE_COUNTS").reshape(16384, 64).T
        assert np.array_equal(sc, arr[run + "_sample_counts"].reshape(64, 16384))
        assert np.array_equal(rend.read("RANK_STATS"), arr[run + "_rank_stats"])
        lv, lval = hash_map(rend.read("LBUFFER"), 12, 32)
        assert np.array_equal(lv, arr[run + "_lbuf_vids"])
        assert np.array_equal(lval[:, :10], arr[run + "_lbuf_vals"][:, :10])   # sums and counts
        assert np.array_equal(lval, arr[run + "_lbuf_vals"])                 # and last-deposit times
        nv, nval = hash_map(rend.read("NBUFFER"), 3, 32)
        assert np.array_equal(nv, arr[run + "_nbuf_vids"]) and np.array_equal(nval, arr[run + "_nbuf_vals"])
        assert_out_close(rend.read("OUTPUT", np.float32, W * H * 16), arr[run + "_out"])
    finally:
        rend.release(scene)
        rend.close()


def test_frames_vs_oracle_random(ve, scenes, oracle, small_scene):
    """Random cameras, sizes, budgets and bounce limits: GPU == oracle on every
    integer structure, resolve within tolerance."""
    rng = np.random.default_rng(11)
    words, depth = small_scene["words"], small_scene["depth"]
    mats = scenes.palette(3)
    for case in range(3):
        W, H = int(rng.integers(24, 150)), int(rng.integers(16, 120))
        mb = int(rng.integers(1, 5))
        eye = rng.uniform(-40, 170, 3)
        tgt = rng.uniform(30, 100, 3)
        e, d = scenes.look_at(eye, tgt)
        ubo = ve.camera_ubo(e, d, float(rng.uniform(40, 100)), W / H)
        knobs = dict(normal_precision=int(rng.integers(1, 4)), shading_budget=int(rng.integers(50, 4000)),
                     update_time=int(rng.integers(3, 10)), bounce_controlchecks=int(rng.integers(30, 140)))
        print("case random frame", case, W, H, mb, knobs, flush=True)
        rend = ve.Renderer(W, H, max_bounces=mb)
        of = oracle.Frame(W, H, 16384, mb)
        try:
            rend.set_materials(mats)
            of.set_materials(mats)
            rend.set_sky_constant((90, 150, 240))
            of.set_sky_rgb8((90, 150, 240))
            rend.set_camera(ubo)
            scene = rend.upload(words, depth, small_scene["info"].size)
            fc_g, fc_o = ve.FrameConfig(**knobs), oracle.FrameConfig(**knobs)
            for f in range(3):
                rend.frame(scene, fc_g, ve.FrameInputs(W, H, f, frame_seed(f + case), 2 * f + 1))
                of.frame(words, depth, ubo, fc_o, f, frame_seed(f + case), 2 * f + 1)
                cg, co = rend.counters(), of.counters()
                assert [cg[0], cg[1], cg[7], cg[4]] == [co[0], co[1], co[7], co[4]]
            nc, ns = int(cg[0]), int(cg[1])
            assert np.array_equal(rend.read("CANDIDATES", np.uint32, nc * 16), of.read("CANDIDATES"))
            assert np.array_equal(rend.read("SHADE_LIST", np.uint32, ns * 8), of.read("SHADE_LIST"))
            gl, ol = hash_map(rend.read("LBUFFER"), 12, 32), hash_map(of.read("LBUFFER"), 12, 32)
            assert np.array_equal(gl[0], ol[0]) and np.array_equal(gl[1], ol[1])
            assert_out_close(rend.read("OUTPUT", np.float32, W * H * 16), of.read("OUTPUT", np.float32))
            rend.release(scene)
        finally:
            rend.close()
            of.close()


@pytest.mark.parametrize("radius", [0, 1, 10, 28, 29, 30])
def test_timeout_reconstruction(ve, scenes, oracle, small_scene, radius):
    """reconstructTimeout (resolve.comp:193-219) on a view where a low primary
    check budget times many rays out, radii from 0 to the reference's UI
    maximum of 30 (src/UI/control.cpp:47).  Deterministic knobs (one
    sample per voxel per frame, no stale reset, caps above the sample count),
    so the lBuffer the resolve reads is bit-exact."""
    W, H = 160, 96
    words, depth = small_scene["words"], small_scene["depth"]
    e, d = scenes.look_at((-30.0, 70.0, -20.0), (90.0, 40.0, 100.0))
    ubo = ve.camera_ubo(e, d, 75.0, W / H)
    knobs = dict(DET_KNOBS, primary_controlchecks=14, reconstruction_radius=radius)
    print("case timeout radius", radius, flush=True)
    rend = ve.Renderer(W, H, max_bounces=2)
    of = oracle.Frame(W, H, 16384, 2)
    try:
        rend.set_materials(scenes.palette(1))
        of.set_materials(scenes.palette(1))
        rend.set_sky_constant((140, 180, 230))
        of.set_sky_rgb8((140, 180, 230))
        rend.set_camera(ubo)
        scene = rend.upload(words, depth, small_scene["info"].size)
        fc_g, fc_o = ve.FrameConfig(**knobs), oracle.FrameConfig(**knobs)
        for f in range(2):
            rend.frame(scene, fc_g, ve.FrameInputs(W, H, f, frame_seed(f), f + 1))
            of.frame(words, depth, ubo, fc_o, f, frame_seed(f), f + 1)
        g, _ = rend.gbuffer(W, H)
        g = np.asarray(g).reshape(-1, 4)
        n_timeout = int(np.count_nonzero((g[:, 3] == 0) & (g[:, 0] >= 14)))
        assert n_timeout > 100
        assert int(rend.counters()[ve.CNT["TIMEOUTS"]]) == n_timeout
        assert int(rend.counters()[ve.CNT["OVERFLOW"]]) == 0
        assert_out_close(rend.read("OUTPUT", np.float32, W * H * 16), of.read("OUTPUT", np.float32))
        rend.release(scene)
    finally:
        rend.close()
        of.close()


def aces_rgba8(rgba32f):
    """final.frag:11-19 in float32, in the shader's operation order, then the
    unorm8 conversion (round to nearest, ties to even)."""
    x = np.asarray(rgba32f, np.float32).reshape(-1, 4)[:, :3]
    f = np.float32
    num = x * (f(2.51) * x + f(0.03))
    den = x * (f(2.43) * x + f(0.59)) + f(0.14)
    v = np.clip(num / den, f(0.0), f(1.0)).astype(np.float32)
    out = np.empty((len(x), 4), np.uint8)
    out[:, :3] = np.rint(v * f(255.0)).astype(np.uint8)
    out[:, 3] = 255
    return out


def test_display_pass_is_aces(ve, scenes, small_scene):
    """The display pass (Renderer::run's final blit) maps the resolve output
    through final.frag's ACES curve into RGBA8, pixel for pixel; also over a
    synthetic HDR image written into the output buffer."""
    W, H = 96, 64
    rend = ve.Renderer(W, H, max_bounces=2)
    try:
        rend.set_materials(scenes.palette(1))
        rend.set_sky_constant((140, 180, 230))
        e, d, fov = scenes.camera_small()
        rend.set_camera(ve.camera_ubo(e, d, fov, W / H))
        sc = rend.upload(small_scene["words"], small_scene["depth"], small_scene["info"].size)
        fc = ve.FrameConfig(**DET_KNOBS)
        inp = ve.FrameInputs(W, H, 0, frame_seed(0), 1)
        rend.frame(sc, fc, inp)
        rend.run_pass(sc, "DISPLAY", fc, inp)
        out = rend.read("OUTPUT", np.float32, W * H * 16)
        disp = rend.read("DISPLAY", np.uint8, W * H * 4).reshape(-1, 4)
        assert np.array_equal(disp, aces_rgba8(out))
        rng = np.random.default_rng(3)
        hdr = (rng.random((W * H, 4), np.float32) * 8.0).astype(np.float32)
        hdr[:16, :3] = [0.0, 1e-6, 1e4]
        full = np.zeros(rend.size("OUTPUT") // 4, np.float32)
        full[: W * H * 4] = hdr.ravel()
        rend.write("OUTPUT", full)
        rend.run_pass(sc, "DISPLAY", fc, inp)
        disp = rend.read("DISPLAY", np.uint8, W * H * 4).reshape(-1, 4)
        assert np.array_equal(disp, aces_rgba8(hdr))
        rend.release(sc)
    finally:
        rend.close()
