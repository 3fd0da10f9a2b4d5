"""Cross-frame overlap (the next frame's primary and schedule passes on a
second stream, into a second G-buffer set, and its normal pass on a third,
into a second nBuffer set, beside the previous frame's bounce passes)
renders exactly what the reference's serial pass order renders.

A queue of frames with a moving camera, a material change and a scene edit
between two frames, a lone pass call and a display pass is enqueued with no
other host synchronisation in between, so the primary passes really run ahead; every image, counter,
list and ray total is then compared bit for bit, and the nBuffer / lBuffer
contents voxel by voxel, against a context making the same calls with the
overlap switched off.  The knobs are the order-independent ones, so two
serial runs agree the same way (a hash bucket's slot order follows which
concurrent claimant came first, as in the reference)."""
import numpy as np
import pytest

from frame_common import DET_KNOBS, frame_seed, hash_map

pytestmark = pytest.mark.gpu

W, H = 160, 120
FRAMES = 12
READS = ("OUTPUT", "GBUFFER", "GBUFFER_POS", "CANDIDATES", "SHADE_LIST", "COUNTERS", "RANK_STATS", "SAMPLE_COUNTS",
         "CLAIM_MAP", "DISPLAY")
HASHED = {"NBUFFER": (3, 32), "LBUFFER": (12, 32)}   # words per slot, slots per bucket


def _render(ve, scenes, overlap):
    depth, ops = scenes.scene_small(1)
    cpu = ve.OctreeCPU(depth)
    cpu.insert_ops(ops)
    octree = ve.Octree(depth).set(cpu)
    rend = ve.Renderer(W, H, hash_width=4096, max_bounces=3)
    try:
        rend.set_overlap(overlap)
        rend.set_materials(scenes.palette(1))
        rend.set_sky_constant((140, 180, 230))
        scene = rend.upload(octree.data, depth, octree.size)
        fc = ve.FrameConfig(**DET_KNOBS)
        e, d, fov = scenes.camera_small()
        for f in range(FRAMES):
            eye = np.asarray(e, np.float32) + np.float32(1.5 * f) * np.array([1.0, 0.0, 0.5], np.float32)
            rend.set_camera(ve.camera_ubo(eye, d, fov, W / H))
            inputs = ve.FrameInputs(W, H, f, frame_seed(f), f + 1)
            if f == 6:   # the next primary pass must see the edited words
                cpu.insertSphere(np.array([64.0, 70.0, 64.0], np.float32), 12.0, 5)
                octree.applyEdits(cpu)
                rend.update(scene, octree)
            if f == 4:   # the previous frame's resolve must still see the old materials
                rend.set_materials(scenes.palette_spec(1))
            if f == 8:   # a lone pass between frames
                rend.run_pass(scene, "PRIMARY", fc, inputs)
            rend.frame(scene, fc, inputs)
            if f % 3 == 2:
                rend.run_pass(scene, "DISPLAY", fc, inputs)
        print("case overlap", overlap, "frames enqueued", flush=True)
        out = {name: rend.read(name).copy() for name in READS}
        for name, (stride, slots) in HASHED.items():
            out[name] = hash_map(rend.read(name), stride, slots)
        out["totals"] = rend.ray_totals()
        rend.release(scene)
        return out
    finally:
        rend.close()


def test_overlap_matches_serial_order(ve, scenes):
    serial = _render(ve, scenes, False)
    overlapped = _render(ve, scenes, True)
    assert serial["COUNTERS"][ve.CNT["PRIMARY_RAYS"]] == W * H
    assert serial["totals"]["primary"] == (FRAMES + 1) * W * H   # every frame plus the lone pass
    assert overlapped["totals"] == serial["totals"]
    for name in READS:
        assert np.array_equal(overlapped[name], serial[name]), name
    for name in HASHED:
        assert np.array_equal(overlapped[name][0], serial[name][0]), name
        assert np.array_equal(overlapped[name][1], serial[name][1]), name



@pytest.mark.parametrize("bounce_cc", [120, 20])
def test_reused_bounce0_hits_equal_traversal(ve, scenes, small_scene, bounce_cc):
    """Bounce-0 rays take their hit record from the primary pass when the
    pixel's primary ray hit the voxel within the bounce budget; a context with
    node counting on traverses every ray as the reference does.  With a bounce
    budget below the primary one, the rays whose primary hit needed more steps
    fall back to a traversal (and time out the same way)."""
    knobs = dict(DET_KNOBS, bounce_controlchecks=bounce_cc)
    res = []
    for traverse_all in (False, True):
        rend = ve.Renderer(W, H, hash_width=4096, max_bounces=3)
        try:
            rend.set_materials(scenes.palette(1))
            rend.set_sky_constant((140, 180, 230))
            rend.enable_node_stats(traverse_all)
            sc = rend.upload(small_scene["words"], small_scene["depth"], small_scene["info"].size)
            e, d, fov = scenes.camera_small()
            fc = ve.FrameConfig(**knobs)
            reused = []
            for f in range(4):
                rend.set_camera(ve.camera_ubo(e, d, fov, W / H))
                rend.frame(sc, fc, ve.FrameInputs(W, H, f, frame_seed(f), f + 1))
                c = rend.counters()
                reused.append((int(c[ve.CNT["REUSED"]]), int(c[ve.CNT["SHADE"]])))
            print("case reuse bounce_cc", bounce_cc, "traverse_all", traverse_all, reused, flush=True)
            out = {name: rend.read(name).copy() for name in ("OUTPUT", "COUNTERS", "SHADE_LIST")}
            out["LBUFFER"] = hash_map(rend.read("LBUFFER"), 12, 32)
            totals = rend.ray_totals()
            rend.release(sc)
            res.append((out, reused, totals))
        finally:
            rend.close()
    (a, ra, ta), (b, rb, tb) = res
    assert all(r == 0 for r, _ in rb)                      # node counting: every ray traversed
    assert all(0 < r <= s for r, s in ra)
    if bounce_cc >= 70:
        assert all(r == s for r, s in ra)                  # every shade-list pixel hit within 70 steps
    else:
        assert any(r < s for r, s in ra)                   # some primary hits took more than 21 steps
    assert ta["traced"] == tb["traced"] and ta["primary"] == tb["primary"] and ta["reused"] == sum(r for r, _ in ra)
    for name in ("OUTPUT", "SHADE_LIST"):
        assert np.array_equal(a[name], b[name]), name
    ca, cb = a["COUNTERS"].copy(), b["COUNTERS"].copy()
    ca[ve.CNT["REUSED"]] = cb[ve.CNT["REUSED"]] = 0
    assert np.array_equal(ca, cb)
    assert np.array_equal(a["LBUFFER"][0], b["LBUFFER"][0]) and np.array_equal(a["LBUFFER"][1], b["LBUFFER"][1])


@pytest.mark.parametrize("min_samples", [1, 48])
def test_frame_equals_pass_by_pass(ve, scenes, small_scene, min_samples):
    """ve_frame -- primary pass overlapped with the previous frame, bounce-0
    hits taken from the primary pass, bounce iterations 1.. fused per ray with
    their deposits grouped per lBuffer bucket and applied in iteration order
    -- renders what the reference's one dispatch per pass and per bounce
    iteration renders (Renderer::runWavefrontFrame, renderer.cpp:282-649),
    frame after frame.  Four bounces, so two iterations' deposits are held
    back and applied in order.  With 48 samples per visible voxel a voxel's
    bucket holds back more deposits than it keeps in place, so the overflow
    list is grouped and applied too."""
    bounces = 4
    fc_kw = dict(DET_KNOBS, min_samples=min_samples)
    out = []
    for by_pass in (False, True):
        rend = ve.Renderer(W, H, hash_width=4096, max_bounces=bounces)
        try:
            rend.set_materials(scenes.palette(1))
            rend.set_sky_constant((140, 180, 230))
            sc = rend.upload(small_scene["words"], small_scene["depth"], small_scene["info"].size)
            e, d, fov = scenes.camera_small()
            fc = ve.FrameConfig(**fc_kw)
            for f in range(5):
                eye = np.asarray(e, np.float32) + np.float32(2.0 * f) * np.array([0.5, 0.0, 1.0], np.float32)
                rend.set_camera(ve.camera_ubo(eye, d, fov, W / H))
                inputs = ve.FrameInputs(W, H, f, frame_seed(f), f + 1)
                print("case frame paths min_samples", min_samples, "by_pass", by_pass, "frame", f, flush=True)
                if by_pass:
                    for name in ("RESET", "PRIMARY", "NORMAL", "SCHEDULE", "RANK", "THRESHOLD", "EMIT"):
                        rend.run_pass(sc, name, fc, inputs)
                    for _ in range(bounces):
                        rend.run_pass(sc, "TRACE", fc, inputs)
                        rend.run_pass(sc, "SHADE", fc, inputs)
                    rend.run_pass(sc, "RESOLVE", fc, inputs)
                else:
                    rend.frame(sc, fc, inputs)
            res = {name: rend.read(name).copy() for name in ("OUTPUT", "COUNTERS", "CANDIDATES", "SHADE_LIST")}
            for name, (stride, slots) in HASHED.items():
                res[name] = hash_map(rend.read(name), stride, slots)
            res["totals"] = rend.ray_totals()
            rend.release(sc)
            out.append(res)
        finally:
            rend.close()
    fused, by_pass = out
    assert fused["totals"]["reused"] > 0 and by_pass["totals"]["reused"] == 0
    for k in ("primary", "traced", "shadow"):
        assert fused["totals"][k] == by_pass["totals"][k], k
    ca, cb = fused["COUNTERS"].copy(), by_pass["COUNTERS"].copy()
    assert int(cb[ve.CNT["BOUNCES_RUN"]]) >= 3
    for k in ("LIVE", "LIVE_B", "REUSED"):   # bookkeeping of the per-iteration dispatch / of the reuse
        ca[ve.CNT[k]] = cb[ve.CNT[k]] = 0
    ca[14] = cb[14] = 0                    # deposits held back by the fused pass
    assert np.array_equal(ca, cb)
    for name in ("OUTPUT", "CANDIDATES", "SHADE_LIST"):
        assert np.array_equal(fused[name], by_pass[name]), name
    for name in HASHED:
        assert np.array_equal(fused[name][0], by_pass[name][0]), name
        assert np.array_equal(fused[name][1], by_pass[name][1]), name


def _row_voxels(row, chunks, seed=5):
    """Random voxels of a 512^3 grid (y in 64..447) whose ids fall in claim
    row `row`: voxelKey + 1 then claimHash % 16384 (internal.glsl:263-291,
    schedule.comp:36-43) in 32-bit arithmetic."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(chunks):
        x, y, z = rng.integers(0, 512, size=(3, 1 << 24), dtype=np.uint32)
        y = np.uint32(64) + y % np.uint32(384)
        v = x | (y << np.uint32(10)) | (z << np.uint32(20))
        for s, m in ((16, 0x7FEB352D), (15, 0x846CA68B)):
            v ^= v >> np.uint32(s)
            v *= np.uint32(m)
        v ^= v >> np.uint32(16)
        v = (v & np.uint32(0x7FFFFFFF)) + np.uint32(1)
        v ^= v >> np.uint32(16)
        v *= np.uint32(0x7FEB352D)
        v ^= v >> np.uint32(15)
        sel = v % np.uint32(16384) == row
        out.append(np.stack([x[sel], y[sel], z[sel]]))
    return np.unique(np.concatenate(out, axis=1), axis=1).astype(np.int64)


def test_overflow_flag_per_frame(ve, scenes):
    """A claim row crowded past the schedule's exact-order capacity (more
    than 2048 distinct visible voxels in one row) raises the overflow
    counter.  With the overlap on, the frame's schedule pass runs ahead of
    the frame's reset and reports it into its buffer set's side word, which
    the frame's resolve folds into the counter and clears: the flag reads the
    same as the serial context's frame by frame, and frames that see no
    crowded row read it clear in both buffer sets."""
    from test_gpu_claim_overflow import claim_rows, voxel_ids
    P = _row_voxels(777, 6)
    assert np.all(claim_rows(voxel_ids(*(P.astype(np.uint64)))) == 777)
    ops = scenes._ops(P[0], P[1], P[2], scenes.voxel_materials(P[0], P[1], P[2], 1))
    words, info = ve.octree_build(ops, 9)
    W = H = 1024
    crowded = [True, True, False, True, False]
    flags = {}
    for overlap in (False, True):
        rend = ve.Renderer(W, H, hash_width=4096, max_bounces=1)
        try:
            rend.set_overlap(overlap)
            rend.set_materials(scenes.palette(1))
            rend.set_sky_constant((140, 180, 230))
            sc = rend.upload(words, 9, info.size)
            fc = ve.FrameConfig(**DET_KNOBS)
            seq = []
            for f, crowd in enumerate(crowded):
                if crowd:
                    e, d = scenes.look_at((256.0 + 4.0 * f, 900.0, -500.0), (256.0, 250.0, 256.0))
                else:   # up and away from above the grid: sky only
                    e, d = scenes.look_at((256.0, 600.0, 256.0), (300.0, 1000.0, 330.0))
                rend.set_camera(ve.camera_ubo(e, d, 50.0, 1.0))
                rend.frame(sc, fc, ve.FrameInputs(W, H, f, frame_seed(f), f + 1))
                c = rend.counters()
                seq.append(int(c[ve.CNT["OVERFLOW"]]))
                # the one row's first 64 claimants (CLAIM_PROBES), or nothing
                assert int(c[ve.CNT["CANDIDATES"]]) == (64 if crowd else 0)
            print("case overflow flag overlap", overlap, seq, flush=True)
            rend.release(sc)
            flags[overlap] = seq
        finally:
            rend.close()
    assert flags[False] == flags[True] == [1 if c else 0 for c in crowded]
