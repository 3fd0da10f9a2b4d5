"""The batched form (ve_batch_frame) against each view rendered alone
(ve_frame): uneven views on different scenes, sizes, bounce counts and knobs,
their primary passes in one persistent launch pulling tiles from a global
queue -- the other passes on each view's stream, or (whole frames) as one
launch per kernel over all the views; every structure of every view
bit-exact frame by frame, for several launch shapes and groupings, with and
without graph replay (per view, or of the whole-frame batch's launches)."""
import numpy as np
import pytest

from frame_common import frame_seed, hash_map

pytestmark = pytest.mark.gpu

BUFS = ("GBUFFER", "GBUFFER_POS", "COUNTERS", "LBUFFER", "NBUFFER", "EMISSIVES", "OUTPUT")
DET = dict(restir_spatial=0, weight_pixels=0.0, weight_variance=0.0, top_fraction=0.0, mid_fraction=1.0)


@pytest.fixture(scope="module")
def views(ve, scenes, small_scene):
    ld, lops = scenes.scene_lights(1)
    lw, linfo = ve.octree_build(lops, ld)
    fd, fops = scenes.scene_fbm(9, 7)
    fw, finfo = ve.octree_build(fops, fd)
    small = (small_scene["depth"], small_scene["words"], small_scene["info"])
    return [
        # scene, palette, W, H, max_bounces, knobs, orbit seed
        (small, scenes.palette(1), 160, 120, 3, dict(DET), 3),
        ((ld, lw, linfo), scenes.palette_lights(1), 97, 61, 4, dict(DET), 5),
        ((fd, fw, finfo), scenes.palette(1), 256, 200, 2, dict(DET, min_samples=2), 7),
        (small, scenes.palette_spec(1), 64, 64, 3, dict(DET, sample_cap_direct=8, sample_cap_indirect=16), 9),
        ((fd, fw, finfo), scenes.palette(1), 200, 150, 3, {}, 11),
        (small, scenes.palette(1), 48, 40, 1, dict(DET), 13),   # one bounce: no fused bounce pass
    ]


def _make(ve, vs, graph, bounce_blocks=0):
    out = []
    for (sc, mats, W, H, mb, _, _) in vs:
        r = ve.Renderer(W, H, hash_width=16384, max_bounces=mb)
        r.set_overlap(False)
        if graph:
            r.set_graph(True)
        if bounce_blocks:
            r.set_tuning(bounce_blocks=bounce_blocks)
        r.set_materials(mats)
        r.set_sky_constant((120, 160, 210))
        out.append((r, r.upload(sc[1], sc[0], sc[2].size)))
    return out


@pytest.mark.parametrize("tile_w,grab,waves,groups,graph,whole,bounce_blocks", [(8, 4, 16, 1, False, False, 0),
                                                                                (32, 1, 4, 4, True, False, 0),
                                                                                (64, 16, 8, 64, False, False, 0),
                                                                                (16, 2, 32, 2, True, False, 0),
                                                                                (8, 4, 32, 1, False, True, 0),
                                                                                (32, 2, 8, 4, False, True, 0),
                                                                                (8, 4, 32, 1, False, True, 3),
                                                                                (16, 4, 16, 1, True, True, 0)])
def test_batch_matches_views_alone(ve, scenes, views, tile_w, grab, waves, groups, graph, whole, bounce_blocks):
    """bounce_blocks: the batched views carry that bounce-grid tuning (the
    views alone keep the full grid); results may not depend on it."""
    batch = ve.Batch(tile_w=tile_w, grab=grab, waves_per_cu=waves, groups=groups, whole_frame=whole)
    a, b = _make(ve, views, graph, bounce_blocks), _make(ve, views, False)
    try:
        for f in range(3):
            fcs, ins = [], []
            for i, (sc, _, W, H, mb, knobs, seed) in enumerate(views):
                eye, d, fov = scenes.orbit_camera(1 << sc[0], f, 12, seed)
                ubo = ve.camera_ubo(eye, d, fov, W / H)
                fc, fi = ve.FrameConfig(**knobs), ve.FrameInputs(W, H, f, frame_seed(f + i), f + 1)
                a[i][0].set_camera(ubo)
                b[i][0].set_camera(ubo)
                b[i][0].frame(b[i][1], fc, fi)
                fcs.append(fc)
                ins.append(fi)
            batch.frame([r for r, _ in a], [s for _, s in a], fcs, ins)
            print("case batch", tile_w, grab, waves, groups, graph, whole, bounce_blocks, "frame", f, flush=True)
            for i in range(len(views)):
                for name in BUFS:
                    assert np.array_equal(a[i][0].read(name), b[i][0].read(name)), "view %d frame %d %s" % (i, f, name)
                ra, rb = hash_map(a[i][0].read("RBUFFER"), 8, 16), hash_map(b[i][0].read("RBUFFER"), 8, 16)
                assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1], rb[1]), "view %d frame %d RBUFFER" % (i, f)
        assert batch.primary_ms() > 0.0
        if graph and whole:   # one batch graph, captured on the first frame and replayed for each
            assert batch.graph_stats() == dict(captures=1, replays=3)
        elif graph:
            assert all(r.graph_stats()["replays"] == 3 for r, _ in a)
    finally:
        for r, _ in a + b:
            r.close()
        batch.close()


def test_batch_rejects_overlapped_context(ve, small_scene):
    batch = ve.Batch()
    r = ve.Renderer(32, 32)
    try:
        s = r.upload(small_scene["words"], small_scene["depth"], small_scene["info"].size)
        with pytest.raises(ve.VeError) as e:
            batch.frame([r], [s], [ve.FrameConfig()], [ve.FrameInputs(32, 32, 0, 1, 1)])
        assert e.value.code == ve.VE_ESTATE
    finally:
        r.close()
        batch.close()
