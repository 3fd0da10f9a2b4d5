"""Autotuner of the batched form's primary launch (ve_batch_config).

Runs the batched configuration's viewports (serving.make_viewports: 64 views
over the resident scenes, resolutions 128x128..1920x1080) through ve_batch
and times the persistent primary launch (ve_batch_primary_ms) for every
tile shape (tile_w 8/16/32/64, tiles of tile_w x 64/tile_w pixels), queue
pull size (grab: tiles a wave takes per atomic) and grid size (waves per
compute unit); then, for the best shape, the traversal jump-table depth of
the scenes (the octree levels a ray skips through one table lookup, 6..8)
and the number of launch groups, timed on whole ticks (every view's frame
done), and the same ticks as whole-frame batches (every later pass one
launch over all the views).  Writes every measurement and the chosen shape
to tuning.json beside this file, which serving.batch_tuning() reads.

    python -m voxelengine_amd.autotune            (through tests/ve_loader)
    python voxelengine3.0_amd/autotune.py [--quick] [--out PATH]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "tuning.json")

TILE_WS = (8, 16, 32, 64)
GRABS = (1, 2, 4, 8, 16)
WAVES = (4, 8, 16, 32)
GROUPS = (1, 2, 4, 8, 16)
JUMPS = (6, 7, 8)


def _views(ve, S, serving, res_scenes, count, seed, jump_level):
    """One context per viewport (overlap off), the scenes uploaded once at the
    given jump-table depth and shared by the views."""
    up = ve.Renderer(64, 64)
    up.set_tuning(0, jump_level)
    scenes = [up.upload(w, d, info.size) for d, w, info in res_scenes]
    views = serving.make_viewports(count, len(res_scenes), seed)
    mats = S.palette(1)
    rends, args = [], []
    for v in views:
        r = ve.Renderer(v["W"], v["H"], hash_width=16384, max_bounces=v["bounces"] + 1)
        r.set_overlap(False)
        r.set_materials(mats)
        r.set_sky_constant((140, 180, 230))
        rends.append(r)
        args.append((v, ve.FrameConfig(**serving.session_knobs(v["spp"]))))
    return up, scenes, rends, args


def _tick(ve, S, batch, rends, scenes, args, res_scenes, k):
    fcs, ins, scs = [], [], []
    for r, (v, fc) in zip(rends, args):
        d = res_scenes[v["scene"]][0]
        eye, dd, fov = S.orbit_camera(1 << d, k, 240, v["seed"])
        r.set_camera(ve.camera_ubo(eye, dd, fov, v["W"] / v["H"]))
        fcs.append(fc)
        ins.append(ve.FrameInputs(v["W"], v["H"], k, (v["seed"] + 7919 * k) & 0xFFFFFFFF, k + 1))
        scs.append(scenes[v["scene"]])
    t0 = time.perf_counter()
    batch.frame(rends, scs, fcs, ins)
    for r in rends:
        r.sync()
    return batch.primary_ms(), (time.perf_counter() - t0) * 1e3


def tune(ve, res_scenes, count=64, seed=0, ticks=3, quick=False, log=print):
    from voxelengine_amd import scenes as S, serving
    out = dict(workload="serving.make_viewports(%d, %d scenes, seed %d): resolution U[128..1920]x[128..1080], "
                        "spp {1,2,4}, bounces {1,2,4}" % (count, len(res_scenes), seed),
               stage1=[], stage2=[])
    up, scenes, rends, args = _views(ve, S, serving, res_scenes, count, seed, 8)
    try:
        # the views' own primary passes, one launch each, for comparison
        total = 0.0
        for r, s, (v, fc) in zip(rends, [scenes[a[0]["scene"]] for a in args], args):
            r.enable_timing(2)
            for k in range(2):
                d = res_scenes[v["scene"]][0]
                eye, dd, fov = S.orbit_camera(1 << d, k, 240, v["seed"])
                r.set_camera(ve.camera_ubo(eye, dd, fov, v["W"] / v["H"]))
                r.frame(s, fc, ve.FrameInputs(v["W"], v["H"], k, 1, k + 1))
            total += r.timings()["PRIMARY"]
            r.enable_timing(False)
        out["per_view_primary_ms_sum"] = round(total, 4)
        log("per-view primary passes: %.3f ms per tick" % total)
        shapes = [(tw, g, w) for tw in TILE_WS for g in GRABS for w in WAVES]
        if quick:
            shapes = [(tw, g, w) for tw in (8, 32) for g in (1, 4) for w in (8, 16)]
        for tw, g, w in shapes:
            batch = ve.Batch(max_views=count, tile_w=tw, grab=g, waves_per_cu=w, groups=1)
            try:
                _tick(ve, S, batch, rends, scenes, args, res_scenes, 0)
                ms = [_tick(ve, S, batch, rends, scenes, args, res_scenes, k + 1)[0] for k in range(ticks)]
            finally:
                batch.close()
            row = dict(tile_w=tw, grab=g, waves_per_cu=w, primary_ms=round(float(np.median(ms)), 4))
            out["stage1"].append(row)
            log("stage 1 %s" % row)
        best = min(out["stage1"], key=lambda r: r["primary_ms"])
    finally:
        for r in rends:
            r.close()
        up.close()
    for jl in ((8,) if quick else JUMPS):
        up, scenes, rends, args = _views(ve, S, serving, res_scenes, count, seed, jl)
        try:
            for gr, whole in [(g, False) for g in ((1, 4) if quick else GROUPS)] + [(1, True)]:
                batch = ve.Batch(max_views=count, tile_w=best["tile_w"], grab=best["grab"],
                                 waves_per_cu=best["waves_per_cu"], groups=gr, whole_frame=whole)
                try:
                    _tick(ve, S, batch, rends, scenes, args, res_scenes, 0)
                    res = [_tick(ve, S, batch, rends, scenes, args, res_scenes, k + 1) for k in range(ticks)]
                finally:
                    batch.close()
                # whole-frame batches: primary_ms is the whole batch's launches
                row = dict(jump_level=jl, groups=gr, whole_frame=whole,
                           primary_ms=round(float(np.median([a for a, _ in res])), 4),
                           tick_ms=round(float(np.median([b for _, b in res])), 4))
                out["stage2"].append(row)
                log("stage 2 %s" % row)
        finally:
            for r in rends:
                r.close()
            up.close()
    b2 = min((r for r in out["stage2"] if not r["whole_frame"]), key=lambda r: r["tick_ms"])
    bw = min((r for r in out["stage2"] if r["whole_frame"]), key=lambda r: r["tick_ms"])
    out["best"] = dict(tile_w=best["tile_w"], grab=best["grab"], waves_per_cu=best["waves_per_cu"],
                       groups=b2["groups"], jump_level=b2["jump_level"], primary_ms=best["primary_ms"],
                       tick_ms=b2["tick_ms"])
    out["best_whole_frame"] = dict(jump_level=bw["jump_level"], tick_ms=bw["tick_ms"])
    return out


def main():
    os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")
    root = os.path.dirname(HERE)
    sys.path.insert(0, os.path.join(root, "tests"))
    sys.path.insert(0, root)
    import torch
    assert torch.cuda.is_available(), "autotune needs a GPU"
    from ve_loader import load_package
    ve = load_package()
    from voxelengine_amd import scenes as S, serving
    quick = "--quick" in sys.argv
    d, ops = S.scene_extreme(1)
    w, info = ve.octree_build(ops, d)
    cd, cops = S.scene_caves(1)
    cw, cinfo = ve.octree_build(cops, cd, max_words=1 << 28)
    d9, d11 = serving.served_scenes(((9, 7), (11, 8)))
    res = [(d, w, info), (cd, cw, cinfo), d9, d11]
    t0 = time.time()
    r = tune(ve, res, quick=quick, log=lambda m: print(m, flush=True))
    r["seconds"] = round(time.time() - t0, 1)
    path = OUT if not quick else os.path.join(os.getcwd(), "tuning_quick.json")
    if "--out" in sys.argv:
        path = sys.argv[sys.argv.index("--out") + 1]
    with open(path, "w") as fh:
        json.dump({"batched_primary": r}, fh, indent=1)
    print(json.dumps(r["best"]), flush=True)


if __name__ == "__main__":
    main()
