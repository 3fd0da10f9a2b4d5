"""Benchmark: Mrays/s of the SVO wavefront frame on one MI355X.

Workload (the depth-10 configuration): 1024^3 octree of the seeded fbm terrain
shell + 64 seeded spheres (10.4 M voxels), 1920x1080, 1 spp, 2 bounces
(MAX_BOUNCES = 3), 64-frame seeded orbit at radius 0.8N (eye height 0.75N,
40 degree vertical FOV, so the grid fills the frame) with temporal
accumulation (persistent lBuffer).  One step = one whole
Renderer::runWavefrontFrame: primary rays, normals, schedule, rank, threshold,
emit, the trace/shade bounce loop and resolve.

value = (primary rays + traced bounce rays + NEE shadow rays) of the timed
frames / wall time of the timed region, counting every live ray of each
trace pass as the reference does (the bounce-0 camera rays, whose hit records
come from the primary pass's traversal, are listed as reused_rays_per_frame,
and mrays_per_s_traversed_only leaves them out) (inputs resident in HBM; the host only
enqueues; frames as an application renders them: the context's default
cross-frame overlap, where a frame's primary, schedule and normal passes run
on other streams beside the previous frame's bounce passes, and its resolve
beside the next frame's first passes, and no instrumentation).  The same
number of frames then runs again with per-pass HIP events bound to the
passes' kernel dispatches (pass_ms), and once more with the overlap off (the
reference's strict pass order) for pass_ms_serial.  The roofline entry is the
primary kernel (the reference's RaycastCounted traversal): algorithmic node
bytes per launch (4 bytes per node word the reference's traversal reads,
counted by the instrumented kernel on the same frames) over its mean launch
time from those HIP events, in the overlapped pipeline (launch_ms, sharing
the device with the previous frame's bounce passes) and alone
(isolated_launch_ms / isolated_frac).  cpu_baseline: the C oracle (a scalar
port of the same passes) on the host, whole frames of the same workload.

`served` and `batched` (skipped with --no-served): the served form (Poisson
sessions over resident octrees on the session server, p50/p99 frame latency,
frames/s and Mrays/s per offered load, concurrent server vs the plain
one-frame-at-a-time loop) and the batched form (64 uneven viewports, frame k
of each submitted together); see voxelengine3.0_amd/serving.py.  The
process runs with GPU_MAX_HW_QUEUES=16 unless the environment sets it: the
concurrent server gives every session its own stream, and HIP's default of
4 hardware queues makes sessions share them; the headline frames (two
streams) and the plain loop (one stream) measure the same either way.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

# before anything initialises HIP: hardware queues for the session server's streams
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
W, H = 1920, 1080
BOUNCES = 2
ORBIT_FRAMES = 64
SKY = (140, 180, 230)


def knobs():
    # reference FrameConfig with the accumulation settings of a still camera
    # (voxelengine.cpp:201-207), central-difference normals, full budget, one
    # sample per visible voxel per frame
    return dict(primary_controlchecks=70, bounce_controlchecks=120, normal_precision=1, shading_budget=196608,
                weight_pixels=0.0, weight_variance=0.0, top_fraction=0.0, mid_fraction=1.0, min_samples=1,
                sample_cap_direct=1024, sample_cap_indirect=4096, disable_stale_reset=1, update_time=0xFFFFFFFF)


def profiled_traffic(kernel="k_primary"):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "k_primary_extreme_1080p.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("TRAFFIC %s " % kernel):
                    return int(line.split()[2])
    except OSError:
        pass
    return None


def frame_seed(f):
    return (1 + 7919 * f) & 0xFFFFFFFF


def host_cpu():
    """(threads this process may use, capped at the 16 a one-GPU share
    allows; nproc; CPU model)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return max(1, min(n, 16)), os.cpu_count(), model


def cpu_baseline(words, depth, ubos, fc_kw, budget_s=20.0):
    """C oracle (a scalar port of the same passes) on the host cores: one
    renderer per thread, each accumulating its own orbit sequence of whole
    frames of the same workload, until ~budget_s seconds have passed (at least
    one frame each); value = all threads' rays / wall time."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    from voxelengine_amd import scenes
    threads, nproc, model = host_cpu()
    words = np.ascontiguousarray(words, np.uint32)
    fc = O.FrameConfig(**fc_kw)
    t0 = time.perf_counter()

    def worker(t):
        of = O.Frame(W, H, 16384, BOUNCES + 1)
        try:
            of.set_materials(scenes.palette(1))
            of.set_sky_rgb8(SKY)
            rays = frames = 0
            while True:
                i = t * 7 + frames
                f = i % len(ubos)
                of.frame(words, depth, ubos[f], fc, i, frame_seed(i), i + 1)   # the C call releases the GIL
                c = of.counters()
                rays += int(c[6]) + int(c[4]) + int(c[5])
                frames += 1
                if time.perf_counter() - t0 > budget_s:
                    return rays, frames
        finally:
            of.close()

    with ThreadPoolExecutor(threads) as ex:
        res = list(ex.map(worker, range(threads)))
    dt = time.perf_counter() - t0
    rays, frames = sum(r for r, _ in res), sum(f for _, f in res)
    return dict(value=rays / dt / 1e6, unit="Mrays/s", cores=threads, kind="port", nproc=nproc, cpu=model,
                sample="%d whole frames of the same 1080p workload, %d threads (one renderer each), %.1f s"
                       % (frames, threads, dt))


# offered loads (frames/s) of the served sweep, per mode (each row also
# reports the load its Poisson realisation actually offered); the sweep then
# bisects between the highest load within the p99 bound and the next one
SERVED_LAMS = {"placed": [50, 200, 600, 1200, 1600, 2000, 2200, 2300, 2400], 0: [50, 600, 1200, 1600, 2000, 2200],
               1: [50, 200, 600, 900, 1100, 1200, 1300], "batched": [50, 600, 1200, 1600, 2000, 2200],
               "whole": [50, 600, 1200, 1600, 2000, 2200]}


def divergence(ve, caves, frames=32, warmup=4):
    """The divergence configuration: the depth-12 cave scene (fbm ground with
    value-noise caves and 2048 pillars), 2560x1440, 4 samples per visible
    voxel (bottom rank tier), MAX_BOUNCES 5 (4 bounces), a 32-frame seeded
    orbit; whole frames as an application renders them."""
    from voxelengine_amd import scenes
    depth, words, info = caves
    W2, H2 = 2560, 1440
    r = ve.Renderer(W2, H2, hash_width=16384, max_bounces=5)
    try:
        scene = r.upload(words, depth, info.size)
        r.set_materials(scenes.palette(1))
        r.set_sky_constant((140, 180, 230))
        fc = ve.FrameConfig(**dict(knobs(), top_fraction=0.0, mid_fraction=0.0, min_samples=4))
        ubos = []
        for f in range(frames):
            eye, d, fov = scenes.orbit_camera(1 << depth, f, frames, 1)
            ubos.append(ve.camera_ubo(eye, d, fov, W2 / H2))

        def frame(i):
            r.set_camera(ubos[i % frames])
            r.frame(scene, fc, ve.FrameInputs(W2, H2, i, frame_seed(i), i + 1))

        for i in range(warmup):
            frame(i)
        r.sync()
        r.reset_ray_totals()
        t0 = time.perf_counter()
        for i in range(warmup, warmup + frames):
            frame(i)
        r.sync()
        dt = time.perf_counter() - t0
        t = r.ray_totals()
        rays = t["primary"] + t["traced"] + t["shadow"]
        c = r.counters()
        # per-ray traversal figures: the same frames again with node counting
        # (which traverses every bounce-0 ray too, as the reference does)
        n_st = 4
        r.enable_node_stats(True)
        r.reset_ray_totals()
        for i in range(warmup, warmup + n_st):
            frame(i)
        r.sync()
        st = r.node_stats()
        ts = r.ray_totals()
        r.enable_node_stats(False)
        prim, traced = ts["primary"], ts["traced"]
        per_ray = dict(primary_steps_per_ray=round(st["primary_steps"] / prim, 3),
                       primary_node_bytes_per_ray=round(4.0 * st["primary_reads"] / prim, 2),
                       trace_steps_per_ray=round(st["trace_steps"] / max(traced, 1), 3),
                       trace_node_bytes_per_ray=round(4.0 * st["trace_reads"] / max(traced, 1), 2),
                       steps_per_ray=round((st["primary_steps"] + st["trace_steps"]) / (prim + traced), 3),
                       node_bytes_per_ray=round(4.0 * (st["primary_reads"] + st["trace_reads"]) / (prim + traced), 2),
                       per_ray_frames=n_st)
        out = dict(workload="svo-depth12-caves-1440p-4spp-4bounce-orbit32", voxels=int(info.num_voxels),
                   candidates_per_frame=int(c[ve.CNT["CANDIDATES"]]), claim_order_exact=bool(c[ve.CNT["OVERFLOW"]] == 0),
                   node_words=int(info.size), frames=frames, ms_per_frame=round(dt / frames * 1e3, 3),
                   mrays_per_s=round(rays / dt / 1e6, 1), rays_per_frame=round(rays / frames, 1),
                   traced_rays_per_frame=round(t["traced"] / frames, 1), **per_ray)
        r.release(scene)
        return out
    finally:
        r.close()


def served_and_batched(ve, extreme, caves):
    """Served: Poisson sessions over three resident fbm octrees (depth 9/10/11),
    3 s of measured arrivals per offered load (after 0.5 s of warm-up), on the
    concurrent server with each frame placed on the least-loaded of 16 shared
    streams ("placed"), the concurrent server with one stream per session (0),
    the plain loop (one stream, 1), the batching server (frames arriving
    while a batch's primary launch runs join the next batch) and the
    whole-frame batching server (every pass of a batch's frames as one launch
    per kernel); batched: 64 uneven viewports over the depth-10 scene, the
    depth-12 cave scene and fbm depth 9 / 11 octrees, 6 frames each, the same
    five ways."""
    from voxelengine_amd import serving

    def log(m):
        print(m, file=sys.stderr, flush=True)

    modes = ("placed", 0, 1, "batched", "whole")
    sweep = serving.sweep(ve, SERVED_LAMS, 3.0, 0, modes, log=log)
    best = {q: serving.max_sustained(sweep[q]) for q in modes}
    plain = best[1]

    def ratio(q):
        return round(best[q] / plain, 3) if plain else None

    served = {
        "workload": "poisson sessions: fbm shells depth 9/10/11 (seeds 11/12/13), res U[256..2560]x[256..1440], "
                    "spp 1-8, bounces 1-4, 8-256 frames at 30 fps each; 0.5 s warm-up + 3 s measured per load",
        "concurrent": sweep["placed"],
        "per_session_streams": sweep[0],
        "plain": sweep[1],
        "batched": sweep["batched"],
        "whole_frame_batches": sweep["whole"],
        "max_frames_per_s_at_p99_33ms": {"concurrent": best["placed"], "per_session_streams": best[0], "plain": plain,
                                         "batched": best["batched"], "whole_frame_batches": best["whole"],
                                         "ratio": ratio("placed"), "ratio_per_session_streams": ratio(0),
                                         "ratio_batched": ratio("batched"), "ratio_whole_frame_batches": ratio("whole")},
    }
    d9, d11 = serving.served_scenes(((9, 7), (11, 8)))
    b = serving.batched(ve, [extreme, caves, d9, d11], 64, 6, 0, modes, log=log)
    pf = b[1]["frames_per_s"] if b[1].get("frames") else None

    def bratio(q):
        return round(b[q]["frames_per_s"] / pf, 3) if pf else None

    batched = {
        "workload": "64 viewports on {depth-10 extreme, depth-12 caves, fbm depth 9 seed 7, fbm depth 11 seed 8}, res "
                    "U[128..1920]x[128..1080], spp {1,2,4}, bounces {1,2,4}, 6 frames per viewport",
        "concurrent": b["placed"],
        "per_session_streams": b[0],
        "plain": b[1],
        "batched": b["batched"],
        "whole_frame_batches": b["whole"],
        "ratio_frames_per_s": bratio("placed"),
        "ratio_per_session_streams": bratio(0),
        "ratio_batched_frames_per_s": bratio("batched"),
        "ratio_whole_frame_batches": bratio("whole"),
    }
    return served, batched


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=128)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-served", action="store_true",
                    help="skip the served (Poisson sessions) and batched (64 viewports) measurements")
    ap.add_argument("--no-divergence", action="store_true",
                    help="skip the depth-12 divergence configuration (and, with it, the batched set's cave scene)")
    ap.add_argument("--batched-only", metavar="MODE",
                    help="run only the batched configuration (64 viewports, 6 frames each) on one server mode "
                         "(0, 1, placed, batched or whole) and print its summary (for kernel traces)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 256 MiB device copy (4 launches), the reference kernel for scaling "
                         "the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(4):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": x.numel() * 4, "launches": 4}))
        return
    if args.batched_only:
        from ve_loader import load_package
        ve = load_package()
        from voxelengine_amd import scenes as S, serving
        mode = int(args.batched_only) if args.batched_only.isdigit() else args.batched_only
        depth, ops = S.scene_extreme(1)
        words, info = ve.octree_build(ops, depth)
        cdepth, cops = S.scene_caves(1)
        cwords, cinfo = ve.octree_build(cops, cdepth, max_words=1 << 28)
        del ops, cops
        d9, d11 = serving.served_scenes(((9, 7), (11, 8)))
        b = serving.batched(ve, [(depth, words, info), (cdepth, cwords, cinfo), d9, d11], 64, 6, 0, (mode,))
        print(json.dumps({"batched_mode": str(mode), **b[mode]}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X, and the "
              "multi-GPU form (whole sessions per GPU, octree broadcast over xGMI) is not built; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    from ve_loader import load_package
    ve = load_package()
    from voxelengine_amd import scenes

    t0 = time.perf_counter()
    depth, ops = scenes.scene_extreme(1)
    words, info = ve.octree_build(ops, depth)
    t_build = time.perf_counter() - t0
    del ops

    r = ve.Renderer(W, H, hash_width=16384, max_bounces=BOUNCES + 1)
    scene = r.upload(words, depth, info.size)
    r.set_materials(scenes.palette(1))
    r.set_sky_constant(SKY)
    fc = ve.FrameConfig(**knobs())
    ubos = []
    for f in range(ORBIT_FRAMES):
        eye, d, fov = scenes.orbit_camera(1 << depth, f, ORBIT_FRAMES, 1)
        ubos.append(ve.camera_ubo(eye, d, fov, W / H))

    def frame(i):
        f = i % ORBIT_FRAMES
        r.set_camera(ubos[f])
        r.frame(scene, fc, ve.FrameInputs(W, H, i, frame_seed(i), i + 1))

    for i in range(args.warmup):
        frame(i)
    r.sync()
    c = r.counters()
    assert c[ve.CNT["OVERFLOW"]] == 0, "claim-row overflow"

    # timed region: K frames as an application renders them
    r.reset_ray_totals()
    torch.cuda.synchronize()
    r.sync()
    t0 = time.perf_counter()
    for i in range(args.warmup, args.warmup + args.steps):
        frame(i)
    r.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    totals = r.ray_totals()
    rays = totals["primary"] + totals["traced"] + totals["shadow"]
    value = rays / elapsed / 1e6

    # the same K frames with HIP events on every pass, then with the overlap off
    def timed_frames(overlap):
        r.set_overlap(overlap)
        r.enable_timing(args.steps)
        r.sync()
        for i in range(args.warmup, args.warmup + args.steps):
            frame(i)
        r.sync()
        later = r.timings_iter(1)   # the fused kernel of bounce iterations >= 1, under TRACE
        ms = r.timings()
        r.enable_timing(False)
        return ms, later

    pass_ms, later_ms = timed_frames(True)
    pass_ms_serial, later_ms_serial = timed_frames(False)
    r.set_overlap(True)

    # algorithmic node traffic of the primary kernel on the same frames
    r.enable_node_stats(True)
    for i in range(args.warmup, args.warmup + min(args.steps, ORBIT_FRAMES)):
        f = i % ORBIT_FRAMES
        r.set_camera(ubos[f])
        r.run_pass(scene, "PRIMARY", fc, ve.FrameInputs(W, H, i, frame_seed(i), i + 1))
    st = r.node_stats()
    # the fused bounce kernel's node traffic: whole frames with node counting
    n_bst = min(args.steps, 8)
    r.enable_node_stats(True)
    for i in range(args.warmup, args.warmup + n_bst):
        frame(i)
    r.sync()
    bst = r.node_stats()
    cb = r.counters()
    r.enable_node_stats(False)
    n_stat = min(args.steps, ORBIT_FRAMES)
    bytes_per_launch = st["primary_reads"] * 4.0 / n_stat
    prim_ms = pass_ms["PRIMARY"]
    achieved = bytes_per_launch / (prim_ms * 1e-3) / 1e9 if prim_ms > 0 else 0.0
    iso_ms = pass_ms_serial["PRIMARY"]
    iso_achieved = bytes_per_launch / (iso_ms * 1e-3) / 1e9 if iso_ms > 0 else 0.0
    # the fused bounce kernel (trace + shade of iterations >= 1 per ray)
    b_bytes = bst["bounce_reads"] * 4.0 / n_bst
    b_ms, b_ms_iso = later_ms["TRACE"], later_ms_serial["TRACE"]
    b_rays = totals["traced"] / args.steps - int(cb[ve.CNT["SHADE"]])
    bounce_kernel = {
        "kernel": "k_bounce (trace.comp + shade.comp, bounce iterations >= 1, fused per ray)",
        "bound": "hbm",
        "algorithmic_bytes_per_launch": round(b_bytes),
        "rays_per_launch": round(b_rays, 1),
        "node_reads_per_ray": round(bst["bounce_reads"] / n_bst / max(b_rays, 1.0), 3),
        "steps_per_ray": round(bst["bounce_steps"] / n_bst / max(b_rays, 1.0), 3),
        "launch_ms": round(b_ms, 4),
        "achieved": round(b_bytes / (b_ms * 1e-3) / 1e9, 1) if b_ms > 0 else None,
        "frac": round(b_bytes / (b_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if b_ms > 0 else None,
        "isolated_launch_ms": round(b_ms_iso, 4),
        "isolated_frac": round(b_bytes / (b_ms_iso * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if b_ms_iso > 0 else None,
        "traffic": profiled_traffic("k_bounce"),
        "grays_per_s": round(b_rays / (b_ms * 1e-3) / 1e9, 3) if b_ms > 0 else None,
    }

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(words, depth, ubos, knobs())

    # the divergence configuration (depth-12 cave scene), also resident in
    # the batched set
    r.release(scene)
    r.close()
    del ubos
    from voxelengine_amd import scenes as S
    div = caves = None
    if not (args.no_divergence and args.no_served):
        t0 = time.perf_counter()
        cdepth, cops = S.scene_caves(1)
        cwords, cinfo = ve.octree_build(cops, cdepth, max_words=1 << 28)
        t_caves = time.perf_counter() - t0
        del cops
        caves = (cdepth, cwords, cinfo)
    if not args.no_divergence:
        div = divergence(ve, caves)
        div["scene_build_s"] = round(t_caves, 2)

    served = batched = None
    if not args.no_served:
        served, batched = served_and_batched(ve, (depth, words, info), caves)

    out = {
        "metric": "Mrays/s @1080p 10.4M-voxel SVO 2-bounce",
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded fbm terrain shell + 64 spheres, seeded orbit camera)",
        "config": {
            "workload": "svo-depth10-extreme-1080p-2bounce-orbit64",
            "voxels": int(info.num_voxels),
            "node_words": int(info.size),
            "resolution": "%dx%d" % (W, H),
            "bounces": BOUNCES,
            "spp": 1,
            "frames_per_s": round(args.steps / elapsed, 2),
            "rays_per_frame": round(rays / args.steps, 1),
            "primary_rays_per_frame": W * H,
            "traced_rays_per_frame": round(totals["traced"] / args.steps, 1),
            # of the traced rays, the bounce-0 camera rays whose hit record the
            # emit pass took from the primary pass (the same traversal)
            "reused_rays_per_frame": round(totals["reused"] / args.steps, 1),
            "mrays_per_s_traversed_only": round((rays - totals["reused"]) / elapsed / 1e6, 3),
            "overlap": "primary, schedule and normal passes of frame k+1 beside the bounce passes of frame k; "
                       "resolve of frame k beside the first passes of frame k+1",
            "pass_ms": {k: round(v, 4) for k, v in pass_ms.items()},
            "pass_ms_serial": {k: round(v, 4) for k, v in pass_ms_serial.items()},
            "scene_build_s": round(t_build, 2),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_primary (generate_primary.comp + RaycastCounted)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic(),
            "traffic_source": "profiles/k_primary_extreme_1080p.md (rocprofv3 FETCH_SIZE+WRITE_SIZE, copy-calibrated)",
            "algorithmic_bytes_per_launch": round(bytes_per_launch),
            "launch_ms": round(prim_ms, 4),
            "isolated_launch_ms": round(iso_ms, 4),
            "isolated_achieved": round(iso_achieved, 1),
            "isolated_frac": round(iso_achieved / HBM_PEAK_GBS, 4),
            "node_reads_per_ray": round(st["primary_reads"] / n_stat / (W * H), 3),
            "steps_per_ray": round(st["primary_steps"] / n_stat / (W * H), 3),
        },
        "bounce_kernel": bounce_kernel,
        "cpu_baseline": cpu,
        "divergence": div,
        "served": served,
        "batched": batched,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
