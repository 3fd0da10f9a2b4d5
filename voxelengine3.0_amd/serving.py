"""Served workload: Poisson session arrivals driven against the session
server (ve_server_*), reporting frame latency percentiles, frames/s and
Mrays/s.  The same request schedule runs on the concurrent server (one
stream per session, or frames placed on shared streams by load) and on the
plain loop (every frame on one stream, in arrival order), so they are
compared on identical work.

Workload (the served configuration): three resident fbm terrain-shell
octrees (depth 9, 10, 11; seeds 11, 12, 13); each session picks one, a
resolution uniform in 256x256..2560x1440, 1-8 samples per visible voxel,
1-4 bounces and a length of 8-256 frames, and requests its frames at a
steady 30 frames/s from its arrival (a frame interval of 33 ms).  Sessions
arrive as a Poisson process whose rate makes the offered load `lam`
frames/s.

Modes: 0 = the concurrent server (one stream per session, each frame
submitted as it falls due); "placed" = the concurrent server over 16 shared
streams, each frame joining the stream whose queued frames hold the least
estimated work (ve_server_set_placement); 1 = the plain loop (every frame on
one stream, in arrival order); "batched" = the concurrent server with continuous batching:
frames go in through ve_server_submit_batch and wait while the previous
batch's primary launch runs, so the frames falling due meanwhile share one
persistent primary launch over a global tile queue; every other pass runs
on the session's own stream; "whole" = the same batching with whole-frame
batches: after the shared primary launch, every other pass of the batch's
frames runs as one launch per kernel over all of them.
"""
import json
import os
import time

import numpy as np

from . import scenes as S

CLIENT_FPS = 30.0
LEN_RANGE = (8, 256)


def served_scenes(specs=((9, 11), (10, 12), (11, 13))):
    """(depth, words, info) of the resident octrees."""
    from . import octree_build
    out = []
    for depth, seed in specs:
        d, ops = S.scene_fbm(depth, seed)
        words, info = octree_build(ops, d)
        out.append((d, words, info))
    return out


def session_knobs(spp):
    """FrameConfig of a session with `spp` samples per visible voxel: rank
    weights zero and empty top/mid tiers put every voxel in the bottom tier,
    which takes min_samples (rank.comp:116-118)."""
    return dict(primary_controlchecks=70, bounce_controlchecks=120, normal_precision=1, shading_budget=196608,
                weight_pixels=0.0, weight_variance=0.0, top_fraction=0.0, mid_fraction=0.0, min_samples=int(spp),
                sample_cap_direct=1024, sample_cap_indirect=4096, disable_stale_reset=1, update_time=0xFFFFFFFF)


LAM_MAX = 4000.0


def make_workload(lam, duration_s, seed=0, n_scenes=3):
    """Sessions of a Poisson arrival process with offered load `lam` frames/s
    over `duration_s` seconds.  The window starts in steady state: the
    sessions already running at t=0 are drawn too (with their remaining
    lengths).  Frames requested after the window are dropped.  The sessions
    of a lower load are a random thinning of those of LAM_MAX with the same
    seed, so the workloads of one sweep are nested."""
    rng = np.random.default_rng(seed)
    keep_rng = np.random.default_rng(seed + 0x5EED)
    mean_len = 0.5 * (LEN_RANGE[0] + LEN_RANGE[1])
    rate = LAM_MAX / mean_len                  # sessions per second before thinning
    keep = min(1.0, lam / LAM_MAX)
    sessions = []

    lengths = np.arange(LEN_RANGE[0], LEN_RANGE[1] + 1)

    def draw(t0, residual):
        if residual:
            # a session running at a random instant: its length is drawn in
            # proportion to itself, and the part still to come is uniform
            length = int(rng.choice(lengths, p=lengths / lengths.sum()))
            length = int(rng.integers(1, length + 1))
        else:
            length = int(rng.integers(LEN_RANGE[0], LEN_RANGE[1] + 1))
        sc = int(rng.integers(0, n_scenes))
        W, H = int(rng.integers(256, 2561)), int(rng.integers(256, 1441))
        s = dict(arrival_s=t0, scene=sc, W=W, H=H, spp=int(rng.integers(1, 9)), bounces=int(rng.integers(1, 5)),
                 length=length, seed=int(rng.integers(1, 1 << 30)), running=bool(residual))
        s["frames_s"] = [t0 + k / CLIENT_FPS for k in range(length) if t0 + k / CLIENT_FPS < duration_s]
        if keep_rng.uniform() < keep and s["frames_s"]:
            sessions.append(s)

    for _ in range(int(rng.poisson(rate * mean_len / CLIENT_FPS))):   # running at t=0, at random phases
        draw(float(rng.uniform(0.0, 1.0 / CLIENT_FPS)), True)
    t = float(rng.exponential(1.0 / rate))
    while t < duration_s:
        draw(t, False)
        t += float(rng.exponential(1.0 / rate))
    return sessions


TUNING_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
DEFAULT_BATCH = dict(tile_w=8, grab=4, waves_per_cu=16, groups=4)


def batch_tuning():
    """The batched primary's launch shape chosen by autotune.py (tuning.json),
    else the library defaults."""
    try:
        with open(TUNING_PATH) as fh:
            best = json.load(fh)["batched_primary"]["best"]
        return {k: int(best[k]) for k in DEFAULT_BATCH}
    except (OSError, KeyError, ValueError):
        return dict(DEFAULT_BATCH)


PLACED_STREAMS = 16   # one per hardware queue (bench.py sets GPU_MAX_HW_QUEUES=16)


def make_server(ve, mode, max_sessions=64):
    """A session server for a mode (0, 1, "placed", "batched" or "whole").  The batching server
    launches each session's other passes directly: replaying them from frame
    graphs measured slower here (2498 against 2561 frames/s on the batched
    configuration, and more host time per frame in the served sweep)."""
    if mode == "batched":
        srv = ve.Server(queues=0, max_sessions=max_sessions)
        srv.set_batch(**batch_tuning())
        return srv
    if mode == "whole":
        srv = ve.Server(queues=0, max_sessions=max_sessions)
        srv.set_batch(**dict(batch_tuning(), groups=1, whole_frame=True))
        return srv
    if mode == "placed":
        srv = ve.Server(queues=PLACED_STREAMS, max_sessions=max_sessions)
        srv.set_placement(True)
        return srv
    return ve.Server(queues=int(mode), max_sessions=max_sessions)


def _ubo(ve, s, depth, k):
    n = 1 << depth
    eye, d, fov = S.orbit_camera(n, k, 240, s["seed"])
    return ve.camera_ubo(eye, d, fov, s["W"] / s["H"])


def run_workload(ve, srv, scene_ids, scene_depths, sessions, materials, sky=(140, 180, 230)):
    """Replay `sessions` in real time against `srv`; returns per-frame records
    (latency measured from each frame's scheduled arrival to the end of its
    last kernel on the device).  The sessions already running when the window
    starts are opened before the clock starts, as a server in steady state
    holds them; every later session is opened at its arrival.  Python's
    cyclic garbage collector is held off during the replay (the loop stands
    in for a native serving thread; a collection pass would pause
    submissions for tens of milliseconds)."""
    import gc
    events = []
    for i, s in enumerate(sessions):
        for k, ts in enumerate(s["frames_s"]):
            events.append((ts, i, k))
    events.sort()
    fcs = [ve.FrameConfig(**session_knobs(s["spp"])) for s in sessions]
    sid = [None] * len(sessions)
    for i, s in enumerate(sessions):
        if s.get("running"):
            sid[i] = srv.open(scene_ids[s["scene"]], s["W"], s["H"], s["bounces"] + 1, materials, sky)
    records = []
    gc.collect()
    gc_was_on = gc.isenabled()
    gc.disable()
    try:
        return _replay(ve, srv, scene_depths, scene_ids, sessions, events, fcs, sid, records, materials, sky)
    finally:
        if gc_was_on:
            gc.enable()


def _replay(ve, srv, scene_depths, scene_ids, sessions, events, fcs, sid, records, materials, sky):
    t_base = srv.now_ms() + 50.0              # schedule starts 50 ms from now
    run_workload.t_base = t_base
    n_events, e = len(events), 0
    outstanding = 0
    batched = getattr(srv, "batched", False)
    while e < n_events or outstanding:
        now = srv.now_ms()
        due, closing = [], []
        while e < n_events and t_base + events[e][0] * 1e3 <= now:
            ts, i, k = events[e]
            s = sessions[i]
            if sid[i] is None:
                sid[i] = srv.open(scene_ids[s["scene"]], s["W"], s["H"], s["bounces"] + 1, materials, sky)
            fi = ve.FrameInputs(s["W"], s["H"], k, (s["seed"] + 7919 * k) & 0xFFFFFFFF, k + 1)
            req = (sid[i], fcs[i], fi, _ubo(ve, s, scene_depths[s["scene"]], k), i, t_base + ts * 1e3)
            if batched:
                if any(d[0] == sid[i] for d in due):   # one frame per session per batch
                    _submit_due(srv, due, closing)
                due.append(req)
                if k == len(s["frames_s"]) - 1:
                    closing.append(sid[i])
            else:
                srv.submit(req[0], req[1], req[2], req[3], tag=req[4], arrival_ms=req[5])
                if k == len(s["frames_s"]) - 1:
                    srv.close_session(sid[i])
            outstanding += 1
            e += 1
        _submit_due(srv, due, closing)
        # never block on the device while arrivals are still to come: a wait
        # for the oldest frame in flight would hold back the frames falling
        # due meanwhile
        got = srv.poll(wait=e >= n_events and outstanding > 0)
        for d in got:
            outstanding -= 1
            records.append(d)
        if not got and e < n_events:
            time.sleep(max(0.0, min(0.0002, (t_base + events[e][0] * 1e3 - srv.now_ms()) * 1e-3)))
    return records


def _submit_due(srv, due, closing):
    """Submit the collected frames as one batch, then close the sessions whose
    last frame was among them."""
    if due:
        srv.submit_batch([d[0] for d in due], [d[1] for d in due], [d[2] for d in due], [d[3] for d in due],
                         tags=[d[4] for d in due], arrival_ms=[d[5] for d in due])
    for c in closing:
        srv.close_session(c)
    del due[:], closing[:]


def summarize(records, after_ms=None):
    """Latency percentiles, frames/s and Mrays/s; with after_ms, only frames
    that arrived from then on (the warm-up before it is left out)."""
    if after_ms is not None:
        records = [r for r in records if r["arrival_ms"] >= after_ms]
    if not records:
        return dict(frames=0)
    lat = np.array([r["done_ms"] - r["arrival_ms"] for r in records])
    t0 = min(r["arrival_ms"] for r in records)
    t1 = max(r["done_ms"] for r in records)
    rays = sum(r["rays"] for r in records)
    span_s = (t1 - t0) * 1e-3
    return dict(frames=len(records), p50_ms=round(float(np.percentile(lat, 50)), 3),
                p99_ms=round(float(np.percentile(lat, 99)), 3), frames_per_s=round(len(records) / span_s, 1),
                mrays_per_s=round(rays / span_s / 1e6, 1))


P99_MS = 33.0   # the latency bound of the served comparison (one 30 frames/s client interval)


def sweep(ve, lams, duration_s=3.0, seed=0, queues_list=(0, 1), log=None, res_scenes=None, max_sessions=96,
          warmup_s=0.5, refine=3):
    """Run the served workload at each offered load on the concurrent server
    and on the plain loop: warmup_s + duration_s seconds of arrivals, of
    which the first warmup_s (session contexts being created for the first
    time) are left out of the statistics.  `lams` is one list for every mode
    or a dict {queues: list}.  Then, per mode, `refine` bisection steps
    between the highest load whose p99 latency stays within P99_MS and the
    next load of the grid (the workloads of one seed are nested, so the
    realised load grows with lam).  Returns {queues: [summary per lam]},
    rows in the order run."""
    res_scenes = res_scenes or served_scenes()
    mats = S.palette(1)
    out = {}

    def run(q, lam):
        srv = make_server(ve, q, max_sessions)
        srv.batched = q in ("batched", "whole")
        try:
            ids = [srv.add_scene(w, d, info.size) for d, w, info in res_scenes]
            sessions = make_workload(lam, warmup_s + duration_s, seed, len(res_scenes))
            rec = run_workload(ve, srv, ids, [d for d, _, _ in res_scenes], sessions, mats)
            t_measure = run_workload.t_base + warmup_s * 1e3
        finally:
            srv.close()
        offered = sum(1 for x in sessions for t in x["frames_s"] if t >= warmup_s) / duration_s
        row = dict(lam=lam, offered_frames_per_s=round(offered, 1), sessions=len(sessions),
                   **summarize(rec, t_measure))
        if log:
            log("served mode=%s %s" % (q, row))
        return row

    def ok(row):
        return bool(row.get("frames")) and row["p99_ms"] <= P99_MS

    for q in queues_list:
        rows = [run(q, lam) for lam in (lams[q] if isinstance(lams, dict) else lams)]
        passing = [r["lam"] for r in rows if ok(r)]
        if passing:
            lo = max(passing)
            above = [r["lam"] for r in rows if r["lam"] > lo]
            hi = min(above) if above else None
            for _ in range(refine if hi is not None else 0):
                mid = int(round((lo + hi) / 20.0)) * 10
                if mid <= lo or mid >= hi:
                    break
                r = run(q, mid)
                rows.append(r)
                if ok(r):
                    lo = mid
                else:
                    hi = mid
        out[q] = rows
    return out


def make_viewports(count, n_scenes, seed=0):
    """The batched configuration's viewports: each on a random resident scene,
    resolution uniform in 128x128..1920x1080, spp in {1, 2, 4}, bounces in
    {1, 2, 4}, an independent seeded orbit camera."""
    rng = np.random.default_rng(seed)
    return [dict(scene=int(rng.integers(0, n_scenes)), W=int(rng.integers(128, 1921)), H=int(rng.integers(128, 1081)),
                 spp=int(rng.choice([1, 2, 4])), bounces=int(rng.choice([1, 2, 4])),
                 seed=int(rng.integers(1, 1 << 30))) for _ in range(count)]


def run_batched(ve, srv, scene_ids, scene_depths, views, frames, materials, sky=(140, 180, 230)):
    """Every viewport renders `frames` frames; frame k of all viewports is
    submitted together and the batch waits for all of them before frame k+1
    (a renderer producing one image per view per tick).  Returns the summary
    over the whole run."""
    sids = [srv.open(scene_ids[v["scene"]], v["W"], v["H"], v["bounces"] + 1, materials, sky) for v in views]
    fcs = [ve.FrameConfig(**session_knobs(v["spp"])) for v in views]
    records = []
    batched = getattr(srv, "batched", False)
    for k in range(frames):
        ins = [ve.FrameInputs(v["W"], v["H"], k, (v["seed"] + 7919 * k) & 0xFFFFFFFF, k + 1) for v in views]
        ubos = [_ubo(ve, v, scene_depths[v["scene"]], k) for v in views]
        if batched:
            srv.submit_batch(sids, fcs, ins, ubos, tags=list(range(len(views))))
        else:
            for i in range(len(views)):
                srv.submit(sids[i], fcs[i], ins[i], ubos[i], tag=i)
        got = 0
        while got < len(views):
            batch = srv.poll(wait=True)
            got += len(batch)
            records += batch
    for s in sids:
        srv.close_session(s)
    srv.poll()
    return summarize(records)


def batched(ve, res_scenes, count=64, frames=6, seed=0, queues_list=(0, 1, "batched"), log=None):
    """The batched configuration on the servers of `queues_list` (make_server's
    modes): the batching one runs a tick's primary passes in one persistent
    launch."""
    mats = S.palette(1)
    views = make_viewports(count, len(res_scenes), seed)
    out = {}
    for q in queues_list:
        srv = make_server(ve, q, max(64, count))
        srv.batched = q in ("batched", "whole")
        try:
            ids = [srv.add_scene(w, d, info.size) for d, w, info in res_scenes]
            run_batched(ve, srv, ids, [d for d, _, _ in res_scenes], views, 1, mats)   # warm: contexts created
            out[q] = run_batched(ve, srv, ids, [d for d, _, _ in res_scenes], views, frames, mats)
        finally:
            srv.close()
        if log:
            log("batched mode=%s %s" % (q, out[q]))
    return out


def max_sustained(rows, p99_ms=P99_MS):
    """Highest achieved frames/s over the loads whose p99 latency stays within p99_ms."""
    ok = [r["frames_per_s"] for r in rows if r.get("frames") and r["p99_ms"] <= p99_ms]
    return max(ok) if ok else 0.0


__all__ = ["served_scenes", "session_knobs", "make_workload", "run_workload", "summarize", "sweep", "max_sustained",
           "make_viewports", "run_batched", "batched", "make_server", "batch_tuning", "CLIENT_FPS"]
