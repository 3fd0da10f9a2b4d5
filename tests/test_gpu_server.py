"""Session server (ve_server_*, the served/batched form) on the GPU.

Each session's frames must equal the same frames rendered alone through
ve_frame on a private context (bit-exact under the order-independent knobs),
whether the sessions overlap on their own streams, share one stream (the
plain loop), wait for admission under a tight memory budget, or reuse a
pooled context that an earlier session left behind (queues=2: three
sessions on two shared streams)."""
import numpy as np
import pytest

from frame_common import DET_KNOBS, frame_seed

pytestmark = pytest.mark.gpu

SKY = (140, 180, 230)


def _views(scenes):
    # (W, H, max_bounces, eye, target, fov)
    return [(96, 64, 2, (-30.0, 90.0, -40.0), (64.0, 40.0, 64.0), 70.0),
            (80, 80, 3, (150.0, 80.0, 20.0), (60.0, 30.0, 70.0), 60.0),
            (64, 48, 2, (64.0, 120.0, 160.0), (64.0, 30.0, 64.0), 80.0)]


def _ubos(ve, scenes, v, frames):
    W, H, _, eye, tgt, fov = v
    out = []
    for f in range(frames):
        e = np.asarray(eye, np.float32) + np.float32(2.0 * f)
        ee, d = scenes.look_at(e, tgt)
        out.append(ve.camera_ubo(ee, d, fov, W / H))
    return out


def _alone(ve, scenes, small_scene, v, frames, fc):
    W, H, mb = v[:3]
    r = ve.Renderer(W, H, max_bounces=mb)
    try:
        r.set_materials(scenes.palette(1))
        r.set_sky_constant(SKY)
        sc = r.upload(small_scene["words"], small_scene["depth"], small_scene["info"].size)
        outs, rays = [], []
        for f, ubo in enumerate(_ubos(ve, scenes, v, frames)):
            r.reset_ray_totals()
            r.set_camera(ubo)
            r.frame(sc, fc, ve.FrameInputs(W, H, f, frame_seed(f), f + 1))
            outs.append(r.read("OUTPUT", np.float32, W * H * 16).copy())
            t = r.ray_totals()
            rays.append(t["primary"] + t["traced"] + t["shadow"])
        r.release(sc)
        return outs, rays
    finally:
        r.close()


def _served(ve, scenes, small_scene, views, frames, fc, place=False, **server_kw):
    """All views as concurrent sessions, frames submitted round-robin; the
    output of each session is read after each of its frames completes."""
    srv = ve.Server(**server_kw)
    if place:
        srv.set_placement(True)
    try:
        scene = srv.add_scene(small_scene["words"], small_scene["depth"], small_scene["info"].size)
        sids = [srv.open(scene, v[0], v[1], v[2], scenes.palette(1), SKY) for v in views]
        ubos = [_ubos(ve, scenes, v, frames) for v in views]
        outs = {s: [] for s in sids}
        rays = {s: [] for s in sids}
        done = []
        for f in range(frames):
            for k, (s, v) in enumerate(zip(sids, views)):
                srv.submit(s, fc, ve.FrameInputs(v[0], v[1], f, frame_seed(f), f + 1), ubos[k][f], tag=f)
            # each session's output buffer holds its latest frame: collect this
            # round before the next overwrites it
            got = 0
            while got < len(sids):
                batch = srv.poll(wait=True)
                got += len(batch)
                done += batch
            for s, v in zip(sids, views):
                outs[s].append(srv.read_output(s, v[0], v[1]))
        for d in done:
            rays[d["session"]].append(d["rays"])
        st = srv.stats()
        for s in sids:
            srv.close_session(s)
        srv.poll()
        return sids, outs, rays, done, st, srv.stats()
    finally:
        srv.close()


@pytest.mark.parametrize("queues", [0, 1, 2])
def test_sessions_alone_equal_in_mix(ve, scenes, small_scene, queues):
    print("case server mix queues", queues, flush=True)
    fc = ve.FrameConfig(**DET_KNOBS)
    views = _views(scenes)
    frames = 3
    ref = [_alone(ve, scenes, small_scene, v, frames, fc) for v in views]
    sids, outs, rays, done, st, st_after = _served(ve, scenes, small_scene, views, frames, fc, queues=queues)
    assert st["admitted"] == len(views) and st["waiting"] == 0
    for k, s in enumerate(sids):
        for f in range(frames):
            assert np.array_equal(outs[s][f], ref[k][0][f]), (k, f)
        assert rays[s] == ref[k][1]
    for d in done:
        assert d["arrival_ms"] <= d["start_ms"] + 1.0 and d["start_ms"] <= d["done_ms"]
    if queues == 1:   # one stream: frames never overlap on the device
        ivs = sorted((d["start_ms"], d["done_ms"]) for d in done)
        for a, b in zip(ivs, ivs[1:]):
            assert b[0] >= a[1] - 1e-3
    assert st_after["admitted"] == 0 and st_after["pooled"] == len(views)


def test_placement_by_load_keeps_each_session_in_order(ve, scenes, small_scene):
    """Frames placed on the least-loaded of two shared streams: a session's
    frames move between streams (each waiting for its predecessor on the
    device), yet every frame equals the frame rendered alone -- per round,
    and with all rounds submitted at once (the outputs then compared after
    the last frame, the ray counts frame by frame)."""
    print("case server placement", flush=True)
    fc = ve.FrameConfig(**DET_KNOBS)
    views = _views(scenes)
    frames = 3
    ref = [_alone(ve, scenes, small_scene, v, frames, fc) for v in views]
    sids, outs, rays, done, st, st_after = _served(ve, scenes, small_scene, views, frames, fc, place=True, queues=2)
    for k, s in enumerate(sids):
        for f in range(frames):
            assert np.array_equal(outs[s][f], ref[k][0][f]), (k, f)
        assert rays[s] == ref[k][1]
    srv = ve.Server(queues=2)
    srv.set_placement(True)
    try:
        scene = srv.add_scene(small_scene["words"], small_scene["depth"], small_scene["info"].size)
        sids = [srv.open(scene, v[0], v[1], v[2], scenes.palette(1), SKY) for v in views]
        ubos = [_ubos(ve, scenes, v, frames) for v in views]
        for f in range(frames):
            for k, (s, v) in enumerate(zip(sids, views)):
                srv.submit(s, fc, ve.FrameInputs(v[0], v[1], f, frame_seed(f), f + 1), ubos[k][f], tag=f)
        done = []
        while len(done) < frames * len(sids):
            done += srv.poll(wait=True)
        for k, (s, v) in enumerate(zip(sids, views)):
            assert np.array_equal(srv.read_output(s, v[0], v[1]), ref[k][0][-1]), k
            got = sorted((d["frame"], d["rays"]) for d in done if d["session"] == s)
            assert [r for _, r in got] == ref[k][1], k
        with pytest.raises(ve.VeError) as e:
            srv.set_placement(False)   # sessions are open
        assert e.value.code == ve.VE_ESTATE
    finally:
        srv.close()
    one = ve.Server(queues=0)
    try:
        with pytest.raises(ve.VeError) as e:
            one.set_placement(True)
        assert e.value.code == ve.VE_EINVAL
    finally:
        one.close()


def test_admission_waits_for_budget(ve, scenes, small_scene):
    """A budget that holds one session: the second is admitted only when the
    first closes, reuses its pooled context, and still renders as alone."""
    print("case server admission", flush=True)
    fc = ve.FrameConfig(**DET_KNOBS)
    v = _views(scenes)[0]
    W, H, mb = v[:3]
    ref_outs, _ = _alone(ve, scenes, small_scene, v, 2, fc)
    probe = ve.Server()
    need = 0
    try:
        sc = probe.add_scene(small_scene["words"], small_scene["depth"], small_scene["info"].size)
        s0 = probe.open(sc, W, H, mb, scenes.palette(1), SKY)
        need = probe.stats()["bytes_in_use"]
        probe.close_session(s0)
    finally:
        probe.close()
    assert need > 0
    srv = ve.Server(hbm_budget=need + need // 2)
    try:
        sc = srv.add_scene(small_scene["words"], small_scene["depth"], small_scene["info"].size)
        a = srv.open(sc, W, H, mb, scenes.palette(1), SKY)
        b = srv.open(sc, W, H, mb, scenes.palette(1), SKY)
        st = srv.stats()
        assert st["admitted"] == 1 and st["waiting"] == 1
        ubos = _ubos(ve, scenes, v, 2)
        for s in (a, b):
            for f in range(2):
                srv.submit(s, fc, ve.FrameInputs(W, H, f, frame_seed(f), f + 1), ubos[f])
        first = []
        while len(first) < 2:
            first += srv.poll(wait=True)
        assert {d["session"] for d in first} == {a}
        out_a = srv.read_output(a, W, H)
        srv.close_session(a)          # frees the slot: b is admitted on a's context
        second = []
        while len(second) < 2:
            second += srv.poll(wait=True)
        assert {d["session"] for d in second} == {b}
        assert min(d["start_ms"] for d in second) >= max(d["done_ms"] for d in first) - 1e-3
        assert srv.stats()["pooled"] == 0 and srv.stats()["admitted"] == 1
        assert np.array_equal(out_a, ref_outs[-1])
        assert np.array_equal(srv.read_output(b, W, H), ref_outs[-1])   # pooled context fully reset
    finally:
        srv.close()


def test_closed_waiting_session_still_renders(ve, scenes, small_scene):
    """ve_server_close closes after the session's submitted frames, also for a
    session still waiting for a slot: it keeps its place, its frames run once
    it is admitted (when the running session retires), and they render as
    alone; every submitted frame is reported."""
    print("case server close while waiting", flush=True)
    fc = ve.FrameConfig(**DET_KNOBS)
    v = _views(scenes)[0]
    W, H, mb = v[:3]
    ref_outs, _ = _alone(ve, scenes, small_scene, v, 2, fc)
    srv = ve.Server(max_sessions=1)
    try:
        sc = srv.add_scene(small_scene["words"], small_scene["depth"], small_scene["info"].size)
        a = srv.open(sc, W, H, mb, scenes.palette(1), SKY)
        b = srv.open(sc, W, H, mb, scenes.palette(1), SKY)
        assert srv.stats()["waiting"] == 1
        ubos = _ubos(ve, scenes, v, 2)
        for s in (a, b):
            for f in range(2):
                srv.submit(s, fc, ve.FrameInputs(W, H, f, frame_seed(f), f + 1), ubos[f])
        srv.close_session(b)          # still waiting: its two frames stay queued
        srv.close_session(a)
        done = []
        while len(done) < 4:
            got = srv.poll(wait=True)
            assert got or srv.stats()["frames_in_flight"] > 0 or srv.stats()["waiting"] > 0
            done += got
        assert sorted((d["session"], d["frame"]) for d in done) == [(a, 0), (a, 1), (b, 0), (b, 1)]
        db = [d for d in done if d["session"] == b]
        da = [d for d in done if d["session"] == a]
        assert min(d["start_ms"] for d in db) >= max(d["done_ms"] for d in da) - 1e-3
        st = srv.stats()
        assert st["admitted"] == 0 and st["waiting"] == 0 and st["frames_in_flight"] == 0
    finally:
        srv.close()


def test_server_rejects_bad_requests(ve, scenes, small_scene):
    srv = ve.Server(hbm_budget=1 << 20)
    try:
        sc = srv.add_scene(small_scene["words"], small_scene["depth"], small_scene["info"].size)
        with pytest.raises(ve.VeError) as e:
            srv.open(sc, 64, 64, 2, scenes.palette(1), SKY)   # alone above the whole budget
        assert e.value.code == ve.VE_EBUDGET
        with pytest.raises(ve.VeError) as e:
            srv.open(sc + 5, 64, 64, 2, scenes.palette(1), SKY)
        assert e.value.code == ve.VE_ENOTFOUND
        with pytest.raises(ve.VeError) as e:
            srv.submit(12345, ve.FrameConfig(), ve.FrameInputs(64, 64, 0, 1, 1), np.zeros(16, np.float32))
        assert e.value.code == ve.VE_ENOTFOUND
    finally:
        srv.close()
