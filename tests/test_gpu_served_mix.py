"""Sessions of the served and batched workloads on the session server, at
their real sizes: the resident fbm shells (depth 9 seeds 7 and 11, depth 10
seed 12, depth 11 seed 13) and the depth-12 cave scene, resolutions from
128x128 to 2560x1440, 1-8 samples per visible voxel, 1-4 bounces.

Each session's frames are checked three ways (Renderer::runWavefrontFrame,
renderer.cpp:282-649, per session):
  - in the mix (every session on its own stream, frames interleaved) ==
    alone (the session on a private context), bit-exact on the G-buffer,
    lBuffer, shade list and output;
  - alone == the C oracle on every structure (compare_frame), two frames each.
"""
import numpy as np
import pytest

from frame_common import compare_frame

pytestmark = pytest.mark.gpu

SKY = (140, 180, 230)
FRAMES = 2
# (scene key, W, H, spp, bounces)
SESSIONS = [("d9s11", 2560, 1440, 1, 1), ("d10s12", 1280, 720, 8, 4), ("d11s13", 1024, 576, 4, 2),
            ("d9s7", 128, 128, 2, 1), ("d10s12", 640, 360, 3, 3)]
CAVE_SESSION = ("caves", 512, 288, 2, 4)
SCENE_SPECS = {"d9s11": (9, 11), "d10s12": (10, 12), "d11s13": (11, 13), "d9s7": (9, 7)}


@pytest.fixture(scope="module")
def resident(ve, scenes):
    out = {}
    for key, (depth, seed) in SCENE_SPECS.items():
        d, ops = scenes.scene_fbm(depth, seed)
        words, info = ve.octree_build(ops, d)
        out[key] = (d, words, info)
    return out


def _knobs(spp):
    from voxelengine_amd import serving
    return serving.session_knobs(spp)


def _ubo(ve, scenes, depth, W, H, seed, k):
    eye, d, fov = scenes.orbit_camera(1 << depth, k, 240, seed)
    return ve.camera_ubo(eye, d, fov, W / H)


def _inputs(ve, W, H, seed, k):
    return ve.FrameInputs(W, H, k, (seed + 7919 * k) & 0xFFFFFFFF, k + 1)


def _snapshot(read, W, H):
    return dict(gbuf=read("GBUFFER", W * H * 16).copy(), out=read("OUTPUT", W * H * 16).copy(),
                lbuf=read("LBUFFER", 16384 * 384 * 4).copy(), shade=read("SHADE_LIST", 196608 * 8).copy(),
                counters=read("COUNTERS", 64).copy())


def _alone(ve, scenes, scene, sess, seed):
    d, words, info = scene
    _, W, H, spp, bounces = sess
    r = ve.Renderer(W, H, hash_width=16384, max_bounces=bounces + 1)
    try:
        r.set_materials(scenes.palette(1))
        r.set_sky_constant(SKY)
        sc = r.upload(words, d, info.size)
        fc = ve.FrameConfig(**_knobs(spp))
        snaps = []
        for k in range(FRAMES):
            r.set_camera(_ubo(ve, scenes, d, W, H, seed, k))
            r.frame(sc, fc, _inputs(ve, W, H, seed, k))
            snaps.append(_snapshot(lambda b, n: r.read(b, np.uint8, n), W, H))
        r.release(sc)
        return snaps
    finally:
        r.close()


@pytest.mark.parametrize("batched", [False, True, "whole"])
def test_sessions_alone_equal_in_mix(ve, scenes, resident, batched):
    """Five sessions overlap on their own streams, frame k of each submitted
    together -- one by one, or as one batch (their primary passes in one
    persistent launch, the rest replayed from each session's frame graph; or,
    whole frames, every pass one launch over the batch's sessions); every
    frame equals the session rendered alone."""
    seeds = [101 + 17 * i for i in range(len(SESSIONS))]
    alone = []
    for i, s in enumerate(SESSIONS):
        print("case served alone", i, s, flush=True)
        alone.append(_alone(ve, scenes, resident[s[0]], s, seeds[i]))
    srv = ve.Server(queues=0)
    try:
        ids = {key: srv.add_scene(w, d, info.size) for key, (d, w, info) in resident.items()}
        if batched == "whole":
            srv.set_batch(tile_w=8, grab=4, waves_per_cu=32, groups=1, whole_frame=True)
        elif batched:
            srv.set_graph(True)
            srv.set_batch(tile_w=16, grab=4, waves_per_cu=16)
        sids = [srv.open(ids[s[0]], s[1], s[2], s[4] + 1, scenes.palette(1), SKY) for s in SESSIONS]
        for k in range(FRAMES):
            print("case served mix frame", k, batched, flush=True)
            args = [(sids[i], ve.FrameConfig(**_knobs(s[3])), _inputs(ve, s[1], s[2], seeds[i], k),
                     _ubo(ve, scenes, resident[s[0]][0], s[1], s[2], seeds[i], k)) for i, s in enumerate(SESSIONS)]
            if batched:
                srv.submit_batch([a[0] for a in args], [a[1] for a in args], [a[2] for a in args],
                                 [a[3] for a in args], tags=list(range(len(args))))
            else:
                for i, a in enumerate(args):
                    srv.submit(*a, tag=i)
            got = 0
            while got < len(SESSIONS):
                got += len(srv.poll(wait=True))
            for i, s in enumerate(SESSIONS):
                mix = _snapshot(lambda b, n: srv.read(sids[i], b, np.uint8, n), s[1], s[2])
                for key in ("gbuf", "out", "lbuf", "shade", "counters"):
                    assert np.array_equal(mix[key], alone[i][k][key]), "session %d frame %d: %s" % (i, k, key)
        for sid in sids:
            srv.close_session(sid)
        srv.poll()
    finally:
        srv.close()


def _oracle_check(ve, scenes, oracle, scene, sess, seed):
    d, words, info = scene
    _, W, H, spp, bounces = sess
    r = ve.Renderer(W, H, hash_width=16384, max_bounces=bounces + 1)
    of = oracle.Frame(W, H, 16384, bounces + 1)
    try:
        mats = scenes.palette(1)
        r.set_materials(mats)
        of.set_materials(mats)
        r.set_sky_constant(SKY)
        of.set_sky_rgb8(SKY)
        sc = r.upload(words, d, info.size)
        fc_g, fc_o = ve.FrameConfig(**_knobs(spp)), oracle.FrameConfig(**_knobs(spp))
        for k in range(FRAMES):
            ubo = _ubo(ve, scenes, d, W, H, seed, k)
            inp = _inputs(ve, W, H, seed, k)
            r.set_camera(ubo)
            r.frame(sc, fc_g, inp)
            of.frame(words, d, ubo, fc_o, inp.frame_idx, inp.frame_seed, inp.time, n_words=info.size)
            compare_frame(ve, r, of, W, H, "%s frame %d" % (sess, k))
        r.release(sc)
    finally:
        r.close()
        of.close()


@pytest.mark.parametrize("i", range(len(SESSIONS)))
def test_session_matches_oracle(ve, scenes, oracle, resident, i):
    s = SESSIONS[i]
    print("case served oracle", i, s, flush=True)
    _oracle_check(ve, scenes, oracle, resident[s[0]], s, 101 + 17 * i)


def test_cave_session_matches_oracle(ve, scenes, oracle):
    """The depth-12 cave scene of the batched set (146.6 M node words)."""
    d, ops = scenes.scene_caves(1)
    words, info = ve.octree_build(ops, d, max_words=1 << 28)
    del ops
    print("case served oracle caves", CAVE_SESSION, flush=True)
    _oracle_check(ve, scenes, oracle, (d, words, info), CAVE_SESSION, 977)
