"""Pass timing (ve_enable_timing / ve_timings / ve_timings_iter) around whole
frames, overlapped and serial, as bench.py uses it: every pass that launches
kernels in a frame reports a positive mean, the threshold (folded into the
emit kernel in a whole frame) reports none, and reading the timings leaves no
error behind for the next frame's launch checks."""
import pytest

from frame_common import frame_seed

pytestmark = pytest.mark.gpu


def test_timing_rounds_and_frames_after_them(ve, scenes, small_scene):
    W, H = 160, 120
    r = ve.Renderer(W, H, hash_width=16384, max_bounces=3)
    try:
        sc = r.upload(small_scene["words"], small_scene["depth"], small_scene["info"].size)
        r.set_materials(scenes.palette(1))
        r.set_sky_constant((120, 160, 210))
        fc = ve.FrameConfig(primary_controlchecks=70, bounce_controlchecks=100)
        k = 0

        def frames(n):
            nonlocal k
            for _ in range(n):
                eye, d, fov = scenes.orbit_camera(1 << small_scene["depth"], k, 16, 3)
                r.set_camera(ve.camera_ubo(eye, d, fov, W / H))
                r.frame(sc, fc, ve.FrameInputs(W, H, k, frame_seed(k), k + 1))
                k += 1

        frames(2)
        for overlap in (True, False, True):
            print("case timing overlap", overlap, flush=True)
            r.set_overlap(overlap)
            r.enable_timing(4)
            frames(4)
            later = r.timings_iter(1)
            ms = r.timings()
            r.enable_timing(False)
            for name in ("PRIMARY", "NORMAL", "RESET", "SCHEDULE", "RANK", "EMIT", "TRACE", "SHADE", "RESOLVE"):
                assert ms[name] > 0.0, (overlap, name)
            assert ms["THRESHOLD"] == 0.0 and ms["DISPLAY"] == 0.0
            assert later["TRACE"] > 0.0
            frames(2)   # launch checks after the timing reads
        r.sync()
    finally:
        r.close()
