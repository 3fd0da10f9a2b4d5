"""Wall time of the reference's own GLSL passes on Mesa llvmpipe (the CPU
baseline of the reference method): whole Renderer::runWavefrontFrame frames
(renderer.cpp:282-649, replayed by replay.GLPipeline) of the small depth-7
configuration (256x256, 1 bounce) or the depth-10 configuration (1920x1080,
2 bounces, 10.4 M voxels, orbit camera), with the benchmark's knobs.

    LP_NUM_THREADS=8 python oracle/glsl/time_llvmpipe.py small|extreme [frames] [warmup]

Prints one JSON line: median and mean ms per frame after the warm-up frames
(each frame ends with glFinish; shader compilation happens before the first
warm-up frame), Mrays/s counted as the benchmark counts rays (W*H primary
rays + every ray in each trace dispatch's window; the palette has no
emissive material, so there are no shadow rays), the thread count and the
CPU model.  The ray-window readbacks between dispatches are part of the
reference's frame loop (renderer.cpp:517-606) and are timed with it.
Needs the reference checkout (the shaders are read from it at run time).
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)

os.environ.setdefault("LP_NUM_THREADS", "8")

import numpy as np  # noqa: E402

from oracle.glsl import gl as G  # noqa: E402
from oracle.glsl import replay as R  # noqa: E402
from ve_loader import load_package  # noqa: E402

KNOBS = dict(primary_controlchecks=70, bounce_controlchecks=120, normal_precision=1, shading_budget=196608,
             weight_pixels=0.0, weight_variance=0.0, top_fraction=0.0, mid_fraction=1.0, min_samples=1,
             sample_cap_direct=1024, sample_cap_indirect=4096, disable_stale_reset=1, update_time=0xFFFFFFFF)
SKY = (140, 180, 230)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    which = sys.argv[1] if len(sys.argv) > 1 else "small"
    frames = int(sys.argv[2]) if len(sys.argv) > 2 else 8
    warmup = int(sys.argv[3]) if len(sys.argv) > 3 else 2
    ve = load_package()
    from voxelengine_amd import scenes
    from oracle import oracle as O
    if which == "small":
        depth, ops = scenes.scene_small(1)
        W, H, max_bounces = 256, 256, 2
        cams = [scenes.camera_small()]
        workload = "svo-depth7-small-256x256-1bounce"
    else:
        depth, ops = scenes.scene_extreme(1)
        W, H, max_bounces = 1920, 1080, 3
        cams = [scenes.orbit_camera(1 << depth, f, 64, 1) for f in range(64)]
        workload = "svo-depth10-extreme-1080p-2bounce-orbit64"
    words, info = ve.octree_build(ops, depth)
    del ops
    pl = R.GLPipeline(words, depth, W, H, max_bounces=max_bounces, materials=scenes.palette(1), sky_rgb8=SKY)
    fc = O.FrameConfig(**KNOBS).as_dict()
    times, rays = [], []
    for i in range(warmup + frames):
        eye, d, fov = cams[i % len(cams)]
        pl.set_camera(ve.camera_ubo(eye, d, fov, W / H))
        t0 = time.perf_counter()
        st = pl.frame(fc, i, (1 + 7919 * i) & 0xFFFFFFFF, i + 1)
        G.GL.get().glFinish()
        dt = time.perf_counter() - t0
        print("frame %d: %.1f ms, %s" % (i, dt * 1e3, st), file=sys.stderr, flush=True)
        if i >= warmup:
            times.append(dt)
            rays.append(W * H + st["window"])
    t = np.array(times)
    print(json.dumps(dict(workload=workload, driver="Mesa llvmpipe", threads=int(os.environ["LP_NUM_THREADS"]),
                          cpu=cpu_model(), nproc=os.cpu_count(), frames=frames, warmup=warmup,
                          ms_per_frame_median=round(float(np.median(t)) * 1e3, 1),
                          ms_per_frame_mean=round(float(t.mean()) * 1e3, 1),
                          mrays_per_s=round(float(np.sum(rays)) / float(t.sum()) / 1e6, 3),
                          rays_per_frame=round(float(np.mean(rays)), 1))))


if __name__ == "__main__":
    main()
