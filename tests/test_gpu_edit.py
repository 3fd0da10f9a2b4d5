"""Incremental edits on the device (Octree::applyEdits + glBufferSubData,
src/renderer/octree.cpp:228-315, and the dig / build actions of the main loop,
src/voxelengine.cpp:221-265): after ve_scene_update carries only the changed
words, the primary pass on the edited scene equals the oracle's
RaycastCounted on the edited node words and a fresh upload of them; both the
in-place path and the grow path (node array larger than the allocation)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _primary(ve, renderer, scene, ubo, W, H, cc=70):
    renderer.set_camera(ubo)
    fc = ve.FrameConfig(primary_controlchecks=cc)
    renderer.run_pass(scene, "PRIMARY", fc, ve.FrameInputs(W, H, 0, 1, 1))
    return renderer.gbuffer(W, H)


def test_dig_and_build_update_the_device_scene(ve, renderer, oracle, scenes):
    depth, ops = scenes.scene_small(1)
    cpu = ve.OctreeCPU(depth)
    cpu.insert_ops(ops)
    o = ve.Octree(depth).set(cpu)
    W, H = 160, 120
    e, d, fov = scenes.camera_small()
    ubo = ve.camera_ubo(e, d, fov, W / H)
    scene = renderer.upload(o.data, depth, o.size)
    try:
        rng = np.random.default_rng(4)
        for step in range(4):
            # pick along the camera ray like the main loop, then dig or build
            hit, pos, _ = cpu.raycast(e, d)
            if step == 0:
                assert hit
            c = (np.array(pos) if hit else rng.integers(20, 108, 3)) + rng.integers(-6, 7, 3)
            if step % 2 == 0:
                R = 14
                g = np.stack(np.meshgrid(*[np.arange(c[i] - R, c[i] + R + 1) for i in range(3)], indexing="ij"),
                             -1).reshape(-1, 3)
                g = g[(((g - c) ** 2).sum(1) <= R * R) & (g >= 0).all(1)]
                cpu.remove_many(g)
            else:
                cpu.insertSphere(c.astype(np.float32), 10.0, 3 + step)
            before = o.size
            o.applyEdits(cpu)
            print("case edit step", step, "changed", len(o.changed), "size", before, "->", o.size, flush=True)
            renderer.update(scene, o)
            g1, p1 = _primary(ve, renderer, scene, ubo, W, H)
            go, po, _ = oracle.generate_primary(o.data[: o.size], depth, ubo, W, H, 70)
            assert np.array_equal(g1, go) and np.array_equal(p1, po)
        fresh = renderer.upload(o.data, depth, o.size)
        try:
            g2, p2 = _primary(ve, renderer, fresh, ubo, W, H)
        finally:
            renderer.release(fresh)
        assert np.array_equal(g1, g2) and np.array_equal(p1, p2)
        # grow: a large build appends blocks beyond the device allocation
        cpu.insertBox((0, 100, 0), (128, 104, 128), 2)
        o.applyEdits(cpu)
        renderer.update(scene, o)
        g3, p3 = _primary(ve, renderer, scene, ubo, W, H)
        go, po, _ = oracle.generate_primary(o.data[: o.size], depth, ubo, W, H, 70)
        assert np.array_equal(g3, go) and np.array_equal(p3, po)
    finally:
        renderer.release(scene)
