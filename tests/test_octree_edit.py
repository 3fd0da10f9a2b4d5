"""Editable octree: OctreeCPU::insert/remove/lookup/raycast and
Octree::set/applyEdits/findGpuSlot (src/octree_cpu.cpp:40-215,
src/renderer/octree.cpp:127-341) in the library's ve_octree_cpu.

Pins: Octree::set equals the sort-based batch builder and the C oracle's
pointer-tree restatement on random inserts; a hand-derived known answer for
remove -> applyEdits -> insert with last-freed-first block reuse; after random
edit sequences every inserted / removed voxel decodes as expected from the
words, every word that differs from the previous image is in the reported
changed slots, and block pointers stay inside the array; the picking ray
equals the node-word Raycast of the oracle on the same tree."""
import numpy as np
import pytest


def _ops(scenes, rows):
    a = np.zeros(len(rows), scenes.INSERT_DTYPE)
    for i, r in enumerate(rows):
        a[i] = r
    return a


def decode(words, depth, x, y, z):
    """Material found by descending the node words (0 = empty)."""
    off = 0
    for d in range(depth + 1):
        bit = 1 << (depth - d)
        idx = ((x & bit) != 0) << 2 | ((y & bit) != 0) << 1 | ((z & bit) != 0)
        w = int(words[off + idx])
        if d == depth or not (w & 1):
            return 0 if (w & 1) else (w >> 1) & 0x7F
        off = w >> 4
    return 0


def test_set_matches_batch_builders(ve, scenes, oracle):
    rng = np.random.default_rng(5)
    for depth in (3, 5, 7):
        n = 1 << depth
        ops = np.zeros(3000, scenes.INSERT_DTYPE)
        ops["x"], ops["y"], ops["z"] = (rng.integers(0, n + 2, (3, 3000)))
        ops["material"] = rng.integers(0, 6, 3000)
        cpu = ve.OctreeCPU(depth)
        cpu.insert_ops(ops)
        o = ve.Octree(depth).set(cpu)
        words, info = ve.octree_build(ops, depth)
        ow, oinfo = oracle.octree_build(ops, depth)
        assert (o.size, o.capacity, o.numVoxels) == (info.size, info.capacity, info.num_voxels)
        assert np.array_equal(o.data, words) and np.array_equal(o.data, ow)
        assert o.full


def test_edit_known_answer(ve, scenes):
    # depth 2: (0,0,0) m3 and (3,3,3) m5 -> the known image (root child 0's
    # block 9, root child 7's block 17).  Removing (3,3,3) frees block 17 and
    # zeroes slot 8; inserting (2,2,2) m7 under root child 7 again takes block
    # 17 back (last freed first), zeroes it, and writes the leaf at 17 + digit 0.
    cpu = ve.OctreeCPU(2)
    cpu.insert_ops(_ops(scenes, [(0, 0, 0, 3), (3, 3, 3, 5)]))
    o = ve.Octree(2).set(cpu)
    assert o.data[0] == 0x15 and o.data[1] == 0x93 and o.data[9] == 6 and o.data[8] == 0x113 and o.data[24] == 10
    cpu.remove(3, 3, 3)
    o.applyEdits(cpu)
    assert o.data[0] == 0x13 and o.data[8] == 0 and o.data[24] == 10   # stale leaf, unreachable
    assert sorted(set(o.changed.tolist())) == [0, 8] and not o.full
    cpu.insert(2, 2, 2, 7)
    o.applyEdits(cpu)
    assert o.data[0] == 0x15 and o.data[8] == 0x113 and o.data[17] == 14 and o.data[24] == 0
    assert set(range(17, 25)) | {0, 8} == set(o.changed.tolist())
    assert o.size == 25 and decode(o.data, 2, 2, 2, 2) == 7 and decode(o.data, 2, 3, 3, 3) == 0
    assert o.findGpuSlot(cpu, 2, 2, 2) == 17 and o.findGpuSlot(cpu, 0, 0, 0) == 9
    assert o.findGpuSlot(cpu, 3, 3, 3) is None
    assert cpu.lookup(2, 2, 2) == (True, 7) and cpu.lookup(3, 3, 3) == (False, 0)


def test_random_edit_sequences(ve, scenes):
    rng = np.random.default_rng(9)
    for depth in (4, 6):
        n = 1 << depth
        cpu = ve.OctreeCPU(depth)
        truth = {}
        ops = np.zeros(1500, scenes.INSERT_DTYPE)
        ops["x"], ops["y"], ops["z"] = rng.integers(0, n, (3, 1500))
        ops["material"] = rng.integers(1, 9, 1500)
        cpu.insert_ops(ops)
        for r in ops:
            truth[(int(r["x"]), int(r["y"]), int(r["z"]))] = int(r["material"])
        o = ve.Octree(depth).set(cpu)
        for step in range(12):
            prev = o.data.copy()
            prev_size = o.size
            if step % 3 == 2:   # dig a ball, like the main loop's left click
                c = rng.integers(0, n, 3)
                rr = int(rng.integers(1, max(2, n // 4)))
                g = np.stack(np.meshgrid(*[np.arange(max(0, c[i] - rr), min(n, c[i] + rr + 1)) for i in range(3)],
                                         indexing="ij"), -1).reshape(-1, 3)
                g = g[((g - c) ** 2).sum(1) <= rr * rr]
                cpu.remove_many(g)
                for p in map(tuple, g.tolist()):
                    truth.pop(p, None)
            else:
                k = int(rng.integers(1, 200))
                new = np.zeros(k, scenes.INSERT_DTYPE)
                new["x"], new["y"], new["z"] = rng.integers(0, n, (3, k))
                new["material"] = rng.integers(1, 9, k)
                cpu.insert_ops(new)
                for r in new:
                    truth[(int(r["x"]), int(r["y"]), int(r["z"]))] = int(r["material"])
            o.applyEdits(cpu)
            assert o.size >= prev_size
            if not o.full:
                cur = o.data[: len(prev)]
                diff = set(np.nonzero(cur != prev)[0].tolist())
                assert diff <= set(o.changed.tolist())
            w = o.data[: o.size]
            internal = (w & 1) == 1
            assert np.all((w[internal] >> 4) + 8 <= o.size)
            for (x, y, z), m in truth.items():
                assert decode(o.data, depth, x, y, z) == m
            for p in map(tuple, rng.integers(0, n, (300, 3)).tolist()):
                assert decode(o.data, depth, *p) == truth.get(p, 0)


def test_picking_ray_matches_node_word_raycast(ve, scenes, oracle):
    depth, ops = scenes.scene_small(1)
    cpu = ve.OctreeCPU(depth)
    cpu.insert_ops(ops)
    o = ve.Octree(depth).set(cpu)
    rng = np.random.default_rng(2)
    hits = 0
    for _ in range(300):
        eye = rng.uniform(-60, 190, 3).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        h1, p1, m1 = cpu.raycast(eye, d, max_steps=1000)
        h2, p2, m2, _ = oracle.raycast(o.data, depth, eye, d, 1000, n_words=o.size)
        assert h1 == h2 and (not h1 or (p1 == p2 and m1 == m2))
        hits += h1
    assert hits > 50


def test_word_limit_is_an_error_not_an_overflow(ve, scenes):
    """A tree that needs more words than max_words: set() and applyEdits()
    fail with VE_ECAPACITY, the image never grows past the limit, and after
    removes bring the tree back under it the dirty path is written."""
    cpu = ve.OctreeCPU(4, max_words=16)
    rng = np.random.default_rng(3)
    ops = np.zeros(64, scenes.INSERT_DTYPE)
    ops["x"], ops["y"], ops["z"] = rng.integers(0, 16, (3, 64))
    ops["material"] = 1
    cpu.insert_ops(ops)
    o = ve.Octree(4, max_words=16)
    with pytest.raises(ve.VeError) as e:
        o.set(cpu)
    assert e.value.code == ve.VE_ECAPACITY
    # the image the library holds never grew past the limit
    import ctypes as C
    w, cap = C.POINTER(C.c_uint32)(), C.c_uint32()
    assert ve.lib().ve_octree_words(cpu.h, C.byref(w), C.byref(cap)) == 0
    assert cap.value <= 16
    with pytest.raises(ve.VeError) as e1:      # no incremental sync of an incomplete image
        o.applyEdits(cpu)
    assert e1.value.code == ve.VE_ECAPACITY

    # one voxel fits in 9 + 8 * depth words; an edit that needs more fails,
    # removing it again lets the next apply succeed
    n_fit = 9 + 8 * 3
    cpu2 = ve.OctreeCPU(4, max_words=n_fit)
    cpu2.insert(1, 2, 3, 4)
    o3 = ve.Octree(4).set(cpu2)
    assert o3.size == n_fit and decode(o3.data, 4, 1, 2, 3) == 4
    cpu2.insert(14, 13, 12, 5)                 # a second path: 3 more blocks
    with pytest.raises(ve.VeError) as e3:
        o3.applyEdits(cpu2)
    assert e3.value.code == ve.VE_ECAPACITY
    assert o3.capacity <= n_fit and decode(o3.data, 4, 1, 2, 3) == 4
    cpu2.remove(1, 2, 3)                       # frees the first path's blocks
    o3.applyEdits(cpu2)
    assert decode(o3.data, 4, 14, 13, 12) == 5 and decode(o3.data, 4, 1, 2, 3) == 0
    assert o3.size <= n_fit
