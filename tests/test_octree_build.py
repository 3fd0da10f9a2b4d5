"""Octree build parity: OctreeCPU::insert* + Octree::set (src/octree_cpu.cpp:40-64,
src/renderer/octree.cpp:127-217).

Pins: a hand-derived known answer for a tiny tree, the C oracle's pointer-tree
restatement against the product's sort-based builder on random inserts
(duplicates, material 0, out-of-range), decode invariants (every inserted
voxel is found at its Morton path with its last material), the 3-bit count
wrap and the DFS block order.
"""
import numpy as np
import pytest


def _ops(scenes, rows):
    a = np.zeros(len(rows), scenes.INSERT_DTYPE)
    for i, r in enumerate(rows):
        a[i] = r
    return a


def test_tiny_known_answer(ve, scenes):
    # depth 2 (4^3 grid).  Voxels (0,0,0) m=3 and (3,3,3) m=5.
    # Root children: digit of level 1 (bit 1): (0,0,0)->0, (3,3,3)->7.
    # slot0 = root: count 2, next 1 -> 1 | 2<<1 | 1<<4 = 0x15.
    # Octree::set visits root child 0 first: slot1 internal, block 9,
    # count 1 -> 1 | 1<<1 | 9<<4 = 0x93; its child digit (bit 0) 0 -> slot 9 leaf m=3 -> 3<<1 = 6.
    # then root child 7: slot8 internal, block 17 -> 1 | 1<<1 | 17<<4 = 0x113;
    # child digit 7 -> slot 24 leaf m=5 -> 10.
    words, info = ve.octree_build(_ops(scenes, [(0, 0, 0, 3), (3, 3, 3, 5)]), 2)
    exp = np.zeros(32, np.uint32)
    exp[0], exp[1], exp[9], exp[8], exp[24] = 0x15, 0x93, 6, 0x113, 10
    assert info.size == 25 and info.capacity == 32 and info.num_voxels == 2
    assert np.array_equal(words, exp)


def test_dfs_child7_first(ve, scenes):
    # Two voxels under the same root child: the subtree's children are
    # visited 7..0 (linearizeNode pushes 0..7 on a stack), so the internal
    # node of the higher digit gets the lower block.
    ops = _ops(scenes, [(0, 0, 0, 1), (1, 1, 1, 2)])   # depth 3: both under root child 0
    words, info = ve.octree_build(ops, 3)
    # slot1: root child 0 internal with block 9; its children: digit of bit 1 -> both 0 -> one child.
    assert words[1] & 1 and (words[1] >> 4) == 9
    # level-3 node at slot 9 has children digits 0 (0,0,0) and 7 (1,1,1): block 17
    assert (words[9] >> 4) == 17 and ((words[9] >> 1) & 7) == 2
    assert words[17] == 1 << 1 and words[24] == 2 << 1


def test_count_wraps_at_8(ve, scenes):
    ops = _ops(scenes, [(x, y, z, 1) for x in range(2) for y in range(2) for z in range(2)])
    words, _ = ve.octree_build(ops, 1)
    assert words[0] == 1 | (0 << 1) | (1 << 4)   # 8 children -> 3-bit count 0
    assert all(words[1 + i] == 2 for i in range(8))


def test_empty_and_out_of_range(ve, scenes):
    words, info = ve.octree_build(_ops(scenes, [(9, 0, 0, 1)]), 3)   # ignored by insert()
    assert info.size == 8 and info.capacity == 8 and not words.any()


def test_material_zero_keeps_node(ve, scenes):
    ops = _ops(scenes, [(1, 1, 1, 4), (1, 1, 1, 0), (2, 2, 2, 6)])
    words, info = ve.octree_build(ops, 2)
    assert info.num_voxels == 1
    # root child 0 still exists (count 2 at the root) but its leaf slot is 0
    assert (words[0] >> 1) & 7 == 2


def _lookup(words, depth, x, y, z):
    off = 0
    for d in range(depth):
        p2 = 1 << (depth - d)
        off += (((x & p2) != 0) << 2) | (((y & p2) != 0) << 1) | ((z & p2) != 0)
        raw = int(words[off])
        if not raw & 1:
            return 0
        off = raw >> 4
    off += ((x & 1) << 2) | ((y & 1) << 1) | (z & 1)
    return (int(words[off]) >> 1) & 0x7F


@pytest.mark.parametrize("seed,depth,n", [(1, 3, 40), (2, 4, 300), (3, 5, 2000), (4, 6, 5000)])
def test_random_vs_pointer_tree(ve, scenes, oracle, seed, depth, n):
    rng = np.random.default_rng(seed)
    N = 1 << depth
    a = np.zeros(n, scenes.INSERT_DTYPE)
    a["x"] = rng.integers(0, N + 2, n)   # a few out of range
    a["y"] = rng.integers(0, N, n)
    a["z"] = rng.integers(0, N, n)
    a["material"] = rng.integers(0, 128, n)   # includes 0 and >127 wrap via & 0x7F
    a = np.concatenate([a, a[: n // 3]])       # duplicates: last write wins
    w1, i1 = ve.octree_build(a, depth)
    w2, i2 = oracle.octree_build(a, depth)
    assert (i1.size, i1.capacity, i1.num_voxels) == (i2["size"], i2["capacity"], i2["num_voxels"])
    assert np.array_equal(w1, w2)
    # decode invariant against a brute-force voxel map
    last = {}
    for r in a:
        if r["x"] < N and r["y"] < N and r["z"] < N:
            last[(int(r["x"]), int(r["y"]), int(r["z"]))] = int(r["material"]) & 0x7F
    for (x, y, z), m in list(last.items())[:400]:
        assert _lookup(w1, depth, x, y, z) == m


def test_small_scene_vs_pointer_tree(ve, oracle, small_scene):
    w2, i2 = oracle.octree_build(small_scene["ops"], small_scene["depth"])
    info = small_scene["info"]
    assert (info.size, info.capacity, info.num_voxels) == (i2["size"], i2["capacity"], i2["num_voxels"])
    assert np.array_equal(small_scene["words"], w2)


def test_capacity_limit(ve, scenes):
    ops = _ops(scenes, [(x, 0, 0, 1) for x in range(16)])
    with pytest.raises(ve.VeError) as e:
        ve.octree_build(ops, 4, max_words=32)
    assert e.value.code == ve.VE_ECAPACITY
