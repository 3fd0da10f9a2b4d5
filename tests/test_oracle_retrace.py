"""Pin the oracle's Raycast (internal.glsl:351-400) and its FP64 twin against
brute-force voxel marching on random small grids (depth 2-4): every occupied
unit voxel's slab interval along the exact ray, in double precision, the
first one entered wins.

The reference's walk nudges each step 0.001 past the cell it leaves
(internal.glsl:245-261), so it can pass a voxel that the exact ray only clips
at an edge or corner, and two voxels entered through a shared edge are a tie.
Those are the only disagreements allowed, and only within EPS of ray length;
the rays compared have every direction component at least MIN_COMPONENT in
size, so a nudge moves at most 0.001 / MIN_COMPONENT = 0.02 along the ray.
Axis-parallel rays through cell centres have neither: they must agree
exactly.  A plausible slip (a swapped locate bit, a wrong nudge sign, a
missed bound) fails many rays by far more than EPS.
"""
import numpy as np
import pytest

MIN_COMPONENT = 0.05
EPS = 0.05


def _grid(rng, depth, density):
    n = 1 << depth
    occ = rng.random((n, n, n)) < density
    mat = rng.integers(1, 8, (n, n, n))
    xs, ys, zs = np.nonzero(occ)
    ops = np.zeros(len(xs), dtype=[("x", "<u4"), ("y", "<u4"), ("z", "<u4"), ("material", "<u4")])
    ops["x"], ops["y"], ops["z"], ops["material"] = xs, ys, zs, mat[xs, ys, zs]
    return ops, np.stack([xs, ys, zs], 1).astype(np.float64)


def _brute(o, d, cells, n):
    """First occupied unit cell along p(t) = o + t d from the walk's start
    (o + 4d, or where the ray enters the grid): (index, entry t, exit t) of
    every candidate, sorted by entry; empty if none."""
    o = np.asarray(o, np.float64) + 4.0 * np.asarray(d, np.float64)
    d = np.asarray(d, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d
        lo = (cells - o) * inv
        hi = (cells + 1.0 - o) * inv
        g_lo, g_hi = (0.0 - o) * inv, (n - o) * inv
    lo = np.where(d == 0, np.where((o >= cells) & (o <= cells + 1), -np.inf, np.inf), lo)
    hi = np.where(d == 0, np.where((o >= cells) & (o <= cells + 1), np.inf, -np.inf), hi)
    tn = np.minimum(lo, hi).max(1)
    tf = np.maximum(lo, hi).min(1)
    g_lo = np.where(d == 0, -np.inf, g_lo)
    g_hi = np.where(d == 0, np.inf, g_hi)
    t0 = max(0.0, float(np.minimum(g_lo, g_hi).max()))
    te = np.maximum(tn, t0)
    ok = tf >= te
    idx = np.nonzero(ok)[0]
    order = np.argsort(te[idx], kind="stable")
    return idx[order], te[idx][order], tf[idx][order]


def _interval(o, d, c):
    o = np.asarray(o, np.float64) + 4.0 * np.asarray(d, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        lo = (c - o) / d
        hi = (c + 1.0 - o) / d
    return np.minimum(lo, hi).max(), np.maximum(lo, hi).min()


def _explained(o, d, got, cand, te, tf, cells):
    """A walk result `got` (cell index or None) that differs from the exact
    first voxel is a knife edge: the exact first voxel is only clipped, or
    the exact ray passes within EPS of `got` and enters it (or would) within
    EPS of the exact first voxel's entry (a tie through a shared edge)."""
    if len(cand) and tf[0] - te[0] < EPS:
        return True
    if got is None:
        return False
    g_tn, g_tf = _interval(o, d, cells[got])
    if g_tf < g_tn - EPS:
        return False
    return not len(cand) or g_tn - te[0] < EPS


def _cell_of(pos, cells):
    hit = np.nonzero((cells == np.asarray(pos, np.float64)).all(1))[0]
    return int(hit[0]) if len(hit) else None


@pytest.mark.parametrize("depth,density,seed", [(2, 0.3, 1), (3, 0.15, 2), (4, 0.08, 3), (4, 0.02, 4)])
def test_raycast_matches_brute_force(oracle, depth, density, seed):
    rng = np.random.default_rng(seed)
    n = 1 << depth
    ops, cells = _grid(rng, depth, density)
    words, info = oracle.octree_build(ops, depth)
    size = info["size"]
    compared = knife = 0
    for _ in range(1500):
        o = rng.uniform(-0.5 * n, 1.5 * n, 3).astype(np.float32)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        if np.abs(d).min() < MIN_COMPONENT:
            continue
        d = d.astype(np.float32)
        hit, pos, mat, _ = oracle.raycast(words, depth, o, d, 100000, n_words=size)
        cand, te, tf = _brute(o, d, cells, n)
        got = _cell_of(pos, cells) if hit else None
        if hit:
            assert got is not None, "walk hit an empty cell %s" % (pos,)
        compared += 1
        exact = int(cand[0]) if len(cand) else None
        if got != exact:
            knife += 1
            assert _explained(o, d, got, cand, te, tf, cells), (o, d, pos, exact)
    assert compared > 1000
    assert knife <= compared // 50, (knife, compared)


@pytest.mark.parametrize("depth,seed", [(3, 5), (4, 6)])
def test_axis_rays_through_cell_centres_are_exact(oracle, depth, seed):
    rng = np.random.default_rng(seed)
    n = 1 << depth
    ops, cells = _grid(rng, depth, 0.1)
    words, info = oracle.octree_build(ops, depth)
    size = info["size"]
    occ = np.zeros((n, n, n), bool)
    occ[ops["x"], ops["y"], ops["z"]] = True
    for axis in range(3):
        for sign in (1.0, -1.0):
            for _ in range(40):
                o = rng.integers(0, n, 3).astype(np.float64) + 0.5
                o[axis] = -8.0 if sign > 0 else n + 8.0   # the walk starts 4 units along: still outside
                d = np.zeros(3)
                d[axis] = sign
                hit, pos, _, _ = oracle.raycast(words, depth, o.astype(np.float32), d.astype(np.float32), 100000,
                                                n_words=size)
                idx = [slice(None)] * 3
                for a in range(3):
                    if a != axis:
                        idx[a] = int(o[a])
                column = occ[tuple(idx)]
                ks = np.nonzero(column)[0]
                if sign < 0:
                    ks = ks[::-1]
                if len(ks) == 0:
                    assert not hit
                    continue
                want = [int(o[a]) for a in range(3)]
                want[axis] = int(ks[0])
                assert hit and tuple(pos) == tuple(want), (axis, sign, o, pos, want)


@pytest.mark.parametrize("depth,density,seed", [(3, 0.12, 7), (4, 0.05, 8)])
def test_fp64_twin_matches_brute_force(oracle, depth, density, seed):
    """The FP64 twin (and the FP32 walk beside it) over camera rays, against
    the brute force on the FP64 ray."""
    rng = np.random.default_rng(seed)
    n = 1 << depth
    ops, cells = _grid(rng, depth, density)
    words, info = oracle.octree_build(ops, depth)
    size = info["size"]
    keys = np.array([oracle.voxel_key(*c) + 1 for c in cells.astype(np.uint32)], np.uint64)
    W, H = 24, 16
    compared = knife64 = knife32 = 0
    for cam in range(12):
        eye = rng.uniform(-0.6 * n, 1.6 * n, 3)
        tgt = rng.uniform(0.3 * n, 0.7 * n, 3)
        fwd = (tgt - eye) / np.linalg.norm(tgt - eye)
        ubo = oracle.camera_ubo(eye.astype(np.float32), fwd.astype(np.float32), 60.0, W / H)
        pix = np.arange(W * H, dtype=np.uint32)
        r = oracle.retrace_f64(words, depth, ubo, W, H, 100000, pix, n_words=size)
        u = ubo.astype(np.float64)
        for k in pix:
            x, y = k % W, k // W
            nx, ny = (x + 0.5) / W * 2 - 1, (y + 0.5) / H * 2 - 1
            d = u[4:7] + nx * u[8:11] - ny * u[12:15]
            d /= np.linalg.norm(d)
            if np.abs(d).min() < MIN_COMPONENT:
                continue
            cand, te, tf = _brute(u[0:3], d, cells, n)
            exact = int(keys[cand[0]]) if len(cand) else 0
            compared += 1
            for field in ("key64", "key32"):
                got_key = int(r[field][k])
                if got_key == exact:
                    continue
                got = int(np.nonzero(keys == got_key)[0][0]) if got_key else None
                assert _explained(u[0:3], d, got, cand, te, tf, cells), (field, k, got_key, exact)
                if field == "key64":
                    knife64 += 1
                else:
                    knife32 += 1
    assert compared > 2000
    assert knife64 <= compared // 50 and knife32 <= compared // 50, (knife64, knife32, compared)


def test_brute_force_check_rejects_a_transposed_grid(oracle):
    """The same comparison with the grid's x and z swapped on the brute-force
    side (a transposed operand) leaves over a tenth of the rays unexplained."""
    rng = np.random.default_rng(9)
    depth, n = 4, 16
    ops, cells = _grid(rng, depth, 0.08)
    words, info = oracle.octree_build(ops, depth)
    swapped = cells[:, ::-1].copy()
    compared = unexplained = 0
    for _ in range(600):
        o = rng.uniform(-0.5 * n, 1.5 * n, 3).astype(np.float32)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        if np.abs(d).min() < MIN_COMPONENT:
            continue
        d = d.astype(np.float32)
        hit, pos, _, _ = oracle.raycast(words, depth, o, d, 100000, n_words=info["size"])
        cand, te, tf = _brute(o, d, swapped, n)
        got = _cell_of(pos, swapped) if hit else None
        exact = int(cand[0]) if len(cand) else None
        compared += 1
        if (hit and got is None) or (got != exact and not _explained(o, d, got, cand, te, tf, swapped)):
            unexplained += 1
    assert unexplained > compared // 10, (unexplained, compared)   # against none unexplained above
