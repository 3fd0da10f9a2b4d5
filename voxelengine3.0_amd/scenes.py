"""Seeded synthetic scenes and cameras for the benchmark and parity workloads.

The reference ships no scene generator (its demo builds a Cornell box by hand,
src/voxelengine.cpp:129-142, and its Noise/ code is unseeded and unused), so
this module defines the workloads exactly.  Everything is integer hashing plus
float64 arithmetic on the host, so the same seed gives the same voxels on any
machine.  Scenes are returned as ordered insert lists — the calls that
OctreeCPU::insertFunction / insertSphere (src/octree_cpu.cpp:221-262) would
make, in the same loop order — so the octree build sees exactly the
reference's insertion sequence (later inserts of a position win).

Value noise: lattice values hash(ix, iz, octave, seed) in [0, 1) with
smoothstep interpolation; fbm = 4 octaves, lacunarity 2, gain 0.5, normalised
by the amplitude sum; base frequency 4/N.
"""
import math

import numpy as np

U32 = np.uint32
M32 = 0xFFFFFFFF

INSERT_DTYPE = np.dtype([("x", "<u4"), ("y", "<u4"), ("z", "<u4"), ("material", "<u4")])


def _mix32(h):
    h = h.astype(U32)
    h ^= h >> U32(16)
    h *= U32(0x7FEB352D)
    h ^= h >> U32(15)
    h *= U32(0x846CA68B)
    h ^= h >> U32(16)
    return h


def hash4(a, b, c, d):
    """32-bit integer hash of four uint32 lanes (numpy arrays or scalars)."""
    with np.errstate(over="ignore"):
        a = np.asarray(a, np.int64).astype(U32)
        b = np.asarray(b, np.int64).astype(U32)
        c = np.asarray(c, np.int64).astype(U32)
        d = np.asarray(d, np.int64).astype(U32)
        h = _mix32(a * U32(0x9E3779B1) + U32(0x632BE5AB))
        h = _mix32(h ^ (b * U32(0x85EBCA77)))
        h = _mix32(h ^ (c * U32(0xC2B2AE3D)))
        h = _mix32(h ^ (d * U32(0x27D4EB2F)))
    return h


def unit(h):
    """uint32 hash -> float64 in [0, 1) with 24-bit resolution."""
    return (np.asarray(h, U32) >> U32(8)).astype(np.float64) * (1.0 / 16777216.0)


def value_noise2(x, z, octave, seed):
    ix = np.floor(x)
    iz = np.floor(z)
    fx = x - ix
    fz = z - iz
    sx = fx * fx * (3.0 - 2.0 * fx)
    sz = fz * fz * (3.0 - 2.0 * fz)
    ix = ix.astype(np.int64)
    iz = iz.astype(np.int64)
    v00 = unit(hash4(ix, iz, octave, seed))
    v10 = unit(hash4(ix + 1, iz, octave, seed))
    v01 = unit(hash4(ix, iz + 1, octave, seed))
    v11 = unit(hash4(ix + 1, iz + 1, octave, seed))
    a = v00 + sx * (v10 - v00)
    b = v01 + sx * (v11 - v01)
    return a + sz * (b - a)


def fbm2(x, z, n, seed, octaves=4, lacunarity=2.0, gain=0.5):
    f = 4.0 / n
    amp = 1.0
    total = np.zeros(np.broadcast(x, z).shape)
    norm = 0.0
    for o in range(octaves):
        total = total + amp * value_noise2(x * f, z * f, o, seed)
        norm += amp
        amp *= gain
        f *= lacunarity
    return total / norm


def heightfield(n, seed):
    """H(x, z) = N/4 + N/4 * fbm(x + .5, z + .5) sampled at column centres
    (insertFunction samples cell centres, octree_cpu.cpp:250-262)."""
    c = np.arange(n, dtype=np.float64) + 0.5
    X, Z = np.meshgrid(c, c, indexing="ij")
    return n / 4.0 + n / 4.0 * fbm2(X, Z, n, seed)


def voxel_materials(x, y, z, seed):
    return (U32(1) + (hash4(x, y, z, seed ^ 0x5EED) & U32(7))).astype(U32)


def _ops(x, y, z, m):
    out = np.empty(len(x), INSERT_DTYPE)
    out["x"] = x
    out["y"] = y
    out["z"] = z
    out["material"] = m
    return out


def terrain_solid(n, seed):
    """Config-0 rule: solid iff y < N/4 + N/4*fbm(x, z); x-, y-, z-major loop
    order of insertFunction."""
    H = heightfield(n, seed)
    hi = np.clip(np.ceil(H), 0, n).astype(np.int64)   # y < H  <=>  y <= ceil(H)-1
    return _columns(n, hi, np.zeros_like(hi), seed)


def terrain_shell(n, seed, thickness=1):
    """Watertight surface shell of the same heightfield: each column keeps the
    voxels from min(top - thickness, lowest 4-neighbour top) up to its top, so
    no ray can see through a slope."""
    H = heightfield(n, seed)
    hi = np.clip(np.ceil(H), 0, n).astype(np.int64)
    lo = hi - thickness
    pad = np.pad(hi, 1, mode="edge")
    nb = np.minimum.reduce([pad[2:, 1:-1], pad[:-2, 1:-1], pad[1:-1, 2:], pad[1:-1, :-2]])
    lo = np.clip(np.minimum(lo, nb), 0, None)
    return _columns(n, hi, lo, seed)


def _columns(n, hi, lo, seed):
    # loop order x, y, z (insertFunction): sort columns' voxels by (x, y, z)
    cnt = (hi - lo).ravel()
    col = np.repeat(np.arange(n * n), cnt)
    start = np.repeat(np.cumsum(cnt) - cnt, cnt)
    y = lo.ravel()[col] + (np.arange(cnt.sum()) - start)
    x = col // n
    z = col % n
    order = np.lexsort((z, y, x))
    x, y, z = x[order], y[order], z[order]
    return _ops(x, y, z, voxel_materials(x, y, z, seed))


def sphere_ops(centre, radius, material, n):
    """OctreeCPU::insertSphere (octree_cpu.cpp:232-248) with its FP32 test
    distance((x+.5, y+.5, z+.5), centre) <= radius."""
    c = np.asarray(centre, np.float32)
    r = np.float32(radius)
    lo = [max(0, int(math.floor(np.float32(c[i] - r)))) for i in range(3)]
    hi = [min(n - 1, int(math.ceil(np.float32(c[i] + r)))) for i in range(3)]
    if any(hi[i] < lo[i] for i in range(3)):
        return _ops(*(np.zeros(0, np.int64),) * 3, np.zeros(0, U32))
    X, Y, Z = np.meshgrid(np.arange(lo[0], hi[0] + 1), np.arange(lo[1], hi[1] + 1),
                          np.arange(lo[2], hi[2] + 1), indexing="ij")
    dx = (X.astype(np.float32) + np.float32(0.5)) - c[0]
    dy = (Y.astype(np.float32) + np.float32(0.5)) - c[1]
    dz = (Z.astype(np.float32) + np.float32(0.5)) - c[2]
    d = np.sqrt((dx * dx + dy * dy) + dz * dz)   # glm::distance: sqrt((x*x + y*y) + z*z)
    m = d <= r
    return _ops(X[m], Y[m], Z[m], np.full(int(m.sum()), material, U32))


def random_spheres(n, seed, count, rmin, rmax, power, H):
    """`count` seeded spheres: centre x, z uniform in [0.1N, 0.9N), resting a
    quarter radius above the terrain; radius rmin + (rmax-rmin)*u^power."""
    i = np.arange(count)
    u = unit(hash4(i, 1, 0xB00, seed))
    cx = 0.1 * n + 0.8 * n * unit(hash4(i, 2, 0xB00, seed))
    cz = 0.1 * n + 0.8 * n * unit(hash4(i, 3, 0xB00, seed))
    r = rmin + (rmax - rmin) * u ** power
    out = []
    for k in range(count):
        ix = min(n - 1, int(cx[k]))
        iz = min(n - 1, int(cz[k]))
        cy = H[ix, iz] + 0.25 * r[k]
        out.append(sphere_ops((cx[k], cy, cz[k]), r[k], 1 + (k & 7), n))
    return np.concatenate(out) if out else np.zeros(0, INSERT_DTYPE)


# Sphere-radius exponent that brings the depth-10 scene to 10.4 M voxels
# (+-1%); chosen once by measuring the count (recorded by the tests).
EXTREME_SPHERE_POWER = 2.28


def scene_small(seed=1):
    """128^3 fbm terrain (depth 7): solid iff y < 32 + 32*fbm(x, z)."""
    return 7, terrain_solid(128, seed)


def scene_extreme(seed=1, power=EXTREME_SPHERE_POWER):
    """1024^3 (depth 10): fbm terrain shell + 64 seeded solid spheres."""
    n = 1024
    H = heightfield(n, seed)
    ops = [terrain_shell(n, seed, 1), random_spheres(n, seed, 64, 8.0, 64.0, power, H)]
    return 10, np.concatenate(ops)


LIGHT_BOXES = (((28, 78, 40), (33, 83, 45)), ((88, 74, 70), (91, 77, 73)), ((52, 98, 56), (70, 99, 74)))


def value_noise3(x, y, z, seed):
    """3D value noise: smoothstep-weighted trilinear blend (x, then y, then z)
    of lattice values unit(hash4(i, j, k, seed)), float64."""
    ix, iy, iz = np.floor(x), np.floor(y), np.floor(z)
    sx, sy, sz = [f * f * (3.0 - 2.0 * f) for f in (x - ix, y - iy, z - iz)]
    ix, iy, iz = ix.astype(np.int64), iy.astype(np.int64), iz.astype(np.int64)
    v = {(a, b, c): unit(hash4(ix + a, iy + b, iz + c, seed)) for a in (0, 1) for b in (0, 1) for c in (0, 1)}
    a00 = v[0, 0, 0] + sx * (v[1, 0, 0] - v[0, 0, 0])
    a10 = v[0, 1, 0] + sx * (v[1, 1, 0] - v[0, 1, 0])
    a01 = v[0, 0, 1] + sx * (v[1, 0, 1] - v[0, 0, 1])
    a11 = v[0, 1, 1] + sx * (v[1, 1, 1] - v[0, 1, 1])
    b0 = a00 + sy * (a10 - a00)
    b1 = a01 + sy * (a11 - a01)
    return b0 + sz * (b1 - b0)


# The divergence configuration: a depth-12 fbm ground whose top CAVE_BAND
# voxels are hollowed by 3D value noise over a CAVE_CELL-voxel lattice above
# CAVE_THRESHOLD (only the surface shell and the cave walls are voxels), plus
# PILLAR_COUNT seeded 1x1 pillars standing on the ground, U[16, 512] tall.
CAVE_BAND, CAVE_CELL, CAVE_THRESHOLD = 48, 32, 0.55
PILLAR_COUNT, PILLAR_MIN, PILLAR_MAX = 2048, 16, 512


def shell_columns(n, seed, thickness=1):
    """(hi, lo) per column [x][z] of terrain_shell: hi = one past the top
    voxel, lo = the first voxel of the watertight shell."""
    H = heightfield(n, seed)
    hi = np.clip(np.ceil(H), 0, n).astype(np.int64)
    pad = np.pad(hi, 1, mode="edge")
    nb = np.minimum.reduce([pad[2:, 1:-1], pad[:-2, 1:-1], pad[1:-1, 2:], pad[1:-1, :-2]])
    lo = np.clip(np.minimum(hi - thickness, nb), 0, None)
    return hi, lo


def caves_numpy(hi, lo, band, cell, threshold, noise_seed, material_seed):
    """The cave-terrain rule of ve_gen_caves restated over whole arrays (small
    grids only): voxels sorted by (x, z, y)."""
    n = hi.shape[0]
    X, Y, Z = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    HI = hi[:, None, :].repeat(n, 1)          # [x][y][z]
    LO = lo[:, None, :].repeat(n, 1)
    v = value_noise3((X + 0.5) / cell, (Y + 0.5) / cell, (Z + 0.5) / cell, noise_seed)
    cave = (Y >= HI - band) & (Y < HI) & (v > threshold)
    solid = (Y < HI) & ~cave
    pad = np.pad(cave, 1)
    touch = (pad[2:, 1:-1, 1:-1] | pad[:-2, 1:-1, 1:-1] | pad[1:-1, 2:, 1:-1] | pad[1:-1, :-2, 1:-1] |
             pad[1:-1, 1:-1, 2:] | pad[1:-1, 1:-1, :-2])
    keep = solid & ((Y >= LO) | touch)
    x, y, z = X[keep], Y[keep], Z[keep]
    order = np.lexsort((y, z, x))
    x, y, z = x[order], y[order], z[order]
    return _ops(x, y, z, voxel_materials(x, y, z, material_seed))


def pillar_ops(n, seed, hi, count=PILLAR_COUNT, hmin=PILLAR_MIN, hmax=PILLAR_MAX):
    """`count` seeded 1x1 pillars: column x, z uniform in [0.05N, 0.95N),
    standing on the ground (from hi up), height uniform in [hmin, hmax]."""
    i = np.arange(count)
    px = (0.05 * n + 0.9 * n * unit(hash4(i, 1, 0x9111, seed))).astype(np.int64)
    pz = (0.05 * n + 0.9 * n * unit(hash4(i, 2, 0x9111, seed))).astype(np.int64)
    h = hmin + (hash4(i, 3, 0x9111, seed) % U32(hmax - hmin + 1)).astype(np.int64)
    base = hi[px, pz]
    top = np.minimum(base + h, n)
    cnt = np.maximum(top - base, 0)
    k = np.repeat(i, cnt)
    y = base[k] + (np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt))
    x, z = px[k], pz[k]
    return _ops(x, y, z, voxel_materials(x, y, z, seed))


def scene_caves(seed=1, depth=12, patch=None, pillars=PILLAR_COUNT, threads=0):
    """The divergence configuration at `depth` (12: 4096^3): fbm ground with
    value-noise caves under its surface and `pillars` pillars.  With `patch`,
    the ground covers only the patch x patch columns at the origin (its
    heightfield that of a patch-sized grid) and the pillars stand on the
    empty floor elsewhere: the same rules on a small voxel count."""
    from . import gen_caves
    n = 1 << depth
    m = n if patch is None else int(patch)
    hi_p, lo_p = shell_columns(m, seed)
    caves = gen_caves(hi_p, lo_p, CAVE_BAND, CAVE_CELL, CAVE_THRESHOLD, seed ^ 0xCA7E, seed, threads)
    hi = hi_p if m == n else np.pad(hi_p, ((0, n - m), (0, n - m)))
    return depth, np.concatenate([caves, pillar_ops(n, seed, hi, count=pillars)])


def scene_lights(seed=1):
    """The depth-7 terrain of scene_small plus three emissive blocks above it
    (a 5^3 cube, a 3^3 cube and an 18x1x18 ceiling panel) using materials 9
    and 10 of palette_lights; the blocks are inserted after the terrain."""
    depth, ops = scene_small(seed)
    extra = []
    for k, (lo, hi) in enumerate(LIGHT_BOXES):
        X, Y, Z = np.meshgrid(*[np.arange(lo[i], hi[i]) for i in range(3)], indexing="ij")
        o = np.empty(X.size, INSERT_DTYPE)
        o["x"], o["y"], o["z"], o["material"] = X.ravel(), Y.ravel(), Z.ravel(), 9 + (k % 2)
        extra.append(o)
    return depth, np.concatenate([ops] + extra)


def scene_fbm(depth, seed):
    """fbm terrain shell at any depth (batched / served resident scenes)."""
    n = 1 << depth
    return depth, terrain_shell(n, seed, 1)


# ---------------------------------------------------------------- materials
MATERIAL_DTYPE = np.dtype([("color", "<f4", 4), ("specular_color", "<f4", 4), ("roughness", "<f4"),
                           ("specular", "<f4"), ("metallic", "<f4"), ("emissive", "<i4"),
                           ("emissive_intensity", "<f4"), ("_pad", "<f4", 3)])
assert MATERIAL_DTYPE.itemsize == 64


def palette(seed=1, count=8):
    """Material pool with entry 0 unused and `count` seeded diffuse albedos
    (roughness 1, metallic 0, non-emissive), std140 64-byte entries
    (src/renderer/material.hpp:17-25)."""
    m = np.zeros(count + 1, MATERIAL_DTYPE)
    for i in range(1, count + 1):
        rgb = 0.2 + 0.7 * unit(hash4(i, np.arange(3), 0xC01, seed))
        m[i]["color"] = [rgb[0], rgb[1], rgb[2], 0.0]
        m[i]["specular_color"] = [1.0, 1.0, 1.0, 1.0]
        m[i]["roughness"] = 1.0
        m[i]["specular"] = 0.0
        m[i]["metallic"] = 0.0
    return m


def palette_spec(seed=1):
    """palette(seed) with three specular entries: 3 = the reference demo's
    metallic_m (albedo 0.8, roughness 0.15, specular 1, metallic 0.7:
    metallic * (1 - roughness) > 0.5, so shade.comp samples GGX VNDF),
    5 = a specular dielectric (specular 0.9, metallic 0.2, roughness 0.5) and
    7 = a rough metal (specular 1, metallic 0.9, roughness 0.5): the last two
    take the cosine sample with a Blinn-Phong lobe of weight specular *
    metallic (internal.glsl:120-149)."""
    m = palette(seed)
    for i, col, rough, spec, metal in ((3, (0.8, 0.8, 0.8), 0.15, 1.0, 0.7), (5, None, 0.5, 0.9, 0.2),
                                          (7, None, 0.5, 1.0, 0.9)):
        if col is not None:
            m[i]["color"] = [col[0], col[1], col[2], 0.0]
        m[i]["roughness"] = rough
        m[i]["specular"] = spec
        m[i]["metallic"] = metal
    return m


def palette_lights(seed=1):
    """palette(seed) plus two emissive entries: 9 warm white (intensity 4),
    10 blue (intensity 6)."""
    m = np.zeros(11, MATERIAL_DTYPE)
    m[:9] = palette(seed)
    for i, (rgb, inten) in ((9, ((1.0, 0.9, 0.7), 4.0)), (10, ((0.4, 0.6, 1.0), 6.0))):
        m[i]["color"] = [rgb[0], rgb[1], rgb[2], 0.0]
        m[i]["specular_color"] = [1.0, 1.0, 1.0, 1.0]
        m[i]["roughness"] = 1.0
        m[i]["emissive"] = 1
        m[i]["emissive_intensity"] = inten
    return m


# ------------------------------------------------------------------ cameras
def look_at(eye, target):
    d = np.asarray(target, np.float64) - np.asarray(eye, np.float64)
    d /= np.linalg.norm(d)
    return np.asarray(eye, np.float32), d.astype(np.float32)


def camera_small():
    """Pinhole at (64, 96, -40) looking at the grid centre, 70 degree FOV."""
    return look_at((64.0, 96.0, -40.0), (64.0, 64.0, 64.0)) + (70.0,)


def orbit_camera(n, frame, frames, seed, radius_frac=0.8, height_frac=0.75, fov=40.0):
    """Seeded orbit around the grid centre: radius radius_frac*N in xz, eye
    height height_frac*N, looking at (N/2, 0.3N, N/2), phase offset from the
    seed.  The 40 degree vertical FOV makes the grid fill a 16:9 frame (about
    73% of the pixels hit geometry on the depth-10 scene)."""
    phase = 2.0 * math.pi * float(unit(hash4(seed, 7, 0x0B, 0)))
    a = phase + 2.0 * math.pi * frame / frames
    c = n / 2.0
    eye = (c + radius_frac * n * math.cos(a), height_frac * n, c + radius_frac * n * math.sin(a))
    return look_at(eye, (c, n * 0.3, c)) + (fov,)
