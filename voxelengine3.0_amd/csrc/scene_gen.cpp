// Synthetic cave terrain: the generator of the depth-12 divergence scene (a
// heightfield ground with 3D value-noise caves in a band under its surface).
// A data source for tests and benchmarks, not a reference pass: the scene's
// voxels go through ve_octree_build like any insert list.
//
// The rule, for an n x n heightfield given per column as hi (one past the top
// ground voxel) and lo (the first voxel of the watertight surface shell):
//   cave(x, y, z)  = hi - band <= y < hi and v((x+.5)/cell, (y+.5)/cell, (z+.5)/cell) > threshold
//   solid(x, y, z) = 0 <= y < hi and not cave
//   emitted        = solid and (y >= lo or a 6-neighbour inside the grid is a cave)
// where v is value noise: smoothstep-weighted trilinear blend of lattice
// values unit(hash4(i, j, k, seed)) (the 3D form of the package's 2D value
// noise, in double precision, blended along x, then y, then z).  Material of
// every voxel: 1 + (hash4(x, y, z, material_seed ^ 0x5EED) & 7).
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <thread>
#include <vector>

#include "../../include/ve_svo.h"
#include "ve_internal.h"

namespace {

inline uint32_t mix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}

inline uint32_t hash4(uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    uint32_t h = mix32(a * 0x9E3779B1u + 0x632BE5ABu);
    h = mix32(h ^ (b * 0x85EBCA77u));
    h = mix32(h ^ (c * 0xC2B2AE3Du));
    h = mix32(h ^ (d * 0x27D4EB2Fu));
    return h;
}

inline double unit(uint32_t h) { return (double)(h >> 8) * (1.0 / 16777216.0); }

inline double smooth(double f) { return f * f * (3.0 - 2.0 * f); }

struct Params {
    uint32_t n, band, cell;
    double threshold;
    uint32_t seed, mseed;
    const int32_t *hi, *lo;
};

// Cave bits of column (x, z): bit k for y = hi - band + k, k < band.
uint64_t column_mask(const Params &P, uint32_t x, uint32_t z)
{
    const int64_t hi = P.hi[(size_t)x * P.n + z];
    const int64_t y0 = std::max<int64_t>(0, hi - (int64_t)P.band);
    if (hi <= 0) return 0;
    const double c = (double)P.cell;
    const double px = ((double)x + 0.5) / c, pz = ((double)z + 0.5) / c;
    const double fx0 = std::floor(px), fz0 = std::floor(pz);
    const double sx = smooth(px - fx0), sz = smooth(pz - fz0);
    const int64_t ix = (int64_t)fx0, iz = (int64_t)fz0;
    uint64_t mask = 0;
    int64_t cur_iy = INT64_MIN;
    double v[2][2][2] = {};   // [dx][dy][dz]
    for (int64_t y = y0; y < hi; y++) {
        const double py = ((double)y + 0.5) / c;
        const double fy0 = std::floor(py);
        const int64_t iy = (int64_t)fy0;
        if (iy != cur_iy) {
            for (int dx = 0; dx < 2; dx++)
                for (int dy = 0; dy < 2; dy++)
                    for (int dz = 0; dz < 2; dz++)
                        v[dx][dy][dz] = unit(hash4((uint32_t)(ix + dx), (uint32_t)(iy + dy), (uint32_t)(iz + dz), P.seed));
            cur_iy = iy;
        }
        const double sy = smooth(py - fy0);
        const double a00 = v[0][0][0] + sx * (v[1][0][0] - v[0][0][0]);
        const double a10 = v[0][1][0] + sx * (v[1][1][0] - v[0][1][0]);
        const double a01 = v[0][0][1] + sx * (v[1][0][1] - v[0][0][1]);
        const double a11 = v[0][1][1] + sx * (v[1][1][1] - v[0][1][1]);
        const double b0 = a00 + sy * (a10 - a00);
        const double b1 = a01 + sy * (a11 - a01);
        const double val = b0 + sz * (b1 - b0);
        if (val > P.threshold) mask |= 1ull << (uint32_t)(y - (hi - (int64_t)P.band));
    }
    return mask;
}

struct Masks {
    const Params &P;
    const std::vector<uint64_t> &m;
    bool cave(int64_t x, int64_t y, int64_t z) const
    {
        if (x < 0 || z < 0 || x >= (int64_t)P.n || z >= (int64_t)P.n) return false;
        const size_t col = (size_t)x * P.n + (size_t)z;
        const int64_t hi = P.hi[col], base = hi - (int64_t)P.band;
        if (y < base || y >= hi || y < 0) return false;
        return (m[col] >> (uint32_t)(y - base)) & 1u;
    }
};

// Visits the emitted voxels of slab x in (z, y) order.
template <typename F>
void slab(const Params &P, const Masks &M, uint32_t x, F &&emit)
{
    for (uint32_t z = 0; z < P.n; z++) {
        const size_t col = (size_t)x * P.n + z;
        const int64_t hi = P.hi[col], lo = P.lo[col];
        // lowest voxel that can touch a cave: under this column's band, or
        // beside the band of a neighbour column
        int64_t from = std::min<int64_t>(lo, hi - (int64_t)P.band - 1);
        if (x > 0) from = std::min<int64_t>(from, P.hi[col - P.n] - (int64_t)P.band);
        if (x + 1 < P.n) from = std::min<int64_t>(from, P.hi[col + P.n] - (int64_t)P.band);
        if (z > 0) from = std::min<int64_t>(from, P.hi[col - 1] - (int64_t)P.band);
        if (z + 1 < P.n) from = std::min<int64_t>(from, P.hi[col + 1] - (int64_t)P.band);
        from = std::max<int64_t>(0, from);
        for (int64_t y = from; y < hi; y++) {
            if (M.cave(x, y, z)) continue;
            const bool keep = y >= lo || M.cave(x, y + 1, z) || M.cave(x, y - 1, z) || M.cave((int64_t)x + 1, y, z) ||
                              M.cave((int64_t)x - 1, y, z) || M.cave(x, y, (int64_t)z + 1) || M.cave(x, y, (int64_t)z - 1);
            if (keep) emit((uint32_t)y, z);
        }
    }
}

template <typename F>
void parallel_for(uint32_t n, uint32_t threads, F &&f)
{
    std::vector<std::thread> pool;
    const uint32_t t = std::max(1u, std::min(threads, n));
    for (uint32_t k = 0; k < t; k++)
        pool.emplace_back([&, k] {
            for (uint32_t i = k; i < n; i += t) f(i);
        });
    for (auto &th : pool) th.join();
}

}  // namespace

extern "C" int ve_gen_caves(uint32_t n, const int32_t *hi, const int32_t *lo, uint32_t band, uint32_t cell,
                            double threshold, uint32_t noise_seed, uint32_t material_seed, uint32_t threads,
                            ve_insert *out, size_t capacity, size_t *count)
{
    ve_clear_global_error();
    if (!hi || !lo || !count) return ve_global_error(VE_EINVAL, "null argument");
    if (n == 0 || band == 0 || band > 63 || cell == 0)
        return ve_global_error(VE_EINVAL, "n, band and cell must be positive, band at most 63");
    for (size_t i = 0; i < (size_t)n * n; i++)
        if (hi[i] < 0 || (uint32_t)hi[i] > n || lo[i] < 0 || lo[i] > hi[i]) {
            char msg[96];
            std::snprintf(msg, sizeof msg, "column %zu: need 0 <= lo <= hi <= n", i);
            return ve_global_error(VE_EINVAL, msg);
        }
    if (threads == 0) threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const Params P{n, band, cell, threshold, noise_seed, material_seed, hi, lo};
    std::vector<uint64_t> masks((size_t)n * n);
    parallel_for(n, threads, [&](uint32_t x) {
        for (uint32_t z = 0; z < n; z++) masks[(size_t)x * n + z] = column_mask(P, x, z);
    });
    const Masks M{P, masks};
    std::vector<size_t> per_slab(n, 0);
    parallel_for(n, threads, [&](uint32_t x) {
        size_t c = 0;
        slab(P, M, x, [&](uint32_t, uint32_t) { c++; });
        per_slab[x] = c;
    });
    size_t total = 0;
    std::vector<size_t> start(n);
    for (uint32_t x = 0; x < n; x++) {
        start[x] = total;
        total += per_slab[x];
    }
    *count = total;
    if (!out || capacity < total) return VE_OK;   // count only
    parallel_for(n, threads, [&](uint32_t x) {
        ve_insert *o = out + start[x];
        slab(P, M, x, [&](uint32_t y, uint32_t z) {
            *o++ = ve_insert{x, y, z, 1u + (hash4(x, y, z, material_seed ^ 0x5EEDu) & 7u)};
        });
    });
    return VE_OK;
}
