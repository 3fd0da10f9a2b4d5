// Device-side building blocks: FP32 helpers with the reference's evaluation
// order, hashing, and the octree DDA traversal.
//
// FP policy.  The reference's GLSL runs on the GL driver; the parity target
// is Mesa llvmpipe, whose lowering was measured with probe shaders:
//   - division and sqrt are IEEE correctly rounded (hipcc default for f32);
//   - no multiply-add contraction (this file is compiled with
//     -ffp-contract=off, and no fma is written by hand);
//   - dot(a, b) = a.x*b.x + (a.y*b.y + a.z*b.z);
//   - normalize(v) = v * (1 / sqrt(dot(v, v)));
//   - min/max return the non-NaN operand and the second operand on ties.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ve_internal.h"

namespace ve {

__device__ __forceinline__ float gmin(float a, float b) { return (a < b || __builtin_isnan(b)) ? a : b; }
__device__ __forceinline__ float gmax(float a, float b) { return (a > b || __builtin_isnan(b)) ? a : b; }

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 mul(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 muls(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
__device__ __forceinline__ f3 normalize(f3 v)
{
    float inv = 1.0f / __fsqrt_rn(dot(v, v));
    return muls(v, inv);
}
__device__ __forceinline__ f3 rcp3(f3 d) { return f3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z}; }

// voxelKey + hash (internal.glsl:263-291); the 10-bit packing aliases when a
// coordinate reaches 1024, exactly as in the reference.
__device__ __forceinline__ uint32_t voxel_key(uint32_t x, uint32_t y, uint32_t z)
{
    uint32_t h = x | (y << 10) | (z << 20);
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h & 0x7FFFFFFFu;
}

__device__ __forceinline__ uint32_t locate(uint32_t x, uint32_t y, uint32_t z, uint32_t p2)
{
    return ((uint32_t)((x & p2) != 0u) << 2) | ((uint32_t)((y & p2) != 0u) << 1) | (uint32_t)((z & p2) != 0u);
}

__device__ __forceinline__ bool in_bounds(f3 v, float n)
{
    return 0.0f <= v.x && v.x <= n && 0.0f <= v.y && v.y <= n && 0.0f <= v.z && v.z <= n;
}

struct Hit {
    uint32_t hit, id, material, px, py, pz, size;
};

// Primary-ray direction of pixel (x, y) (generate_primary.comp:44-52).
__device__ __forceinline__ f3 primary_dir(const float *cam, uint32_t x, uint32_t y, uint32_t W, uint32_t H)
{
    float nx = ((float)x + 0.5f) / (float)W * 2.0f - 1.0f;
    float ny = ((float)y + 0.5f) / (float)H * 2.0f - 1.0f;
    f3 d;
    d.x = (cam[4] + nx * cam[8]) - ny * cam[12];
    d.y = (cam[5] + nx * cam[9]) - ny * cam[13];
    d.z = (cam[6] + nx * cam[10]) - ny * cam[14];
    return normalize(d);
}

// RaycastCounted / Raycast (internal.glsl:297-400), step for step.
// `checks` receives q; `reads` accumulates node words fetched (the
// reference's algorithmic traffic, 4 bytes each).
template <bool STATS>
__device__ __forceinline__ Hit raycast(const uint32_t *__restrict__ words, uint32_t n_words, uint32_t depth, f3 o,
                                       f3 d, f3 inv, uint32_t controlchecks, uint32_t &checks, uint32_t &reads)
{
    Hit miss{0, 0, 0, 0, 0, 0, 0};
    const uint32_t len = 1u << depth;
    const float fl = (float)len;
    uint32_t q = 0;
    o = add(o, muls(d, 4.0f));   // internal.glsl:361
    f3 r;
    if (in_bounds(o, fl)) {
        r = o;
    } else {   // intersect(ray, vec3(0), vec3(N)), internal.glsl:245-253
        float t1x = ((0.0f - o.x) + 0.001f) * inv.x, t2x = ((fl - o.x) - 0.001f) * inv.x;
        float t1y = ((0.0f - o.y) + 0.001f) * inv.y, t2y = ((fl - o.y) - 0.001f) * inv.y;
        float t1z = ((0.0f - o.z) + 0.001f) * inv.z, t2z = ((fl - o.z) - 0.001f) * inv.z;
        float te = gmax(gmax(gmin(t1x, t2x), gmin(t1y, t2y)), gmin(t1z, t2z));
        float tx = gmin(gmin(gmax(t1x, t2x), gmax(t1y, t2y)), gmax(t1z, t2z));
        q++;
        if (tx < te || tx < 0.0f) {
            checks = q;
            return miss;
        }
        r = add(muls(d, te), o);
    }
    uint32_t tpx = 0, tpy = 0, tpz = 0, tsize = 0;
    while (in_bounds(r, fl) && q++ <= controlchecks) {
        const uint32_t ux = (uint32_t)r.x, uy = (uint32_t)r.y, uz = (uint32_t)r.z;
        uint32_t offset = 0, dd = 0;
        bool found = false;
        for (; dd < depth; dd++) {
            offset += locate(ux, uy, uz, len >> dd);
            const uint32_t raw = offset < n_words ? words[offset] : 0u;
            if (STATS) reads++;
            if (!(raw & 1u)) {
                tsize = len >> dd;
                found = true;
                break;
            }
            offset = raw >> 4;
        }
        if (!found) {
            offset += locate(ux, uy, uz, 1u);
            const uint32_t raw = offset < n_words ? words[offset] : 0u;
            if (STATS) reads++;
            tsize = 1u;
            const uint32_t mat = (raw >> 1) & 0x7Fu;
            if (mat != 0u) {
                checks = q;
                return Hit{1u, offset, mat, ux, uy, uz, 1u};
            }
        }
        tpx = ux & ~(tsize - 1u);
        tpy = uy & ~(tsize - 1u);
        tpz = uz & ~(tsize - 1u);
        // intersect_inside(ray, pos, pos + size), internal.glsl:255-261
        const float bx = (float)tpx, by = (float)tpy, bz = (float)tpz, fs = (float)tsize;
        float t1x = ((bx - o.x) - 0.001f) * inv.x, t2x = (((bx + fs) - o.x) + 0.001f) * inv.x;
        float t1y = ((by - o.y) - 0.001f) * inv.y, t2y = (((by + fs) - o.y) + 0.001f) * inv.y;
        float t1z = ((bz - o.z) - 0.001f) * inv.z, t2z = (((bz + fs) - o.z) + 0.001f) * inv.z;
        float tx = gmin(gmin(gmax(t1x, t2x), gmax(t1y, t2y)), gmax(t1z, t2z));
        r = add(muls(d, tx), o);
    }
    checks = q;
    return miss;
}

}  // namespace ve
