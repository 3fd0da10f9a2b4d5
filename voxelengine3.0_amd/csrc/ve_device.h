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

// Hardware IEEE min/max (v_min_f32 / v_max_f32): a NaN operand yields the
// other operand like gmin/gmax; they differ only in which zero a +0/-0 tie
// returns.  Used for the traversal's slab tests only, where the result t is
// consumed as d*t + o (a signed zero gives the same sum unless o is also
// zero, and then +0 and -0 both pass inBounds and convert to lattice 0) and
// in comparisons (-0 < 0 is false): the lattice positions, hits and check
// counts are the same.  Shading code keeps gmin/gmax.
__device__ __forceinline__ float tmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ float tmax(float a, float b) { return __builtin_fmaxf(a, b); }

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 mul(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 muls(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
// sqrtf compiles to a correctly rounded square root with this build's flags
// (the IEEE result llvmpipe and the C restatement use); __fsqrt_rn does not
// (about 15% of results differ by an ulp on gfx950), so it is not used.
__device__ __forceinline__ f3 normalize(f3 v)
{
    float inv = 1.0f / sqrtf(dot(v, v));
    return muls(v, inv);
}
__device__ __forceinline__ f3 rcp3(f3 d) { return f3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z}; }

// The per-frame inputs: kernel arguments, or a graph replay's device block.
__device__ __forceinline__ uint32_t fp_frame_idx(const FrameParams &p) { return p.dyn ? p.dyn->frame_idx : p.frame_idx; }
__device__ __forceinline__ uint32_t fp_frame_seed(const FrameParams &p) { return p.dyn ? p.dyn->frame_seed : p.frame_seed; }
__device__ __forceinline__ uint32_t fp_time(const FrameParams &p) { return p.dyn ? p.dyn->time : p.time; }
__device__ __forceinline__ const float *fp_cam(const FrameParams &p) { return p.dyn ? p.dyn->cam : p.cam; }

// A pointer read from device memory (a view record of a launch over many
// views) is generic to the compiler, where the pointer arguments of a kernel
// are known to address device memory: flat instructions then replace the
// global ones, and every wait on a load also waits for LDS traffic.  The
// buffers such records name are device memory, so their pointer slots are
// read as global-address-space pointers.
template <typename T>
__device__ __forceinline__ T *gptr(T *const &slot)
{
    typedef __attribute__((address_space(1))) T *GlobalPtr;
    return (T *)*static_cast<const GlobalPtr *>(static_cast<const void *>(&slot));
}
// Block b of a screen pass (4 tiles of 8x8 pixels per block) -> its tile
// group.  The hardware deals blocks to the 8 XCDs round-robin (XCD b % 8).
// chunk 0: XCD x takes the x-th contiguous eighth of the groups (one band of
// rows: its L2 holds that band's nodes); chunk n: XCD x takes runs of n
// consecutive groups in turn (runs x, x + 8, ...), so every XCD gets work
// from every band.  The grid is padded to a multiple of 8 * max(chunk, 1)
// blocks, which makes the map a bijection.
__device__ __forceinline__ uint32_t xcd_group(uint32_t b, uint32_t nb_padded, uint32_t chunk)
{
    const uint32_t x = b & 7u, k = b >> 3;
    if (chunk == 0u) return x * (nb_padded >> 3) + k;
    return ((k / chunk) * 8u + x) * chunk + k % chunk;
}

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

// locate() for p2 = 1 << bit: the three coordinate bits, extracted directly.
__device__ __forceinline__ uint32_t locate_bit(uint32_t x, uint32_t y, uint32_t z, uint32_t bit)
{
    return (__builtin_amdgcn_ubfe(x, bit, 1u) << 2) | (__builtin_amdgcn_ubfe(y, bit, 1u) << 1) |
           __builtin_amdgcn_ubfe(z, bit, 1u);
}

// Node word `offset` of a validated array: the upload and every update check
// that each internal node's child block lies inside the array, so a descent
// never computes an offset past it (the reference's `offset < size ? w : 0`
// guard is then always true); the clamp only keeps the load in bounds.
__device__ __forceinline__ uint32_t node_word(const uint32_t *__restrict__ words, uint32_t n_words, uint32_t offset)
{
    return words[min(offset, n_words - 1u)];
}

__device__ __forceinline__ bool in_bounds(f3 v, float n)
{
    return (0.0f <= v.x) & (v.x <= n) & (0.0f <= v.y) & (v.y <= n) & (0.0f <= v.z) & (v.z <= n);
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

// texture(skybox, d).rgb under the reference skybox's sampler state
// (skybox.cpp:113-117: GL_LINEAR, CLAMP_TO_EDGE, level 0, no seamless
// filtering), evaluated as the GL driver the reference runs on does:
//   face: major axis of the GL cube-map face table, ties toward z, then y;
//   s, t = (sc / |ma| + 1) / 2, (tc / |ma| + 1) / 2;
//   u = s * edge - 1/2 in 24.8 fixed point, rounded to nearest even: integer
//   part = left texel, fraction = 8-bit weight; texels clamped to the face;
//   bytes blended as a + (((b - a) w + 128) >> 8), along s, then t;
//   result c * (1/255).
__device__ __forceinline__ void cube_axis(float st, uint32_t edge, uint32_t &i0, uint32_t &i1, int &w)
{
    const float u = st * (float)(edge * 256u) - 128.0f;
    const int x = (int)__builtin_rintf(u);
    const int lo = x >> 8, top = (int)edge - 1;
    w = x & 255;
    i0 = (uint32_t)min(max(lo, 0), top);
    i1 = (uint32_t)min(max(lo + 1, 0), top);
}

__device__ __forceinline__ int lerp8(int a, int b, int w) { return a + (((b - a) * w + 128) >> 8); }

__device__ __noinline__ f3 sky_cube(const uint32_t *__restrict__ tex, uint32_t edge, float x, float y, float z)
{
    const float ax = fabsf(x), ay = fabsf(y), az = fabsf(z);
    uint32_t face;
    float sc, tc, ma;
    if (az >= ax && az >= ay) {
        face = z >= 0.0f ? 4u : 5u;
        sc = z >= 0.0f ? x : -x;
        tc = -y;
        ma = az;
    } else if (ay >= ax) {
        face = y >= 0.0f ? 2u : 3u;
        sc = x;
        tc = y >= 0.0f ? z : -z;
        ma = ay;
    } else {
        face = x >= 0.0f ? 0u : 1u;
        sc = x >= 0.0f ? -z : z;
        tc = -y;
        ma = ax;
    }
    const float s = (sc / ma + 1.0f) * 0.5f, t = (tc / ma + 1.0f) * 0.5f;
    uint32_t i0, i1, j0, j1;
    int wa, wb;
    cube_axis(s, edge, i0, i1, wa);
    cube_axis(t, edge, j0, j1, wb);
    const uint32_t *f = tex + (size_t)face * edge * edge;
    const uint32_t t00 = f[j0 * edge + i0], t10 = f[j0 * edge + i1], t01 = f[j1 * edge + i0], t11 = f[j1 * edge + i1];
    float c[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int sh = 8 * k;
        const int top = lerp8((int)((t00 >> sh) & 255u), (int)((t10 >> sh) & 255u), wa);
        const int bot = lerp8((int)((t01 >> sh) & 255u), (int)((t11 >> sh) & 255u), wa);
        c[k] = (float)lerp8(top, bot, wb) * (1.0f / 255.0f);
    }
    return f3{c[0], c[1], c[2]};
}

// The sky seen along d: the cube map, or the constant colour.
__device__ __forceinline__ f3 sky_color(const FrameParams &p, float x, float y, float z)
{
    if (p.sky_edge == 0u) return f3{p.sky[0], p.sky[1], p.sky[2]};
    return sky_cube(p.sky_tex, p.sky_edge, x, y, z);
}

// Ray state after the setup of internal.glsl:352-366.
struct DDA {
    f3 o, d, inv, r;
    uint32_t q;
    uint32_t mx, my, mz;   // all ones where the axis' inverse direction is positive (far face at pos + size)
    float ex, ey, ez;      // the far face's nudge: +0.001 there, -0.001 on the other faces
    bool exact;            // an infinite inverse component: exits take the reference's full min/max form
};

enum : int { DDA_CONTINUE = 0, DDA_MISS = 1, DDA_HIT = 2 };

// Ray setup up to the first loop test (internal.glsl:352-366).
__device__ __forceinline__ int dda_start(DDA &s, f3 o, f3 d, f3 inv, uint32_t depth, uint32_t &checks)
{
    const float fl = (float)(1u << depth);
    s.d = d;
    s.inv = inv;
    s.q = 0;
    s.mx = inv.x > 0.0f ? 0xFFFFFFFFu : 0u;
    s.my = inv.y > 0.0f ? 0xFFFFFFFFu : 0u;
    s.mz = inv.z > 0.0f ? 0xFFFFFFFFu : 0u;
    s.ex = inv.x > 0.0f ? 0.001f : -0.001f;
    s.ey = inv.y > 0.0f ? 0.001f : -0.001f;
    s.ez = inv.z > 0.0f ? 0.001f : -0.001f;
    s.exact = __builtin_isinf(inv.x) || __builtin_isinf(inv.y) || __builtin_isinf(inv.z);
    s.o = add(o, muls(d, 4.0f));   // internal.glsl:361
    if (in_bounds(s.o, fl)) {
        s.r = s.o;
        return DDA_CONTINUE;
    }
    // intersect(ray, vec3(0), vec3(N)), internal.glsl:245-253, with the
    // association llvmpipe's compiler gives it: t1 = (0.001 - o) * inv,
    // t2 = ((N - 0.001) - o) * inv
    const f3 oo = s.o;
    const float fe = fl - 0.001f;
    const float t1x = (0.001f - oo.x) * inv.x, t2x = (fe - oo.x) * inv.x;
    const float t1y = (0.001f - oo.y) * inv.y, t2y = (fe - oo.y) * inv.y;
    const float t1z = (0.001f - oo.z) * inv.z, t2z = (fe - oo.z) * inv.z;
    const float te = tmax(tmax(tmin(t1x, t2x), tmin(t1y, t2y)), tmin(t1z, t2z));
    const float tx = tmin(tmin(tmax(t1x, t2x), tmax(t1y, t2y)), tmax(t1z, t2z));
    s.q = 1;
    if (tx < te || tx < 0.0f) {
        checks = s.q;
        return DDA_MISS;
    }
    s.r = add(muls(d, te), oo);
    return DDA_CONTINUE;
}

// RaycastCounted / Raycast (internal.glsl:297-400): the while loop of
// internal.glsl:370-398 after dda_start.  `checks` receives q.
//
// Every DDA step of the reference descends from the root.  The words read at
// depth dd depend only on coordinate bits D..D-dd (locate() uses bit D-dd and
// the child offsets come from words on the same path), so when the new
// lattice position differs from the previous one first in bit hb, depths
// 0..D-hb-1 read exactly the words they read last step.  The descent resumes
// at depth D-hb from the `next` pointer cached for depth D-hb-1 (in LDS, one
// column per thread: entry dd at cache[dd * stride]), or reuses the last
// outcome outright when it ended above that depth.  A descent that would
// resume at or above depth jl (the step left its cell of the depth-jl grid)
// takes the outcome through depth jl from the scene's jump table in one load
// and continues below it (positions on the grid's far faces, whose
// coordinate bit D takes part in the reference's locate(), descend from the
// root instead).  Results are identical to the full descent; only the
// repeated loads are skipped.  `reads` still
// counts the words the reference's full descent reads (its algorithmic
// traffic, 4 bytes each).  The loop carries only the traversal state; the hit
// record is formed at the exit.
template <bool STATS>
__device__ __forceinline__ Hit raycast(const uint32_t *__restrict__ words, const uint32_t *__restrict__ jump,
                                       uint32_t jl, uint32_t n_words, uint32_t depth, f3 o, f3 d, f3 inv,
                                       uint32_t controlchecks, uint32_t &checks, uint32_t &reads, uint32_t *cache,
                                       uint32_t stride)
{
    DDA s;
    if (dda_start(s, o, d, inv, depth, checks) != DDA_CONTINUE) return Hit{0, 0, 0, 0, 0, 0, 0};
    const uint32_t len = 1u << depth;
    const float fl = (float)len;
    const f3 oo = s.o;
    f3 r = s.r;
    uint32_t q = s.q, px = 0xFFFFFFFFu, py = 0xFFFFFFFFu, pz = 0xFFFFFFFFu, pstop = 0;
    for (;;) {
        // inBounds(r) && q++ <= controlchecks, without the short-circuit branch
        const bool inb = in_bounds(r, fl);
        const bool go = inb & (q <= controlchecks);
        q += inb ? 1u : 0u;
        if (!go) break;
        const uint32_t ux = (uint32_t)r.x, uy = (uint32_t)r.y, uz = (uint32_t)r.z;
        const uint32_t diff = (ux ^ px) | (uy ^ py) | (uz ^ pz);
        // first depth whose word may differ: D - (highest differing bit); the
        // first step (no previous position) starts at the root
        const uint32_t hb = 31u - (uint32_t)__clz(diff);
        const uint32_t resume = diff == 0u ? 0xFFFFFFFFu : (hb >= depth ? 0u : depth - hb);
        uint32_t dd;
        if (resume > pstop) {
            dd = pstop;   // same words down to the one that ended the last descent
        } else {
            uint32_t offset, from = resume;
            if (resume > jl) {
                offset = cache[(resume - 1u) * stride];
            } else if (((ux | uy | uz) >> depth) == 0u) {
                const uint32_t sh = depth - jl;
                const uint32_t e = jump[((ux >> sh) << (2u * jl)) | ((uy >> sh) << jl) | (uz >> sh)];
                offset = e;
                from = jl + 1u;
                if (e & JUMP_ENDED) from = depth + 1u;   // ended at depth (e & 0xFF): no loop, no leaf
                else cache[jl * stride] = e;
            } else {
                offset = 0u;
                from = 0u;
            }
            for (dd = from; dd < depth; dd++) {
                offset += locate_bit(ux, uy, uz, depth - dd);
                const uint32_t raw = node_word(words, n_words, offset);
                if (!(raw & 1u)) break;
                offset = raw >> 4;
                cache[dd * stride] = offset;
            }
            if (from == depth + 1u) dd = offset & 0xFFu;
            if (dd == depth) {   // bottom level: the leaf slot
                offset += locate_bit(ux, uy, uz, 0u);
                const uint32_t raw = node_word(words, n_words, offset);
                const uint32_t mat = (raw >> 1) & 0x7Fu;
                if (mat != 0u) {
                    if (STATS) reads += depth + 1u;
                    checks = q;
                    return Hit{1u, offset, mat, ux, uy, uz, 1u};
                }
            }
            pstop = dd;
        }
        if (STATS) reads += dd + 1u;   // the reference re-reads depths 0..dd every step
        px = ux;
        py = uy;
        pz = uz;
        const uint32_t tsize = len >> dd;
        const uint32_t tpx = ux & ~(tsize - 1u), tpy = uy & ~(tsize - 1u), tpz = uz & ~(tsize - 1u);
        // intersect_inside(ray, pos, pos + size), internal.glsl:255-261, with
        // the association llvmpipe's compiler gives it: tx = min over axes of
        // max(t1, t2), t1 = ((pos - 0.001) - o) * inv, t2 = (((pos + 0.001) +
        // size) - o) * inv.  The second numerator exceeds the first, so for a
        // finite inv the max is t2 where inv > 0 and t1 where inv < 0: only the
        // far face is evaluated ((pos -+ 0.001) + 0 is pos -+ 0.001 exactly).
        // An infinite component (zero direction) can make a numerator zero
        // (pos - 0.001 rounds to pos from 32768 up) and a product NaN, so
        // such rays take the full form.
        const float pxf = (float)tpx, pyf = (float)tpy, pzf = (float)tpz;
        float tx;
        if (!s.exact) {
            const float ax = (((pxf + s.ex) + (float)(tsize & s.mx)) - oo.x) * inv.x,
                        ay = (((pyf + s.ey) + (float)(tsize & s.my)) - oo.y) * inv.y,
                        az = (((pzf + s.ez) + (float)(tsize & s.mz)) - oo.z) * inv.z;
            tx = tmin(tmin(ax, ay), az);
        } else {
            const float szf = (float)tsize;
            const float ax = tmax(((pxf - 0.001f) - oo.x) * inv.x, (((pxf + 0.001f) + szf) - oo.x) * inv.x),
                        ay = tmax(((pyf - 0.001f) - oo.y) * inv.y, (((pyf + 0.001f) + szf) - oo.y) * inv.y),
                        az = tmax(((pzf - 0.001f) - oo.z) * inv.z, (((pzf + 0.001f) + szf) - oo.z) * inv.z);
            tx = tmin(tmin(ax, ay), az);
        }
        r = add(muls(d, tx), oo);
    }
    checks = q;
    return Hit{0, 0, 0, 0, 0, 0, 0};
}

}  // namespace ve

namespace ve {

// ---- GLSL built-ins with llvmpipe's composition -------------------------------
// log = log2 * ln2, exp = exp2(x * log2e), pow = exp2(log2(x) * y); the log2 /
// exp2 / sin / cos cores are evaluated in double and rounded once, which is
// the same definition the C oracle uses.
// The double-precision cores are kept out of line: inlined at every call
// site they made the bounce kernels larger than the instruction cache.
__device__ __noinline__ float m_log2(float x) { return (float)log2((double)x); }
__device__ __noinline__ float m_exp2(float x) { return (float)exp2((double)x); }
__device__ __forceinline__ float m_log(float x) { return m_log2(x) * 0.6931471805599453f; }
__device__ __forceinline__ float m_exp(float x) { return m_exp2(x * 1.4426950408889634f); }
__device__ __forceinline__ float m_pow(float x, float y) { return m_exp2(m_log2(x) * y); }
__device__ __noinline__ float m_cos(float x) { return (float)cos((double)x); }
__device__ __noinline__ float m_sin(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
__device__ __forceinline__ float gsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
__device__ __forceinline__ float length3(f3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 divs(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ f3 vmin3(f3 a, f3 b) { return f3{gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b)
{
    return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}

// ---- Halton sampler (internal.glsl:57-89) -------------------------------------
// Radical inverse of i in base B, digit by digit as the reference computes it.
// B is a compile-time constant, so i / B is a multiply-high, not a division.
template <uint32_t B>
__device__ __forceinline__ float halton_b(uint32_t i)
{
    float f = 1.0f, r = 0.0f;
    for (int k = 0; k < 24 && i > 0u; k++) {
        f /= (float)B;
        const uint32_t q = i / B;
        r += f * (float)(i - q * B);
        i = q;
    }
    return r;
}

// The base of dimension `dim` is the (dim mod 16)-th prime (internal.glsl:
// 57-60).  Out of line: one copy of the sixteen loops serves every draw.
__device__ __noinline__ float halton_dim(uint32_t i, uint32_t dim)
{
    switch (dim & 15u) {
    case 0: return halton_b<2>(i);    case 1: return halton_b<3>(i);    case 2: return halton_b<5>(i);
    case 3: return halton_b<7>(i);    case 4: return halton_b<11>(i);   case 5: return halton_b<13>(i);
    case 6: return halton_b<17>(i);   case 7: return halton_b<19>(i);   case 8: return halton_b<23>(i);
    case 9: return halton_b<29>(i);   case 10: return halton_b<31>(i);  case 11: return halton_b<37>(i);
    case 12: return halton_b<41>(i);  case 13: return halton_b<43>(i);  case 14: return halton_b<47>(i);
    default: return halton_b<53>(i);
    }
}

__device__ __forceinline__ float rand_lds(uint32_t &state)
{
    const uint32_t dim = state & 0xFFu;
    const uint32_t sample_idx = state >> 8;
    state = (sample_idx << 8) | ((dim + 1u) & 0xFFu);
    const uint32_t effective = sample_idx ^ ((dim >> 4) * 0x9E3779B9u);
    return halton_dim(effective, dim);
}

__device__ __forceinline__ float rand_normal(uint32_t &state)   // internal.glsl:101-105
{
    const float theta = (2.0f * 3.1415926f) * rand_lds(state);
    const float rho = sqrtf(-2.0f * m_log(rand_lds(state)));
    return rho * m_cos(theta);
}

__device__ __forceinline__ f3 random_direction(uint32_t &state)   // internal.glsl:107-112
{
    const float x = rand_normal(state);
    const float y = rand_normal(state);
    const float z = rand_normal(state);
    return normalize(f3{x, y, z});
}

// EvaluateBRDF (internal.glsl:120-149)
__device__ __forceinline__ f3 evaluate_brdf(f3 Vd, f3 L, f3 N, f3 albedo, float roughness, float metallic,
                                            float spec_int)
{
    const float NdotL = gmax(dot(N, L), 0.001f);
    if (NdotL < 0.001f) return f3{0, 0, 0};
    const float rc = gmax(roughness, 0.01f);
    const float spec_exp = 2.0f / (rc * rc);
    const f3 H = normalize(sub(L, Vd));
    const float NdotH = gmax(dot(N, H), 0.001f);
    const float spec_norm = (spec_exp + 2.0f) * 0.15915494f;
    const float specular = spec_norm * m_pow(NdotH, spec_exp);
    const float kS = spec_int * metallic;
    const float kD = 1.0f - kS * 0.5f;
    // albedo * kD * 0.31831 as llvmpipe's compiler reassociates it: the
    // constant factor moves next to the first operand
    const f3 diffuse = muls(muls(albedo, 0.31831f), kD);
    const float sp = specular * kS;
    return muls(add(diffuse, f3{sp, sp, sp}), NdotL);
}

__device__ __forceinline__ float smith_g1(float NdotX, float a)
{
    const float a2 = a * a;
    return 2.0f * NdotX / (NdotX + sqrtf(a2 + (1.0f - a2) * NdotX * NdotX));
}

// sampleGGXVNDF (internal.glsl:182-219)
// (out of line, with its two random numbers drawn by the caller, so that no
// reference escapes into the call)
__device__ __noinline__ f3 sample_ggx_vndf_at(f3 N, f3 Vv, float roughness, float r1, float r2)
{
    const float a = gmax(roughness * roughness, 0.001f);
    const f3 up = fabsf(N.z) < 0.999f ? f3{0, 0, 1} : f3{1, 0, 0};
    const f3 T = normalize(cross3(up, N));
    const f3 B = cross3(N, T);
    const f3 Vl = f3{dot(Vv, T), dot(Vv, B), dot(Vv, N)};
    const f3 Vh = normalize(f3{a * Vl.x, a * Vl.y, Vl.z});
    const float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    const f3 T1 = lensq > 0.0f ? muls(f3{-Vh.y, Vh.x, 0.0f}, 1.0f / sqrtf(lensq)) : f3{1, 0, 0};
    const f3 T2 = cross3(Vh, T1);
    const float r = sqrtf(r1);
    const float phi = 6.2831853f * r2;
    const float t1 = r * m_cos(phi);
    float t2 = r * m_sin(phi);
    const float sc = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - sc) * sqrtf(gmax(1.0f - t1 * t1, 0.0f)) + sc * t2;
    const f3 Nh = add(add(muls(T1, t1), muls(T2, t2)), muls(Vh, sqrtf(gmax(0.0f, 1.0f - t1 * t1 - t2 * t2))));
    const f3 Hl = normalize(f3{a * Nh.x, a * Nh.y, gmax(0.0f, Nh.z)});
    return add(add(muls(T, Hl.x), muls(B, Hl.y)), muls(N, Hl.z));
}

__device__ __forceinline__ f3 sample_ggx_vndf(f3 N, f3 Vv, float roughness, uint32_t &state)
{
    const float r1 = rand_lds(state);   // internal.glsl draws r1 then r2
    const float r2 = rand_lds(state);
    return sample_ggx_vndf_at(N, Vv, roughness, r1, r2);
}

__device__ __noinline__ f3 ggx_throughput(f3 N, f3 Vv, f3 L, f3 H, float roughness, f3 F0)
{
    const float NdotL = gmax(dot(N, L), 0.001f);
    const float VdotH = gmax(dot(Vv, H), 0.001f);
    const float a = roughness * roughness;
    const float p5 = m_pow(1.0f - VdotH, 5.0f);
    const f3 F = add(F0, muls(sub(f3{1, 1, 1}, F0), p5));
    return muls(F, smith_g1(NdotL, a));
}

__device__ __forceinline__ f3 reflect3(f3 I, f3 N) { return sub(I, muls(N, 2.0f * dot(N, I))); }

__device__ __forceinline__ f3 unpack_normal(uint32_t p)
{
    return f3{(float)((p >> 22) & 0x3FFu) / 511.0f - 1.0f, (float)((p >> 12) & 0x3FFu) / 511.0f - 1.0f,
              (float)((p >> 2) & 0x3FFu) / 511.0f - 1.0f};
}

__device__ __forceinline__ uint32_t claim_hash(uint32_t v)   // schedule.comp:36-43
{
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    return v;
}

__device__ __forceinline__ uint32_t bit_reverse(uint32_t x) { return __brev(x); }

}  // namespace ve
