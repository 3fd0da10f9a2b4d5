// gfx950 kernels of the SVO ray pipeline.
//
// Pass map (reference file:line -> kernel):
//   generate_primary.comp:38-85  -> k_primary
//
// Launch shape.  Screen passes run one 64-lane wavefront per 8x8 pixel tile
// (a tile's rays are coherent, so the lanes walk the same octree nodes), four
// wavefronts per 256-thread block.  Blocks are handed out round-robin to the
// 8 XCDs by the hardware; the block -> tile map (xcd_group) gives each XCD
// runs of consecutive tiles, so neighbouring rays share an L2, while every
// XCD gets runs from every part of the image (even load: sky and terrain).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>

#include "ve_device.h"
#include "ve_launch.h"

namespace ve {

constexpr int TILE = 8;
constexpr int WAVES_PER_BLOCK = 4;
constexpr int NUM_XCD = 8;

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// The tile/wave mapping exits whole waves early, so the path cache below is
// only ever touched by live threads; no barrier is needed (each thread owns
// its column).
template <bool STATS>
__global__ __launch_bounds__(256) void k_primary(const uint32_t *__restrict__ words, FrameParams p,
                                                 uint4 *__restrict__ gbuf, uint32_t *__restrict__ gpos,
                                                 unsigned long long *__restrict__ stats, uint32_t *__restrict__ counters,
                                                 unsigned long long *__restrict__ totals, uint2 *__restrict__ ghit)
{
    if (!p.count_primary_in_reset && blockIdx.x == 0 && threadIdx.x == 0) {   // one primary ray per pixel
        counters[VE_CNT_PRIMARY_RAYS] = p.W * p.H;
        totals[0] += (unsigned long long)p.W * p.H;
    }
    const uint32_t tiles_x = (p.W + TILE - 1) / TILE;
    const uint32_t tiles_y = (p.H + TILE - 1) / TILE;
    const uint32_t group = xcd_group(blockIdx.x, gridDim.x, p.tile_chunk);
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t tile = group * WAVES_PER_BLOCK + wave;
    if (tile >= tiles_x * tiles_y) return;
    const uint32_t x = (tile % tiles_x) * TILE + (lane & 7u);
    const uint32_t y = (tile / tiles_x) * TILE + (lane >> 3);
    __shared__ uint32_t path_cache[VE_MAX_DEPTH * 256];
    uint32_t checks = 0, reads = 0;
    if (x < p.W && y < p.H) {
        const f3 dir = primary_dir(fp_cam(p), x, y, p.W, p.H);
        const f3 org = mk(fp_cam(p)[0], fp_cam(p)[1], fp_cam(p)[2]);
        const Hit h = raycast<STATS>(words, p.jump, p.jump_level, p.n_words, p.depth, org, dir, rcp3(dir), p.primary_cc, checks, reads,
                                     path_cache + threadIdx.x, 256u);
        const size_t pix = (size_t)y * p.W + x;
        if (h.hit) {
            const uint32_t bits = p.depth - 1u;
            gbuf[pix] = make_uint4(checks, h.material, 0u, voxel_key(h.px, h.py, h.pz) + 1u);
            gpos[pix] = (h.px >> 1) | ((h.py >> 1) << bits) | ((h.pz >> 1) << (2u * bits));
            if (p.reuse_primary_hits) ghit[pix] = make_uint2(h.px | (h.py << 16), h.pz);
        } else {
            gbuf[pix] = make_uint4(checks, 0u, 0u, 0u);
            gpos[pix] = 0u;
        }
    }
    if (STATS) {
        unsigned long long r = wave_sum(reads), c = wave_sum(checks);
        if (lane == 0) {
            atomicAdd(&stats[0], r);
            atomicAdd(&stats[2], c);
        }
    }
}

static uint32_t padded_blocks(uint32_t work_groups, uint32_t chunk)
{
    uint32_t nb = work_groups == 0 ? 1 : work_groups;
    const uint32_t unit = NUM_XCD * (chunk ? chunk : 1u);
    return (nb + unit - 1) / unit * unit;
}

void launch_primary(hipStream_t s, const uint32_t *words, const FrameParams &p, DeviceBuffers &b, bool stats)
{
    const uint32_t tiles = ((p.W + TILE - 1) / TILE) * ((p.H + TILE - 1) / TILE);
    const uint32_t nb = padded_blocks((tiles + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK, p.tile_chunk);
    if (stats)
        launch(k_primary<true>, dim3(nb), dim3(256), 0, s, words, p, b.gbuf, b.gpos, b.node_stats,
                           b.counters, b.ray_totals, b.ghit);
    else
        launch(k_primary<false>, dim3(nb), dim3(256), 0, s, words, p, b.gbuf, b.gpos, b.node_stats,
                           b.counters, b.ray_totals, b.ghit);
}

// Primary rays of many views in one launch (the batched form): a persistent
// grid whose waves pull runs of tiles from a global counter over the
// concatenated tile lists of all views, so views of every size share the
// chip and a small view's tail overlaps the next view's work.  A tile is
// tile_w x (64 / tile_w) pixels, one per lane; each pixel's ray is the one
// k_primary casts for it (same direction, same traversal), so every view's
// G-buffer equals its own primary pass bit for bit.
template <uint32_t TW>
__global__ __launch_bounds__(256) void k_primary_batch(const BatchView *__restrict__ views,
                                                       const uint32_t *__restrict__ bases, uint32_t n_views,
                                                       uint32_t total, uint32_t *__restrict__ queue, uint32_t grab)
{
    constexpr uint32_t TH = 64u / TW;
    __shared__ uint32_t path_cache[VE_MAX_DEPTH * 256];
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t v = 0;
    for (;;) {
        uint32_t t0 = 0;
        if (lane == 0u) t0 = atomicAdd(queue, grab);
        t0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl((int)t0, 0, 64));
        if (t0 >= total) break;   // every wave reaches the end of the list
        const uint32_t t1 = min(t0 + grab, total);
        if (t0 < bases[v]) v = 0;   // runs are handed out in increasing order per wave
        for (uint32_t t = t0; t < t1; t++) {
            while (v + 1u < n_views && t >= bases[v + 1u]) v++;
            const BatchView &bv = views[v];
            const uint32_t local = t - bases[v];
            const uint32_t x = (local % bv.tiles_x) * TW + (lane % TW);
            const uint32_t y = (local / bv.tiles_x) * TH + lane / TW;
            if (x >= bv.W || y >= bv.H) continue;
            uint32_t checks = 0, reads = 0;
            const f3 dir = primary_dir(bv.cam, x, y, bv.W, bv.H);
            const f3 org = mk(bv.cam[0], bv.cam[1], bv.cam[2]);
            const Hit h = raycast<false>(gptr(bv.words), gptr(bv.jump), bv.jump_level, bv.n_words, bv.depth, org, dir,
                                         rcp3(dir), bv.cc, checks, reads, path_cache + threadIdx.x, 256u);
            const size_t pix = (size_t)y * bv.W + x;
            uint4 *gbuf = gptr(bv.gbuf);
            uint32_t *gpos = gptr(bv.gpos);
            if (h.hit) {
                const uint32_t bits = bv.depth - 1u;
                gbuf[pix] = make_uint4(checks, h.material, 0u, voxel_key(h.px, h.py, h.pz) + 1u);
                gpos[pix] = (h.px >> 1) | ((h.py >> 1) << bits) | ((h.pz >> 1) << (2u * bits));
                if (bv.reuse) gptr(bv.ghit)[pix] = make_uint2(h.px | (h.py << 16), h.pz);
            } else {
                gbuf[pix] = make_uint4(checks, 0u, 0u, 0u);
                gpos[pix] = 0u;
            }
        }
    }
}

void launch_primary_batch(hipStream_t s, const BatchView *views, const uint32_t *bases, uint32_t n, uint32_t total,
                          uint32_t *queue, uint32_t tile_w, uint32_t grab, uint32_t blocks)
{
    const dim3 g(blocks ? blocks : 1u), b(256);
    switch (tile_w) {
    case 16: launch(k_primary_batch<16>, g, b, 0, s, views, bases, n, total, queue, grab); break;
    case 32: launch(k_primary_batch<32>, g, b, 0, s, views, bases, n, total, queue, grab); break;
    case 64: launch(k_primary_batch<64>, g, b, 0, s, views, bases, n, total, queue, grab); break;
    default: launch(k_primary_batch<8>, g, b, 0, s, views, bases, n, total, queue, grab); break;
    }
}

// Jump table: one thread per cell of the grid at depth `level` (cells of
// 2^(depth - level) voxels, index x | y | z bits of the cell coordinates):
// the descent of raycast from the root through depth `level` depends only on
// the coordinate bits above the cell, so its outcome is the same for every
// position in the cell -- the child block of the node at depth `level`, or
// the depth at which a non-internal word ended it (JUMP_ENDED | depth, with
// JUMP_OCC when that word carries a material).
__global__ void k_build_jump(const uint32_t *__restrict__ words, uint32_t n_words, uint32_t depth, uint32_t level,
                             uint32_t *__restrict__ jump)
{
    const uint32_t cells = 1u << (3u * level);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells) return;
    const uint32_t sh = depth - level;
    const uint32_t ux = (i >> (2u * level)) << sh, uy = ((i >> level) & ((1u << level) - 1u)) << sh,
                   uz = (i & ((1u << level) - 1u)) << sh;
    uint32_t offset = 0;
    for (uint32_t dd = 0; dd <= level; dd++) {
        offset += locate_bit(ux, uy, uz, depth - dd);
        const uint32_t raw = node_word(words, n_words, offset);
        if (!(raw & 1u)) {   // the ending word's material bits decide an occupancy lookup
            jump[i] = JUMP_ENDED | (((raw >> 1) & 0x7Fu) != 0u ? JUMP_OCC : 0u) | dd;
            return;
        }
        offset = raw >> 4;
    }
    jump[i] = offset;
}

void launch_build_jump(hipStream_t s, const uint32_t *words, uint32_t n_words, uint32_t depth, uint32_t level,
                       uint32_t *jump)
{
    const uint32_t cells = 1u << (3u * level);
    launch(k_build_jump, dim3((cells + 255u) / 256u), dim3(256), 0, s, words, n_words, depth, level, jump);
}

// ---------------------------------------------------------------------------
// Allocation-free primary pass
// ---------------------------------------------------------------------------
// The jump table of ve_pass_primary_noalloc lives in a file-scope device
// array instead of the scene: a depth-min(6, depth-1) table (at most 8^6
// words, 1 MiB) in `scratch`, after 64 words reserved for the kernel's
// counter arguments.  Each device holds its own copy; its address comes from
// hipGetSymbolAddress on the first call on that device (the array's name is
// a host-side stand-in, never a device address).  Every call on a device
// rebuilds the table over the same memory, so calls that overlap on one
// device race.
constexpr uint32_t NOALLOC_JUMP_LEVEL = 6;
constexpr size_t NOALLOC_HEAD_WORDS = 64;   // counters (16), node stats (8), ray totals (8), spare
constexpr size_t NOALLOC_SCRATCH_WORDS = NOALLOC_HEAD_WORDS + ((size_t)1 << (3u * NOALLOC_JUMP_LEVEL));
static __device__ __align__(256) uint32_t scratch[NOALLOC_SCRATCH_WORDS];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex noalloc_mutex;
static uint32_t *noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

static int noalloc_scratch(int dev, uint32_t **out)
{
    std::lock_guard<std::mutex> lock(noalloc_mutex);
    if (!noalloc_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            char buf[256];
            snprintf(buf, sizeof buf, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
            return ve_global_error(VE_EDEVICE, buf);
        }
        noalloc_ptr[dev] = static_cast<uint32_t *>(ptr);
    }
    *out = noalloc_ptr[dev];
    return VE_OK;
}

}  // namespace ve

extern "C" int ve_pass_primary_noalloc(const uint32_t *words, uint32_t n_words, uint32_t depth, const float ubo[16],
                                       uint32_t width, uint32_t height, int primary_controlchecks, uint32_t *gbuffer,
                                       uint32_t *gbuffer_pos, void *stream)
{
    using namespace ve;
    ve_clear_global_error();
    if (!words || !ubo || !gbuffer || !gbuffer_pos) return ve_global_error(VE_EINVAL, "null argument");
    if (n_words == 0u || n_words > (1u << 28)) return ve_global_error(VE_EINVAL, "n_words outside [1, 2^28]");
    if (depth < 1u || depth > VE_MAX_DEPTH) return ve_global_error(VE_EINVAL, "depth out of range [1,16]");
    if (width < 1u || height < 1u || width > 16384u || height > 16384u)
        return ve_global_error(VE_EINVAL, "framebuffer size outside [1, 16384]");
    const uint32_t jl = depth - 1u < NOALLOC_JUMP_LEVEL ? depth - 1u : NOALLOC_JUMP_LEVEL;
    const size_t need = NOALLOC_HEAD_WORDS + ((size_t)1 << (3u * jl));
    if (need > NOALLOC_SCRATCH_WORDS) {
        char buf[160];
        snprintf(buf, sizeof buf, "jump table needs %zu scratch words, above the %zu available", need,
                 NOALLOC_SCRATCH_WORDS);
        return ve_global_error(VE_ECAPACITY, buf);
    }
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return ve_global_error(VE_EDEVICE, hipGetErrorString(e));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return ve_global_error(VE_EDEVICE, "device ordinal above 63");
    uint32_t *sp = nullptr;
    const int rc = noalloc_scratch(dev, &sp);
    if (rc != VE_OK) return rc;
    FrameParams p;
    std::memset(&p, 0, sizeof p);
    p.W = width;
    p.H = height;
    p.depth = depth;
    p.n_words = n_words;
    uint32_t *const jump = sp + NOALLOC_HEAD_WORDS;
    p.jump = jump;
    p.jump_level = jl;
    p.primary_cc = (uint32_t)primary_controlchecks;
    p.count_primary_in_reset = 1;   // the ray counters are the context's, not written here
    p.reuse_primary_hits = 0;       // no hit records kept
    p.tile_chunk = DEFAULT_TILE_CHUNK;
    std::memcpy(p.cam, ubo, sizeof p.cam);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    launch_build_jump(s, words, n_words, depth, jl, jump);
    const uint32_t tiles = ((width + TILE - 1) / TILE) * ((height + TILE - 1) / TILE);
    const uint32_t nb = padded_blocks((tiles + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK, p.tile_chunk);
    launch(k_primary<false>, dim3(nb), dim3(256), 0, s, words, p, reinterpret_cast<uint4 *>(gbuffer), gbuffer_pos,
           reinterpret_cast<unsigned long long *>(sp + 16), sp, reinterpret_cast<unsigned long long *>(sp + 24),
           static_cast<uint2 *>(nullptr));
    e = hipGetLastError();
    if (e != hipSuccess) return ve_global_error(VE_EDEVICE, hipGetErrorString(e));
    return VE_OK;
}
