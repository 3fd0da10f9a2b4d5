// Internal definitions shared by the host library and the kernels.
#pragma once

#include <stdint.h>

#include "../../include/ve_svo.h"

// ---- error reporting --------------------------------------------------------
// Calls without a context record their message in a thread-local slot that
// ve_last_error(NULL) returns.
void ve_clear_global_error();
int ve_global_error(int code, const char *msg);

// ---- constants mirrored from the reference ----------------------------------
// src/renderer/renderer.hpp:90-103 and src/shd/wavefront.glsl
namespace ve {
constexpr uint32_t LBUFFER_STRIDE = 12;        // wavefront.glsl:289
constexpr uint32_t LBUFFER_SLOTS = 32;         // renderer.hpp:90
constexpr uint32_t NBUFFER_STRIDE = 3;         // normal.comp:31
constexpr uint32_t NBUFFER_SLOTS = 32;         // renderer.hpp:91
constexpr uint32_t RBUFFER_STRIDE = 8;         // wavefront.glsl:274
constexpr uint32_t RBUFFER_SLOTS = 16;         // renderer.hpp:92
constexpr uint32_t SHADE_BUDGET_MAX = 196608;  // renderer.hpp:95
constexpr uint32_t CLAIM_WIDTH = 16384;        // renderer.hpp:96
constexpr uint32_t CLAIM_PROBES = 64;          // renderer.hpp:97
constexpr uint32_t EMISSIVE_CAPACITY = 4096;   // wavefront.glsl:194 (the shader's value governs)
constexpr uint32_t EMISSIVE_CLAIM_W = 8192;    // wavefront.glsl:195
constexpr uint32_t EMISSIVE_CLAIM_PROBES = 8;  // wavefront.glsl:196
constexpr uint32_t CANDIDATE_CAPACITY = 1048576;  // wavefront.glsl:259
constexpr uint32_t HISTOGRAM_BINS = 256;       // wavefront.glsl:260
constexpr uint32_t RR_START_BOUNCE = 2;        // wavefront.glsl:112
constexpr uint32_t TRUSTED_THRESHOLD = 128;    // resolve.comp:82
constexpr uint32_t RANK_STATS_WORDS = 264;     // wavefront.glsl:243-255
constexpr uint32_t NUM_COUNTERS = 16;
// past the counters: the overflow side words of early schedule / normal
// passes, {schedule, normal} per buffer set (FrameParams::early_set)
constexpr uint32_t COUNTER_SIDE_WORDS = 4;
constexpr uint32_t SCHED_SLOTS = 128;          // ordering table: voxels per claim row before overflow
// Voxels past a claim row's table go to a per-frame overflow list, and their
// rows are ranked exactly by a separate pass: up to SCHED_OVER_CAP records
// per frame and SCHED_ROW_OVER of them in one row (beyond either, the claim
// order is no longer the reference's and VE_CNT_OVERFLOW is set).
constexpr uint32_t SCHED_OVER_CAP = 4096, SCHED_ROW_OVER = 2048;
constexpr uint32_t MATERIAL_SLOTS = 127;       // shade.comp:90
constexpr uint32_t DEFERRED_BLOCKS = 256;      // deposit-record passes' grid (entries strided over it)
constexpr uint32_t ORDER_LDS_EVENTS = 512;     // per-bucket event lists up to this length sort in LDS
constexpr uint32_t DEP_CAP = 128;              // deposits per lBuffer bucket listed in place (the rest overflow)
// One deposit of the frame (depositSampleDual, shade.comp:162-257), held
// until the frame's bounces are done and then replayed per lBuffer bucket in
// the reference's order.  `meta`: the ray's shade-list index (bits 0-17), the
// bounce iteration (bits 18-21) and the deposit's call site in shade.comp
// (bits 22-23: DEP_MISS, DEP_LAST, DEP_RR).
struct Deferred {
    uint32_t vid, meta;
    float dl[3], il[3];
};
constexpr uint32_t DEP_MISS = 0, DEP_LAST = 1, DEP_RR = 2;   // shade.comp:558-564, 840-845, 877-889
constexpr uint32_t RAY_CHUNK = 49152;          // renderer.hpp:94: rays per emit chunk (renderer.cpp:476-606)
constexpr uint32_t MAX_BOUNCE_ITERS = 8;       // renderer.cpp:515: bounce iterations per chunk
constexpr uint32_t RANK_BLOCKS = 1024;         // rank pass grid (candidates strided over it)
// The rank blocks flush their counts, min/max and histogram into RANK_COPIES
// copies (block b into copy b % RANK_COPIES), so no address takes more than
// RANK_BLOCKS / RANK_COPIES of those atomics; copy layout: count, ~min, max,
// histogram[256].
constexpr uint32_t RANK_COPIES = 32, RANK_COPY_WORDS = 3 + HISTOGRAM_BINS;
// The fused bounce pass's waves flush their traced-ray counts and deepest
// iteration into BOUNCE_COPIES copies of {rays, iterations run}.
constexpr uint32_t BOUNCE_COPIES = 32;
// Jump tables are built at depth min(JUMP_LEVEL, scene depth - 1): 8^8
// entries (64 MiB, resident in the MALL), rebuilt in well under a
// millisecond after an edit.  Measured on the depth-10 frame: level 6 (1 MiB)
// 3015, 7 3110, 8 3207, 9 (512 MiB) 3244 Mrays/s.
constexpr uint32_t JUMP_LEVEL = 8;
// Screen passes deal runs of 16 tile groups (64 tiles) to the XCDs in turn:
// measured on the depth-10 1080p orbit, the primary pass takes 0.159 ms
// against 0.194 ms with one contiguous band per XCD (the bands of sky rows
// finish early and leave their XCDs idle), the normal and resolve passes
// gain 10-12% the same way.
constexpr uint32_t DEFAULT_TILE_CHUNK = 16;
constexpr uint32_t JUMP_ENDED = 0x80000000u;   // entry flag: the descent ended at depth (entry & 0xFF)
constexpr uint32_t JUMP_OCC = 0x100u;          // with JUMP_ENDED: the word that ended it has a material
constexpr uint32_t RESOLVE_TIMEOUT_BLOCKS = 1024; // 4 waves each, striding over the timed-out pixels
// Timed-out pixels are reconstructed per 16x16 tile from LDS for radii up to
// RECON_TILED_MAX (the default is 10); larger radii gather per pixel.
constexpr uint32_t RECON_TILE = 16;
constexpr int RECON_TILED_MAX = 12;
constexpr uint32_t RECON_TILE_BLOCKS = 4096;  // blocks striding over the listed tiles
constexpr uint32_t RECON_DENSE = 24;          // timed-out pixels from which a tile stages its window in LDS
constexpr uint32_t RECON_TILED_MIN = 65536;   // timed-out pixels in a frame from which dense tiles are tiled

// Device layout of one hash bucket (one column h of the reference's R32UI
// images).  The reference interleaves each slot's fields (slot s, field f at
// image row s*stride + f, and every row is a separate 64 KiB-strided image
// line); here a bucket is one contiguous block with the 32 owner words first,
// so finding a voxel id is one 128-byte read instead of up to 32 strided
// probes.  ve_read / ve_write convert to and from the reference layout.
//   nBuffer bucket: owner[32] | normal[32] | time[32]                 (96 words)
//   lBuffer bucket: owner[32] | time[32]   | payload[32][10]          (384 words)
//                   payload = dCnt dR dG dB dL2 iCnt iR iG iB iL2
//   rBuffer bucket: owner[16] | time[16]   | payload[16][6]           (128 words)
constexpr uint32_t NB_NRM = 32, NB_TIME = 64, NB_WORDS = 96;
constexpr uint32_t LB_TIME = 32, LB_PAY = 64, LB_PAYW = 10, LB_WORDS = 384;
constexpr uint32_t RB_TIME = 16, RB_PAY = 32, RB_PAYW = 6, RB_WORDS = 128;

// Ray record, 96 bytes, field for field the reference's std430 Ray
// (wavefront.glsl:128-143).  `pad0` carries the hit voxel extent between
// trace and shade, as the reference does.
struct Ray {
    float ox, oy, oz, pad0;
    float dx, dy, dz, pad1;
    float tx, ty, tz;
    uint32_t primary_vid;
    float dlx, dly, dlz;
    uint32_t rng;
    float ilx, ily, ilz;
    uint32_t bounce;
    uint32_t hit_vid;
    uint32_t hit_material;
    uint32_t pad2, pad3;
};
static_assert(sizeof(Ray) == 96, "Ray must match the reference's 96-byte std430 record");

// std140 material entry (material.hpp:17-25 / shade.comp:84-90)
struct Material {
    float color[4];
    float specular_color[4];
    float roughness, specular, metallic;
    int32_t emissive;
    float emissive_intensity;
    float pad[3];
};
static_assert(sizeof(Material) == 64, "std140 material entry is 64 bytes");

// 32-byte EmissiveVoxel (wavefront.glsl:176-181)
struct EmissiveVoxel {
    float ex, ey, ez;
    uint32_t vid;
    uint32_t px, py, pz;
    uint32_t size;
};
static_assert(sizeof(EmissiveVoxel) == 32, "EmissiveVoxel is 32 bytes");

// Frame parameters handed to every kernel by value.
// The inputs that change from frame to frame (Renderer::frameID, frameSeed,
// glTime and the camera UBO).  A frame replayed from a captured graph reads
// them from a device copy (FrameParams::dyn) instead of its kernel arguments.
struct FrameDyn {
    uint32_t frame_idx, frame_seed, time, pad;
    float cam[16];
};

struct FrameParams {
    uint32_t W, H;
    uint32_t depth;
    uint32_t n_words;
    uint32_t hash_width;
    uint32_t max_bounces;
    uint32_t frame_idx, frame_seed, time;
    uint32_t primary_cc, bounce_cc;
    uint32_t update_time;
    int32_t disable_stale_reset;
    int32_t normal_precision;
    uint32_t shade_budget;
    uint32_t cap_direct, cap_indirect;
    float w_pixels, w_variance, pixels_norm, variance_norm, top_fraction, mid_fraction;
    int32_t min_samples;
    int32_t restir_enabled, restir_m, restir_temporal, restir_spatial, restir_neighbors;
    float restir_radius;
    int32_t restir_max_m;
    float firefly_k, firefly_floor;
    int32_t render_type;
    int32_t reconstruction_radius;
    // 1 inside a whole frame: the reset pass counts the frame's primary rays
    // (the primary pass may run ahead on another stream); 0 for passes run
    // one by one, where the primary pass counts them itself
    int32_t count_primary_in_reset;
    // 1 inside a whole frame: two one-wave passes ride on the next kernel of
    // the stream instead of a launch of their own -- the threshold on the emit
    // kernel's first wave, the deposit records' clear on the resolve's first
    // thread (passes run one by one keep their own launches)
    int32_t fold_small_passes;
    // 1 inside a whole frame (without node counting): the primary pass keeps
    // each hit's lattice position and the emit pass fills in the hit record
    // of a bounce-0 ray whose camera ray is the pixel's primary ray
    int32_t reuse_primary_hits;
    // 1 + the frame's buffer set when its schedule and normal passes run on
    // the auxiliary streams ahead of the previous frame's later passes, and
    // its resolve on a third one beside the next frame's (0 otherwise): the
    // early passes report overflow into that set's side words past the
    // counters, which the kernel clearing the deposit records folds into the
    // overflow counter; the reset pass keeps the schedule's candidate count
    // and the resolve's timeout counts
    int32_t early_set;
    // sky: a cube map of sky_edge^2 texels per face (faces +X -X +Y -Y +Z -Z,
    // texel r | g << 8 | b << 16), or, with sky_edge 0, the constant `sky`
    uint32_t sky_edge;
    const uint32_t *sky_tex;
    // the scene's jump table (see raycast)
    const uint32_t *jump;
    uint32_t jump_level;
    // screen passes: tile groups per XCD run (0 = one band per XCD)
    uint32_t tile_chunk;
    // fused bounce pass: grid in blocks of 256 (0 = one thread per shade-list
    // entry); lanes past the grid's rays take further rays from a queue
    uint32_t bounce_blocks;
    // non-null: frame_idx, frame_seed, time and cam come from this device block
    const FrameDyn *dyn;
    float cam[16];
    float sky[3];
};
}  // namespace ve
