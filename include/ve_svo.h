/*
 * ve_svo.h — C ABI of the MI355X sparse-voxel-octree ray pipeline.
 *
 * Drop-in for VoxelEngine3.0's GL pass API: the octree build
 * (OctreeCPU::insert* + Octree::set), the per-frame wavefront sequence
 * Renderer::runWavefrontFrame (generate_primary, normal, schedule, rank,
 * threshold, emit, trace/shade bounces, resolve) and the buffer readbacks.
 *
 * Conventions
 *   - Every int-returning call returns VE_OK (0) or a negative VE_E* code; the
 *     message is available from ve_last_error(ctx) (ctx may be NULL for calls
 *     that take no context).  This mirrors the reference's `success` flag that
 *     Renderer::run returns (src/renderer/renderer.cpp:977-984,711).
 *   - The library owns all device memory; the caller owns host buffers.
 *   - One ve_ctx = one HIP stream on one device.  A context is not
 *     thread-safe (like a GL context); separate contexts may run concurrently.
 *   - Node words are the reference's exact 32-bit format
 *     (src/renderer/octree.hpp:16-37): internal = bit0 1 | count<<1 (3 bits)
 *     | next<<4 (28 bits); leaf = bit0 0 | material<<1 (7 bits) | normal<<8
 *     (24 bits, always 0); empty = 0.
 */
#ifndef VE_SVO_H
#define VE_SVO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VE_OK            0
#define VE_EINVAL      (-1)  /* bad argument (null pointer, zero size, depth out of range) */
#define VE_ENOMEM      (-2)  /* host or device allocation failed */
#define VE_EDEVICE     (-3)  /* HIP runtime error (message names the call) */
#define VE_ECAPACITY   (-4)  /* node array exceeds the 28-bit `next` field or the word limit */
#define VE_EFORMAT     (-5)  /* malformed node array (a `next` outside the array) */
#define VE_ESTATE      (-6)  /* call out of order (e.g. trace before emit) */
#define VE_EBUDGET     (-7)  /* admission refused: device-memory budget exceeded */
#define VE_ENOTFOUND   (-8)  /* unknown handle / buffer id */

#define VE_MAX_DEPTH   16

/* One OctreeCPU::insert(position, material) call (src/octree_cpu.cpp:40-64).
 * A list of these, in call order, is what insertBox / insertSphere /
 * insertFunction expand to; later inserts at the same position win. */
typedef struct ve_insert {
    uint32_t x, y, z, material;
} ve_insert;

typedef struct ve_octree_info {
    uint32_t depth;
    uint32_t size;        /* used words (Octree::size, src/renderer/octree.cpp:131,299-300) */
    uint32_t capacity;    /* power-of-two word capacity (Octree::resizeDataIfNeeded, octree.cpp:95-119) */
    uint32_t num_voxels;  /* leaves with material != 0 (Octree::numVoxels) */
} ve_octree_info;

/* ---- octree_cpu build -------------------------------------------------------
 * Replaces OctreeCPU::insert (src/octree_cpu.cpp:40-64) followed by
 * Octree::set / linearizeNode (src/renderer/octree.cpp:127-217).  Produces the
 * node words in the reference's exact DFS block order.  `max_words` is the
 * driver's texture-buffer limit (GL_MAX_TEXTURE_BUFFER_SIZE, octree.cpp:102-105);
 * a tree larger than it (or than the 28-bit next field) fails with
 * VE_ECAPACITY instead of the reference's out-of-bounds write.
 * *words_out has `capacity` entries (zero past `size`); free it with ve_free. */
int ve_octree_build(const ve_insert *ops, size_t n_ops, uint32_t depth, uint32_t max_words,
                    uint32_t **words_out, ve_octree_info *info_out);
void ve_free(void *p);

/* ---- synthetic cave terrain -------------------------------------------------
 * No counterpart in the reference: the generator of the depth-12 divergence
 * scene (a heightfield ground with 3D value-noise caves under its surface),
 * a data source for tests and benchmarks whose voxels go through
 * ve_octree_build.  Per column (x, z) of an n x n heightfield (hi = one past
 * the top ground voxel, lo = first voxel of the watertight surface shell,
 * row-major [x][z]):
 *   cave   = hi - band <= y < hi and value_noise((x+.5)/cell, (y+.5)/cell, (z+.5)/cell) > threshold
 *   solid  = 0 <= y < hi and not cave
 *   output = solid and (y >= lo or a 6-neighbour inside the grid is a cave)
 * value_noise: smoothstep trilinear blend (x, then y, then z) of lattice values
 * (hash4(i, j, k, noise_seed) >> 8) / 2^24, in double precision; material
 * 1 + (hash4(x, y, z, material_seed ^ 0x5EED) & 7); hash4 as in
 * voxelengine3.0_amd/scenes.py.  band <= 63.  With out null or capacity below
 * the voxel count, only *count is set.  threads 0 = up to 16. */
int ve_gen_caves(uint32_t n, const int32_t *hi, const int32_t *lo, uint32_t band, uint32_t cell, double threshold,
                 uint32_t noise_seed, uint32_t material_seed, uint32_t threads, ve_insert *out, size_t capacity,
                 size_t *count);

/* ---- renderer ---------------------------------------------------------------- */
typedef struct ve_ctx ve_ctx;
typedef struct ve_scene ve_scene;

/* Renderer construction parameters (Renderer::Renderer, renderer.cpp:20-169).
 * hash_width is the driver's GL_MAX_TEXTURE_SIZE, which sizes the l/n/rBuffer
 * hash images (renderer.cpp:27-40); 16384 on Mesa llvmpipe. */
typedef struct ve_renderer_config {
    int      device;          /* HIP device ordinal */
    uint32_t max_width;       /* largest framebuffer this context will render */
    uint32_t max_height;
    uint32_t hash_width;      /* W_l: width of lBuffer / nBuffer / rBuffer (0 = 16384) */
    uint32_t max_bounces;     /* MAX_BOUNCES of wavefront.glsl:108 (0 = 8) */
} ve_renderer_config;

int  ve_create(ve_ctx **ctx, const ve_renderer_config *cfg);
void ve_destroy(ve_ctx *ctx);
const char *ve_last_error(const ve_ctx *ctx);

/* Octree::set upload (octree.cpp:127-166): copies `n_words` node words to the
 * device.  The words are validated (every `next` block must lie inside the
 * array) so a malformed array fails with VE_EFORMAT instead of faulting.
 * Besides the words, a scene holds a traversal jump table of
 * 8^min(8, depth - 1) words (64 MiB from depth 9 up), built here and rebuilt by
 * every ve_scene_update; results do not depend on it. */
int ve_scene_upload(ve_ctx *ctx, const uint32_t *words, uint32_t n_words, uint32_t depth,
                    ve_scene **scene);
int ve_scene_release(ve_ctx *ctx, ve_scene *scene);

/* No counterpart in the reference: launch-shape knobs, which change speed
 * only, never results.  tile_chunk: how the 8x8-pixel tiles of the screen
 * passes are dealt to the chip's 8 XCDs -- 0 gives each XCD one contiguous
 * band of rows; n > 0 (at most 4096; default 16) deals runs of n groups of
 * 4 tiles to the XCDs in turn.  jump_level: depth of the traversal jump
 * table of scenes uploaded on this context from now on (0 = default 8;
 * 1..8: 8^level words).  bounce_blocks: grid of the fused bounce pass in
 * blocks of 256 threads (0 = one thread per shade-list entry; at most 65536);
 * with fewer threads than rays, a thread whose ray ends takes the next one
 * from a queue, so lanes idle less while the longest paths run. */
typedef struct ve_tuning {
    uint32_t tile_chunk;
    uint32_t jump_level;
    uint32_t bounce_blocks;
} ve_tuning;
int ve_set_tuning(ve_ctx *ctx, const ve_tuning *t);
int ve_get_tuning(ve_ctx *ctx, ve_tuning *t);

/* 64-byte std140 CameraUniform (camera.cpp:39-50): position, cameraPlane
 * (forward), cameraPlaneRight, cameraPlaneUp, each a vec4 with w = 0. */
int ve_set_camera(ve_ctx *ctx, const float ubo[16]);
/* Camera::UpdateUBO's arithmetic on the host (camera.cpp:43-44), FP32. */
void ve_camera_ubo(const float position[3], const float direction[3], float fov_deg,
                   float aspect, float ubo_out[16]);

/* Material UBO (material.hpp:17-25): 64-byte std140 entries, entry 0 unused,
 * up to 127 entries (shade.comp:84-90). */
int ve_set_materials(ve_ctx *ctx, const void *std140_entries, uint32_t n_entries);

/* Sky (Skybox, skybox.cpp:71-122): six faces +X -X +Y -Y +Z -Z of edge x edge
 * RGB8 texels, rows tightly packed in glTexImage2D order (the reference
 * uploads rows 4-byte aligned, the same bytes when edge is a multiple of 4),
 * sampled along the ray on every miss (shade.comp:559) and along the pixel's
 * primary direction for sky pixels (resolve.comp:218-288) with the
 * reference's sampler state: GL_LINEAR, CLAMP_TO_EDGE, no mipmaps, no
 * seamless filtering.  A constant sky is edge = 1 (any uniform cube is kept
 * as a constant).  edge <= 16384.  Waits for the frames already enqueued on
 * the context before replacing a cube map. */
int ve_set_sky(ve_ctx *ctx, const uint8_t *const faces_rgb8[6], int edge);

/* Per-frame knobs = core::FrameConfig (src/renderer/core.hpp:83-158). */
typedef struct ve_frame_config {
    int32_t  primary_controlchecks;   /* 70 */
    int32_t  bounce_controlchecks;    /* 100 */
    int32_t  reconstruction_radius;   /* 10 */
    int32_t  normal_precision;        /* 6 */
    int32_t  disable_stale_reset;     /* 0 */
    uint32_t update_time;
    int32_t  shading_budget;          /* 131072, clamped to [1, 196608] */
    int32_t  sample_cap_direct;       /* 16 */
    int32_t  sample_cap_indirect;     /* 64 */
    float    weight_pixels;           /* 0.1 */
    float    weight_variance;         /* 0.9 */
    float    pixels_norm;             /* 100 */
    float    variance_norm;           /* 10000 */
    float    top_fraction;            /* 0.30 */
    float    mid_fraction;            /* 0.40 */
    int32_t  min_samples;             /* 1 */
    int32_t  restir_enabled;          /* 1 */
    int32_t  restir_m;                /* 4 */
    int32_t  restir_temporal;         /* 1 */
    int32_t  restir_spatial;          /* 1 */
    int32_t  restir_spatial_neighbors;/* 1 */
    float    restir_spatial_radius;   /* 2.75 */
    int32_t  restir_max_m;            /* 24 */
    float    firefly_k;               /* 1.0 */
    float    firefly_floor;           /* 255 */
    int32_t  render_type;             /* 0 = DEFAULT (core::RenderType) */
} ve_frame_config;

void ve_frame_config_defaults(ve_frame_config *fc);

/* Per-frame inputs the reference takes from the wall clock
 * (renderer.cpp:283,293 and voxelengine.cpp:200,211), made explicit. */
typedef struct ve_frame_inputs {
    uint32_t width, height;
    uint32_t frame_idx;    /* Renderer::frameID */
    uint32_t frame_seed;   /* schedule scramble seed */
    uint32_t time;         /* glTime */
} ve_frame_inputs;

/* Renderer::runWavefrontFrame (renderer.cpp:282-649): the whole frame,
 * enqueued on the context's stream with no host round-trips. */
int ve_frame(ve_ctx *ctx, ve_scene *scene, const ve_frame_config *fc, const ve_frame_inputs *in);

/* One glDispatchCompute equivalent, for per-pass parity tests. */
typedef enum ve_pass_id {
    VE_PASS_PRIMARY = 0,   /* generate_primary.comp */
    VE_PASS_NORMAL,        /* normal.comp */
    VE_PASS_RESET,         /* per-frame meta/claim reset (renderer.cpp:343-375) */
    VE_PASS_SCHEDULE,      /* schedule.comp */
    VE_PASS_RANK,          /* rank.comp */
    VE_PASS_THRESHOLD,     /* threshold.comp */
    VE_PASS_EMIT,          /* emit.comp (whole shade list, one ray batch) */
    VE_PASS_TRACE,         /* trace.comp over the live rays */
    VE_PASS_SHADE,         /* shade.comp over the live rays */
    VE_PASS_RESOLVE,       /* resolve.comp */
    VE_PASS_DISPLAY,       /* final.frag: ACES tonemap of the resolve output into RGBA8 (VE_BUF_DISPLAY);
                              Renderer::run's blit (renderer.cpp:683-704), not part of ve_frame */
    VE_PASS_COUNT
} ve_pass_id;

int ve_pass(ve_ctx *ctx, ve_scene *scene, ve_pass_id pass, const ve_frame_config *fc,
            const ve_frame_inputs *in);

/* No counterpart in the reference: the VE_PASS_PRIMARY dispatch of ve_pass
 * (generate_primary.comp) for callers that cannot allocate device memory --
 * code under HIP graph capture, serving loops that forbid hipMalloc.  It needs
 * no context or scene: `words` (n_words node words) and the outputs are the
 * caller's device buffers (gbuffer: uint32 x4 per pixel as VE_BUF_GBUFFER,
 * gbuffer_pos: uint32 per pixel as VE_BUF_GBUFFER_POS), `ubo` is the host
 * camera of ve_set_camera, `stream` a hipStream_t (NULL = the null stream).
 * The traversal jump table is rebuilt on every call in a static per-device
 * scratch array of 1 MiB + 256 bytes (a depth-min(6, depth-1) table; the
 * call fails with VE_ECAPACITY if a table would not fit), so the call makes
 * no hipMalloc, hipFree or other allocation; results equal ve_pass's bit for
 * bit.  Calls that overlap on one device share that scratch and race.
 * Asynchronous: returns after enqueueing; errors of the arguments return
 * VE_EINVAL with ve_last_error(NULL). */
int ve_pass_primary_noalloc(const uint32_t *words, uint32_t n_words, uint32_t depth, const float ubo[16],
                            uint32_t width, uint32_t height, int primary_controlchecks, uint32_t *gbuffer,
                            uint32_t *gbuffer_pos, void *stream);

/* Buffers readable / writable by id (device copies, synchronous). */
typedef enum ve_buffer_id {
    VE_BUF_GBUFFER = 0,    /* uint32 x4 per pixel: checks, material, 0, voxelKey+1 */
    VE_BUF_GBUFFER_POS,    /* uint32 per pixel: packed half-coordinates */
    VE_BUF_NBUFFER,        /* uint32 [hash_width][3*32]  (bucket-major) */
    VE_BUF_LBUFFER,        /* uint32 [hash_width][12*32] (bucket-major) */
    VE_BUF_RBUFFER,        /* uint32 [hash_width][8*16]  (bucket-major) */
    VE_BUF_CANDIDATES,     /* uint32 x4 per candidate: vid, pix, claim_h, claim_p */
    VE_BUF_SHADE_LIST,     /* uint32 x2 per entry: vid, pix */
    VE_BUF_RAYS,           /* 96-byte Ray records of the current batch */
    VE_BUF_OUTPUT,         /* float x4 per pixel (resolve RGBA32F) */
    VE_BUF_COUNTERS,       /* uint32 x16: see VE_CNT_* */
    VE_BUF_RANK_STATS,     /* uint32 x264: RankStats (wavefront.glsl:243-255) */
    VE_BUF_SAMPLE_COUNTS,  /* uint32 [16384][64] per-voxel pixel counts (claim-map layout) */
    VE_BUF_CLAIM_MAP,      /* uint32 [16384][64] */
    VE_BUF_EMISSIVES,      /* 32-byte EmissiveVoxel records */
    VE_BUF_DISPLAY,        /* uint8 x4 per pixel: tonemapped RGBA8 of the last display pass */
    VE_BUF_COUNT
} ve_buffer_id;

/* VE_BUF_COUNTERS layout */
#define VE_CNT_CANDIDATES     0   /* candidate_count */
#define VE_CNT_SHADE          1   /* shade_count (after budget) */
#define VE_CNT_EMISSIVES      2   /* emissive_count */
#define VE_CNT_LIVE           3   /* live rays entering the next trace */
#define VE_CNT_RAYS_TRACED    4   /* rays traced this frame by trace (all bounces) */
#define VE_CNT_SHADOW_RAYS    5   /* NEE shadow rays this frame */
#define VE_CNT_PRIMARY_RAYS   6   /* W*H of the last primary pass */
#define VE_CNT_BOUNCES_RUN    7   /* trace/shade iterations with live rays */
#define VE_CNT_LIVE_B         8   /* live-ray count of the other ray buffer (internal) */
#define VE_CNT_OVERFLOW       9   /* nonzero: a claim row overflowed the ordering table */
#define VE_CNT_REUSED        10   /* bounce-0 rays this frame whose hit record came from the primary pass */
#define VE_CNT_TIMEOUTS      11   /* pixels whose primary ray timed out (last resolve) */
#define VE_CNT_EM_BASE       12   /* emissive registry size before this frame's registrations (internal) */
#define VE_CNT_EM_NEW        13   /* registrations made by this frame's first shade (internal) */
#define VE_CNT_DEFERRED      14   /* deposits of later bounce iterations held back in the fused bounce pass (internal) */
#define VE_CNT_TIMEOUT_TILES 15   /* 16x16 screen tiles holding a timed-out pixel (last resolve; internal) */

int ve_read(ve_ctx *ctx, ve_buffer_id id, void *dst, size_t bytes);
int ve_write(ve_ctx *ctx, ve_buffer_id id, const void *src, size_t bytes);
int ve_buffer_size(ve_ctx *ctx, ve_buffer_id id, size_t *bytes);

/* Block until the context's streams are idle. */
int ve_sync(ve_ctx *ctx);

/* Cross-frame overlap (on by default for contexts made by ve_create): each
 * ve_frame renders its primary and schedule passes into a second G-buffer
 * set on an auxiliary stream, and its normal pass into a second nBuffer set
 * on another, beside the previous frame's bounce passes; the context's
 * stream waits for the schedule pass before the rank pass and for the
 * normal pass before the trace pass.  The frame's resolve runs on a
 * third auxiliary stream beside the next frame's first passes; any call on
 * the context other than the next ve_frame first waits for it.  Every
 * frame's results are those of the serial order.  Off: all passes in order
 * on the context's stream (Renderer::run's sequence).  No counterpart in
 * the reference, whose GL passes run in submission order; VE_EINVAL for a
 * context without the auxiliary streams (server sessions). */
int ve_set_overlap(ve_ctx *ctx, int enable);

/* No counterpart in the reference: graph replay.  While enabled (and the
 * overlap is off, and timing is off), ve_frame captures the frame's launches
 * once into a HIP graph and replays it for later frames; the frame's varying
 * inputs (frame_idx, frame_seed, time, camera) reach the kernels through a
 * small device block copied before each replay, so results equal those of
 * direct launches bit for bit.  A frame whose other parameters differ
 * (size, knobs, scene, sky) is captured anew.  ve_graph_stats counts the
 * captures and replays. */
int ve_set_graph(ve_ctx *ctx, int enable);
int ve_graph_stats(ve_ctx *ctx, uint64_t *captures, uint64_t *replays);

/* ---- batched form -------------------------------------------------------------
 * No counterpart in the reference: frame k of many views at once.  Their
 * primary passes (generate_primary.comp) run in one persistent launch whose
 * waves pull tiles from a global queue over all the views' tiles; each
 * view's other passes then run on its own context's stream, replayed from its
 * frame graph when ve_set_graph is on.  Every view's results equal those of
 * ve_frame on its context bit for bit.  The contexts need the cross-frame
 * overlap, timing and node counting off, and each view's camera is its
 * context's (ve_set_camera).  Launch shape: tiles of tile_w x (64 / tile_w)
 * pixels (tile_w 8, 16, 32 or 64; 0 = 8), `grab` tiles per queue pull (0 = 4),
 * a grid of waves_per_cu waves per compute unit (0 = 16); at most max_views
 * views per call (0 = 256).  The views are split into `groups` launches (0 =
 * 4) of about equal work, largest views first: a group's views start their
 * other passes as soon as its launch ends, beside the next group's launch.
 * voxelengine3.0_amd/autotune.py picks them.  whole_frame = 1: the batch runs
 * whole frames -- after the one primary launch (groups is then 1), every
 * other pass of all the views runs as one launch per kernel on the batch's
 * stream, each block finding its view in a table of the views' records, so a
 * frame of n views takes the same ~22 launches as a frame of one; the views'
 * streams wait for the last of them.  Results are again those of ve_frame,
 * bit for bit. */
typedef struct ve_batch ve_batch;
typedef struct ve_batch_config {
    int32_t  device;
    uint32_t max_views;
    uint32_t tile_w;
    uint32_t grab;
    uint32_t waves_per_cu;
    uint32_t groups;
    uint32_t whole_frame;
} ve_batch_config;
int  ve_batch_create(ve_batch **out, const ve_batch_config *cfg);
void ve_batch_destroy(ve_batch *b);
const char *ve_batch_last_error(const ve_batch *b);
/* Enqueues one frame of each view (fcs[i], ins[i] as for ve_frame); no host
 * synchronisation. */
int  ve_batch_frame(ve_batch *b, uint32_t n, ve_ctx *const *ctxs, ve_scene *const *scenes,
                    const ve_frame_config *fcs, const ve_frame_inputs *ins);
/* Milliseconds of the last batched primary launch (waits for it; for
 * whole-frame batches, of the whole batch). */
int  ve_batch_primary_ms(ve_batch *b, float *ms);
/* Whole-frame batches whose every view's context replays graphs
 * (ve_set_graph): the batch's launches (the primary launch and one launch
 * per later kernel) are captured into one HIP graph per launch shape (view
 * count, grids, dynamic LDS) and replayed; the views' records are data in the
 * batch's device block, copied before each replay.  No counterpart in the
 * reference.  Counts of captures and replays. */
int  ve_batch_graph_stats(ve_batch *b, uint64_t *captures, uint64_t *replays);

/* Optional per-pass timing: while enabled, every ve_frame binds HIP events to
 * the first and last kernel dispatch of each pass (no host synchronisation,
 * no marker packets); ve_timings waits for the streams and returns, per pass,
 * the mean milliseconds per frame over the frames recorded since the last
 * call (trace and shade: all bounce iterations of a frame together).  An
 * `enable` above 1 is the number of frames about to be recorded; their events
 * are created up front. */
int ve_enable_timing(ve_ctx *ctx, int enable);
int ve_timings(ve_ctx *ctx, float ms_out[VE_PASS_COUNT]);
/* No counterpart in the reference: the same means split by bounce iteration,
 * read before ve_timings (which starts a new recording).  iter 0: the first
 * bounce's trace and shade passes and every other pass; iter 1: the fused
 * kernel of the later iterations under VE_PASS_TRACE and the frame's ordered
 * deposits under VE_PASS_SHADE (0 elsewhere). */
int ve_timings_iter(ve_ctx *ctx, uint32_t iter, float ms_out[VE_PASS_COUNT]);

/* Cumulative ray counts since the context was created (or ve_reset_ray_totals):
 * primary rays (W*H per primary pass), rays traced by the bounce passes (every
 * live ray of each trace dispatch, as the reference counts them), and NEE
 * shadow rays.  Kept on the device; reading synchronises the stream. */
int ve_ray_totals(ve_ctx *ctx, uint64_t *primary, uint64_t *traced, uint64_t *shadow);
/* Of the traced rays, those that were not traversed: inside ve_frame a
 * bounce-0 ray is the pixel's camera ray, which the primary pass already
 * traversed to the same voxel within the bounce budget, so the emit pass
 * fills in its hit record from the primary pass (identical to a traversal).
 * No counterpart in the reference, which traverses it again. */
int ve_ray_totals_reused(ve_ctx *ctx, uint64_t *reused);
int ve_reset_ray_totals(ve_ctx *ctx);

/* Node-read instrumentation: when enabled, the primary and trace kernels add
 * the number of node words they fetch to a device counter (4 bytes each is the
 * algorithmic traffic of the traversal).  Off by default: it adds atomics. */
int ve_enable_node_stats(ve_ctx *ctx, int enable);
int ve_node_stats(ve_ctx *ctx, uint64_t *primary_reads, uint64_t *trace_reads,
                  uint64_t *primary_steps, uint64_t *trace_steps);
/* No counterpart in the reference: the share of trace_reads / trace_steps made
 * by the fused kernel of bounce iterations 1 and later. */
int ve_node_stats_bounce(ve_ctx *ctx, uint64_t *reads, uint64_t *steps);

/* ---- editable octree: OctreeCPU + Octree::set / applyEdits -------------------
 * The persistent pointer tree of OctreeCPU (src/octree_cpu.hpp:9-60) with the
 * node-word image Octree keeps for it (src/renderer/octree.hpp).  Inserts and
 * removes mark paths dirty; ve_octree_set linearises everything
 * (Octree::set, octree.cpp:127-217); ve_octree_apply_edits rewrites only the
 * dirty paths, reusing freed 8-word blocks last-freed-first
 * (Octree::applyEdits, octree.cpp:228-315), and lists the slots it wrote so
 * ve_scene_update can carry just those words to the device.  Host memory only;
 * not thread-safe. */
typedef struct ve_octree_cpu ve_octree_cpu;

/* max_words: the word limit (GL_MAX_TEXTURE_BUFFER_SIZE, octree.cpp:102-105);
 * 0 = 2^27, values above 2^28 (what the 28-bit `next` field addresses) are
 * clamped to 2^28, values below 9 are VE_EINVAL. */
int  ve_octree_cpu_create(ve_octree_cpu **out, uint32_t depth, uint32_t max_words);
void ve_octree_cpu_destroy(ve_octree_cpu *t);
/* OctreeCPU::insert per entry, in order (octree_cpu.cpp:40-64) */
int  ve_octree_cpu_insert(ve_octree_cpu *t, const ve_insert *ops, size_t n);
/* OctreeCPU::remove per (x,y,z) triple (octree_cpu.cpp:66-98) */
int  ve_octree_cpu_remove(ve_octree_cpu *t, const uint32_t *xyz, size_t n);
/* OctreeCPU::lookup (octree_cpu.cpp:100-111): *present = 1 with the leaf's material */
int  ve_octree_cpu_lookup(const ve_octree_cpu *t, uint32_t x, uint32_t y, uint32_t z, int *present,
                          uint32_t *material);
/* OctreeCPU::raycast (octree_cpu.cpp:128-215), the picking ray of the main loop */
int  ve_octree_cpu_raycast(const ve_octree_cpu *t, const float origin[3], const float direction[3],
                           uint32_t max_steps, int *hit, uint32_t position[3], uint32_t *material);
/* VE_ECAPACITY when the tree needs more words than max_words (the reference
 * writes past its array there): nothing is written past the image, which is
 * incomplete until a set succeeds (ve_octree_apply_edits fails until then). */
int  ve_octree_set(ve_octree_cpu *t, ve_octree_info *info);
/* VE_ECAPACITY when a new subtree finds no block under max_words: as in the
 * reference (octree.cpp:290-294) its slot keeps its old word, and here its
 * path also stays dirty, so a later apply_edits (after removes) writes it. */
int  ve_octree_apply_edits(ve_octree_cpu *t, ve_octree_info *info);
/* The node-word image (capacity words; info.size of them in use). */
int  ve_octree_words(const ve_octree_cpu *t, const uint32_t **words, uint32_t *capacity);
/* Slots written by the last set / apply_edits; *full = 1 when the whole image
 * must be uploaded again (a set, or a capacity resize). */
int  ve_octree_changed(const ve_octree_cpu *t, const uint32_t **slots, size_t *n, int *full);
/* Octree::findGpuSlot (octree.cpp:322-341): word slot of the leaf at a position, or 0xFFFFFFFF. */
uint32_t ve_octree_find_slot(const ve_octree_cpu *t, uint32_t x, uint32_t y, uint32_t z);
/* Bring a device node array up to date after ve_octree_apply_edits: `words`
 * is the new image (n_words in use) and `slots` the words that changed since
 * the scene was uploaded or last updated (words appended since then and not
 * listed are zero).  Grows the device array when n_words exceeds it, and
 * rebuilds the scene's jump table.  Ordered on the context's stream after the
 * frames already enqueued there; returns when the update is complete. */
int  ve_scene_update(ve_ctx *ctx, ve_scene *scene, const uint32_t *words, uint32_t n_words,
                     const uint32_t *slots, size_t n_slots);

/* ---- serving: resident scenes, sessions, continuous batching ----------------
 * No single counterpart in the reference, whose one Renderer renders one view
 * on one GL context (src/voxelengine.cpp:158-290).  A server keeps octrees
 * resident, gives every session its own accumulation state (one ve_ctx:
 * l/n/rBuffer, claim maps, ray buffers), admits sessions while a device-memory
 * budget and a slot count allow (later ones wait in arrival order and are
 * admitted as running sessions close), and launches each submitted frame at
 * once on the session's queue, so frames of different sessions run
 * concurrently on the device.  A frame renders exactly as ve_frame on that
 * session alone would: sessions share only the read-only octrees.
 *
 * `queues` chooses the concurrency: 0 gives every session its own HIP stream;
 * 1 puts every session on one stream, which is the plain one-view-at-a-time
 * loop (frames in submission order, no overlap) used as the baseline.
 * Single-threaded like ve_ctx. */
typedef struct ve_server ve_server;

typedef struct ve_server_config {
    int32_t  device;
    uint32_t max_sessions;   /* admitted at once; 0 = 64 */
    uint64_t hbm_budget;     /* bytes of session state; 0 = 3/4 of the free device memory at creation */
    uint32_t queues;         /* 0 = one stream per session; N >= 1 = sessions share N streams */
} ve_server_config;

typedef struct ve_session_desc {
    uint32_t scene;          /* id from ve_server_add_scene */
    uint32_t width, height;
    uint32_t max_bounces;    /* MAX_BOUNCES, as in ve_renderer_config */
    uint32_t hash_width;     /* 0 = 16384 */
    const void *materials;   /* std140 64-byte entries, copied at open */
    uint32_t n_materials;
    uint8_t  sky_rgb8[3];    /* constant sky */
} ve_session_desc;

typedef struct ve_frame_done {
    uint32_t session;
    uint32_t frame;          /* 0-based count of this session's frames */
    uint64_t tag;            /* the caller's tag from ve_server_submit */
    double   arrival_ms;     /* server clock: when the request arrived */
    double   start_ms;       /* device clock mapped to the server clock: first kernel began */
    double   done_ms;        /* ... last kernel ended */
    uint64_t rays;           /* primary + traced + shadow rays of the frame */
} ve_frame_done;

typedef struct ve_server_stats {
    uint32_t admitted;       /* sessions holding device state */
    uint32_t waiting;        /* sessions queued for admission */
    uint32_t frames_in_flight;
    uint32_t pooled;         /* idle contexts kept for reuse */
    uint64_t bytes_in_use;   /* admitted sessions + pooled contexts */
    uint64_t budget;
} ve_server_stats;

int  ve_server_create(ve_server **out, const ve_server_config *cfg);
void ve_server_destroy(ve_server *srv);
const char *ve_server_last_error(const ve_server *srv);
/* Milliseconds since the server was created (the clock of ve_frame_done). */
double ve_server_now_ms(const ve_server *srv);
/* Upload a read-only octree (as ve_scene_upload) shared by every session. */
int  ve_server_add_scene(ve_server *srv, const uint32_t *words, uint32_t n_words, uint32_t depth, uint32_t *scene_id);
/* Open a session; it is admitted now or queued (VE_OK either way).  Fails with
 * VE_EBUDGET only when the session alone exceeds the whole budget. */
int  ve_server_open(ve_server *srv, const ve_session_desc *desc, uint32_t *session_id);
/* Request one frame (camera UBO as ve_camera_ubo, FrameConfig, inputs whose
 * width/height equal the session's).  Launched at once when the session is
 * admitted, else when it is.  arrival_ms < 0 means now. */
int  ve_server_submit(ve_server *srv, uint32_t session, const ve_frame_config *fc, const ve_frame_inputs *in,
                      const float camera_ubo[16], uint64_t tag, double arrival_ms);
/* Close after the session's submitted frames; its state returns to the pool. */
int  ve_server_close(ve_server *srv, uint32_t session);
/* Collect finished frames (at most max).  wait != 0 blocks until at least one
 * finishes when any is in flight.  Also admits waiting sessions. */
int  ve_server_poll(ve_server *srv, ve_frame_done *out, uint32_t max, uint32_t *n_out, int wait);
/* RGBA32F output of the session's last finished frame (waits for its frames). */
int  ve_server_read_output(ve_server *srv, uint32_t session, float *rgba, size_t bytes);
/* Any buffer of an admitted session's state, as ve_read on a lone context
 * (same ids and layouts; waits for the session's frames). */
int  ve_server_read(ve_server *srv, uint32_t session, ve_buffer_id id, void *dst, size_t bytes);
int  ve_server_get_stats(ve_server *srv, ve_server_stats *out);
/* No counterpart in the reference: the batched and graph-replayed forms of the
 * server.  ve_server_set_graph: sessions replay their frame graphs
 * (ve_set_graph).  ve_server_set_batch: the launch shape of the batched
 * primary pass (ve_batch_config; device and max_views are the server's).
 * ve_server_submit_batch: one frame for each of n distinct sessions (camera
 * UBOs in ubos[16 * i]; tags and arrival times optional, an arrival below 0
 * meaning now) -- the primary passes of the admitted sessions whose earlier
 * frames are done run in one launch (a session still busy goes alone on its
 * stream), the rest of each frame on its session's stream;
 * frames of sessions still waiting for admission queue as with
 * ve_server_submit.  Each frame's result equals ve_server_submit's.  A frame
 * of ve_server_submit_batch waits while the previous batch's primary launch
 * runs (ve_server_poll launches it once that launch has ended, and at once
 * when asked to wait): the frames falling due meanwhile share the next launch.
 * ve_server_set_tuning: ve_set_tuning for every session's context, present
 * and future (jump_level is the server's scenes' and is ignored here). */
int  ve_server_set_graph(ve_server *srv, int enable);
int  ve_server_set_tuning(ve_server *srv, const ve_tuning *t);
/* No counterpart in the reference: placement of ve_server_submit's frames on a
 * server with N >= 2 shared streams.  0 (default): a session keeps the stream
 * it was given.  1: every frame joins the shared stream whose uncollected
 * frames hold the least estimated device time (a moving average of its
 * session's frames, else pixels x (2 + samples x bounces) at the server's
 * recent time per unit); a
 * session's frames still run in order (a frame placed elsewhere waits for its
 * session's previous one on the device).  Chosen before the first session
 * opens (VE_ESTATE after, VE_EINVAL with fewer than two streams). */
int  ve_server_set_placement(ve_server *srv, int by_load);
int  ve_server_set_batch(ve_server *srv, const ve_batch_config *cfg);
int  ve_server_submit_batch(ve_server *srv, uint32_t n, const uint32_t *sessions, const ve_frame_config *fcs,
                            const ve_frame_inputs *ins, const float *ubos, const uint64_t *tags,
                            const double *arrival_ms);

#ifdef __cplusplus
}
#endif
#endif
