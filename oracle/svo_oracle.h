/*
 * svo_oracle.h — CPU restatement of VoxelEngine3.0's SVO ray pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP product path:
 * only tests/, the smoke check and the benchmark's CPU-baseline leg load it.
 * It follows the reference algorithm step by step (each function cites the
 * reference file:line it restates) in plain sequential C, with the FP32
 * evaluation order that Mesa llvmpipe uses for the reference's GLSL:
 * IEEE division and sqrt, no multiply-add contraction, dot(a,b) evaluated as
 * a.x*b.x + (a.y*b.y + a.z*b.z), normalize(v) = v * (1/sqrt(dot(v,v))).
 * Sequential execution reproduces llvmpipe's single-threaded dispatch order
 * (work groups in linear dispatch order, invocations in SIMD-8 vectors), which
 * fixes the order of every atomic append.
 */
#ifndef SVO_ORACLE_H
#define SVO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct so_insert { uint32_t x, y, z, material; } so_insert;

/* OctreeCPU::insert + Octree::set/linearizeNode with a real pointer tree
 * (src/octree_cpu.cpp:40-64, src/renderer/octree.cpp:127-217).  Returns 0 or
 * -1 on allocation failure / -4 when the tree exceeds max_words.
 * *words has *capacity entries; free with so_free. */
int so_octree_build(const so_insert *ops, size_t n, uint32_t depth, uint32_t max_words,
                    uint32_t **words, uint32_t *capacity, uint32_t *size, uint32_t *num_voxels);
void so_free(void *p);

/* Camera::UpdateUBO (src/renderer/camera.cpp:39-50). */
void so_camera_ubo(const float pos[3], const float dir[3], float fov_deg, float aspect, float ubo[16]);

typedef struct so_hit {
    uint32_t hit, id, material, pos[3], size;
} so_hit;

/* RaycastCounted / Raycast (src/shd/internal.glsl:297-400).  `checks` gets q;
 * node_reads (optional) accumulates the node words fetched. */
so_hit so_raycast(const uint32_t *words, uint32_t n_words, uint32_t depth, const float origin[3],
                  const float dir[3], const float inv_dir[3], uint32_t controlchecks,
                  uint32_t *checks, uint64_t *node_reads);

uint32_t so_voxel_key(const uint32_t pos[3]);   /* internal.glsl:263-291 */

/* Select an alternative FP lowering (tests only: they show the golden
 * fixtures reject each one).  0 restores llvmpipe's lowering. */
void so_set_fp_variant(int variant);
void so_fp_probe(int op, const float *in, float *out, size_t n);

/* generate_primary.comp:main over a W x H frame (generate_primary.comp:38-85).
 * gbuf: 4 uint32 per pixel, pos: 1 uint32 per pixel, row-major y*W+x. */
void so_generate_primary(const uint32_t *words, uint32_t n_words, uint32_t depth, const float ubo[16],
                         uint32_t W, uint32_t H, uint32_t controlchecks, uint32_t *gbuf, uint32_t *pos,
                         uint64_t *node_reads, uint64_t *steps, int threads);

/* FP64 twin of Raycast (internal.glsl:351-400) for the primary rays of the
 * given pixels (index y*W+x), beside the FP32 walk.  key*: voxelKey+1 of the
 * hit or 0; steps*: q; diverge_step: first loop test at which the two walks
 * are in different cells (0xFFFFFFFF: never); there, face_dist is the FP64
 * position's distance to the face separating the two cells and pos_diff the
 * largest per-axis difference of the two positions. */
typedef struct so_retrace {
    uint32_t key32, key64, steps32, steps64, diverge_step, pad;
    double face_dist, pos_diff;
} so_retrace;
void so_retrace_f64(const uint32_t *words, uint32_t n_words, uint32_t depth, const float ubo[16], uint32_t W,
                    uint32_t H, uint32_t controlchecks, const uint32_t *pixels, size_t n, so_retrace *out);

/* ---- full wavefront frame --------------------------------------------------- */
typedef struct so_frame_config {
    int32_t primary_controlchecks, bounce_controlchecks, reconstruction_radius, normal_precision;
    int32_t disable_stale_reset;
    uint32_t update_time;
    int32_t shading_budget, sample_cap_direct, sample_cap_indirect;
    float weight_pixels, weight_variance, pixels_norm, variance_norm, top_fraction, mid_fraction;
    int32_t min_samples;
    int32_t restir_enabled, restir_m, restir_temporal, restir_spatial, restir_spatial_neighbors;
    float restir_spatial_radius;
    int32_t restir_max_m;
    float firefly_k, firefly_floor;
    int32_t render_type;
} so_frame_config;

typedef struct so_state so_state;

/* Persistent renderer state: l/n/rBuffer, RankStats thresholds, emissive
 * registry — everything that survives across frames (renderer.cpp:114-161). */
so_state *so_state_create(uint32_t hash_width, uint32_t max_bounces, uint32_t W, uint32_t H);
void so_state_destroy(so_state *s);
void so_state_set_materials(so_state *s, const void *std140, uint32_t n);
void so_state_set_sky(so_state *s, const float rgb[3]);
/* Cube-map sky (skybox.cpp:87-122): six faces +X -X +Y -Y +Z -Z of edge x edge
 * RGB8 texels, rows in upload order, sampled as GL_LINEAR / CLAMP_TO_EDGE on
 * every miss; replaces the constant colour of so_state_set_sky.  Returns 0, or
 * -1 when out of memory. */
int so_state_set_sky_cube(so_state *s, const uint8_t *faces_rgb8, uint32_t edge);
/* texture(skybox, dir).rgb of the current sky for n directions (3 floats each). */
void so_sky_sample(const so_state *s, const float *dirs, size_t n, float *rgb);

/* Renderer::runWavefrontFrame (renderer.cpp:282-649) for one frame. */
int so_frame(so_state *s, const uint32_t *words, uint32_t n_words, uint32_t depth, const float ubo[16],
             const so_frame_config *fc, uint32_t frame_idx, uint32_t frame_seed, uint32_t time);

/* resolve.comp alone (renderer.cpp:612-648) on the current state: the last
 * frame's G-buffer and the l/nBuffer, with `fc`'s render type, primary
 * control checks and reconstruction radius. */
int so_resolve(so_state *s, const uint32_t *words, uint32_t n_words, uint32_t depth, const float ubo[16],
               const so_frame_config *fc);

/* The 96-byte Ray records emit.comp writes for shade-list entries
 * [base, base + n) of the last frame (its camera, frame index and shade
 * list), in ring order; returns the number written. */
size_t so_emit_records(const so_state *s, uint32_t base, uint32_t n, void *out);

/* Buffer access by the same ids as the product's ve_buffer_id. */
size_t so_buffer(so_state *s, int id, void **ptr);
void so_counters(so_state *s, uint32_t out[16]);

#ifdef __cplusplus
}
#endif
#endif
