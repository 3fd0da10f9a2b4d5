/*
 * svo_oracle.c — CPU restatement of VoxelEngine3.0's SVO ray pipeline.
 *
 * TEST INFRASTRUCTURE ONLY (see svo_oracle.h).  Built with
 * -ffp-contract=off so every float operation rounds exactly once, in the
 * order written, as llvmpipe evaluates the reference GLSL.
 */
#define _GNU_SOURCE
#include "svo_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

void so_free(void *p) { free(p); }

/* ======================================================================== */
/* Octree build: pointer tree (src/octree_cpu.cpp) + linearize (octree.cpp)  */
/* ======================================================================== */

typedef struct onode {
    struct onode *children[8];
    uint8_t children_count;
    uint32_t material;
} onode;

/* bump allocator: the pointer tree is freed as a whole */
typedef struct arena {
    char **blocks;
    size_t nblocks, cap_blocks, used;
} arena;
#define ARENA_BLOCK (1u << 20)

static onode *arena_node(arena *a)
{
    const size_t per = ARENA_BLOCK / sizeof(onode);
    if (a->nblocks == 0 || a->used == per) {
        if (a->nblocks == a->cap_blocks) {
            size_t nc = a->cap_blocks ? a->cap_blocks * 2 : 64;
            char **nb = realloc(a->blocks, nc * sizeof(char *));
            if (!nb) return NULL;
            a->blocks = nb;
            a->cap_blocks = nc;
        }
        char *blk = calloc(per, sizeof(onode));
        if (!blk) return NULL;
        a->blocks[a->nblocks++] = blk;
        a->used = 0;
    }
    return (onode *)a->blocks[a->nblocks - 1] + a->used++;
}

static void arena_free(arena *a)
{
    for (size_t i = 0; i < a->nblocks; i++) free(a->blocks[i]);
    free(a->blocks);
}

/* OctreeCPU::locate (octree_cpu.cpp:33-38): utils_p2r[level] = 2^(depth-level) */
static uint32_t cpu_locate(uint32_t x, uint32_t y, uint32_t z, uint32_t p2)
{
    return ((uint32_t)((x & p2) != 0) << 2) | ((uint32_t)((y & p2) != 0) << 1) | (uint32_t)((z & p2) != 0);
}

/* OctreeCPU::insert / insertR (octree_cpu.cpp:40-64), iterative form: the
 * recursion creates missing nodes on the way down and bumps childrenCount of
 * a parent whose child did not exist before. */
static int cpu_insert(arena *a, onode **root, uint32_t depth, const so_insert *op)
{
    uint32_t bound = 1u << depth;
    if (op->x >= bound || op->y >= bound || op->z >= bound) return 0;
    if (!*root) {
        *root = arena_node(a);
        if (!*root) return -1;
    }
    onode *node = *root;
    for (uint32_t level = 1; level <= depth; level++) {
        uint32_t idx = cpu_locate(op->x, op->y, op->z, 1u << (depth - level));
        if (!node->children[idx]) {
            node->children[idx] = arena_node(a);
            if (!node->children[idx]) return -1;
            node->children_count++;
        }
        node = node->children[idx];
    }
    node->material = op->material;     /* level > depth: leaf */
    node->children_count = 0;
    return 0;
}

typedef struct lin_entry { onode *n; uint32_t off; uint32_t lv; } lin_entry;

int so_octree_build(const so_insert *ops, size_t n, uint32_t depth, uint32_t max_words,
                    uint32_t **words_out, uint32_t *capacity_out, uint32_t *size_out, uint32_t *num_voxels_out)
{
    arena a = {0};
    onode *root = NULL;
    for (size_t i = 0; i < n; i++)
        if (cpu_insert(&a, &root, depth, &ops[i])) { arena_free(&a); return -1; }

    /* fresh Octree: capacity 8, size 8 (octree.cpp:4-22, octree.hpp:68) */
    uint32_t capacity = 8, size = 8, num_voxels = 0;
    uint32_t *data = calloc(capacity, 4);
    if (!data) { arena_free(&a); return -1; }
    if (!root) goto done;   /* Octree::set returns early (octree.cpp:128) */

#define RESIZE(req)                                                                 \
    do {                                                                            \
        uint32_t need_ = (req);                                                     \
        if (need_ > capacity) {                                                     \
            uint32_t nc_ = capacity;                                                \
            while (nc_ < need_) nc_ *= 2;                                           \
            if (max_words && nc_ > max_words) nc_ = max_words;                      \
            if (need_ > nc_) { free(data); arena_free(&a); return -4; }             \
            uint32_t *nd_ = realloc(data, (size_t)nc_ * 4);                         \
            if (!nd_) { free(data); arena_free(&a); return -1; }                    \
            memset(nd_ + capacity, 0, (size_t)(nc_ - capacity) * 4);                \
            data = nd_;                                                             \
            capacity = nc_;                                                         \
        }                                                                           \
    } while (0)

    size = 9;                 /* octree.cpp:131 */
    RESIZE(size);
    data[0] = 1u | ((root->children_count & 7u) << 1) | (1u << 4);   /* octree.cpp:142-148 */

    size_t stk_cap = 1024, sp = 0;
    lin_entry *stk = malloc(stk_cap * sizeof(lin_entry));
    if (!stk) { free(data); arena_free(&a); return -1; }
    for (int i = 0; i < 8; i++) {   /* octree.cpp:151-152: root children 0..7, one DFS each */
        sp = 0;
        stk[sp++] = (lin_entry){root->children[i], (uint32_t)(i + 1), 2};
        while (sp) {   /* linearizeNode, octree.cpp:170-217 */
            lin_entry e = stk[--sp];
            if (!e.n) continue;
            if (e.lv > depth || (e.n->material != 0 && e.n->children_count == 0)) {
                if (e.n->material == 0) continue;
                RESIZE(e.off + 1);
                data[e.off] = (e.n->material & 0x7Fu) << 1;
                num_voxels++;
            } else {
                RESIZE(e.off + 1);
                uint32_t block = size;
                if (block >= (1u << 28)) {   /* next is a 28-bit field (octree.hpp:23) */
                    free(stk); free(data); arena_free(&a); return -4;
                }
                size += 8;
                RESIZE(size);
                data[e.off] = 1u | ((e.n->children_count & 7u) << 1) | (block << 4);
                if (sp + 8 > stk_cap) {
                    stk_cap *= 2;
                    lin_entry *ns = realloc(stk, stk_cap * sizeof(lin_entry));
                    if (!ns) { free(stk); free(data); arena_free(&a); return -1; }
                    stk = ns;
                }
                for (int c = 0; c < 8; c++) stk[sp++] = (lin_entry){e.n->children[c], block + (uint32_t)c, e.lv + 1};
            }
        }
    }
    free(stk);
#undef RESIZE
done:
    arena_free(&a);
    *words_out = data;
    *capacity_out = capacity;
    *size_out = size;
    *num_voxels_out = num_voxels;
    return 0;
}

/* ======================================================================== */
/* Camera (src/renderer/camera.cpp:39-50) with glm 0.9.9.8 scalar semantics  */
/* ======================================================================== */

static float glm_dot3(const float a[3], const float b[3])
{
    /* glm compute_dot<vec3>: tmp = a*b; tmp.x + tmp.y + tmp.z */
    float t0 = a[0] * b[0], t1 = a[1] * b[1], t2 = a[2] * b[2];
    return (t0 + t1) + t2;
}

static void glm_cross(const float x[3], const float y[3], float r[3])
{
    r[0] = x[1] * y[2] - y[1] * x[2];
    r[1] = x[2] * y[0] - y[2] * x[0];
    r[2] = x[0] * y[1] - y[0] * x[1];
}

static void glm_normalize(const float v[3], float r[3])
{
    float inv = 1.0f / sqrtf(glm_dot3(v, v));   /* glm::inversesqrt = 1/sqrt */
    r[0] = v[0] * inv;
    r[1] = v[1] * inv;
    r[2] = v[2] * inv;
}

void so_camera_ubo(const float pos[3], const float dir[3], float fov_deg, float aspect, float ubo[16])
{
    const float up_axis[3] = {0.0f, 1.0f, 0.0f};
    float half = fov_deg / 2.0f;
    float t = tanf(half * 0.01745329251994329576923690768489f);   /* glm::radians */
    float c[3], n[3], right[3], u[3];
    glm_cross(up_axis, dir, c);
    glm_normalize(c, n);
    for (int i = 0; i < 3; i++) right[i] = (n[i] * t) * aspect;
    glm_cross(right, dir, c);
    glm_normalize(c, u);
    for (int i = 0; i < 3; i++) {
        ubo[0 + i] = pos[i];
        ubo[4 + i] = dir[i];
        ubo[8 + i] = right[i];
        ubo[12 + i] = u[i] * t;
    }
    ubo[3] = ubo[7] = ubo[11] = ubo[15] = 0.0f;
}

/* ======================================================================== */
/* Traversal (src/shd/internal.glsl)                                         */
/* ======================================================================== */

/* GLSL min/max as llvmpipe evaluates them (probed): a NaN operand yields the
 * other operand, and equal operands (+0/-0) yield the second one. */
static inline float gmin(float a, float b) { return (a < b || isnan(b)) ? a : b; }
static inline float gmax(float a, float b) { return (a > b || isnan(b)) ? a : b; }

/* FP-policy variants, used only by the tests that show the golden fixtures
 * reject each plausible alternative lowering.  0 = llvmpipe's (the oracle). */
static int g_fp_variant = 0;
void so_set_fp_variant(int v) { g_fp_variant = v; }
enum { FPV_LLVMPIPE = 0, FPV_FMA = 1, FPV_DOT_LEFT = 2, FPV_NORM_DIV = 3, FPV_DIV_RCP = 4, FPV_SRC_ORDER = 5 };

/* a*b + c, contracted only under the FMA variant */
static inline float madd(float a, float b, float c) { return g_fp_variant == FPV_FMA ? fmaf(a, b, c) : a * b + c; }

/* GLSL normalize / dot as llvmpipe lowers them. */
static inline float gl_dot3(const float a[3], const float b[3])
{
    if (g_fp_variant == FPV_DOT_LEFT) return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
    return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}

static inline void gl_normalize(float v[3])
{
    float len = sqrtf(gl_dot3(v, v));
    if (g_fp_variant == FPV_NORM_DIV) {
        v[0] /= len;
        v[1] /= len;
        v[2] /= len;
        return;
    }
    float inv = 1.0f / len;
    v[0] *= inv;
    v[1] *= inv;
    v[2] *= inv;
}

static inline int in_bounds(const float v[3], float n)   /* internal.glsl:29 */
{
    return 0.0f <= v[0] && v[0] <= n && 0.0f <= v[1] && v[1] <= n && 0.0f <= v[2] && v[2] <= n;
}

static inline uint32_t gl_locate(const uint32_t p[3], uint32_t p2)   /* internal.glsl:30 */
{
    return ((uint32_t)((p[0] & p2) != 0) << 2) | ((uint32_t)((p[1] & p2) != 0) << 1) | (uint32_t)((p[2] & p2) != 0);
}

static inline uint32_t fetch(const uint32_t *w, uint32_t n, uint32_t i) { return i < n ? w[i] : 0u; }

/* log(x) = log2(x) * ln2, exp(x) = exp2(x * log2e), pow(x, y) =
 * exp2(log2(x) * y) (probed on llvmpipe: these compositions match 100%).
 * The log2 / exp2 / sin / cos cores are evaluated in double and rounded
 * once to float; the GPU kernels use the same definition, so the oracle and
 * the product agree bit for bit, and both stay within a few ulp of
 * llvmpipe's polynomial cores. */
static inline float m_log2(float x) { return (float)log2((double)x); }
static inline float m_exp2(float x) { return (float)exp2((double)x); }
static inline float m_log(float x) { return m_log2(x) * 0.6931471805599453f; }
static inline float m_exp(float x) { return m_exp2(x * 1.4426950408889634f); }
static inline float m_pow(float x, float y) { return m_exp2(m_log2(x) * y); }
static inline float m_cos(float x) { return (float)cos((double)x); }
static inline float m_sin(float x) { return (float)sin((double)x); }
/* Elementwise FP primitives of the GLSL lowering, for comparison with
 * llvmpipe probe outputs.  op 0: dot(v,v) of vec3 in[3i..3i+2]; 1: normalize
 * (3 outputs); 2: a/b; 3: a*b+c; 4: min(a,b); 5: max(a,b); 6-14: inversesqrt,
 * sqrt, pow, exp, log, exp2, log2, sin, cos of the probe shaders' arguments. */
void so_fp_probe(int op, const float *in, float *out, size_t n)
{
    for (size_t i = 0; i < n; i++) {
        const float *a = in + 3 * i;
        switch (op) {
        case 0: out[i] = gl_dot3(a, a); break;
        case 1: {
            float v[3] = {a[0], a[1], a[2]};
            gl_normalize(v);
            out[3 * i] = v[0]; out[3 * i + 1] = v[1]; out[3 * i + 2] = v[2];
            break;
        }
        case 2: out[i] = g_fp_variant == FPV_DIV_RCP ? a[0] * (1.0f / a[1]) : a[0] / a[1]; break;
        case 3: out[i] = madd(a[0], a[1], a[2]); break;
        case 4: out[i] = gmin(a[0], a[1]); break;
        case 5: out[i] = gmax(a[0], a[1]); break;
        /* the built-in probes of tests/golden/make_golden.py, same arguments */
        case 6: out[i] = 1.0f / sqrtf(fabsf(a[0]) + 1e-6f); break;                     /* inversesqrt */
        case 7: out[i] = sqrtf(fabsf(a[0])); break;
        case 8: out[i] = m_pow(fabsf(a[0]) * 0.125f, fabsf(a[1]) * 16.0f); break;
        case 9: out[i] = m_exp(a[0] * a[2]); break;
        case 10: out[i] = m_log(fabsf(a[0]) + 1e-6f); break;
        case 11: out[i] = m_exp2(a[0] * 2.0f); break;
        case 12: out[i] = m_log2(fabsf(a[0]) + 1e-6f); break;
        case 13: out[i] = m_sin(a[0] * a[2]); break;
        case 14: out[i] = m_cos(a[0] * a[2]); break;
        default: out[i] = 0.0f;
        }
    }
}

uint32_t so_voxel_key(const uint32_t pos[3])   /* internal.glsl:263-291 */
{
    uint32_t x = pos[0] | (pos[1] << 10) | (pos[2] << 20);
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x & 0x7FFFFFFFu;
}

/* float -> uint conversion of a value in [0, 2^24] (GLSL uint(float)). */
static inline uint32_t f2u(float f) { return (uint32_t)f; }

/* The positions a traversal tests, in order: one entry per test of the loop
 * condition of internal.glsl:370 (cell = the truncated position, or NONE
 * when the position is out of bounds and ends the walk). */
typedef struct trail {
    uint32_t cap, n;
    uint32_t (*cell)[3];
    double (*pos)[3];
} trail;

static inline void trail_add(trail *t, const double r[3], int inb)
{
    if (!t || t->n >= t->cap) return;
    for (int i = 0; i < 3; i++) {
        t->pos[t->n][i] = r[i];
        t->cell[t->n][i] = inb ? (uint32_t)r[i] : 0xFFFFFFFFu;
    }
    t->n++;
}

static so_hit raycast32(const uint32_t *words, uint32_t n_words, uint32_t depth, const float origin_in[3],
                        const float dir[3], const float inv[3], uint32_t controlchecks, uint32_t *checks_out,
                        uint64_t *node_reads, trail *tr);

so_hit so_raycast(const uint32_t *words, uint32_t n_words, uint32_t depth, const float origin_in[3],
                  const float dir[3], const float inv[3], uint32_t controlchecks, uint32_t *checks_out,
                  uint64_t *node_reads)
{
    return raycast32(words, n_words, depth, origin_in, dir, inv, controlchecks, checks_out, node_reads, NULL);
}

static so_hit raycast32(const uint32_t *words, uint32_t n_words, uint32_t depth, const float origin_in[3],
                        const float dir[3], const float inv[3], uint32_t controlchecks, uint32_t *checks_out,
                        uint64_t *node_reads, trail *tr)
{
    so_hit miss = {0};
    uint32_t p2c[17];
    const uint32_t len = 1u << depth;
    for (uint32_t i = 0; i <= depth; i++) p2c[i] = len >> i;
    const float fl = (float)len;
    uint32_t q = 0;
    uint64_t reads = 0;
    float o[3], r[3];
    for (int i = 0; i < 3; i++) o[i] = origin_in[i] + dir[i] * 4.0f;   /* internal.glsl:361 */

    if (in_bounds(o, fl)) {
        r[0] = o[0]; r[1] = o[1]; r[2] = o[2];
    } else {   /* intersect(ray, 0, N), internal.glsl:245-253 */
        /* llvmpipe's compiler reassociates the constant nudges: (0 - o) + 0.001
         * becomes 0.001 - o (the same value) and (N - o) - 0.001 becomes
         * (N - 0.001) - o */
        float tmn[3], tmx[3];
        for (int i = 0; i < 3; i++) {
            float t1 = (0.001f - o[i]) * inv[i];
            float t2 = (g_fp_variant == FPV_SRC_ORDER ? (fl - o[i]) - 0.001f : (fl - 0.001f) - o[i]) * inv[i];
            tmn[i] = gmin(t1, t2);
            tmx[i] = gmax(t1, t2);
        }
        float te = gmax(gmax(tmn[0], tmn[1]), tmn[2]);
        float tx = gmin(gmin(tmx[0], tmx[1]), tmx[2]);
        q++;
        if (tx < te || tx < 0.0f) {
            if (checks_out) *checks_out = q;
            return miss;
        }
        for (int i = 0; i < 3; i++) r[i] = madd(dir[i], te, o[i]);
    }

    uint32_t tpos[3] = {0, 0, 0}, tsize = 0;
    for (;;) {
        const int inb = in_bounds(r, fl);
        if (tr) {
            const double rd[3] = {r[0], r[1], r[2]};
            trail_add(tr, rd, inb);
        }
        if (!(inb && q++ <= controlchecks)) break;   /* internal.glsl:370 */
        uint32_t ur[3] = {f2u(r[0]), f2u(r[1]), f2u(r[2])};
        uint32_t offset = 0, d = 0;
        int found = 0;
        for (; d < depth; d++) {
            offset += gl_locate(ur, p2c[d]);
            uint32_t raw = fetch(words, n_words, offset);
            reads++;
            if (!(raw & 1u)) {
                tsize = p2c[d];
                for (int i = 0; i < 3; i++) tpos[i] = ur[i] & ~(tsize - 1u);
                found = 1;
                break;
            }
            offset = (raw & 0xFFFFFFF0u) >> 4;
        }
        if (!found) {
            offset += gl_locate(ur, p2c[d]);
            uint32_t raw = fetch(words, n_words, offset);
            reads++;
            tsize = p2c[d];
            for (int i = 0; i < 3; i++) tpos[i] = ur[i] & ~(tsize - 1u);
            uint32_t mat = (raw & 254u) >> 1;
            if (mat != 0) {
                so_hit h;
                h.hit = 1;
                h.id = offset;
                h.material = mat;
                h.pos[0] = tpos[0]; h.pos[1] = tpos[1]; h.pos[2] = tpos[2];
                h.size = tsize;
                if (checks_out) *checks_out = q;
                if (node_reads) *node_reads += reads;
                return h;
            }
        }
        /* intersect_inside(ray, target.position, target.position + size), internal.glsl:255-261 */
        float tmx[3];
        for (int i = 0; i < 3; i++) {
            /* reassociated by llvmpipe's compiler: (pos - 0.001) - o and
             * ((pos + 0.001) + size) - o */
            float bmin = (float)tpos[i];
            float t1 = ((bmin - 0.001f) - o[i]) * inv[i];
            float t2 = (((bmin + 0.001f) + (float)tsize) - o[i]) * inv[i];
            if (g_fp_variant == FPV_SRC_ORDER) {   /* GLSL source order (a test variant) */
                t1 = ((bmin - o[i]) - 0.001f) * inv[i];
                t2 = (((bmin + (float)tsize) - o[i]) + 0.001f) * inv[i];
            }
            tmx[i] = gmax(t1, t2);
        }
        float tx = gmin(gmin(tmx[0], tmx[1]), tmx[2]);
        for (int i = 0; i < 3; i++) r[i] = madd(dir[i], tx, o[i]);
    }
    if (checks_out) *checks_out = q;
    if (node_reads) *node_reads += reads;
    return miss;
}

/* Primary ray direction of pixel (x, y), generate_primary.comp:44-52 */
static void primary_dir(const float ubo[16], uint32_t x, uint32_t y, uint32_t W, uint32_t H, float d[3])
{
    float nx, ny;
    if (g_fp_variant == FPV_DIV_RCP) {
        nx = ((float)x + 0.5f) * (1.0f / (float)W) * 2.0f - 1.0f;
        ny = ((float)y + 0.5f) * (1.0f / (float)H) * 2.0f - 1.0f;
    } else {
        nx = ((float)x + 0.5f) / (float)W * 2.0f - 1.0f;
        ny = ((float)y + 0.5f) / (float)H * 2.0f - 1.0f;
    }
    for (int i = 0; i < 3; i++) d[i] = madd(-ny, ubo[12 + i], madd(nx, ubo[8 + i], ubo[4 + i]));
    gl_normalize(d);
}

/* ---- FP64 twin ------------------------------------------------------------
 * Raycast (internal.glsl:351-400) and the primary ray of
 * generate_primary.comp:44-58 in double precision: the camera widened exactly
 * from float, the same 0.001 nudges, traversal and step limit.  Where the
 * FP32 walk and this one visit different cells, the first such step shows
 * whether FP32 rounding decided it: there the FP64 position lies within a
 * rounding-sized distance of the face separating the two cells. */
static inline int in_bounds64(const double v[3], double n)
{
    return 0.0 <= v[0] && v[0] <= n && 0.0 <= v[1] && v[1] <= n && 0.0 <= v[2] && v[2] <= n;
}

static so_hit raycast64(const uint32_t *words, uint32_t n_words, uint32_t depth, const double origin_in[3],
                        const double dir[3], const double inv[3], uint32_t controlchecks, uint32_t *checks_out,
                        trail *tr)
{
    so_hit miss = {0};
    uint32_t p2c[17];
    const uint32_t len = 1u << depth;
    for (uint32_t i = 0; i <= depth; i++) p2c[i] = len >> i;
    const double fl = (double)len;
    uint32_t q = 0;
    double o[3], r[3];
    for (int i = 0; i < 3; i++) o[i] = origin_in[i] + dir[i] * 4.0;
    if (in_bounds64(o, fl)) {
        r[0] = o[0]; r[1] = o[1]; r[2] = o[2];
    } else {
        double te = -INFINITY, tx = INFINITY;
        for (int i = 0; i < 3; i++) {
            const double t1 = (0.001 - o[i]) * inv[i], t2 = ((fl - 0.001) - o[i]) * inv[i];
            const double mn = t1 < t2 ? t1 : t2, mx = t1 > t2 ? t1 : t2;
            if (mn > te) te = mn;
            if (mx < tx) tx = mx;
        }
        q++;
        if (tx < te || tx < 0.0) {
            if (checks_out) *checks_out = q;
            return miss;
        }
        for (int i = 0; i < 3; i++) r[i] = dir[i] * te + o[i];
    }
    for (;;) {
        const int inb = in_bounds64(r, fl);
        trail_add(tr, r, inb);
        if (!(inb && q++ <= controlchecks)) break;
        const uint32_t ur[3] = {(uint32_t)r[0], (uint32_t)r[1], (uint32_t)r[2]};
        uint32_t offset = 0, d = 0, tsize = 0, tpos[3];
        int found = 0;
        for (; d < depth; d++) {
            offset += gl_locate(ur, p2c[d]);
            const uint32_t raw = fetch(words, n_words, offset);
            if (!(raw & 1u)) {
                found = 1;
                break;
            }
            offset = (raw & 0xFFFFFFF0u) >> 4;
        }
        if (!found) offset += gl_locate(ur, p2c[d]);
        tsize = p2c[d];
        for (int i = 0; i < 3; i++) tpos[i] = ur[i] & ~(tsize - 1u);
        if (!found) {
            const uint32_t mat = (fetch(words, n_words, offset) & 254u) >> 1;
            if (mat != 0) {
                so_hit h = {1u, offset, mat, {tpos[0], tpos[1], tpos[2]}, tsize};
                if (checks_out) *checks_out = q;
                return h;
            }
        }
        double tx = INFINITY;
        for (int i = 0; i < 3; i++) {
            const double bmin = (double)tpos[i];
            const double t1 = ((bmin - 0.001) - o[i]) * inv[i], t2 = (((bmin + 0.001) + (double)tsize) - o[i]) * inv[i];
            const double mx = t1 > t2 ? t1 : t2;
            if (mx < tx) tx = mx;
        }
        for (int i = 0; i < 3; i++) r[i] = dir[i] * tx + o[i];
    }
    if (checks_out) *checks_out = q;
    return miss;
}

/* Distance from FP64 position p to the face that separates its cell from
 * cell c (the largest over the axes where they differ); an out-of-bounds
 * side counts the grid's own faces. */
static double separating_face(const double p[3], const uint32_t c64[3], const uint32_t c32[3], double n)
{
    double dist = 0.0;
    if (c64[0] == 0xFFFFFFFFu || c32[0] == 0xFFFFFFFFu) {   /* one walk left the grid: its nearest face */
        dist = INFINITY;
        for (int i = 0; i < 3; i++) {
            const double a = fabs(p[i]), b = fabs(n - p[i]);
            if (a < dist) dist = a;
            if (b < dist) dist = b;
        }
        return dist;
    }
    for (int i = 0; i < 3; i++) {
        if (c64[i] == c32[i]) continue;
        const double face = c32[i] > c64[i] ? (double)c64[i] + 1.0 : (double)c64[i];
        const double di = fabs(face - p[i]);
        if (di > dist) dist = di;
    }
    return dist;
}

void so_retrace_f64(const uint32_t *words, uint32_t n_words, uint32_t depth, const float ubo[16], uint32_t W,
                    uint32_t H, uint32_t controlchecks, const uint32_t *pixels, size_t n, so_retrace *out)
{
    const uint32_t cap = controlchecks + 4u;
    trail t32 = {cap, 0, malloc(sizeof(uint32_t[3]) * cap), malloc(sizeof(double[3]) * cap)};
    trail t64 = {cap, 0, malloc(sizeof(uint32_t[3]) * cap), malloc(sizeof(double[3]) * cap)};
    const double fl = (double)(1u << depth);
    for (size_t k = 0; k < n; k++) {
        const uint32_t x = pixels[k] % W, y = pixels[k] / W;
        so_retrace *o = out + k;
        float d[3], inv[3];
        primary_dir(ubo, x, y, W, H, d);
        for (int i = 0; i < 3; i++) inv[i] = 1.0f / d[i];
        t32.n = t64.n = 0;
        uint32_t c32 = 0, c64 = 0;
        const so_hit h32 = raycast32(words, n_words, depth, ubo, d, inv, controlchecks, &c32, NULL, &t32);
        const double nx = ((double)x + 0.5) / (double)W * 2.0 - 1.0, ny = ((double)y + 0.5) / (double)H * 2.0 - 1.0;
        double dd[3], ii[3], org[3], l2 = 0.0;
        for (int i = 0; i < 3; i++) {
            dd[i] = (double)ubo[4 + i] + nx * (double)ubo[8 + i] - ny * (double)ubo[12 + i];
            l2 += dd[i] * dd[i];
            org[i] = (double)ubo[i];
        }
        for (int i = 0; i < 3; i++) {
            dd[i] /= sqrt(l2);
            ii[i] = 1.0 / dd[i];
        }
        const so_hit h64 = raycast64(words, n_words, depth, org, dd, ii, controlchecks, &c64, &t64);
        o->key32 = h32.hit ? so_voxel_key(h32.pos) + 1u : 0u;
        o->key64 = h64.hit ? so_voxel_key(h64.pos) + 1u : 0u;
        o->steps32 = c32;
        o->steps64 = c64;
        o->diverge_step = 0xFFFFFFFFu;
        o->face_dist = -1.0;
        o->pos_diff = 0.0;
        const uint32_t m = t32.n < t64.n ? t32.n : t64.n;
        uint32_t s = 0;
        while (s < m && t32.cell[s][0] == t64.cell[s][0] && t32.cell[s][1] == t64.cell[s][1] &&
               t32.cell[s][2] == t64.cell[s][2])
            s++;
        if (s < m) {
            o->diverge_step = s;
            o->face_dist = separating_face(t64.pos[s], t64.cell[s], t32.cell[s], fl);
            for (int i = 0; i < 3; i++) {
                const double e = fabs(t64.pos[s][i] - t32.pos[s][i]);
                if (e > o->pos_diff) o->pos_diff = e;
            }
        } else if (t32.n != t64.n) {   /* one walk stopped at the step limit or a hit the other passed */
            o->diverge_step = s;
        }
    }
    free(t32.cell); free(t32.pos); free(t64.cell); free(t64.pos);
}

typedef struct prim_job {
    const uint32_t *words;
    uint32_t n_words, depth;
    const float *ubo;
    uint32_t W, H, cc, y0, y1;
    uint32_t *gbuf, *pos;
    uint64_t reads, steps;
} prim_job;

static void *primary_rows(void *arg)
{
    prim_job *j = arg;
    const uint32_t bits = j->depth - 1u;
    for (uint32_t y = j->y0; y < j->y1; y++)
        for (uint32_t x = 0; x < j->W; x++) {
            float d[3], inv[3];
            primary_dir(j->ubo, x, y, j->W, j->H, d);
            for (int i = 0; i < 3; i++) inv[i] = 1.0f / d[i];
            uint32_t checks = 0;
            so_hit h = so_raycast(j->words, j->n_words, j->depth, j->ubo, d, inv, j->cc, &checks, &j->reads);
            j->steps += checks;
            size_t p = (size_t)y * j->W + x;
            uint32_t *g = j->gbuf + 4 * p;
            if (h.hit) {
                g[0] = checks; g[1] = h.material; g[2] = 0; g[3] = so_voxel_key(h.pos) + 1u;
                j->pos[p] = (h.pos[0] >> 1) | ((h.pos[1] >> 1) << bits) | ((h.pos[2] >> 1) << (2u * bits));
            } else {
                g[0] = checks; g[1] = g[2] = g[3] = 0;
                j->pos[p] = 0;
            }
        }
    return NULL;
}

void so_generate_primary(const uint32_t *words, uint32_t n_words, uint32_t depth, const float ubo[16],
                         uint32_t W, uint32_t H, uint32_t controlchecks, uint32_t *gbuf, uint32_t *pos,
                         uint64_t *node_reads, uint64_t *steps, int threads)
{
    if (threads < 1) threads = 1;
    if ((uint32_t)threads > H) threads = (int)H;
    if (threads < 1) threads = 1;
    prim_job *jobs = calloc((size_t)threads, sizeof(prim_job));
    pthread_t *tids = calloc((size_t)threads, sizeof(pthread_t));
    for (int t = 0; t < threads; t++) {
        jobs[t] = (prim_job){words, n_words, depth, ubo, W, H, controlchecks,
                             (uint32_t)((uint64_t)H * t / threads), (uint32_t)((uint64_t)H * (t + 1) / threads),
                             gbuf, pos, 0, 0};
        if (threads > 1) pthread_create(&tids[t], NULL, primary_rows, &jobs[t]);
    }
    if (threads == 1) primary_rows(&jobs[0]);
    uint64_t r = 0, s = 0;
    for (int t = 0; t < threads; t++) {
        if (threads > 1) pthread_join(tids[t], NULL);
        r += jobs[t].reads;
        s += jobs[t].steps;
    }
    if (node_reads) *node_reads = r;
    if (steps) *steps = s;
    free(jobs);
    free(tids);
}

/* ======================================================================== */
/* GLSL built-ins as llvmpipe composes them                                  */
/* ======================================================================== */
static inline float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
static inline float gsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

typedef struct v3 { float x, y, z; } v3;
static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vmuls(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 vdivs(v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
static inline float vlength(v3 a) { return sqrtf(vdot(a, a)); }
static inline v3 vnorm(v3 a) { float inv = 1.0f / sqrtf(vdot(a, a)); return vmuls(a, inv); }
static inline v3 vmin(v3 a, v3 b) { return V(gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)); }
static inline v3 vcross(v3 a, v3 b) { return V(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }

/* ======================================================================== */
/* Frame state                                                               */
/* ======================================================================== */

#define LSTRIDE 12
#define LSLOTS 32
#define NSTRIDE 3
#define NSLOTS 32
#define RSTRIDE 8
#define RSLOTS 16
#define RAY_CAP 49152u
#define SHADE_MAX 196608u
#define CLAIM_W 16384u
#define CLAIM_P 64u
#define EM_CAP 4096u
#define EM_CLAIM_W 8192u
#define EM_CLAIM_P 8
#define CAND_CAP 1048576u
#define HBINS 256u

typedef struct ray_t {
    float ox, oy, oz, pad0, dx, dy, dz, pad1, tx, ty, tz;
    uint32_t primary_vid;
    float dlx, dly, dlz;
    uint32_t rng;
    float ilx, ily, ilz;
    uint32_t bounce, hit_vid, hit_material, pad2, pad3;
} ray_t;

typedef struct mat_t {
    float color[4], spec[4], roughness, specular, metallic;
    int32_t emissive;
    float emissive_intensity, pad[3];
} mat_t;

typedef struct emis_t { float ex, ey, ez; uint32_t vid, px, py, pz, size; } emis_t;

struct so_state {
    uint32_t hw, max_bounces, W, H;
    uint32_t *gbuf, *gpos;
    float *out;
    uint32_t *nbuf, *lbuf, *rbuf;
    uint32_t *claim, *samples;
    uint32_t *cands;   /* 4 per candidate */
    uint32_t cand_count;
    uint32_t *shade_list;   /* 2 per entry */
    uint32_t shade_count;
    ray_t *ring;
    uint32_t ring_head, ring_tail;
    emis_t *emissives;
    uint32_t emissive_count;
    uint32_t *emissive_claim;
    uint32_t rank_stats[264];
    mat_t materials[128];
    float sky[3];
    uint8_t *sky_tex;     /* cube-map faces (6 * edge * edge RGB8), NULL for a constant sky */
    uint32_t sky_edge;
    uint32_t counters[16];
    /* frame constants */
    const uint32_t *words;
    uint32_t n_words, depth;
    float cam[16];
    const so_frame_config *fc;
    uint32_t frame_idx, frame_seed, time;
};

so_state *so_state_create(uint32_t hash_width, uint32_t max_bounces, uint32_t W, uint32_t H)
{
    so_state *s = calloc(1, sizeof *s);
    if (!s) return NULL;
    s->hw = hash_width;
    s->max_bounces = max_bounces;
    s->W = W;
    s->H = H;
    size_t P = (size_t)W * H;
    s->gbuf = calloc(P * 4, 4);
    s->gpos = calloc(P, 4);
    s->out = calloc(P * 4, 4);
    s->nbuf = calloc((size_t)hash_width * NSTRIDE * NSLOTS, 4);
    s->lbuf = calloc((size_t)hash_width * LSTRIDE * LSLOTS, 4);
    s->rbuf = calloc((size_t)hash_width * RSTRIDE * RSLOTS, 4);
    s->claim = calloc((size_t)CLAIM_W * CLAIM_P, 4);
    s->samples = calloc((size_t)CLAIM_W * CLAIM_P, 4);
    s->cands = calloc((size_t)CAND_CAP * 4, 4);
    s->shade_list = calloc((size_t)SHADE_MAX * 2, 4);
    s->ring = calloc(RAY_CAP, sizeof(ray_t));
    s->emissives = calloc(EM_CAP, sizeof(emis_t));
    s->emissive_claim = calloc((size_t)EM_CLAIM_W * EM_CLAIM_P, 4);
    s->rank_stats[0] = 0x7F800000u;   /* +inf (rank.comp:109-111) */
    s->rank_stats[1] = 0xFF800000u;   /* -inf */
    if (!s->gbuf || !s->gpos || !s->out || !s->nbuf || !s->lbuf || !s->rbuf || !s->claim || !s->samples ||
        !s->cands || !s->shade_list || !s->ring || !s->emissives || !s->emissive_claim) {
        so_state_destroy(s);
        return NULL;
    }
    return s;
}

void so_state_destroy(so_state *s)
{
    if (!s) return;
    free(s->gbuf); free(s->gpos); free(s->out); free(s->nbuf); free(s->lbuf); free(s->rbuf);
    free(s->claim); free(s->samples); free(s->cands); free(s->shade_list); free(s->ring);
    free(s->emissives); free(s->emissive_claim);
    free(s->sky_tex);
    free(s);
}

void so_state_set_materials(so_state *s, const void *std140, uint32_t n)
{
    memset(s->materials, 0, sizeof s->materials);
    if (n > 127) n = 127;
    memcpy(s->materials, std140, (size_t)n * 64);
}

void so_state_set_sky(so_state *s, const float rgb[3])
{
    memcpy(s->sky, rgb, sizeof s->sky);
    free(s->sky_tex);
    s->sky_tex = NULL;
    s->sky_edge = 0;
}

int so_state_set_sky_cube(so_state *s, const uint8_t *faces_rgb8, uint32_t edge)
{
    size_t n = (size_t)6 * edge * edge * 3;
    uint8_t *t = malloc(n);
    if (!t || edge == 0) {
        free(t);
        return -1;
    }
    memcpy(t, faces_rgb8, n);
    free(s->sky_tex);
    s->sky_tex = t;
    s->sky_edge = edge;
    return 0;
}

/* ---- cube-map sampling: texture(samplerCube, dir) ------------------------
 * The sampler state of skybox.cpp:113-117 (GL_LINEAR min/mag, CLAMP_TO_EDGE,
 * level 0 only, seamless filtering off), as the GL driver the reference runs
 * on evaluates it:
 *  - face: the major axis of the GL spec's cube-map face table (GL 4.5
 *    section 8.13, table 8.19), ties resolved toward z, then y;
 *  - s = (sc / |ma| + 1) / 2, t = (tc / |ma| + 1) / 2;
 *  - texel coordinate u = s * edge - 1/2 taken to 24.8 fixed point with
 *    round-to-nearest-even: integer part = left texel, fraction = an 8-bit
 *    weight; both texels clamped to [0, edge - 1];
 *  - 8-bit texels blended in integers: lerp(a, b, w) = a + (((b - a) w + 128)
 *    >> 8), along s first, then t;
 *  - the blended byte c returned as c * (1/255) in single precision.
 * The driver-specific steps (tie order, fixed-point weights, byte lerps, the
 * final scaling) are pinned by the driver's own outputs over random faces and
 * directions, edge and corner cases included. */
static int lerp8(int a, int b, int w) { return a + (((b - a) * w + 128) >> 8); }

static void cube_axis(float st, uint32_t edge, int *i0, int *i1, int *w)
{
    float u = st * (float)(edge * 256u) - 128.0f;
    int x = (int)nearbyintf(u);
    int lo = x >> 8, hi = lo + 1, top = (int)edge - 1;
    *w = x & 255;
    *i0 = lo < 0 ? 0 : lo > top ? top : lo;
    *i1 = hi < 0 ? 0 : hi > top ? top : hi;
}

static v3 sky_sample(const so_state *s, v3 d)
{
    if (!s->sky_tex) return V(s->sky[0], s->sky[1], s->sky[2]);
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z), sc, tc, ma;
    int face;
    if (az >= ax && az >= ay) {
        face = d.z >= 0.0f ? 4 : 5;
        sc = d.z >= 0.0f ? d.x : -d.x;
        tc = -d.y;
        ma = az;
    } else if (ay >= ax) {
        face = d.y >= 0.0f ? 2 : 3;
        sc = d.x;
        tc = d.y >= 0.0f ? d.z : -d.z;
        ma = ay;
    } else {
        face = d.x >= 0.0f ? 0 : 1;
        sc = d.x >= 0.0f ? -d.z : d.z;
        tc = -d.y;
        ma = ax;
    }
    float fs = (sc / ma + 1.0f) * 0.5f, ft = (tc / ma + 1.0f) * 0.5f;
    int i0, i1, wa, j0, j1, wb;
    cube_axis(fs, s->sky_edge, &i0, &i1, &wa);
    cube_axis(ft, s->sky_edge, &j0, &j1, &wb);
    const uint8_t *f = s->sky_tex + (size_t)face * s->sky_edge * s->sky_edge * 3;
    float out[3];
    for (int c = 0; c < 3; c++) {
        int t00 = f[((size_t)j0 * s->sky_edge + i0) * 3 + c], t10 = f[((size_t)j0 * s->sky_edge + i1) * 3 + c];
        int t01 = f[((size_t)j1 * s->sky_edge + i0) * 3 + c], t11 = f[((size_t)j1 * s->sky_edge + i1) * 3 + c];
        int v = lerp8(lerp8(t00, t10, wa), lerp8(t01, t11, wa), wb);
        out[c] = (float)v * (1.0f / 255.0f);
    }
    return V(out[0], out[1], out[2]);
}

void so_sky_sample(const so_state *s, const float *dirs, size_t n, float *rgb)
{
    for (size_t i = 0; i < n; i++) {
        v3 c = sky_sample(s, V(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
        rgb[3 * i] = c.x;
        rgb[3 * i + 1] = c.y;
        rgb[3 * i + 2] = c.z;
    }
}

/* nBuffer / lBuffer / rBuffer are stored bucket-major: bucket h, field y at
 * h * height + y (the reference's image (x = h, y)). */
#define NB(s, h, y) (s)->nbuf[(size_t)(h) * (NSTRIDE * NSLOTS) + (y)]
#define LB(s, h, y) (s)->lbuf[(size_t)(h) * (LSTRIDE * LSLOTS) + (y)]
#define RB(s, h, y) (s)->rbuf[(size_t)(h) * (RSTRIDE * RSLOTS) + (y)]

/* ======================================================================== */
/* Halton sampler (internal.glsl:57-89)                                      */
/* ======================================================================== */
static const uint32_t HALTON_BASES[16] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53};

static float halton(uint32_t i, uint32_t b)
{
    float f = 1.0f, r = 0.0f;
    for (int k = 0; k < 24 && i > 0u; k++) {
        f /= (float)b;
        uint32_t q = i / b;
        r += f * (float)(i - q * b);
        i = q;
    }
    return r;
}

static float rand_lds(uint32_t *state)
{
    uint32_t dim = *state & 0xFFu;
    uint32_t sample_idx = *state >> 8;
    *state = (sample_idx << 8) | ((dim + 1u) & 0xFFu);
    uint32_t base = HALTON_BASES[dim & 15u];
    uint32_t effective = sample_idx ^ ((dim >> 4) * 0x9E3779B9u);
    return halton(effective, base);
}

static float rand_normal(uint32_t *state)   /* internal.glsl:101-105 */
{
    float theta = (2.0f * 3.1415926f) * rand_lds(state);
    float rho = sqrtf(-2.0f * m_log(rand_lds(state)));
    return rho * m_cos(theta);
}

static v3 random_direction(uint32_t *state)   /* internal.glsl:107-112 */
{
    float x = rand_normal(state);
    float y = rand_normal(state);
    float z = rand_normal(state);
    return vnorm(V(x, y, z));
}

/* EvaluateBRDF (internal.glsl:120-149) */
static v3 evaluate_brdf(v3 Vd, v3 L, v3 N, v3 albedo, float roughness, float metallic, float spec_int)
{
    float NdotL = gmax(vdot(N, L), 0.001f);
    if (NdotL < 0.001f) return V(0, 0, 0);
    float rc = gmax(roughness, 0.01f);
    float spec_exp = 2.0f / (rc * rc);
    v3 H = vnorm(vsub(L, Vd));
    float NdotH = gmax(vdot(N, H), 0.001f);
    float spec_norm = (spec_exp + 2.0f) * 0.15915494f;
    float specular = spec_norm * m_pow(NdotH, spec_exp);
    float kS = spec_int * metallic;
    float kD = 1.0f - kS * 0.5f;
    /* albedo * kD * 0.31831, reassociated by llvmpipe's compiler (a constant
     * factor of a product chain moves next to its first operand) */
    v3 diffuse = g_fp_variant == FPV_SRC_ORDER ? vmuls(vmuls(albedo, kD), 0.31831f) : vmuls(vmuls(albedo, 0.31831f), kD);
    float sp = specular * kS;
    return vmuls(vadd(diffuse, V(sp, sp, sp)), NdotL);
}

static float smith_g1(float NdotX, float a)
{
    float a2 = a * a;
    return 2.0f * NdotX / (NdotX + sqrtf(a2 + (1.0f - a2) * NdotX * NdotX));
}

/* sampleGGXVNDF (internal.glsl:182-219) */
static v3 sample_ggx_vndf(v3 N, v3 Vv, float roughness, uint32_t *state)
{
    float a = gmax(roughness * roughness, 0.001f);
    v3 up = fabsf(N.z) < 0.999f ? V(0, 0, 1) : V(1, 0, 0);
    v3 T = vnorm(vcross(up, N));
    v3 B = vcross(N, T);
    v3 Vl = V(vdot(Vv, T), vdot(Vv, B), vdot(Vv, N));
    v3 Vh = vnorm(V(a * Vl.x, a * Vl.y, Vl.z));
    float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    v3 T1 = lensq > 0.0f ? vmuls(V(-Vh.y, Vh.x, 0.0f), 1.0f / sqrtf(lensq)) : V(1, 0, 0);
    v3 T2 = vcross(Vh, T1);
    float r1 = rand_lds(state);
    float r2 = rand_lds(state);
    float r = sqrtf(r1);
    float phi = 6.2831853f * r2;
    float t1 = r * m_cos(phi);
    float t2 = r * m_sin(phi);
    float sc = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - sc) * sqrtf(gmax(1.0f - t1 * t1, 0.0f)) + sc * t2;
    v3 Nh = vadd(vadd(vmuls(T1, t1), vmuls(T2, t2)), vmuls(Vh, sqrtf(gmax(0.0f, 1.0f - t1 * t1 - t2 * t2))));
    v3 Hl = vnorm(V(a * Nh.x, a * Nh.y, gmax(0.0f, Nh.z)));
    return vadd(vadd(vmuls(T, Hl.x), vmuls(B, Hl.y)), vmuls(N, Hl.z));
}

static v3 ggx_throughput(v3 N, v3 Vv, v3 L, v3 H, float roughness, v3 F0)
{
    float NdotL = gmax(vdot(N, L), 0.001f);
    float VdotH = gmax(vdot(Vv, H), 0.001f);
    float a = roughness * roughness;
    float p5 = m_pow(1.0f - VdotH, 5.0f);
    v3 F = vadd(F0, vmuls(vsub(V(1, 1, 1), F0), p5));
    return vmuls(F, smith_g1(NdotL, a));
}

static v3 reflect3(v3 I, v3 N) { return vsub(I, vmuls(N, 2.0f * vdot(N, I))); }

/* ======================================================================== */
/* normal.comp                                                               */
/* ======================================================================== */

static int lookup_occupied(const so_state *s, uint32_t px, uint32_t py, uint32_t pz, uint32_t len)
{
    if (px >= len || py >= len || pz >= len) return 0;
    uint32_t offset = 0, raw;
    uint32_t p[3] = {px, py, pz};
    for (uint32_t d = 0; d < s->depth; d++) {
        offset += gl_locate(p, len >> d);
        raw = fetch(s->words, s->n_words, offset);
        if (!(raw & 1u)) return ((raw >> 1) & 0x7Fu) != 0;
        offset = (raw >> 4) & 0x0FFFFFFFu;
    }
    offset += gl_locate(p, len >> s->depth);
    raw = fetch(s->words, s->n_words, offset);
    return ((raw >> 1) & 0x7Fu) != 0;
}

static v3 refined_normal(const so_state *s, const uint32_t hp[3], uint32_t len)   /* normal.comp:73-87 */
{
    v3 acc = V(0, 0, 0);
    int r = s->fc->normal_precision;
    for (int dx = -r; dx <= r; dx++)
        for (int dy = -r; dy <= r; dy++)
            for (int dz = -r; dz <= r; dz++) {
                if (dx == 0 && dy == 0 && dz == 0) continue;
                if (dx * dx + dy * dy + dz * dz > r * r) continue;
                int nx = (int)hp[0] + dx, ny = (int)hp[1] + dy, nz = (int)hp[2] + dz;
                int oob = nx < 0 || ny < 0 || nz < 0 || nx >= (int)len || ny >= (int)len || nz >= (int)len;
                int empty = oob || !lookup_occupied(s, (uint32_t)nx, (uint32_t)ny, (uint32_t)nz, len);
                if (empty) acc = vadd(acc, V((float)dx, (float)dy, (float)dz));
            }
    return vlength(acc) > 0.001f ? vnorm(acc) : V(0, 1, 0);
}

static uint32_t pack_normal(v3 n)   /* normal.comp:90-95, round() = round-half-even */
{
    int x = (int)rintf((n.x + 1.0f) * 511.0f), y = (int)rintf((n.y + 1.0f) * 511.0f), z = (int)rintf((n.z + 1.0f) * 511.0f);
    x = x < 0 ? 0 : (x > 1022 ? 1022 : x);
    y = y < 0 ? 0 : (y > 1022 ? 1022 : y);
    z = z < 0 ? 0 : (z > 1022 ? 1022 : z);
    return ((uint32_t)x << 22) | ((uint32_t)y << 12) | ((uint32_t)z << 2);
}

/* normal.comp:124-177 for the eight invocations of one SIMD-8 vector (a row
 * of eight pixels of a work group), in llvmpipe's lockstep order: per
 * iteration of the claim loop every active lane issues its compare-and-swap
 * (lane order), then the lanes that found their slot read its normal,
 * compute and store it if it is empty (lane order) and stamp the time; the
 * lanes that found no slot then all scan the slot times before any of them
 * writes, so lanes of one vector that evict in one bucket pick the same
 * least recently used slot and the last of them owns it. */
static void normal_vec(so_state *s, uint32_t y, uint32_t x0, uint32_t nl)
{
    uint32_t vid[8], h[8], pos[8][3];
    int active[8], evict[8];
    uint32_t len = 1u << s->depth, bits = s->depth - 1u, mask = (1u << bits) - 1u;
    for (uint32_t l = 0; l < nl; l++) {
        size_t pix = (size_t)y * s->W + x0 + l;
        vid[l] = s->gbuf[4 * pix + 3];
        active[l] = vid[l] != 0;
        evict[l] = 0;
        if (!active[l]) continue;
        uint32_t pk = s->gpos[pix];
        pos[l][0] = (pk & mask) * 2u;
        pos[l][1] = ((pk >> bits) & mask) * 2u;
        pos[l][2] = ((pk >> (2u * bits)) & mask) * 2u;
        h[l] = vid[l] % (s->hw - 1u) + 1u;
    }
    for (uint32_t sl = 0; sl < NSLOTS; sl++) {
        int found[8];
        for (uint32_t l = 0; l < nl; l++) {   /* imageAtomicCompSwap(owner, 0, vid), lane order */
            found[l] = 0;
            if (!active[l]) continue;
            uint32_t *own = &NB(s, h[l], sl * NSTRIDE);
            uint32_t prev = *own;
            if (prev == 0) *own = vid[l];
            found[l] = prev == 0 || prev == vid[l];
        }
        uint32_t existing[8];
        for (uint32_t l = 0; l < nl; l++)
            if (found[l]) existing[l] = NB(s, h[l], sl * NSTRIDE + 1);
        for (uint32_t l = 0; l < nl; l++)   /* compSwap(normal, 0, n) in lane order */
            if (found[l] && existing[l] == 0) {
                uint32_t n = pack_normal(refined_normal(s, pos[l], len));
                if (NB(s, h[l], sl * NSTRIDE + 1) == 0) NB(s, h[l], sl * NSTRIDE + 1) = n;
            }
        for (uint32_t l = 0; l < nl; l++)
            if (found[l]) {
                NB(s, h[l], sl * NSTRIDE + 2) = s->time;
                active[l] = 0;
            }
    }
    uint32_t oslot[8];
    for (uint32_t l = 0; l < nl; l++) {   /* the time scans of every evicting lane first */
        if (!active[l]) continue;
        evict[l] = 1;
        uint32_t oldest = 0xFFFFFFFFu;
        oslot[l] = 0;
        for (uint32_t sl = 0; sl < NSLOTS; sl++) {
            uint32_t t = NB(s, h[l], sl * NSTRIDE + 2);
            if (t < oldest) { oldest = t; oslot[l] = sl; }
        }
    }
    for (uint32_t l = 0; l < nl; l++)
        if (evict[l]) NB(s, h[l], oslot[l] * NSTRIDE) = vid[l];
    for (uint32_t l = 0; l < nl; l++)
        if (evict[l]) NB(s, h[l], oslot[l] * NSTRIDE + 1) = pack_normal(refined_normal(s, pos[l], len));
    for (uint32_t l = 0; l < nl; l++)
        if (evict[l]) NB(s, h[l], oslot[l] * NSTRIDE + 2) = s->time;
}

static void normal_pass(so_state *s)   /* 8x8 work groups in dispatch order, one SIMD-8 vector per row */
{
    uint32_t gw = (s->W + 7u) / 8u, gh = (s->H + 7u) / 8u;
    for (uint32_t gy = 0; gy < gh; gy++)
        for (uint32_t gx = 0; gx < gw; gx++)
            for (uint32_t ly = 0; ly < 8u; ly++) {
                uint32_t y = gy * 8u + ly, x0 = gx * 8u;
                if (y >= s->H || x0 >= s->W) continue;
                normal_vec(s, y, x0, s->W - x0 < 8u ? s->W - x0 : 8u);
            }
}

/* 8x8 work groups in dispatch order (x fastest), invocations in local order */
#define FOR_GROUPS_2D(s, X, Y, BODY)                                                     \
    do {                                                                                 \
        uint32_t gw_ = ((s)->W + 7u) / 8u, gh_ = ((s)->H + 7u) / 8u;                    \
        for (uint32_t gy_ = 0; gy_ < gh_; gy_++)                                         \
            for (uint32_t gx_ = 0; gx_ < gw_; gx_++)                                     \
                for (uint32_t ly_ = 0; ly_ < 8u; ly_++)                                  \
                    for (uint32_t lx_ = 0; lx_ < 8u; lx_++) {                            \
                        uint32_t X = gx_ * 8u + lx_, Y = gy_ * 8u + ly_;                 \
                        if (X >= (s)->W || Y >= (s)->H) continue;                        \
                        BODY;                                                            \
                    }                                                                    \
    } while (0)

/* ======================================================================== */
/* schedule.comp (SIMD-8 lockstep order of the claim loop)                   */
/* ======================================================================== */

static uint32_t claim_hash(uint32_t v)
{
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    return v;
}

static uint32_t bit_reverse(uint32_t x)
{
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0F0F0F0Fu) | ((x & 0x0F0F0F0Fu) << 4);
    x = ((x >> 8) & 0x00FF00FFu) | ((x & 0x00FF00FFu) << 8);
    return (x >> 16) | (x << 16);
}

static uint32_t find_msb(uint32_t v) { uint32_t r = 0; while (v >>= 1) r++; return r; }

static void schedule_pass(so_state *s)
{
    uint32_t gw = (s->W + 7u) / 8u, gh = (s->H + 7u) / 8u, total = gw * gh;
    uint32_t log2g = total <= 1u ? 1u : find_msb(total - 1u) + 1u;
    for (uint32_t g = 0; g < total; g++) {
        uint32_t sc = bit_reverse(g ^ s->frame_seed) >> (32u - log2g);
        if (sc >= total) sc = g;
        uint32_t gx = sc % gw, gy = sc / gw;
        for (uint32_t ly = 0; ly < 8; ly++) {   /* one SIMD-8 vector per tile row */
            int active[8];
            uint32_t vid[8], h[8], pixi[8];
            for (uint32_t l = 0; l < 8; l++) {
                uint32_t x = gx * 8u + l, y = gy * 8u + ly;
                active[l] = 0;
                if (x >= s->W || y >= s->H) continue;
                vid[l] = s->gbuf[4 * ((size_t)y * s->W + x) + 3];
                if (vid[l] == 0) continue;
                h[l] = claim_hash(vid[l]) % CLAIM_W;
                pixi[l] = y * s->W + x;
                active[l] = 1;
            }
            for (uint32_t p = 0; p < CLAIM_P; p++) {
                uint32_t prev[8];
                int any = 0;
                for (uint32_t l = 0; l < 8; l++) {
                    if (!active[l]) continue;
                    uint32_t *c = &s->claim[(size_t)h[l] * CLAIM_P + p];
                    prev[l] = *c;
                    if (*c == 0) *c = vid[l];
                    any = 1;
                }
                if (!any) break;
                for (uint32_t l = 0; l < 8; l++) {
                    if (!active[l]) continue;
                    if (prev[l] == 0) {
                        uint32_t idx = s->cand_count++;
                        if (idx < CAND_CAP) {
                            uint32_t *cd = &s->cands[4 * (size_t)idx];
                            cd[0] = vid[l]; cd[1] = pixi[l]; cd[2] = h[l]; cd[3] = p;
                        } else {
                            s->cand_count--;
                        }
                        s->samples[(size_t)h[l] * CLAIM_P + p]++;
                        active[l] = 0;
                    } else if (prev[l] == vid[l]) {
                        s->samples[(size_t)h[l] * CLAIM_P + p]++;
                        active[l] = 0;
                    }
                }
            }
        }
    }
}

/* ======================================================================== */
/* rank.comp + threshold.comp                                                */
/* ======================================================================== */

static float indirect_variance(const so_state *s, uint32_t vid)   /* rank.comp:66-87 */
{
    uint32_t h = vid % (s->hw - 1u) + 1u;
    for (uint32_t sl = 0; sl < LSLOTS; sl++) {
        uint32_t i = sl * LSTRIDE;
        uint32_t owner = LB(s, h, i);
        if (owner == 0) return 0.0f;
        if (owner != vid) continue;
        uint32_t c = LB(s, h, i + 6);
        if (c < 2u) return 0.0f;
        uint32_t rs = LB(s, h, i + 7), gs = LB(s, h, i + 8), bs = LB(s, h, i + 9), l2s = LB(s, h, i + 10);
        float fc = (float)c;
        float meanL = (float)(rs + gs + bs) / (3.0f * fc);
        float meanL2 = (float)l2s / fc;
        return gmax(meanL2 - meanL * meanL, 0.0f);
    }
    return 0.0f;
}

static void rank_pass(so_state *s, uint32_t budget)
{
    const so_frame_config *fc = s->fc;
    uint32_t n = s->cand_count < CAND_CAP ? s->cand_count : CAND_CAP;
    float tHigh, tMid;
    memcpy(&tHigh, &s->rank_stats[0], 4);
    memcpy(&tMid, &s->rank_stats[1], 4);
    for (uint32_t v0 = 0; v0 < n; v0 += 8) {
        int samples[8];
        int cnt = 0;
        for (uint32_t l = 0; l < 8; l++) {
            uint32_t t = v0 + l;
            samples[l] = 0;
            if (t >= n) continue;
            const uint32_t *cd = &s->cands[4 * (size_t)t];
            uint32_t pixels = s->samples[(size_t)cd[2] * CLAIM_P + cd[3]];
            float var = indirect_variance(s, cd[0]);
            float pn = gclamp((float)pixels / gmax(fc->pixels_norm, 1.0f), 0.0f, 1.0f);
            float vn = gclamp(var / gmax(fc->variance_norm, 1.0f), 0.0f, 1.0f);
            float imp = fc->weight_pixels * pn + fc->weight_variance * vn;
            int sm;
            if (imp >= tHigh) sm = 2;
            else if (imp >= tMid) sm = 1;
            else sm = fc->min_samples;
            samples[l] = sm;
            cnt = 1;
            uint32_t bin = (uint32_t)(imp * 256.0f);
            if (bin >= HBINS) bin = HBINS - 1u;
            s->rank_stats[8 + bin]++;
            uint32_t ib;
            memcpy(&ib, &imp, 4);
            if (ib < s->rank_stats[4]) s->rank_stats[4] = ib;
            if (ib > s->rank_stats[5]) s->rank_stats[5] = ib;
            s->rank_stats[6]++;
        }
        (void)cnt;
        /* the sample loop runs in lockstep: iteration i for every lane with samples > i */
        int done[8] = {0};
        for (int i = 0; i < 2 || i < fc->min_samples; i++) {
            for (uint32_t l = 0; l < 8; l++) {
                uint32_t t = v0 + l;
                if (t >= n || done[l] || i >= samples[l]) continue;
                uint32_t idx = s->shade_count++;
                if (idx < budget) {
                    s->shade_list[2 * (size_t)idx] = s->cands[4 * (size_t)t];
                    s->shade_list[2 * (size_t)idx + 1] = s->cands[4 * (size_t)t + 1];
                } else {
                    s->shade_count--;
                    done[l] = 1;
                }
            }
        }
    }
}

static void threshold_pass(so_state *s)   /* threshold.comp:31-78 */
{
    const so_frame_config *fc = s->fc;
    uint32_t total = 0;
    for (uint32_t b = 0; b < HBINS; b++) total += s->rank_stats[8 + b];
    if (total == 0) {
        s->rank_stats[0] = 0x7F800000u;
        s->rank_stats[1] = 0xFF800000u;
        return;
    }
    float topF = gclamp(fc->top_fraction, 0.0f, 1.0f);
    float midF = gclamp(fc->mid_fraction, 0.0f, 1.0f - topF);
    uint32_t topCount = (uint32_t)(topF * (float)total);
    uint32_t topMidCount = (uint32_t)((topF + midF) * (float)total);
    uint32_t cum = 0, thb = HBINS, tmb = HBINS;
    for (int b = (int)HBINS - 1; b >= 0; b--) {
        cum += s->rank_stats[8 + b];
        if (thb == HBINS && cum >= topCount) thb = (uint32_t)b;
        if (tmb == HBINS && cum >= topMidCount) tmb = (uint32_t)b;
        if (thb != HBINS && tmb != HBINS) break;
    }
    if (thb == HBINS) thb = HBINS - 1u;
    if (tmb == HBINS) tmb = HBINS - 1u;
    float th = (float)thb / 256.0f, tm = (float)tmb / 256.0f;
    memcpy(&s->rank_stats[0], &th, 4);
    memcpy(&s->rank_stats[1], &tm, 4);
}

/* ======================================================================== */
/* emit / trace / shade                                                      */
/* ======================================================================== */

static v3 pixel_dir(const so_state *s, uint32_t px, uint32_t py)
{
    float d[3];
    primary_dir(s->cam, px, py, s->W, s->H, d);
    return V(d[0], d[1], d[2]);
}

/* The Ray record emit.comp:55-92 writes for shade-list entry li. */
static ray_t emit_record(const so_state *s, uint32_t li)
{
    uint32_t vid = s->shade_list[2 * (size_t)li], pix = s->shade_list[2 * (size_t)li + 1];
    uint32_t px = pix % s->W, py = pix / s->W;
    v3 d = pixel_dir(s, px, py);
    ray_t r;
    memset(&r, 0, sizeof r);
    r.ox = s->cam[0]; r.oy = s->cam[1]; r.oz = s->cam[2];
    r.dx = d.x; r.dy = d.y; r.dz = d.z;
    r.tx = r.ty = r.tz = 1.0f;
    r.primary_vid = vid;
    uint32_t sample_idx = s->frame_idx + pix * 2654435761u + li * 1442695040u;
    r.rng = (sample_idx << 8) & 0xFFFFFF00u;
    return r;
}

static void emit_pass(so_state *s, uint32_t chunk_base, uint32_t chunk_size)   /* emit.comp:39-100 */
{
    for (uint32_t t = 0; t < chunk_size; t++) {
        /* one atomicAdd(ring_head) per invocation, in invocation order */
        uint32_t slot = s->ring_head++;
        s->ring[slot % RAY_CAP] = emit_record(s, chunk_base + t);
    }
}

size_t so_emit_records(const so_state *s, uint32_t base, uint32_t n, void *out)
{
    if (base >= s->shade_count) return 0;
    if (n > s->shade_count - base) n = s->shade_count - base;
    ray_t *o = (ray_t *)out;
    for (uint32_t t = 0; t < n; t++) o[t] = emit_record(s, base + t);
    return n;
}

static void trace_pass(so_state *s, uint32_t base, uint32_t size)   /* trace.comp:25-70 */
{
    for (uint32_t t = 0; t < size; t++) {
        ray_t *r = &s->ring[(base + t) % RAY_CAP];
        if (r->bounce >= s->max_bounces) continue;
        float o[3] = {r->ox, r->oy, r->oz}, d[3] = {r->dx, r->dy, r->dz};
        float inv[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
        uint32_t checks;
        so_hit h = so_raycast(s->words, s->n_words, s->depth, o, d, inv, (uint32_t)s->fc->bounce_controlchecks,
                              &checks, NULL);
        s->counters[4]++;
        if (h.hit) {
            r->ox = (float)h.pos[0]; r->oy = (float)h.pos[1]; r->oz = (float)h.pos[2];
            r->pad0 = (float)h.size;
            r->hit_vid = so_voxel_key(h.pos) + 1u;
            r->hit_material = h.material;
        } else {
            r->hit_vid = 0;
            r->hit_material = 0;
        }
    }
}

static void pack_rgb_l2(v3 c, uint32_t out[4])   /* shade.comp:154-160 */
{
    out[0] = (uint32_t)(gclamp(c.x, 0.0f, 1024.0f) * 255.0f);
    out[1] = (uint32_t)(gclamp(c.y, 0.0f, 1024.0f) * 255.0f);
    out[2] = (uint32_t)(gclamp(c.z, 0.0f, 1024.0f) * 255.0f);
    uint32_t L = (out[0] + out[1] + out[2]) / 3u;
    out[3] = L * L;
}

static void reset_lslot(so_state *s, uint32_t h, uint32_t i, uint32_t vid, const uint32_t d[4], const uint32_t e[4])
{
    LB(s, h, i) = vid;
    LB(s, h, i + 1) = 1; LB(s, h, i + 2) = d[0]; LB(s, h, i + 3) = d[1]; LB(s, h, i + 4) = d[2]; LB(s, h, i + 5) = d[3];
    LB(s, h, i + 6) = 1; LB(s, h, i + 7) = e[0]; LB(s, h, i + 8) = e[1]; LB(s, h, i + 9) = e[2]; LB(s, h, i + 10) = e[3];
    LB(s, h, i + 11) = s->time;
}

static void firefly(const so_state *s, uint32_t h, uint32_t i, uint32_t cnt_field, uint32_t v[4])
{
    uint32_t c = LB(s, h, i + cnt_field);
    if (c <= 4u) return;
    uint32_t rs = LB(s, h, i + cnt_field + 1), gs = LB(s, h, i + cnt_field + 2), bs = LB(s, h, i + cnt_field + 3);
    float meanL = (float)(rs + gs + bs) / (float)(3u * c);
    float sampleL = (float)(v[0] + v[1] + v[2]) / 3.0f;
    float thresh = gmax(meanL * s->fc->firefly_k, s->fc->firefly_floor);
    if (sampleL > thresh) {
        float scale = thresh / sampleL;
        v[0] = (uint32_t)((float)v[0] * scale);
        v[1] = (uint32_t)((float)v[1] * scale);
        v[2] = (uint32_t)((float)v[2] * scale);
        uint32_t L = (v[0] + v[1] + v[2]) / 3u;
        v[3] = L * L;
    }
}

/* depositSampleDual (shade.comp:162-257) for the lanes of one SIMD-8 vector
 * that reach the same call site together, in llvmpipe's lockstep order: per
 * iteration of the slot loop every active lane issues its compare-and-swap
 * (lane order), then every lane loads the slot's owner and time, then the
 * lanes whose id matches deposit (each step of the deposit body for all of
 * them before the next step: halving loads before halving stores, firefly
 * loads before any add) and leave the loop, the others fold the loaded time
 * into their eviction candidate; lanes that never match reset their
 * candidate slot after the loop, in lane order.  Lanes of one vector thus
 * see each other's claims of this iteration but not the time stamps their
 * deposits write later in it: a lane can evict the slot another lane of its
 * vector has just claimed (its loaded time is still the old one).  With one
 * lane this is the sequential deposit. */
static void deposit_vec(so_state *s, uint32_t n, const uint32_t *vid, const v3 *dir_rgb, const v3 *ind_rgb)
{
    uint32_t d[8][4], e[8][4], h[8], cand_t[8], cand_i[8];
    int active[8];
    for (uint32_t l = 0; l < n; l++) {
        active[l] = vid[l] != 0;
        if (!active[l]) continue;
        pack_rgb_l2(dir_rgb[l], d[l]);
        pack_rgb_l2(ind_rgb[l], e[l]);
        h[l] = vid[l] % (s->hw - 1u) + 1u;
        cand_t[l] = 0xFFFFFFFFu;
        cand_i[l] = 0;
    }
    for (uint32_t i = 0; i < LSTRIDE * LSLOTS; i += LSTRIDE) {
        uint32_t id[8], st[8];
        int match[8];
        for (uint32_t l = 0; l < n; l++)   /* imageAtomicCompSwap(owner, 0, vid), lane order */
            if (active[l] && LB(s, h[l], i) == 0) LB(s, h[l], i) = vid[l];
        for (uint32_t l = 0; l < n; l++) {   /* owner and time loads */
            if (!active[l]) continue;
            id[l] = LB(s, h[l], i);
            st[l] = LB(s, h[l], i + 11);
        }
        int any = 0;
        for (uint32_t l = 0; l < n; l++) {
            match[l] = 0;
            if (!active[l]) continue;
            if (id[l] != vid[l]) {
                if (st[l] < cand_t[l]) { cand_t[l] = st[l]; cand_i[l] = i; }
                continue;
            }
            match[l] = 1;
            any = 1;
        }
        if (!any) continue;
        /* stale reset (shade.comp:183-186): those lanes leave */
        for (uint32_t l = 0; l < n; l++) {
            if (!match[l]) continue;
            if (!s->fc->disable_stale_reset && st[l] != 0 && st[l] + s->fc->update_time < s->time) {
                reset_lslot(s, h[l], i, vid[l], d[l], e[l]);
                match[l] = 0;
                active[l] = 0;
            }
        }
        /* halving at the caps (shade.comp:188-194): every count load, then
         * the halving lanes' field loads, then their stores */
        for (int ch = 0; ch < 2; ch++) {
            uint32_t f0 = ch == 0 ? 1u : 6u, cap = ch == 0 ? (uint32_t)s->fc->sample_cap_direct
                                                          : (uint32_t)s->fc->sample_cap_indirect;
            uint32_t cnt[8], old[8][5];
            for (uint32_t l = 0; l < n; l++)
                if (match[l]) cnt[l] = LB(s, h[l], i + f0);
            for (uint32_t l = 0; l < n; l++)
                if (match[l] && cnt[l] >= cap)
                    for (uint32_t f = 0; f < 5; f++) old[l][f] = LB(s, h[l], i + f0 + f);
            for (uint32_t f = 0; f < 5; f++)
                for (uint32_t l = 0; l < n; l++)
                    if (match[l] && cnt[l] >= cap) LB(s, h[l], i + f0 + f) = old[l][f] >> 1;
        }
        /* firefly clamps (shade.comp:201-234) read the sums before any lane adds */
        for (uint32_t l = 0; l < n; l++)
            if (match[l]) firefly(s, h[l], i, 1, d[l]);
        for (uint32_t l = 0; l < n; l++)
            if (match[l]) firefly(s, h[l], i, 6, e[l]);
        for (uint32_t l = 0; l < n; l++) {   /* the adds commute; time last */
            if (!match[l]) continue;
            LB(s, h[l], i + 1) += 1; LB(s, h[l], i + 2) += d[l][0]; LB(s, h[l], i + 3) += d[l][1];
            LB(s, h[l], i + 4) += d[l][2]; LB(s, h[l], i + 5) += d[l][3];
            LB(s, h[l], i + 6) += 1; LB(s, h[l], i + 7) += e[l][0]; LB(s, h[l], i + 8) += e[l][1];
            LB(s, h[l], i + 9) += e[l][2]; LB(s, h[l], i + 10) += e[l][3];
        }
        for (uint32_t l = 0; l < n; l++)
            if (match[l]) {
                LB(s, h[l], i + 11) = s->time;
                active[l] = 0;
            }
    }
    for (uint32_t l = 0; l < n; l++)   /* bucket full: reset the candidate (shade.comp:254-256) */
        if (active[l] && cand_t[l] != 0xFFFFFFFFu) reset_lslot(s, h[l], cand_i[l], vid[l], d[l], e[l]);
}

static v3 unpack_normal(uint32_t p)
{
    return V((float)((p >> 22) & 0x3FFu) / 511.0f - 1.0f, (float)((p >> 12) & 0x3FFu) / 511.0f - 1.0f,
             (float)((p >> 2) & 0x3FFu) / 511.0f - 1.0f);
}

static v3 lookup_normal(const so_state *s, uint32_t vid, v3 rd)   /* shade.comp:261-275 */
{
    uint32_t h = vid % (s->hw - 1u) + 1u;
    for (uint32_t sl = 0; sl < NSLOTS; sl++) {
        if (NB(s, h, sl * NSTRIDE) == vid) {
            uint32_t nf = NB(s, h, sl * NSTRIDE + 1);
            if (nf != 0) return vnorm(unpack_normal(nf));
        }
    }
    v3 a = V(fabsf(rd.x), fabsf(rd.y), fabsf(rd.z));
    if (a.x > a.y && a.x > a.z) return V(-gsign(rd.x), 0, 0);
    if (a.y > a.z) return V(0, -gsign(rd.y), 0);
    return V(0, 0, -gsign(rd.z));
}

/* ======================================================================== */
/* Next-event estimation: ReSTIR DI (bounce 0) and single-pick RIS          */
/* (shade.comp:276-539, 620-836)                                             */
/* ======================================================================== */
typedef struct reservoir {
    uint32_t light_idx, light_vid, face_seed;
    float wsum;
    uint32_t m;
    float target_pdf;
} reservoir;

static const v3 LUMA = {0.2126f, 0.7152f, 0.0722f};

static void r_update(reservoir *r, uint32_t idx, uint32_t lvid, uint32_t seed, float target, float w,
                     uint32_t *rng)   /* shade.comp:320-336 */
{
    if (w <= 0.0f) { r->m += 1u; return; }
    r->wsum += w;
    r->m += 1u;
    if (rand_lds(rng) * r->wsum < w) {
        r->light_idx = idx; r->light_vid = lvid; r->face_seed = seed; r->target_pdf = target;
    }
}

static void r_combine(reservoir *a, const reservoir *b, float target_at_us, uint32_t *rng)   /* shade.comp:352-365 */
{
    if (b->m == 0u || b->target_pdf <= 0.0f || target_at_us <= 0.0f) return;
    float w = target_at_us * b->wsum / b->target_pdf;
    a->wsum += w;
    a->m += b->m;
    if (rand_lds(rng) * a->wsum < w) {
        a->light_idx = b->light_idx; a->light_vid = b->light_vid; a->face_seed = b->face_seed;
        a->target_pdf = target_at_us;
    }
}

static void r_cap(reservoir *r, uint32_t max_m)   /* shade.comp:372-377 */
{
    if (r->m > max_m && max_m > 0u) {
        r->wsum = r->wsum * (float)max_m / (float)r->m;
        r->m = max_m;
    }
}

static uint32_t pack_face_seed(uint32_t face, float u, float v)   /* shade.comp:382-386 */
{
    uint32_t ui = f2u(gclamp(u, 0.0f, 0.9999f) * 16383.0f), vi = f2u(gclamp(v, 0.0f, 0.9999f) * 16383.0f);
    return (face & 0xFu) | ((ui & 0x3FFFu) << 4) | ((vi & 0x3FFFu) << 18);
}

/* face frame of the six-face construction (shade.comp:398-413, 761-771) */
static void face_frame(uint32_t face, v3 *n, v3 *t, v3 *b)
{
    if (face < 2u) { *n = V(face == 0u ? 1.0f : -1.0f, 0, 0); *t = V(0, 1, 0); *b = V(0, 0, 1); }
    else if (face < 4u) { *n = V(0, face == 2u ? 1.0f : -1.0f, 0); *t = V(1, 0, 0); *b = V(0, 0, 1); }
    else { *n = V(0, 0, face == 4u ? 1.0f : -1.0f); *t = V(1, 0, 0); *b = V(0, 1, 0); }
}

static void light_sample_point(uint32_t seed, const emis_t *e, v3 *P, v3 *N)   /* shade.comp:393-420 */
{
    uint32_t face = seed & 0xFu;
    float u = (float)((seed >> 4) & 0x3FFFu) / 16383.0f, v = (float)((seed >> 18) & 0x3FFFu) / 16383.0f;
    if (face > 5u) face = 5u;
    v3 fn, ft, fb;
    face_frame(face, &fn, &ft, &fb);
    float half = (float)e->size * 0.5f;
    v3 c = vadd(V((float)e->px, (float)e->py, (float)e->pz), V(half, half, half));
    *P = vadd(vadd(vadd(c, vmuls(fn, half)), vmuls(ft, (u - 0.5f) * 2.0f * half)), vmuls(fb, (v - 0.5f) * 2.0f * half));
    *N = fn;
}

static float target_pdf(const emis_t *e, uint32_t seed, v3 sp, v3 nrm)   /* shade.comp:423-438 */
{
    v3 P, N;
    light_sample_point(seed, e, &P, &N);
    v3 to = vsub(P, sp);
    float d2 = vdot(to, to);
    if (d2 < 1e-6f) return 0.0f;
    float d = sqrtf(d2);
    v3 ld = vdivs(to, d);
    float cx = vdot(nrm, ld);
    if (cx <= 0.0f) return 0.0f;
    float cy = vdot(V(-ld.x, -ld.y, -ld.z), N);
    if (cy <= 0.0f) return 0.0f;
    float lum = gmax(vdot(V(e->ex, e->ey, e->ez), LUMA), 1e-6f);
    float area = 6.0f * (float)(e->size * e->size);
    return lum * cy * area / d2;
}

/* rBufferLoad (shade.comp:466-485) over the reference-layout image */
static int rbuf_load(const so_state *s, uint32_t vid, reservoir *r)
{
    memset(r, 0, sizeof *r);
    uint32_t row = vid % (s->hw - 1u) + 1u;
    for (uint32_t k = 0; k < RSLOTS; k++) {
        if (RB(s, row, k * RSTRIDE) != vid) continue;
        r->light_idx = RB(s, row, k * RSTRIDE + 1);
        r->light_vid = RB(s, row, k * RSTRIDE + 2);
        r->face_seed = RB(s, row, k * RSTRIDE + 3);
        memcpy(&r->wsum, &RB(s, row, k * RSTRIDE + 4), 4);
        r->m = RB(s, row, k * RSTRIDE + 5);
        memcpy(&r->target_pdf, &RB(s, row, k * RSTRIDE + 6), 4);
        uint32_t t = RB(s, row, k * RSTRIDE + 7);
        if (!s->fc->disable_stale_reset && t != 0u && t + s->fc->update_time < s->time) {
            memset(r, 0, sizeof *r);
            return 0;
        }
        return 1;
    }
    return 0;
}

static void rbuf_store_at(so_state *s, uint32_t row, uint32_t k, uint32_t vid, const reservoir *r)
{
    RB(s, row, k * RSTRIDE) = vid;
    RB(s, row, k * RSTRIDE + 1) = r->light_idx;
    RB(s, row, k * RSTRIDE + 2) = r->light_vid;
    RB(s, row, k * RSTRIDE + 3) = r->face_seed;
    memcpy(&RB(s, row, k * RSTRIDE + 4), &r->wsum, 4);
    RB(s, row, k * RSTRIDE + 5) = r->m;
    memcpy(&RB(s, row, k * RSTRIDE + 6), &r->target_pdf, 4);
    RB(s, row, k * RSTRIDE + 7) = s->time;
}

static void rbuf_store(so_state *s, uint32_t vid, const reservoir *r)   /* rBufferStore, shade.comp:501-519 */
{
    uint32_t row = vid % (s->hw - 1u) + 1u, oldest = 0, oldest_t = 0xFFFFFFFFu;
    for (uint32_t k = 0; k < RSLOTS; k++) {
        uint32_t prev = RB(s, row, k * RSTRIDE);
        if (prev == 0u) RB(s, row, k * RSTRIDE) = vid;
        if (prev == 0u || prev == vid) { rbuf_store_at(s, row, k, vid, r); return; }
        uint32_t t = RB(s, row, k * RSTRIDE + 7);
        if (t < oldest_t) { oldest_t = t; oldest = k; }
    }
    rbuf_store_at(s, row, oldest, vid, r);
}

static int light_valid(const reservoir *r, const so_state *s, uint32_t count)   /* shade.comp:446-449 */
{
    if (r->m == 0u || r->light_idx >= count) return 0;
    return s->emissives[r->light_idx].vid == r->light_vid;
}

static uint32_t neighbour_vid(const uint32_t hp[3], uint32_t size, const int off[3])   /* shade.comp:527-535 */
{
    int q[3];
    for (int i = 0; i < 3; i++) {
        q[i] = (int)hp[i] + off[i] * (int)size;
        if (q[i] < 0) return 0u;
    }
    uint32_t sn[3];
    for (int i = 0; i < 3; i++) sn[i] = ((uint32_t)q[i] & ~(size - 1u)) & 0x3FFu;
    return so_voxel_key(sn) + 1u;
}

/* RaycastVisibilityToVoxel (internal.glsl:406-410), bounce check budget */
static int visible_to(so_state *s, v3 o, v3 d, uint32_t target)
{
    float org[3] = {o.x, o.y, o.z}, dir[3] = {d.x, d.y, d.z};
    float inv[3] = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    uint32_t checks = 0;
    s->counters[5]++;
    so_hit h = so_raycast(s->words, s->n_words, s->depth, org, dir, inv, (uint32_t)s->fc->bounce_controlchecks,
                          &checks, NULL);
    return h.hit && so_voxel_key(h.pos) + 1u == target;
}

/* Per-lane NEE state of one SIMD-8 vector: the reservoir to store once every
 * lane of the vector has finished its reads (shade.comp:812-814). */
typedef struct nee_store { int pending; uint32_t vid; reservoir res; } nee_store;

static void nee_restir(so_state *s, ray_t *r, const mat_t *m, v3 nrm, const uint32_t hp[3], uint32_t hsize,
                       uint32_t n_lights, uint32_t count, nee_store *st)   /* shade.comp:646-815 */
{
    v3 sp = vadd(vadd(V((float)hp[0], (float)hp[1], (float)hp[2]), V(0.5f, 0.5f, 0.5f)), nrm);
    reservoir res;
    memset(&res, 0, sizeof res);
    uint32_t pvid = r->primary_vid;
    int nm = s->fc->restir_m > 1 ? s->fc->restir_m : 1;
    float inv_src = (float)n_lights * 6.0f;
    for (int k = 0; k < nm; k++) {
        uint32_t pick = f2u(rand_lds(&r->rng) * (float)n_lights);
        if (pick >= n_lights) pick = n_lights - 1u;
        const emis_t *c = &s->emissives[pick];
        uint32_t face = f2u(rand_lds(&r->rng) * 6.0f);
        if (face >= 6u) face = 5u;
        float uu = rand_lds(&r->rng);
        float vv = rand_lds(&r->rng);
        uint32_t seed = pack_face_seed(face, uu, vv);
        float tg = target_pdf(c, seed, sp, nrm);
        r_update(&res, pick, c->vid, seed, tg, tg * inv_src, &r->rng);
    }
    if (s->fc->restir_temporal) {
        reservoir tr;
        if (rbuf_load(s, pvid, &tr) && light_valid(&tr, s, count)) {
            r_cap(&tr, (uint32_t)s->fc->restir_max_m);
            r_combine(&res, &tr, tr.target_pdf, &r->rng);
        }
    }
    if (s->fc->restir_spatial && s->fc->restir_spatial_neighbors > 0) {
        v3 up = fabsf(nrm.z) < 0.999f ? V(0, 0, 1) : V(1, 0, 0);
        v3 T = vnorm(vcross(up, nrm));
        v3 B = vcross(nrm, T);
        int K = s->fc->restir_spatial_neighbors < 6 ? s->fc->restir_spatial_neighbors : 6;
        for (int k = 0; k < K; k++) {
            float ang = 6.2831853f * rand_lds(&r->rng);
            float rr = s->fc->restir_spatial_radius * sqrtf(rand_lds(&r->rng));
            v3 off = vmuls(vadd(vmuls(T, m_cos(ang)), vmuls(B, m_sin(ang))), rr);
            int lo[3] = {(int)rintf(off.x), (int)rintf(off.y), (int)rintf(off.z)};
            if (lo[0] == 0 && lo[1] == 0 && lo[2] == 0) continue;
            uint32_t nv = neighbour_vid(hp, hsize, lo);
            if (nv == 0u || nv == pvid) continue;
            reservoir nb;
            if (!rbuf_load(s, nv, &nb) || !light_valid(&nb, s, count)) continue;
            float t_us = target_pdf(&s->emissives[nb.light_idx], nb.face_seed, sp, nrm);
            if (t_us <= 0.0f) continue;
            r_cap(&nb, (uint32_t)s->fc->restir_max_m);
            r_combine(&res, &nb, t_us, &r->rng);
        }
    }
    if (res.m > 0u && res.target_pdf > 0.0f) {
        const emis_t *e = &s->emissives[res.light_idx];
        if (e->vid == res.light_vid) {
            v3 P, N;
            light_sample_point(res.face_seed, e, &P, &N);
            v3 to = vsub(P, sp);
            float d2 = vdot(to, to);
            float d = sqrtf(gmax(d2, 1e-12f));
            if (d > 1e-3f) {
                v3 ld = vdivs(to, d);
                float cx = gmax(vdot(nrm, ld), 0.0f);
                float cy = gmax(vdot(V(-ld.x, -ld.y, -ld.z), N), 0.0f);
                if (cx > 0.0f && cy > 0.0f && visible_to(s, sp, ld, res.light_vid)) {
                    v3 bc = evaluate_brdf(V(r->dx, r->dy, r->dz), ld, nrm, V(m->color[0], m->color[1], m->color[2]),
                                          m->roughness, m->metallic, m->specular);
                    float W = res.wsum / ((float)res.m * res.target_pdf);
                    float sz2 = (float)(e->size * e->size);
                    float geom = cy / gmax(d2, 1e-6f);
                    v3 c = vmuls(vmuls(vmuls(vmul(V(e->ex, e->ey, e->ez), bc), geom), sz2), W);
                    c = vmin(c, V(20, 20, 20));
                    v3 add = vmul(V(r->tx, r->ty, r->tz), c);
                    r->dlx += add.x; r->dly += add.y; r->dlz += add.z;
                }
            }
        }
    }
    st->pending = res.m > 0u;
    st->vid = pvid;
    st->res = res;
}

static void nee_ris(so_state *s, ray_t *r, const mat_t *m, v3 nrm, const uint32_t hp[3],
                    uint32_t n_lights)   /* shade.comp:816-836 (the RIS_M = 4 path) */
{
    v3 sp = vadd(vadd(V((float)hp[0], (float)hp[1], (float)hp[2]), V(0.5f, 0.5f, 0.5f)), nrm);
    float wsum = 0.0f, chosen_t = 0.0f;
    uint32_t chosen = 0u;
    for (int k = 0; k < 4; k++) {
        uint32_t pick = f2u(rand_lds(&r->rng) * (float)n_lights);
        if (pick >= n_lights) pick = n_lights - 1u;
        const emis_t *c = &s->emissives[pick];
        float lum = gmax(vdot(V(c->ex, c->ey, c->ez), LUMA), 1e-6f);
        float tg = lum * (float)(c->size * c->size);
        wsum += tg;
        if (rand_lds(&r->rng) * wsum < tg) { chosen = pick; chosen_t = tg; }
    }
    float avg = wsum / 4.0f;
    const emis_t *e = &s->emissives[chosen];
    uint32_t face = f2u(rand_lds(&r->rng) * 6.0f);
    if (face >= 6u) face = 5u;
    v3 fn, ft, fb;
    face_frame(face, &fn, &ft, &fb);
    float half = (float)e->size * 0.5f;
    v3 c = vadd(V((float)e->px, (float)e->py, (float)e->pz), V(half, half, half));
    float ru = rand_lds(&r->rng);
    float rv = rand_lds(&r->rng);
    v3 lp = vadd(vadd(vadd(c, vmuls(fn, half)), vmuls(ft, (ru - 0.5f) * 2.0f * half)), vmuls(fb, (rv - 0.5f) * 2.0f * half));
    v3 ld = vsub(lp, sp);
    float d2 = vdot(ld, ld);
    float d = sqrtf(d2);
    if (d > 1e-3f && chosen_t > 0.0f) {
        ld = vdivs(ld, d);
        float cx = gmax(vdot(nrm, ld), 0.0f);
        float cy = gmax(vdot(V(-ld.x, -ld.y, -ld.z), fn), 0.0f);
        if (cx > 0.0f && cy > 0.0f && visible_to(s, sp, ld, e->vid)) {
            v3 bc = evaluate_brdf(V(r->dx, r->dy, r->dz), ld, nrm, V(m->color[0], m->color[1], m->color[2]),
                                  m->roughness, m->metallic, m->specular);
            float area = 6.0f * (float)e->size * (float)e->size;
            float ris = (float)n_lights * avg / gmax(chosen_t, 1e-6f);
            float geom = cy / gmax(d2, 1e-6f);
            v3 cc = vmuls(vmuls(vmuls(vmul(V(e->ex, e->ey, e->ez), bc), geom), area), ris);
            cc = vmin(cc, V(20, 20, 20));
            v3 w = vmul(V(r->tx, r->ty, r->tz), cc);
            if (r->bounce == 0u) { r->dlx += w.x; r->dly += w.y; r->dlz += w.z; }
            else { r->ilx += w.x; r->ily += w.y; r->ilz += w.z; }
        }
    }
}

/* shade.comp:main, one SIMD-8 vector at a time so side effects (deposits,
 * emissive registry, survivor appends) happen in the order llvmpipe issues
 * them: every lane's miss deposit, then hits, then last-bounce deposits,
 * then roulette deposits, then survivor appends, each in lane order. */
static void shade_pass(so_state *s, uint32_t base, uint32_t size)
{
    const uint32_t MB = s->max_bounces;
    for (uint32_t v0 = 0; v0 < size; v0 += 8) {
        ray_t r[8];
        uint32_t slot[8];
        int live[8];
        for (uint32_t l = 0; l < 8; l++) {
            live[l] = 0;
            if (v0 + l >= size) continue;
            slot[l] = (base + v0 + l) % RAY_CAP;
            r[l] = s->ring[slot[l]];
            live[l] = r[l].bounce < MB;
        }
        /* case 1: misses */
        uint32_t dn = 0, dvid[8];
        v3 ddir[8], dind[8];
        for (uint32_t l = 0; l < 8; l++) {
            if (!live[l] || r[l].hit_vid != 0) continue;
            v3 sky = sky_sample(s, V(r[l].dx, r[l].dy, r[l].dz));   /* shade.comp:559 */
            r[l].ilx = r[l].ilx + sky.x * r[l].tx;
            r[l].ily = r[l].ily + sky.y * r[l].ty;
            r[l].ilz = r[l].ilz + sky.z * r[l].tz;
            dvid[dn] = r[l].primary_vid;
            ddir[dn] = V(r[l].dlx, r[l].dly, r[l].dlz);
            dind[dn++] = V(r[l].ilx, r[l].ily, r[l].ilz);
            r[l].bounce = MB;
            s->ring[slot[l]] = r[l];
            live[l] = 0;
        }
        deposit_vec(s, dn, dvid, ddir, dind);
        v3 normal[8];
        uint32_t hp[8][3], hsize[8];
        const mat_t *mat[8];
        for (uint32_t l = 0; l < 8; l++) {
            if (!live[l]) continue;
            uint32_t mi = r[l].hit_material < 127u ? r[l].hit_material : 126u;
            mat[l] = &s->materials[mi];
            hp[l][0] = (uint32_t)r[l].ox; hp[l][1] = (uint32_t)r[l].oy; hp[l][2] = (uint32_t)r[l].oz;
            hsize[l] = (uint32_t)r[l].pad0;
            normal[l] = lookup_normal(s, r[l].hit_vid, V(r[l].dx, r[l].dy, r[l].dz));
        }
        /* emissive contribution + registry (bounce 0) */
        for (uint32_t l = 0; l < 8; l++) {
            if (!live[l] || !mat[l]->emissive) continue;
            v3 col = V(mat[l]->color[0], mat[l]->color[1], mat[l]->color[2]);
            v3 em = vmin(vmul(vmuls(col, mat[l]->emissive_intensity), V(r[l].tx, r[l].ty, r[l].tz)), V(20, 20, 20));
            if (r[l].bounce == 0) { r[l].dlx += em.x; r[l].dly += em.y; r[l].dlz += em.z; }
            else { r[l].ilx += em.x; r[l].ily += em.y; r[l].ilz += em.z; }
            if (r[l].bounce == 0) {
                uint32_t vid = r[l].hit_vid;
                uint32_t ch = claim_hash(vid) % EM_CLAIM_W;
                for (int p = 0; p < EM_CLAIM_P; p++) {
                    uint32_t *c = &s->emissive_claim[(size_t)ch * EM_CLAIM_P + p];
                    uint32_t prev = *c;
                    if (prev == 0) *c = vid;
                    if (prev == vid) break;
                    if (prev == 0) {
                        uint32_t idx = s->emissive_count++;
                        if (idx < EM_CAP) {
                            emis_t *e = &s->emissives[idx];
                            v3 ee = vmuls(col, mat[l]->emissive_intensity);
                            e->ex = ee.x; e->ey = ee.y; e->ez = ee.z;
                            e->vid = vid;
                            e->px = hp[l][0]; e->py = hp[l][1]; e->pz = hp[l][2];
                            e->size = hsize[l] > 1u ? hsize[l] : 1u;
                        } else {
                            s->emissive_count--;
                        }
                        break;
                    }
                }
            }
        }
        /* next-event estimation (shade.comp:631-836): every lane of the vector
         * sees the registry after this vector's registrations; reservoir reads
         * of all lanes come before any lane's store */
        {
            nee_store st[8];
            memset(st, 0, sizeof st);
            uint32_t count = s->emissive_count;
            uint32_t n_lights = count < EM_CAP ? count : EM_CAP;
            for (uint32_t l = 0; l < 8; l++) {
                if (!live[l] || mat[l]->emissive || r[l].bounce >= 2u || count == 0) continue;
                if (s->fc->restir_enabled && r[l].bounce == 0u)
                    nee_restir(s, &r[l], mat[l], normal[l], hp[l], hsize[l], n_lights, count, &st[l]);
                else
                    nee_ris(s, &r[l], mat[l], normal[l], hp[l], n_lights);
            }
            for (uint32_t l = 0; l < 8; l++)
                if (st[l].pending) rbuf_store(s, st[l].vid, &st[l].res);
        }
        /* last allowed bounce: deposit and die */
        dn = 0;
        for (uint32_t l = 0; l < 8; l++) {
            if (!live[l] || r[l].bounce + 1u < MB) continue;
            dvid[dn] = r[l].primary_vid;
            ddir[dn] = V(r[l].dlx, r[l].dly, r[l].dlz);
            dind[dn++] = V(r[l].ilx, r[l].ily, r[l].ilz);
            r[l].bounce = MB;
            s->ring[slot[l]] = r[l];
            live[l] = 0;
        }
        deposit_vec(s, dn, dvid, ddir, dind);
        /* BRDF sampling + Russian roulette */
        v3 sdir[8], ntp[8];
        for (uint32_t l = 0; l < 8; l++) {
            if (!live[l]) continue;
            const mat_t *m = mat[l];
            v3 albedo = V(m->color[0], m->color[1], m->color[2]);
            v3 rd = V(r[l].dx, r[l].dy, r[l].dz);
            int use_ggx = m->metallic * (1.0f - m->roughness) > 0.5f;
            uint32_t rng = r[l].rng;
            v3 bt;
            if (use_ggx) {
                v3 Vv = V(-rd.x, -rd.y, -rd.z);
                v3 H = sample_ggx_vndf(normal[l], Vv, m->roughness, &rng);
                sdir[l] = reflect3(rd, H);
                if (vdot(sdir[l], normal[l]) <= 0.0f) {
                    sdir[l] = vnorm(vadd(normal[l], random_direction(&rng)));
                    bt = evaluate_brdf(rd, sdir[l], normal[l], albedo, m->roughness, m->metallic, m->specular);
                } else {
                    v3 F0 = vadd(vmuls(V(0.04f, 0.04f, 0.04f), 1.0f - m->metallic), vmuls(albedo, m->metallic));
                    bt = ggx_throughput(normal[l], Vv, sdir[l], H, m->roughness, F0);
                }
            } else {
                sdir[l] = vnorm(vadd(normal[l], random_direction(&rng)));
                bt = evaluate_brdf(rd, sdir[l], normal[l], albedo, m->roughness, m->metallic, m->specular);
            }
            r[l].rng = rng;
            ntp[l] = vmul(V(r[l].tx, r[l].ty, r[l].tz), bt);
        }
        dn = 0;
        for (uint32_t l = 0; l < 8; l++) {
            if (!live[l]) continue;
            if (r[l].bounce + 1u >= 2u) {   /* RR_START_BOUNCE */
                float p = gclamp(gmax(ntp[l].x, gmax(ntp[l].y, ntp[l].z)), 0.05f, 0.95f);
                float xi = rand_lds(&r[l].rng);
                if (xi > p) {
                    dvid[dn] = r[l].primary_vid;
                    ddir[dn] = V(r[l].dlx, r[l].dly, r[l].dlz);
                    dind[dn++] = V(r[l].ilx, r[l].ily, r[l].ilz);
                    r[l].bounce = MB;
                    s->ring[slot[l]] = r[l];
                    live[l] = 0;
                    continue;
                }
                ntp[l] = vdivs(ntp[l], p);
            }
        }
        deposit_vec(s, dn, dvid, ddir, dind);
        /* survivors: mark the slot dead, append the continuation at ring_head */
        for (uint32_t l = 0; l < 8; l++) {
            if (!live[l]) continue;
            ray_t sv = r[l];
            v3 o = vadd(V((float)hp[l][0] + 0.5f, (float)hp[l][1] + 0.5f, (float)hp[l][2] + 0.5f), normal[l]);
            sv.ox = o.x; sv.oy = o.y; sv.oz = o.z;
            sv.dx = sdir[l].x; sv.dy = sdir[l].y; sv.dz = sdir[l].z;
            sv.tx = ntp[l].x; sv.ty = ntp[l].y; sv.tz = ntp[l].z;
            sv.bounce = r[l].bounce + 1u;
            sv.hit_vid = 0;
            sv.hit_material = 0;
            ray_t dead = sv;
            dead.bounce = MB;
            s->ring[slot[l]] = dead;
            uint32_t ns = s->ring_head++;
            s->ring[ns % RAY_CAP] = sv;
        }
    }
}

/* ======================================================================== */
/* resolve.comp                                                              */
/* ======================================================================== */

static v3 voxel_avg(const so_state *s, uint32_t vid, uint32_t *count, float *variance)   /* resolve.comp:109-148 */
{
    if (vid == 0) { *count = 0; *variance = 0; return V(0, 0, 0); }
    uint32_t h = vid % (s->hw - 1u) + 1u;
    for (uint32_t i = 0; i < LSTRIDE * LSLOTS; i += LSTRIDE) {
        if (LB(s, h, i) != vid) continue;
        uint32_t dC = LB(s, h, i + 1), iC = LB(s, h, i + 6);
        *count = dC + iC;
        if (*count == 0) { *variance = 0; return V(0, 0, 0); }
        v3 da = V(0, 0, 0), ia = V(0, 0, 0);
        if (dC > 0) {
            float inv = 1.0f / (float)dC;
            da = vdivs(vmuls(V((float)LB(s, h, i + 2), (float)LB(s, h, i + 3), (float)LB(s, h, i + 4)), inv), 255.0f);
        }
        *variance = 0.0f;
        if (iC > 0) {
            uint32_t ir = LB(s, h, i + 7), ig = LB(s, h, i + 8), ib = LB(s, h, i + 9), il2 = LB(s, h, i + 10);
            float inv = 1.0f / (float)iC;
            ia = vdivs(vmuls(V((float)ir, (float)ig, (float)ib), inv), 255.0f);
            float meanL = (float)(ir + ig + ib) * inv / 3.0f;
            *variance = gmax((float)il2 * inv - meanL * meanL, 0.0f) / 65025.0f;
        }
        return vadd(da, ia);
    }
    *count = 0;
    *variance = 0;
    return V(0, 0, 0);
}

static int lookup_normal_resolve(const so_state *s, uint32_t vid, v3 *n)   /* resolve.comp:152-166 */
{
    *n = V(0, 0, 0);
    if (vid == 0) return 0;
    uint32_t h = vid % (s->hw - 1u) + 1u;
    for (uint32_t sl = 0; sl < NSLOTS; sl++) {
        if (NB(s, h, sl * NSTRIDE) == vid) {
            uint32_t f = NB(s, h, sl * NSTRIDE + 1);
            if (f == 0) return 0;
            *n = vnorm(unpack_normal(f));
            return 1;
        }
    }
    return 0;
}

/* texture(skybox, pixelDirection(pix)) (resolve.comp:170-180) */
static v3 sky_rgb(const so_state *s, int px, int py) { return sky_sample(s, pixel_dir(s, (uint32_t)px, (uint32_t)py)); }

static v3 reconstruct_timeout(const so_state *s, int px, int py)   /* resolve.comp:193-219 */
{
    v3 sum = V(0, 0, 0);
    float wsum = 0.0f;
    int R = s->fc->reconstruction_radius;
    for (int dy = -R; dy <= R; dy++)
        for (int dx = -R; dx <= R; dx++) {
            if (dx == 0 && dy == 0) continue;
            int qx = px + dx, qy = py + dy;
            if (qx < 0 || qy < 0 || qx >= (int)s->W || qy >= (int)s->H) continue;
            const uint32_t *nd = &s->gbuf[4 * ((size_t)qy * s->W + qx)];
            if (nd[3] == 0 || nd[0] >= (uint32_t)s->fc->primary_controlchecks) continue;
            uint32_t nc;
            float nv;
            v3 na = voxel_avg(s, nd[3], &nc, &nv);
            if (nc == 0) continue;
            float dist = (float)(dx * dx + dy * dy);
            float w = (float)(nc < 32u ? nc : 32u) / (1.0f + dist * 0.15f);
            sum = vadd(sum, vmuls(na, w));
            wsum += w;
        }
    return wsum > 0.0001f ? vdivs(sum, wsum) : sky_rgb(s, px, py);
}

static v3 resolve_default(const so_state *s, int px, int py, uint32_t vid, uint32_t checks)   /* resolve.comp:224-261 */
{
    if (vid == 0) {
        if (checks >= (uint32_t)s->fc->primary_controlchecks) return reconstruct_timeout(s, px, py);
        return sky_rgb(s, px, py);
    }
    uint32_t cc;
    float cv;
    v3 ca = voxel_avg(s, vid, &cc, &cv);
    if (cc >= 128u && cv < 0.02f) return ca;
    v3 sum = vmuls(ca, (float)cc);
    float wsum = (float)cc;
    static const int OFF[4][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}};
    for (int k = 0; k < 4; k++) {
        int qx = px + OFF[k][0], qy = py + OFF[k][1];
        if (qx < 0 || qy < 0 || qx >= (int)s->W || qy >= (int)s->H) continue;
        uint32_t nvid = s->gbuf[4 * ((size_t)qy * s->W + qx) + 3];
        if (nvid == 0 || nvid == vid) continue;
        uint32_t nc;
        float nv;
        v3 na = voxel_avg(s, nvid, &nc, &nv);
        if (nc == 0) continue;
        float dc = vlength(vsub(na, ca));
        float wc = m_exp(-dc * 6.0f);
        float wv = 1.0f / (1.0f + nv * 32.0f);
        float w = wc * wv * (float)(nc < 32u ? nc : 32u);
        sum = vadd(sum, vmuls(na, w));
        wsum += w;
    }
    return wsum > 0.0001f ? vdivs(sum, wsum) : V(0, 0, 0);
}

static float raw_variance(const so_state *s, uint32_t vid)   /* resolve.comp:302-322 */
{
    if (vid == 0) return 0.0f;
    uint32_t h = vid % (s->hw - 1u) + 1u;
    for (uint32_t i = 0; i < LSTRIDE * LSLOTS; i += LSTRIDE) {
        uint32_t id = LB(s, h, i);
        if (id == 0) return 0.0f;
        if (id != vid) continue;
        uint32_t c = LB(s, h, i + 6);
        if (c < 2u) return 0.0f;
        uint32_t rs = LB(s, h, i + 7), gs = LB(s, h, i + 8), bs = LB(s, h, i + 9), l2s = LB(s, h, i + 10);
        float fc = (float)c;
        float meanL = (float)(rs + gs + bs) / (3.0f * fc);
        float meanL2 = (float)l2s / fc;
        return gmax(meanL2 - meanL * meanL, 0.0f) / 65025.0f;
    }
    return 0.0f;
}

static void resolve_pixel(so_state *s, uint32_t x, uint32_t y)   /* resolve.comp:362-382 */
{
    size_t pix = (size_t)y * s->W + x;
    const uint32_t *g = &s->gbuf[4 * pix];
    uint32_t checks = g[0], mi = g[1], vid = g[3];
    int timeout = vid == 0 && checks >= (uint32_t)s->fc->primary_controlchecks;
    v3 col;
    switch (s->fc->render_type) {
    case 1: col = V((float)checks / 65.0f, (float)checks / 65.0f, (float)checks / 65.0f); break;
    case 2:
        if (vid == 0) col = timeout ? reconstruct_timeout(s, (int)x, (int)y) : sky_rgb(s, (int)x, (int)y);
        else if (mi == 0 || mi >= 127u) col = V(0, 0, 0);
        else col = V(s->materials[mi].color[0], s->materials[mi].color[1], s->materials[mi].color[2]);
        break;
    case 3:
        if (vid == 0) col = timeout ? reconstruct_timeout(s, (int)x, (int)y) : sky_rgb(s, (int)x, (int)y);
        else {
            v3 n;
            col = lookup_normal_resolve(s, vid, &n) ? vadd(vmuls(n, 0.5f), V(0.5f, 0.5f, 0.5f)) : V(0, 0, 0);
        }
        break;
    case 4:
        if (vid == 0) col = timeout ? reconstruct_timeout(s, (int)x, (int)y) : sky_rgb(s, (int)x, (int)y);
        else {
            uint32_t hid = vid - 1u;
            col = V((float)((hid + 1u) % 255u) / 255.0f, (float)(((hid + 1u) / 255u) % 255u) / 255.0f,
                    (float)(((hid + 1u) / (255u * 255u)) % 255u) / 255.0f);
        }
        break;
    case 5:
        if (vid == 0) col = timeout ? reconstruct_timeout(s, (int)x, (int)y) : sky_rgb(s, (int)x, (int)y);
        else {
            float sc = raw_variance(s, vid) * 65025.0f / gmax(s->fc->variance_norm, 1.0f);
            sc = gclamp(sc, 0.0f, 1.0f);
            col = V(sc, sc, sc);
        }
        break;
    default: col = resolve_default(s, (int)x, (int)y, vid, checks); break;
    }
    float *o = &s->out[4 * pix];
    o[0] = col.x; o[1] = col.y; o[2] = col.z; o[3] = 1.0f;
}

/* ======================================================================== */
/* Renderer::runWavefrontFrame (renderer.cpp:282-649)                        */
/* ======================================================================== */

int so_frame(so_state *s, const uint32_t *words, uint32_t n_words, uint32_t depth, const float ubo[16],
             const so_frame_config *fc, uint32_t frame_idx, uint32_t frame_seed, uint32_t time)
{
    s->words = words;
    s->n_words = n_words;
    s->depth = depth;
    memcpy(s->cam, ubo, sizeof s->cam);
    s->fc = fc;
    s->frame_idx = frame_idx;
    s->frame_seed = frame_seed;
    s->time = time;
    memset(s->counters, 0, sizeof s->counters);
    uint32_t budget = fc->shading_budget < 1 ? 1u : (uint32_t)fc->shading_budget;
    if (budget > SHADE_MAX) budget = SHADE_MAX;

    /* 1. generate_primary */
    so_generate_primary(words, n_words, depth, ubo, s->W, s->H, (uint32_t)fc->primary_controlchecks, s->gbuf, s->gpos,
                        NULL, NULL, 1);
    s->counters[6] = s->W * s->H;
    /* 2. normal */
    normal_pass(s);
    /* per-frame reset (renderer.cpp:343-375) */
    s->ring_head = s->ring_tail = 0;
    s->shade_count = 0;
    s->cand_count = 0;
    s->rank_stats[4] = 0x7F7FFFFFu;
    s->rank_stats[5] = 0;
    s->rank_stats[6] = 0;
    s->rank_stats[7] = 0;
    memset(&s->rank_stats[8], 0, 256 * 4);
    memset(s->claim, 0, (size_t)CLAIM_W * CLAIM_P * 4);
    memset(s->samples, 0, (size_t)CLAIM_W * CLAIM_P * 4);
    /* 3. schedule */
    schedule_pass(s);
    uint32_t cand = s->cand_count > CAND_CAP ? CAND_CAP : s->cand_count;
    s->counters[0] = cand;
    /* 3b. rank + threshold */
    uint32_t scheduled = 0;
    if (cand > 0) {
        rank_pass(s, budget);
        threshold_pass(s);
        scheduled = s->shade_count;
    }
    if (scheduled > budget) scheduled = budget;
    s->counters[1] = scheduled;
    /* 4+5. chunked emit / trace / shade */
    uint32_t bounces_run = 0;
    for (uint32_t base = 0; base < scheduled; base += RAY_CAP) {
        uint32_t csize = scheduled - base < RAY_CAP ? scheduled - base : RAY_CAP;
        s->ring_head = s->ring_tail = 0;
        emit_pass(s, base, csize);
        for (uint32_t b = 0; b < 8u; b++) {
            uint32_t head = s->ring_head, tail = s->ring_tail;
            if (head == tail) break;
            uint32_t win = head - tail;
            trace_pass(s, tail, win);
            shade_pass(s, tail, win);
            s->ring_tail = head;
            if (b + 1 > bounces_run) bounces_run = b + 1;
        }
    }
    s->counters[7] = bounces_run;
    s->counters[2] = s->emissive_count;
    /* 6. resolve */
    FOR_GROUPS_2D(s, X, Y, resolve_pixel(s, X, Y));
    return 0;
}

int so_resolve(so_state *s, const uint32_t *words, uint32_t n_words, uint32_t depth, const float ubo[16],
               const so_frame_config *fc)
{
    s->words = words;
    s->n_words = n_words;
    s->depth = depth;
    memcpy(s->cam, ubo, sizeof s->cam);
    s->fc = fc;
    FOR_GROUPS_2D(s, X, Y, resolve_pixel(s, X, Y));
    return 0;
}

size_t so_buffer(so_state *s, int id, void **ptr)
{
    size_t P = (size_t)s->W * s->H;
    switch (id) {
    case 0: *ptr = s->gbuf; return P * 16;
    case 1: *ptr = s->gpos; return P * 4;
    case 2: *ptr = s->nbuf; return (size_t)s->hw * NSTRIDE * NSLOTS * 4;
    case 3: *ptr = s->lbuf; return (size_t)s->hw * LSTRIDE * LSLOTS * 4;
    case 4: *ptr = s->rbuf; return (size_t)s->hw * RSTRIDE * RSLOTS * 4;
    case 5: *ptr = s->cands; return (size_t)(s->cand_count < CAND_CAP ? s->cand_count : CAND_CAP) * 16;
    case 6: *ptr = s->shade_list; return (size_t)s->shade_count * 8;
    case 7: *ptr = s->ring; return (size_t)RAY_CAP * sizeof(ray_t);
    case 8: *ptr = s->out; return P * 16;
    case 10: *ptr = s->rank_stats; return sizeof s->rank_stats;
    case 11: *ptr = s->samples; return (size_t)CLAIM_W * CLAIM_P * 4;
    case 12: *ptr = s->claim; return (size_t)CLAIM_W * CLAIM_P * 4;
    case 13: *ptr = s->emissives; return (size_t)EM_CAP * sizeof(emis_t);
    default: *ptr = NULL; return 0;
    }
}

void so_counters(so_state *s, uint32_t out[16]) { memcpy(out, s->counters, sizeof s->counters); }
