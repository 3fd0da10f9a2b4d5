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
enum { FPV_LLVMPIPE = 0, FPV_FMA = 1, FPV_DOT_LEFT = 2, FPV_NORM_DIV = 3, FPV_DIV_RCP = 4 };

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

/* Elementwise FP primitives of the GLSL lowering, for comparison with
 * llvmpipe probe outputs.  op 0: dot(v,v) of vec3 in[3i..3i+2]; 1: normalize
 * (3 outputs); 2: a/b; 3: a*b+c; 4: min(a,b); 5: max(a,b). */
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

so_hit so_raycast(const uint32_t *words, uint32_t n_words, uint32_t depth, const float origin_in[3],
                  const float dir[3], const float inv[3], uint32_t controlchecks, uint32_t *checks_out,
                  uint64_t *node_reads)
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
        float tmn[3], tmx[3];
        for (int i = 0; i < 3; i++) {
            float t1 = ((0.0f - o[i]) + 0.001f) * inv[i];
            float t2 = ((fl - o[i]) - 0.001f) * inv[i];
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
    while (in_bounds(r, fl) && q++ <= controlchecks) {   /* internal.glsl:370 */
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
            float bmin = (float)tpos[i];
            float bmax = bmin + (float)tsize;
            float t1 = ((bmin - o[i]) - 0.001f) * inv[i];
            float t2 = ((bmax - o[i]) + 0.001f) * inv[i];
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
