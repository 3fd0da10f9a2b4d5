// C ABI of the SVO ray pipeline: context, scene upload, per-pass dispatch,
// whole-frame sequence and buffer access.  Host code only; every kernel is
// launched through ve_launch.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/ve_svo.h"
#include "ve_internal.h"
#include "ve_launch.h"
#include "ve_ctx.h"

using namespace ve;

namespace {
thread_local std::string g_error;
}

void ve_clear_global_error() { g_error.clear(); }
int ve_global_error(int code, const char *msg)
{
    g_error = msg;
    return code;
}

namespace {

int ctx_error(ve_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf; else g_error = buf;
    return code;
}

#define HIP_OK(ctx, call)                                                                      \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return ctx_error(ctx, VE_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// An overlapped frame's resolve pass runs on aux3 beside the next frame's
// first passes; every call other than the next frame makes the main stream
// wait for it first, so that the main stream again follows everything the
// context has enqueued.
int join_resolve(ve_ctx *c)
{
    if (!c->resolve_pending) return VE_OK;
    HIP_OK(c, hipSetDevice(c->device));
    HIP_OK(c, hipStreamWaitEvent(c->stream, c->resolve_pending, 0));
    c->resolve_pending = nullptr;
    return VE_OK;
}

#define JOIN_RESOLVE(ctx)                         \
    do {                                          \
        if (int rc_ = join_resolve(ctx)) return rc_; \
    } while (0)

// Every device buffer of a context, in one list, carved out of one
// allocation at 256-byte offsets: with `base` null the list only adds up the
// bytes (admission control prices a context before creating it), with `base`
// set it hands out the pointers.  One hipMalloc and one clearing fill per
// context instead of one of each per buffer.
struct Alloc {
    uint8_t *base = nullptr;
    size_t bytes = 0;
    template <typename T>
    void operator()(T **p, size_t count)
    {
        if (base) *p = reinterpret_cast<T *>(base + bytes);
        bytes += (count * sizeof(T) + 255) & ~(size_t)255;
    }
};

void plan_buffers(Alloc &a, DeviceBuffers &b, uint32_t max_w, uint32_t max_h, uint32_t hash_width, bool overlap)
{
    const size_t P = (size_t)max_w * max_h;
    const size_t groups_max = (size_t)((max_w + 7) / 8) * ((max_h + 7) / 8);
    for (int k = 0; k < (overlap ? 2 : 1); k++) {
        a(&b.gbuf_set[k], P);
        a(&b.gpos_set[k], P);
        a(&b.ghit_set[k], P);
    }
    b.gbuf = b.gbuf_set[0];
    b.gpos = b.gpos_set[0];
    b.ghit = b.ghit_set[0];
    a(&b.output, P); a(&b.display, P); a(&b.pix_avg, P); a(&b.pix_cnt, P); a(&b.timeouts, P);
    a(&b.timeout_tiles, (size_t)((max_w + RECON_TILE - 1) / RECON_TILE) * ((max_h + RECON_TILE - 1) / RECON_TILE));
    a(&b.timeout_list, (size_t)((max_w + RECON_TILE - 1) / RECON_TILE) * ((max_h + RECON_TILE - 1) / RECON_TILE));
    for (int k = 0; k < (overlap ? 2 : 1); k++) a(&b.nbuf_set[k], (size_t)hash_width * NBUFFER_STRIDE * NBUFFER_SLOTS);
    b.nbuf = b.nbuf_set[0];
    a(&b.lbuf, (size_t)hash_width * LBUFFER_STRIDE * LBUFFER_SLOTS);
    a(&b.rbuf, (size_t)hash_width * RBUFFER_STRIDE * RBUFFER_SLOTS);
    a(&b.claim, (size_t)CLAIM_WIDTH * CLAIM_PROBES);
    a(&b.samples, (size_t)CLAIM_WIDTH * CLAIM_PROBES);
    a(&b.candidates, CANDIDATE_CAPACITY);
    a(&b.shade_list, SHADE_BUDGET_MAX);
    a(&b.rays, SHADE_BUDGET_MAX);
    a(&b.rays_next, SHADE_BUDGET_MAX);
    a(&b.counters, NUM_COUNTERS + COUNTER_SIDE_WORDS);
    a(&b.rank_stats, RANK_STATS_WORDS);
    a(&b.rank_part, (size_t)RANK_COPIES * RANK_COPY_WORDS);
    a(&b.bounce_part, (size_t)BOUNCE_COPIES * 2 + 1);
    a(&b.emissives, EMISSIVE_CAPACITY);
    a(&b.emissive_claim, (size_t)EMISSIVE_CLAIM_W * EMISSIVE_CLAIM_PROBES);
    a(&b.em_bits, SHADE_BUDGET_MAX / 32);
    a(&b.em_reg, EMISSIVE_CAPACITY);
    a(&b.hit_bits, SHADE_BUDGET_MAX / 32);
    a(&b.deferred, SHADE_BUDGET_MAX);
    a(&b.ray_li[0], SHADE_BUDGET_MAX);
    a(&b.ray_li[1], SHADE_BUDGET_MAX);
    a(&b.dep_death, SHADE_BUDGET_MAX);
    a(&b.dep_wpos, SHADE_BUDGET_MAX);
    a(&b.dep_cnt, hash_width);
    a(&b.dep_slots, (size_t)hash_width * DEP_CAP);
    a(&b.dep_over, SHADE_BUDGET_MAX);
    a(&b.dep_over_n, 2);
    a(&b.dep_spill, SHADE_BUDGET_MAX);
    a(&b.materials, MATERIAL_SLOTS);
    a(&b.node_stats, 8);
    a(&b.ray_totals, 4);
    a(&b.sched_tab, (size_t)CLAIM_WIDTH * SCHED_SLOTS);
    a(&b.sched_over, SCHED_OVER_CAP);
    a(&b.sched_over_n, 2);
    a(&b.sched_row_over, CLAIM_WIDTH);
    a(&b.sched_rows, CLAIM_WIDTH);
    a(&b.cand_at, groups_max * 64);
    a(&b.vec_cnt, groups_max * 8);
    a(&b.vec_chunk, (groups_max * 8 + 255) / 256);
    a(&b.rank_chunk, (std::min<size_t>(P, CANDIDATE_CAPACITY) + 255) / 256);
    a(&b.nsamp, std::min<size_t>(P, CANDIDATE_CAPACITY));
    b.ord_cap = 64u;
    while (b.ord_cap < ORDER_LDS_EVENTS && (size_t)b.ord_cap * hash_width < P) b.ord_cap *= 2u;
    b.ord_over_cap = b.ord_spill_cap = (uint32_t)P;
    a(&b.ord_cnt, hash_width);
    a(&b.ord_ev, (size_t)hash_width * b.ord_cap);
    a(&b.ord_over, P);
    a(&b.ord_over_n, 2);
    a(&b.ord_spill, P);
    a(&b.ord_work, (size_t)hash_width * NBUFFER_SLOTS);
    a(&b.dyn, 1);
}

// RankStats thresholds start at (+inf, -inf): every voxel in the mid tier on
// the first frame (rank.comp:109-111, threshold.comp:183-189).
const uint32_t INITIAL_THRESHOLDS[2] = {0x7F800000u, 0xFF800000u};

struct BufView {
    void *ptr;
    size_t bytes;
};

BufView buffer_view(ve_ctx *c, ve_buffer_id id)
{
    const size_t P = (size_t)c->max_w * c->max_h;
    switch (id) {
    case VE_BUF_GBUFFER: return {c->b.gbuf, P * 16};
    case VE_BUF_GBUFFER_POS: return {c->b.gpos, P * 4};
    case VE_BUF_NBUFFER: return {c->b.nbuf, (size_t)c->hash_width * NBUFFER_STRIDE * NBUFFER_SLOTS * 4};
    case VE_BUF_LBUFFER: return {c->b.lbuf, (size_t)c->hash_width * LBUFFER_STRIDE * LBUFFER_SLOTS * 4};
    case VE_BUF_RBUFFER: return {c->b.rbuf, (size_t)c->hash_width * RBUFFER_STRIDE * RBUFFER_SLOTS * 4};
    case VE_BUF_CANDIDATES: return {c->b.candidates, (size_t)CANDIDATE_CAPACITY * 16};
    case VE_BUF_SHADE_LIST: return {c->b.shade_list, (size_t)SHADE_BUDGET_MAX * 8};
    case VE_BUF_RAYS: return {c->b.rays, (size_t)SHADE_BUDGET_MAX * sizeof(Ray)};
    case VE_BUF_OUTPUT: return {c->b.output, P * 16};
    case VE_BUF_COUNTERS: return {c->b.counters, NUM_COUNTERS * 4};
    case VE_BUF_RANK_STATS: return {c->b.rank_stats, RANK_STATS_WORDS * 4};
    case VE_BUF_SAMPLE_COUNTS: return {c->b.samples, (size_t)CLAIM_WIDTH * CLAIM_PROBES * 4};
    case VE_BUF_CLAIM_MAP: return {c->b.claim, (size_t)CLAIM_WIDTH * CLAIM_PROBES * 4};
    case VE_BUF_EMISSIVES: return {c->b.emissives, (size_t)EMISSIVE_CAPACITY * sizeof(EmissiveVoxel)};
    case VE_BUF_DISPLAY: return {c->b.display, P * 4};
    default: return {nullptr, 0};
    }
}

void fill_params(ve_ctx *c, const ve_scene *s, const ve_frame_config *fc, const ve_frame_inputs *in, FrameParams &p)
{
    std::memset(&p, 0, sizeof p);
    p.W = in->width;
    p.H = in->height;
    p.depth = s ? s->depth : 0;
    p.n_words = s ? s->n_words : 0;
    p.jump = s ? s->jump : nullptr;
    p.jump_level = s ? s->jump_level : 0;
    p.tile_chunk = c->tuning.tile_chunk;
    p.bounce_blocks = c->tuning.bounce_blocks;
    p.hash_width = c->hash_width;
    p.max_bounces = c->max_bounces;
    p.frame_idx = in->frame_idx;
    p.frame_seed = in->frame_seed;
    p.time = in->time;
    p.primary_cc = (uint32_t)fc->primary_controlchecks;
    p.bounce_cc = (uint32_t)fc->bounce_controlchecks;
    p.update_time = fc->update_time;
    p.disable_stale_reset = fc->disable_stale_reset;
    p.normal_precision = fc->normal_precision;
    uint32_t budget = fc->shading_budget < 1 ? 1u : (uint32_t)fc->shading_budget;   // renderer.cpp:285-287
    p.shade_budget = std::min(budget, SHADE_BUDGET_MAX);
    p.cap_direct = (uint32_t)fc->sample_cap_direct;
    p.cap_indirect = (uint32_t)fc->sample_cap_indirect;
    p.w_pixels = fc->weight_pixels;
    p.w_variance = fc->weight_variance;
    p.pixels_norm = fc->pixels_norm;
    p.variance_norm = fc->variance_norm;
    p.top_fraction = fc->top_fraction;
    p.mid_fraction = fc->mid_fraction;
    p.min_samples = fc->min_samples;
    p.restir_enabled = fc->restir_enabled;
    p.restir_m = fc->restir_m;
    p.restir_temporal = fc->restir_temporal;
    p.restir_spatial = fc->restir_spatial;
    p.restir_neighbors = fc->restir_spatial_neighbors;
    p.restir_radius = fc->restir_spatial_radius;
    p.restir_max_m = fc->restir_max_m;
    p.firefly_k = fc->firefly_k;
    p.firefly_floor = fc->firefly_floor;
    p.render_type = fc->render_type;
    p.reconstruction_radius = fc->reconstruction_radius;
    std::memcpy(p.cam, c->cam, sizeof p.cam);
    std::memcpy(p.sky, c->sky, sizeof p.sky);
    p.sky_edge = c->sky_edge;
    p.sky_tex = c->sky_tex;
}

int check_frame(ve_ctx *c, const ve_scene *s, const ve_frame_config *fc, const ve_frame_inputs *in)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    if (!s || !s->words) return ctx_error(c, VE_EINVAL, "null scene");
    if (!fc || !in) return ctx_error(c, VE_EINVAL, "null frame config / inputs");
    if (in->width == 0 || in->height == 0) return ctx_error(c, VE_EINVAL, "zero-sized viewport");
    if (in->width > c->max_w || in->height > c->max_h)
        return ctx_error(c, VE_EINVAL, "viewport %ux%u exceeds the context's %ux%u", in->width, in->height, c->max_w, c->max_h);
    return VE_OK;
}

// Reference-layout index (slot * stride + field) -> device bucket index for
// the three hash images (see the bucket layouts in ve_internal.h).
uint32_t bucket_index(ve_buffer_id id, uint32_t y)
{
    switch (id) {
    case VE_BUF_NBUFFER: {
        const uint32_t s = y / NBUFFER_STRIDE, f = y % NBUFFER_STRIDE;
        return f * NBUFFER_SLOTS + s;
    }
    case VE_BUF_LBUFFER: {
        const uint32_t s = y / LBUFFER_STRIDE, f = y % LBUFFER_STRIDE;
        if (f == 0) return s;
        if (f == 11) return LB_TIME + s;
        return LB_PAY + s * LB_PAYW + (f - 1);
    }
    default: {   // VE_BUF_RBUFFER
        const uint32_t s = y / RBUFFER_STRIDE, f = y % RBUFFER_STRIDE;
        if (f == 0) return s;
        if (f == 7) return RB_TIME + s;
        return RB_PAY + s * RB_PAYW + (f - 1);
    }
    }
}

bool is_hash_buffer(ve_buffer_id id) { return id == VE_BUF_NBUFFER || id == VE_BUF_LBUFFER || id == VE_BUF_RBUFFER; }

// Permute a whole hash image between the device and the reference layout.
void permute_hash(ve_buffer_id id, const uint32_t *src, uint32_t *dst, size_t words, bool to_reference)
{
    const uint32_t per = id == VE_BUF_NBUFFER ? NB_WORDS : (id == VE_BUF_LBUFFER ? LB_WORDS : RB_WORDS);
    const size_t buckets = words / per;
    std::vector<uint32_t> map(per);
    for (uint32_t y = 0; y < per; y++) map[y] = bucket_index(id, y);
    for (size_t h = 0; h < buckets; h++) {
        const uint32_t *a = src + h * per;
        uint32_t *b = dst + h * per;
        if (to_reference)
            for (uint32_t y = 0; y < per; y++) b[y] = a[map[y]];
        else
            for (uint32_t y = 0; y < per; y++) b[map[y]] = a[y];
    }
}

}  // namespace

extern "C" {

const char *ve_last_error(const ve_ctx *ctx) { return ctx ? ctx->err.c_str() : g_error.c_str(); }

void ve_frame_config_defaults(ve_frame_config *fc)
{
    // core::FrameConfig defaults (src/renderer/core.hpp:83-158)
    std::memset(fc, 0, sizeof *fc);
    fc->primary_controlchecks = 70;
    fc->bounce_controlchecks = 100;
    fc->reconstruction_radius = 10;
    fc->normal_precision = 6;
    fc->disable_stale_reset = 0;
    fc->update_time = 0;
    fc->shading_budget = 131072;
    fc->sample_cap_direct = 16;
    fc->sample_cap_indirect = 64;
    fc->weight_pixels = 0.1f;
    fc->weight_variance = 0.9f;
    fc->pixels_norm = 100.0f;
    fc->variance_norm = 10000.0f;
    fc->top_fraction = 0.30f;
    fc->mid_fraction = 0.40f;
    fc->min_samples = 1;
    fc->restir_enabled = 1;
    fc->restir_m = 4;
    fc->restir_temporal = 1;
    fc->restir_spatial = 1;
    fc->restir_spatial_neighbors = 1;
    fc->restir_spatial_radius = 2.75f;
    fc->restir_max_m = 24;
    fc->firefly_k = 1.0f;
    fc->firefly_floor = 255.0f;
    fc->render_type = 0;
}

void ve_camera_ubo(const float pos[3], const float dir[3], float fov_deg, float aspect, float ubo[16])
{
    // Camera::UpdateUBO (camera.cpp:39-50) with glm's scalar arithmetic:
    // cross, normalize = v * (1/sqrt((x*x + y*y) + z*z)), radians = deg * pi/180.
    auto cross = [](const float x[3], const float y[3], float r[3]) {
        r[0] = x[1] * y[2] - y[1] * x[2];
        r[1] = x[2] * y[0] - y[2] * x[0];
        r[2] = x[0] * y[1] - y[0] * x[1];
    };
    auto normalize = [](const float v[3], float r[3]) {
        float t0 = v[0] * v[0], t1 = v[1] * v[1], t2 = v[2] * v[2];
        float inv = 1.0f / std::sqrt((t0 + t1) + t2);
        r[0] = v[0] * inv;
        r[1] = v[1] * inv;
        r[2] = v[2] * inv;
    };
    const float up_axis[3] = {0.0f, 1.0f, 0.0f};
    const float t = std::tan((fov_deg / 2.0f) * 0.01745329251994329576923690768489f);
    float c[3], n[3], right[3], u[3];
    cross(up_axis, dir, c);
    normalize(c, n);
    for (int i = 0; i < 3; i++) right[i] = (n[i] * t) * aspect;
    cross(right, dir, c);
    normalize(c, u);
    for (int i = 0; i < 3; i++) {
        ubo[0 + i] = pos[i];
        ubo[4 + i] = dir[i];
        ubo[8 + i] = right[i];
        ubo[12 + i] = u[i] * t;
    }
    ubo[3] = ubo[7] = ubo[11] = ubo[15] = 0.0f;
}

int ve_create(ve_ctx **out, const ve_renderer_config *cfg)
{
    ve_clear_global_error();
    std::string err;
    const int rc = ve::ctx_create(out, cfg, nullptr, true, err);
    return rc == VE_OK ? VE_OK : ve_global_error(rc, err.c_str());
}

void ve_destroy(ve_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->aux) (void)hipStreamSynchronize(c->aux);
    if (c->aux2) (void)hipStreamSynchronize(c->aux2);
    if (c->aux3) (void)hipStreamSynchronize(c->aux3);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &a : c->allocations) (void)hipFree(a.first);
    if (c->sky_tex) (void)hipFree(c->sky_tex);
    for (auto &fe : c->frame_events)
        for (hipEvent_t e : fe)
            if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {c->ev_primary, c->ev_sched, c->ev_emit, c->ev_fence, c->ev_normal, c->ev_dep, c->ev_resolve[0],
                         c->ev_resolve[1]})
        if (e) (void)hipEventDestroy(e);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    for (hipEvent_t e : c->dyn_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->dyn_host) (void)hipHostFree(c->dyn_host);
    if (c->aux) (void)hipStreamDestroy(c->aux);
    if (c->aux2) (void)hipStreamDestroy(c->aux2);
    if (c->aux3) (void)hipStreamDestroy(c->aux3);
    if (c->stream && c->owns_stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int ve_scene_upload(ve_ctx *c, const uint32_t *words, uint32_t n_words, uint32_t depth, ve_scene **scene)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    std::string err;
    c->aux_fence = true;   // the primary pass reads scenes; it follows the upload
    JOIN_RESOLVE(c);
    const int rc = ve::scene_upload(c->device, c->stream, words, n_words, depth, scene, err, c->tuning.jump_level);
    return rc == VE_OK ? VE_OK : ctx_error(c, rc, "%s", err.c_str());
}

int ve_scene_update(ve_ctx *c, ve_scene *s, const uint32_t *words, uint32_t n_words, const uint32_t *slots,
                    size_t n_slots)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    if (!s || !words || n_words == 0 || (n_slots && !slots)) return ctx_error(c, VE_EINVAL, "null argument");
    if (n_words < s->n_words) return ctx_error(c, VE_EINVAL, "node array shrank from %u to %u words", s->n_words, n_words);
    HIP_OK(c, hipSetDevice(c->device));
    JOIN_RESOLVE(c);
    c->aux_fence = true;   // the next overlapped primary pass reads the edited words
    if (n_words > s->alloc_words) {   // grow: a new array with headroom, whole image copied
        for (uint32_t i = 0; i < n_words; i++)
            if ((words[i] & 1u) && (uint64_t)(words[i] >> 4) + 8u > n_words)
                return ctx_error(c, VE_EFORMAT, "node %u points to block %u outside the %u-word array", i,
                                 words[i] >> 4, n_words);
        const uint32_t alloc = (uint32_t)std::min<uint64_t>((uint64_t)n_words + n_words / 2 + 8, 0xFFFFFFF0ull);
        void *d = nullptr;
        hipError_t e = hipMalloc(&d, (size_t)alloc * 4);
        if (e != hipSuccess) return ctx_error(c, VE_ENOMEM, "hipMalloc node array: %s", hipGetErrorString(e));
        e = hipMemsetAsync(static_cast<uint32_t *>(d) + n_words, 0, (size_t)(alloc - n_words) * 4, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(d, words, (size_t)n_words * 4, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            (void)hipFree(d);
            return ctx_error(c, VE_EDEVICE, "node upload: %s", hipGetErrorString(e));
        }
        (void)hipFree(s->words);
        s->words = static_cast<uint32_t *>(d);
        s->alloc_words = alloc;
        s->n_words = n_words;
        s->device_bytes = (size_t)alloc * 4 + ((size_t)4 << (3u * s->jump_level));
        launch_build_jump(c->stream, s->words, s->n_words, s->depth, s->jump_level, s->jump);
        HIP_OK(c, hipGetLastError());
        HIP_OK(c, hipStreamSynchronize(c->stream));
        return VE_OK;
    }
    std::vector<uint2> pairs;
    pairs.reserve(n_slots);
    for (size_t k = 0; k < n_slots; k++) {
        const uint32_t i = slots[k], w = i < n_words ? words[i] : 0u;
        if (i >= n_words) return ctx_error(c, VE_EINVAL, "changed slot %u beyond the %u-word array", i, n_words);
        if ((w & 1u) && (uint64_t)(w >> 4) + 8u > n_words)
            return ctx_error(c, VE_EFORMAT, "node %u points to block %u outside the %u-word array", i, w >> 4, n_words);
        pairs.push_back(make_uint2(i, w));
    }
    if (!pairs.empty()) {
        if (pairs.size() > s->staging_cap) {
            if (s->staging) (void)hipFree(s->staging);
            s->staging = nullptr;
            s->staging_cap = 0;
            const size_t cap = std::max<size_t>(pairs.size(), 4096);
            HIP_OK(c, hipMalloc(&s->staging, cap * sizeof(uint2)));
            s->staging_cap = cap;
        }
        HIP_OK(c, hipMemcpyAsync(s->staging, pairs.data(), pairs.size() * sizeof(uint2), hipMemcpyHostToDevice,
                                 c->stream));
        launch_scatter_words(c->stream, s->words, s->staging, (uint32_t)pairs.size());
    }
    s->n_words = n_words;
    launch_build_jump(c->stream, s->words, s->n_words, s->depth, s->jump_level, s->jump);   // after the new words
    HIP_OK(c, hipGetLastError());
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return VE_OK;
}

int ve_scene_release(ve_ctx *c, ve_scene *s)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    if (!s) return ctx_error(c, VE_ENOTFOUND, "null scene");
    (void)hipSetDevice(c->device);
    JOIN_RESOLVE(c);
    (void)hipStreamSynchronize(c->stream);
    ve::scene_free(s);
    return VE_OK;
}

int ve_set_tuning(ve_ctx *c, const ve_tuning *t)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    if (!t) return ctx_error(c, VE_EINVAL, "null tuning");
    if (t->tile_chunk > 4096u) return ctx_error(c, VE_EINVAL, "tile_chunk %u above 4096", t->tile_chunk);
    if (t->jump_level > JUMP_LEVEL) return ctx_error(c, VE_EINVAL, "jump_level %u above %u", t->jump_level, JUMP_LEVEL);
    if (t->bounce_blocks > 65536u) return ctx_error(c, VE_EINVAL, "bounce_blocks %u above 65536", t->bounce_blocks);
    c->tuning = *t;
    return VE_OK;
}

int ve_get_tuning(ve_ctx *c, ve_tuning *t)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    if (!t) return ctx_error(c, VE_EINVAL, "null tuning");
    *t = c->tuning;
    return VE_OK;
}

int ve_set_camera(ve_ctx *c, const float ubo[16])
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    if (!ubo) return ctx_error(c, VE_EINVAL, "null camera");
    std::memcpy(c->cam, ubo, sizeof c->cam);
    return VE_OK;
}

int ve_set_materials(ve_ctx *c, const void *entries, uint32_t n)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    if (!entries || n == 0 || n > MATERIAL_SLOTS) return ctx_error(c, VE_EINVAL, "material count %u outside [1,127]", n);
    HIP_OK(c, hipSetDevice(c->device));
    JOIN_RESOLVE(c);
    HIP_OK(c, hipMemcpyAsync(c->b.materials, entries, (size_t)n * sizeof(Material), hipMemcpyHostToDevice, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return VE_OK;
}

int ve_set_sky(ve_ctx *c, const uint8_t *const faces[6], int edge)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    if (!faces || edge < 1) return ctx_error(c, VE_EINVAL, "null faces or edge < 1");
    if (edge > 16384) return ctx_error(c, VE_EINVAL, "cube-map edge %d above 16384", edge);
    for (int f = 0; f < 6; f++)
        if (!faces[f]) return ctx_error(c, VE_EINVAL, "null face");
    // A uniform sky (every texel of every face the same RGB8 value) filters to
    // that texel for every direction: kept as a constant, c * (1/255) like
    // the sampler's own result.
    const uint8_t *c0 = faces[0];
    const size_t texels = (size_t)edge * (size_t)edge;
    bool uniform = true;
    for (int f = 0; f < 6 && uniform; f++)
        for (size_t i = 0; i < texels && uniform; i++) uniform = std::memcmp(faces[f] + 3 * i, c0, 3) == 0;
    if (uniform) {
        for (int i = 0; i < 3; i++) c->sky[i] = (float)c0[i] * (1.0f / 255.0f);
        c->sky_edge = 0;
        return VE_OK;
    }
    std::vector<uint32_t> tex(6 * texels);
    for (int f = 0; f < 6; f++)
        for (size_t i = 0; i < texels; i++) {
            const uint8_t *t = faces[f] + 3 * i;
            tex[f * texels + i] = (uint32_t)t[0] | (uint32_t)t[1] << 8 | (uint32_t)t[2] << 16;
        }
    HIP_OK(c, hipSetDevice(c->device));
    JOIN_RESOLVE(c);
    // frames already enqueued may still sample the current texture
    HIP_OK(c, hipStreamSynchronize(c->stream));
    if (tex.size() > c->sky_tex_words) {
        if (c->sky_tex) HIP_OK(c, hipFree(c->sky_tex));
        c->sky_tex = nullptr;
        c->sky_tex_words = 0;
        c->sky_edge = 0;
        HIP_OK(c, hipMalloc(&c->sky_tex, tex.size() * sizeof(uint32_t)));
        c->sky_tex_words = tex.size();
    }
    HIP_OK(c, hipMemcpy(c->sky_tex, tex.data(), tex.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    c->sky_edge = (uint32_t)edge;
    return VE_OK;
}

}  // extern "C"

namespace {

// Apply the deposits that pass-by-pass shade dispatches recorded, with the
// parameters of their frame.
int flush_deposits(ve_ctx *c)
{
    if (c->dep_iters == 0u) return VE_OK;
    const int rc = launch_deposits(c->stream, c->dep_params, c->b, c->dep_iters);
    c->dep_iters = 0u;
    if (rc) return ctx_error(c, rc, "deposits: device library call failed");
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ctx_error(c, VE_EDEVICE, "deposit launch: %s", hipGetErrorString(e));
    return VE_OK;
}

// Enqueue one pass; every launch goes on the context's stream.  The primary,
// schedule, normal and resolve passes run on `side_stream` when given (an
// overlapped frame's auxiliary streams).
int run_pass(ve_ctx *c, ve_scene *s, ve_pass_id pass, const FrameParams &p, hipStream_t side_stream = nullptr)
{
    int rc = VE_OK;
    // passes that read or clear what a frame's deposits change, or the
    // records themselves, come after the pending deposits
    if (pass == VE_PASS_RESET || pass == VE_PASS_SCHEDULE || pass == VE_PASS_RANK || pass == VE_PASS_THRESHOLD ||
        pass == VE_PASS_EMIT || pass == VE_PASS_RESOLVE)
        if ((rc = flush_deposits(c))) return rc;
    switch (pass) {
    case VE_PASS_PRIMARY:
        launch_primary(side_stream ? side_stream : c->stream, s->words, p, c->b, c->node_stats);
        break;
    case VE_PASS_NORMAL: launch_normal(side_stream ? side_stream : c->stream, s->words, p, c->b, c->nbuf_from); break;
    case VE_PASS_RESET: launch_reset(c->stream, p, c->b); break;
    case VE_PASS_SCHEDULE: rc = launch_schedule(side_stream ? side_stream : c->stream, p, c->b); break;
    case VE_PASS_RANK: rc = launch_rank(c->stream, p, c->b); break;
    case VE_PASS_THRESHOLD: launch_threshold(c->stream, p, c->b); break;
    case VE_PASS_EMIT:
        launch_emit(c->stream, p, c->b);
        c->bounce_iter = 0;
        break;
    case VE_PASS_TRACE:
        if (c->bounce_iter >= 8u) return ctx_error(c, VE_ESTATE, "trace after 8 bounce iterations (emit first)");
        launch_trace(c->stream, s->words, p, c->b, c->bounce_iter, c->node_stats);
        break;
    case VE_PASS_SHADE:
        if (c->bounce_iter >= 8u) return ctx_error(c, VE_ESTATE, "shade after 8 bounce iterations (emit first)");
        launch_shade(c->stream, s->words, p, c->b, c->bounce_iter);
        c->bounce_iter++;
        if (!p.count_primary_in_reset) {   // a lone dispatch: its deposits wait for the frame's others
            c->dep_iters = std::max(c->dep_iters, c->bounce_iter);
            c->dep_params = p;
        }
        break;
    case VE_PASS_RESOLVE: launch_resolve(side_stream ? side_stream : c->stream, p, c->b); break;
    case VE_PASS_DISPLAY: launch_display(c->stream, p, c->b); break;
    default: return ctx_error(c, VE_EINVAL, "unknown pass %d", (int)pass);
    }
    if (rc) return ctx_error(c, rc, "pass %d: device library call failed", (int)pass);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ctx_error(c, VE_EDEVICE, "pass %d launch: %s", (int)pass, hipGetErrorString(e));
    return VE_OK;
}

// Pass-timing events only carry timestamps (ve_timings reads them after a
// stream synchronisation) and are bound to the pass's kernel dispatches, so
// they need no system-scope release of their own.
constexpr unsigned TIMING_EVENT_FLAGS = hipEventDisableSystemFence;

// Event slot of the frame being recorded (timing mode only).
hipEvent_t *frame_slot(ve_ctx *c)
{
    if (c->frames_recorded >= c->frame_events.size()) {
        std::array<hipEvent_t, 2 * VE_PASS_COUNT * MAX_ITERS> fe{};
        for (auto &e : fe)
            if (hipEventCreateWithFlags(&e, TIMING_EVENT_FLAGS) != hipSuccess) return nullptr;
        c->frame_events.push_back(fe);
    }
    return c->frame_events[c->frames_recorded].data();
}

int timed_pass(ve_ctx *c, ve_scene *s, ve_pass_id pass, const FrameParams &p, uint32_t iter, hipEvent_t *ev,
               hipStream_t side_stream = nullptr)
{
    if (!ev) return run_pass(c, s, pass, p, side_stream);
    hipEvent_t *pair = ev + 2 * (pass * MAX_ITERS + iter);
    pass_timer = PassTimer{pair[0], pair[1]};
    const int rc = run_pass(c, s, pass, p, side_stream);
    pass_timer = PassTimer{};
    return rc;
}

// renderer.cpp:282-649 without its host readbacks: counts stay on the device
// and every kernel sizes its own work from them.  With cross-frame overlap
// the primary and schedule passes go to the auxiliary stream, into the other
// G-buffer set, once the previous frame's first shade pass is done, and the
// normal pass to the second auxiliary stream, into the other nBuffer set:
// they then run beside that frame's fused bounce and deposit passes, which
// leave most of the device idle (their time is the tail of the longest
// rays).  The main stream waits for the schedule pass before the rank pass
// and for the normal pass before the trace pass.  The resolve
// pass goes to the third auxiliary stream after the frame's deposits and
// runs beside the next frame's first passes; the next frame's deposits wait
// for it, and so does the primary pass two frames on, the next writer of the
// sets it reads.
int frame_passes(ve_ctx *c, ve_scene *s, const FrameParams &frame)
{
    hipEvent_t *ev = nullptr;
    if (c->timing && !(ev = frame_slot(c))) return ctx_error(c, VE_EDEVICE, "hipEventCreate failed");
    int rc;
    FrameParams p = frame;
    if (c->overlap) {
        // deposits that lone shade passes left pending go on the stream
        // before the fence that the auxiliary streams wait for
        if ((rc = flush_deposits(c))) return rc;
        const uint32_t k = c->gsel ^ 1u;
        c->gsel = k;
        c->b.gbuf = c->b.gbuf_set[k];
        c->b.gpos = c->b.gpos_set[k];
        c->b.ghit = c->b.ghit_set[k];
        uint32_t *nbuf_prev = c->b.nbuf;
        c->b.nbuf = c->b.nbuf_set[k];
        p.early_set = (int32_t)(1u + k);
        if (c->emit_done) HIP_OK(c, hipStreamWaitEvent(c->aux, c->emit_done, 0));
        if (c->aux_fence) {
            HIP_OK(c, hipEventRecord(c->ev_fence, c->stream));
            HIP_OK(c, hipStreamWaitEvent(c->aux, c->ev_fence, 0));
            c->aux_fence = false;
        }
        if (c->resolve_rec[k]) HIP_OK(c, hipStreamWaitEvent(c->aux, c->ev_resolve[k], 0));   // the sets' last reader
        if ((rc = timed_pass(c, s, VE_PASS_PRIMARY, p, 0, ev, c->aux))) return rc;
        HIP_OK(c, hipEventRecord(c->ev_primary, c->aux));
        if ((rc = timed_pass(c, s, VE_PASS_SCHEDULE, p, 0, ev, c->aux))) return rc;
        HIP_OK(c, hipEventRecord(c->ev_sched, c->aux));
        HIP_OK(c, hipStreamWaitEvent(c->aux2, c->ev_primary, 0));
        c->nbuf_from = nbuf_prev;   // the replay copies the buckets it leaves unchanged
        rc = timed_pass(c, s, VE_PASS_NORMAL, p, 0, ev, c->aux2);
        c->nbuf_from = nullptr;
        if (rc) return rc;
        HIP_OK(c, hipEventRecord(c->ev_normal, c->aux2));
        if ((rc = timed_pass(c, s, VE_PASS_RESET, p, 0, ev))) return rc;
        HIP_OK(c, hipStreamWaitEvent(c->stream, c->ev_sched, 0));
    } else {
        if ((rc = timed_pass(c, s, VE_PASS_RESET, p, 0, ev))) return rc;
        if (!c->skip_primary && (rc = timed_pass(c, s, VE_PASS_PRIMARY, p, 0, ev))) return rc;
        if ((rc = timed_pass(c, s, VE_PASS_NORMAL, p, 0, ev))) return rc;
        if ((rc = timed_pass(c, s, VE_PASS_SCHEDULE, p, 0, ev))) return rc;
    }
    // the threshold (threshold.comp) runs on the emit kernel's first wave
    // when the frame folds its small passes
    if ((rc = timed_pass(c, s, VE_PASS_RANK, p, 0, ev))) return rc;
    if (!p.fold_small_passes && (rc = timed_pass(c, s, VE_PASS_THRESHOLD, p, 0, ev))) return rc;
    const uint32_t iters = std::min<uint32_t>(c->max_bounces, (uint32_t)MAX_ITERS);   // renderer.cpp:515
    // the next frame's early passes follow this frame's first shade pass (or,
    // without bounces, its emit pass): the kernel's own dispatch marks its
    // completion, the timing event when the pass is timed, ev_emit otherwise
    auto anchored = [&](ve_pass_id pass) -> int {
        if (ev) {
            if (int r = timed_pass(c, s, pass, p, 0, ev)) return r;
            c->emit_done = ev[2 * (pass * MAX_ITERS) + 1];
            return VE_OK;
        }
        pass_timer = PassTimer{nullptr, c->ev_emit};
        const int r = run_pass(c, s, pass, p);
        pass_timer = PassTimer{};
        if (r == VE_OK) c->emit_done = c->ev_emit;
        return r;
    };
    if (c->overlap && iters == 0) {
        if ((rc = anchored(VE_PASS_EMIT))) return rc;
    } else if ((rc = timed_pass(c, s, VE_PASS_EMIT, p, 0, ev))) {
        return rc;
    }
    if (c->overlap) HIP_OK(c, hipStreamWaitEvent(c->stream, c->ev_normal, 0));   // shading and resolve read normals
    // iteration 0 (emissive registration, reservoirs) as its own passes; the
    // later ones fused per ray, timed as trace / shade of iteration 1
    if (iters > 0) {
        if ((rc = timed_pass(c, s, VE_PASS_TRACE, p, 0, ev))) return rc;
        if (c->overlap) {
            if ((rc = anchored(VE_PASS_SHADE))) return rc;
        } else if ((rc = timed_pass(c, s, VE_PASS_SHADE, p, 0, ev))) {
            return rc;
        }
    }
    if (iters > 1) {
        if (ev) pass_timer = PassTimer{ev[2 * (VE_PASS_TRACE * MAX_ITERS + 1)], ev[2 * (VE_PASS_TRACE * MAX_ITERS + 1) + 1]};
        launch_bounce(c->stream, s->words, p, c->b, 1u, iters, c->node_stats);
    }
    if (c->resolve_pending) {   // the previous frame's resolve reads the lBuffer
        HIP_OK(c, hipStreamWaitEvent(c->stream, c->resolve_pending, 0));
        c->resolve_pending = nullptr;
    }
    if (iters > 0) {   // the frame's deposits, timed with the last shade pass
        const uint32_t slot = 2 * (VE_PASS_SHADE * MAX_ITERS + (iters > 1 ? 1 : 0));
        if (ev) pass_timer = PassTimer{iters > 1 ? ev[slot] : nullptr, ev[slot + 1]};
        rc = launch_deposits(c->stream, p, c->b, iters);
        pass_timer = PassTimer{};
        if (rc) return ctx_error(c, rc, "deposits: device library call failed");
        HIP_OK(c, hipGetLastError());
    }
    if (c->overlap) {   // the resolve beside the next frame's first passes
        launch_dep_done(c->stream, p, c->b);
        HIP_OK(c, hipGetLastError());
        HIP_OK(c, hipEventRecord(c->ev_dep, c->stream));
        HIP_OK(c, hipStreamWaitEvent(c->aux3, c->ev_dep, 0));
        if ((rc = timed_pass(c, s, VE_PASS_RESOLVE, p, 0, ev, c->aux3))) return rc;
        const uint32_t k = c->gsel;
        HIP_OK(c, hipEventRecord(c->ev_resolve[k], c->aux3));
        c->resolve_rec[k] = true;
        c->resolve_pending = c->ev_resolve[k];
    } else if ((rc = timed_pass(c, s, VE_PASS_RESOLVE, p, 0, ev))) {
        return rc;
    }
    if (ev) {
        c->iters_recorded = std::min<uint32_t>(iters, 2u);
        c->frames_recorded++;
    }
    return VE_OK;
}

// One frame as a graph replay: the per-frame inputs go to the device block
// b.dyn (a copy on the context's stream from a pinned staging slot), the
// kernels read them through FrameParams::dyn, and the frame's launches --
// everything else in their parameters unchanged -- are replayed from the
// captured graph; a frame whose other parameters differ is captured anew.
// The staging slot is rewritten only once its previous copy has run.
int frame_graph(ve_ctx *c, ve_scene *s, const FrameParams &p)
{
    if (!c->dyn_host) {
        void *h = nullptr;
        HIP_OK(c, hipHostMalloc(&h, ve_ctx::DYN_RING * sizeof(FrameDyn), hipHostMallocDefault));
        c->dyn_host = static_cast<FrameDyn *>(h);
        for (auto &e : c->dyn_ev) HIP_OK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    const uint32_t j = c->dyn_next++ % ve_ctx::DYN_RING;
    HIP_OK(c, hipEventSynchronize(c->dyn_ev[j]));   // returns at once for a slot never copied from
    FrameDyn &d = c->dyn_host[j];
    d.frame_idx = p.frame_idx;
    d.frame_seed = p.frame_seed;
    d.time = p.time;
    d.pad = 0;
    std::memcpy(d.cam, p.cam, sizeof d.cam);
    HIP_OK(c, hipMemcpyAsync(c->b.dyn, &d, sizeof d, hipMemcpyHostToDevice, c->stream));
    HIP_OK(c, hipEventRecord(c->dyn_ev[j], c->stream));
    FrameParams q = p;
    q.frame_idx = q.frame_seed = q.time = 0;
    std::memset(q.cam, 0, sizeof q.cam);
    q.dyn = c->b.dyn;
    // the node array's address is a kernel argument beside the parameters
    std::vector<uint8_t> key(sizeof q + sizeof(void *) + 2);
    std::memcpy(key.data(), &q, sizeof q);
    const void *words = s->words;
    std::memcpy(key.data() + sizeof q, &words, sizeof words);
    key[sizeof q + sizeof words] = c->skip_primary ? 1 : 0;
    key[sizeof q + sizeof words + 1] = c->node_stats ? 1 : 0;
    if (!c->gexec || key != c->gkey) {
        if (c->gexec) {
            HIP_OK(c, hipGraphExecDestroy(c->gexec));
            c->gexec = nullptr;
        }
        HIP_OK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        const int rc = frame_passes(c, s, q);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : ctx_error(c, VE_EDEVICE, "graph capture: %s", hipGetErrorString(e));
        }
        const hipError_t ei = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            c->gexec = nullptr;
            return ctx_error(c, VE_EDEVICE, "graph instantiate: %s", hipGetErrorString(ei));
        }
        c->gkey = std::move(key);
        c->graph_captures++;
    }
    HIP_OK(c, hipGraphLaunch(c->gexec, c->stream));
    c->graph_replays++;
    return VE_OK;
}

}  // namespace

extern "C" {

int ve_set_graph(ve_ctx *c, int enable)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    c->graph = enable != 0;
    return VE_OK;
}

int ve_graph_stats(ve_ctx *c, uint64_t *captures, uint64_t *replays)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    if (captures) *captures = c->graph_captures;
    if (replays) *replays = c->graph_replays;
    return VE_OK;
}

int ve_pass(ve_ctx *c, ve_scene *s, ve_pass_id pass, const ve_frame_config *fc, const ve_frame_inputs *in)
{
    int rc = check_frame(c, s, fc, in);
    if (rc) return rc;
    HIP_OK(c, hipSetDevice(c->device));
    FrameParams p;
    fill_params(c, s, fc, in, p);
    // an overlapped primary pass must not run ahead of a pass that may touch
    // the G-buffers; the display pass reads only the resolved image
    if (pass != VE_PASS_DISPLAY) c->aux_fence = true;
    JOIN_RESOLVE(c);
    return run_pass(c, s, pass, p);
}

int ve_frame(ve_ctx *c, ve_scene *s, const ve_frame_config *fc, const ve_frame_inputs *in)
{
    int rc = check_frame(c, s, fc, in);
    if (rc) return rc;
    HIP_OK(c, hipSetDevice(c->device));
    FrameParams p;
    fill_params(c, s, fc, in, p);
    p.count_primary_in_reset = 1;
    p.fold_small_passes = 1;
    p.reuse_primary_hits = c->node_stats ? 0 : 1;   // node counting traces every ray as the reference does
    if (c->graph && !c->overlap && !c->timing) return frame_graph(c, s, p);
    rc = frame_passes(c, s, p);
    // a frame left half-enqueued may still read either G-buffer set: the next
    // primary pass waits for everything on the stream
    if (rc) c->aux_fence = true;
    return rc;
}

int ve_buffer_size(ve_ctx *c, ve_buffer_id id, size_t *bytes)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    BufView v = buffer_view(c, id);
    if (!v.ptr) return ctx_error(c, VE_ENOTFOUND, "unknown buffer id %d", (int)id);
    if (bytes) *bytes = v.bytes;
    return VE_OK;
}

int ve_read(ve_ctx *c, ve_buffer_id id, void *dst, size_t bytes)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    BufView v = buffer_view(c, id);
    if (!v.ptr) return ctx_error(c, VE_ENOTFOUND, "unknown buffer id %d", (int)id);
    if (!dst || bytes > v.bytes) return ctx_error(c, VE_EINVAL, "read of %zu bytes from a %zu-byte buffer", bytes, v.bytes);
    HIP_OK(c, hipSetDevice(c->device));
    JOIN_RESOLVE(c);
    if (int rc = flush_deposits(c)) return rc;
    if (is_hash_buffer(id)) {   // returned in the reference's slot-interleaved layout
        std::vector<uint32_t> tmp(v.bytes / 4), ref(v.bytes / 4);
        HIP_OK(c, hipMemcpyAsync(tmp.data(), v.ptr, v.bytes, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(c, hipStreamSynchronize(c->stream));
        permute_hash(id, tmp.data(), ref.data(), tmp.size(), true);
        std::memcpy(dst, ref.data(), bytes);
        return VE_OK;
    }
    HIP_OK(c, hipMemcpyAsync(dst, v.ptr, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return VE_OK;
}

int ve_write(ve_ctx *c, ve_buffer_id id, const void *src, size_t bytes)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    BufView v = buffer_view(c, id);
    if (!v.ptr) return ctx_error(c, VE_ENOTFOUND, "unknown buffer id %d", (int)id);
    if (!src || bytes > v.bytes) return ctx_error(c, VE_EINVAL, "write of %zu bytes to a %zu-byte buffer", bytes, v.bytes);
    c->aux_fence = true;
    HIP_OK(c, hipSetDevice(c->device));
    JOIN_RESOLVE(c);
    if (int rc = flush_deposits(c)) return rc;
    if (is_hash_buffer(id)) {   // accepted in the reference layout; whole image only
        if (bytes != v.bytes) return ctx_error(c, VE_EINVAL, "hash images are written whole (%zu bytes)", v.bytes);
        std::vector<uint32_t> dev(v.bytes / 4);
        permute_hash(id, static_cast<const uint32_t *>(src), dev.data(), dev.size(), false);
        HIP_OK(c, hipMemcpyAsync(v.ptr, dev.data(), v.bytes, hipMemcpyHostToDevice, c->stream));
        HIP_OK(c, hipStreamSynchronize(c->stream));
        return VE_OK;
    }
    HIP_OK(c, hipMemcpyAsync(v.ptr, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return VE_OK;
}

int ve_sync(ve_ctx *c)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    HIP_OK(c, hipSetDevice(c->device));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    if (c->aux) HIP_OK(c, hipStreamSynchronize(c->aux));
    if (c->aux2) HIP_OK(c, hipStreamSynchronize(c->aux2));
    if (c->aux3) HIP_OK(c, hipStreamSynchronize(c->aux3));
    return VE_OK;
}

int ve_set_overlap(ve_ctx *c, int enable)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    if (enable && !c->aux) return ctx_error(c, VE_EINVAL, "context has no auxiliary stream for overlap");
    JOIN_RESOLVE(c);
    c->overlap = enable != 0;
    c->emit_done = nullptr;
    c->aux_fence = true;
    return VE_OK;
}

int ve_enable_timing(ve_ctx *c, int enable)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    c->timing = enable != 0;
    c->frames_recorded = 0;
    // enable > 1 is the number of frames to record: create their events now,
    // outside any timed region
    HIP_OK(c, hipSetDevice(c->device));
    while (enable > 1 && c->frame_events.size() < (size_t)enable) {
        std::array<hipEvent_t, 2 * VE_PASS_COUNT * MAX_ITERS> fe{};
        for (auto &e : fe) HIP_OK(c, hipEventCreateWithFlags(&e, TIMING_EVENT_FLAGS));
        c->frame_events.push_back(fe);
    }
    return VE_OK;
}

// Elapsed time of a pass's event pair.  A pass that launched nothing in the
// frame (the threshold of a whole frame, which rides on the emit kernel)
// left its events unrecorded: the query fails, and the error it leaves with
// the runtime is cleared so that the next launch check does not report it.
static bool pair_ms(const hipEvent_t *pair, float &ms)
{
    if (hipEventElapsedTime(&ms, pair[0], pair[1]) == hipSuccess) return true;
    (void)hipGetLastError();
    return false;
}

int ve_timings_iter(ve_ctx *c, uint32_t iter, float ms_out[VE_PASS_COUNT])
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    if (!ms_out) return ctx_error(c, VE_EINVAL, "null output");
    if (iter > 1) return ctx_error(c, VE_EINVAL, "iter %u above 1", iter);
    HIP_OK(c, hipSetDevice(c->device));
    if (c->aux) HIP_OK(c, hipStreamSynchronize(c->aux));
    if (c->aux2) HIP_OK(c, hipStreamSynchronize(c->aux2));
    if (c->aux3) HIP_OK(c, hipStreamSynchronize(c->aux3));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < VE_PASS_COUNT; i++) {
        double sum = 0.0;
        const bool split = i == VE_PASS_TRACE || i == VE_PASS_SHADE;
        if (i != VE_PASS_DISPLAY && (iter == 0 || (split && iter < c->iters_recorded)))
            for (size_t f = 0; f < c->frames_recorded; f++) {
                const hipEvent_t *pair = &c->frame_events[f][2 * (i * MAX_ITERS + iter)];
                float ms = 0.0f;
                if (pair_ms(pair, ms)) sum += ms;
            }
        ms_out[i] = c->frames_recorded ? (float)(sum / (double)c->frames_recorded) : 0.0f;
    }
    return VE_OK;
}

int ve_timings(ve_ctx *c, float ms_out[VE_PASS_COUNT])
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    if (!ms_out) return ctx_error(c, VE_EINVAL, "null output");
    HIP_OK(c, hipSetDevice(c->device));
    if (c->aux) HIP_OK(c, hipStreamSynchronize(c->aux));
    if (c->aux2) HIP_OK(c, hipStreamSynchronize(c->aux2));
    if (c->aux3) HIP_OK(c, hipStreamSynchronize(c->aux3));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    double sum[VE_PASS_COUNT] = {0};
    for (size_t f = 0; f < c->frames_recorded; f++)
        for (int i = 0; i < VE_PASS_COUNT; i++) {
            if (i == VE_PASS_DISPLAY) continue;   // not part of ve_frame: no events recorded
            const uint32_t n_it = (i == VE_PASS_TRACE || i == VE_PASS_SHADE) ? c->iters_recorded : 1u;
            for (uint32_t it = 0; it < n_it; it++) {
                const hipEvent_t *pair = &c->frame_events[f][2 * (i * MAX_ITERS + it)];
                float ms = 0.0f;
                if (pair_ms(pair, ms)) sum[i] += ms;
            }
        }
    for (int i = 0; i < VE_PASS_COUNT; i++)
        ms_out[i] = c->frames_recorded ? (float)(sum[i] / (double)c->frames_recorded) : 0.0f;
    c->frames_recorded = 0;
    return VE_OK;
}

int ve_ray_totals(ve_ctx *c, uint64_t *primary, uint64_t *traced, uint64_t *shadow)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    unsigned long long v[3];
    HIP_OK(c, hipSetDevice(c->device));
    HIP_OK(c, hipMemcpyAsync(v, c->b.ray_totals, sizeof v, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    if (primary) *primary = v[0];
    if (traced) *traced = v[1];
    if (shadow) *shadow = v[2];
    return VE_OK;
}

int ve_ray_totals_reused(ve_ctx *c, uint64_t *reused)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    if (!reused) return ctx_error(c, VE_EINVAL, "null output");
    unsigned long long v = 0;
    HIP_OK(c, hipSetDevice(c->device));
    HIP_OK(c, hipMemcpyAsync(&v, c->b.ray_totals + 3, sizeof v, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    *reused = v;
    return VE_OK;
}

int ve_reset_ray_totals(ve_ctx *c)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    HIP_OK(c, hipSetDevice(c->device));
    HIP_OK(c, hipMemsetAsync(c->b.ray_totals, 0, 4 * sizeof(unsigned long long), c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return VE_OK;
}

int ve_enable_node_stats(ve_ctx *c, int enable)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    c->node_stats = enable != 0;
    HIP_OK(c, hipSetDevice(c->device));
    HIP_OK(c, hipMemsetAsync(c->b.node_stats, 0, 8 * sizeof(unsigned long long), c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return VE_OK;
}

int ve_node_stats(ve_ctx *c, uint64_t *pr, uint64_t *tr, uint64_t *ps, uint64_t *ts)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    unsigned long long v[6];
    HIP_OK(c, hipSetDevice(c->device));
    HIP_OK(c, hipMemcpyAsync(v, c->b.node_stats, sizeof v, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    if (pr) *pr = v[0];
    if (tr) *tr = v[1] + v[4];
    if (ps) *ps = v[2];
    if (ts) *ts = v[3] + v[5];
    return VE_OK;
}

int ve_node_stats_bounce(ve_ctx *c, uint64_t *reads, uint64_t *steps)
{
    if (!c) return ve_global_error(VE_EINVAL, "null context");
    unsigned long long v[6];
    HIP_OK(c, hipSetDevice(c->device));
    HIP_OK(c, hipMemcpyAsync(v, c->b.node_stats, sizeof v, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    if (reads) *reads = v[4];
    if (steps) *steps = v[5];
    return VE_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// internal entry points shared with the session server (ve_ctx.h)
// ---------------------------------------------------------------------------
namespace ve {

size_t ctx_device_bytes(uint32_t max_w, uint32_t max_h, uint32_t hash_width)
{
    Alloc a;
    DeviceBuffers b;
    plan_buffers(a, b, max_w, max_h, hash_width ? hash_width : 16384u, false);
    return a.bytes;
}

int ctx_create(ve_ctx **out, const ve_renderer_config *cfg, hipStream_t shared, bool overlap, std::string &err)
{
    if (!out || !cfg) return err = "null argument", VE_EINVAL;
    if (cfg->max_width == 0 || cfg->max_height == 0) return err = "zero-sized framebuffer", VE_EINVAL;
    if (cfg->max_bounces > 64) return err = "max_bounces above 64", VE_EINVAL;
    const uint32_t hw = cfg->hash_width ? cfg->hash_width : 16384u;
    if (hw < 2) return err = "hash_width below 2", VE_EINVAL;
    const size_t groups_max = (size_t)((cfg->max_width + 7) / 8) * ((cfg->max_height + 7) / 8);
    if (groups_max * 64 >= (1ull << 31)) return err = "framebuffer too large", VE_EINVAL;
    ve_ctx *c = new (std::nothrow) ve_ctx();
    if (!c) return err = "context allocation", VE_ENOMEM;
    c->device = cfg->device;
    c->max_w = cfg->max_width;
    c->max_h = cfg->max_height;
    c->hash_width = hw;
    c->max_bounces = cfg->max_bounces ? cfg->max_bounces : 8u;
    hipError_t e = hipSetDevice(c->device);
    if (e == hipSuccess) {
        if (shared) {
            c->stream = shared;
            c->owns_stream = false;
        } else {
            e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);   // default priority
            if (e == hipSuccess && overlap) {
                c->overlap = true;
                e = hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking);   // default priority
                // ordering only, between two streams of one device
                const unsigned fl = hipEventDisableTiming | hipEventReleaseToDevice;
                if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_primary, fl);
                // bound to the emit kernel's dispatch, like the pass-timing events
                if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_emit, hipEventDisableSystemFence);
                if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fence, fl);
                if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->aux2, hipStreamNonBlocking);   // default priority
                if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_sched, fl);
                if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_normal, fl);
                if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->aux3, hipStreamNonBlocking);   // default priority
                if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_dep, fl);
                for (int k = 0; k < 2 && e == hipSuccess; k++) e = hipEventCreateWithFlags(&c->ev_resolve[k], fl);
            }
        }
    }
    if (e != hipSuccess) {
        err = hipGetErrorString(e);
        ve_destroy(c);
        return VE_EDEVICE;
    }
    Alloc a;
    plan_buffers(a, c->b, c->max_w, c->max_h, c->hash_width, c->overlap);   // size
    void *block = nullptr;
    e = hipMalloc(&block, a.bytes);
    if (e != hipSuccess) {
        err = std::string("hipMalloc(") + std::to_string(a.bytes) + " bytes): " + hipGetErrorString(e);
        ve_destroy(c);
        return VE_ENOMEM;
    }
    c->allocations.push_back({block, a.bytes});
    Alloc carve;
    carve.base = static_cast<uint8_t *>(block);
    plan_buffers(carve, c->b, c->max_w, c->max_h, c->hash_width, c->overlap);
    // cleared on the context's stream, ahead of everything that uses it
    e = hipMemsetAsync(block, 0, a.bytes, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->b.rank_stats, INITIAL_THRESHOLDS, sizeof INITIAL_THRESHOLDS, hipMemcpyHostToDevice,
                           c->stream);
    if (e != hipSuccess) {
        err = hipGetErrorString(e);
        ve_destroy(c);
        return VE_EDEVICE;
    }
    c->device_bytes = a.bytes;
    *out = c;
    return VE_OK;
}

int frame_begin(ve_ctx *c, ve_scene *s, const ve_frame_config *fc, const ve_frame_inputs *in, FrameParams &p)
{
    int rc = check_frame(c, s, fc, in);
    if (rc) return rc;
    if (c->overlap) return ctx_error(c, VE_ESTATE, "batched frames need the cross-frame overlap off");
    if (c->timing || c->node_stats) return ctx_error(c, VE_ESTATE, "batched frames run without timing or node counting");
    HIP_OK(c, hipSetDevice(c->device));
    fill_params(c, s, fc, in, p);
    p.count_primary_in_reset = 1;
    p.fold_small_passes = 1;
    p.reuse_primary_hits = 1;
    return VE_OK;
}

int frame_rest(ve_ctx *c, ve_scene *s, const FrameParams &p)
{
    c->skip_primary = true;
    const int rc = c->graph ? frame_graph(c, s, p) : frame_passes(c, s, p);
    c->skip_primary = false;
    return rc;
}

int ctx_reset(ve_ctx *c)
{
    HIP_OK(c, hipSetDevice(c->device));
    for (auto &a : c->allocations) HIP_OK(c, hipMemsetAsync(a.first, 0, a.second, c->stream));
    HIP_OK(c, hipMemcpyAsync(c->b.rank_stats, INITIAL_THRESHOLDS, sizeof INITIAL_THRESHOLDS, hipMemcpyHostToDevice,
                             c->stream));
    c->bounce_iter = 0;
    c->dep_iters = 0;
    c->frames_recorded = 0;
    c->timing = false;
    c->gsel = 0;
    c->b.gbuf = c->b.gbuf_set[0];
    c->b.gpos = c->b.gpos_set[0];
    c->b.ghit = c->b.ghit_set[0];
    c->b.nbuf = c->b.nbuf_set[0];
    c->resolve_rec[0] = c->resolve_rec[1] = false;
    c->resolve_pending = nullptr;
    c->emit_done = nullptr;
    c->aux_fence = true;
    c->node_stats = false;
    std::memset(c->cam, 0, sizeof c->cam);
    std::memset(c->sky, 0, sizeof c->sky);
    c->sky_edge = 0;
    return VE_OK;
}

int scene_upload(int device, hipStream_t stream, const uint32_t *words, uint32_t n_words, uint32_t depth,
                 ve_scene **scene, std::string &err, uint32_t jump_level)
{
    if (!words || !scene || n_words == 0) return err = "null words / scene or empty array", VE_EINVAL;
    char buf[256];
    if (depth < 1 || depth > 15) {
        snprintf(buf, sizeof buf, "depth %u outside [1,15]", depth);
        return err = buf, VE_EINVAL;
    }
    // Every internal node's 8-word child block must lie inside the array, so
    // the traversal can never fetch outside it.
    for (uint32_t i = 0; i < n_words; i++) {
        const uint32_t w = words[i];
        if ((w & 1u) && (uint64_t)(w >> 4) + 8u > n_words) {
            snprintf(buf, sizeof buf, "node %u points to block %u outside the %u-word array", i, w >> 4, n_words);
            return err = buf, VE_EFORMAT;
        }
    }
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return err = hipGetErrorString(e), VE_EDEVICE;
    ve_scene *s = new (std::nothrow) ve_scene();
    if (!s) return err = "scene allocation", VE_ENOMEM;
    void *d = nullptr;
    e = hipMalloc(&d, (size_t)n_words * 4);
    if (e != hipSuccess) {
        delete s;
        return err = std::string("hipMalloc node array: ") + hipGetErrorString(e), VE_ENOMEM;
    }
    const uint32_t jl = std::min(jump_level ? std::min(jump_level, JUMP_LEVEL) : JUMP_LEVEL, depth - 1u);
    void *j = nullptr;
    e = hipMalloc(&j, (size_t)4 << (3u * jl));
    if (e != hipSuccess) {
        (void)hipFree(d);
        delete s;
        return err = std::string("hipMalloc jump table: ") + hipGetErrorString(e), VE_ENOMEM;
    }
    e = hipMemcpyAsync(d, words, (size_t)n_words * 4, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) {
        launch_build_jump(stream, static_cast<const uint32_t *>(d), n_words, depth, jl, static_cast<uint32_t *>(j));
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        (void)hipFree(d);
        (void)hipFree(j);
        delete s;
        return err = std::string("node upload: ") + hipGetErrorString(e), VE_EDEVICE;
    }
    s->words = static_cast<uint32_t *>(d);
    s->n_words = n_words;
    s->alloc_words = n_words;
    s->depth = depth;
    s->jump = static_cast<uint32_t *>(j);
    s->jump_level = jl;
    s->device_bytes = (size_t)n_words * 4 + ((size_t)4 << (3u * jl));
    *scene = s;
    return VE_OK;
}

void scene_free(ve_scene *s)
{
    if (!s) return;
    (void)hipFree(s->words);
    if (s->jump) (void)hipFree(s->jump);
    if (s->staging) (void)hipFree(s->staging);
    delete s;
}

hipStream_t ctx_stream(ve_ctx *c) { return c->stream; }
size_t ctx_bytes(const ve_ctx *c) { return c->device_bytes; }
void ctx_shape(const ve_ctx *c, uint32_t *w, uint32_t *h, uint32_t *hash_width, uint32_t *max_bounces)
{
    *w = c->max_w;
    *h = c->max_h;
    *hash_width = c->hash_width;
    *max_bounces = c->max_bounces;
}
unsigned long long *ctx_ray_totals(ve_ctx *c) { return c->b.ray_totals; }

}  // namespace ve
