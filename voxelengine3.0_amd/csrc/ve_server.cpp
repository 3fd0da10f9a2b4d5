// Session server: resident octrees, per-session renderer state admitted by a
// device-memory budget, frames of different sessions overlapping on the
// device (continuous batching).  Host code only: every frame is the same
// ve_frame sequence a lone context runs, enqueued on the session's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "../../include/ve_svo.h"
#include "ve_ctx.h"

namespace {

using Clock = std::chrono::steady_clock;

// Pinned host slot that receives a context's cumulative ray totals after a
// frame (a 24-byte device-to-host copy on the frame's stream).
struct RaySlot {
    unsigned long long v[4];
};

struct Frame {
    uint64_t tag = 0;
    uint32_t index = 0;
    double arrival_ms = 0;
    ve_frame_config fc{};
    ve_frame_inputs in{};
    float ubo[16] = {0};
    hipEvent_t start = nullptr, end = nullptr;
    RaySlot *slot = nullptr;
    uint64_t seq = 0;   // launch order across the server
    int queue = -1;     // shared stream it was placed on (placement by load)
    double cost = 0;    // its estimated device time there, ms
    double units = 0;   // its size in work units (pixels x (2 + samples x bounces))
};

enum class State { Waiting, Active, Closed };

struct Session {
    uint32_t id = 0;
    ve_session_desc desc{};
    std::vector<uint8_t> materials;
    State state = State::Waiting;
    bool close_requested = false;
    ve_ctx *ctx = nullptr;
    uint32_t frames = 0;                  // submitted so far
    std::deque<Frame> pending;            // submitted before admission
    std::deque<Frame> in_flight;          // launched, not yet collected
    bool queued = false;                  // a frame waits in the server's next batch
    hipEvent_t reset_ev = nullptr;        // placement by load: after the context's reset on its stream
    double est_ms = 0;                    // placement by load: moving average of its frames' device time
    unsigned long long last_totals[3] = {0, 0, 0};
};

struct Pooled {
    ve_ctx *ctx;
    uint64_t last_used;
};

}  // namespace

struct ve_server {
    int device = 0;
    uint32_t max_sessions = 64;
    uint64_t budget = 0;
    uint64_t used = 0;
    std::vector<hipStream_t> queues;       // shared streams (empty: one per session)
    bool place = false;                    // each frame on the shared stream with the least work queued
    std::vector<double> qload;             // per shared stream: estimated device time of its uncollected frames, ms
    double ms_per_unit = 0;                // moving average over collected frames: device time per work unit
    uint32_t next_queue = 0;
    std::string err;
    Clock::time_point t0;
    hipEvent_t ref = nullptr;              // device-clock origin
    double ref_ms = 0;                     // server time of `ref`
    std::vector<ve_scene *> scenes;
    std::map<uint32_t, Session> sessions;
    std::deque<uint32_t> waiting;
    std::vector<Pooled> pool;
    uint32_t next_id = 1;
    uint32_t admitted = 0;
    uint64_t seq = 0, use_tick = 0;
    std::vector<hipEvent_t> free_events;
    std::vector<RaySlot *> free_slots;
    std::vector<RaySlot *> slot_chunks;
    std::deque<ve_frame_done> done;        // finished, not yet returned
    bool graph = false;                    // sessions replay their frame graphs (ve_set_graph)
    bool tuned = false;                    // ve_server_set_tuning called: `tuning` goes to every context
    ve_tuning tuning{};
    ve_batch *batch = nullptr;             // ve_server_submit_batch's primary launches
    ve_batch_config batch_cfg{};
    // frames of ve_server_submit_batch waiting for the next batch: {session, frame}
    std::vector<std::pair<uint32_t, Frame>> ready;
};

namespace {

int srv_error(ve_server *s, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (s) s->err = buf;
    return code;
}

#define SRV_HIP(srv, call)                                                                       \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) return srv_error(srv, VE_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

double now_ms(const ve_server *s)
{
    return std::chrono::duration<double, std::milli>(Clock::now() - s->t0).count();
}

int get_event(ve_server *s, hipEvent_t *e)
{
    if (!s->free_events.empty()) {
        *e = s->free_events.back();
        s->free_events.pop_back();
        return VE_OK;
    }
    SRV_HIP(s, hipEventCreate(e));
    return VE_OK;
}

int get_slot(ve_server *s, RaySlot **slot)
{
    if (s->free_slots.empty()) {
        constexpr size_t CHUNK = 1024;
        void *p = nullptr;
        SRV_HIP(s, hipHostMalloc(&p, CHUNK * sizeof(RaySlot), hipHostMallocDefault));
        s->slot_chunks.push_back(static_cast<RaySlot *>(p));
        for (size_t i = 0; i < CHUNK; i++) s->free_slots.push_back(static_cast<RaySlot *>(p) + i);
    }
    *slot = s->free_slots.back();
    s->free_slots.pop_back();
    return VE_OK;
}

// Enqueue one frame on the session's stream: start marker, ve_frame, the
// ray-total copy and the end marker.  No host synchronisation.
int launch(ve_server *s, Session &ss, Frame &f)
{
    ve_ctx *c = ss.ctx;
    int rc;
    if ((rc = get_event(s, &f.start)) || (rc = get_event(s, &f.end)) || (rc = get_slot(s, &f.slot))) return rc;
    if (s->place && s->queues.size() > 1) {
        // join the shortest queue: the shared stream whose uncollected frames
        // hold the least estimated device time;
        // the session's previous work, if still running elsewhere, is waited
        // for on the device
        size_t best = 0;
        for (size_t i = 1; i < s->queues.size(); i++)
            if (s->qload[i] < s->qload[best]) best = i;
        hipStream_t q = s->queues[best];
        if (c->stream != q) {
            hipEvent_t dep = !ss.in_flight.empty() ? ss.in_flight.back().end : ss.reset_ev;
            if (dep && hipEventQuery(dep) == hipErrorNotReady) SRV_HIP(s, hipStreamWaitEvent(q, dep, 0));
            c->stream = q;
        }
        // a frame's device time: its session's recent frames', else its size
        // in work units at the server's recent time per unit
        const uint32_t mb = ss.desc.max_bounces ? ss.desc.max_bounces : 8u;
        f.queue = (int)best;
        f.units = (double)f.in.width * f.in.height * (2.0 + (double)std::max(f.fc.min_samples, 1) * mb);
        f.cost = ss.est_ms > 0 ? ss.est_ms : f.units * (s->ms_per_unit > 0 ? s->ms_per_unit : 1e-6);
        s->qload[best] += f.cost;
    }
    std::memcpy(c->cam, f.ubo, sizeof c->cam);
    SRV_HIP(s, hipEventRecord(f.start, c->stream));
    rc = ve_frame(c, s->scenes[ss.desc.scene], &f.fc, &f.in);
    if (rc != VE_OK) return srv_error(s, rc, "session %u frame %u: %s", ss.id, f.index, c->err.c_str());
    SRV_HIP(s, hipMemcpyAsync(f.slot->v, ve::ctx_ray_totals(c), 3 * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, c->stream));
    SRV_HIP(s, hipEventRecord(f.end, c->stream));
    f.seq = s->seq++;
    return VE_OK;
}

bool fits(const ve_ctx *c, const ve_session_desc &d, uint32_t hw)
{
    uint32_t w, h, chw, mb;
    ve::ctx_shape(c, &w, &h, &chw, &mb);
    const uint64_t area = (uint64_t)w * h, need = (uint64_t)d.width * d.height;
    return w >= d.width && h >= d.height && chw == hw && mb == (d.max_bounces ? d.max_bounces : 8u) &&
           area <= 2 * need + 65536;
}

void destroy_pooled(ve_server *s, size_t i)
{
    s->used -= ve::ctx_bytes(s->pool[i].ctx);
    ve_destroy(s->pool[i].ctx);
    s->pool.erase(s->pool.begin() + (long)i);
}

// Placement by load: an event after the context's set-up on its stream,
// which the session's first frame waits for wherever it is placed.
int mark_reset(ve_server *s, Session &ss)
{
    if (!s->place) return VE_OK;
    if (!ss.reset_ev) {
        int rc = get_event(s, &ss.reset_ev);
        if (rc) return rc;
    }
    SRV_HIP(s, hipEventRecord(ss.reset_ev, ss.ctx->stream));
    return VE_OK;
}

// Give the session a context: a pooled one that fits (cleared), else a new
// one if the budget allows after dropping idle pooled contexts.  Returns
// VE_EBUDGET when it has to wait.
int acquire(ve_server *s, Session &ss)
{
    const uint32_t hw = ss.desc.hash_width ? ss.desc.hash_width : 16384u;
    for (size_t i = 0; i < s->pool.size(); i++) {
        if (!fits(s->pool[i].ctx, ss.desc, hw)) continue;
        ss.ctx = s->pool[i].ctx;
        ss.ctx->graph = s->graph;
        if (s->tuned) ss.ctx->tuning = s->tuning;
        s->pool.erase(s->pool.begin() + (long)i);
        if (ve::ctx_reset(ss.ctx) != VE_OK) return srv_error(s, VE_EDEVICE, "context reset: %s", ss.ctx->err.c_str());
        return mark_reset(s, ss);
    }
    const uint64_t need = ve::ctx_device_bytes(ss.desc.width, ss.desc.height, hw);
    while (s->used + need > s->budget && !s->pool.empty()) {
        size_t oldest = 0;
        for (size_t i = 1; i < s->pool.size(); i++)
            if (s->pool[i].last_used < s->pool[oldest].last_used) oldest = i;
        destroy_pooled(s, oldest);
    }
    if (s->used + need > s->budget) return VE_EBUDGET;
    ve_renderer_config cfg{s->device, ss.desc.width, ss.desc.height, hw, ss.desc.max_bounces};
    hipStream_t q = nullptr;
    if (!s->queues.empty()) q = s->queues[s->next_queue++ % s->queues.size()];
    std::string err;
    const int rc = ve::ctx_create(&ss.ctx, &cfg, q, false, err);
    if (rc != VE_OK) return srv_error(s, rc, "session %u: %s", ss.id, err.c_str());
    s->used += ve::ctx_bytes(ss.ctx);
    ss.ctx->graph = s->graph;
    if (s->tuned) ss.ctx->tuning = s->tuning;
    return mark_reset(s, ss);
}

// Admit waiting sessions in arrival order while slots and memory allow, and
// launch the frames they queued.
int admit(ve_server *s)
{
    while (!s->waiting.empty() && s->admitted < s->max_sessions) {
        Session &ss = s->sessions[s->waiting.front()];
        int rc = acquire(s, ss);
        if (rc == VE_EBUDGET) break;
        if (rc != VE_OK) return rc;
        s->waiting.pop_front();
        ve_ctx *c = ss.ctx;
        SRV_HIP(s, hipMemcpyAsync(c->b.materials, ss.materials.data(), ss.materials.size(), hipMemcpyHostToDevice,
                                  c->stream));
        // a constant sky, as ve_set_sky keeps a uniform cube
        for (int i = 0; i < 3; i++) c->sky[i] = (float)ss.desc.sky_rgb8[i] * (1.0f / 255.0f);
        c->sky_edge = 0;
        ss.state = State::Active;
        s->admitted++;
        while (!ss.pending.empty()) {
            Frame f = ss.pending.front();
            ss.pending.pop_front();
            if ((rc = launch(s, ss, f))) return rc;
            ss.in_flight.push_back(f);
        }
    }
    return VE_OK;
}

// A closed session with nothing left in flight returns its context to the pool.
void retire(ve_server *s, Session &ss)
{
    if (ss.state != State::Active || !ss.close_requested || !ss.in_flight.empty() || !ss.pending.empty() || ss.queued)
        return;
    s->pool.push_back({ss.ctx, s->use_tick++});
    ss.ctx = nullptr;
    if (ss.reset_ev) {
        s->free_events.push_back(ss.reset_ev);
        ss.reset_ev = nullptr;
    }
    ss.state = State::Closed;
    s->admitted--;
}

void finish(ve_server *s, Session &ss, Frame &f)
{
    float a = 0, b = 0;
    (void)hipEventElapsedTime(&a, s->ref, f.start);
    (void)hipEventElapsedTime(&b, s->ref, f.end);
    ve_frame_done d{};
    d.session = ss.id;
    d.frame = f.index;
    d.tag = f.tag;
    d.arrival_ms = f.arrival_ms;
    d.start_ms = s->ref_ms + a;
    d.done_ms = s->ref_ms + b;
    unsigned long long r = 0;
    for (int i = 0; i < 3; i++) {
        r += f.slot->v[i] - ss.last_totals[i];
        ss.last_totals[i] = f.slot->v[i];
    }
    d.rays = r;
    s->done.push_back(d);
    if (f.queue >= 0) {
        s->qload[(size_t)f.queue] = std::max(0.0, s->qload[(size_t)f.queue] - f.cost);
        const double el = std::max(0.0, (double)(b - a));
        ss.est_ms = ss.est_ms > 0 ? 0.75 * ss.est_ms + 0.25 * el : el;
        if (f.units > 0)
            s->ms_per_unit = s->ms_per_unit > 0 ? 0.95 * s->ms_per_unit + 0.05 * (el / f.units) : el / f.units;
    }
    s->free_events.push_back(f.start);
    s->free_events.push_back(f.end);
    s->free_slots.push_back(f.slot);
}

// Collect every finished frame (in order per session).  Returns the number found.
size_t collect(ve_server *s)
{
    size_t n = 0;
    for (auto &kv : s->sessions) {
        Session &ss = kv.second;
        while (!ss.in_flight.empty() && hipEventQuery(ss.in_flight.front().end) == hipSuccess) {
            finish(s, ss, ss.in_flight.front());
            ss.in_flight.pop_front();
            n++;
        }
        retire(s, ss);
    }
    for (auto it = s->sessions.begin(); it != s->sessions.end();)
        it = (it->second.state == State::Closed) ? s->sessions.erase(it) : std::next(it);
    return n;
}

// Launch the frames waiting in the next batch: every frame's checks and
// parameters, the batch's primary launches, then each frame's other passes
// on its session's stream (or, whole-frame batches, as one launch per kernel
// over all the batch's frames on the batch stream).  Unless forced (or the batch is full) this waits
// while the previous batch's primary launch still runs: frames arriving
// meanwhile join the next batch, so the batches grow with the load and the
// batch stream never queues more than one of them.
int flush_ready(ve_server *s, bool force)
{
    if (s->ready.empty()) return VE_OK;
    if (!force && s->ready.size() < s->batch_cfg.max_views && !ve::batch_idle(s->batch)) return VE_OK;
    std::vector<std::pair<uint32_t, Frame>> rd;
    rd.swap(s->ready);
    const uint32_t m = (uint32_t)rd.size();
    std::vector<Session *> act(m);
    std::vector<ve_ctx *> ctxs(m);
    std::vector<ve_scene *> scs(m);
    std::vector<ve::FrameParams> ps(m);
    std::vector<hipEvent_t> starts(m);
    int rc;
    for (uint32_t i = 0; i < m; i++) {
        act[i] = &s->sessions.at(rd[i].first);
        act[i]->queued = false;
    }
    for (uint32_t i = 0; i < m; i++) {
        Session &ss = *act[i];
        Frame &f = rd[i].second;
        if ((rc = get_event(s, &f.start)) || (rc = get_event(s, &f.end)) || (rc = get_slot(s, &f.slot))) return rc;
        ctxs[i] = ss.ctx;
        scs[i] = s->scenes[ss.desc.scene];
        std::memcpy(ss.ctx->cam, f.ubo, sizeof ss.ctx->cam);
        if ((rc = ve::frame_begin(ss.ctx, scs[i], &f.fc, &f.in, ps[i])))
            return srv_error(s, rc, "session %u frame %u: %s", ss.id, f.index, ss.ctx->err.c_str());
        starts[i] = f.start;
    }
    const bool whole = ve::batch_whole(s->batch);
    if ((rc = ve::batch_primary(s->batch, m, ctxs.data(), scs.data(), ps.data(), starts.data(), whole)))
        return srv_error(s, rc, "batch: %s", ve::batch_error_text(s->batch));
    for (uint32_t i = 0; i < m; i++) {
        Session &ss = *act[i];
        Frame &f = rd[i].second;
        ve_ctx *c = ss.ctx;
        // whole frames: the session's stream already waits for the batch's last launch
        if (!whole && (rc = ve::frame_rest(c, scs[i], ps[i])))
            return srv_error(s, rc, "session %u frame %u: %s", ss.id, f.index, c->err.c_str());
        SRV_HIP(s, hipMemcpyAsync(f.slot->v, ve::ctx_ray_totals(c), 3 * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, c->stream));
        SRV_HIP(s, hipEventRecord(f.end, c->stream));
        f.seq = s->seq++;
        ss.in_flight.push_back(f);
    }
    return VE_OK;
}

// Frames of a session must run in order: one waiting in the next batch goes
// first.
int flush_for(ve_server *s, const Session &ss) { return ss.queued ? flush_ready(s, true) : VE_OK; }

}  // namespace

extern "C" {

int ve_server_create(ve_server **out, const ve_server_config *cfg)
{
    if (!out || !cfg) return ve_global_error(VE_EINVAL, "null argument");
    ve_server *s = new (std::nothrow) ve_server();
    if (!s) return ve_global_error(VE_ENOMEM, "server allocation");
    s->device = cfg->device;
    s->max_sessions = cfg->max_sessions ? cfg->max_sessions : 64u;
    s->t0 = Clock::now();
    hipError_t e = hipSetDevice(s->device);
    size_t free_b = 0, total_b = 0;
    if (e == hipSuccess) e = hipMemGetInfo(&free_b, &total_b);
    for (uint32_t i = 0; e == hipSuccess && i < cfg->queues; i++) {
        hipStream_t q = nullptr;
        e = hipStreamCreateWithFlags(&q, hipStreamNonBlocking);   // default priority
        if (e == hipSuccess) s->queues.push_back(q);
        s->qload.push_back(0.0);
    }
    // device-clock origin: a marker on an idle queue, observed from the host
    hipStream_t q0 = nullptr;
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&q0, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&s->ref);
    if (e == hipSuccess) e = hipEventRecord(s->ref, q0);
    if (e == hipSuccess) e = hipEventSynchronize(s->ref);
    s->ref_ms = now_ms(s);
    if (q0) (void)hipStreamDestroy(q0);
    if (e != hipSuccess) {
        ve_global_error(VE_EDEVICE, hipGetErrorString(e));
        ve_server_destroy(s);
        return VE_EDEVICE;
    }
    s->budget = cfg->hbm_budget ? cfg->hbm_budget : (uint64_t)(free_b / 4 * 3);
    *out = s;
    return VE_OK;
}

void ve_server_destroy(ve_server *s)
{
    if (!s) return;
    (void)hipSetDevice(s->device);
    for (auto &kv : s->sessions) {
        if (kv.second.ctx) ve_destroy(kv.second.ctx);   // waits for its stream
        for (Frame &f : kv.second.in_flight) {
            (void)hipEventDestroy(f.start);
            (void)hipEventDestroy(f.end);
        }
        if (kv.second.reset_ev) (void)hipEventDestroy(kv.second.reset_ev);
    }
    for (auto &p : s->pool) ve_destroy(p.ctx);
    if (s->batch) ve_batch_destroy(s->batch);
    for (ve_scene *sc : s->scenes) ve::scene_free(sc);
    for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
    for (RaySlot *c : s->slot_chunks) (void)hipHostFree(c);
    for (hipStream_t q : s->queues) (void)hipStreamDestroy(q);
    if (s->ref) (void)hipEventDestroy(s->ref);
    delete s;
}

const char *ve_server_last_error(const ve_server *s) { return s ? s->err.c_str() : ve_last_error(nullptr); }

double ve_server_now_ms(const ve_server *s) { return s ? now_ms(s) : 0.0; }

int ve_server_add_scene(ve_server *s, const uint32_t *words, uint32_t n_words, uint32_t depth, uint32_t *scene_id)
{
    if (!s || !scene_id) return ve_global_error(VE_EINVAL, "null argument");
    hipStream_t q = nullptr;
    SRV_HIP(s, hipSetDevice(s->device));
    SRV_HIP(s, hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
    ve_scene *sc = nullptr;
    std::string err;
    const int rc = ve::scene_upload(s->device, q, words, n_words, depth, &sc, err);
    (void)hipStreamDestroy(q);
    if (rc != VE_OK) return srv_error(s, rc, "%s", err.c_str());
    *scene_id = (uint32_t)s->scenes.size();
    s->scenes.push_back(sc);
    return VE_OK;
}

int ve_server_open(ve_server *s, const ve_session_desc *d, uint32_t *session_id)
{
    if (!s || !d || !session_id) return ve_global_error(VE_EINVAL, "null argument");
    if (d->scene >= s->scenes.size()) return srv_error(s, VE_ENOTFOUND, "unknown scene %u", d->scene);
    if (d->width == 0 || d->height == 0) return srv_error(s, VE_EINVAL, "zero-sized session");
    if (d->max_bounces > 64) return srv_error(s, VE_EINVAL, "max_bounces above 64");
    if (!d->materials || d->n_materials == 0 || d->n_materials > ve::MATERIAL_SLOTS)
        return srv_error(s, VE_EINVAL, "material count %u outside [1,127]", d->n_materials);
    const uint32_t hw = d->hash_width ? d->hash_width : 16384u;
    if (hw < 2) return srv_error(s, VE_EINVAL, "hash_width below 2");
    const uint64_t need = ve::ctx_device_bytes(d->width, d->height, hw);
    if (need > s->budget)
        return srv_error(s, VE_EBUDGET, "session needs %llu bytes, budget is %llu", (unsigned long long)need,
                         (unsigned long long)s->budget);
    SRV_HIP(s, hipSetDevice(s->device));
    Session ss;
    ss.id = s->next_id++;
    ss.desc = *d;
    ss.materials.assign(static_cast<const uint8_t *>(d->materials),
                        static_cast<const uint8_t *>(d->materials) + (size_t)d->n_materials * sizeof(ve::Material));
    ss.desc.materials = nullptr;
    const uint32_t id = ss.id;
    s->sessions[id] = std::move(ss);
    s->waiting.push_back(id);
    *session_id = id;
    return admit(s);
}

int ve_server_submit(ve_server *s, uint32_t session, const ve_frame_config *fc, const ve_frame_inputs *in,
                     const float ubo[16], uint64_t tag, double arrival_ms)
{
    if (!s || !fc || !in || !ubo) return ve_global_error(VE_EINVAL, "null argument");
    auto it = s->sessions.find(session);
    if (it == s->sessions.end() || it->second.close_requested)
        return srv_error(s, VE_ENOTFOUND, "unknown or closing session %u", session);
    Session &ss = it->second;
    if (in->width != ss.desc.width || in->height != ss.desc.height)
        return srv_error(s, VE_EINVAL, "frame %ux%u on a %ux%u session", in->width, in->height, ss.desc.width,
                         ss.desc.height);
    SRV_HIP(s, hipSetDevice(s->device));
    Frame f;
    f.tag = tag;
    f.index = ss.frames++;
    f.arrival_ms = arrival_ms < 0 ? now_ms(s) : arrival_ms;
    f.fc = *fc;
    f.in = *in;
    std::memcpy(f.ubo, ubo, sizeof f.ubo);
    if (ss.state != State::Active) {
        ss.pending.push_back(f);
        return VE_OK;
    }
    int rc = flush_for(s, ss);
    if (rc != VE_OK) return rc;
    rc = launch(s, ss, f);
    if (rc != VE_OK) return rc;
    ss.in_flight.push_back(f);
    return VE_OK;
}

int ve_server_set_graph(ve_server *s, int enable)
{
    if (!s) return ve_global_error(VE_EINVAL, "null server");
    s->graph = enable != 0;
    for (auto &kv : s->sessions)
        if (kv.second.ctx) kv.second.ctx->graph = s->graph;
    for (auto &p : s->pool) p.ctx->graph = s->graph;
    return VE_OK;
}

int ve_server_set_placement(ve_server *s, int by_load)
{
    if (!s) return ve_global_error(VE_EINVAL, "null server");
    if (by_load && s->queues.size() < 2) return srv_error(s, VE_EINVAL, "placement by load needs two or more shared streams");
    if (!s->sessions.empty()) return srv_error(s, VE_ESTATE, "placement is chosen before the first session opens");
    s->place = by_load != 0;
    return VE_OK;
}

int ve_server_set_tuning(ve_server *s, const ve_tuning *t)
{
    if (!s || !t) return ve_global_error(VE_EINVAL, "null argument");
    if (t->tile_chunk > 4096u || t->bounce_blocks > 65536u)
        return srv_error(s, VE_EINVAL, "tile_chunk above 4096 or bounce_blocks above 65536");
    s->tuning = *t;
    s->tuning.jump_level = 0u;
    s->tuned = true;
    for (auto &kv : s->sessions)
        if (kv.second.ctx) kv.second.ctx->tuning = s->tuning;
    for (auto &p : s->pool) p.ctx->tuning = s->tuning;
    return VE_OK;
}

int ve_server_set_batch(ve_server *s, const ve_batch_config *cfg)
{
    if (!s || !cfg) return ve_global_error(VE_EINVAL, "null argument");
    ve_batch_config c = *cfg;
    c.device = s->device;
    c.max_views = std::max(c.max_views ? c.max_views : 256u, s->max_sessions);   // a batch holds every admitted session
    ve_batch *b = nullptr;
    int rc = ve_batch_create(&b, &c);
    if (rc != VE_OK) return srv_error(s, rc, "batch: %s", ve_last_error(nullptr));
    if (s->batch && (rc = flush_ready(s, true)) != VE_OK) {
        ve_batch_destroy(b);
        return rc;
    }
    if (s->batch) ve_batch_destroy(s->batch);
    s->batch = b;
    s->batch_cfg = c;
    return VE_OK;
}

int ve_server_submit_batch(ve_server *s, uint32_t n, const uint32_t *sessions, const ve_frame_config *fcs,
                           const ve_frame_inputs *ins, const float *ubos, const uint64_t *tags, const double *arrival_ms)
{
    if (!s || (n && (!sessions || !fcs || !ins || !ubos))) return ve_global_error(VE_EINVAL, "null argument");
    SRV_HIP(s, hipSetDevice(s->device));
    if (!s->batch) {
        ve_batch_config c{};
        c.device = s->device;
        if (int rc = ve_server_set_batch(s, &c)) return rc;
    }
    const double now = now_ms(s);
    for (uint32_t i = 0; i < n; i++) {
        auto it = s->sessions.find(sessions[i]);
        if (it == s->sessions.end() || it->second.close_requested)
            return srv_error(s, VE_ENOTFOUND, "unknown or closing session %u", sessions[i]);
        if (ins[i].width != it->second.desc.width || ins[i].height != it->second.desc.height)
            return srv_error(s, VE_EINVAL, "frame %ux%u on a %ux%u session", ins[i].width, ins[i].height,
                             it->second.desc.width, it->second.desc.height);
    }
    int rc;
    for (uint32_t i = 0; i < n; i++) {
        Session &ss = s->sessions.find(sessions[i])->second;
        Frame f;
        f.tag = tags ? tags[i] : 0;
        f.index = ss.frames++;
        f.arrival_ms = arrival_ms && arrival_ms[i] >= 0 ? arrival_ms[i] : now;
        f.fc = fcs[i];
        f.in = ins[i];
        std::memcpy(f.ubo, ubos + 16 * (size_t)i, sizeof f.ubo);
        if (ss.state != State::Active) {
            ss.pending.push_back(f);
            continue;
        }
        if ((rc = flush_for(s, ss))) return rc;   // its earlier frame first
        // a session still busy with an earlier frame goes alone on its stream:
        // the batch's launch would otherwise wait for that frame
        if (!ss.in_flight.empty() && hipEventQuery(ss.in_flight.back().end) != hipSuccess) {
            if ((rc = launch(s, ss, f))) return rc;
            ss.in_flight.push_back(f);
            continue;
        }
        s->ready.emplace_back(ss.id, f);
        ss.queued = true;
        if (s->ready.size() >= s->batch_cfg.max_views && (rc = flush_ready(s, true))) return rc;
    }
    return flush_ready(s, false);
}

int ve_server_close(ve_server *s, uint32_t session)
{
    if (!s) return ve_global_error(VE_EINVAL, "null server");
    auto it = s->sessions.find(session);
    if (it == s->sessions.end()) return srv_error(s, VE_ENOTFOUND, "unknown session %u", session);
    Session &ss = it->second;
    ss.close_requested = true;
    if (ss.state == State::Waiting && ss.pending.empty()) {   // never admitted and nothing queued
        s->waiting.erase(std::find(s->waiting.begin(), s->waiting.end(), session));
        s->sessions.erase(it);
        return admit(s);
    }
    // a waiting session keeps its place: its queued frames run once it is
    // admitted, and it retires after them like an admitted one
    SRV_HIP(s, hipSetDevice(s->device));
    collect(s);
    return admit(s);
}

int ve_server_poll(ve_server *s, ve_frame_done *out, uint32_t max, uint32_t *n_out, int wait)
{
    if (!s || !n_out || (max && !out)) return ve_global_error(VE_EINVAL, "null argument");
    SRV_HIP(s, hipSetDevice(s->device));
    collect(s);
    int rc = admit(s);
    if (rc != VE_OK) return rc;
    if ((rc = flush_ready(s, wait && s->done.empty())) != VE_OK) return rc;
    if (wait && s->done.empty()) {
        // block on the oldest frame in flight
        const Frame *oldest = nullptr;
        for (auto &kv : s->sessions)
            if (!kv.second.in_flight.empty() && (!oldest || kv.second.in_flight.front().seq < oldest->seq))
                oldest = &kv.second.in_flight.front();
        if (oldest) {
            SRV_HIP(s, hipEventSynchronize(oldest->end));
            collect(s);
            if ((rc = admit(s)) != VE_OK) return rc;
        }
    }
    uint32_t n = 0;
    while (n < max && !s->done.empty()) {
        out[n++] = s->done.front();
        s->done.pop_front();
    }
    *n_out = n;
    return VE_OK;
}

int ve_server_read_output(ve_server *s, uint32_t session, float *rgba, size_t bytes)
{
    if (!s || !rgba) return ve_global_error(VE_EINVAL, "null argument");
    auto it = s->sessions.find(session);
    if (it == s->sessions.end() || !it->second.ctx)
        return srv_error(s, VE_ENOTFOUND, "session %u holds no device state", session);
    Session &ss = it->second;
    const size_t need = (size_t)ss.desc.width * ss.desc.height * 16;
    if (bytes > need) return srv_error(s, VE_EINVAL, "read of %zu bytes from a %zu-byte output", bytes, need);
    SRV_HIP(s, hipSetDevice(s->device));
    if (int rc = flush_for(s, ss)) return rc;
    SRV_HIP(s, hipMemcpyAsync(rgba, ss.ctx->b.output, bytes, hipMemcpyDeviceToHost, ss.ctx->stream));
    SRV_HIP(s, hipStreamSynchronize(ss.ctx->stream));
    return VE_OK;
}

int ve_server_read(ve_server *s, uint32_t session, ve_buffer_id id, void *dst, size_t bytes)
{
    if (!s || !dst) return ve_global_error(VE_EINVAL, "null argument");
    auto it = s->sessions.find(session);
    if (it == s->sessions.end() || !it->second.ctx)
        return srv_error(s, VE_ENOTFOUND, "session %u holds no device state", session);
    SRV_HIP(s, hipSetDevice(s->device));
    int rc = flush_for(s, it->second);
    if (rc != VE_OK) return rc;
    rc = ve_read(it->second.ctx, id, dst, bytes);   // ordered after the session's frames
    if (rc != VE_OK) return srv_error(s, rc, "session %u: %s", session, it->second.ctx->err.c_str());
    return VE_OK;
}

int ve_server_get_stats(ve_server *s, ve_server_stats *o)
{
    if (!s || !o) return ve_global_error(VE_EINVAL, "null argument");
    std::memset(o, 0, sizeof *o);
    o->admitted = s->admitted;
    o->waiting = (uint32_t)s->waiting.size();
    for (auto &kv : s->sessions) o->frames_in_flight += (uint32_t)kv.second.in_flight.size();
    o->frames_in_flight += (uint32_t)s->ready.size();
    o->pooled = (uint32_t)s->pool.size();
    o->bytes_in_use = s->used;
    o->budget = s->budget;
    return VE_OK;
}

}  // extern "C"
