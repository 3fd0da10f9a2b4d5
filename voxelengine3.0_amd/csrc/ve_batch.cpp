// The batched form: frame k of many views at once.  Their primary passes run
// in one persistent launch that pulls tiles from a global queue over all the
// views (launch_primary_batch); every other pass of each view runs on its own
// context's stream after that launch -- replayed from the context's captured
// frame graph when graph replay is on.  No counterpart in the reference,
// which renders one view per Renderer::run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ve_ctx.h"
#include "ve_internal.h"
#include "ve_launch.h"

struct ve_batch {
    int device = 0;
    uint32_t max_views = 0;
    ve_batch_config cfg{};
    uint32_t blocks = 0;                     // persistent grid
    hipStream_t stream = nullptr;
    // one block on the device and in each staging set, filled by one copy
    // (Layout): queue counters, tile bases, primary views and, for whole
    // frames, the view records and block offsets of the other passes
    size_t block_bytes = 0;
    unsigned char *dev = nullptr;
    // pinned staging sets: a set is rewritten only once its copy has run (a
    // query, never a host wait: a serving loop must not block); a new set is
    // added when none is free
    struct Staging {
        unsigned char *host = nullptr;
        hipEvent_t copied = nullptr;
    };
    std::vector<Staging> staging;
    std::vector<hipEvent_t> ctx_ev;          // per view: its stream's work before the batch
    hipEvent_t done[64] = {};                // per group: its primary launch finished
    hipEvent_t t0 = nullptr, t1 = nullptr;   // around the primary launch
    bool launched = false;                   // t1 recorded at least once
    std::vector<ve::FrameParams> params;
    // whole-frame batches of graph-replaying contexts: the batch's launches
    // captured once per launch shape (the views' records and block offsets
    // are data in the device block, copied before each replay)
    hipGraphExec_t wexec = nullptr;
    std::vector<uint32_t> wkey;
    uint64_t w_captures = 0, w_replays = 0;
    std::string err;
};

namespace {

int batch_error(ve_batch *b, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    return code;
}

#define B_HIP(b, call)                                                                              \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) return batch_error(b, VE_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

constexpr uint32_t MAX_GROUPS = 64;

bool tile_width_ok(uint32_t w)
{
    for (uint32_t v : ve::BATCH_TILE_WIDTHS)
        if (v == w) return true;
    return false;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Byte offsets in the staging / device block of a batch of n views: queue
// counters (MAX_GROUPS, zero), tile bases (n + MAX_GROUPS), n primary views,
// then (whole frames) n view records and MK_COUNT rows of n + 1 block offsets.
struct Layout {
    size_t views, recs, pre, end;
    explicit Layout(uint32_t n)
    {
        views = align256((2 * (size_t)MAX_GROUPS + n) * sizeof(uint32_t));
        recs = align256(views + (size_t)n * sizeof(ve::BatchView));
        pre = align256(recs + (size_t)n * sizeof(ve::ViewRec));
        end = pre + (size_t)ve::MK_COUNT * (n + 1u) * sizeof(uint32_t);
    }
};

}  // namespace

namespace ve {

// The batch's primary launches for views whose parameters frame_begin filled.
// They wait for everything already on the views' streams.  The views, largest
// first, are split into groups of about equal work, one persistent launch
// each on the batch stream; a view's stream waits for its group's launch only,
// so the views of early groups run their other passes beside the later
// launches.  `starts` (optional, one per view) are recorded on the batch
// stream before the first launch.  With `whole`, the launch is one group and
// every other pass of every view follows it on the batch stream, one launch
// per kernel over all the views (launch_frames_multi); the views' streams
// then wait for the last of them.
int batch_primary(ve_batch *b, uint32_t n, ve_ctx *const *ctxs, ve_scene *const *scenes, const FrameParams *ps,
                  hipEvent_t *starts, bool whole)
{
    if (n > b->max_views) return batch_error(b, VE_EINVAL, "%u views above the batch's %u", n, b->max_views);
    const uint32_t tw = b->cfg.tile_w, th = 64u / tw;
    std::vector<uint32_t> tiles(n), order(n);
    uint64_t all = 0;
    for (uint32_t i = 0; i < n; i++) {
        tiles[i] = ((ps[i].W + tw - 1) / tw) * ((ps[i].H + th - 1) / th);
        all += tiles[i];
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) { return tiles[a] > tiles[c]; });
    const uint32_t groups = whole ? 1u : std::max(1u, std::min(std::min(b->cfg.groups, n), MAX_GROUPS));
    // group g: sorted views [first[g], first[g + 1]), cut where the running work passes (g + 1) / groups
    std::vector<uint32_t> first(groups + 1, n);
    first[0] = 0;
    uint64_t run = 0;
    uint32_t g = 1;
    for (uint32_t k = 0; k < n && g < groups; k++) {
        run += tiles[order[k]];
        if (run * groups >= all * g && k + 1 < n) first[g++] = k + 1;
    }
    for (; g < groups; g++) first[g] = n;   // fewer cuts than groups: the rest are empty
    ve_batch::Staging *st = nullptr;
    for (auto &x : b->staging)
        if (hipEventQuery(x.copied) == hipSuccess) {
            st = &x;
            break;
        }
    if (!st) {
        ve_batch::Staging x;
        void *h = nullptr;
        B_HIP(b, hipHostMalloc(&h, b->block_bytes, hipHostMallocDefault));
        x.host = static_cast<unsigned char *>(h);
        B_HIP(b, hipEventCreateWithFlags(&x.copied, hipEventDisableTiming));
        b->staging.push_back(x);
        st = &b->staging.back();
    }
    const Layout L(n);
    uint32_t *qh = reinterpret_cast<uint32_t *>(st->host);
    uint32_t *bh = qh + MAX_GROUPS;   // group g's bases at first[g] + g, first[g + 1] - first[g] + 1 of them
    BatchView *vh = reinterpret_cast<BatchView *>(st->host + L.views);
    uint32_t *queue_dev = reinterpret_cast<uint32_t *>(b->dev), *bases_dev = queue_dev + MAX_GROUPS;
    BatchView *views_dev = reinterpret_cast<BatchView *>(b->dev + L.views);
    uint32_t grids[MK_COUNT] = {0};
    uint32_t lds = 0;
    if (whole) {   // every view's record and its place in each pass's grid
        ViewRec *rh = reinterpret_cast<ViewRec *>(st->host + L.recs);
        uint32_t *ph = reinterpret_cast<uint32_t *>(st->host + L.pre);
        for (uint32_t i = 0; i < n; i++) {
            ViewRec &r = rh[i];
            r.p = ps[i];
            r.b = ctxs[i]->b;
            r.words = scenes[i]->words;
            r.iters = std::min<uint32_t>(ctxs[i]->max_bounces, (uint32_t)MAX_ITERS);   // renderer.cpp:515
            lds = std::max(lds, timeouts_lds(ps[i]));
        }
        for (uint32_t k = 0; k < MK_COUNT; k++) {
            uint32_t *row = ph + (size_t)k * (n + 1u), at = 0;
            for (uint32_t i = 0; i < n; i++) {
                row[i] = at;
                const uint64_t next = (uint64_t)at + multi_blocks(k, ps[i], rh[i].iters);
                if (next > 0x7FFFFFFFull) return batch_error(b, VE_ECAPACITY, "whole-frame batch grid above 2^31 blocks");
                at = (uint32_t)next;
            }
            row[n] = at;
            grids[k] = at;
        }
    }
    std::memset(qh, 0, MAX_GROUPS * sizeof(uint32_t));
    std::vector<uint32_t> totals(groups, 0);
    for (uint32_t gi = 0; gi < groups; gi++) {
        uint32_t total = 0;
        for (uint32_t j = first[gi]; j < first[gi + 1]; j++) {
            const uint32_t i = order[j];
            const FrameParams &p = ps[i];
            BatchView &v = vh[j];
            v.words = scenes[i]->words;
            v.jump = p.jump;
            v.n_words = p.n_words;
            v.depth = p.depth;
            v.jump_level = p.jump_level;
            v.cc = p.primary_cc;
            v.W = p.W;
            v.H = p.H;
            v.tiles_x = (p.W + tw - 1) / tw;
            v.reuse = (uint32_t)p.reuse_primary_hits;
            std::memcpy(v.cam, p.cam, sizeof v.cam);
            v.gbuf = ctxs[i]->b.gbuf;
            v.gpos = ctxs[i]->b.gpos;
            v.ghit = ctxs[i]->b.ghit;
            bh[j + gi] = total;
            total += tiles[i];
        }
        bh[first[gi + 1] + gi] = total;
        totals[gi] = total;
    }
    // wait only for streams with work still queued (an idle one needs no wait)
    for (uint32_t i = 0; i < n; i++) {
        const hipError_t q = hipStreamQuery(ctxs[i]->stream);
        if (q == hipSuccess) continue;
        if (q != hipErrorNotReady) return batch_error(b, VE_EDEVICE, "hipStreamQuery: %s", hipGetErrorString(q));
        B_HIP(b, hipEventRecord(b->ctx_ev[i], ctxs[i]->stream));
        B_HIP(b, hipStreamWaitEvent(b->stream, b->ctx_ev[i], 0));
    }
    B_HIP(b, hipMemcpyAsync(b->dev, st->host, whole ? L.end : L.recs, hipMemcpyHostToDevice, b->stream));
    B_HIP(b, hipEventRecord(st->copied, b->stream));
    if (starts)
        for (uint32_t i = 0; i < n; i++) B_HIP(b, hipEventRecord(starts[i], b->stream));
    B_HIP(b, hipEventRecord(b->t0, b->stream));
    auto primary_blocks = [&](uint32_t gi) {   // no more blocks than the group's runs of tiles need (one run per wave)
        const uint64_t runs = ((uint64_t)totals[gi] + b->cfg.grab - 1) / b->cfg.grab;
        return (uint32_t)std::min<uint64_t>(b->blocks, (runs + 3) / 4);
    };
    if (whole) {
        // one group: the primary launch, then every later pass over all the views
        auto launches = [&]() -> int {
            launch_primary_batch(b->stream, views_dev, bases_dev, n, totals[0], queue_dev, tw, b->cfg.grab,
                                 primary_blocks(0));
            B_HIP(b, hipGetLastError());
            launch_frames_multi(b->stream, reinterpret_cast<const ViewRec *>(b->dev + L.recs),
                                reinterpret_cast<const uint32_t *>(b->dev + L.pre), n, grids, lds);
            B_HIP(b, hipGetLastError());
            return VE_OK;
        };
        bool graph = true;   // replayed from a captured graph when every view's context replays graphs
        for (uint32_t i = 0; i < n; i++) graph = graph && ctxs[i]->graph;
        if (graph) {
            std::vector<uint32_t> key{n, tw, b->cfg.grab, primary_blocks(0), totals[0], lds};
            key.insert(key.end(), grids, grids + MK_COUNT);
            if (!b->wexec || key != b->wkey) {
                if (b->wexec) {
                    B_HIP(b, hipGraphExecDestroy(b->wexec));
                    b->wexec = nullptr;
                }
                B_HIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
                const int rc = launches();
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(b->stream, &g);
                if (rc || e != hipSuccess) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc ? rc : batch_error(b, VE_EDEVICE, "graph capture: %s", hipGetErrorString(e));
                }
                const hipError_t ei = hipGraphInstantiate(&b->wexec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ei != hipSuccess) {
                    b->wexec = nullptr;
                    return batch_error(b, VE_EDEVICE, "graph instantiate: %s", hipGetErrorString(ei));
                }
                b->wkey = std::move(key);
                b->w_captures++;
            }
            B_HIP(b, hipGraphLaunch(b->wexec, b->stream));
            b->w_replays++;
        } else if (int rc = launches()) {
            return rc;
        }
        B_HIP(b, hipEventRecord(b->t1, b->stream));   // the whole batch (its primary launch is inside it)
        B_HIP(b, hipEventRecord(b->done[0], b->stream));
        for (uint32_t i = 0; i < n; i++) B_HIP(b, hipStreamWaitEvent(ctxs[i]->stream, b->done[0], 0));
        b->launched = true;
        return VE_OK;
    }
    for (uint32_t gi = 0; gi < groups; gi++) {
        const uint32_t m = first[gi + 1] - first[gi];
        if (m == 0) continue;
        launch_primary_batch(b->stream, views_dev + first[gi], bases_dev + first[gi] + gi, m, totals[gi],
                             queue_dev + gi, tw, b->cfg.grab, primary_blocks(gi));
        B_HIP(b, hipGetLastError());
        B_HIP(b, hipEventRecord(b->done[gi], b->stream));
        for (uint32_t j = first[gi]; j < first[gi + 1]; j++)
            B_HIP(b, hipStreamWaitEvent(ctxs[order[j]]->stream, b->done[gi], 0));
    }
    B_HIP(b, hipEventRecord(b->t1, b->stream));
    b->launched = true;
    return VE_OK;
}

bool batch_idle(ve_batch *b) { return !b->launched || hipEventQuery(b->t1) != hipErrorNotReady; }

bool batch_whole(const ve_batch *b) { return b->cfg.whole_frame != 0; }

const char *batch_error_text(const ve_batch *b) { return b->err.c_str(); }

}  // namespace ve

extern "C" {

int ve_batch_create(ve_batch **out, const ve_batch_config *cfg)
{
    if (!out || !cfg) return ve_global_error(VE_EINVAL, "null argument");
    ve_batch_config c = *cfg;
    if (c.max_views == 0) c.max_views = 256;
    if (c.tile_w == 0) c.tile_w = 8;
    if (c.grab == 0) c.grab = 4;
    if (c.waves_per_cu == 0) c.waves_per_cu = 16;
    if (c.groups == 0) c.groups = 4;
    if (!tile_width_ok(c.tile_w)) return ve_global_error(VE_EINVAL, "tile_w must be 8, 16, 32 or 64");
    if (c.max_views > 65536 || c.grab > 1024 || c.waves_per_cu > 32 || c.groups > MAX_GROUPS)
        return ve_global_error(VE_EINVAL, "max_views above 65536, grab above 1024, waves_per_cu above 32 or groups above 64");
    ve_batch *b = new (std::nothrow) ve_batch();
    if (!b) return ve_global_error(VE_ENOMEM, "batch allocation");
    b->device = c.device;
    b->max_views = c.max_views;
    b->cfg = c;
    hipDeviceProp_t prop{};
    hipError_t e = hipSetDevice(c.device);
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, c.device);
    if (e == hipSuccess) {
        const uint32_t cus = prop.multiProcessorCount > 0 ? (uint32_t)prop.multiProcessorCount : 256u;
        b->blocks = (cus * c.waves_per_cu + 3u) / 4u;   // 4 waves per block
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);   // default priority
    b->block_bytes = Layout(c.max_views).end;
    if (e == hipSuccess) e = hipMalloc(&b->dev, b->block_bytes);
    b->ctx_ev.assign(c.max_views, nullptr);
    for (uint32_t i = 0; i < c.max_views && e == hipSuccess; i++)
        e = hipEventCreateWithFlags(&b->ctx_ev[i], hipEventDisableTiming);
    for (uint32_t g = 0; g < MAX_GROUPS && e == hipSuccess; g++)
        e = hipEventCreateWithFlags(&b->done[g], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreate(&b->t0);
    if (e == hipSuccess) e = hipEventCreate(&b->t1);
    if (e != hipSuccess) {
        ve_global_error(VE_EDEVICE, hipGetErrorString(e));
        ve_batch_destroy(b);
        return VE_EDEVICE;
    }
    *out = b;
    return VE_OK;
}

void ve_batch_destroy(ve_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (hipEvent_t e : b->ctx_ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->done)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {b->t0, b->t1})
        if (e) (void)hipEventDestroy(e);
    if (b->wexec) (void)hipGraphExecDestroy(b->wexec);
    for (auto &x : b->staging) {
        if (x.copied) (void)hipEventDestroy(x.copied);
        if (x.host) (void)hipHostFree(x.host);
    }
    if (b->dev) (void)hipFree(b->dev);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char *ve_batch_last_error(const ve_batch *b) { return b ? b->err.c_str() : ve_last_error(nullptr); }

int ve_batch_frame(ve_batch *b, uint32_t n, ve_ctx *const *ctxs, ve_scene *const *scenes, const ve_frame_config *fcs,
                   const ve_frame_inputs *ins)
{
    if (!b) return ve_global_error(VE_EINVAL, "null batch");
    if (n == 0) return VE_OK;
    if (!ctxs || !scenes || !fcs || !ins) return batch_error(b, VE_EINVAL, "null argument");
    if (n > b->max_views) return batch_error(b, VE_EINVAL, "%u views above the batch's %u", n, b->max_views);
    b->params.resize(n);
    for (uint32_t i = 0; i < n; i++) {
        if (!ctxs[i]) return batch_error(b, VE_EINVAL, "null context %u", i);
        if (ctxs[i]->device != b->device) return batch_error(b, VE_EINVAL, "view %u is on another device", i);
        for (uint32_t j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i]) return batch_error(b, VE_EINVAL, "context of view %u repeated", i);
        const int rc = ve::frame_begin(ctxs[i], scenes[i], &fcs[i], &ins[i], b->params[i]);
        if (rc) return batch_error(b, rc, "view %u: %s", i, ctxs[i]->err.c_str());
    }
    B_HIP(b, hipSetDevice(b->device));
    const bool whole = b->cfg.whole_frame != 0;
    int rc = ve::batch_primary(b, n, ctxs, scenes, b->params.data(), nullptr, whole);
    if (rc || whole) return rc;
    for (uint32_t i = 0; i < n; i++)
        if ((rc = ve::frame_rest(ctxs[i], scenes[i], b->params[i])))
            return batch_error(b, rc, "view %u: %s", i, ctxs[i]->err.c_str());
    return VE_OK;
}

int ve_batch_graph_stats(ve_batch *b, uint64_t *captures, uint64_t *replays)
{
    if (!b) return ve_global_error(VE_EINVAL, "null batch");
    if (captures) *captures = b->w_captures;
    if (replays) *replays = b->w_replays;
    return VE_OK;
}

int ve_batch_primary_ms(ve_batch *b, float *ms)
{
    if (!b || !ms) return ve_global_error(VE_EINVAL, "null argument");
    B_HIP(b, hipSetDevice(b->device));
    B_HIP(b, hipEventSynchronize(b->t1));
    B_HIP(b, hipEventElapsedTime(ms, b->t0, b->t1));
    return VE_OK;
}

}  // extern "C"
