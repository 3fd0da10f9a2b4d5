// Internal definitions of the C ABI's opaque handles and the host entry
// points the session server (ve_server.cpp) shares with ve_api.cpp.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ve_svo.h"
#include "ve_internal.h"
#include "ve_launch.h"

// A resident octree: the node words in device memory (Octree::set's buffer
// texture, octree.cpp:157-166).
struct ve_scene {
    uint32_t *words = nullptr;
    uint32_t n_words = 0;
    uint32_t alloc_words = 0;          // words allocated (>= n_words)
    uint32_t depth = 0;
    size_t device_bytes = 0;
    uint2 *staging = nullptr;          // (slot, word) pairs of ve_scene_update
    size_t staging_cap = 0;
    // Jump table: for each cell of the grid at depth jump_level (8^level
    // cells), the outcome of the descent from the root through that depth
    // (rebuilt whenever the words change).
    uint32_t *jump = nullptr;
    uint32_t jump_level = 0;
};

constexpr int MAX_ITERS = 8;   // host bounce-loop cap (renderer.cpp:515)

// One renderer: the per-view device state of Renderer (renderer.cpp:20-169)
// on one HIP stream.
struct ve_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = true;            // false: a stream shared with other contexts
    uint32_t max_w = 0, max_h = 0;
    uint32_t hash_width = 16384;
    uint32_t max_bounces = 8;
    std::string err;
    float cam[16] = {0};
    float sky[3] = {0, 0, 0};            // constant sky (sky_edge 0)
    uint32_t sky_edge = 0;               // cube-map sky: texels per face edge
    uint32_t *sky_tex = nullptr;         // 6 x sky_edge^2 texels, r | g << 8 | b << 16
    size_t sky_tex_words = 0;            // allocated
    bool node_stats = false;
    ve_tuning tuning{ve::DEFAULT_TILE_CHUNK, 0u, 0u};
    bool timing = false;
    uint32_t bounce_iter = 0;           // next trace/shade iteration of per-pass dispatch
    // Deposits recorded by pass-by-pass shade dispatches and not yet applied
    // (the reference's chunk order needs all of a frame's deposits): the
    // iterations recorded and the parameters of the frame that recorded them.
    uint32_t dep_iters = 0;
    ve::FrameParams dep_params{};
    // timing: one event pair per pass and bounce iteration per recorded frame
    std::vector<std::array<hipEvent_t, 2 * VE_PASS_COUNT * MAX_ITERS>> frame_events;
    size_t frames_recorded = 0;
    uint32_t iters_recorded = 0;
    ve::DeviceBuffers b;
    // Cross-frame overlap (contexts that own their stream): the primary pass,
    // which reads only the camera and the octree, runs on `aux` into the other
    // G-buffer set while the previous frame's latency-bound bounce passes
    // still run on `stream`, and the schedule pass follows it there: it reads
    // only the G-buffer and writes the claim map and candidate list, which no
    // pass after the previous frame's emit pass reads.  They start once the
    // previous frame's first shade pass is done (its emit pass, without
    // bounces).
    bool overlap = false;
    hipStream_t aux = nullptr;
    uint32_t gsel = 0;                      // G-buffer (and nBuffer) set of the latest frame
    hipEvent_t ev_primary = nullptr;        // on aux: primary pass written
    hipEvent_t ev_sched = nullptr;          // on aux: schedule pass written
    hipEvent_t ev_emit = nullptr;           // on stream: first shade (or emit) pass of the latest frame done
    hipEvent_t emit_done = nullptr;         // the event that marks it (ev_emit, or its timing event)
    hipEvent_t ev_fence = nullptr;          // on stream: everything enqueued so far
    // The normal pass updates the nBuffer, which shading and the resolve
    // read: inside an overlapped frame it runs on aux2 after the frame's
    // primary pass, into the nBuffer set of the frame's G-buffer set, which
    // its replay fills from the previous frame's set.  The main stream waits
    // for it (ev_normal) before the trace pass, and for the schedule pass
    // (ev_sched) before the rank pass.
    hipStream_t aux2 = nullptr;
    hipEvent_t ev_normal = nullptr;         // on aux2: normal pass written
    const uint32_t *nbuf_from = nullptr;    // while an overlapped frame enqueues its normal pass: the previous set
    // The resolve pass reads the lBuffer, the frame's G-buffer and nBuffer
    // sets and the materials, and writes only the image and its own scratch:
    // inside an overlapped frame it runs on aux3 after the frame's deposits
    // (ev_dep), beside the next frame's first passes.  The next deposits wait
    // for it on the main stream, the primary pass two frames on (the next
    // writer of its sets) on aux, and any other call on the context joins
    // it to the main stream first (resolve_pending).
    hipStream_t aux3 = nullptr;
    hipEvent_t ev_dep = nullptr;            // on stream: the frame's deposits applied
    hipEvent_t ev_resolve[2] = {};          // on aux3: the latest resolve of each buffer set done
    bool resolve_rec[2] = {false, false};   // ev_resolve[k] recorded
    hipEvent_t resolve_pending = nullptr;   // the resolve the main stream has not waited for yet
    bool aux_fence = true;                  // the next primary must follow all of stream's work
    // Graph replay (ve_set_graph): the frame's launches are captured once into
    // a HIP graph and replayed; the per-frame inputs travel in the device
    // block b.dyn, copied from a pinned staging ring before each replay.  A
    // new capture is made when anything else in the frame's parameters changes.
    static constexpr uint32_t DYN_RING = 8;
    bool graph = false;
    hipGraphExec_t gexec = nullptr;
    std::vector<uint8_t> gkey;              // parameters of the captured frame (per-frame inputs cleared)
    ve::FrameDyn *dyn_host = nullptr;       // pinned, DYN_RING slots
    hipEvent_t dyn_ev[DYN_RING] = {};       // the copy from each slot
    uint32_t dyn_next = 0;
    uint64_t graph_captures = 0, graph_replays = 0;
    // the frame's primary pass was run by a batch launch (ve_batch_frame)
    bool skip_primary = false;
    std::vector<std::pair<void *, size_t>> allocations;
    size_t device_bytes = 0;
};

namespace ve {

// Device bytes a context of this shape allocates.
size_t ctx_device_bytes(uint32_t max_w, uint32_t max_h, uint32_t hash_width);
// ve_create on a given stream (null: the context creates its own).
// `overlap`: give the context a second stream and G-buffer set for
// cross-frame overlap (ignored with a shared stream).
int ctx_create(ve_ctx **out, const ve_renderer_config *cfg, hipStream_t shared, bool overlap, std::string &err);
// Back to the state of a fresh context (all buffers zero, initial RankStats),
// enqueued on the context's stream without waiting.
int ctx_reset(ve_ctx *c);
// A frame whose primary pass runs in a batch (ve_batch_frame): the checks and
// parameters of ve_frame (the context must have the overlap, timing and node
// counting off), then every other pass on the context's stream, replayed
// from its graph when graph replay is on.
int frame_begin(ve_ctx *c, ve_scene *s, const ve_frame_config *fc, const ve_frame_inputs *in, FrameParams &p);
int frame_rest(ve_ctx *c, ve_scene *s, const FrameParams &p);
// The batched primary launch of ve_batch_frame for views prepared by
// frame_begin (ve_batch.cpp); `starts` (optional) are recorded just before it.
// With `whole`, every other pass of the views follows on the batch stream,
// one launch per kernel over all of them, and the views' streams wait for it.
int batch_primary(ve_batch *b, uint32_t n, ve_ctx *const *ctxs, ve_scene *const *scenes, const FrameParams *ps,
                  hipEvent_t *starts, bool whole = false);
// The batch runs whole frames (ve_batch_config::whole_frame).
bool batch_whole(const ve_batch *b);
const char *batch_error_text(const ve_batch *b);
// No primary launch of the batch still running (a query, no host wait).
bool batch_idle(ve_batch *b);
int scene_upload(int device, hipStream_t stream, const uint32_t *words, uint32_t n_words, uint32_t depth,
                 ve_scene **scene, std::string &err, uint32_t jump_level = JUMP_LEVEL);
void scene_free(ve_scene *s);
hipStream_t ctx_stream(ve_ctx *c);
size_t ctx_bytes(const ve_ctx *c);
void ctx_shape(const ve_ctx *c, uint32_t *w, uint32_t *h, uint32_t *hash_width, uint32_t *max_bounces);
unsigned long long *ctx_ray_totals(ve_ctx *c);

}  // namespace ve
