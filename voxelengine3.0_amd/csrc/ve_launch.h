// Host-side launch wrappers for the kernels in ve_kernels.hip.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ve_internal.h"

namespace ve {

// Timing of one pass without marker packets in the stream: the events are
// bound to the pass's own dispatches.  The first kernel launched records
// `start` as it begins; every kernel records `stop` as it ends, so the last
// one's stands.  Set around a timed pass, null otherwise.
struct PassTimer {
    hipEvent_t start = nullptr, stop = nullptr;
};
extern thread_local PassTimer pass_timer;

template <typename K, typename... A>
inline void launch(K kernel, dim3 grid, dim3 block, uint32_t shared_bytes, hipStream_t s, A... args)
{
    hipExtLaunchKernelGGL(kernel, grid, block, shared_bytes, s, pass_timer.start, pass_timer.stop, 0u, args...);
    pass_timer.start = nullptr;
}

// Device buffers of one context (all owned by the context).
struct DeviceBuffers {
    uint4 *gbuf = nullptr;          // P x (checks, material, 0, vid): the current set
    uint32_t *gpos = nullptr;       // P: the current set
    uint2 *ghit = nullptr;          // P: the primary hit's lattice position (x | y << 16, z), current set
    uint4 *gbuf_set[2] = {nullptr, nullptr};     // second set only with cross-frame overlap
    uint32_t *gpos_set[2] = {nullptr, nullptr};
    uint2 *ghit_set[2] = {nullptr, nullptr};
    float4 *output = nullptr;       // P
    uchar4 *display = nullptr;      // P: tonemapped RGBA8
    float4 *pix_avg = nullptr;      // P: per-pixel voxel average (rgb, indirect variance)
    uint32_t *pix_cnt = nullptr;    // P: per-pixel voxel sample count
    uint32_t *timeouts = nullptr;   // P: timed-out pixels (resolve)
    uint32_t *timeout_tiles = nullptr;  // per 16x16 tile: holds a timed-out pixel (zero between frames)
    uint32_t *timeout_list = nullptr;   // the flagged 16x16 tiles, in listing order
    uint32_t *nbuf = nullptr;       // hash_width x 96: the current set
    uint32_t *nbuf_set[2] = {nullptr, nullptr};  // second set only with cross-frame overlap
    uint32_t *lbuf = nullptr;       // hash_width x 384
    uint32_t *rbuf = nullptr;       // hash_width x 128
    uint32_t *claim = nullptr;      // CLAIM_WIDTH x CLAIM_PROBES
    uint32_t *samples = nullptr;    // CLAIM_WIDTH x CLAIM_PROBES
    uint4 *candidates = nullptr;    // CANDIDATE_CAPACITY
    uint2 *shade_list = nullptr;    // SHADE_BUDGET_MAX
    Ray *rays = nullptr;            // SHADE_BUDGET_MAX (current bounce)
    Ray *rays_next = nullptr;       // SHADE_BUDGET_MAX (survivors)
    uint32_t *counters = nullptr;   // NUM_COUNTERS + COUNTER_SIDE_WORDS
    uint32_t *rank_stats = nullptr; // RANK_STATS_WORDS
    uint32_t *rank_part = nullptr;  // RANK_COPIES x RANK_COPY_WORDS (zero between rank passes)
    uint32_t *bounce_part = nullptr;  // BOUNCE_COPIES x 2 + the ray queue (zero between fused bounce passes)
    EmissiveVoxel *emissives = nullptr;  // EMISSIVE_CAPACITY
    uint32_t *emissive_claim = nullptr;  // EMISSIVE_CLAIM_W x EMISSIVE_CLAIM_PROBES
    uint32_t *em_bits = nullptr;         // SHADE_BUDGET_MAX / 32: first-shade rays that may register a light
    uint32_t *em_reg = nullptr;          // EMISSIVE_CAPACITY: ray index of each registration this frame
    uint32_t *hit_bits = nullptr;        // SHADE_BUDGET_MAX / 32: bounce-0 rays whose hit record emit filled in
    // the frame's deposits, replayed per lBuffer bucket in the reference's order
    Deferred *deferred = nullptr;        // SHADE_BUDGET_MAX: in append order
    uint32_t *ray_li[2] = {nullptr, nullptr};  // SHADE_BUDGET_MAX each: shade-list index of each ray of rays / rays_next
    uint32_t *dep_death = nullptr;       // SHADE_BUDGET_MAX: bounce iteration at which each list entry's ray deposited
    uint32_t *dep_wpos = nullptr;        // SHADE_BUDGET_MAX: its position in the reference's ray window then
    uint32_t *dep_cnt = nullptr;         // hash_width: deposits per lBuffer bucket (zero between frames)
    uint32_t *dep_slots = nullptr;       // hash_width x DEP_CAP: record indices of each bucket's deposits
    uint2 *dep_over = nullptr;           // SHADE_BUDGET_MAX: {bucket, record} past a bucket's DEP_CAP
    uint32_t *dep_over_n = nullptr;      // 2: overflow entries, spill words handed out (zero between frames)
    uint64_t *dep_spill = nullptr;       // SHADE_BUDGET_MAX: sort keys of buckets past DEP_CAP
    Material *materials = nullptr;       // MATERIAL_SLOTS
    unsigned long long *node_stats = nullptr;  // 4: primary reads, trace reads, primary steps, trace steps
    unsigned long long *ray_totals = nullptr;  // 4: primary, traced, shadow, traced with a reused hit (cumulative)
    // schedule / rank ordering scratch
    uint4 *sched_tab = nullptr;          // CLAIM_WIDTH x SCHED_SLOTS: {voxel id, ~lowest execution rank, pixels, 0}
    uint4 *sched_over = nullptr;         // SCHED_OVER_CAP: {voxel id, ~rank, pixels, claim row} past a full row table
    uint32_t *sched_over_n = nullptr;    // 2: overflow records, flagged rows (zero between frames)
    uint32_t *sched_row_over = nullptr;  // CLAIM_WIDTH: 1 where the row has overflow records (zero between frames)
    uint32_t *sched_rows = nullptr;      // CLAIM_WIDTH: the flagged rows
    uint32_t *cand_at = nullptr;         // per execution rank: claim row/probe of the voxel it first claims
    uint32_t *vec_cnt = nullptr;         // per SIMD-8 vector: entries appended
    uint32_t *vec_chunk = nullptr;       // per 256 SIMD-8 vectors: sum of vec_cnt (schedule)
    uint32_t *rank_chunk = nullptr;      // per 256 candidates: shade-list entries they append (rank)
    int32_t *nsamp = nullptr;            // per candidate: samples
    // ordered per-bucket replay of the normal pass (events of one frame)
    uint32_t ord_cap = 0;                // events per bucket in place: a power of two >= max(64, pixels / hash_width)
    uint32_t ord_over_cap = 0;           // = ord_spill_cap = max_w * max_h
    uint32_t ord_spill_cap = 0;
    uint32_t *ord_cnt = nullptr;         // hash_width: events per nBuffer bucket (zero between passes)
    uint4 *ord_ev = nullptr;             // hash_width x ord_cap: {execution rank, 0, voxel id, packed position}
    uint4 *ord_over = nullptr;           // ord_over_cap: {execution rank, bucket, voxel id, packed position}
    uint32_t *ord_over_n = nullptr;      // 2: overflow events, spill words handed out (zero between passes)
    uint64_t *ord_spill = nullptr;       // ord_spill_cap: sort keys of buckets past ord_cap
    uint32_t *ord_work = nullptr;        // hash_width x 32: packed position of each slot's new normal, or NONE
    FrameDyn *dyn = nullptr;             // 1: per-frame inputs of a graph replay
};

void launch_primary(hipStream_t s, const uint32_t *words, const FrameParams &p, DeviceBuffers &b, bool stats);

// ---- one launch per pass over many views (the whole-frame batched form) ----
// A view of such a launch: its frame parameters, its context's buffers, its
// scene and the frame's bounce iterations (renderer.cpp:515).  The table and
// the block offsets live in device memory.
struct ViewRec {
    FrameParams p;
    DeviceBuffers b;
    const uint32_t *words;
    uint32_t iters;
};
struct MultiTab {
    const ViewRec *views;
    const uint32_t *pre;       // pre[v]: first block of view v; pre[n]: the grid
    uint32_t n;
};
// The kernels after the primary pass, in the order a frame runs them
// (renderer.cpp:282-649 with the threshold and the deposit-record clear folded
// into the emit and resolve kernels).
enum MultiKernel : uint32_t {
    MK_RESET, MK_NRM_EVENTS, MK_NRM_REPLAY, MK_NRM_COMPUTE, MK_SCHED_CLAIM, MK_SCHED_RANK, MK_SCHED_RANK_OVER,
    MK_SCHED_VEC_COUNT, MK_SCHED_WRITE, MK_RANK, MK_RANK_WRITE, MK_EMIT, MK_TRACE, MK_EM_REGISTER, MK_SHADE,
    MK_BOUNCE, MK_DEP_PREP, MK_DEP_WPOS, MK_DEP_REPLAY, MK_RESOLVE_AVG, MK_RESOLVE, MK_RESOLVE_TIMEOUTS, MK_COUNT
};
// Blocks of kernel k in one view's whole frame (the grid ve_frame launches it
// with; 0 when that frame skips it).  Frames of such launches run with the
// small passes folded and the full bounce grid.
uint32_t multi_blocks(uint32_t k, const FrameParams &p, uint32_t iters);
// Dynamic LDS bytes of one view's timeout reconstruction launch.
uint32_t timeouts_lds(const FrameParams &p);
// Every pass after the primary pass of n views, one launch per kernel on `s`:
// `views` (n records) and `pre` (MK_COUNT rows of n + 1 block offsets, row k
// for kernel k) in device memory; grids[k] = row k's last entry (launches
// with an empty grid are skipped); `lds` = the largest timeouts_lds.
void launch_frames_multi(hipStream_t s, const ViewRec *views, const uint32_t *pre, uint32_t n, const uint32_t *grids,
                         uint32_t lds);

// One view of a batched primary pass: what generate_primary.comp reads for
// it, and where its G-buffer goes.
struct BatchView {
    const uint32_t *words;
    const uint32_t *jump;
    uint32_t n_words, depth, jump_level, cc;
    uint32_t W, H, tiles_x, reuse;
    float cam[16];
    uint4 *gbuf;
    uint32_t *gpos;
    uint2 *ghit;
};
constexpr uint32_t BATCH_TILE_WIDTHS[4] = {8u, 16u, 32u, 64u};
// The primary passes of n views in one persistent launch of `blocks` blocks
// of 4 waves: bases[v] = first tile of view v in the concatenated tile list
// (bases[n] = total), tiles of tile_w x (64 / tile_w) pixels, each wave
// taking runs of `grab` tiles from the counter *queue (zero at launch).
void launch_primary_batch(hipStream_t s, const BatchView *views, const uint32_t *bases, uint32_t n, uint32_t total,
                          uint32_t *queue, uint32_t tile_w, uint32_t grab, uint32_t blocks);
void launch_fill(hipStream_t s, uint32_t *p, uint32_t v, size_t n);
// prev: the nBuffer set the pass starts from when it writes into another
// one (b.nbuf), else null
void launch_normal(hipStream_t s, const uint32_t *words, const FrameParams &p, DeviceBuffers &b,
                   const uint32_t *prev = nullptr);
void launch_reset(hipStream_t s, const FrameParams &p, DeviceBuffers &b);
int launch_schedule(hipStream_t s, const FrameParams &p, DeviceBuffers &b);
int launch_rank(hipStream_t s, const FrameParams &p, DeviceBuffers &b);
void launch_emit(hipStream_t s, const FrameParams &p, DeviceBuffers &b);
void launch_threshold(hipStream_t s, const FrameParams &p, DeviceBuffers &b);
void launch_trace(hipStream_t s, const uint32_t *words, const FrameParams &p, DeviceBuffers &b, uint32_t iter,
                  bool stats);
void launch_shade(hipStream_t s, const uint32_t *words, const FrameParams &p, DeviceBuffers &b, uint32_t iter);
// Bounce iterations first..iters-1 fused per ray (one launch), then their
// held-back deposits in iteration order.
void launch_bounce(hipStream_t s, const uint32_t *words, const FrameParams &p, DeviceBuffers &b, uint32_t first,
                   uint32_t iters, bool stats);
// The frame's deposits (records of every shade pass and the fused bounce
// pass, bounce iterations 0..iters-1), replayed into the lBuffer in the
// reference's order.
int launch_deposits(hipStream_t s, const FrameParams &p, DeviceBuffers &b, uint32_t iters);
void launch_resolve(hipStream_t s, const FrameParams &p, DeviceBuffers &b);
void launch_dep_done(hipStream_t s, const FrameParams &p, DeviceBuffers &b);
void launch_display(hipStream_t s, const FrameParams &p, DeviceBuffers &b);
void launch_scatter_words(hipStream_t s, uint32_t *words, const uint2 *pairs, uint32_t n);
// The jump table of a node array (8^level entries).
void launch_build_jump(hipStream_t s, const uint32_t *words, uint32_t n_words, uint32_t depth, uint32_t level,
                       uint32_t *jump);

}  // namespace ve
