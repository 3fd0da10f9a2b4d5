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

// The deposits held back by the fused bounce pass, per lBuffer bucket: its
// first BIN_CAP in `binned` (row h), the rest listed in `over` and then
// grouped per bucket into `over_sorted` at `over_off[h]`.  A record's `iter`
// holds the bounce iteration in bits 0-7 and the bucket-local index above.
struct BinLists {
    uint32_t hash_width;
    uint32_t *count;
    Deferred *binned, *over, *over_sorted;
    uint32_t *over_off, *over_n;
};

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
    uint32_t *nbuf = nullptr;       // hash_width x 96
    uint32_t *lbuf = nullptr;       // hash_width x 384
    uint32_t *rbuf = nullptr;       // hash_width x 128
    uint32_t *claim = nullptr;      // CLAIM_WIDTH x CLAIM_PROBES
    uint32_t *samples = nullptr;    // CLAIM_WIDTH x CLAIM_PROBES
    uint4 *candidates = nullptr;    // CANDIDATE_CAPACITY
    uint2 *shade_list = nullptr;    // SHADE_BUDGET_MAX
    Ray *rays = nullptr;            // SHADE_BUDGET_MAX (current bounce)
    Ray *rays_next = nullptr;       // SHADE_BUDGET_MAX (survivors)
    uint32_t *counters = nullptr;   // NUM_COUNTERS
    uint32_t *rank_stats = nullptr; // RANK_STATS_WORDS
    uint32_t *rank_part = nullptr;  // RANK_COPIES x RANK_COPY_WORDS (zero between rank passes)
    uint32_t *bounce_part = nullptr;  // BOUNCE_COPIES x 2 (zero between fused bounce passes)
    EmissiveVoxel *emissives = nullptr;  // EMISSIVE_CAPACITY
    uint32_t *emissive_claim = nullptr;  // EMISSIVE_CLAIM_W x EMISSIVE_CLAIM_PROBES
    uint32_t *em_bits = nullptr;         // SHADE_BUDGET_MAX / 32: first-shade rays that may register a light
    uint32_t *em_reg = nullptr;          // EMISSIVE_CAPACITY: ray index of each registration this frame
    uint32_t *hit_bits = nullptr;        // SHADE_BUDGET_MAX / 32: bounce-0 rays whose hit record emit filled in
    // deposits held back by the fused bounce pass (see BinLists)
    Deferred *binned = nullptr;          // hash_width x BIN_CAP
    Deferred *deferred = nullptr;        // SHADE_BUDGET_MAX: in append order
    Deferred *bin_over = nullptr;        // SHADE_BUDGET_MAX: past BIN_CAP in a bucket, in placing order
    Deferred *bin_over_sorted = nullptr; // SHADE_BUDGET_MAX: the overflow grouped by lBuffer bucket
    uint32_t *bin_count = nullptr;       // hash_width: deposits per bucket (zero between frames)
    uint32_t *bin_off = nullptr;         // hash_width: overflow region of a bucket past BIN_CAP
    uint32_t *bin_over_n = nullptr;      // 2: overflow entries, overflow entries placed (zero between frames)
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
};

void launch_primary(hipStream_t s, const uint32_t *words, const FrameParams &p, DeviceBuffers &b, bool stats);
void launch_fill(hipStream_t s, uint32_t *p, uint32_t v, size_t n);
void launch_normal(hipStream_t s, const uint32_t *words, const FrameParams &p, DeviceBuffers &b);
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
int launch_deferred(hipStream_t s, const FrameParams &p, DeviceBuffers &b, uint32_t first, uint32_t iters);
void launch_resolve(hipStream_t s, const FrameParams &p, DeviceBuffers &b);
void launch_display(hipStream_t s, const FrameParams &p, DeviceBuffers &b);
void launch_scatter_words(hipStream_t s, uint32_t *words, const uint2 *pairs, uint32_t n);
// The jump table of a node array (8^level entries).
void launch_build_jump(hipStream_t s, const uint32_t *words, uint32_t n_words, uint32_t depth, uint32_t level,
                       uint32_t *jump);

}  // namespace ve
