// Host-side launch wrappers for the kernels in ve_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ve_internal.h"

namespace ve {

// Device buffers of one context (all owned by the context).
struct DeviceBuffers {
    uint4 *gbuf = nullptr;          // P x (checks, material, 0, vid)
    uint32_t *gpos = nullptr;       // P
    float4 *output = nullptr;       // P
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
    EmissiveVoxel *emissives = nullptr;  // EMISSIVE_CAPACITY
    uint32_t *emissive_claim = nullptr;  // EMISSIVE_CLAIM_W x EMISSIVE_CLAIM_PROBES
    Material *materials = nullptr;       // MATERIAL_SLOTS
    unsigned long long *node_stats = nullptr;  // 4: primary reads, trace reads, primary steps, trace steps
    // schedule ordering scratch
    uint32_t *first_rank = nullptr;      // per claim slot: first claimant's execution rank
    uint64_t *order_keys = nullptr;      // CANDIDATE_CAPACITY sort keys
    uint64_t *order_keys_alt = nullptr;
    uint32_t *order_vals = nullptr;
    uint32_t *order_vals_alt = nullptr;
    void *sort_temp = nullptr;
    size_t sort_temp_bytes = 0;
};

void launch_primary(hipStream_t s, const uint32_t *words, const FrameParams &p, DeviceBuffers &b, bool stats);

}  // namespace ve
