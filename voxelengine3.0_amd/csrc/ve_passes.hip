// gfx950 kernels for the wavefront passes after the primary pass:
//   normal.comp     -> k_normal
//   schedule.comp   -> k_sched_claim, k_sched_rank(_over), k_sched_vec_count, k_sched_write
//   rank.comp       -> k_rank, k_rank_write
//   threshold.comp  -> k_threshold
//   emit.comp       -> k_emit
//   trace.comp      -> k_trace
//   shade.comp      -> k_shade
//   resolve.comp    -> k_resolve
//
// Ordering.  The reference appends candidates and shade-list entries with
// atomic counters, so their order is whatever order the driver runs the
// invocations in.  These kernels reproduce one fixed order — that of a
// single-threaded llvmpipe run (work groups in dispatch order, invocations in
// SIMD-8 vectors, lockstep loop iterations) — without serialising anything:
// every pixel gets its execution rank e, a voxel's representative pixel is its
// lowest-ranked pixel, a voxel's claim-map probe index is its rank among the
// voxels of the same claim row, and positions in the candidate and shade lists
// come from exclusive scans over SIMD-8 vectors.  The result is deterministic
// and equal to the reference's output on the same inputs.
#include <hip/hip_runtime.h>

#include "ve_device.h"
#include "ve_launch.h"

namespace ve {

thread_local PassTimer pass_timer;


constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr int TILE = 8;
constexpr int NUM_XCD = 8;

// ---------------------------------------------------------------------------
// Launches over many views
// ---------------------------------------------------------------------------
// The frame kernels take a MULTI template parameter.  MULTI = false: one view,
// arguments as given, blocks = blockIdx.x of gridDim.x.  MULTI = true: one
// launch covers the same pass of many views (ViewRec, MultiTab); its grid is
// the concatenation of each view's own grid, and every block finds its view,
// takes that view's arguments from its record and runs as block x of the
// view's n.

__device__ __forceinline__ FrameParams global_params(const FrameParams &v)
{
    FrameParams p = v;
    p.sky_tex = gptr(v.sky_tex);
    p.jump = gptr(v.jump);
    p.dyn = gptr(v.dyn);
    return p;
}

// The block's view: the last v with pre[v] <= blockIdx.x (views with an empty
// grid share their start with the next view).  Called by whole waves before
// any branch; every wave of a block computes the same answer.
__device__ __forceinline__ const ViewRec &multi_view(const MultiTab &mt, uint32_t &x, uint32_t &n)
{
    const uint32_t b = blockIdx.x, lane = threadIdx.x & 63u;
    uint32_t cnt = 0;
    for (uint32_t base = 0; base < mt.n; base += 64u) {
        const uint32_t i = base + lane;
        const uint64_t m = __ballot(i < mt.n && mt.pre[i] <= b);
        cnt += (uint32_t)__popcll(m);
        if (m != ~0ull) break;
    }
    const uint32_t v = (uint32_t)__builtin_amdgcn_readfirstlane(cnt - 1u);
    const uint32_t first = mt.pre[v];
    x = b - first;
    n = mt.pre[v + 1u] - first;
    return mt.views[v];
}

// Bounce iteration `iter` of the trace/shade loop (renderer.cpp:515-602); the
// ping-pong direction alternates with `iter`.
__host__ __device__ inline Ray *bounce_in(const DeviceBuffers &b, uint32_t iter) { return (iter & 1u) ? b.rays_next : b.rays; }
__host__ __device__ inline Ray *bounce_out(const DeviceBuffers &b, uint32_t iter) { return (iter & 1u) ? b.rays : b.rays_next; }
__host__ __device__ inline uint32_t *live_in_ptr(const DeviceBuffers &b, uint32_t iter)
{
    return b.counters + ((iter & 1u) ? VE_CNT_LIVE_B : VE_CNT_LIVE);
}
__host__ __device__ inline uint32_t *live_out_ptr(const DeviceBuffers &b, uint32_t iter)
{
    return b.counters + ((iter & 1u) ? VE_CNT_LIVE : VE_CNT_LIVE_B);
}
__host__ __device__ inline uint32_t *li_in_ptr(const DeviceBuffers &b, uint32_t iter) { return b.ray_li[iter & 1u]; }
__host__ __device__ inline uint32_t *li_out_ptr(const DeviceBuffers &b, uint32_t iter) { return b.ray_li[(iter & 1u) ^ 1u]; }
// SIMD-8 vectors of the schedule pass (one 8x8 work group = 8 vectors)
__host__ __device__ inline uint32_t sched_vectors(const FrameParams &p) { return ((p.W + 7u) / 8u) * ((p.H + 7u) / 8u) * 8u; }

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool tile_pixel(uint32_t vbx, uint32_t vbn, const FrameParams &p, uint32_t &x,
                                           uint32_t &y)
{
    const uint32_t tiles_x = (p.W + TILE - 1) / TILE;
    const uint32_t tiles_y = (p.H + TILE - 1) / TILE;
    const uint32_t tile = xcd_group(vbx, vbn, p.tile_chunk) * 4u + (threadIdx.x >> 6);
    if (tile >= tiles_x * tiles_y) return false;
    const uint32_t lane = threadIdx.x & 63u;
    x = (tile % tiles_x) * TILE + (lane & 7u);
    y = (tile / tiles_x) * TILE + (lane >> 3);
    return x < p.W && y < p.H;
}

// Per-frame clear of the three ordering tables in one pass.
__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t v, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// First slot s in [from, n) of a bucket's owner words with owner == vid, or
// (with STOP_AT_EMPTY) owner == 0; n when there is none.  The whole owner row
// (n <= 32 words) is read at once, its loads in flight together, and scanned
// in registers.  `owner` receives the owner word found.
template <bool STOP_AT_EMPTY>
__device__ __forceinline__ uint32_t find_owner(const uint32_t *__restrict__ owners, uint32_t from, uint32_t n,
                                               uint32_t vid, uint32_t &owner)
{
    uint32_t ov[32];
#pragma unroll
    for (uint32_t base = 0; base < 32; base += 4) {
        uint4 o = make_uint4(0u, 0u, 0u, 0u);
        if (base < n) o = *reinterpret_cast<const uint4 *>(owners + base);
        ov[base] = o.x;
        ov[base + 1] = o.y;
        ov[base + 2] = o.z;
        ov[base + 3] = o.w;
    }
    uint32_t is_vid = 0u, is_free = 0u;
#pragma unroll
    for (uint32_t s = 0; s < 32; s++) {
        is_vid |= (uint32_t)(ov[s] == vid) << s;
        is_free |= (uint32_t)(STOP_AT_EMPTY && ov[s] == 0u) << s;
    }
    const uint32_t in_range = (from >= 32u ? 0u : 0xFFFFFFFFu << from) & (n >= 32u ? 0xFFFFFFFFu : (1u << n) - 1u);
    const uint32_t hit = (is_vid | is_free) & in_range;
    if (hit == 0u) {
        owner = 0xFFFFFFFFu;
        return n;
    }
    const uint32_t s = (uint32_t)__ffs(hit) - 1u;
    owner = (is_vid >> s) & 1u ? vid : 0u;
    return s;
}

// The reference's claim loop `for s: prev = CAS(owner[s], 0, vid); if (prev
// == 0 || prev == vid) break;` — a CAS on a slot owned by another id changes
// nothing, so scanning with plain loads and issuing the CAS only at the first
// free slot gives the same result; a lost race continues the scan.
// Returns the slot (n when the bucket is full) and whether this call claimed it.
__device__ __forceinline__ uint32_t claim_slot(uint32_t *__restrict__ owners, uint32_t n, uint32_t vid, bool &claimed)
{
    claimed = false;
    uint32_t from = 0;
    while (from < n) {
        uint32_t own;
        const uint32_t s = find_owner<true>(owners, from, n, vid, own);
        if (s >= n) return n;
        if (own == vid) return s;
        const uint32_t prev = atomicCAS(&owners[s], 0u, vid);
        if (prev == 0u) {
            claimed = true;
            return s;
        }
        if (prev == vid) return s;
        from = s + 1;
    }
    return n;
}

// ---------------------------------------------------------------------------
// normal.comp
// ---------------------------------------------------------------------------
// Occupancy of the lattice points around a centre (hx, hy, hz) as the
// reference's point lookup reports it (normal.comp:55-70: descend from the
// root; a non-internal node, or the leaf slot at the bottom, answers by its
// material), with the centre's descent shared: depth d of a
// descent chooses its child by coordinate bit D-d, so a neighbour whose
// coordinates first differ from the centre's in bit b reads the same nodes as
// the centre down to depth D-b-1.  The centre's descent records the group
// offset of every depth (LDS column `cache`, entry d at cache[d * stride]);
// a neighbour resumes at depth D-b from there, or takes the centre's outcome
// when the centre's descent ended above that depth.
struct SharedDescent {
    uint32_t hx, hy, hz;
    uint32_t stop;   // depth at which the centre's descent ended
    bool occ;        // its outcome
    uint32_t from;   // first depth whose block base the cache holds (jl + 1 after a jump-table start)
};

// The jump table's outcome through depth jl for (x, y, z): true when the
// descent ended there (occ set), else `offset` = the block base of depth jl + 1.
__device__ __forceinline__ bool jump_start(const uint32_t *__restrict__ jump, uint32_t jl, uint32_t depth, uint32_t x,
                                           uint32_t y, uint32_t z, uint32_t &offset, uint32_t &stop, bool &occ)
{
    const uint32_t sh = depth - jl;
    const uint32_t e = jump[((x >> sh) << (2u * jl)) | ((y >> sh) << jl) | (z >> sh)];
    if (e & JUMP_ENDED) {
        stop = e & 0xFFu;
        occ = (e & JUMP_OCC) != 0u;
        return true;
    }
    offset = e;
    return false;
}

// The centre voxel's descent (normal.comp's lookupOccupied): from the jump
// table through depth jl, then word by word, keeping each depth's block base
// for the neighbours' lookups.
__device__ __forceinline__ SharedDescent centre_descent(const uint32_t *__restrict__ w, const uint32_t *__restrict__ jump,
                                                        uint32_t jl, uint32_t n, uint32_t depth, uint32_t hx,
                                                        uint32_t hy, uint32_t hz, uint32_t len, uint32_t *cache,
                                                        uint32_t stride)
{
    SharedDescent c{hx, hy, hz, depth, false, jl + 1u};
    uint32_t offset = 0;
    if (jump_start(jump, jl, depth, hx, hy, hz, offset, c.stop, c.occ)) return c;
    for (uint32_t d = jl + 1u; d <= depth; d++) {
        cache[d * stride] = offset;
        offset += locate_bit(hx, hy, hz, depth - d);
        const uint32_t raw = node_word(w, n, offset);
        if (d == depth || !(raw & 1u)) {
            c.stop = d;
            c.occ = ((raw >> 1) & 0x7Fu) != 0u;
            break;
        }
        offset = (raw >> 4) & 0x0FFFFFFFu;
    }
    return c;
}

__device__ __forceinline__ bool neighbour_occupied(const uint32_t *__restrict__ w, const uint32_t *__restrict__ jump,
                                                   uint32_t jl, uint32_t n, uint32_t depth, const SharedDescent &c,
                                                   uint32_t x, uint32_t y, uint32_t z, uint32_t len,
                                                   const uint32_t *cache, uint32_t stride)
{
    const uint32_t diff = (x ^ c.hx) | (y ^ c.hy) | (z ^ c.hz);
    uint32_t d0 = depth - (31u - (uint32_t)__clz(diff));   // first depth choosing by a differing bit
    if (c.stop < d0) return c.occ;
    uint32_t offset;
    if (d0 < c.from) {   // the paths part above the cached depths: this voxel's own table entry
        uint32_t stop;
        bool occ;
        if (jump_start(jump, jl, depth, x, y, z, offset, stop, occ)) return occ;
        d0 = jl + 1u;
    } else {
        offset = cache[d0 * stride];
    }
    for (uint32_t d = d0; d <= depth; d++) {
        offset += locate_bit(x, y, z, depth - d);
        const uint32_t raw = node_word(w, n, offset);
        if (d == depth || !(raw & 1u)) return ((raw >> 1) & 0x7Fu) != 0u;
        offset = (raw >> 4) & 0x0FFFFFFFu;
    }
    return false;
}

__device__ f3 refined_normal(const uint32_t *__restrict__ w, const uint32_t *__restrict__ jump, uint32_t jl, uint32_t n,
                             uint32_t depth, uint32_t hx, uint32_t hy, uint32_t hz, uint32_t len, int r,
                             uint32_t *cache, uint32_t stride)   // normal.comp:73-87
{
    const SharedDescent c = centre_descent(w, jump, jl, n, depth, hx, hy, hz, len, cache, stride);
    f3 acc{0.0f, 0.0f, 0.0f};
    for (int dx = -r; dx <= r; dx++)
        for (int dy = -r; dy <= r; dy++)
            for (int dz = -r; dz <= r; dz++) {
                if (dx == 0 && dy == 0 && dz == 0) continue;
                if (dx * dx + dy * dy + dz * dz > r * r) continue;
                const int nx = (int)hx + dx, ny = (int)hy + dy, nz = (int)hz + dz;
                const bool oob = nx < 0 || ny < 0 || nz < 0 || nx >= (int)len || ny >= (int)len || nz >= (int)len;
                const bool empty =
                    oob || !neighbour_occupied(w, jump, jl, n, depth, c, (uint32_t)nx, (uint32_t)ny, (uint32_t)nz, len,
                                               cache, stride);
                if (empty) acc = add(acc, f3{(float)dx, (float)dy, (float)dz});
            }
    return length3(acc) > 0.001f ? normalize(acc) : f3{0.0f, 1.0f, 0.0f};
}

__device__ __forceinline__ uint32_t pack_normal(f3 v)   // normal.comp:90-95, round = half-to-even
{
    int x = (int)rintf((v.x + 1.0f) * 511.0f), y = (int)rintf((v.y + 1.0f) * 511.0f), z = (int)rintf((v.z + 1.0f) * 511.0f);
    x = min(max(x, 0), 1022);
    y = min(max(y, 0), 1022);
    z = min(max(z, 0), 1022);
    return ((uint32_t)x << 22) | ((uint32_t)y << 12) | ((uint32_t)z << 2);
}

// ---------------------------------------------------------------------------
// Ordered per-bucket replay
// ---------------------------------------------------------------------------
// The nBuffer and lBuffer are 32-slot hash buckets whose claims and
// least-recently-used evictions the reference makes with atomics from many
// invocations at once.  When a bucket fills up, which slot a voxel takes or
// loses depends on the order of those atomics.  Its order on the GL driver
// the parity is defined against (single-threaded llvmpipe: work groups in
// dispatch order, the invocations of a work group in SIMD-8 vectors, each
// vector's lanes in lockstep) is replayed exactly: the work for each bucket is
// listed as events with their execution order, every bucket's events are
// sorted by it, and one wave per bucket replays them against the bucket,
// lanes holding its slots.

// Ascending sort of n 64-bit keys by one wave: the bitonic network of the
// next power of two in the form whose every compare-exchange puts the
// smaller key at the lower index (the first step of each merge pairs i with
// its mirror in the block), so the positions past n act as +infinity and
// every exchange that touches them is skipped.  GLOBAL: the keys are in
// device memory (loads and stores bypass the non-coherent L1 between steps).
template <bool GLOBAL>
__device__ __forceinline__ uint64_t key_ld(const uint64_t *k, uint32_t i)
{
    if (GLOBAL) return __hip_atomic_load(k + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return k[i];
}
template <bool GLOBAL>
__device__ __forceinline__ void key_st(uint64_t *k, uint32_t i, uint64_t v)
{
    if (GLOBAL) __hip_atomic_store(k + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else k[i] = v;
}
// Orders a wave's LDS accesses before its next ones (the sorts and replays
// are per wave, several waves to a block, so no block barrier).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <bool GLOBAL>
__device__ void wave_sort_keys(uint64_t *keys, uint32_t n)
{
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t np = 1;
    while (np < n) np <<= 1;
    for (uint32_t k = 2; k <= np; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = lane; i < np / 2u; i += 64u) {
                const uint32_t a = (i / j) * 2u * j + (i % j);
                const uint32_t b = j == (k >> 1) ? a ^ (k - 1u) : a + j;   // mirror, then half-cleaners
                if (b < n) {
                    const uint64_t x = key_ld<GLOBAL>(keys, a), y = key_ld<GLOBAL>(keys, b);
                    if (y < x) {
                        key_st<GLOBAL>(keys, a, y);
                        key_st<GLOBAL>(keys, b, x);
                    }
                }
            }
            if (GLOBAL) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
            wave_lds_sync();
        }
    }
}

// The same network over one key per lane (n <= 64; absent keys ~0): lane i
// receives the i-th smallest.
__device__ __forceinline__ uint64_t wave_sort64(uint64_t key)
{
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (uint32_t k = 2; k <= 64; k <<= 1)
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            const uint32_t partner = j == (k >> 1) ? lane ^ (k - 1u) : lane ^ j;
            const uint64_t other = __shfl((unsigned long long)key, (int)partner, 64);
            key = (lane < partner) ? (other < key ? other : key) : (other < key ? key : other);
        }
    return key;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}

// Execution rank of pixel (x, y) in an 8x8-group dispatch: work groups in
// dispatch order (x fastest), invocations in local order.
__device__ __forceinline__ uint32_t exec_rank(uint32_t x, uint32_t y, uint32_t W)
{
    const uint32_t gw = (W + 7u) / 8u;
    return (((y >> 3) * gw + (x >> 3)) << 6) | ((y & 7u) << 3) | (x & 7u);
}

// A bucket's events in sorted order: buckets of up to 64 events are sorted
// in registers (lane i holds the i-th), up to ORDER_LDS_EVENTS in LDS, larger
// ones (a voxel covering many work groups of a close view) in device memory.
static_assert(DEP_CAP <= ORDER_LDS_EVENTS, "in-place deposit lists sort in LDS");

struct SortedKeys {
    uint32_t n;
    uint64_t reg;        // keys == nullptr (n <= 64): this lane's key
    uint64_t *keys;      // else: all n keys, sorted (LDS, or device memory when glob)
    bool glob;
    __device__ __forceinline__ uint64_t at(uint32_t q) const   // q = window base + lane
    {
        if (keys == nullptr) return reg;
        if (!glob) return keys[q];
        return __hip_atomic_load(keys + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

// Sort keys (high word from key_hi(i), low word i) of the bucket's n events.
template <typename KeyHi>
__device__ __forceinline__ SortedKeys sort_bucket(uint32_t n, KeyHi key_hi, uint64_t *lds_keys, uint64_t *spill)
{
    const uint32_t lane = threadIdx.x & 63u;
    SortedKeys sk{n, ~0ull, nullptr, false};
    if (n <= 64u) {
        sk.reg = wave_sort64(lane < n ? ((uint64_t)key_hi(lane) << 32) | lane : ~0ull);
        return sk;
    }
    const bool in_lds = n <= ORDER_LDS_EVENTS;
    sk.keys = in_lds ? lds_keys : spill;
    sk.glob = !in_lds;
    for (uint32_t i = lane; i < n; i += 64u) {
        const uint64_t k = ((uint64_t)key_hi(i) << 32) | i;
        if (in_lds) sk.keys[i] = k;
        else __hip_atomic_store(sk.keys + i, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (!in_lds) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
    wave_lds_sync();
    if (in_lds) wave_sort_keys<false>(sk.keys, n);
    else wave_sort_keys<true>(sk.keys, n);
    return sk;
}

// ---------------------------------------------------------------------------
// normal.comp
// ---------------------------------------------------------------------------
// Replay of normal.comp:124-177 per nBuffer bucket.  Per pixel: claim the
// first slot that is free or holds the voxel (computing the normal if the
// slot has none) and stamp the time; when every slot holds another voxel,
// take the least recently stamped one.  Within one SIMD-8 vector the claim
// loops of the lanes give the same slots as running the lanes one after
// another, but the lanes that find no slot all scan the time stamps before
// any of them writes: they pick the same slot, and the last of them owns it.
//
// 1. k_nrm_events: one wave per 8x8 work group lists its hit pixels per
//    bucket as events {execution rank, -, voxel id, packed position}.  A pixel
//    whose closest earlier pixel of the group in the same bucket carries the
//    same voxel (and position) is left out: it finds the slot that pixel just
//    claimed, or evicts with it, to the same effect.  One atomic per distinct
//    bucket, all of a wave's in one instruction; each bucket has room for
//    `cap` events, more go to an overflow list {rank, bucket, id, position}.
// 2. k_nrm_replay: one wave per bucket (four per block) sorts its events and
//    replays them, lanes holding the slots, events streamed through
//    registers; each slot's new normal (if any) is noted in its work word.
// 3. k_nrm_compute: one lane per slot with a new normal computes it.
struct OrderLists {
    uint32_t cap;          // events per bucket in place
    uint32_t over_cap;     // overflow list capacity
    uint32_t *cnt;         // per bucket: events (zero between passes)
    uint4 *ev;             // bucket-major, cap each
    uint4 *over;           // overflow events
    uint32_t *over_n;      // [0] overflow entries, [1] spill words handed out (zero between passes)
    uint64_t *spill;       // sort keys of buckets whose events overflowed
    uint32_t spill_cap;
};

__host__ __device__ inline OrderLists nrm_lists(const DeviceBuffers &b)
{
    return OrderLists{b.ord_cap, b.ord_over_cap, b.ord_cnt, b.ord_ev, b.ord_over, b.ord_over_n, b.ord_spill,
                      b.ord_spill_cap};
}
__device__ __forceinline__ OrderLists nrm_lists_global(const DeviceBuffers &b)
{
    return OrderLists{b.ord_cap, b.ord_over_cap, gptr(b.ord_cnt), gptr(b.ord_ev), gptr(b.ord_over),
                      gptr(b.ord_over_n), gptr(b.ord_spill), b.ord_spill_cap};
}

template <bool MULTI>
__global__ __launch_bounds__(256) void k_nrm_events(FrameParams p, const uint4 *__restrict__ gbuf,
                                                    const uint32_t *__restrict__ gpos, OrderLists ol,
                                                    uint32_t *__restrict__ overflow,
                                                    MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        p = global_params(V.p);
        gbuf = gptr(V.b.gbuf);
        gpos = gptr(V.b.gpos);
        ol = nrm_lists_global(V.b);
        overflow = gptr(V.b.counters) + VE_CNT_OVERFLOW;
    }
    uint32_t x = 0, y = 0;
    const bool valid = tile_pixel(vbx, vbn, p, x, y);
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (1ull << lane) - 1ull;
    const size_t pix = valid ? (size_t)y * p.W + x : 0;
    const uint32_t vid = valid ? gbuf[pix].w : 0u;
    const bool hit = vid != 0u;
    const uint32_t pk = hit ? gpos[pix] : 0u;
    const uint32_t h = hit ? vid % (p.hash_width - 1u) + 1u : 0u;
    uint64_t pending = __ballot(hit);
    bool keep = false, leads = false;
    uint32_t rank = 0, group = 0, leader_of = lane;
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t hb = (uint32_t)__builtin_amdgcn_readlane((int)h, leader);
        const bool mine = hit && h == hb;
        const uint64_t same = __ballot(mine);
        pending &= ~same;
        const uint64_t before = same & below;
        const int j = before ? 63 - __clzll((long long)before) : (int)lane;
        const uint32_t pv = __shfl(vid, j, 64), pp = __shfl(pk, j, 64);
        const bool k = mine && !(before && pv == vid && pp == pk);
        const uint64_t kept = __ballot(k);
        if (mine) {
            keep = k;
            rank = (uint32_t)__popcll(kept & below);
            leader_of = (uint32_t)leader;
        }
        if ((int)lane == leader) {
            leads = true;
            group = (uint32_t)__popcll(kept);
        }
    }
    uint32_t b = 0;
    if (leads) b = atomicAdd(&ol.cnt[h], group);
    b = __shfl(b, (int)leader_of, 64);
    if (!keep) return;
    const uint32_t at = b + rank, e = exec_rank(x, y, p.W);
    if (at < ol.cap) {
        ol.ev[(size_t)h * ol.cap + at] = make_uint4(e, 0u, vid, pk);
    } else {
        const uint32_t o = atomicAdd(&ol.over_n[0], 1u);
        if (o < ol.over_cap) ol.over[o] = make_uint4(e, h, vid, pk);
        else atomicOr(overflow, 2u);   // past the list: no longer the reference's order
    }
}

// The sorted events of bucket h: its in-place ones plus, when it has more
// than `cap`, its entries of the overflow list (found by scanning it, keys in
// spill words handed out for the bucket).  A key's low word indexes the
// bucket's row, or with bit 31 set the overflow list.
__device__ __forceinline__ SortedKeys sort_listed(const OrderLists &ol, uint32_t h, uint32_t n_all, uint64_t *lds_keys)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint4 *row = ol.ev + (size_t)h * ol.cap;
    if (n_all <= ol.cap) return sort_bucket(n_all, [&](uint32_t i) { return row[i].x; }, lds_keys, nullptr);
    uint32_t base = 0;
    if (lane == 0u) base = atomicAdd(&ol.over_n[1], n_all);
    base = __shfl(base, 0, 64);
    if (base + n_all > ol.spill_cap) return SortedKeys{0u, ~0ull, nullptr, false};   // reported by the caller
    SortedKeys sk{n_all, ~0ull, ol.spill + base, true};
    for (uint32_t i = lane; i < ol.cap; i += 64u)
        __hip_atomic_store(sk.keys + i, ((uint64_t)row[i].x << 32) | i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t L = min(ol.over_n[0], ol.over_cap);
    uint32_t put = ol.cap;
    for (uint32_t j0 = 0; j0 < L; j0 += 64u) {
        const uint32_t j = j0 + lane;
        const bool m = j < L && ol.over[j].y == h;
        const uint64_t mb = __ballot(m);
        const uint32_t at = put + (uint32_t)__popcll(mb & ((1ull << lane) - 1ull));
        if (m && at < n_all)
            __hip_atomic_store(sk.keys + at, ((uint64_t)ol.over[j].x << 32) | 0x80000000u | j, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        put += (uint32_t)__popcll(mb);
    }
    sk.n = min(put, n_all);   // fewer when the overflow list lost events (reported by the caller)
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
    wave_sort_keys<true>(sk.keys, sk.n);
    return sk;
}

__device__ __forceinline__ uint4 listed_event(const OrderLists &ol, uint32_t h, uint64_t key)
{
    const uint32_t lo = (uint32_t)key;
    return (lo & 0x80000000u) ? ol.over[lo & 0x7FFFFFFFu] : ol.ev[(size_t)h * ol.cap + lo];
}

constexpr uint32_t REPLAY_WAVES = 4;   // buckets per block of the replays (one per wave)

// One slot's words of an nBuffer bucket, from another set.
__device__ __forceinline__ void nrm_copy_slot(uint32_t *bucket, const uint32_t *from, uint32_t sl)
{
    bucket[sl] = from[sl];
    bucket[NB_TIME + sl] = from[NB_TIME + sl];
    bucket[NB_NRM + sl] = from[NB_NRM + sl];
}

template <bool MULTI>
__global__ __launch_bounds__(64 * REPLAY_WAVES) void k_nrm_replay(FrameParams p, uint32_t *nbuf, const uint32_t *prev,
                                                                  OrderLists ol, uint32_t *__restrict__ work,
                                                                  uint32_t *__restrict__ overflow,
                                                                  MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        p = global_params(V.p);
        nbuf = gptr(V.b.nbuf);
        prev = nbuf;
        ol = nrm_lists_global(V.b);
        work = gptr(V.b.ord_work);
        overflow = gptr(V.b.counters) + VE_CNT_OVERFLOW;
    }
    __shared__ uint64_t lds_keys_all[REPLAY_WAVES][ORDER_LDS_EVENTS];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t h = vbx * REPLAY_WAVES + wv;
    if (h >= p.hash_width) return;
    const bool slot_lane = lane < NBUFFER_SLOTS;
    const uint32_t n = ol.cnt[h];
    if (slot_lane) work[h * NBUFFER_SLOTS + lane] = NONE;   // no new normal unless the replay gives one
    // a pass into another nBuffer set (prev != nbuf) starts from the previous
    // set's bucket and writes all of it
    uint32_t *bucket = nbuf + (size_t)h * NB_WORDS;
    const uint32_t *from = prev + (size_t)h * NB_WORDS;
    const bool copy = prev != nbuf;
    if (n == 0u) {
        if (copy && slot_lane) nrm_copy_slot(bucket, from, lane);
        return;
    }
    if (lane == 0u) ol.cnt[h] = 0u;   // the next pass's count starts from zero
    const SortedKeys sk = sort_listed(ol, h, n, lds_keys_all[wv]);
    if (sk.n < n) {   // events lost past the lists: no longer the reference's order
        if (lane == 0u) atomicOr(overflow, 2u);
        if (sk.n == 0u) {
            if (copy && slot_lane) nrm_copy_slot(bucket, from, lane);
            return;
        }
    }
    uint32_t owner = slot_lane ? from[lane] : NONE;
    uint32_t tstamp = slot_lane ? from[NB_TIME + lane] : NONE;
    const uint32_t nrm = slot_lane ? from[NB_NRM + lane] : 0u;
    if (copy && slot_lane) bucket[NB_NRM + lane] = nrm;
    bool has_src = false;   // the slot takes the normal at packed position `src`
    uint32_t src = 0u;
    uint32_t cur_vec = NONE, ev_vid = 0u, ev_pos = 0u, last_vid = 0u, last_pos = 0u;
    auto evict = [&]() {   // the lockstep eviction of one vector's slotless lanes
        const uint32_t best = wave_min_u32(slot_lane ? tstamp : NONE);
        const uint64_t at = __ballot(slot_lane && tstamp == best);
        if (lane == (uint32_t)__ffsll((unsigned long long)at) - 1u) {
            owner = ev_vid;
            has_src = true;
            src = ev_pos;
            tstamp = fp_time(p);
        }
        ev_vid = 0u;
    };
    for (uint32_t w0 = 0; w0 < sk.n; w0 += 64u) {
        const uint32_t q = w0 + lane, cnt_w = min(64u, sk.n - w0);
        uint4 e = make_uint4(0u, 0u, 0u, 0u);   // this lane's event of the window, in sorted order
        if (q < sk.n) e = listed_event(ol, h, sk.at(q));
        for (uint32_t j = 0; j < cnt_w; j++) {
            const uint32_t vec = (uint32_t)__builtin_amdgcn_readlane((int)e.x, (int)j) >> 3;
            const uint32_t vid = (uint32_t)__builtin_amdgcn_readlane((int)e.z, (int)j);
            const uint32_t pos = (uint32_t)__builtin_amdgcn_readlane((int)e.w, (int)j);
            if (vec != cur_vec) {
                if (ev_vid != 0u) evict();
                cur_vec = vec;
            }
            if (vid == last_vid && pos == last_pos) continue;   // finds what the last event left: no change
            last_vid = vid;
            last_pos = pos;
            const uint64_t m = __ballot(slot_lane && (owner == vid || owner == 0u));
            if (m) {
                if (lane == (uint32_t)__ffsll((unsigned long long)m) - 1u) {
                    if (owner == 0u) owner = vid;
                    if (nrm == 0u && !has_src) {   // normal.comp:151-156: computed only into an empty field
                        has_src = true;
                        src = pos;
                    }
                    tstamp = fp_time(p);
                }
            } else {   // lockstep: the vector's last slotless lane wins the evicted slot
                ev_vid = vid;
                ev_pos = pos;
            }
        }
    }
    if (ev_vid != 0u) evict();
    if (slot_lane) {
        bucket[lane] = owner;
        bucket[NB_TIME + lane] = tstamp;
        if (has_src) work[h * NBUFFER_SLOTS + lane] = src;
    }
}

// Work word NONE: the slot keeps its normal.  (A packed position of NONE
// would need every coordinate bit set past depth 11, where the reference's
// packing has already overflowed; such a slot keeps its old normal.)
template <bool MULTI>
__global__ __launch_bounds__(256) void k_nrm_compute(const uint32_t *__restrict__ words, FrameParams p,
                                                     uint32_t *__restrict__ nbuf, const uint32_t *__restrict__ work,
                                                     uint32_t *__restrict__ over_n,
                                                     MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        words = gptr(V.words);
        p = global_params(V.p);
        nbuf = gptr(V.b.nbuf);
        work = gptr(V.b.ord_work);
        over_n = gptr(V.b.ord_over_n);
    }
    __shared__ uint32_t path_cache[VE_MAX_DEPTH * 256];
    const uint32_t t = vbx * blockDim.x + threadIdx.x;
    if (t == 0) over_n[0] = over_n[1] = 0u;   // the replay, the lists' last reader, is done
    if (t >= p.hash_width * NBUFFER_SLOTS) return;
    const uint32_t src = work[t];
    if (src == NONE) return;
    const uint32_t h = t / NBUFFER_SLOTS, sl = t % NBUFFER_SLOTS;
    const uint32_t len = 1u << p.depth, bits = p.depth - 1u, mask = (1u << bits) - 1u;
    const uint32_t hx = (src & mask) * 2u, hy = ((src >> bits) & mask) * 2u, hz = ((src >> (2u * bits)) & mask) * 2u;
    nbuf[(size_t)h * NB_WORDS + NB_NRM + sl] =
        pack_normal(refined_normal(words, p.jump, p.jump_level, p.n_words, p.depth, hx, hy, hz, len,
                                   p.normal_precision, path_cache + threadIdx.x, 256u));
}

// ---------------------------------------------------------------------------
// schedule.comp
// ---------------------------------------------------------------------------
// Work group g of the 8x8 dispatch handles the bit-reversed tile of
// schedule.comp:58-71; its invocation l has execution rank e = g*64 + l.
__device__ __forceinline__ uint32_t sched_tile(uint32_t g, uint32_t total, uint32_t seed)
{
    const uint32_t log2g = total <= 1u ? 1u : (32u - __clz(total - 1u));
    uint32_t sc = bit_reverse(g ^ seed) >> (32u - log2g);
    return sc >= total ? g : sc;
}

__device__ __forceinline__ uint32_t sched_pixel(uint32_t e, const FrameParams &p, uint32_t &pix)
{
    const uint32_t gw = (p.W + 7u) / 8u, gh = (p.H + 7u) / 8u;
    const uint32_t tile = sched_tile(e >> 6, gw * gh, fp_frame_seed(p));
    const uint32_t x = (tile % gw) * 8u + (e & 7u), y = (tile / gw) * 8u + ((e >> 3) & 7u);
    pix = y * p.W + x;
    return x < p.W && y < p.H;
}

// Home slot of voxel id v inside its claim row's region (a mixer unrelated to
// claim_hash, which chose the row).
__device__ __forceinline__ uint32_t sched_home(uint32_t v)
{
    v *= 0x9E3779B1u;
    v ^= v >> 15;
    v *= 0x85EBCA77u;
    v ^= v >> 13;
    return v & (SCHED_SLOTS - 1u);
}

// One wave per work group g. Lanes with equal voxel ids are merged first (the
// lowest lane of an id carries its lane count), then every distinct id of the
// wave enters its claim row's region at once, by linear probing from its home
// slot. A slot is {id, ~lowest execution rank, pixels, 0}: the complement
// makes a zeroed slot "free" and lets one 64-bit CAS both claim the slot and
// keep the lowest rank; the pixel count is added on the same cache line.
template <bool MULTI>
__global__ __launch_bounds__(64) void k_sched_claim(FrameParams p, const uint4 *__restrict__ gbuf, uint4 *__restrict__ tab,
                                                    uint32_t *__restrict__ cand_at, uint32_t *__restrict__ overflow,
                                                    uint4 *__restrict__ over, uint32_t *__restrict__ over_n,
                                                    uint32_t *__restrict__ row_over, uint32_t *__restrict__ rows,
                                                    MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        p = global_params(V.p);
        gbuf = gptr(V.b.gbuf);
        tab = gptr(V.b.sched_tab);
        cand_at = gptr(V.b.cand_at);
        overflow = gptr(V.b.counters) + VE_CNT_OVERFLOW;
        over = gptr(V.b.sched_over);
        over_n = gptr(V.b.sched_over_n);
        row_over = gptr(V.b.sched_row_over);
        rows = gptr(V.b.sched_rows);
    }
    const uint32_t lane = threadIdx.x;
    const uint32_t e = vbx * 64u + lane;
    cand_at[e] = NONE;
    uint32_t pix;
    uint32_t vid = 0u;
    if (sched_pixel(e, p, pix)) vid = gbuf[pix].w;
    uint64_t pending = __ballot(vid != 0u);
    uint32_t lanes = 0u;
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lv = (uint32_t)__builtin_amdgcn_readlane((int)vid, leader);
        const uint64_t same = __ballot(vid == lv);
        pending &= ~same;
        if (lane == (uint32_t)leader) lanes = (uint32_t)__popcll(same);
    }
    if (lanes == 0u) return;
    const uint32_t h = claim_hash(vid) % CLAIM_WIDTH;
    uint4 *row = tab + (size_t)h * SCHED_SLOTS;
    const unsigned long long want = ((unsigned long long)~e << 32) | vid;
    uint32_t s = sched_home(vid);
    for (uint32_t probe = 0; probe < SCHED_SLOTS; probe++, s = (s + 1u) & (SCHED_SLOTS - 1u)) {
        unsigned long long *kf = reinterpret_cast<unsigned long long *>(row + s);
        // the home slot is claimed outright (usually free); past it, an id
        // never leaves its slot, so a plain read passes other ids' slots
        // without an atomic and only a free-looking slot is claimed
        unsigned long long cur = probe == 0u ? atomicCAS(kf, 0ull, want)
                                             : __hip_atomic_load(kf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur != 0ull && (uint32_t)cur != vid) continue;   // another id's slot
        if (cur == 0ull && probe != 0u) {
            cur = atomicCAS(kf, 0ull, want);
            if (cur != 0ull && (uint32_t)cur != vid) continue;   // another id took it first
        }
        // free (now ours) or this id's: keep the lower rank (higher complement)
        while (cur != 0ull && (uint32_t)(cur >> 32) < ~e) {
            const unsigned long long prev = atomicCAS(kf, cur, want);
            if (prev == cur) break;
            cur = prev;
        }
        atomicAdd(&row[s].z, lanes);
        return;
    }
    // the row's table is full: an overflow record, and the row is ranked by
    // k_sched_rank_over
    const uint32_t i = atomicAdd(&over_n[0], 1u);
    if (i < SCHED_OVER_CAP)
        over[i] = make_uint4(vid, ~e, lanes, h);
    else
        atomicOr(overflow, 1u);
    if (atomicExch(&row_over[h], 1u) == 0u) rows[atomicAdd(&over_n[1], 1u)] = h;
}

// One wave per claim row: a voxel's probe index in the reference's claim map is
// the number of voxels of its row claimed before it (lower execution rank).
// Writes the reference-layout claim and pixel-count maps (all 64 probes, so
// no separate clear is needed), marks each kept voxel's first pixel, and
// clears the row region for the next frame.
template <bool MULTI>
__global__ __launch_bounds__(64) void k_sched_rank(uint4 *__restrict__ tab, uint32_t *__restrict__ claim,
                                                   uint32_t *__restrict__ samples, uint32_t *__restrict__ cand_at,
                                                   const uint32_t *__restrict__ row_over,
                                                   MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        tab = gptr(V.b.sched_tab);
        claim = gptr(V.b.claim);
        samples = gptr(V.b.samples);
        cand_at = gptr(V.b.cand_at);
        row_over = gptr(V.b.sched_row_over);
    }
    constexpr uint32_t PER_LANE = SCHED_SLOTS / 64u;
    __shared__ uint32_t firsts[SCHED_SLOTS];
    const uint32_t h = vbx, lane = threadIdx.x;
    if (row_over[h] != 0u) return;   // k_sched_rank_over ranks it
    uint4 *row = tab + (size_t)h * SCHED_SLOTS;
    uint4 m[PER_LANE];
#pragma unroll
    for (uint32_t k = 0; k < PER_LANE; k++) m[k] = row[lane + 64u * k];
    // compact the occupied slots: member j of the row sits at firsts[j]
    uint32_t at[PER_LANE], n = 0;
#pragma unroll
    for (uint32_t k = 0; k < PER_LANE; k++) {
        const uint64_t occ = __ballot(m[k].x != 0u);
        at[k] = n + (uint32_t)__popcll(occ & ((1ull << lane) - 1ull));
        n += (uint32_t)__popcll(occ);
        if (m[k].x != 0u) {
            firsts[at[k]] = ~m[k].y;
            row[lane + 64u * k] = make_uint4(0u, 0u, 0u, 0u);
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < PER_LANE; k++) {
        if (m[k].x == 0u) continue;
        const uint32_t f = ~m[k].y;
        uint32_t rank = 0;
        for (uint32_t j = 0; j < n; j++) rank += firsts[j] < f;
        if (rank < CLAIM_PROBES) {
            claim[(size_t)h * CLAIM_PROBES + rank] = m[k].x;
            samples[(size_t)h * CLAIM_PROBES + rank] = m[k].z;
            cand_at[f] = (rank << 26) | (h << 6) | rank;
        }
    }
    // probes past the row's population read 0, as after the per-frame clear
    for (uint32_t s = n + lane; s < CLAIM_PROBES; s += 64u) {
        claim[(size_t)h * CLAIM_PROBES + s] = 0u;
        samples[(size_t)h * CLAIM_PROBES + s] = 0u;
    }
}

constexpr uint32_t SCHED_OVER_BLOCKS = 256;  // blocks striding over the flagged rows

// The rows whose table filled up: members are the table's voxels plus the
// row's overflow records, where one voxel may have several records (one per
// work group that saw it first in the row's probe order); a voxel's first
// execution rank is its records' lowest and its pixel count their sum. Then
// the same ranking as k_sched_rank. One block per flagged row; the table row
// and the flag are cleared for the next frame.
template <bool MULTI>
__global__ __launch_bounds__(256) void k_sched_rank_over(uint4 *__restrict__ tab, uint32_t *__restrict__ claim,
                                                         uint32_t *__restrict__ samples, uint32_t *__restrict__ cand_at,
                                                         const uint4 *__restrict__ over, const uint32_t *__restrict__ over_n,
                                                         uint32_t *__restrict__ row_over, const uint32_t *__restrict__ rows,
                                                         uint32_t *__restrict__ overflow,
                                                         MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        tab = gptr(V.b.sched_tab);
        claim = gptr(V.b.claim);
        samples = gptr(V.b.samples);
        cand_at = gptr(V.b.cand_at);
        over = gptr(V.b.sched_over);
        over_n = gptr(V.b.sched_over_n);
        row_over = gptr(V.b.sched_row_over);
        rows = gptr(V.b.sched_rows);
        overflow = gptr(V.b.counters) + VE_CNT_OVERFLOW;
    }
    __shared__ uint32_t tv[SCHED_SLOTS], tf[SCHED_SLOTS], tc[SCHED_SLOTS];
    __shared__ uint32_t rv[SCHED_ROW_OVER], rf[SCHED_ROW_OVER], rc[SCHED_ROW_OVER];
    __shared__ uint32_t nt, nr, nrep;
    const uint32_t t = threadIdx.x;
    const uint32_t no = min(over_n[0], SCHED_OVER_CAP), nrows = over_n[1];
    for (uint32_t r = vbx; r < nrows; r += vbn) {
        const uint32_t h = rows[r];
        uint4 *row = tab + (size_t)h * SCHED_SLOTS;
        if (t == 0u) {
            nt = 0u;
            nr = 0u;
            nrep = 0u;
        }
        __syncthreads();
        for (uint32_t k = t; k < SCHED_SLOTS; k += blockDim.x) {
            const uint4 m = row[k];
            if (m.x != 0u) {
                const uint32_t j = atomicAdd(&nt, 1u);
                tv[j] = m.x;
                tf[j] = ~m.y;
                tc[j] = m.z;
            }
            row[k] = make_uint4(0u, 0u, 0u, 0u);
        }
        for (uint32_t i = t; i < no; i += blockDim.x) {
            const uint4 o = over[i];
            if (o.w != h) continue;
            const uint32_t j = atomicAdd(&nr, 1u);
            if (j < SCHED_ROW_OVER) {
                rv[j] = o.x;
                rf[j] = ~o.y;
                rc[j] = o.z;
            }
        }
        __syncthreads();
        const uint32_t n_t = nt, n_r = min(nr, SCHED_ROW_OVER);
        if (t == 0u && nr > SCHED_ROW_OVER) atomicOr(overflow, 1u);
        // representative record of each voxel: its lowest first rank (ranks
        // are distinct, so exactly one record per voxel)
        for (uint32_t i = t; i < n_r; i += blockDim.x) {
            bool rep = true;
            for (uint32_t j = 0; j < n_r && rep; j++) rep = !(rv[j] == rv[i] && rf[j] < rf[i]);
            if (rep) atomicAdd(&nrep, 1u);
        }
        __syncthreads();
        const uint32_t n_all = n_t + nrep;
        for (uint32_t m = t; m < n_t + n_r; m += blockDim.x) {
            uint32_t vid, f, cnt;
            if (m < n_t) {
                vid = tv[m];
                f = tf[m];
                cnt = tc[m];
            } else {
                const uint32_t i = m - n_t;
                vid = rv[i];
                f = rf[i];
                cnt = 0u;
                bool rep = true;
                for (uint32_t j = 0; j < n_r; j++) {
                    if (rv[j] != vid) continue;
                    cnt += rc[j];
                    if (rf[j] < f) rep = false;
                }
                if (!rep) continue;
            }
            uint32_t rank = 0;
            for (uint32_t j = 0; j < n_t; j++) rank += tf[j] < f;
            for (uint32_t j = 0; j < n_r; j++) {
                if (rf[j] >= f) continue;
                bool rep = true;   // count each other voxel once, at its representative
                for (uint32_t q = 0; q < n_r && rep; q++) rep = !(rv[q] == rv[j] && rf[q] < rf[j]);
                rank += rep;
            }
            if (rank < CLAIM_PROBES) {
                claim[(size_t)h * CLAIM_PROBES + rank] = vid;
                samples[(size_t)h * CLAIM_PROBES + rank] = cnt;
                cand_at[f] = (rank << 26) | (h << 6) | rank;
            }
        }
        for (uint32_t q = n_all + t; q < CLAIM_PROBES; q += blockDim.x) {
            claim[(size_t)h * CLAIM_PROBES + q] = 0u;
            samples[(size_t)h * CLAIM_PROBES + q] = 0u;
        }
        if (t == 0u) row_over[h] = 0u;
        __syncthreads();
    }
}

// Sum over a 256-thread block (every thread gets it); `lds` holds 4 words.
__device__ __forceinline__ uint32_t block_sum256(uint32_t x, uint32_t *lds)
{
    for (int off = 32; off > 0; off >>= 1) x += (uint32_t)__shfl_xor((int)x, off, 64);
    if ((threadIdx.x & 63u) == 0u) lds[threadIdx.x >> 6] = x;
    __syncthreads();
    const uint32_t t = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    return t;
}

// Exclusive prefix over a 256-thread block, in thread order; `lds` holds 4 words.
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t x, uint32_t *lds)
{
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t inc = x;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)inc, off, 64);
        if (lane >= off) inc += y;
    }
    if (lane == 63u) lds[w] = inc;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t k = 0; k < w; k++) base += lds[k];
    __syncthreads();
    return base + inc - x;
}

// Candidates first-claimed in each SIMD-8 vector, and their sum per block of
// 256 vectors (k_sched_write turns these into the vectors' offsets).
template <bool MULTI>
__global__ __launch_bounds__(256) void k_sched_vec_count(const uint32_t *__restrict__ cand_at, uint32_t n_vec,
                                                         uint32_t *__restrict__ vec_cnt,
                                                         uint32_t *__restrict__ vec_chunk,
                                                         uint32_t *__restrict__ over_n,
                                                         MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        cand_at = gptr(V.b.cand_at);
        n_vec = sched_vectors(V.p);
        vec_cnt = gptr(V.b.vec_cnt);
        vec_chunk = gptr(V.b.vec_chunk);
        over_n = gptr(V.b.sched_over_n);
    }
    __shared__ uint32_t lds[4];
    const uint32_t v = vbx * blockDim.x + threadIdx.x;
    if (v == 0u) {   // the overflow records of this frame are consumed
        over_n[0] = 0u;
        over_n[1] = 0u;
    }
    uint32_t c = 0;
    if (v < n_vec) {
        for (uint32_t l = 0; l < 8; l++) c += cand_at[v * 8u + l] != NONE;
        vec_cnt[v] = c;
    }
    const uint32_t t = block_sum256(c, lds);
    if (threadIdx.x == 0) vec_chunk[vbx] = t;
}

// Inside a vector, wins happen probe by probe (lockstep loop) and, within a
// probe, in lane order: sort the vector's <= 8 winners by (probe, lane).
// A vector's first candidate index: the block's offset (the sum of the
// earlier blocks' totals) plus the prefix inside the block.
template <bool MULTI>
__global__ __launch_bounds__(256) void k_sched_write(FrameParams p, const uint32_t *__restrict__ cand_at, uint32_t n_vec,
                                                     const uint32_t *__restrict__ vec_cnt,
                                                     const uint32_t *__restrict__ vec_chunk,
                                                     const uint32_t *__restrict__ claim, uint4 *__restrict__ candidates,
                                                     uint32_t *__restrict__ counters,
                                                     MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        p = global_params(V.p);
        cand_at = gptr(V.b.cand_at);
        n_vec = sched_vectors(V.p);
        vec_cnt = gptr(V.b.vec_cnt);
        vec_chunk = gptr(V.b.vec_chunk);
        claim = gptr(V.b.claim);
        candidates = gptr(V.b.candidates);
        counters = gptr(V.b.counters);
    }
    __shared__ uint32_t lds[4];
    const uint32_t v = vbx * blockDim.x + threadIdx.x;
    uint32_t before = 0;
    for (uint32_t i = threadIdx.x; i < vbx; i += blockDim.x) before += vec_chunk[i];
    const uint32_t off = block_sum256(before, lds);
    const uint32_t cnt = v < n_vec ? vec_cnt[v] : 0u;
    const uint32_t vbase = off + block_excl_scan256(cnt, lds);
    if (v >= n_vec) return;
    if (v == n_vec - 1u) counters[VE_CNT_CANDIDATES] = min(vbase + cnt, CANDIDATE_CAPACITY);
    // the vector's claim-map words are loaded together up front (independent
    // loads in flight at once), then written in (probe, lane) order
    uint32_t at[8], word[8];
    for (uint32_t l = 0; l < 8; l++) at[l] = cand_at[v * 8u + l];
    for (uint32_t l = 0; l < 8; l++)
        word[l] = at[l] != NONE ? claim[(size_t)((at[l] >> 6) & 0xFFFFFu) * CLAIM_PROBES + (at[l] & 63u)] : 0u;
    uint32_t key[8];   // (probe, lane), sorted
    uint32_t m = 0;
    for (uint32_t l = 0; l < 8; l++) {
        if (at[l] == NONE) continue;
        const uint32_t k = ((at[l] >> 26) << 3) | l;
        uint32_t j = m++;
        while (j > 0 && key[j - 1] > k) {
            key[j] = key[j - 1];
            j--;
        }
        key[j] = k;
    }
    uint32_t pos = vbase;
    for (uint32_t j = 0; j < m; j++, pos++) {
        if (pos >= CANDIDATE_CAPACITY) break;
        const uint32_t l = key[j] & 7u, e = v * 8u + l;
        uint32_t a = 0u, w = 0u;
        for (uint32_t q = 0; q < 8; q++)   // select, not an indexed register array
            if (q == l) {
                a = at[q];
                w = word[q];
            }
        uint32_t pix;
        sched_pixel(e, p, pix);
        candidates[pos] = make_uint4(w, pix, (a >> 6) & 0xFFFFFu, a & 63u);
    }
}

// ---------------------------------------------------------------------------
// rank.comp
// ---------------------------------------------------------------------------
__device__ float indirect_variance(const uint32_t *__restrict__ lbuf, uint32_t hw, uint32_t vid)   // rank.comp:66-87
{
    const uint32_t h = vid % (hw - 1u) + 1u;
    const uint32_t *bucket = lbuf + (size_t)h * LB_WORDS;
    uint32_t own;
    const uint32_t s = find_owner<true>(bucket, 0, LBUFFER_SLOTS, vid, own);
    if (s >= LBUFFER_SLOTS || own == 0u) return 0.0f;   // not present yet
    const uint32_t *pl = bucket + LB_PAY + s * LB_PAYW;
    const uint32_t c = pl[5];
    if (c < 2u) return 0.0f;
    const float fc = (float)c;
    const float meanL = (float)(pl[6] + pl[7] + pl[8]) / (3.0f * fc);
    const float meanL2 = (float)pl[9] / fc;
    return gmax(meanL2 - meanL * meanL, 0.0f);
}

// Per candidate: samples (0/1/2/minSamples), histogram, min/max; per SIMD-8
// vector: the number of shade-list entries it appends.
// Candidates are strided over a fixed grid (the count is known only on the
// device), so each block merges many candidates into its histogram and
// min/max before the few global atomics of its flush.
template <bool MULTI>
__global__ __launch_bounds__(256) void k_rank(FrameParams p, const uint4 *__restrict__ candidates,
                                              const uint32_t *__restrict__ samples_map, const uint32_t *__restrict__ lbuf,
                                              const uint32_t *__restrict__ rank_stats,
                                              uint32_t *__restrict__ rank_part, const uint32_t *__restrict__ counters,
                                              int32_t *__restrict__ nsamp, uint32_t *__restrict__ vec_tot,
                                              uint32_t *__restrict__ chunk_tot,
                                              MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        p = global_params(V.p);
        candidates = gptr(V.b.candidates);
        samples_map = gptr(V.b.samples);
        lbuf = gptr(V.b.lbuf);
        rank_stats = gptr(V.b.rank_stats);
        rank_part = gptr(V.b.rank_part);
        counters = gptr(V.b.counters);
        nsamp = gptr(V.b.nsamp);
        vec_tot = gptr(V.b.vec_cnt);
        chunk_tot = gptr(V.b.rank_chunk);
    }
    __shared__ uint32_t hist[HISTOGRAM_BINS];
    __shared__ uint32_t sum_lds[4];
    __shared__ uint32_t bmin, bmax;
    for (uint32_t i = threadIdx.x; i < HISTOGRAM_BINS; i += blockDim.x) hist[i] = 0u;
    if (threadIdx.x == 0) {
        bmin = 0xFFFFFFFFu;
        bmax = 0u;
    }
    __syncthreads();
    uint32_t my_min = 0xFFFFFFFFu, my_max = 0u, active = 0u;
    const uint32_t n = counters[VE_CNT_CANDIDATES];
    const float tHigh = __uint_as_float(rank_stats[0]);
    const float tMid = __uint_as_float(rank_stats[1]);
    for (uint32_t base = vbx * blockDim.x; base < n; base += vbn * blockDim.x) {
        const uint32_t t = base + threadIdx.x;
        active += min(n - base, (uint32_t)blockDim.x);
        int32_t s = 0;
        if (t < n) {
            const uint4 cd = candidates[t];
            const uint32_t pixels = samples_map[(size_t)cd.z * CLAIM_PROBES + cd.w];
            // with a zero weight the variance term is 0 * vn with vn in [0, 1]
            // (clamp maps NaN to 0), so skipping the lookup is exact
            const float var = p.w_variance != 0.0f ? indirect_variance(lbuf, p.hash_width, cd.x) : 0.0f;
            const float pn = gclamp((float)pixels / gmax(p.pixels_norm, 1.0f), 0.0f, 1.0f);
            const float vn = gclamp(var / gmax(p.variance_norm, 1.0f), 0.0f, 1.0f);
            const float imp = p.w_pixels * pn + p.w_variance * vn;
            s = imp >= tHigh ? 2 : (imp >= tMid ? 1 : p.min_samples);
            uint32_t bin = (uint32_t)(imp * 256.0f);
            if (bin >= HISTOGRAM_BINS) bin = HISTOGRAM_BINS - 1u;
            atomicAdd(&hist[bin], 1u);
            // min/max of the float bits as uints (rank.comp:141-143)
            my_min = min(my_min, __float_as_uint(imp));
            my_max = max(my_max, __float_as_uint(imp));
            nsamp[t] = s;
        }
        // vector total = sum of max(s, 0) over the 8 lanes of the vector
        uint32_t tot = s > 0 ? (uint32_t)s : 0u;
        tot += __shfl_xor(tot, 1, 8);
        tot += __shfl_xor(tot, 2, 8);
        tot += __shfl_xor(tot, 4, 8);
        if ((t & 7u) == 0u && t < n) vec_tot[t >> 3] = tot;
        // the chunk of 256 candidates this iteration covers: its total, from
        // which k_rank_write offsets its vectors
        const uint32_t c = block_sum256(s > 0 ? (uint32_t)s : 0u, sum_lds);
        if (threadIdx.x == 0) chunk_tot[base / 256u] = c;
    }
    for (int off = 32; off > 0; off >>= 1) {
        my_min = min(my_min, (uint32_t)__shfl_xor((int)my_min, off, 64));
        my_max = max(my_max, (uint32_t)__shfl_xor((int)my_max, off, 64));
    }
    if ((threadIdx.x & 63u) == 0u) {
        atomicMin(&bmin, my_min);
        atomicMax(&bmax, my_max);
    }
    __syncthreads();
    uint32_t *part = rank_part + (size_t)(vbx % RANK_COPIES) * RANK_COPY_WORDS;
    if (threadIdx.x == 0 && active) {
        atomicAdd(&part[0], active);
        atomicMax(&part[1], ~bmin);   // the complement: a zeroed copy holds "no minimum"
        atomicMax(&part[2], bmax);
    }
    for (uint32_t i = threadIdx.x; i < HISTOGRAM_BINS; i += blockDim.x)
        if (hist[i]) atomicAdd(&part[3 + i], hist[i]);
}

// RankStats from the rank blocks' copies (rank.comp:141-145: count, min and
// max of the importance bits, histogram), which are cleared for the next
// pass.  One block of HISTOGRAM_BINS threads, after k_rank.
__device__ void rank_fold_copies(uint32_t *__restrict__ rank_stats, uint32_t *__restrict__ rank_part)
{
    const uint32_t i = threadIdx.x;
    uint32_t h = 0;
    for (uint32_t c = 0; c < RANK_COPIES; c++) {
        uint32_t *part = rank_part + (size_t)c * RANK_COPY_WORDS;
        h += part[3 + i];
        part[3 + i] = 0u;
    }
    rank_stats[8 + i] += h;
    if (i == 0) {
        uint32_t cnt = 0, nmin = 0, mx = 0;
        for (uint32_t c = 0; c < RANK_COPIES; c++) {
            uint32_t *part = rank_part + (size_t)c * RANK_COPY_WORDS;
            cnt += part[0];
            nmin = max(nmin, part[1]);
            mx = max(mx, part[2]);
            part[0] = part[1] = part[2] = 0u;
        }
        rank_stats[6] += cnt;
        rank_stats[4] = min(rank_stats[4], ~nmin);
        rank_stats[5] = max(rank_stats[5], mx);
    }
}

// Lockstep order of rank.comp's sample loop inside a vector: iteration i
// appends, in lane order, every lane with more than i samples.
template <bool MULTI>
__global__ __launch_bounds__(256) void k_rank_write(FrameParams p, const uint4 *__restrict__ candidates,
                                                    const int32_t *__restrict__ nsamp,
                                                    const uint32_t *__restrict__ vec_tot,
                                                    const uint32_t *__restrict__ chunk_tot, uint2 *__restrict__ shade_list,
                                                    uint32_t *__restrict__ counters, uint32_t *__restrict__ rank_stats,
                                                    uint32_t *__restrict__ rank_part,
                                                    MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        p = global_params(V.p);
        candidates = gptr(V.b.candidates);
        nsamp = gptr(V.b.nsamp);
        vec_tot = gptr(V.b.vec_cnt);
        chunk_tot = gptr(V.b.rank_chunk);
        shade_list = gptr(V.b.shade_list);
        counters = gptr(V.b.counters);
        rank_stats = gptr(V.b.rank_stats);
        rank_part = gptr(V.b.rank_part);
    }
    static_assert(HISTOGRAM_BINS == 256, "the fold takes one bin per thread of a 256-thread block");
    __shared__ uint32_t lds[4];
    if (vbx == 0) rank_fold_copies(rank_stats, rank_part);
    const uint32_t n = counters[VE_CNT_CANDIDATES];
    if (vbx * blockDim.x >= n) return;   // the whole block
    const uint32_t t = vbx * blockDim.x + threadIdx.x;
    const uint32_t v = t >> 3, l = t & 7u, v0 = v * 8u;
    // the vector's first shade-list index: the totals of the earlier 256-
    // candidate chunks plus the vectors before it in this one
    uint32_t before = 0;
    for (uint32_t i = threadIdx.x; i < vbx; i += blockDim.x) before += chunk_tot[i];
    const uint32_t off = block_sum256(before, lds);
    const uint32_t mine = (l == 0u && t < n) ? vec_tot[v] : 0u;
    const uint32_t excl = block_excl_scan256(mine, lds);
    const uint32_t vbase = off + (uint32_t)__shfl((int)excl, (int)((threadIdx.x & 63u) & ~7u), 64);
    const uint32_t vtot = (uint32_t)__shfl((int)mine, (int)((threadIdx.x & 63u) & ~7u), 64);
    if (t >= n) return;
    int32_t s[8];
    int32_t smax = 0;
    for (uint32_t k = 0; k < 8; k++) {
        s[k] = v0 + k < n ? nsamp[v0 + k] : 0;
        smax = max(smax, s[k]);
    }
    if (t == n - 1u) counters[VE_CNT_SHADE] = min(vbase + vtot, p.shade_budget);
    const uint4 cd = candidates[t];
    uint32_t pos = vbase;
    for (int32_t i = 0; i < smax; i++) {
        uint32_t before = 0, all = 0;
        for (uint32_t k = 0; k < 8; k++) {
            const uint32_t in_iter = s[k] > i;
            all += in_iter;
            before += (k < l) & in_iter;
        }
        if (i < s[l]) {
            const uint32_t idx = pos + before;
            if (idx < p.shade_budget) shade_list[idx] = make_uint2(cd.x, cd.y);
        }
        pos += all;
    }
}

// threshold.comp:31-78. The reference walks the histogram from the top bin
// down and stops where the running count first reaches topCount (resp.
// topMidCount); that bin is the largest b whose suffix count (bins b..255)
// reaches it, i.e. (number of bins with suffix >= count) - 1, as suffix
// counts fall with b. One wave: four bins per lane, a suffix scan across lanes.
static_assert(HISTOGRAM_BINS == 256, "threshold_wave holds four bins per lane");
__device__ void threshold_wave(const FrameParams &p, uint32_t *__restrict__ rank_stats,
                               const uint32_t *__restrict__ counters)
{
    if (counters[VE_CNT_CANDIDATES] == 0u) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint4 h = reinterpret_cast<const uint4 *>(rank_stats + 8)[lane];
    const uint32_t mine = h.x + h.y + h.z + h.w;
    uint32_t suf = mine;   // bins 4*lane .. 255
    for (uint32_t o = 1; o < 64u; o <<= 1) {
        const uint32_t t = __shfl_down(suf, o, 64);
        if (lane + o < 64u) suf += t;
    }
    const uint32_t total = __shfl(suf, 0, 64);
    if (total == 0u) {
        if (lane == 0u) {
            rank_stats[0] = 0x7F800000u;
            rank_stats[1] = 0xFF800000u;
        }
        return;
    }
    const float topF = gclamp(p.top_fraction, 0.0f, 1.0f);
    const float midF = gclamp(p.mid_fraction, 0.0f, 1.0f - topF);
    const uint32_t topCount = (uint32_t)(topF * (float)total);
    const uint32_t topMidCount = (uint32_t)((topF + midF) * (float)total);
    const uint32_t s3 = suf - mine + h.w, s2 = s3 + h.z, s1 = s2 + h.y, s0 = suf;
    const uint32_t n_top = (uint32_t)(__popcll(__ballot(s0 >= topCount)) + __popcll(__ballot(s1 >= topCount)) +
                                      __popcll(__ballot(s2 >= topCount)) + __popcll(__ballot(s3 >= topCount)));
    const uint32_t n_mid = (uint32_t)(__popcll(__ballot(s0 >= topMidCount)) + __popcll(__ballot(s1 >= topMidCount)) +
                                      __popcll(__ballot(s2 >= topMidCount)) + __popcll(__ballot(s3 >= topMidCount)));
    // no bin reaching the count leaves the top bin, as the reference's fallback
    const uint32_t thb = n_top ? n_top - 1u : HISTOGRAM_BINS - 1u;
    const uint32_t tmb = n_mid ? n_mid - 1u : HISTOGRAM_BINS - 1u;
    if (lane == 0u) {
        rank_stats[0] = __float_as_uint((float)thb / 256.0f);
        rank_stats[1] = __float_as_uint((float)tmb / 256.0f);
    }
}

__global__ __launch_bounds__(64) void k_threshold(FrameParams p, uint32_t *__restrict__ rank_stats,
                                                  const uint32_t *__restrict__ counters)
{
    threshold_wave(p, rank_stats, counters);
}

// ---------------------------------------------------------------------------
// emit.comp: one ray per shade-list entry, ray index = list index
// ---------------------------------------------------------------------------
// Bounce-0 rays that hit an emissive voxel the claim map does not hold yet
// set their bit in `em_bits`; k_em_register then registers them in ray order.
__device__ __forceinline__ bool emissive_unknown(const uint32_t *__restrict__ em_claim, uint32_t vid)
{
    const uint32_t ch = claim_hash(vid) % EMISSIVE_CLAIM_W;
    for (uint32_t q = 0; q < EMISSIVE_CLAIM_PROBES; q++) {
        const uint32_t c = em_claim[(size_t)ch * EMISSIVE_CLAIM_PROBES + q];
        if (c == vid) return false;
        if (c == 0u) return true;
    }
    return false;   // every probe taken by other ids: the reference cannot register it either
}

// A bounce-0 ray starts at the camera through its pixel: it is the pixel's
// primary ray, traversed by the primary pass with the same origin, direction
// and DDA.  When that traversal hit this entry's voxel within the bounce
// budget (a hit at step q is found by any budget >= q - 1), the trace pass
// would reach the same hit, so emit writes the ray's post-trace record (hit
// position, size 1, id, material, emissive flag) and marks it in `hit_bits`;
// the trace pass then leaves it alone.  The ray still counts as traced (the
// reference traverses it again); `VE_CNT_REUSED` and ray total 3 count it.
template <bool MULTI>
__global__ __launch_bounds__(256) void k_emit(FrameParams p, const uint2 *__restrict__ shade_list,
                                              Ray *__restrict__ rays, uint32_t *__restrict__ counters,
                                              const uint4 *__restrict__ gbuf, const uint2 *__restrict__ ghit,
                                              const Material *__restrict__ mats, const uint32_t *__restrict__ em_claim,
                                              uint32_t *__restrict__ em_bits, uint32_t *__restrict__ hit_bits,
                                              unsigned long long *__restrict__ totals, uint32_t *__restrict__ ray_li,
                                              uint32_t *__restrict__ death, uint32_t *__restrict__ rank_stats,
                                              MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        p = global_params(V.p);
        shade_list = gptr(V.b.shade_list);
        rays = gptr(V.b.rays);
        counters = gptr(V.b.counters);
        gbuf = gptr(V.b.gbuf);
        ghit = gptr(V.b.ghit);
        mats = gptr(V.b.materials);
        em_claim = gptr(V.b.emissive_claim);
        em_bits = gptr(V.b.em_bits);
        hit_bits = gptr(V.b.hit_bits);
        totals = gptr(V.b.ray_totals);
        ray_li = gptr(V.b.ray_li[0]);
        death = gptr(V.b.dep_death);
        rank_stats = gptr(V.b.rank_stats);
    }
    // the threshold pass, folded: it reads the rank pass's histogram, done
    // before this launch, and writes thresholds only the next frame's rank
    // pass reads
    if (p.fold_small_passes && vbx == 0 && threadIdx.x < 64u) threshold_wave(p, rank_stats, counters);
    __shared__ uint32_t block_reused;
    if (threadIdx.x == 0) block_reused = 0u;
    __syncthreads();
    const uint32_t t = vbx * blockDim.x + threadIdx.x;
    const uint32_t n = counters[VE_CNT_SHADE];
    if (t == 0) counters[VE_CNT_LIVE] = n;
    bool reused = false;
    if (t < n) {
        const uint2 en = shade_list[t];
        const uint32_t px = en.y % p.W, py = en.y / p.W;
        const f3 d = primary_dir(fp_cam(p), px, py, p.W, p.H);
        Ray r;
        r.ox = fp_cam(p)[0]; r.oy = fp_cam(p)[1]; r.oz = fp_cam(p)[2]; r.pad0 = 0.0f;
        r.dx = d.x; r.dy = d.y; r.dz = d.z; r.pad1 = 0.0f;
        r.tx = r.ty = r.tz = 1.0f;
        r.primary_vid = en.x;
        r.dlx = r.dly = r.dlz = 0.0f;
        const uint32_t sample_idx = fp_frame_idx(p) + en.y * 2654435761u + t * 1442695040u;
        r.rng = (sample_idx << 8) & 0xFFFFFF00u;
        r.ilx = r.ily = r.ilz = 0.0f;
        r.bounce = 0u;
        r.hit_vid = 0u;
        r.hit_material = 0u;
        r.pad2 = r.pad3 = 0u;
        if (p.reuse_primary_hits && p.max_bounces > 0u) {
            const uint4 g = gbuf[en.y];
            if (g.w == en.x && g.w != 0u && g.x - 1u <= p.bounce_cc) {
                const uint2 hp = ghit[en.y];
                r.ox = (float)(hp.x & 0xFFFFu);
                r.oy = (float)(hp.x >> 16);
                r.oz = (float)hp.y;
                r.pad0 = 1.0f;
                r.hit_vid = g.w;
                r.hit_material = g.y;
                reused = true;
                if (mats[g.y < MATERIAL_SLOTS ? g.y : MATERIAL_SLOTS - 1u].emissive && emissive_unknown(em_claim, g.w))
                    atomicOr(&em_bits[t >> 5], 1u << (t & 31u));
            }
        }
        rays[t] = r;
        ray_li[t] = t;      // shade-list index, carried through the survivor compaction
        death[t] = NONE;    // the iteration of its deposit, recorded after the bounces
    }
    const uint64_t m = __ballot(reused);
    const uint32_t lane = threadIdx.x & 63u;
    if ((lane & 31u) == 0u && t < SHADE_BUDGET_MAX) hit_bits[t >> 5] = (uint32_t)(m >> lane);
    if (lane == 0u && m) atomicAdd(&block_reused, (uint32_t)__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0 && block_reused) {
        atomicAdd(&counters[VE_CNT_REUSED], block_reused);
        atomicAdd(&totals[3], (unsigned long long)block_reused);
    }
}

// ---------------------------------------------------------------------------
// trace.comp
// ---------------------------------------------------------------------------
// trace.comp for ray t of the batch: the hit record (position, size, id,
// material) or a miss; first-shade rays that hit an unregistered emissive
// voxel are flagged for k_em_register.
template <bool STATS>
__device__ __forceinline__ void trace_ray(const uint32_t *__restrict__ words, const FrameParams &p, uint32_t iter,
                                          uint32_t t, Ray &r, const Material *__restrict__ mats,
                                          const uint32_t *__restrict__ em_claim, uint32_t *__restrict__ em_bits,
                                          uint32_t *cache, uint32_t &checks, uint32_t &reads)
{
    if (r.bounce >= p.max_bounces) return;
    const f3 o{r.ox, r.oy, r.oz}, d{r.dx, r.dy, r.dz};
    const Hit h = raycast<STATS>(words, p.jump, p.jump_level, p.n_words, p.depth, o, d, rcp3(d), p.bounce_cc, checks, reads,
                                 cache, 256u);
    if (h.hit) {
        r.ox = (float)h.px;
        r.oy = (float)h.py;
        r.oz = (float)h.pz;
        r.pad0 = (float)h.size;
        r.hit_vid = voxel_key(h.px, h.py, h.pz) + 1u;
        r.hit_material = h.material;
        if (iter == 0u && mats[h.material < MATERIAL_SLOTS ? h.material : MATERIAL_SLOTS - 1u].emissive &&
            emissive_unknown(em_claim, r.hit_vid))
            atomicOr(&em_bits[t >> 5], 1u << (t & 31u));
    } else {
        r.hit_vid = 0u;
        r.hit_material = 0u;
    }
}

template <bool STATS, bool MULTI>
__global__ __launch_bounds__(256) void k_trace(const uint32_t *__restrict__ words, FrameParams p, Ray *__restrict__ rays,
                                               const uint32_t *__restrict__ live_ptr, uint32_t iter,
                                               uint32_t *__restrict__ counters, unsigned long long *__restrict__ stats,
                                               unsigned long long *__restrict__ totals,
                                               const Material *__restrict__ mats,
                                               const uint32_t *__restrict__ em_claim, uint32_t *__restrict__ em_bits,
                                               uint32_t *__restrict__ live_out, const uint32_t *__restrict__ hit_bits,
                                               MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        words = gptr(V.words);
        p = global_params(V.p);
        rays = gptr(V.b.rays);
        live_ptr = gptr(V.b.counters) + VE_CNT_LIVE;
        iter = 0u;
        counters = gptr(V.b.counters);
        stats = gptr(V.b.node_stats);
        totals = gptr(V.b.ray_totals);
        mats = gptr(V.b.materials);
        em_claim = gptr(V.b.emissive_claim);
        em_bits = gptr(V.b.em_bits);
        live_out = gptr(V.b.counters) + VE_CNT_LIVE_B;
        hit_bits = gptr(V.b.hit_bits);
    }
    const uint32_t live = *live_ptr;
    const uint32_t t = vbx * blockDim.x + threadIdx.x;
    // the shade pass of this iteration appends its survivors from zero; the
    // counter was last read by the previous iteration
    if (t == 0) *live_out = 0u;
    if (t == 0 && live > 0u) {
        atomicAdd(&counters[VE_CNT_RAYS_TRACED], live);
        atomicMax(&counters[VE_CNT_BOUNCES_RUN], iter + 1u);
        atomicAdd(&totals[1], (unsigned long long)live);
    }
    __shared__ uint32_t path_cache[VE_MAX_DEPTH * 256];
    uint32_t checks = 0, reads = 0;
    // bounce-0 rays whose hit record the emit pass already filled in
    const bool filled = iter == 0u && t < live && ((hit_bits[t >> 5] >> (t & 31u)) & 1u);
    if (t < live && !filled) trace_ray<STATS>(words, p, iter, t, rays[t], mats, em_claim, em_bits,
                                              path_cache + threadIdx.x, checks, reads);
    if (STATS) {
        unsigned long long a = reads, c = checks;
        for (int off = 32; off > 0; off >>= 1) {
            a += __shfl_xor(a, off, 64);
            c += __shfl_xor(c, off, 64);
        }
        if ((threadIdx.x & 63u) == 0u) {
            atomicAdd(&stats[1], a);
            atomicAdd(&stats[3], c);
        }
    }
}

// ---------------------------------------------------------------------------
// shade.comp
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint4 pack_rgb_l2(f3 c)   // shade.comp:154-160
{
    uint4 o;
    o.x = (uint32_t)(gclamp(c.x, 0.0f, 1024.0f) * 255.0f);
    o.y = (uint32_t)(gclamp(c.y, 0.0f, 1024.0f) * 255.0f);
    o.z = (uint32_t)(gclamp(c.z, 0.0f, 1024.0f) * 255.0f);
    const uint32_t L = (o.x + o.y + o.z) / 3u;
    o.w = L * L;
    return o;
}

__device__ __forceinline__ uint32_t lbuf_row_of(uint32_t hash_width, uint32_t vid) { return vid % (hash_width - 1u) + 1u; }
__device__ __forceinline__ uint32_t lbuf_row(const FrameParams &p, uint32_t vid) { return lbuf_row_of(p.hash_width, vid); }

__device__ f3 lookup_normal(const uint32_t *__restrict__ nbuf, uint32_t hw, uint32_t vid, f3 rd)   // shade.comp:261-275
{
    const uint32_t h = vid % (hw - 1u) + 1u;
    const uint32_t *bucket = nbuf + (size_t)h * NB_WORDS;
    uint32_t own, from = 0;
    for (;;) {   // every slot owned by vid, in order, until one has a normal
        const uint32_t s = find_owner<false>(bucket, from, NBUFFER_SLOTS, vid, own);
        if (s >= NBUFFER_SLOTS) break;
        const uint32_t nf = bucket[NB_NRM + s];
        if (nf != 0u) return normalize(unpack_normal(nf));
        from = s + 1;
    }
    const float ax = fabsf(rd.x), ay = fabsf(rd.y), az = fabsf(rd.z);
    if (ax > ay && ax > az) return f3{-gsign(rd.x), 0.0f, 0.0f};
    if (ay > az) return f3{0.0f, -gsign(rd.y), 0.0f};
    return f3{0.0f, 0.0f, -gsign(rd.z)};
}

// Exclusive prefix of one value per thread over a block of up to 1024
// threads (wave scans, then the wave totals from LDS); `total` gets the sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *wsum, uint32_t &total)
{
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6, nwv = (blockDim.x + 63u) >> 6;
    uint32_t incl = v;
    for (uint32_t o = 1; o < 64u; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63u) wsum[wv] = incl;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (uint32_t k = 0; k < nwv; k++) {
        const uint32_t x = wsum[k];
        if (k < wv) before += x;
        total += x;
    }
    __syncthreads();   // wsum may be rewritten by the next scan
    return before + incl - v;
}

// The emissive record of ray r's hit (shade.comp:604-614).
__device__ __forceinline__ EmissiveVoxel emissive_of(const Ray &r, const Material *__restrict__ mats, uint32_t vid)
{
    const Material &m8 = mats[r.hit_material < MATERIAL_SLOTS ? r.hit_material : MATERIAL_SLOTS - 1u];
    const f3 ee = muls(f3{m8.color[0], m8.color[1], m8.color[2]}, m8.emissive_intensity);
    const uint32_t hs = (uint32_t)r.pad0;
    return EmissiveVoxel{ee.x, ee.y, ee.z, vid, (uint32_t)r.ox, (uint32_t)r.oy, (uint32_t)r.oz, hs > 1u ? hs : 1u};
}

// Emissive registry of the first shade (shade.comp:592-619) in the
// reference's order: rays in index order register the ids k_trace flagged.
// A ray's claim probes its row in order: its id found -> nothing; an empty
// probe -> claimed, and registered while the registry is below capacity;
// every probe another id's -> nothing.  One block; the bitmap is scanned 1024
// words at a time and cleared as it goes, and the flagged rays of a chunk are
// taken in order, EM_BATCH at a time:
//   - a few (at most EM_SERIAL): one thread runs the probes on global memory;
//   - more: the outcome is computed in parallel.  Only the first ray of an id
//     can claim (a later one finds the id, or the same full row); claims of
//     different rows are independent, so one thread per distinct row runs its
//     first-of-id rays in ray order over the row's probes held in registers;
//     registration indices are the prefix count of the claims in ray order,
//     cut at the capacity.
// Records the count before this frame's registrations and the ray index of
// each registration, so that shade can give every SIMD-8 vector of rays the
// registry size it would have seen.  The block is kept to one wave per SIMD
// (256 threads, four bitmap words each): the pass sits between the trace and
// the shade on the frame's critical path, while the next frame's primary pass
// may hold most wave slots of the chip -- a 1024-thread block waited up to
// 0.5 ms there for a compute unit with 16 free slots.
constexpr uint32_t EM_THREADS = 256;
constexpr uint32_t EM_CHUNK = 4u * EM_THREADS;   // bitmap words per scan step
constexpr uint32_t EM_BATCH = EM_THREADS;        // flagged rays taken at once, one per thread
constexpr uint32_t EM_KEYS = 2u * EM_BATCH;      // LDS open-addressed tables of the batch's ids and rows
constexpr uint32_t EM_SERIAL = 24;               // up to this many: one thread
static_assert(SHADE_BUDGET_MAX % (32u * 4u) == 0u, "the bitmap is read four words at a time");
template <bool MULTI>
__global__ __launch_bounds__(EM_THREADS) void k_em_register(const Ray *__restrict__ rays, const uint32_t *__restrict__ live_ptr,
                                                      const Material *__restrict__ mats, uint32_t *__restrict__ em_claim,
                                                      EmissiveVoxel *__restrict__ emissives, uint32_t *__restrict__ em_bits,
                                                      uint32_t *__restrict__ em_reg, uint32_t *__restrict__ counters,
                                                            MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        rays = gptr(V.b.rays);
        live_ptr = gptr(V.b.counters) + VE_CNT_LIVE;
        mats = gptr(V.b.materials);
        em_claim = gptr(V.b.emissive_claim);
        emissives = gptr(V.b.emissives);
        em_bits = gptr(V.b.em_bits);
        em_reg = gptr(V.b.em_reg);
        counters = gptr(V.b.counters);
    }
    static_assert(EM_KEYS >= 2u * EM_BATCH, "the key tables stay at most half full");
    __shared__ uint32_t wsum[EM_THREADS / 64u];
    __shared__ uint32_t list_ri[EM_BATCH], e_vid[EM_BATCH], e_slot[EM_BATCH], e_claim[EM_BATCH], seg[EM_BATCH];
    __shared__ uint32_t vkey[EM_KEYS], vfirst[EM_KEYS], rkey[EM_KEYS], rslot[EM_KEYS];
    __shared__ uint32_t slot_row[EM_BATCH], row_cnt[EM_BATCH], row_off[EM_BATCH], row_fill[EM_BATCH];
    __shared__ uint32_t count, nreg, n_slots;
    const uint32_t tid = threadIdx.x;
    const uint32_t live = *live_ptr;
    const uint32_t nw = (live + 31u) / 32u;
    if (tid == 0) {
        count = counters[VE_CNT_EMISSIVES];
        nreg = 0u;
        counters[VE_CNT_EM_BASE] = count;
    }
    for (uint32_t k = tid; k < EM_KEYS; k += blockDim.x) {
        vkey[k] = NONE;
        rkey[k] = NONE;
    }
    __syncthreads();
    for (uint32_t base = 0; base < nw; base += EM_CHUNK) {
        const uint32_t i = base + 4u * tid;   // this thread's four words: i .. i + 3
        uint4 w4 = make_uint4(0u, 0u, 0u, 0u);
        if (i < nw) {
            w4 = *reinterpret_cast<const uint4 *>(em_bits + i);   // within the SHADE_BUDGET_MAX / 32 words
            if (i + 1u >= nw) w4.y = 0u;   // words past the live rays hold no flags; keep them out
            if (i + 2u >= nw) w4.z = 0u;
            if (i + 3u >= nw) w4.w = 0u;
            if (w4.x | w4.y | w4.z | w4.w) *reinterpret_cast<uint4 *>(em_bits + i) = make_uint4(0u, 0u, 0u, 0u);
        }
        const uint32_t ws[4] = {w4.x, w4.y, w4.z, w4.w};
        if (!__syncthreads_or((w4.x | w4.y | w4.z | w4.w) != 0u)) continue;
        uint32_t total;
        const uint32_t first = block_excl_scan((uint32_t)(__popc(w4.x) + __popc(w4.y) + __popc(w4.z) + __popc(w4.w)),
                                               wsum, total);   // this thread's first list place
        for (uint32_t r0 = 0; r0 < total; r0 += EM_BATCH) {
            const uint32_t nb = min(EM_BATCH, total - r0);
            {   // the batch's rays in order
                uint32_t at = first;
                for (uint32_t k = 0; k < 4u && at < r0 + EM_BATCH; k++)
                    for (uint32_t m = ws[k]; m && at < r0 + EM_BATCH; m &= m - 1u, at++)
                        if (at >= r0) list_ri[at - r0] = (i + k) * 32u + (uint32_t)(__ffs(m) - 1);
            }
            if (tid == 0) n_slots = 0u;
            __syncthreads();
            if (nb <= EM_SERIAL) {
                if (tid == 0) {
                    for (uint32_t j = 0; j < nb; j++) {
                        const uint32_t ri = list_ri[j];
                        const Ray &r = rays[ri];
                        const uint32_t vid = r.hit_vid;
                        const uint32_t ch = claim_hash(vid) % EMISSIVE_CLAIM_W;
                        for (uint32_t q = 0; q < EMISSIVE_CLAIM_PROBES; q++) {
                            uint32_t &c = em_claim[(size_t)ch * EMISSIVE_CLAIM_PROBES + q];
                            if (c == vid) break;
                            if (c != 0u) continue;
                            c = vid;
                            if (count < EMISSIVE_CAPACITY) {
                                emissives[count] = emissive_of(r, mats, vid);
                                count++;
                                em_reg[nreg++] = ri;
                            }
                            break;
                        }
                    }
                }
                __syncthreads();
                continue;
            }
            // ids and rows of the batch: the first ray of each id, a slot per row
            uint32_t vk = 0, rk = 0;
            if (tid < nb) {
                const uint32_t vid = rays[list_ri[tid]].hit_vid;
                const uint32_t row = claim_hash(vid) % EMISSIVE_CLAIM_W;
                e_vid[tid] = vid;
                e_claim[tid] = 0u;
                vk = vid % EM_KEYS;
                for (;;) {   // at most EM_BATCH keys in EM_KEYS: a free one is always found
                    const uint32_t prev = atomicCAS(&vkey[vk], NONE, vid);
                    if (prev == NONE) {
                        vfirst[vk] = NONE;
                        break;
                    }
                    if (prev == vid) break;
                    vk = (vk + 1u) % EM_KEYS;
                }
                rk = row % EM_KEYS;
                for (;;) {
                    const uint32_t prev = atomicCAS(&rkey[rk], NONE, row);
                    if (prev == NONE) {
                        const uint32_t sl = atomicAdd(&n_slots, 1u);
                        slot_row[sl] = row;
                        row_cnt[sl] = 0u;
                        row_fill[sl] = 0u;
                        rslot[rk] = sl;
                        break;
                    }
                    if (prev == row) break;
                    rk = (rk + 1u) % EM_KEYS;
                }
            }
            __syncthreads();
            if (tid < nb) atomicMin(&vfirst[vk], tid);
            __syncthreads();
            const bool keeper = tid < nb && vfirst[vk] == tid;
            uint32_t sl = 0;
            if (tid < nb) {
                sl = rslot[rk];
                e_slot[tid] = sl;
                if (keeper) atomicAdd(&row_cnt[sl], 1u);
            }
            __syncthreads();
            const uint32_t ns = n_slots;
            uint32_t unused;
            const uint32_t off = block_excl_scan(tid < ns ? row_cnt[tid] : 0u, wsum, unused);
            if (tid < ns) row_off[tid] = off;
            __syncthreads();
            if (keeper) seg[row_off[sl] + atomicAdd(&row_fill[sl], 1u)] = tid;
            __syncthreads();
            if (tid < ns) {   // this row's first-of-id rays in ray order over its probes
                const uint32_t o = row_off[tid], n = row_cnt[tid];
                for (uint32_t a = 1; a < n; a++) {   // a row holds few: insertion sort
                    const uint32_t x = seg[o + a];
                    uint32_t b = a;
                    for (; b > 0u && seg[o + b - 1u] > x; b--) seg[o + b] = seg[o + b - 1u];
                    seg[o + b] = x;
                }
                uint32_t *rowp = em_claim + (size_t)slot_row[tid] * EMISSIVE_CLAIM_PROBES;
                uint32_t pr[EMISSIVE_CLAIM_PROBES];
#pragma unroll
                for (uint32_t q = 0; q < EMISSIVE_CLAIM_PROBES; q++) pr[q] = rowp[q];
                bool changed = false;
                for (uint32_t a = 0; a < n; a++) {
                    const uint32_t j = seg[o + a], vid = e_vid[j];
                    bool done = false;
#pragma unroll
                    for (uint32_t q = 0; q < EMISSIVE_CLAIM_PROBES; q++) {
                        if (done) continue;
                        if (pr[q] == vid) {
                            done = true;
                        } else if (pr[q] == 0u) {
                            pr[q] = vid;
                            e_claim[j] = 1u;
                            changed = true;
                            done = true;
                        }
                    }
                }
                if (changed) {
#pragma unroll
                    for (uint32_t q = 0; q < EMISSIVE_CLAIM_PROBES; q++) rowp[q] = pr[q];
                }
            }
            __syncthreads();
            uint32_t claims;
            const uint32_t cl = tid < nb ? e_claim[tid] : 0u;
            const uint32_t before = block_excl_scan(cl, wsum, claims);
            const uint32_t c0 = count, room = EMISSIVE_CAPACITY - min(c0, EMISSIVE_CAPACITY);
            if (cl && before < room) {   // registered: the claims in ray order, up to the capacity
                const uint32_t ri = list_ri[tid];
                emissives[c0 + before] = emissive_of(rays[ri], mats, e_vid[tid]);
                em_reg[nreg + before] = ri;
            }
            __syncthreads();
            if (tid == 0) {
                const uint32_t add = min(claims, room);
                count = c0 + add;
                nreg += add;
            }
            for (uint32_t k = tid; k < EM_KEYS; k += blockDim.x) {
                vkey[k] = NONE;
                rkey[k] = NONE;
            }
            __syncthreads();
        }
    }
    if (tid == 0) {
        counters[VE_CNT_EMISSIVES] = count;
        counters[VE_CNT_EM_NEW] = nreg;
    }
}

// ---- next-event estimation (shade.comp:276-539, 620-836) -------------------
struct Reservoir {
    uint32_t light_idx, light_vid, face_seed;
    float wsum;
    uint32_t m;
    float target_pdf;
};

__device__ __forceinline__ void r_update(Reservoir &r, uint32_t idx, uint32_t lvid, uint32_t seed, float target,
                                         float w, uint32_t &rng)   // shade.comp:320-336
{
    if (w <= 0.0f) {
        r.m += 1u;
        return;
    }
    r.wsum += w;
    r.m += 1u;
    if (rand_lds(rng) * r.wsum < w) {
        r.light_idx = idx;
        r.light_vid = lvid;
        r.face_seed = seed;
        r.target_pdf = target;
    }
}

__device__ __forceinline__ void r_combine(Reservoir &a, const Reservoir &b, float t_us, uint32_t &rng)   // :352-365
{
    if (b.m == 0u || b.target_pdf <= 0.0f || t_us <= 0.0f) return;
    const float w = t_us * b.wsum / b.target_pdf;
    a.wsum += w;
    a.m += b.m;
    if (rand_lds(rng) * a.wsum < w) {
        a.light_idx = b.light_idx;
        a.light_vid = b.light_vid;
        a.face_seed = b.face_seed;
        a.target_pdf = t_us;
    }
}

__device__ __forceinline__ void r_cap(Reservoir &r, uint32_t max_m)   // shade.comp:372-377
{
    if (r.m > max_m && max_m > 0u) {
        r.wsum = r.wsum * (float)max_m / (float)r.m;
        r.m = max_m;
    }
}

__device__ __forceinline__ uint32_t pack_face_seed(uint32_t face, float u, float v)   // shade.comp:382-386
{
    const uint32_t ui = (uint32_t)(gclamp(u, 0.0f, 0.9999f) * 16383.0f);
    const uint32_t vi = (uint32_t)(gclamp(v, 0.0f, 0.9999f) * 16383.0f);
    return (face & 0xFu) | ((ui & 0x3FFFu) << 4) | ((vi & 0x3FFFu) << 18);
}

__device__ __forceinline__ void face_frame(uint32_t face, f3 &n, f3 &t, f3 &b)   // shade.comp:398-413
{
    if (face < 2u) {
        n = f3{face == 0u ? 1.0f : -1.0f, 0.0f, 0.0f};
        t = f3{0.0f, 1.0f, 0.0f};
        b = f3{0.0f, 0.0f, 1.0f};
    } else if (face < 4u) {
        n = f3{0.0f, face == 2u ? 1.0f : -1.0f, 0.0f};
        t = f3{1.0f, 0.0f, 0.0f};
        b = f3{0.0f, 0.0f, 1.0f};
    } else {
        n = f3{0.0f, 0.0f, face == 4u ? 1.0f : -1.0f};
        t = f3{1.0f, 0.0f, 0.0f};
        b = f3{0.0f, 1.0f, 0.0f};
    }
}

// fCenter + fNormal*h + ((u-0.5)*2*h)*fT + ((v-0.5)*2*h)*fB, summed left to right
__device__ __forceinline__ f3 face_point(const EmissiveVoxel &e, f3 fn, f3 ft, f3 fb, float u, float v)
{
    const float half = (float)e.size * 0.5f;
    const f3 c = add(f3{(float)e.px, (float)e.py, (float)e.pz}, f3{half, half, half});
    return add(add(add(c, muls(fn, half)), muls(ft, (u - 0.5f) * 2.0f * half)), muls(fb, (v - 0.5f) * 2.0f * half));
}

__device__ __forceinline__ void light_sample_point(uint32_t seed, const EmissiveVoxel &e, f3 &P, f3 &N)   // :393-420
{
    uint32_t face = seed & 0xFu;
    const float u = (float)((seed >> 4) & 0x3FFFu) / 16383.0f, v = (float)((seed >> 18) & 0x3FFFu) / 16383.0f;
    if (face > 5u) face = 5u;
    f3 ft, fb;
    face_frame(face, N, ft, fb);
    P = face_point(e, N, ft, fb, u, v);
}

__device__ __forceinline__ float luminance_floor(const EmissiveVoxel &e)
{
    return gmax(dot(f3{e.ex, e.ey, e.ez}, f3{0.2126f, 0.7152f, 0.0722f}), 1e-6f);
}

__device__ float target_pdf(const EmissiveVoxel &e, uint32_t seed, f3 sp, f3 nrm)   // shade.comp:423-438
{
    f3 P, N;
    light_sample_point(seed, e, P, N);
    const f3 to = sub(P, sp);
    const float d2 = dot(to, to);
    if (d2 < 1e-6f) return 0.0f;
    const float d = sqrtf(d2);
    const f3 ld = divs(to, d);
    const float cx = dot(nrm, ld);
    if (cx <= 0.0f) return 0.0f;
    const float cy = dot(f3{-ld.x, -ld.y, -ld.z}, N);
    if (cy <= 0.0f) return 0.0f;
    const float area = 6.0f * (float)(e.size * e.size);
    return luminance_floor(e) * cy * area / d2;
}

// rBufferLoad (shade.comp:466-485) on the device bucket layout
__device__ bool rbuf_load(const FrameParams &p, const uint32_t *__restrict__ rbuf, uint32_t vid, Reservoir &r)
{
    r = Reservoir{0u, 0u, 0u, 0.0f, 0u, 0.0f};
    const uint32_t *bucket = rbuf + (size_t)(vid % (p.hash_width - 1u) + 1u) * RB_WORDS;
    uint32_t own;
    const uint32_t s = find_owner<false>(bucket, 0, RBUFFER_SLOTS, vid, own);
    if (s >= RBUFFER_SLOTS) return false;
    const uint32_t *pl = bucket + RB_PAY + s * RB_PAYW;
    r.light_idx = pl[0];
    r.light_vid = pl[1];
    r.face_seed = pl[2];
    r.wsum = __uint_as_float(pl[3]);
    r.m = pl[4];
    r.target_pdf = __uint_as_float(pl[5]);
    const uint32_t t = bucket[RB_TIME + s];
    if (!p.disable_stale_reset && t != 0u && t + p.update_time < fp_time(p)) {
        r = Reservoir{0u, 0u, 0u, 0.0f, 0u, 0.0f};
        return false;
    }
    return true;
}

// rBufferStore (shade.comp:501-519): claim (or keep) a slot, else evict the
// least recently written one
__device__ void rbuf_store(const FrameParams &p, uint32_t *__restrict__ rbuf, uint32_t vid, const Reservoir &r)
{
    uint32_t *bucket = rbuf + (size_t)(vid % (p.hash_width - 1u) + 1u) * RB_WORDS;
    bool claimed;
    uint32_t s = claim_slot(bucket, RBUFFER_SLOTS, vid, claimed);
    if (s >= RBUFFER_SLOTS) {
        uint32_t oldest = 0xFFFFFFFFu;
        s = 0;
        for (uint32_t q = 0; q < RBUFFER_SLOTS; q++) {
            const uint32_t t = __atomic_load_n(&bucket[RB_TIME + q], __ATOMIC_RELAXED);
            if (t < oldest) {
                oldest = t;
                s = q;
            }
        }
        atomicExch(&bucket[s], vid);
    }
    uint32_t *pl = bucket + RB_PAY + s * RB_PAYW;
    pl[0] = r.light_idx;
    pl[1] = r.light_vid;
    pl[2] = r.face_seed;
    pl[3] = __float_as_uint(r.wsum);
    pl[4] = r.m;
    pl[5] = __float_as_uint(r.target_pdf);
    bucket[RB_TIME + s] = fp_time(p);
}

__device__ __forceinline__ bool light_valid(const Reservoir &r, const EmissiveVoxel *__restrict__ em, uint32_t count)
{
    if (r.m == 0u || r.light_idx >= count) return false;   // shade.comp:446-449
    return em[r.light_idx].vid == r.light_vid;
}

__device__ __forceinline__ uint32_t neighbour_vid(uint32_t hx, uint32_t hy, uint32_t hz, uint32_t size, int ox, int oy,
                                                  int oz)   // shade.comp:527-535
{
    const int qx = (int)hx + ox * (int)size, qy = (int)hy + oy * (int)size, qz = (int)hz + oz * (int)size;
    if (qx < 0 || qy < 0 || qz < 0) return 0u;
    const uint32_t m = ~(size - 1u);
    return voxel_key(((uint32_t)qx & m) & 0x3FFu, ((uint32_t)qy & m) & 0x3FFu, ((uint32_t)qz & m) & 0x3FFu) + 1u;
}

// Context of one shade invocation's next-event estimation.
struct NeeCtx {
    const uint32_t *words;
    const EmissiveVoxel *em;
    uint32_t *rbuf;
    uint32_t *cache;        // this thread's LDS path-cache column
    uint32_t shadow_rays;   // RaycastVisibilityToVoxel calls
};

// RaycastVisibilityToVoxel (internal.glsl:406-410) with the bounce check budget
__device__ bool visible_to(const FrameParams &p, NeeCtx &c, f3 o, f3 d, uint32_t target)
{
    uint32_t checks = 0, reads = 0;
    c.shadow_rays++;
    const Hit h = raycast<false>(c.words, p.jump, p.jump_level, p.n_words, p.depth, o, d, rcp3(d), p.bounce_cc, checks,
                                 reads, c.cache, 256u);
    return h.hit && voxel_key(h.px, h.py, h.pz) + 1u == target;
}

__device__ void nee_restir(const FrameParams &p, NeeCtx &c, Ray &r, const Material &m, f3 nrm, uint32_t hx,
                           uint32_t hy, uint32_t hz, uint32_t hsize, uint32_t n_lights, uint32_t count)   // :646-815
{
    const f3 sp = add(add(f3{(float)hx, (float)hy, (float)hz}, f3{0.5f, 0.5f, 0.5f}), nrm);
    Reservoir res{0u, 0u, 0u, 0.0f, 0u, 0.0f};
    const uint32_t pvid = r.primary_vid;
    const int nm = p.restir_m > 1 ? p.restir_m : 1;
    const float inv_src = (float)n_lights * 6.0f;
    for (int k = 0; k < nm; k++) {
        uint32_t pick = (uint32_t)(rand_lds(r.rng) * (float)n_lights);
        if (pick >= n_lights) pick = n_lights - 1u;
        const EmissiveVoxel cand = c.em[pick];
        uint32_t face = (uint32_t)(rand_lds(r.rng) * 6.0f);
        if (face >= 6u) face = 5u;
        const float uu = rand_lds(r.rng);
        const float vv = rand_lds(r.rng);
        const uint32_t seed = pack_face_seed(face, uu, vv);
        const float tg = target_pdf(cand, seed, sp, nrm);
        r_update(res, pick, cand.vid, seed, tg, tg * inv_src, r.rng);
    }
    if (p.restir_temporal) {
        Reservoir tr;
        if (rbuf_load(p, c.rbuf, pvid, tr) && light_valid(tr, c.em, count)) {
            r_cap(tr, (uint32_t)p.restir_max_m);
            r_combine(res, tr, tr.target_pdf, r.rng);
        }
    }
    if (p.restir_spatial && p.restir_neighbors > 0) {
        const f3 up = fabsf(nrm.z) < 0.999f ? f3{0.0f, 0.0f, 1.0f} : f3{1.0f, 0.0f, 0.0f};
        const f3 T = normalize(cross3(up, nrm));
        const f3 B = cross3(nrm, T);
        const int K = p.restir_neighbors < 6 ? p.restir_neighbors : 6;
        for (int k = 0; k < K; k++) {
            const float ang = 6.2831853f * rand_lds(r.rng);
            const float rr = p.restir_radius * sqrtf(rand_lds(r.rng));
            const f3 off = muls(add(muls(T, m_cos(ang)), muls(B, m_sin(ang))), rr);
            const int lx = (int)rintf(off.x), ly = (int)rintf(off.y), lz = (int)rintf(off.z);
            if (lx == 0 && ly == 0 && lz == 0) continue;
            const uint32_t nv = neighbour_vid(hx, hy, hz, hsize, lx, ly, lz);
            if (nv == 0u || nv == pvid) continue;
            Reservoir nb;
            if (!rbuf_load(p, c.rbuf, nv, nb) || !light_valid(nb, c.em, count)) continue;
            const float t_us = target_pdf(c.em[nb.light_idx], nb.face_seed, sp, nrm);
            if (t_us <= 0.0f) continue;
            r_cap(nb, (uint32_t)p.restir_max_m);
            r_combine(res, nb, t_us, r.rng);
        }
    }
    if (res.m > 0u && res.target_pdf > 0.0f) {
        const EmissiveVoxel e = c.em[res.light_idx];
        if (e.vid == res.light_vid) {
            f3 P, N;
            light_sample_point(res.face_seed, e, P, N);
            const f3 to = sub(P, sp);
            const float d2 = dot(to, to);
            const float d = sqrtf(gmax(d2, 1e-12f));
            if (d > 1e-3f) {
                const f3 ld = divs(to, d);
                const float cx = gmax(dot(nrm, ld), 0.0f);
                const float cy = gmax(dot(f3{-ld.x, -ld.y, -ld.z}, N), 0.0f);
                if (cx > 0.0f && cy > 0.0f && visible_to(p, c, sp, ld, res.light_vid)) {
                    const f3 bc = evaluate_brdf(f3{r.dx, r.dy, r.dz}, ld, nrm, f3{m.color[0], m.color[1], m.color[2]},
                                                m.roughness, m.metallic, m.specular);
                    const float W = res.wsum / ((float)res.m * res.target_pdf);
                    const float sz2 = (float)(e.size * e.size);
                    const float geom = cy / gmax(d2, 1e-6f);
                    f3 ct = muls(muls(muls(mul(f3{e.ex, e.ey, e.ez}, bc), geom), sz2), W);
                    ct = vmin3(ct, f3{20.0f, 20.0f, 20.0f});
                    const f3 w = mul(f3{r.tx, r.ty, r.tz}, ct);
                    r.dlx += w.x;
                    r.dly += w.y;
                    r.dlz += w.z;
                }
            }
        }
    }
    if (res.m > 0u) rbuf_store(p, c.rbuf, pvid, res);
}

__device__ void nee_ris(const FrameParams &p, NeeCtx &c, Ray &r, const Material &m, f3 nrm, uint32_t hx, uint32_t hy,
                        uint32_t hz, uint32_t n_lights)   // shade.comp:816-836
{
    const f3 sp = add(add(f3{(float)hx, (float)hy, (float)hz}, f3{0.5f, 0.5f, 0.5f}), nrm);
    float wsum = 0.0f, chosen_t = 0.0f;
    uint32_t chosen = 0u;
    for (int k = 0; k < 4; k++) {
        uint32_t pick = (uint32_t)(rand_lds(r.rng) * (float)n_lights);
        if (pick >= n_lights) pick = n_lights - 1u;
        const EmissiveVoxel cand = c.em[pick];
        const float tg = luminance_floor(cand) * (float)(cand.size * cand.size);
        wsum += tg;
        if (rand_lds(r.rng) * wsum < tg) {
            chosen = pick;
            chosen_t = tg;
        }
    }
    const float avg = wsum / 4.0f;
    const EmissiveVoxel e = c.em[chosen];
    uint32_t face = (uint32_t)(rand_lds(r.rng) * 6.0f);
    if (face >= 6u) face = 5u;
    f3 fn, ft, fb;
    face_frame(face, fn, ft, fb);
    const float ru = rand_lds(r.rng);
    const float rv = rand_lds(r.rng);
    f3 ld = sub(face_point(e, fn, ft, fb, ru, rv), sp);
    const float d2 = dot(ld, ld);
    const float d = sqrtf(d2);
    if (d > 1e-3f && chosen_t > 0.0f) {
        ld = divs(ld, d);
        const float cx = gmax(dot(nrm, ld), 0.0f);
        const float cy = gmax(dot(f3{-ld.x, -ld.y, -ld.z}, fn), 0.0f);
        if (cx > 0.0f && cy > 0.0f && visible_to(p, c, sp, ld, e.vid)) {
            const f3 bc = evaluate_brdf(f3{r.dx, r.dy, r.dz}, ld, nrm, f3{m.color[0], m.color[1], m.color[2]},
                                        m.roughness, m.metallic, m.specular);
            const float area = 6.0f * (float)e.size * (float)e.size;
            const float ris = (float)n_lights * avg / gmax(chosen_t, 1e-6f);
            const float geom = cy / gmax(d2, 1e-6f);
            f3 ct = muls(muls(muls(mul(f3{e.ex, e.ey, e.ez}, bc), geom), area), ris);
            ct = vmin3(ct, f3{20.0f, 20.0f, 20.0f});
            const f3 w = mul(f3{r.tx, r.ty, r.tz}, ct);
            if (r.bounce == 0u) {
                r.dlx += w.x;
                r.dly += w.y;
                r.dlz += w.z;
            } else {
                r.ilx += w.x;
                r.ily += w.y;
                r.ilz += w.z;
            }
        }
    }
}

// Registry size seen by the SIMD-8 vector of ray `t` in the first shade: the
// count before this frame's registrations plus those made by rays of this or
// an earlier vector (em_reg is ascending).
__device__ __forceinline__ uint32_t registry_seen(const uint32_t *__restrict__ counters,
                                                  const uint32_t *__restrict__ em_reg, uint32_t t)
{
    const uint32_t last = t | 7u;
    uint32_t lo = 0, hi = counters[VE_CNT_EM_NEW];
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (em_reg[mid] <= last) lo = mid + 1u;
        else hi = mid;
    }
    return counters[VE_CNT_EM_BASE] + lo;
}

// shade.comp for ray t of the batch at bounce iteration `iter`: NEE, the next
// direction and Russian roulette.  Returns whether the ray continues (r holds
// the next bounce's ray), ends with a deposit of its direct/indirect light
// into its primary voxel (r's dl/il), or ends without one.
// ANY_BOUNCE false: the ray is past bounce 0 (reservoir reuse and the
// first-shade registry view are then never needed and not compiled in).
// shade_ray outcomes: the ray ends without depositing, survives, or deposits
// and ends at one of shade.comp's three call sites (miss, last bounce,
// Russian roulette), RAY_DEPOSITS + DEP_*
enum : int { RAY_DONE = 0, RAY_SURVIVES = 1, RAY_DEPOSITS = 2 };

template <bool ANY_BOUNCE>
__device__ __forceinline__ int shade_ray(const FrameParams &p, uint32_t iter, uint32_t t, Ray &r,
                                         const uint32_t *__restrict__ nbuf, const Material *__restrict__ mats,
                                         const uint32_t *__restrict__ em_reg, const uint32_t *__restrict__ counters,
                                         NeeCtx &nee)
{
    const uint32_t MB = p.max_bounces;
    if (r.bounce >= MB) return RAY_DONE;
    if (r.hit_vid == 0u) {   // miss: sky * throughput into the indirect channel
        const f3 sky = sky_color(p, r.dx, r.dy, r.dz);
        r.ilx = r.ilx + sky.x * r.tx;
        r.ily = r.ily + sky.y * r.ty;
        r.ilz = r.ilz + sky.z * r.tz;
        return RAY_DEPOSITS + (int)DEP_MISS;
    }
    const uint32_t mi = r.hit_material < MATERIAL_SLOTS ? r.hit_material : MATERIAL_SLOTS - 1u;
    const Material &m = mats[mi];
    const uint32_t hx = (uint32_t)r.ox, hy = (uint32_t)r.oy, hz = (uint32_t)r.oz;
    const uint32_t hsize = (uint32_t)r.pad0;
    const f3 rd{r.dx, r.dy, r.dz};
    const f3 nrm = lookup_normal(nbuf, p.hash_width, r.hit_vid, rd);
    const f3 albedo{m.color[0], m.color[1], m.color[2]};
    if (m.emissive) {   // shade.comp:576-590 (registration: k_em_register)
        const f3 em = vmin3(mul(muls(albedo, m.emissive_intensity), f3{r.tx, r.ty, r.tz}), f3{20, 20, 20});
        if (r.bounce == 0u) { r.dlx += em.x; r.dly += em.y; r.dlz += em.z; }
        else { r.ilx += em.x; r.ily += em.y; r.ilz += em.z; }
    }
    // next-event estimation (shade.comp:631-836)
    const uint32_t count = ANY_BOUNCE && iter == 0u ? registry_seen(counters, em_reg, t) : counters[VE_CNT_EMISSIVES];
    if (!m.emissive && r.bounce < 2u && count > 0u) {
        const uint32_t n_lights = count < EMISSIVE_CAPACITY ? count : EMISSIVE_CAPACITY;
        if (ANY_BOUNCE && p.restir_enabled && r.bounce == 0u)
            nee_restir(p, nee, r, m, nrm, hx, hy, hz, hsize, n_lights, count);
        else
            nee_ris(p, nee, r, m, nrm, hx, hy, hz, n_lights);
    }
    if (r.bounce + 1u >= MB) return RAY_DEPOSITS + (int)DEP_LAST;   // last allowed bounce
    const bool use_ggx = m.metallic * (1.0f - m.roughness) > 0.5f;
    uint32_t rng = r.rng;
    f3 sdir, bt;
    if (use_ggx) {
        const f3 Vv{-rd.x, -rd.y, -rd.z};
        const f3 H = sample_ggx_vndf(nrm, Vv, m.roughness, rng);
        sdir = reflect3(rd, H);
        if (dot(sdir, nrm) <= 0.0f) {
            sdir = normalize(add(nrm, random_direction(rng)));
            bt = evaluate_brdf(rd, sdir, nrm, albedo, m.roughness, m.metallic, m.specular);
        } else {
            const f3 F0 = add(muls(f3{0.04f, 0.04f, 0.04f}, 1.0f - m.metallic), muls(albedo, m.metallic));
            bt = ggx_throughput(nrm, Vv, sdir, H, m.roughness, F0);
        }
    } else {
        sdir = normalize(add(nrm, random_direction(rng)));
        bt = evaluate_brdf(rd, sdir, nrm, albedo, m.roughness, m.metallic, m.specular);
    }
    r.rng = rng;
    f3 ntp = mul(f3{r.tx, r.ty, r.tz}, bt);
    bool dies = false;
    if (r.bounce + 1u >= RR_START_BOUNCE) {   // Russian roulette (shade.comp:879-889)
        const float pr = gclamp(gmax(ntp.x, gmax(ntp.y, ntp.z)), 0.05f, 0.95f);
        const float xi = rand_lds(r.rng);
        if (xi > pr) dies = true;
        else ntp = divs(ntp, pr);
    }
    if (dies) return RAY_DEPOSITS + (int)DEP_RR;
    // survivor (shade.comp:892-908)
    r.ox = ((float)hx + 0.5f) + nrm.x;
    r.oy = ((float)hy + 0.5f) + nrm.y;
    r.oz = ((float)hz + 0.5f) + nrm.z;
    r.dx = sdir.x; r.dy = sdir.y; r.dz = sdir.z;
    r.tx = ntp.x; r.ty = ntp.y; r.tz = ntp.z;
    r.bounce = r.bounce + 1u;
    r.hit_vid = 0u;
    r.hit_material = 0u;
    return RAY_SURVIVES;
}

// Appends a deposit record for each lane with `dep` (the wave's records
// take consecutive slots: one atomic per wave).
__device__ __forceinline__ void record_deposit(bool dep, const Ray &r, uint32_t li, uint32_t iter, int outcome,
                                               uint32_t *__restrict__ counters, Deferred *__restrict__ deferred)
{
    const uint64_t m = __ballot(dep);
    if (!m) return;
    const uint32_t lane = threadIdx.x & 63u;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(&counters[VE_CNT_DEFERRED], (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    const uint32_t slot = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (dep && slot < SHADE_BUDGET_MAX)   // one deposit per ray and frame: always in range
        deferred[slot] = Deferred{r.primary_vid, li | (iter << 18) | ((uint32_t)(outcome - RAY_DEPOSITS) << 22),
                                  {r.dlx, r.dly, r.dlz}, {r.ilx, r.ily, r.ilz}};
}

template <bool MULTI>
__global__ __launch_bounds__(256) void k_shade(FrameParams p, const uint32_t *__restrict__ words, uint32_t iter,
                                               const Ray *__restrict__ rays_in, const uint32_t *__restrict__ live_in,
                                               const uint32_t *__restrict__ li_in, Ray *__restrict__ rays_out,
                                               uint32_t *__restrict__ live_out, uint32_t *__restrict__ li_out,
                                               Deferred *__restrict__ deferred, const uint32_t *__restrict__ nbuf,
                                               uint32_t *__restrict__ rbuf, const Material *__restrict__ mats,
                                               const EmissiveVoxel *__restrict__ emissives,
                                               const uint32_t *__restrict__ em_reg, uint32_t *__restrict__ counters,
                                               unsigned long long *__restrict__ totals,
                                               MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        p = global_params(V.p);
        words = gptr(V.words);
        iter = 0u;
        rays_in = gptr(V.b.rays);
        live_in = gptr(V.b.counters) + VE_CNT_LIVE;
        li_in = gptr(V.b.ray_li[0]);
        rays_out = gptr(V.b.rays_next);
        live_out = gptr(V.b.counters) + VE_CNT_LIVE_B;
        li_out = gptr(V.b.ray_li[1]);
        deferred = gptr(V.b.deferred);
        nbuf = gptr(V.b.nbuf);
        rbuf = gptr(V.b.rbuf);
        mats = gptr(V.b.materials);
        emissives = gptr(V.b.emissives);
        em_reg = gptr(V.b.em_reg);
        counters = gptr(V.b.counters);
        totals = gptr(V.b.ray_totals);
    }
    __shared__ uint32_t path_cache[VE_MAX_DEPTH * 256];
    NeeCtx nee{words, emissives, rbuf, path_cache + threadIdx.x, 0u};
    const uint32_t live = *live_in;
    const uint32_t t = vbx * blockDim.x + threadIdx.x;
    bool survive = false;
    int outcome = RAY_DONE;
    uint32_t li = 0u;
    Ray r;
    if (t < live) {
        r = rays_in[t];
        li = li_in[t];
        outcome = shade_ray<true>(p, iter, t, r, nbuf, mats, em_reg, counters, nee);
        survive = outcome == RAY_SURVIVES;
    }
    record_deposit(outcome >= RAY_DEPOSITS, r, li, iter, outcome, counters, deferred);
    // shadow rays of the wave: one atomic each for the frame counter and the
    // cumulative total
    unsigned long long sr = nee.shadow_rays;
    for (int off = 32; off > 0; off >>= 1) sr += __shfl_xor(sr, off, 64);
    if ((threadIdx.x & 63u) == 0u && sr) {
        atomicAdd(&counters[VE_CNT_SHADOW_RAYS], (uint32_t)sr);
        atomicAdd(&totals[2], sr);
    }
    // stream compaction of survivors: wave ballot + one atomic per wave
    const uint64_t mask = __ballot(survive);
    if (mask) {
        const uint32_t lane = threadIdx.x & 63u;
        const uint32_t before = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        const int leader = __ffsll((unsigned long long)mask) - 1;
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(live_out, (uint32_t)__popcll(mask));
        base = __shfl(base, leader, 64);
        if (survive) {
            rays_out[base + before] = r;
            li_out[base + before] = li;
        }
    }
}

// Bounce iterations first..iters-1 (renderer.cpp:515-602 after the first)
// fused per ray: each lane takes one ray entering iteration `first` through
// trace and shade until it ends, so the phase costs about the longest path
// rather than the sum over iterations of each iteration's longest trace.
// With a grid smaller than the ray list (bounce_blocks), a lane whose ray
// has ended takes the next ray from a queue at the iteration boundary, so a
// wave keeps its lanes busy while its longest path runs.
// Shading in these iterations reads nothing a deposit writes (normals,
// materials, the emissive registry; reservoirs are written at bounce 0
// only), so deposits can wait: each is appended to `deferred` with its
// iteration, and after the kernel they are grouped by lBuffer bucket and
// applied bucket by bucket in LDS, iteration by iteration -- the order the
// reference's one dispatch per iteration gives.  Traversals then run without
// a claim and up to ten adds on global memory per deposit beside them (the
// kernel is kept this lean: extra bookkeeping here has cost the traversal
// loop 20% through its register allocation, and placing each deposit in its
// lBuffer bucket from here 27 VGPRs and a wave per SIMD).  It is held to 5
// waves per SIMD (96 VGPRs with 5 spilled; 109 VGPRs and 4 waves otherwise).
// Ray counts: every ray entering an iteration counts as traced, as in the
// per-iteration trace dispatch; each lane counts its own, and each wave adds
// its sums to one of BOUNCE_COPIES copies at the end (k_dep_prep folds
// them), since per-iteration atomics on one counter from every wave cost the
// traversals a quarter of their time.
template <bool STATS, bool REFILL, bool MULTI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_bounce(const uint32_t *__restrict__ words, FrameParams p,
                                                const Ray *__restrict__ rays_in, const uint32_t *__restrict__ live_ptr,
                                                uint32_t first, uint32_t iters, uint32_t *__restrict__ counters,
                                                unsigned long long *__restrict__ stats,
                                                unsigned long long *__restrict__ totals,
                                                const Material *__restrict__ mats, const uint32_t *__restrict__ nbuf,
                                                uint32_t *__restrict__ rbuf, const EmissiveVoxel *__restrict__ emissives,
                                                const uint32_t *__restrict__ em_reg,
                                                const uint32_t *__restrict__ em_claim, uint32_t *__restrict__ em_bits,
                                                const uint32_t *__restrict__ li_in,
                                                Deferred *__restrict__ deferred, uint32_t *__restrict__ bounce_part,
                                                                                        MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        words = gptr(V.words);
        p = global_params(V.p);
        rays_in = gptr(V.b.rays_next);
        live_ptr = gptr(V.b.counters) + VE_CNT_LIVE_B;
        first = 1u;
        iters = V.iters;
        counters = gptr(V.b.counters);
        stats = gptr(V.b.node_stats);
        totals = gptr(V.b.ray_totals);
        mats = gptr(V.b.materials);
        nbuf = gptr(V.b.nbuf);
        rbuf = gptr(V.b.rbuf);
        emissives = gptr(V.b.emissives);
        em_reg = gptr(V.b.em_reg);
        em_claim = gptr(V.b.emissive_claim);
        em_bits = gptr(V.b.em_bits);
        li_in = gptr(V.b.ray_li[1]);
        deferred = gptr(V.b.deferred);
        bounce_part = gptr(V.b.bounce_part);
    }
    const uint32_t live = *live_ptr;
    const uint32_t tid = vbx * blockDim.x + threadIdx.x, threads = vbn * blockDim.x;
    if (vbx * blockDim.x >= live) return;   // whole block past the live rays
    __shared__ uint32_t path_cache[VE_MAX_DEPTH * 256];
    NeeCtx nee{words, emissives, rbuf, path_cache + threadIdx.x, 0u};
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t checks_sum = 0, reads = 0, entered = 0, reached = 0;
    // rays [0, threads) start on their threads; with REFILL the rest are
    // handed out from the queue to lanes whose ray has ended (no ray's result
    // depends on the lane or the order: deposits are ordered afterwards by
    // their keys)
    uint32_t *queue = bounce_part + 2u * BOUNCE_COPIES;
    bool drained = !REFILL || threads >= live;
    uint32_t t = tid, iter_l = first;
    bool active = t < live;
    Ray r;
    uint32_t li = 0u;
    if (active) {
        r = rays_in[t];
        li = li_in[t];
    }
    for (uint32_t step = 0;; step++) {
        const uint32_t iter = REFILL ? iter_l : first + step;   // without REFILL every lane is at one iteration
        if (!REFILL && iter >= iters) break;
        if (REFILL && !drained) {
            const uint64_t idle = __ballot(!active);
            if (idle) {
                const int leader = __ffsll((unsigned long long)idle) - 1;
                uint32_t base = 0;
                if ((int)lane == leader) base = atomicAdd(queue, (uint32_t)__popcll(idle));
                base = threads + (uint32_t)__shfl((int)base, leader, 64);
                if (!active) {
                    const uint32_t q = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                    if (q < live) {
                        t = q;
                        iter_l = first;
                        r = rays_in[q];
                        li = li_in[q];
                        active = true;
                    }
                }
                drained = base + (uint32_t)__popcll(idle) >= live;
            }
        }
        if (__ballot(active) == 0ull) break;
        if (!active) continue;
        const uint32_t it = REFILL ? iter_l : iter;
        entered++;
        reached = max(reached, it + 1u);
        uint32_t checks = 0;
        trace_ray<STATS>(words, p, it, t, r, mats, em_claim, em_bits, path_cache + threadIdx.x, checks, reads);
        checks_sum += checks;
        const int outcome = shade_ray<false>(p, it, t, r, nbuf, mats, em_reg, counters, nee);
        if (outcome >= RAY_DEPOSITS) {   // at most one deposit per ray: the slot is always in range
            const uint32_t slot = atomicAdd(&counters[VE_CNT_DEFERRED], 1u);
            if (slot < SHADE_BUDGET_MAX)
                deferred[slot] = Deferred{r.primary_vid,
                                          li | (it << 18) | ((uint32_t)(outcome - RAY_DEPOSITS) << 22),
                                          {r.dlx, r.dly, r.dlz}, {r.ilx, r.ily, r.ilz}};
        }
        active = outcome == RAY_SURVIVES;
        if (REFILL) active = active && ++iter_l < iters;
    }
    for (int off = 32; off > 0; off >>= 1) {
        entered += __shfl_xor(entered, off, 64);
        reached = max(reached, (uint32_t)__shfl_xor((int)reached, off, 64));
    }
    if (lane == 0u && entered) {
        uint32_t *part = bounce_part + 2u * ((tid >> 6) % BOUNCE_COPIES);
        atomicAdd(&part[0], entered);
        atomicMax(&part[1], reached);
    }
    unsigned long long sr = nee.shadow_rays;
    for (int off = 32; off > 0; off >>= 1) sr += __shfl_xor(sr, off, 64);
    if (lane == 0u && sr) {
        atomicAdd(&counters[VE_CNT_SHADOW_RAYS], (uint32_t)sr);
        atomicAdd(&totals[2], sr);
    }
    if (STATS) {
        unsigned long long a = reads, c = checks_sum;
        for (int off = 32; off > 0; off >>= 1) {
            a += __shfl_xor(a, off, 64);
            c += __shfl_xor(c, off, 64);
        }
        if (lane == 0u) {   // the fused kernel's share, kept apart from k_trace's
            atomicAdd(&stats[4], a);
            atomicAdd(&stats[5], c);
        }
    }
}

// ---------------------------------------------------------------------------
// The frame's deposits in the reference's order (AccumPass)
// ---------------------------------------------------------------------------
// The reference deposits from its shade dispatches: chunk by chunk of 49152
// shade-list entries (renderer.cpp:476-606), bounce iteration by iteration
// within a chunk, each dispatch over the window of that chunk's rays still
// alive, in SIMD-8 vectors, each vector's lanes in lockstep through its three
// call sites (misses, last bounce, Russian roulette).  A window keeps the
// shade-list order (survivors are appended in invocation order), so a ray's
// place in the window of iteration b is the number of rays of its chunk
// before it in the list still alive at b.  Every deposit of the frame is
// recorded with its ray's list index, iteration and call site; after the
// bounces:
//   k_dep_prep    records each ray's last iteration and lists each deposit
//                 under its lBuffer bucket (and folds the fused pass's ray
//                 counts);
//   k_dep_wpos    window positions, one block per 1024 list entries;
//   k_dep_replay  one wave per bucket sorts its deposits by order key
//                 {chunk, iteration, vector, call site, lane} and replays
//                 depositSampleDual with lane s holding slot s: a vector's
//                 call site with one deposit here as one invocation, several
//                 in lockstep (deposit_lockstep, on an LDS copy);
//   k_dep_done    clears the frame's record and overflow counts.
struct DepLists {
    uint32_t cap;          // deposits per bucket in place
    uint32_t *cnt;         // per bucket (zero between frames)
    uint32_t *slots;       // bucket-major record indices, cap each
    uint2 *over;           // {bucket, record} of the rest
    uint32_t *over_n;      // [0] overflow entries, [1] spill words handed out
    uint64_t *spill;       // sort keys of buckets that overflowed
    uint32_t cap_total;    // over / spill capacity (SHADE_BUDGET_MAX)
};

__host__ __device__ inline DepLists dep_lists(const DeviceBuffers &b)
{
    return DepLists{DEP_CAP, b.dep_cnt, b.dep_slots, b.dep_over, b.dep_over_n, b.dep_spill, SHADE_BUDGET_MAX};
}
__device__ __forceinline__ DepLists dep_lists_global(const DeviceBuffers &b)
{
    return DepLists{DEP_CAP, gptr(b.dep_cnt), gptr(b.dep_slots), gptr(b.dep_over), gptr(b.dep_over_n),
                    gptr(b.dep_spill), SHADE_BUDGET_MAX};
}

template <bool MULTI>
__global__ __launch_bounds__(256) void k_dep_prep(FrameParams p, const Deferred *__restrict__ deferred,
                                                  uint32_t *__restrict__ counters, uint32_t *__restrict__ bounce_part,
                                                  unsigned long long *__restrict__ totals, uint32_t *__restrict__ death,
                                                  DepLists dl,
                                                  MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        p = global_params(V.p);
        deferred = gptr(V.b.deferred);
        counters = gptr(V.b.counters);
        bounce_part = gptr(V.b.bounce_part);
        totals = gptr(V.b.ray_totals);
        death = gptr(V.b.dep_death);
        dl = dep_lists_global(V.b);
    }
    if (vbx == 0 && threadIdx.x == 0) {   // the fused bounce pass's per-wave ray counts
        uint32_t rays = 0, reached = 0;
        for (uint32_t c = 0; c < BOUNCE_COPIES; c++) {
            rays += bounce_part[2u * c];
            reached = max(reached, bounce_part[2u * c + 1u]);
            bounce_part[2u * c] = bounce_part[2u * c + 1u] = 0u;
        }
        bounce_part[2u * BOUNCE_COPIES] = 0u;   // the fused pass's ray queue
        counters[VE_CNT_RAYS_TRACED] += rays;
        totals[1] += rays;
        counters[VE_CNT_BOUNCES_RUN] = max(counters[VE_CNT_BOUNCES_RUN], reached);
    }
    const uint32_t n = min(counters[VE_CNT_DEFERRED], SHADE_BUDGET_MAX);
    for (uint32_t e = vbx * blockDim.x + threadIdx.x; e < n; e += vbn * blockDim.x) {
        const Deferred d = deferred[e];
        death[d.meta & 0x3FFFFu] = (d.meta >> 18) & 0xFu;
        if (d.vid == 0u) continue;
        const uint32_t h = lbuf_row_of(p.hash_width, d.vid);
        const uint32_t at = atomicAdd(&dl.cnt[h], 1u);
        if (at < dl.cap) dl.slots[(size_t)h * dl.cap + at] = e;
        else dl.over[atomicAdd(&dl.over_n[0], 1u)] = make_uint2(h, e);   // at most one entry per record
    }
}

// Block k: list entries [1024k, 1024k + 1024) of chunk c = k / 48: for each
// entry whose ray deposited at iteration b, the number of the chunk's entries
// before it whose rays were alive at b (deposited at b or later).  The
// chunk's earlier entries are counted by every block (at most 47 x 1024
// loads a thread-block, all in flight), the block's own with ballots.
constexpr uint32_t WPOS_BLOCKS = RAY_CHUNK / 1024u;
template <bool MULTI>
__global__ __launch_bounds__(1024) void k_dep_wpos(const uint32_t *__restrict__ counters,
                                                   const uint32_t *__restrict__ death, uint32_t *__restrict__ wpos,
                                                   uint32_t iters,
                                                   MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        counters = gptr(V.b.counters);
        death = gptr(V.b.dep_death);
        wpos = gptr(V.b.dep_wpos);
        iters = V.iters;
    }
    __shared__ uint32_t wsum[16][MAX_BOUNCE_ITERS];
    __shared__ uint32_t before_blk[MAX_BOUNCE_ITERS];
    const uint32_t ns = counters[VE_CNT_SHADE], t = threadIdx.x, lane = t & 63u, w = t >> 6;
    const uint32_t c = vbx / WPOS_BLOCKS, k = vbx % WPOS_BLOCKS;
    const uint32_t lo = c * RAY_CHUNK, i = lo + k * 1024u + t;
    if (lo + k * 1024u >= ns) return;
    const uint64_t below = (1ull << lane) - 1ull;
    // entries of the chunk's earlier blocks alive at each iteration
    uint32_t acc[MAX_BOUNCE_ITERS] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t q = 0; q < k; q++) {
        const uint32_t d = death[lo + q * 1024u + t];   // earlier blocks lie below ns
#pragma unroll
        for (uint32_t b = 0; b < MAX_BOUNCE_ITERS; b++) acc[b] += (d != NONE && d >= b) ? 1u : 0u;
    }
#pragma unroll
    for (uint32_t b = 0; b < MAX_BOUNCE_ITERS; b++) {
        uint32_t v = acc[b];
        for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, 64);
        if (lane == 0u) wsum[w][b] = v;
    }
    __syncthreads();
    if (t < MAX_BOUNCE_ITERS) {
        uint32_t v = 0;
        for (uint32_t q = 0; q < 16; q++) v += wsum[q][t];
        before_blk[t] = v;
    }
    __syncthreads();
    const bool here = i < ns;
    const uint32_t d = here ? death[i] : NONE;
    uint32_t mine_before = 0;
#pragma unroll
    for (uint32_t b = 0; b < MAX_BOUNCE_ITERS; b++) {
        const uint64_t alive = __ballot(here && d != NONE && d >= b);
        if (lane == 0u) wsum[w][b] = (uint32_t)__popcll(alive);
        if (d == b) mine_before = (uint32_t)__popcll(alive & below);
    }
    __syncthreads();
    if (here && d < iters) {
        uint32_t pos = before_blk[d] + mine_before;
        for (uint32_t q = 0; q < w; q++) pos += wsum[q][d];
        wpos[i] = pos;
    }
}

// Slot state of an lBuffer bucket copy (device layout: owner[32] | time[32] |
// payload[32][10], payload = dCnt dR dG dB dL2 iCnt iR iG iB iL2).
__device__ __forceinline__ void lb_reset(uint32_t *bk, uint32_t s, uint32_t vid, uint4 d, uint4 e, uint32_t time)
{   // resetLSlotDual (shade.comp:108-126)
    uint32_t *pl = bk + LB_PAY + s * LB_PAYW;
    bk[s] = vid;
    pl[0] = 1u; pl[1] = d.x; pl[2] = d.y; pl[3] = d.z; pl[4] = d.w;
    pl[5] = 1u; pl[6] = e.x; pl[7] = e.y; pl[8] = e.z; pl[9] = e.w;
    bk[LB_TIME + s] = time;
}

__device__ __forceinline__ void lb_firefly(const uint32_t *pl, uint32_t cf, uint4 &v, float k, float floor_)
{   // shade.comp:201-234 for one channel, on the slot's current sums
    const uint32_t c = pl[cf];
    if (c <= 4u) return;
    const float meanL = (float)(pl[cf + 1] + pl[cf + 2] + pl[cf + 3]) / (float)(3u * c);
    const float sampleL = (float)(v.x + v.y + v.z) / 3.0f;
    const float thresh = gmax(meanL * k, floor_);
    if (sampleL > thresh) {
        const float scale = thresh / sampleL;
        v.x = (uint32_t)((float)v.x * scale);
        v.y = (uint32_t)((float)v.y * scale);
        v.z = (uint32_t)((float)v.z * scale);
        const uint32_t L = (v.x + v.y + v.z) / 3u;
        v.w = L * L;
    }
}

__device__ __forceinline__ void lb_add(uint32_t *pl, uint4 d, uint4 e)
{
    pl[0] += 1u; pl[1] += d.x; pl[2] += d.y; pl[3] += d.z; pl[4] += d.w;
    pl[5] += 1u; pl[6] += e.x; pl[7] += e.y; pl[8] += e.z; pl[9] += e.w;
}

// depositSampleDual for k (2..8) lanes of one vector at one call site, all
// into this bucket, in llvmpipe's lockstep order: per slot iteration every
// active lane's compare-and-swap (lane order), then every lane's owner and
// time loads; lanes whose id matches deposit (all halving loads before the
// halving stores, all firefly reads before any add) and leave; the others
// fold the loaded time into their eviction candidate.  Lanes that never
// match reset their candidate after the loop, in lane order.  One thread.
// The lanes' records sit in the bucket's LDS copy after its slots: voxel
// ids, deposits (direct, indirect; firefly-clamped in place) and eviction
// candidates, so that this one-thread loop keeps no arrays of its own (they
// would live in scratch memory and hold the kernel's registers).
constexpr uint32_t LK_VID = LB_WORDS, LK_D = LB_WORDS + 8u, LK_E = LB_WORDS + 40u, LK_CT = LB_WORDS + 72u,
                   LK_CI = LB_WORDS + 80u, LK_WORDS = LB_WORDS + 88u;

__device__ __forceinline__ uint4 lk_load(const uint32_t *bk, uint32_t at)
{
    return make_uint4(bk[at], bk[at + 1u], bk[at + 2u], bk[at + 3u]);
}

__device__ void deposit_lockstep(const FrameParams &p, uint32_t *bk, uint32_t k)
{
    uint32_t active = (1u << k) - 1u;
    for (uint32_t l = 0; l < k; l++) {
        bk[LK_CT + l] = NONE;
        bk[LK_CI + l] = 0u;
    }
    for (uint32_t i = 0; i < LBUFFER_SLOTS && active; i++) {
        for (uint32_t l = 0; l < k; l++)
            if (((active >> l) & 1u) && bk[i] == 0u) bk[i] = bk[LK_VID + l];
        const uint32_t id = bk[i], st = bk[LB_TIME + i];   // every active lane loads the same slot
        uint32_t match = 0u;
        for (uint32_t l = 0; l < k; l++) {
            if (!((active >> l) & 1u)) continue;
            if (id == bk[LK_VID + l]) match |= 1u << l;
            else if (st < bk[LK_CT + l]) {
                bk[LK_CT + l] = st;
                bk[LK_CI + l] = i;
            }
        }
        uint32_t *pl = bk + LB_PAY + i * LB_PAYW;
        if (match && !p.disable_stale_reset && st != 0u && st + p.update_time < fp_time(p)) {   // shade.comp:183-186
            for (uint32_t l = 0; l < k; l++)
                if ((match >> l) & 1u)
                    lb_reset(bk, i, bk[LK_VID + l], lk_load(bk, LK_D + 4u * l), lk_load(bk, LK_E + 4u * l), fp_time(p));
            active &= ~match;
            match = 0u;
        }
        for (uint32_t ch = 0; ch < 2; ch++) {   // halving at the caps (shade.comp:188-194)
            const uint32_t f0 = ch ? 5u : 0u, cap = ch ? p.cap_indirect : p.cap_direct;
            if (match && pl[f0] >= cap)   // every halving lane loads the same values before any stores
                for (uint32_t f = 0; f < 5; f++) pl[f0 + f] >>= 1;
        }
        for (uint32_t ch = 0; ch < 2; ch++)   // all firefly reads before any add
            for (uint32_t l = 0; l < k; l++)
                if ((match >> l) & 1u) {
                    const uint32_t at = (ch ? LK_E : LK_D) + 4u * l;
                    uint4 v = lk_load(bk, at);
                    lb_firefly(pl, ch ? 5u : 0u, v, p.firefly_k, p.firefly_floor);
                    bk[at] = v.x;
                    bk[at + 1u] = v.y;
                    bk[at + 2u] = v.z;
                    bk[at + 3u] = v.w;
                }
        for (uint32_t l = 0; l < k; l++)
            if ((match >> l) & 1u) {
                lb_add(pl, lk_load(bk, LK_D + 4u * l), lk_load(bk, LK_E + 4u * l));
                bk[LB_TIME + i] = fp_time(p);
            }
        active &= ~match;
    }
    for (uint32_t l = 0; l < k; l++)   // bucket full: reset the candidate (shade.comp:254-256)
        if (((active >> l) & 1u) && bk[LK_CT + l] != NONE)
            lb_reset(bk, bk[LK_CI + l], bk[LK_VID + l], lk_load(bk, LK_D + 4u * l), lk_load(bk, LK_E + 4u * l),
                     fp_time(p));
}

// Lane-held slot of an lBuffer bucket: lane s (< 32) holds slot s's direct
// channel (payload words 0-4), lane s + 32 its indirect channel (words 5-9);
// both keep the slot's owner and time stamp.  depositSampleDual's steps on
// one channel (halving at its cap, the firefly clamp against its running
// mean, the add) read nothing of the other, so the two lanes run a deposit
// side by side.
struct LChan {
    uint32_t owner, tstamp, c[5];
};

__device__ __forceinline__ void lchan_reset(LChan &s, uint32_t vid, uint4 v, uint32_t time)
{   // resetLSlotDual (shade.comp:108-126), one channel
    s.owner = vid;
    s.c[0] = 1u; s.c[1] = v.x; s.c[2] = v.y; s.c[3] = v.z; s.c[4] = v.w;
    s.tstamp = time;
}

// One invocation's depositSampleDual into slot s (found or claimed), on the
// lane that holds the channel: v is the channel's sample, cap its cap.
__device__ __forceinline__ void lchan_deposit(const FrameParams &p, LChan &s, uint32_t vid, uint4 v, uint32_t cap)
{
    const uint32_t st = s.tstamp;
    if (s.owner == 0u) s.owner = vid;
    if (!p.disable_stale_reset && st != 0u && st + p.update_time < fp_time(p)) {   // shade.comp:183-186
        lchan_reset(s, vid, v, fp_time(p));
        return;
    }
    if (s.c[0] >= cap)   // shade.comp:188-194
        for (uint32_t f = 0; f < 5; f++) s.c[f] >>= 1;
    lb_firefly(s.c, 0, v, p.firefly_k, p.firefly_floor);
    s.c[0] += 1u; s.c[1] += v.x; s.c[2] += v.y; s.c[3] += v.z; s.c[4] += v.w;
    s.tstamp = fp_time(p);
}

// The order key of record `rec` (without its list index): chunk, bounce
// iteration, SIMD-8 vector, call site, lane.
__device__ __forceinline__ uint32_t dep_key_meta(uint32_t meta, const uint32_t *__restrict__ wpos)
{
    const uint32_t li = meta & 0x3FFFFu, it = (meta >> 18) & 0xFu, site = (meta >> 22) & 3u;
    const uint32_t w = wpos[li];
    return ((li / RAY_CHUNK) << 22) | (it << 18) | ((w >> 3) << 5) | (site << 3) | (w & 7u);
}

__device__ __forceinline__ uint32_t dep_key(const Deferred *__restrict__ deferred, const uint32_t *__restrict__ wpos,
                                            uint32_t rec)
{
    return dep_key_meta(deferred[rec].meta, wpos);
}

template <bool MULTI>
__global__ __launch_bounds__(64 * REPLAY_WAVES) void k_dep_replay(FrameParams p, uint32_t *__restrict__ lbuf,
                                                                  DepLists dl, const Deferred *__restrict__ deferred,
                                                                  const uint32_t *__restrict__ wpos,
                                                                  uint32_t *__restrict__ counters,
                                                                  MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        p = global_params(V.p);
        lbuf = gptr(V.b.lbuf);
        dl = dep_lists_global(V.b);
        deferred = gptr(V.b.deferred);
        wpos = gptr(V.b.dep_wpos);
        counters = gptr(V.b.counters);
    }
    // a bucket's list sorts in LDS only while it fits in place (n <= DEP_CAP):
    // beyond, its keys go to the global spill
    __shared__ uint64_t lds_keys_all[REPLAY_WAVES][DEP_CAP];
    __shared__ __align__(16) uint32_t bk_all[REPLAY_WAVES][LK_WORDS];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t h = vbx * REPLAY_WAVES + wv;
    if (h >= p.hash_width) return;
    const uint32_t n = dl.cnt[h];
    if (n == 0u) return;
    if (lane == 0u) dl.cnt[h] = 0u;   // the next frame's count starts from zero
    uint32_t *bk = bk_all[wv];
    const uint32_t *row = dl.slots + (size_t)h * dl.cap;
    static_assert(LBUFFER_SLOTS == 32 && LB_PAYW == 10, "one slot per lane pair, five payload words a channel");
    const bool sl = lane < LBUFFER_SLOTS;   // the lanes that hold the direct channels
    const uint32_t slot = lane & 31u, ch = lane >> 5;
    uint32_t *gb = lbuf + (size_t)h * LB_WORDS;
    LChan st;   // issued first: the loads overlap the records' gather and the sort
    st.owner = gb[slot];
    st.tstamp = gb[LB_TIME + slot];
#pragma unroll
    for (uint32_t f = 0; f < 5u; f++) st.c[f] = gb[LB_PAY + slot * LB_PAYW + 5u * ch + f];
    const uint32_t cap = ch ? p.cap_indirect : p.cap_direct;
    // records: row[i] in place, then this bucket's overflow entries (low word bit 31)
    SortedKeys sk{n, ~0ull, nullptr, false};
    // this lane's record of the window [w0, w0 + 64)
    uint32_t w0 = 0, r_grp = NONE, r_vid = 0u;
    uint4 r_d = make_uint4(0u, 0u, 0u, 0u), r_e = r_d;
    auto load_window = [&](uint32_t at) {
        w0 = at;
        const uint32_t q = at + lane;
        r_grp = NONE;
        if (q < n) {
            const uint64_t key = sk.at(q);   // n <= 64: q is this lane, which holds the q-th key
            const uint32_t lo = (uint32_t)key;
            const uint32_t rec = (lo & 0x80000000u) ? dl.over[lo & 0x7FFFFFFFu].y : row[lo];
            const Deferred dd = deferred[rec];
            r_grp = (uint32_t)(key >> 35);   // the order key without its lane bits
            r_vid = dd.vid;
            r_d = pack_rgb_l2(f3{dd.dl[0], dd.dl[1], dd.dl[2]});
            r_e = pack_rgb_l2(f3{dd.il[0], dd.il[1], dd.il[2]});
        }
    };
    if (n <= 64u) {
        // lane i gathers record row[i] whole; after the sort lane q takes the
        // q-th record's fields from the lane that gathered it (no second gather)
        Deferred dd{0u, 0u, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
        uint32_t kh = 0u;
        if (lane < n) {
            dd = deferred[row[lane]];
            kh = dep_key_meta(dd.meta, wpos);
        }
        sk.reg = wave_sort64(lane < n ? ((uint64_t)kh << 32) | lane : ~0ull);
        const int src = (int)((uint32_t)sk.reg & 63u);
        const uint32_t vid = __shfl(dd.vid, src, 64);
        const f3 dv{__shfl(dd.dl[0], src, 64), __shfl(dd.dl[1], src, 64), __shfl(dd.dl[2], src, 64)};
        const f3 ev{__shfl(dd.il[0], src, 64), __shfl(dd.il[1], src, 64), __shfl(dd.il[2], src, 64)};
        if (lane < n) {
            r_grp = (uint32_t)(sk.reg >> 35);   // the order key without its lane bits
            r_vid = vid;
            r_d = pack_rgb_l2(dv);
            r_e = pack_rgb_l2(ev);
        }
    } else {
        if (n <= dl.cap) {
            sk = sort_bucket(n, [&](uint32_t i) { return dep_key(deferred, wpos, row[i]); }, lds_keys_all[wv], nullptr);
        } else {
            uint32_t base = 0;
            if (lane == 0u) base = atomicAdd(&dl.over_n[1], n);
            base = __shfl(base, 0, 64);
            sk.keys = dl.spill + base;   // spill words for all overflowing buckets sum to at most the record count
            sk.glob = true;
            for (uint32_t i = lane; i < dl.cap; i += 64u)
                __hip_atomic_store(sk.keys + i, ((uint64_t)dep_key(deferred, wpos, row[i]) << 32) | i, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t L = dl.over_n[0];
            uint32_t put = dl.cap;
            for (uint32_t j0 = 0; j0 < L; j0 += 64u) {
                const uint32_t j = j0 + lane;
                const bool m = j < L && dl.over[j].x == h;
                const uint64_t mb = __ballot(m);
                const uint32_t at = put + (uint32_t)__popcll(mb & ((1ull << lane) - 1ull));
                if (m && at < n)
                    __hip_atomic_store(sk.keys + at, ((uint64_t)dep_key(deferred, wpos, dl.over[j].y) << 32) | 0x80000000u | j,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                put += (uint32_t)__popcll(mb);
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
            wave_sort_keys<true>(sk.keys, n);
        }
        load_window(0);
    }
    for (uint32_t k = 0; k < n;) {
        if (k + 8u > w0 + 64u && w0 + 64u < n) load_window(k);   // a whole group inside the window
        const uint32_t j = k - w0;
        const uint32_t g = (uint32_t)__builtin_amdgcn_readlane((int)r_grp, (int)j);
        uint32_t m = 1;
        while (m < 8u && k + m < n && (uint32_t)__builtin_amdgcn_readlane((int)r_grp, (int)(j + m)) == g) m++;
        if (m == 1u) {
            const uint32_t vid = (uint32_t)__builtin_amdgcn_readlane((int)r_vid, (int)j);
            const uint4 d = make_uint4((uint32_t)__builtin_amdgcn_readlane((int)r_d.x, (int)j),
                                       (uint32_t)__builtin_amdgcn_readlane((int)r_d.y, (int)j),
                                       (uint32_t)__builtin_amdgcn_readlane((int)r_d.z, (int)j),
                                       (uint32_t)__builtin_amdgcn_readlane((int)r_d.w, (int)j));
            const uint4 e = make_uint4((uint32_t)__builtin_amdgcn_readlane((int)r_e.x, (int)j),
                                       (uint32_t)__builtin_amdgcn_readlane((int)r_e.y, (int)j),
                                       (uint32_t)__builtin_amdgcn_readlane((int)r_e.z, (int)j),
                                       (uint32_t)__builtin_amdgcn_readlane((int)r_e.w, (int)j));
            const uint4 v = ch ? e : d;
            const uint64_t hit = __ballot(sl && (st.owner == vid || st.owner == 0u));
            if (hit) {
                if (slot == (uint32_t)__ffsll((unsigned long long)hit) - 1u) lchan_deposit(p, st, vid, v, cap);
            } else {   // every slot holds another voxel: reset the least recently stamped one
                const uint32_t best = wave_min_u32(sl ? st.tstamp : NONE);
                const uint64_t at = __ballot(sl && st.tstamp == best);
                if (best != NONE && slot == (uint32_t)__ffsll((unsigned long long)at) - 1u)
                    lchan_reset(st, vid, v, fp_time(p));
            }
        } else {   // several lanes of one vector here: lockstep, on an LDS copy
            if (sl) {
                bk[slot] = st.owner;
                bk[LB_TIME + slot] = st.tstamp;
            }
            for (uint32_t f = 0; f < 5u; f++) bk[LB_PAY + slot * LB_PAYW + 5u * ch + f] = st.c[f];
            if (lane >= j && lane < j + m) {   // the group's records, lane j + l into place l
                const uint32_t l = lane - j;
                bk[LK_VID + l] = r_vid;
                bk[LK_D + 4u * l] = r_d.x;
                bk[LK_D + 4u * l + 1u] = r_d.y;
                bk[LK_D + 4u * l + 2u] = r_d.z;
                bk[LK_D + 4u * l + 3u] = r_d.w;
                bk[LK_E + 4u * l] = r_e.x;
                bk[LK_E + 4u * l + 1u] = r_e.y;
                bk[LK_E + 4u * l + 2u] = r_e.z;
                bk[LK_E + 4u * l + 3u] = r_e.w;
            }
            wave_lds_sync();
            if (lane == 0u) deposit_lockstep(p, bk, m);
            wave_lds_sync();
            st.owner = bk[slot];
            st.tstamp = bk[LB_TIME + slot];
            for (uint32_t f = 0; f < 5u; f++) st.c[f] = bk[LB_PAY + slot * LB_PAYW + 5u * ch + f];
            wave_lds_sync();
        }
        k += m;
    }
    if (sl) {
        gb[slot] = st.owner;
        gb[LB_TIME + slot] = st.tstamp;
    }
#pragma unroll
    for (uint32_t f = 0; f < 5u; f++) gb[LB_PAY + slot * LB_PAYW + 5u * ch + f] = st.c[f];
}

__global__ void k_dep_done(uint32_t *__restrict__ counters, uint32_t *__restrict__ over_n, int32_t early_set)
{
    counters[VE_CNT_DEFERRED] = 0u;   // the frame's records are applied
    over_n[0] = over_n[1] = 0u;
    if (early_set) {   // the overflow that the frame's early schedule and normal passes reported
        uint32_t *side = counters + NUM_COUNTERS + 2u * (uint32_t)(early_set - 1);
        counters[VE_CNT_OVERFLOW] |= side[0] | side[1];
        side[0] = side[1] = 0u;
    }
}

// ---------------------------------------------------------------------------
// resolve.comp = AvgPass
// ---------------------------------------------------------------------------
struct Avg {
    f3 c;
    uint32_t count;
    float var;
};

__device__ Avg voxel_avg(const uint32_t *__restrict__ lbuf, uint32_t hw, uint32_t vid)   // resolve.comp:109-148
{
    Avg a{{0, 0, 0}, 0u, 0.0f};
    if (vid == 0u) return a;
    const uint32_t h = vid % (hw - 1u) + 1u;
    const uint32_t *bucket = lbuf + (size_t)h * LB_WORDS;
    uint32_t own;
    const uint32_t s = find_owner<false>(bucket, 0, LBUFFER_SLOTS, vid, own);
    if (s >= LBUFFER_SLOTS) return a;
    const uint32_t *pl = bucket + LB_PAY + s * LB_PAYW;
    const uint32_t dC = pl[0], iC = pl[5];
    a.count = dC + iC;
    if (a.count == 0u) return a;
    f3 da{0, 0, 0}, ia{0, 0, 0};
    if (dC > 0u) {
        const float inv = 1.0f / (float)dC;
        da = divs(muls(f3{(float)pl[1], (float)pl[2], (float)pl[3]}, inv), 255.0f);
    }
    if (iC > 0u) {
        const uint32_t ir = pl[6], ig = pl[7], ib = pl[8], il2 = pl[9];
        const float inv = 1.0f / (float)iC;
        ia = divs(muls(f3{(float)ir, (float)ig, (float)ib}, inv), 255.0f);
        const float meanL = (float)(ir + ig + ib) * inv / 3.0f;
        a.var = gmax((float)il2 * inv - meanL * meanL, 0.0f) / 65025.0f;
    }
    a.c = add(da, ia);
    return a;
}

// AvgPass part 1: the voxel average of every pixel's voxel id, once per pixel.
// lookupVoxelAvg is a pure function of the id, so the bilateral taps and the
// timeout gather of part 2 read these per-pixel values instead of probing the
// lBuffer hash again for each tap (up to 440 taps for a timed-out pixel).
// Lanes of a tile that see the same voxel share one lookup.
template <bool MULTI>
__global__ __launch_bounds__(256) void k_resolve_avg(FrameParams p, const uint4 *__restrict__ gbuf,
                                                     const uint32_t *__restrict__ lbuf, float4 *__restrict__ pix_avg,
                                                     uint32_t *__restrict__ pix_cnt, uint32_t *__restrict__ counters,
                                                     uint32_t *__restrict__ tcount, uint32_t *__restrict__ dep_over_n,
                                                     MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        p = global_params(V.p);
        gbuf = gptr(V.b.gbuf);
        lbuf = gptr(V.b.lbuf);
        pix_avg = gptr(V.b.pix_avg);
        pix_cnt = gptr(V.b.pix_cnt);
        counters = gptr(V.b.counters);
        tcount = gptr(V.b.timeout_tiles);
        dep_over_n = gptr(V.b.dep_over_n);
    }
    if (vbx == 0 && threadIdx.x == 0) {   // k_resolve appends after this pass
        counters[VE_CNT_TIMEOUTS] = 0u;
        counters[VE_CNT_TIMEOUT_TILES] = 0u;
        if (p.fold_small_passes && !p.early_set) {   // k_dep_done, folded: the frame's deposit records are applied
            counters[VE_CNT_DEFERRED] = 0u;
            dep_over_n[0] = dep_over_n[1] = 0u;
        }
    }
    {   // and counts timed-out pixels per 16x16 tile from zero
        const uint32_t n_t = ((p.W + RECON_TILE - 1u) / RECON_TILE) * ((p.H + RECON_TILE - 1u) / RECON_TILE);
        for (uint32_t i = vbx * blockDim.x + threadIdx.x; i < n_t; i += vbn * blockDim.x) tcount[i] = 0u;
    }
    uint32_t x = 0, y = 0;
    const bool valid = tile_pixel(vbx, vbn, p, x, y);
    const size_t pix = valid ? (size_t)y * p.W + x : 0;
    const uint32_t vid = valid ? gbuf[pix].w : 0u;
    // leader = lowest lane holding the same id; leaders look up in parallel
    const uint32_t lane = threadIdx.x & 63u;
    int leader = (int)lane;
    uint64_t pending = __ballot(vid != 0u);
    while (pending) {
        const int l = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lv = __shfl(vid, l, 64);
        const uint64_t same = __ballot(vid == lv);
        pending &= ~same;
        if (vid == lv) leader = l;
    }
    Avg a{{0, 0, 0}, 0u, 0.0f};
    if (vid != 0u && leader == (int)lane) a = voxel_avg(lbuf, p.hash_width, vid);
    Avg mine;
    mine.c.x = __shfl(a.c.x, leader, 64);
    mine.c.y = __shfl(a.c.y, leader, 64);
    mine.c.z = __shfl(a.c.z, leader, 64);
    mine.count = __shfl(a.count, leader, 64);
    mine.var = __shfl(a.var, leader, 64);
    if (valid) {
        pix_avg[pix] = make_float4(mine.c.x, mine.c.y, mine.c.z, mine.var);
        pix_cnt[pix] = mine.count;
    }
}

// AvgPass part 2 (resolve.comp:362-382): per-pixel colour of the selected view.
// Timed-out pixels (reconstructTimeout) are appended to `timeouts` and
// coloured by k_resolve_timeouts.
template <bool MULTI>
__global__ __launch_bounds__(256) void k_resolve(FrameParams p, const uint4 *__restrict__ gbuf,
                                                 const uint32_t *__restrict__ lbuf, const uint32_t *__restrict__ nbuf,
                                                 const Material *__restrict__ mats, const float4 *__restrict__ pix_avg,
                                                 const uint32_t *__restrict__ pix_cnt, float4 *__restrict__ out,
                                                 uint32_t *__restrict__ timeouts, uint32_t *__restrict__ counters,
                                                 uint32_t *__restrict__ tflag, uint32_t *__restrict__ tlist,
                                                 MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        p = global_params(V.p);
        gbuf = gptr(V.b.gbuf);
        lbuf = gptr(V.b.lbuf);
        nbuf = gptr(V.b.nbuf);
        mats = gptr(V.b.materials);
        pix_avg = gptr(V.b.pix_avg);
        pix_cnt = gptr(V.b.pix_cnt);
        out = gptr(V.b.output);
        timeouts = gptr(V.b.timeouts);
        counters = gptr(V.b.counters);
        tflag = gptr(V.b.timeout_tiles);
        tlist = gptr(V.b.timeout_list);
    }
    uint32_t x = 0, y = 0;
    const bool valid = tile_pixel(vbx, vbn, p, x, y);
    const size_t pix = valid ? (size_t)y * p.W + x : 0;
    const uint4 g = valid ? gbuf[pix] : make_uint4(0u, 0u, 0u, 0u);
    const uint32_t checks = g.x, mi = g.y, vid = g.w;
    const bool timeout = valid && p.render_type != 1 && vid == 0u && checks >= p.primary_cc;
    const uint64_t tmask = __ballot(timeout);
    if (tmask != 0ull) {
        const uint32_t lane = threadIdx.x & 63u;
        const int first = __ffsll((unsigned long long)tmask) - 1;
        uint32_t base = 0;
        if ((int)lane == first) base = atomicAdd(&counters[VE_CNT_TIMEOUTS], (uint32_t)__popcll(tmask));
        base = __shfl(base, first, 64);
        if (timeout) timeouts[base + (uint32_t)__popcll(tmask & ((1ull << lane) - 1ull))] = (uint32_t)pix;
        // the 16x16 tiles of the tiled gather that hold a timed-out pixel (a
        // wave's 8x8 tile lies in one of them), listed once each
        if ((int)lane == first && p.reconstruction_radius <= RECON_TILED_MAX) {
            const uint32_t t16 = (y / RECON_TILE) * ((p.W + RECON_TILE - 1u) / RECON_TILE) + x / RECON_TILE;
            if (atomicAdd(&tflag[t16], (uint32_t)__popcll(tmask)) == 0u)
                tlist[atomicAdd(&counters[VE_CNT_TIMEOUT_TILES], 1u)] = t16;
        }
    }
    if (!valid || timeout) return;
    f3 col;
    if (p.render_type == 1) {
        const float v = (float)checks / 65.0f;
        col = f3{v, v, v};
    } else if (vid == 0u) {   // texture(skybox, pixelDirection(pix)) (resolve.comp:170-180,227)
        const f3 d = primary_dir(fp_cam(p), x, y, p.W, p.H);
        col = sky_color(p, d.x, d.y, d.z);
    } else if (p.render_type == 2) {
        col = (mi == 0u || mi >= 127u) ? f3{0, 0, 0} : f3{mats[mi].color[0], mats[mi].color[1], mats[mi].color[2]};
    } else if (p.render_type == 3) {
        col = f3{0, 0, 0};   // resolve.comp:152-166: the first slot owned by vid decides
        const uint32_t h = vid % (p.hash_width - 1u) + 1u;
        const uint32_t *bucket = nbuf + (size_t)h * NB_WORDS;
        uint32_t own;
        const uint32_t s = find_owner<false>(bucket, 0, NBUFFER_SLOTS, vid, own);
        if (s < NBUFFER_SLOTS) {
            const uint32_t f = bucket[NB_NRM + s];
            if (f != 0u) col = add(muls(normalize(unpack_normal(f)), 0.5f), f3{0.5f, 0.5f, 0.5f});
        }
    } else if (p.render_type == 4) {
        const uint32_t hid = vid - 1u;
        col = f3{(float)((hid + 1u) % 255u) / 255.0f, (float)(((hid + 1u) / 255u) % 255u) / 255.0f,
                 (float)(((hid + 1u) / (255u * 255u)) % 255u) / 255.0f};
    } else if (p.render_type == 5) {
        const float var = indirect_variance(lbuf, p.hash_width, vid) / 65025.0f;   // resolve.comp:302-322
        const float sc = gclamp(var * 65025.0f / gmax(p.variance_norm, 1.0f), 0.0f, 1.0f);
        col = f3{sc, sc, sc};
    } else {   // DEFAULT (resolve.comp:224-261)
        const float4 ca = pix_avg[pix];
        const uint32_t cc = pix_cnt[pix];
        const f3 cavg{ca.x, ca.y, ca.z};
        if (cc >= TRUSTED_THRESHOLD && ca.w < 0.02f) {
            col = cavg;
        } else {
            f3 sum = muls(cavg, (float)cc);
            float wsum = (float)cc;
            const int OX[4] = {1, -1, 0, 0}, OY[4] = {0, 0, 1, -1};
            for (int k = 0; k < 4; k++) {
                const int qx = (int)x + OX[k], qy = (int)y + OY[k];
                if (qx < 0 || qy < 0 || qx >= (int)p.W || qy >= (int)p.H) continue;
                const size_t q = (size_t)qy * p.W + qx;
                const uint32_t nvid = gbuf[q].w;
                if (nvid == 0u || nvid == vid) continue;
                const uint32_t nc = pix_cnt[q];
                if (nc == 0u) continue;
                const float4 na = pix_avg[q];
                const f3 navg{na.x, na.y, na.z};
                const float dc = length3(sub(navg, cavg));
                const float wc = m_exp(-dc * 6.0f);
                const float wv = 1.0f / (1.0f + na.w * 32.0f);
                const float w = wc * wv * (float)min(nc, 32u);
                sum = add(sum, muls(navg, w));
                wsum += w;
            }
            col = wsum > 0.0001f ? divs(sum, wsum) : f3{0, 0, 0};
        }
    }
    out[pix] = make_float4(col.x, col.y, col.z, 1.0f);
}

__device__ __forceinline__ float wave_fsum(float v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// reconstructTimeout (resolve.comp:193-219) for the pixels k_resolve listed:
// one wave per timed-out pixel, the (2R+1)^2 - 1 taps spread over the lanes
// (tap t on lane t % 64) and the weighted sums reduced across the wave.  Each
// tap's weight and weighted colour are computed exactly as the reference does;
// only the order of the final float additions differs from its sequential
// loop (all terms are non-negative, so the sums agree to a few ulp).  Waves
// stride over the list and exit once it is exhausted.
// reconstructTimeout (resolve.comp:193-219) for radii up to RECON_TILED_MAX,
// on the 16x16 screen tiles holding RECON_DENSE or more timed-out pixels
// (k_resolve lists and counts them; sparser tiles go to the listed form):
// blocks stride over the listed tiles, and a dense one stages, for the tile
// and its R-pixel halo, each pixel's tap weight base min(count, 32) (0 where
// the tap is skipped: no voxel, timed out itself, no samples, or outside the
// screen) with its voxel average in LDS; each timed-out pixel then sums its
// (2R+1)^2 window over 16 lanes (partial sums reduced across them; the image
// is compared within 1e-4).  Only in frames with RECON_TILED_MIN timed-out
// pixels or more: below that, staging windows costs more than it saves.
__device__ __forceinline__ void resolve_timeout_tiles(FrameParams p, const uint4 *__restrict__ gbuf,
                                                               const float4 *__restrict__ pix_avg,
                                                               const uint32_t *__restrict__ pix_cnt,
                                                               uint32_t *__restrict__ tflag,
                                                               const uint32_t *__restrict__ tlist,
                                                               const uint32_t *__restrict__ counters,
                                                               float4 *__restrict__ out, uint32_t blk, uint32_t nblk)
{
    extern __shared__ float4 tap[];   // (16 + 2R)^2 taps
    __shared__ uint32_t todo[RECON_TILE * RECON_TILE];
    __shared__ uint32_t n_todo;
    if (counters[VE_CNT_TIMEOUTS] < RECON_TILED_MIN) return;   // the listed form did every pixel
    const uint32_t tw = (p.W + RECON_TILE - 1u) / RECON_TILE;
    const uint32_t n_tiles = counters[VE_CNT_TIMEOUT_TILES];
    const int R = min(max(p.reconstruction_radius, 0), (int)RECON_TILED_MAX);
    const int D = (int)RECON_TILE + 2 * R;
    for (uint32_t it = blk; it < n_tiles; it += nblk) {
        const uint32_t tile = tlist[it];
        const int tx0 = (int)((tile % tw) * RECON_TILE), ty0 = (int)((tile / tw) * RECON_TILE);
        if (tflag[tile] < RECON_DENSE) continue;   // sparse tiles: the listed form did them
        if (threadIdx.x == 0u) n_todo = 0u;
        __syncthreads();
        {   // the tile's timed-out pixels, packed onto the first threads
            const int px = tx0 + (int)(threadIdx.x % RECON_TILE), py = ty0 + (int)(threadIdx.x / RECON_TILE);
            if (px < (int)p.W && py < (int)p.H && p.render_type != 1) {
                const uint4 g = gbuf[(size_t)py * p.W + px];
                if (g.w == 0u && g.x >= p.primary_cc) todo[atomicAdd(&n_todo, 1u)] = threadIdx.x;
            }
        }
        __syncthreads();
        const uint32_t n = n_todo;
        {
            // dense tile: stage the window's taps once, sum each pixel's
            // window sequentially in the reference's order (dy, then dx)
            for (int i = (int)threadIdx.x; i < D * D; i += (int)blockDim.x) {
                const int qx = tx0 - R + i % D, qy = ty0 - R + i / D;
                float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (qx >= 0 && qy >= 0 && qx < (int)p.W && qy < (int)p.H) {
                    const size_t q = (size_t)qy * p.W + qx;
                    const uint4 nd = gbuf[q];
                    const uint32_t nc = pix_cnt[q];
                    if (nd.w != 0u && nd.x < p.primary_cc && nc != 0u) {
                        const float4 na = pix_avg[q];
                        v = make_float4(na.x, na.y, na.z, (float)min(nc, 32u));
                    }
                }
                tap[i] = v;
            }
            __syncthreads();
            // 16 lanes per timed-out pixel, each summing every 16th tap of
            // the window; the partial sums meet in a 16-lane reduction
            const uint32_t g = threadIdx.x / 16u, gl = threadIdx.x % 16u;
            const int K = 2 * R + 1, NT = K * K;
            for (uint32_t base = 0; base < n; base += blockDim.x / 16u) {
                const uint32_t k = base + g;
                float sr = 0.0f, sg = 0.0f, sb = 0.0f, ws = 0.0f;
                int lx = 0, ly = 0;
                if (k < n) {
                    const uint32_t me = todo[k];
                    lx = (int)(me % RECON_TILE);
                    ly = (int)(me / RECON_TILE);
                    for (int t = (int)gl; t < NT; t += 16) {
                        const int dx = t % K - R, dy = t / K - R;
                        if (dx == 0 && dy == 0) continue;
                        const float4 e = tap[(ly + R + dy) * D + (lx + R + dx)];
                        if (e.w == 0.0f) continue;
                        const float dist = (float)(dx * dx + dy * dy);
                        const float w = e.w / (1.0f + dist * 0.15f);
                        sr = sr + e.x * w;
                        sg = sg + e.y * w;
                        sb = sb + e.z * w;
                        ws = ws + w;
                    }
                }
                for (int off = 8; off > 0; off >>= 1) {
                    sr += __shfl_xor(sr, off, 16);
                    sg += __shfl_xor(sg, off, 16);
                    sb += __shfl_xor(sb, off, 16);
                    ws += __shfl_xor(ws, off, 16);
                }
                if (k < n && gl == 0u) {
                    const int px = tx0 + lx, py = ty0 + ly;
                    f3 col;
                    if (ws > 0.0001f) {
                        col = f3{sr / ws, sg / ws, sb / ws};
                    } else {   // resolve.comp:218
                        const f3 d = primary_dir(fp_cam(p), (uint32_t)px, (uint32_t)py, p.W, p.H);
                        col = sky_color(p, d.x, d.y, d.z);
                    }
                    out[(size_t)py * p.W + px] = make_float4(col.x, col.y, col.z, 1.0f);
                }
            }
        }
        __syncthreads();   // before the next tile reuses the LDS
    }
}

__device__ __forceinline__ void resolve_timeout_list(FrameParams p, const uint4 *__restrict__ gbuf,
                                                         const float4 *__restrict__ pix_avg,
                                                         const uint32_t *__restrict__ pix_cnt,
                                                         const uint32_t *__restrict__ timeouts,
                                                         const uint32_t *__restrict__ counters,
                                                         const uint32_t *__restrict__ tcount,
                                                         float4 *__restrict__ out, uint32_t blk, uint32_t nblk)
{
    const int R = p.reconstruction_radius < 0 ? 0 : p.reconstruction_radius;
    const uint32_t tw = (p.W + RECON_TILE - 1u) / RECON_TILE;
    const int D = 2 * R + 1, NT = D * D;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t n = counters[VE_CNT_TIMEOUTS];
    // in a frame with many timed-out pixels, those of dense tiles are left to
    // the tiled blocks
    const bool tiled = R <= RECON_TILED_MAX && n >= RECON_TILED_MIN;
    const uint32_t waves = nblk * (blockDim.x >> 6);
    for (uint32_t i = blk * (blockDim.x >> 6) + (threadIdx.x >> 6); i < n; i += waves) {
        const uint32_t pix = timeouts[i];
        const int px = (int)(pix % p.W), py = (int)(pix / p.W);
        if (tiled && tcount[(uint32_t)(py / (int)RECON_TILE) * tw + (uint32_t)px / RECON_TILE] >= RECON_DENSE) continue;
        float sr = 0.0f, sg = 0.0f, sb = 0.0f, ws = 0.0f;
        for (int t = (int)lane; t < NT; t += 64) {
            const int dx = t % D - R, dy = t / D - R;
            const int qx = px + dx, qy = py + dy;
            if ((dx == 0 && dy == 0) || qx < 0 || qy < 0 || qx >= (int)p.W || qy >= (int)p.H) continue;
            const size_t q = (size_t)qy * p.W + qx;
            const uint4 nd = gbuf[q];
            const uint32_t nc = pix_cnt[q];
            const float4 na = pix_avg[q];
            if (nd.w == 0u || nd.x >= p.primary_cc || nc == 0u) continue;
            const float dist = (float)(dx * dx + dy * dy);
            const float w = (float)min(nc, 32u) / (1.0f + dist * 0.15f);
            sr += na.x * w;
            sg += na.y * w;
            sb += na.z * w;
            ws += w;
        }
        sr = wave_fsum(sr);
        sg = wave_fsum(sg);
        sb = wave_fsum(sb);
        ws = wave_fsum(ws);
        if (lane == 0u) {
            f3 col;
            if (ws > 0.0001f) {
                col = f3{sr / ws, sg / ws, sb / ws};
            } else {   // resolve.comp:218
                const f3 d = primary_dir(fp_cam(p), (uint32_t)px, (uint32_t)py, p.W, p.H);
                col = sky_color(p, d.x, d.y, d.z);
            }
            out[pix] = make_float4(col.x, col.y, col.z, 1.0f);
        }
    }
}

// The timed-out pixels' reconstruction in one launch: blocks below
// `list_blocks` take the listed pixels one wave each, the rest the dense
// 16x16 tiles from LDS (with the tiled path on; the two roles write disjoint
// pixels and only read the per-tile counts).
template <bool MULTI>
__global__ __launch_bounds__(256) void k_resolve_timeouts(FrameParams p, const uint4 *__restrict__ gbuf,
                                                          const float4 *__restrict__ pix_avg,
                                                          const uint32_t *__restrict__ pix_cnt,
                                                          const uint32_t *__restrict__ timeouts,
                                                          uint32_t *__restrict__ tflag,
                                                          const uint32_t *__restrict__ tlist,
                                                          const uint32_t *__restrict__ counters,
                                                          float4 *__restrict__ out, uint32_t list_blocks,
                                                          MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        p = global_params(V.p);
        gbuf = gptr(V.b.gbuf);
        pix_avg = gptr(V.b.pix_avg);
        pix_cnt = gptr(V.b.pix_cnt);
        timeouts = gptr(V.b.timeouts);
        tflag = gptr(V.b.timeout_tiles);
        tlist = gptr(V.b.timeout_list);
        counters = gptr(V.b.counters);
        out = gptr(V.b.output);
        list_blocks = RESOLVE_TIMEOUT_BLOCKS;
    }
    if (vbx < list_blocks)
        resolve_timeout_list(p, gbuf, pix_avg, pix_cnt, timeouts, counters, tflag, out, vbx, list_blocks);
    else
        resolve_timeout_tiles(p, gbuf, pix_avg, pix_cnt, tflag, tlist, counters, out, vbx - list_blocks,
                              vbn - list_blocks);
}

// RenderPass (final.frag:11-19): ACES filmic curve per channel, clamped to
// [0,1], written as an RGBA8 unorm texel like the blit into an 8-bit default
// framebuffer (v * 255 rounded to nearest, ties to even), alpha 255.
__device__ __forceinline__ float aces(float x)
{
    return gclamp((x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f), 0.0f, 1.0f);
}

__global__ __launch_bounds__(256) void k_display(uint32_t n, const float4 *__restrict__ in, uchar4 *__restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 c = in[i];
    out[i] = make_uchar4((unsigned char)rintf(aces(c.x) * 255.0f), (unsigned char)rintf(aces(c.y) * 255.0f),
                         (unsigned char)rintf(aces(c.z) * 255.0f), 255);
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
static inline uint32_t cdiv(size_t a, uint32_t b) { return (uint32_t)((a + b - 1) / b); }

static uint32_t screen_blocks(const FrameParams &p)
{
    const uint32_t tiles = ((p.W + TILE - 1) / TILE) * ((p.H + TILE - 1) / TILE);
    uint32_t nb = (tiles + 3u) / 4u;
    nb = nb == 0 ? 1 : nb;
    const uint32_t unit = NUM_XCD * (p.tile_chunk ? p.tile_chunk : 1u);
    return (nb + unit - 1) / unit * unit;
}

void launch_fill(hipStream_t s, uint32_t *p, uint32_t v, size_t n)
{
    if (n) launch(k_fill_u32, dim3(cdiv(n, 256)), dim3(256), 0, s, p, v, n);
}

// The word a frame's schedule (which = 0) or normal (1) pass reports overflow
// into: the overflow counter, or, when those passes run ahead of the reset
// pass on the auxiliary streams (early_set), the buffer set's side word past
// the counters, which the frame's resolve folds into the counter.
static uint32_t *overflow_word(const FrameParams &p, DeviceBuffers &b, uint32_t which)
{
    if (!p.early_set) return b.counters + VE_CNT_OVERFLOW;
    return b.counters + NUM_COUNTERS + 2u * (uint32_t)(p.early_set - 1) + which;
}

void launch_normal(hipStream_t s, const uint32_t *words, const FrameParams &p, DeviceBuffers &b, const uint32_t *prev)
{
    const OrderLists ol = nrm_lists(b);
    uint32_t *ovf = overflow_word(p, b, 1u);
    launch(k_nrm_events<false>, dim3(screen_blocks(p)), dim3(256), 0, s, p, b.gbuf, b.gpos, ol, ovf, MultiTab{});
    launch(k_nrm_replay<false>, dim3(cdiv(p.hash_width, REPLAY_WAVES)), dim3(64 * REPLAY_WAVES), 0, s, p, b.nbuf,
           prev ? prev : b.nbuf, ol, b.ord_work, ovf, MultiTab{});
    launch(k_nrm_compute<false>, dim3(cdiv((size_t)p.hash_width * NBUFFER_SLOTS, 256)), dim3(256), 0, s, words, p, b.nbuf,
           b.ord_work, b.ord_over_n, MultiTab{});
}

// Per-frame clear of RankStats' per-frame half and the counters the frame
// accumulates into (renderer.cpp:343-375); fill kernels, not memsets, so the
// sequence stays capturable.
template <bool MULTI>
__global__ void k_frame_reset(FrameParams p, uint32_t *__restrict__ rank_stats, uint32_t *__restrict__ counters,
                              unsigned long long *__restrict__ totals,
                              MultiTab mt)
{
    uint32_t vbx = blockIdx.x, vbn = gridDim.x;
    if constexpr (MULTI) {
        const ViewRec &V = multi_view(mt, vbx, vbn);
        p = global_params(V.p);
        rank_stats = gptr(V.b.rank_stats);
        counters = gptr(V.b.counters);
        totals = gptr(V.b.ray_totals);
    }
    const uint32_t i = threadIdx.x;
    const uint32_t primary = p.count_primary_in_reset ? p.W * p.H : 0u;   // one primary ray per pixel
    if (i < 4) rank_stats[4 + i] = i == 0 ? 0x7F7FFFFFu : 0u;
    if (i < HISTOGRAM_BINS) rank_stats[8 + i] = 0u;
    // with early_set, the schedule pass may already have written the count and
    // the previous frame's resolve may still be counting its timed-out pixels
    const bool early = i == VE_CNT_CANDIDATES || i == VE_CNT_TIMEOUTS || i == VE_CNT_TIMEOUT_TILES;
    const bool keep = i == VE_CNT_EMISSIVES || (p.early_set && early);
    if (i < NUM_COUNTERS && !keep) counters[i] = i == VE_CNT_PRIMARY_RAYS ? primary : 0u;
    if (i == 0) totals[0] += primary;
}

void launch_reset(hipStream_t s, const FrameParams &p, DeviceBuffers &b)
{
    launch(k_frame_reset<false>, dim3(1), dim3(256), 0, s, p, b.rank_stats, b.counters, b.ray_totals, MultiTab{});
}

int launch_schedule(hipStream_t s, const FrameParams &p, DeviceBuffers &b)
{
    const uint32_t gw = (p.W + 7u) / 8u, gh = (p.H + 7u) / 8u;
    const uint32_t groups = gw * gh;
    const uint32_t n_vec = groups * 8u;
    uint32_t *ovf = overflow_word(p, b, 0u);
    launch(k_sched_claim<false>, dim3(groups), dim3(64), 0, s, p, b.gbuf, b.sched_tab, b.cand_at, ovf, b.sched_over,
           b.sched_over_n, b.sched_row_over, b.sched_rows, MultiTab{});
    launch(k_sched_rank<false>, dim3(CLAIM_WIDTH), dim3(64), 0, s, b.sched_tab, b.claim, b.samples, b.cand_at,
           b.sched_row_over, MultiTab{});
    launch(k_sched_rank_over<false>, dim3(SCHED_OVER_BLOCKS), dim3(256), 0, s, b.sched_tab, b.claim, b.samples, b.cand_at,
           b.sched_over, b.sched_over_n, b.sched_row_over, b.sched_rows, ovf, MultiTab{});
    launch(k_sched_vec_count<false>, dim3(cdiv(n_vec, 256)), dim3(256), 0, s, b.cand_at, n_vec, b.vec_cnt, b.vec_chunk,
           b.sched_over_n, MultiTab{});
    launch(k_sched_write<false>, dim3(cdiv(n_vec, 256)), dim3(256), 0, s, p, b.cand_at, n_vec, b.vec_cnt, b.vec_chunk,
           b.claim, b.candidates, b.counters, MultiTab{});
    return VE_OK;
}

int launch_rank(hipStream_t s, const FrameParams &p, DeviceBuffers &b)
{
    // candidates <= distinct visible voxels <= pixels
    const uint32_t bound = (uint32_t)std::min<size_t>((size_t)p.W * p.H, CANDIDATE_CAPACITY);
    launch(k_rank<false>, dim3(std::min<uint32_t>(cdiv(bound, 256), RANK_BLOCKS)), dim3(256), 0, s, p, b.candidates,
                       b.samples, b.lbuf, b.rank_stats, b.rank_part, b.counters, b.nsamp, b.vec_cnt, b.rank_chunk, MultiTab{});
    launch(k_rank_write<false>, dim3(cdiv(bound, 256)), dim3(256), 0, s, p, b.candidates, b.nsamp, b.vec_cnt,
                       b.rank_chunk, b.shade_list, b.counters, b.rank_stats, b.rank_part, MultiTab{});
    return VE_OK;
}

void launch_threshold(hipStream_t s, const FrameParams &p, DeviceBuffers &b)
{
    launch(k_threshold, dim3(1), dim3(64), 0, s, p, b.rank_stats, b.counters);
}

void launch_emit(hipStream_t s, const FrameParams &p, DeviceBuffers &b)
{
    launch(k_emit<false>, dim3(cdiv(p.shade_budget, 256)), dim3(256), 0, s, p, b.shade_list, b.rays, b.counters, b.gbuf, b.ghit,
           b.materials, b.emissive_claim, b.em_bits, b.hit_bits, b.ray_totals, b.ray_li[0], b.dep_death, b.rank_stats, MultiTab{});
}


void launch_trace(hipStream_t s, const uint32_t *words, const FrameParams &p, DeviceBuffers &b, uint32_t iter, bool stats)
{
    const uint32_t nb = cdiv(p.shade_budget, 256);
    if (stats)
        launch(k_trace<true, false>, dim3(nb), dim3(256), 0, s, words, p, bounce_in(b, iter), live_in_ptr(b, iter),
                           iter, b.counters, b.node_stats, b.ray_totals, b.materials, b.emissive_claim, b.em_bits,
                           live_out_ptr(b, iter), b.hit_bits, MultiTab{});
    else
        launch(k_trace<false, false>, dim3(nb), dim3(256), 0, s, words, p, bounce_in(b, iter),
                           live_in_ptr(b, iter), iter, b.counters, b.node_stats, b.ray_totals, b.materials,
                           b.emissive_claim, b.em_bits, live_out_ptr(b, iter), b.hit_bits, MultiTab{});
    if (iter == 0)
        launch(k_em_register<false>, dim3(1), dim3(EM_THREADS), 0, s, bounce_in(b, iter), live_in_ptr(b, iter),
                           b.materials, b.emissive_claim, b.emissives, b.em_bits, b.em_reg, b.counters, MultiTab{});
}

void launch_shade(hipStream_t s, const uint32_t *words, const FrameParams &p, DeviceBuffers &b, uint32_t iter)
{
    const uint32_t nb = cdiv(p.shade_budget, 256);
    launch(k_shade<false>, dim3(nb), dim3(256), 0, s, p, words, iter, bounce_in(b, iter), live_in_ptr(b, iter),
           li_in_ptr(b, iter), bounce_out(b, iter), live_out_ptr(b, iter), li_out_ptr(b, iter), b.deferred, b.nbuf, b.rbuf,
           b.materials, b.emissives, b.em_reg, b.counters, b.ray_totals, MultiTab{});
}

void launch_bounce(hipStream_t s, const uint32_t *words, const FrameParams &p, DeviceBuffers &b, uint32_t first,
                   uint32_t iters, bool stats)
{
    const uint32_t full = cdiv(p.shade_budget, 256);
    const uint32_t nb = p.bounce_blocks ? std::min(full, p.bounce_blocks) : full;
    auto go = [&](auto kernel) {
        launch(kernel, dim3(nb), dim3(256), 0, s, words, p, bounce_in(b, first), live_in_ptr(b, first), first, iters,
               b.counters, b.node_stats, b.ray_totals, b.materials, b.nbuf, b.rbuf, b.emissives, b.em_reg,
               b.emissive_claim, b.em_bits, li_in_ptr(b, first), b.deferred, b.bounce_part, MultiTab{});
    };
    const bool refill = nb < full;
    if (stats) refill ? go(k_bounce<true, true, false>) : go(k_bounce<true, false, false>);
    else refill ? go(k_bounce<false, true, false>) : go(k_bounce<false, false, false>);
}

int launch_deposits(hipStream_t s, const FrameParams &p, DeviceBuffers &b, uint32_t iters)
{
    const uint32_t chunks = (p.shade_budget + RAY_CHUNK - 1u) / RAY_CHUNK;
    const DepLists dl = dep_lists(b);
    launch(k_dep_prep<false>, dim3(DEFERRED_BLOCKS), dim3(256), 0, s, p, b.deferred, b.counters, b.bounce_part, b.ray_totals,
           b.dep_death, dl, MultiTab{});
    launch(k_dep_wpos<false>, dim3(chunks * WPOS_BLOCKS), dim3(1024), 0, s, b.counters, b.dep_death, b.dep_wpos, iters, MultiTab{});
    launch(k_dep_replay<false>, dim3(cdiv(p.hash_width, REPLAY_WAVES)), dim3(64 * REPLAY_WAVES), 0, s, p, b.lbuf, dl,
           b.deferred, b.dep_wpos, b.counters, MultiTab{});
    if (!p.fold_small_passes) launch(k_dep_done, dim3(1), dim3(1), 0, s, b.counters, b.dep_over_n, 0);   // else the resolve's
    return VE_OK;
}

// An overlapped frame's resolve runs on another stream: the deposit records'
// clear and the fold of the early passes' overflow take a launch of their own.
void launch_dep_done(hipStream_t s, const FrameParams &p, DeviceBuffers &b)
{
    launch(k_dep_done, dim3(1), dim3(1), 0, s, b.counters, b.dep_over_n, (int32_t)p.early_set);
}

// Partial node-array update (Octree::UpdateNode's glBufferSubData per word).
__global__ void k_scatter_words(uint32_t *__restrict__ words, const uint2 *__restrict__ pairs, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) words[pairs[i].x] = pairs[i].y;
}

void launch_scatter_words(hipStream_t s, uint32_t *words, const uint2 *pairs, uint32_t n)
{
    if (n) launch(k_scatter_words, dim3(cdiv(n, 256)), dim3(256), 0, s, words, pairs, n);
}

void launch_display(hipStream_t s, const FrameParams &p, DeviceBuffers &b)
{
    const uint32_t n = p.W * p.H;
    launch(k_display, dim3(cdiv(n, 256)), dim3(256), 0, s, n, b.output, b.display);
}

// Blocks of the tiled timeout reconstruction (dense 16x16 tiles staged in LDS;
// none for radii past RECON_TILED_MAX) and their dynamic LDS bytes.
static uint32_t timeout_tile_blocks(const FrameParams &p)
{
    if (p.reconstruction_radius > (int)RECON_TILED_MAX) return 0u;
    return std::min(cdiv(p.W, RECON_TILE) * cdiv(p.H, RECON_TILE), RECON_TILE_BLOCKS);
}

uint32_t timeouts_lds(const FrameParams &p)
{
    if (p.reconstruction_radius > (int)RECON_TILED_MAX) return 0u;
    const uint32_t D = RECON_TILE + 2u * (uint32_t)std::max(p.reconstruction_radius, 0);
    return (uint32_t)(D * D * sizeof(float4));
}

void launch_resolve(hipStream_t s, const FrameParams &p, DeviceBuffers &b)
{
    launch(k_resolve_avg<false>, dim3(screen_blocks(p)), dim3(256), 0, s, p, b.gbuf, b.lbuf, b.pix_avg, b.pix_cnt,
                       b.counters, b.timeout_tiles, b.dep_over_n, MultiTab{});
    launch(k_resolve<false>, dim3(screen_blocks(p)), dim3(256), 0, s, p, b.gbuf, b.lbuf, b.nbuf, b.materials,
                       b.pix_avg, b.pix_cnt, b.output, b.timeouts, b.counters, b.timeout_tiles, b.timeout_list, MultiTab{});
    // timed-out pixels one wave each; in timeout-heavy frames, the tiles
    // holding RECON_DENSE or more go to the tiled blocks instead
    launch(k_resolve_timeouts<false>, dim3(RESOLVE_TIMEOUT_BLOCKS + timeout_tile_blocks(p)), dim3(256), timeouts_lds(p),
           s, p, b.gbuf, b.pix_avg, b.pix_cnt, b.timeouts, b.timeout_tiles, b.timeout_list, b.counters, b.output,
           RESOLVE_TIMEOUT_BLOCKS, MultiTab{});
}


// ---------------------------------------------------------------------------
// one launch per pass over many views
// ---------------------------------------------------------------------------
uint32_t multi_blocks(uint32_t k, const FrameParams &p, uint32_t iters)
{
    const uint32_t rays = cdiv(p.shade_budget, 256);
    const uint32_t bound = (uint32_t)std::min<size_t>((size_t)p.W * p.H, CANDIDATE_CAPACITY);
    const uint32_t groups = ((p.W + 7u) / 8u) * ((p.H + 7u) / 8u);
    switch (k) {
    case MK_RESET: return 1u;
    case MK_NRM_EVENTS: return screen_blocks(p);
    case MK_NRM_REPLAY: return cdiv(p.hash_width, REPLAY_WAVES);
    case MK_NRM_COMPUTE: return cdiv((size_t)p.hash_width * NBUFFER_SLOTS, 256);
    case MK_SCHED_CLAIM: return groups;
    case MK_SCHED_RANK: return CLAIM_WIDTH;
    case MK_SCHED_RANK_OVER: return SCHED_OVER_BLOCKS;
    case MK_SCHED_VEC_COUNT:
    case MK_SCHED_WRITE: return cdiv(sched_vectors(p), 256);
    case MK_RANK: return std::min<uint32_t>(cdiv(bound, 256), RANK_BLOCKS);
    case MK_RANK_WRITE: return cdiv(bound, 256);
    case MK_EMIT: return rays;
    case MK_TRACE:
    case MK_SHADE: return iters > 0 ? rays : 0u;
    case MK_EM_REGISTER: return iters > 0 ? 1u : 0u;
    case MK_BOUNCE: return iters > 1 ? rays : 0u;
    case MK_DEP_PREP: return iters > 0 ? DEFERRED_BLOCKS : 0u;
    case MK_DEP_WPOS: return iters > 0 ? (p.shade_budget + RAY_CHUNK - 1u) / RAY_CHUNK * WPOS_BLOCKS : 0u;
    case MK_DEP_REPLAY: return iters > 0 ? cdiv(p.hash_width, REPLAY_WAVES) : 0u;
    case MK_RESOLVE_AVG:
    case MK_RESOLVE: return screen_blocks(p);
    case MK_RESOLVE_TIMEOUTS: return RESOLVE_TIMEOUT_BLOCKS + timeout_tile_blocks(p);
    default: return 0u;
    }
}

// The multi launches pass null / zero for every argument their kernels take
// from the view records instead.
void launch_frames_multi(hipStream_t s, const ViewRec *views, const uint32_t *pre, uint32_t n, const uint32_t *grids,
                         uint32_t lds)
{
    auto tab = [&](uint32_t k) { return MultiTab{views, pre + (size_t)k * (n + 1u), n}; };
    const FrameParams z{};
    const OrderLists ol{};
    const DepLists dl{};
    if (grids[MK_RESET]) launch(k_frame_reset<true>, dim3(grids[MK_RESET]), dim3(256), 0, s, z, nullptr, nullptr, nullptr, tab(MK_RESET));
    if (grids[MK_NRM_EVENTS])
        launch(k_nrm_events<true>, dim3(grids[MK_NRM_EVENTS]), dim3(256), 0, s, z, nullptr, nullptr, ol, nullptr,
               tab(MK_NRM_EVENTS));
    if (grids[MK_NRM_REPLAY])
        launch(k_nrm_replay<true>, dim3(grids[MK_NRM_REPLAY]), dim3(64 * REPLAY_WAVES), 0, s, z, nullptr, nullptr, ol,
               nullptr, nullptr, tab(MK_NRM_REPLAY));
    if (grids[MK_NRM_COMPUTE])
        launch(k_nrm_compute<true>, dim3(grids[MK_NRM_COMPUTE]), dim3(256), 0, s, nullptr, z, nullptr, nullptr, nullptr,
               tab(MK_NRM_COMPUTE));
    if (grids[MK_SCHED_CLAIM])
        launch(k_sched_claim<true>, dim3(grids[MK_SCHED_CLAIM]), dim3(64), 0, s, z, nullptr, nullptr, nullptr, nullptr,
               nullptr, nullptr, nullptr, nullptr, tab(MK_SCHED_CLAIM));
    if (grids[MK_SCHED_RANK])
        launch(k_sched_rank<true>, dim3(grids[MK_SCHED_RANK]), dim3(64), 0, s, nullptr, nullptr, nullptr, nullptr, nullptr,
               tab(MK_SCHED_RANK));
    if (grids[MK_SCHED_RANK_OVER])
        launch(k_sched_rank_over<true>, dim3(grids[MK_SCHED_RANK_OVER]), dim3(256), 0, s, nullptr, nullptr, nullptr,
               nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, tab(MK_SCHED_RANK_OVER));
    if (grids[MK_SCHED_VEC_COUNT])
        launch(k_sched_vec_count<true>, dim3(grids[MK_SCHED_VEC_COUNT]), dim3(256), 0, s, nullptr, 0u, nullptr, nullptr,
               nullptr, tab(MK_SCHED_VEC_COUNT));
    if (grids[MK_SCHED_WRITE])
        launch(k_sched_write<true>, dim3(grids[MK_SCHED_WRITE]), dim3(256), 0, s, z, nullptr, 0u, nullptr, nullptr,
               nullptr, nullptr, nullptr, tab(MK_SCHED_WRITE));
    if (grids[MK_RANK])
        launch(k_rank<true>, dim3(grids[MK_RANK]), dim3(256), 0, s, z, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
               nullptr, nullptr, nullptr, tab(MK_RANK));
    if (grids[MK_RANK_WRITE])
        launch(k_rank_write<true>, dim3(grids[MK_RANK_WRITE]), dim3(256), 0, s, z, nullptr, nullptr, nullptr, nullptr,
               nullptr, nullptr, nullptr, nullptr, tab(MK_RANK_WRITE));
    if (grids[MK_EMIT])
        launch(k_emit<true>, dim3(grids[MK_EMIT]), dim3(256), 0, s, z, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
               nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, tab(MK_EMIT));
    if (grids[MK_TRACE])
        launch(k_trace<false, true>, dim3(grids[MK_TRACE]), dim3(256), 0, s, nullptr, z, nullptr, nullptr, 0u, nullptr,
               nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, tab(MK_TRACE));
    if (grids[MK_EM_REGISTER])
        launch(k_em_register<true>, dim3(grids[MK_EM_REGISTER]), dim3(EM_THREADS), 0, s, nullptr, nullptr, nullptr,
               nullptr, nullptr, nullptr, nullptr, nullptr, tab(MK_EM_REGISTER));
    if (grids[MK_SHADE])
        launch(k_shade<true>, dim3(grids[MK_SHADE]), dim3(256), 0, s, z, nullptr, 0u, nullptr, nullptr, nullptr, nullptr,
               nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, tab(MK_SHADE));
    if (grids[MK_BOUNCE])
        launch(k_bounce<false, false, true>, dim3(grids[MK_BOUNCE]), dim3(256), 0, s, nullptr, z, nullptr, nullptr, 0u,
               0u, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
               nullptr, nullptr, tab(MK_BOUNCE));
    if (grids[MK_DEP_PREP])
        launch(k_dep_prep<true>, dim3(grids[MK_DEP_PREP]), dim3(256), 0, s, z, nullptr, nullptr, nullptr, nullptr,
               nullptr, dl, tab(MK_DEP_PREP));
    if (grids[MK_DEP_WPOS])
        launch(k_dep_wpos<true>, dim3(grids[MK_DEP_WPOS]), dim3(1024), 0, s, nullptr, nullptr, nullptr, 0u,
               tab(MK_DEP_WPOS));
    if (grids[MK_DEP_REPLAY])
        launch(k_dep_replay<true>, dim3(grids[MK_DEP_REPLAY]), dim3(64 * REPLAY_WAVES), 0, s, z, nullptr, dl, nullptr,
               nullptr, nullptr, tab(MK_DEP_REPLAY));
    if (grids[MK_RESOLVE_AVG])
        launch(k_resolve_avg<true>, dim3(grids[MK_RESOLVE_AVG]), dim3(256), 0, s, z, nullptr, nullptr, nullptr, nullptr,
               nullptr, nullptr, nullptr, tab(MK_RESOLVE_AVG));
    if (grids[MK_RESOLVE])
        launch(k_resolve<true>, dim3(grids[MK_RESOLVE]), dim3(256), 0, s, z, nullptr, nullptr, nullptr, nullptr, nullptr,
               nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, tab(MK_RESOLVE));
    if (grids[MK_RESOLVE_TIMEOUTS])
        launch(k_resolve_timeouts<true>, dim3(grids[MK_RESOLVE_TIMEOUTS]), dim3(256), lds, s, z, nullptr, nullptr,
               nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0u, tab(MK_RESOLVE_TIMEOUTS));
}

}  // namespace ve
