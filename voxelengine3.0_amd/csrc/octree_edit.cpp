// Editable host octree: the persistent pointer tree of OctreeCPU
// (src/octree_cpu.cpp) together with the node-word image Octree keeps for it
// (src/renderer/octree.cpp): bulk linearisation (Octree::set), incremental
// sync after insert/remove (Octree::applyEdits) with LIFO reuse of freed
// 8-word blocks, leaf-slot lookup (findGpuSlot) and the CPU ray cast used for
// picking (OctreeCPU::raycast).  Host code only; ve_scene_update carries the
// changed words to a device node array.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/ve_svo.h"
#include "ve_internal.h"

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;

struct ENode {
    ENode *child[8] = {};
    uint8_t count = 0;        // childrenCount
    uint32_t material = 0;
    uint32_t gpu_block = NONE;
    bool dirty = false;
};

void free_subtree(ENode *n)
{
    if (!n) return;
    std::vector<ENode *> stk{n};
    while (!stk.empty()) {
        ENode *x = stk.back();
        stk.pop_back();
        for (ENode *c : x->child)
            if (c) stk.push_back(c);
        delete x;
    }
}

uint32_t internal_word(uint32_t next, uint32_t count) { return 1u | ((count & 7u) << 1) | (next << 4); }
uint32_t leaf_word(uint32_t material) { return (material & 0x7Fu) << 1; }

}  // namespace

struct ve_octree_cpu {
    uint32_t depth = 0;
    ENode *root = nullptr;
    std::vector<uint32_t> freed;        // OctreeCPU::freedGpuBlocks
    // Octree side
    std::vector<uint32_t> data;         // node words; data.size() == capacity
    uint32_t size = 8, capacity = 8, num_voxels = 0, max_words = 1u << 27;
    std::vector<uint32_t> free_blocks;  // Octree::freeBlocks (a stack: back() is top())
    std::vector<uint32_t> changed;      // slots written by the last set / apply_edits
    bool full_upload = false;           // the last call rewrote the whole image
    bool incomplete = false;            // the last set hit the word limit: no image until a set succeeds

    ~ve_octree_cpu() { free_subtree(root); }

    // OctreeCPU::locate (octree_cpu.cpp:33-38): level l uses bit depth - l
    uint32_t locate(const uint32_t p[3], uint32_t level) const
    {
        const uint32_t b = 1u << (depth - level);
        return ((uint32_t)((p[0] & b) != 0) << 2) | ((uint32_t)((p[1] & b) != 0) << 1) | (uint32_t)((p[2] & b) != 0);
    }

    void insert_r(const uint32_t p[3], uint32_t material, ENode *&node, uint32_t level)   // octree_cpu.cpp:50-64
    {
        if (!node) node = new ENode();
        node->dirty = true;
        if (level > depth) {
            node->material = material;
            node->count = 0;
            return;
        }
        const uint32_t idx = locate(p, level);
        const bool existed = node->child[idx] != nullptr;
        insert_r(p, material, node->child[idx], level + 1);
        if (!existed && node->child[idx]) node->count++;
    }

    void remove_r(const uint32_t p[3], ENode *&node, uint32_t level)   // octree_cpu.cpp:74-98
    {
        if (!node) return;
        node->dirty = true;
        if (level > depth) {
            delete node;
            node = nullptr;
            return;
        }
        const uint32_t idx = locate(p, level);
        const bool had = node->child[idx] != nullptr;
        remove_r(p, node->child[idx], level + 1);
        if (had && node->child[idx] == nullptr) {
            node->count--;
            if (node->count == 0) {
                if (node->gpu_block != NONE) freed.push_back(node->gpu_block);
                delete node;
                node = nullptr;
            }
        }
    }

    // Octree::resizeDataIfNeeded (octree.cpp:95-119): doubling, clamped to the
    // word limit; a resize re-uploads the whole image.  Returns false when
    // `required` exceeds the limit (the reference then writes past its
    // array); nothing is resized then.
    bool resize(uint32_t required)
    {
        if (required <= capacity) return true;
        if (required > max_words) return false;
        uint64_t cap = capacity;
        while (cap < required) cap *= 2;
        capacity = (uint32_t)std::min<uint64_t>(cap, max_words);
        data.resize(capacity, 0u);
        full_upload = true;
        return true;
    }

    void put(uint32_t slot, uint32_t w)
    {
        data[slot] = w;
        changed.push_back(slot);
    }

    // Octree::linearizeNode (octree.cpp:170-215): explicit-stack DFS,
    // children pushed 0..7 (so 7 is linearised first).  Returns false, having
    // written nothing past the array, when a block would cross the word limit.
    bool linearize(ENode *start, uint32_t off0, uint32_t level0)
    {
        struct E {
            ENode *n;
            uint32_t off, lv;
        };
        std::vector<E> stk{{start, off0, level0}};
        while (!stk.empty()) {
            const E e = stk.back();
            stk.pop_back();
            if (!e.n) continue;
            e.n->dirty = false;
            if (e.lv > depth || (e.n->material != 0 && e.n->count == 0)) {
                if (e.n->material == 0) continue;
                if (!resize(e.off + 1)) return false;
                data[e.off] = leaf_word(e.n->material);
                num_voxels++;
            } else {
                if ((uint64_t)size + 8u > max_words || !resize(size + 8u) || e.off >= capacity) return false;
                const uint32_t block = size;
                size += 8;
                e.n->gpu_block = block;
                data[e.off] = internal_word(block, e.n->count);
                for (uint32_t i = 0; i < 8; i++) stk.push_back({e.n->child[i], block + i, e.lv + 1});
            }
        }
        return true;
    }

    // Octree::applyEditsNode (octree.cpp:252-312).  Returns false when a
    // subtree could not be given a block under the word limit: its slot keeps
    // its old word, as in the reference (octree.cpp:290-294), but the path to
    // it stays dirty so a later apply (after removes, or with a set into a
    // larger limit) writes it.
    bool apply_node(ENode *node, uint32_t slot, uint32_t level)
    {
        if (!node) {
            if (data[slot] != 0u) put(slot, 0u);
            return true;
        }
        if (!node->dirty) return true;
        node->dirty = false;
        if (level > depth || (node->material != 0 && node->count == 0)) {
            put(slot, leaf_word(node->material));
            if (node->material != 0) num_voxels++;
            if (node->gpu_block != NONE) {
                free_blocks.push_back(node->gpu_block);
                node->gpu_block = NONE;
            }
            return true;
        }
        if (node->gpu_block == NONE) {
            uint32_t block;
            if (!free_blocks.empty()) {
                block = free_blocks.back();
                free_blocks.pop_back();
                for (uint32_t i = 0; i < 8; i++) put(block + i, 0u);   // recycled block zeroed
            } else {
                if ((uint64_t)size + 8u > max_words || !resize(size + 8u)) {   // the word limit
                    node->dirty = true;
                    return false;
                }
                block = size;
                size += 8;
            }
            node->gpu_block = block;
        }
        put(slot, internal_word(node->gpu_block, node->count));
        bool ok = true;
        for (uint32_t i = 0; i < 8; i++) ok &= apply_node(node->child[i], node->gpu_block + i, level + 1);
        if (!ok) node->dirty = true;
        return ok;
    }
};

extern "C" {

int ve_octree_cpu_create(ve_octree_cpu **out, uint32_t depth, uint32_t max_words)
{
    ve_clear_global_error();
    if (!out) return ve_global_error(VE_EINVAL, "null argument");
    if (depth < 1 || depth > VE_MAX_DEPTH) return ve_global_error(VE_EINVAL, "depth out of range [1,16]");
    ve_octree_cpu *t = new (std::nothrow) ve_octree_cpu();
    if (!t) return ve_global_error(VE_ENOMEM, "octree allocation");
    t->depth = depth;
    // the 28-bit `next` field addresses at most 2^28 words
    t->max_words = max_words ? std::min<uint32_t>(max_words, 1u << 28) : (1u << 27);
    if (t->max_words < 9u) {
        delete t;
        return ve_global_error(VE_EINVAL, "max_words below 9 (the root and its first block)");
    }
    t->data.assign(t->capacity, 0u);
    *out = t;
    return VE_OK;
}

void ve_octree_cpu_destroy(ve_octree_cpu *t) { delete t; }

int ve_octree_cpu_insert(ve_octree_cpu *t, const ve_insert *ops, size_t n)
{
    if (!t || (n && !ops)) return ve_global_error(VE_EINVAL, "null argument");
    const uint32_t bound = 1u << t->depth;
    for (size_t i = 0; i < n; i++) {   // OctreeCPU::insert (octree_cpu.cpp:40-48)
        const uint32_t p[3] = {ops[i].x, ops[i].y, ops[i].z};
        if (p[0] >= bound || p[1] >= bound || p[2] >= bound) continue;
        if (!t->root) t->root = new ENode();
        t->insert_r(p, ops[i].material, t->root, 1);
    }
    return VE_OK;
}

int ve_octree_cpu_remove(ve_octree_cpu *t, const uint32_t *xyz, size_t n)
{
    if (!t || (n && !xyz)) return ve_global_error(VE_EINVAL, "null argument");
    const uint32_t bound = 1u << t->depth;
    for (size_t i = 0; i < n; i++) {   // OctreeCPU::remove (octree_cpu.cpp:66-72)
        const uint32_t *p = xyz + 3 * i;
        if (p[0] >= bound || p[1] >= bound || p[2] >= bound || !t->root) continue;
        t->remove_r(p, t->root, 1);
    }
    return VE_OK;
}

int ve_octree_cpu_lookup(const ve_octree_cpu *t, uint32_t x, uint32_t y, uint32_t z, int *present, uint32_t *material)
{
    if (!t || !present || !material) return ve_global_error(VE_EINVAL, "null argument");
    *present = 0;
    *material = 0;
    const uint32_t bound = 1u << t->depth, p[3] = {x, y, z};
    if (x >= bound || y >= bound || z >= bound) return VE_OK;
    const ENode *node = t->root;   // OctreeCPU::lookup (octree_cpu.cpp:100-111)
    for (uint32_t level = 1; level <= t->depth + 1; level++) {
        if (!node) return VE_OK;
        if (level > t->depth) {
            *present = 1;
            *material = node->material;
            return VE_OK;
        }
        node = node->child[t->locate(p, level)];
    }
    return VE_OK;
}

// OctreeCPU::raycast (octree_cpu.cpp:128-215): the reference's picking ray on
// the pointer tree, FP32 as written (no fused operations).
int ve_octree_cpu_raycast(const ve_octree_cpu *t, const float origin_in[3], const float dir[3], uint32_t max_steps,
                          int *hit, uint32_t pos[3], uint32_t *material)
{
    if (!t || !origin_in || !dir || !hit || !pos || !material) return ve_global_error(VE_EINVAL, "null argument");
    *hit = 0;
    *material = 0;
    pos[0] = pos[1] = pos[2] = 0;
    if (!t->root) return VE_OK;
    const uint32_t len = 1u << t->depth;
    const float bF = (float)len;
    const float inv[3] = {1.0f / dir[0], 1.0f / dir[1], 1.0f / dir[2]};
    float o[3];
    for (int i = 0; i < 3; i++) o[i] = origin_in[i] + dir[i] * 4.0f;
    auto in_bounds = [&](const float p[3]) {
        return p[0] >= 0.0f && p[0] <= bF && p[1] >= 0.0f && p[1] <= bF && p[2] >= 0.0f && p[2] <= bF;
    };
    // glm::min / glm::max: (b < a ? b : a) / (a < b ? b : a)
    auto fmin_ = [](float a, float b) { return b < a ? b : a; };
    auto fmax_ = [](float a, float b) { return a < b ? b : a; };
    float r[3];
    if (in_bounds(o)) {
        for (int i = 0; i < 3; i++) r[i] = o[i];
    } else {
        float tmi[3], tmx[3];
        for (int i = 0; i < 3; i++) {
            const float t1 = ((0.0f - o[i]) + 0.001f) * inv[i], t2 = ((bF - o[i]) - 0.001f) * inv[i];
            tmi[i] = fmin_(t1, t2);
            tmx[i] = fmax_(t1, t2);
        }
        const float te = fmax_(fmax_(tmi[0], tmi[1]), tmi[2]);
        const float tx = fmin_(fmin_(tmx[0], tmx[1]), tmx[2]);
        if (tx < te || tx < 0.0f) return VE_OK;
        for (int i = 0; i < 3; i++) r[i] = dir[i] * te + o[i];
    }
    for (uint32_t q = 0; in_bounds(r) && q <= max_steps; ++q) {
        const uint32_t ur[3] = {(uint32_t)r[0], (uint32_t)r[1], (uint32_t)r[2]};
        const ENode *node = t->root;
        uint32_t tsz = len, torg[3] = {0, 0, 0};
        for (uint32_t lev = 1; lev <= t->depth; ++lev) {
            const uint32_t csz = 1u << (t->depth - lev);
            const ENode *c = node->child[t->locate(ur, lev)];
            const uint32_t corg[3] = {ur[0] & ~(csz - 1u), ur[1] & ~(csz - 1u), ur[2] & ~(csz - 1u)};
            if (!c) {
                tsz = csz ? csz : 1u;
                std::memcpy(torg, corg, sizeof torg);
                break;
            }
            if (lev == t->depth) {
                if (c->material != 0) {
                    *hit = 1;
                    *material = c->material;
                    std::memcpy(pos, ur, 3 * sizeof(uint32_t));
                    return VE_OK;
                }
                tsz = 1u;
                std::memcpy(torg, ur, sizeof torg);
                break;
            }
            node = c;
            tsz = csz;
            std::memcpy(torg, corg, sizeof torg);
        }
        float tmx[3];
        for (int i = 0; i < 3; i++) {
            const float bmin = (float)torg[i], bmax = bmin + (float)tsz;
            const float t1 = ((bmin - o[i]) - 0.001f) * inv[i], t2 = ((bmax - o[i]) + 0.001f) * inv[i];
            tmx[i] = fmax_(t1, t2);
        }
        const float tx = fmin_(fmin_(tmx[0], tmx[1]), tmx[2]);
        for (int i = 0; i < 3; i++) r[i] = dir[i] * tx + o[i];
    }
    return VE_OK;
}

// Octree::set (octree.cpp:127-166)
int ve_octree_set(ve_octree_cpu *t, ve_octree_info *info)
{
    if (!t) return ve_global_error(VE_EINVAL, "null argument");
    t->changed.clear();
    t->full_upload = true;
    if (t->root) {
        t->data.assign(t->capacity, 0u);
        t->size = 9;   // slot 0 root, slots 1-8 its children
        t->num_voxels = 0;
        t->free_blocks.clear();
        t->freed.clear();
        t->resize(t->size);
        t->root->gpu_block = 1;
        t->root->dirty = false;
        t->data[0] = internal_word(1, t->root->count);
        for (uint32_t i = 0; i < 8; i++)
            if (!t->linearize(t->root->child[i], i + 1, 2)) {
                t->incomplete = true;
                if (info) *info = ve_octree_info{t->depth, t->size, t->capacity, t->num_voxels};
                return ve_global_error(VE_ECAPACITY, "octree needs more node words than the limit");
            }
    }
    t->incomplete = false;
    if (info) *info = ve_octree_info{t->depth, t->size, t->capacity, t->num_voxels};
    return VE_OK;
}

// Octree::applyEdits (octree.cpp:228-250)
int ve_octree_apply_edits(ve_octree_cpu *t, ve_octree_info *info)
{
    if (!t) return ve_global_error(VE_EINVAL, "null argument");
    if (t->incomplete) return ve_global_error(VE_ECAPACITY, "the last set hit the word limit; set again first");
    t->changed.clear();
    t->full_upload = false;
    if (t->root) {
        for (uint32_t b : t->freed) t->free_blocks.push_back(b);
        t->freed.clear();
        if (t->root->dirty) {
            t->put(0, internal_word(1, t->root->count));
            t->root->dirty = false;
            bool ok = true;
            for (uint32_t i = 0; i < 8; i++) ok &= t->apply_node(t->root->child[i], i + 1, 2);
            if (!ok) {
                t->root->dirty = true;
                if (info) *info = ve_octree_info{t->depth, t->size, t->capacity, t->num_voxels};
                return ve_global_error(VE_ECAPACITY, "edits need more node words than the limit; "
                                                     "the subtrees without a block are left out of the image");
            }
        }
    }
    if (info) *info = ve_octree_info{t->depth, t->size, t->capacity, t->num_voxels};
    return VE_OK;
}

int ve_octree_words(const ve_octree_cpu *t, const uint32_t **words, uint32_t *capacity)
{
    if (!t || !words) return ve_global_error(VE_EINVAL, "null argument");
    *words = t->data.data();
    if (capacity) *capacity = t->capacity;
    return VE_OK;
}

int ve_octree_changed(const ve_octree_cpu *t, const uint32_t **slots, size_t *n, int *full)
{
    if (!t || !slots || !n) return ve_global_error(VE_EINVAL, "null argument");
    *slots = t->changed.data();
    *n = t->changed.size();
    if (full) *full = t->full_upload ? 1 : 0;
    return VE_OK;
}

// Octree::findGpuSlot (octree.cpp:322-341)
uint32_t ve_octree_find_slot(const ve_octree_cpu *t, uint32_t x, uint32_t y, uint32_t z)
{
    if (!t || !t->root) return NONE;
    const uint32_t p[3] = {x, y, z};
    const ENode *node = t->root;
    uint32_t block = 1u;
    for (uint32_t level = 1; level <= t->depth; level++) {
        const uint32_t idx = t->locate(p, level);
        const ENode *c = node->child[idx];
        if (!c) return NONE;
        if (level == t->depth) return block + idx;
        if (c->gpu_block == NONE) return NONE;
        block = c->gpu_block;
        node = c;
    }
    return NONE;
}

}  // extern "C"
