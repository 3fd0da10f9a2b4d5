// Octree build: OctreeCPU::insert* followed by Octree::set, as one sort-based
// pass that never materialises the pointer tree.
//
// Reference behaviour reproduced (file:line under the reference tree):
//   * insert/insertR (src/octree_cpu.cpp:40-64): out-of-range positions are
//     ignored; a repeated position keeps its node and takes the last material
//     (material 0 included — the node still exists and still counts as a child).
//   * Octree::set (src/renderer/octree.cpp:127-166): slot 0 = root with next=1,
//     slots 1-8 = the root's children, visited child 0 first.
//   * linearizeNode (octree.cpp:170-217): explicit-stack DFS that pushes
//     children 0..7, so inside each root-child subtree child 7 is visited
//     first; every internal node takes the next 8-word block when it is
//     popped; leaves with material 0 leave their slot 0; count = childrenCount
//     truncated to 3 bits (8 children -> 0).
//   * capacity = power of two >= size, starting from the fresh Octree's 8 and
//     the first resize to 9 (octree.cpp:4-22,95-119,133-137).
//
// The pointer tree's child sets equal the sets of distinct Morton digits of
// the inserted positions, so the build is: stable radix sort of Morton keys
// (insertion order breaks ties, the last insert of a position wins), then one
// recursive walk of the sorted keys that emits blocks in the reference's DFS
// order.  Work O(V log V) host time, no per-voxel allocation.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/ve_svo.h"
#include "ve_internal.h"

namespace {

inline uint64_t morton_key(uint32_t x, uint32_t y, uint32_t z, uint32_t depth)
{
    // digit at level l (1..depth) = locate(pos, 2^(depth-l)) = x_b<<2 | y_b<<1 | z_b
    // with b = depth-l; level 1 is the most significant digit.
    uint64_t k = 0;
    for (int b = int(depth) - 1; b >= 0; --b) {
        uint64_t d = (uint64_t((x >> b) & 1u) << 2) | (uint64_t((y >> b) & 1u) << 1) | uint64_t((z >> b) & 1u);
        k = (k << 3) | d;
    }
    return k;
}

// Stable LSD radix sort of (key, payload) by key, 8-bit digits, skipping
// digits that are constant over the input.
void radix_sort(std::vector<uint64_t> &keys, std::vector<uint32_t> &vals, int key_bits)
{
    const size_t n = keys.size();
    std::vector<uint64_t> k2(n);
    std::vector<uint32_t> v2(n);
    for (int shift = 0; shift < key_bits; shift += 8) {
        size_t cnt[257] = {0};
        for (size_t i = 0; i < n; ++i) cnt[((keys[i] >> shift) & 0xFF) + 1]++;
        bool trivial = false;
        for (int d = 1; d <= 256; ++d)
            if (cnt[d] == n) trivial = true;
        if (trivial) continue;
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (size_t i = 0; i < n; ++i) {
            size_t pos = cnt[(keys[i] >> shift) & 0xFF]++;
            k2[pos] = keys[i];
            v2[pos] = vals[i];
        }
        keys.swap(k2);
        vals.swap(v2);
    }
}

struct Builder {
    const uint64_t *keys;
    const uint32_t *mats;
    uint32_t depth;
    uint32_t *words;
    uint32_t size;
    uint32_t num_voxels;

    inline uint32_t digit(size_t i, uint32_t level) const
    {
        return uint32_t(keys[i] >> (3u * (depth - level))) & 7u;
    }

    // Child ranges of [lo,hi) split by the digit of `level` (keys sorted, so
    // each digit's keys are contiguous): bounds[d]..bounds[d+1] holds digit d.
    void split(size_t lo, size_t hi, uint32_t level, size_t bounds[9]) const
    {
        size_t i = lo;
        for (uint32_t d = 0; d < 8; ++d) {
            bounds[d] = i;
            size_t a = i, e = hi;  // first index in [i,hi) whose digit exceeds d
            while (a < e) {
                size_t m = a + (e - a) / 2;
                if (digit(m, level) <= d) a = m + 1; else e = m;
            }
            i = a;
        }
        bounds[8] = hi;
    }

    // A node at `level` (2..depth+1) covering keys [lo,hi), written to `slot`.
    void visit(size_t lo, size_t hi, uint32_t level, uint32_t slot)
    {
        if (level > depth) {
            uint32_t m = mats[hi - 1];
            if (m == 0) return;
            words[slot] = (m & 0x7Fu) << 1;
            ++num_voxels;
            return;
        }
        size_t b[9];
        split(lo, hi, level, b);
        uint32_t count = 0;
        for (int d = 0; d < 8; ++d) count += b[d + 1] > b[d];
        uint32_t block = size;
        size += 8;
        words[slot] = 1u | ((count & 7u) << 1) | (block << 4);
        for (int d = 7; d >= 0; --d)
            if (b[d + 1] > b[d]) visit(b[d], b[d + 1], level + 1, block + uint32_t(d));
    }
};

uint32_t node_capacity(uint32_t size)
{
    uint32_t cap = 8;
    uint32_t need = size < 9 ? 9 : size;
    while (cap < need) cap *= 2;
    return cap;
}

}  // namespace

extern "C" int ve_octree_build(const ve_insert *ops, size_t n_ops, uint32_t depth, uint32_t max_words,
                               uint32_t **words_out, ve_octree_info *info_out)
{
    ve_clear_global_error();
    if (!words_out || !info_out || (n_ops && !ops)) return ve_global_error(VE_EINVAL, "null argument");
    if (depth < 1 || depth > VE_MAX_DEPTH) return ve_global_error(VE_EINVAL, "depth out of range [1,16]");
    const uint32_t bound = 1u << depth;

    // Keep in-range inserts in call order (insert() ignores the others).
    std::vector<uint64_t> keys;
    std::vector<uint32_t> idx;
    keys.reserve(n_ops);
    idx.reserve(n_ops);
    for (size_t i = 0; i < n_ops; ++i) {
        const ve_insert &o = ops[i];
        if (o.x >= bound || o.y >= bound || o.z >= bound) continue;
        keys.push_back(morton_key(o.x, o.y, o.z, depth));
        idx.push_back(uint32_t(i));
    }
    if (n_ops > 0xFFFFFFFFull) return ve_global_error(VE_EINVAL, "too many inserts");
    radix_sort(keys, idx, int(3 * depth));

    // Unique keys, last insert wins.
    std::vector<uint64_t> ukeys;
    std::vector<uint32_t> umats;
    ukeys.reserve(keys.size());
    umats.reserve(keys.size());
    for (size_t i = 0; i < keys.size(); ++i) {
        if (i + 1 < keys.size() && keys[i + 1] == keys[i]) continue;
        ukeys.push_back(keys[i]);
        umats.push_back(ops[idx[i]].material);
    }
    std::vector<uint64_t>().swap(keys);
    std::vector<uint32_t>().swap(idx);

    if (ukeys.empty()) {
        // No in-range insert: OctreeCPU::root stays null and Octree::set returns
        // before touching the fresh 8-word array (octree.cpp:128, octree.hpp:68).
        uint32_t *words = static_cast<uint32_t *>(std::calloc(8, sizeof(uint32_t)));
        if (!words) return ve_global_error(VE_ENOMEM, "node array allocation");
        info_out->depth = depth;
        info_out->size = 8;
        info_out->capacity = 8;
        info_out->num_voxels = 0;
        *words_out = words;
        return VE_OK;
    }

    // Exact word count before allocating: one block per internal node.
    // Internal nodes at level L (2..depth) = distinct prefixes of L-1 digits.
    uint64_t size64 = 9;
    for (uint32_t level = 2; level <= depth; ++level) {
        const uint32_t sh = 3u * (depth - level + 1);
        uint64_t distinct = 0;
        for (size_t i = 0; i < ukeys.size(); ++i)
            if (i == 0 || (ukeys[i] >> sh) != (ukeys[i - 1] >> sh)) ++distinct;
        size64 += 8 * distinct;
    }
    // next is a 28-bit field (octree.hpp:23): the last block must fit.
    if (size64 - 8 >= (1ull << 28)) return ve_global_error(VE_ECAPACITY, "node array exceeds the 28-bit next field");
    if (max_words && size64 > max_words)
        return ve_global_error(VE_ECAPACITY, "node array exceeds the texture-buffer word limit");
    uint32_t cap = node_capacity(uint32_t(size64));
    if (max_words && cap > max_words) cap = max_words;  // clamp as resizeDataIfNeeded does
    uint32_t *words = static_cast<uint32_t *>(std::calloc(cap, sizeof(uint32_t)));
    if (!words) return ve_global_error(VE_ENOMEM, "node array allocation");

    Builder b{ukeys.data(), umats.data(), depth, words, 9, 0};
    size_t root_b[9];
    uint32_t root_count = 0;
    b.split(0, ukeys.size(), 1, root_b);
    for (int d = 0; d < 8; ++d) root_count += root_b[d + 1] > root_b[d];
    words[0] = 1u | ((root_count & 7u) << 1) | (1u << 4);
    for (int d = 0; d < 8; ++d)  // Octree::set visits the root's children 0..7
            if (root_b[d + 1] > root_b[d]) b.visit(root_b[d], root_b[d + 1], 2, 1u + uint32_t(d));

    info_out->depth = depth;
    info_out->size = b.size;
    info_out->capacity = cap;
    info_out->num_voxels = b.num_voxels;
    *words_out = words;
    return VE_OK;
}

extern "C" void ve_free(void *p) { std::free(p); }
