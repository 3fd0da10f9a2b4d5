// ve_svo.hpp — C++ host classes over the C ABI, shaped like VoxelEngine3.0's
// host API so calling code can switch by name:
//
//   ve::OctreeCPU   OctreeCPU::insert / remove / lookup / raycast / insertBox /
//                   insertSphere / insertFunction (src/octree_cpu.hpp:30-40)
//   ve::Octree      Octree::set / applyEdits / findGpuSlot (src/renderer/octree.hpp:50-59)
//   ve::camera_ubo  Camera::UpdateUBO (src/renderer/camera.cpp:39-50)
//   ve::Renderer    Renderer(...) + bool run(FrameConfig*) (src/renderer/renderer.hpp:11-13)
//
// Header-only; link with libve_svo.so.
#pragma once

#include <cmath>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "ve_svo.h"

namespace ve {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const ve_ctx *ctx = nullptr)
{
    if (rc != VE_OK) throw Error(rc, ve_last_error(ctx));
}

class OctreeCPU {
public:
    // OctreeCPU(depth) (octree_cpu.cpp:4-15): a persistent pointer tree
    explicit OctreeCPU(uint8_t depth_, uint32_t max_words = 1u << 27) : depth(depth_)
    {
        check(ve_octree_cpu_create(&h_, depth_, max_words));
    }
    ~OctreeCPU() { ve_octree_cpu_destroy(h_); }
    OctreeCPU(const OctreeCPU &) = delete;
    OctreeCPU &operator=(const OctreeCPU &) = delete;

    struct RayHit {
        bool hit = false;
        uint32_t position[3] = {0, 0, 0};
        uint32_t material = 0;
    };

    // OctreeCPU::insert (octree_cpu.cpp:40-48): out-of-range positions are
    // ignored; the last insert of a position wins.
    void insert(uint32_t x, uint32_t y, uint32_t z, uint32_t material)
    {
        const ve_insert op{x, y, z, material};
        check(ve_octree_cpu_insert(h_, &op, 1));
        inserts++;
    }
    // OctreeCPU::remove (octree_cpu.cpp:66-72)
    void remove(uint32_t x, uint32_t y, uint32_t z)
    {
        const uint32_t p[3] = {x, y, z};
        check(ve_octree_cpu_remove(h_, p, 1));
    }
    // OctreeCPU::lookup (octree_cpu.cpp:100-111): true when a leaf is present
    bool lookup(uint32_t x, uint32_t y, uint32_t z, uint32_t *material = nullptr) const
    {
        int present = 0;
        uint32_t m = 0;
        check(ve_octree_cpu_lookup(h_, x, y, z, &present, &m));
        if (material) *material = m;
        return present != 0;
    }
    // OctreeCPU::raycast (octree_cpu.cpp:128-215)
    RayHit raycast(const float origin[3], const float direction[3], uint32_t maxSteps = 300) const
    {
        RayHit h;
        int hit = 0;
        check(ve_octree_cpu_raycast(h_, origin, direction, maxSteps, &hit, h.position, &h.material));
        h.hit = hit != 0;
        return h;
    }

    // octree_cpu.cpp:221-230
    void insertBox(const uint32_t mn[3], const uint32_t mx_in[3], uint32_t material)
    {
        const uint32_t bound = 1u << depth;
        uint32_t mx[3];
        for (int i = 0; i < 3; i++) mx[i] = mx_in[i] < bound ? mx_in[i] : bound;
        std::vector<ve_insert> ops;
        for (uint32_t x = mn[0]; x < mx[0]; x++)
            for (uint32_t y = mn[1]; y < mx[1]; y++)
                for (uint32_t z = mn[2]; z < mx[2]; z++) ops.push_back({x, y, z, material});
        flush(ops);
    }

    // octree_cpu.cpp:232-248, FP32 distance test as glm::distance
    void insertSphere(const float c[3], float r, uint32_t material)
    {
        const int32_t bound = 1 << depth;
        int32_t lo[3], hi[3];
        for (int i = 0; i < 3; i++) {
            lo[i] = std::max<int32_t>(0, (int32_t)std::floor(c[i] - r));
            hi[i] = std::min<int32_t>(bound - 1, (int32_t)std::ceil(c[i] + r));
        }
        std::vector<ve_insert> ops;
        for (int32_t x = lo[0]; x <= hi[0]; x++)
            for (int32_t y = lo[1]; y <= hi[1]; y++)
                for (int32_t z = lo[2]; z <= hi[2]; z++) {
                    const float dx = ((float)x + 0.5f) - c[0], dy = ((float)y + 0.5f) - c[1],
                                dz = ((float)z + 0.5f) - c[2];
                    const float t0 = dx * dx, t1 = dy * dy, t2 = dz * dz;
                    if (std::sqrt((t0 + t1) + t2) <= r) ops.push_back({(uint32_t)x, (uint32_t)y, (uint32_t)z, material});
                }
        flush(ops);
    }

    // octree_cpu.cpp:250-262: fn is evaluated at voxel centres
    void insertFunction(const std::function<bool(const float p[3])> &fn, const uint32_t mn[3], const uint32_t mx_in[3],
                        uint32_t material)
    {
        const uint32_t bound = 1u << depth;
        uint32_t mx[3];
        for (int i = 0; i < 3; i++) mx[i] = mx_in[i] < bound ? mx_in[i] : bound;
        std::vector<ve_insert> ops;
        for (uint32_t x = mn[0]; x < mx[0]; x++)
            for (uint32_t y = mn[1]; y < mx[1]; y++)
                for (uint32_t z = mn[2]; z < mx[2]; z++) {
                    const float p[3] = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
                    if (fn(p)) ops.push_back({x, y, z, material});
                }
        flush(ops);
    }

    ve_octree_cpu *handle() const { return h_; }
    uint8_t depth;
    size_t inserts = 0;   // insert calls so far

private:
    void flush(const std::vector<ve_insert> &ops)
    {
        check(ve_octree_cpu_insert(h_, ops.data(), ops.size()));
        inserts += ops.size();
    }
    ve_octree_cpu *h_ = nullptr;
};

class Octree {
public:
    explicit Octree(uint8_t depth_) : depth(depth_) {}

    // Octree::set (octree.cpp:127-166): the whole node image
    void set(OctreeCPU &cpu)
    {
        ve_octree_info info{};
        check(ve_octree_set(cpu.handle(), &info));
        pull(cpu, info);
    }
    // Octree::applyEdits (octree.cpp:228-315): only the dirty paths; the
    // written slots are in `changed` (or `full` when everything must travel)
    void applyEdits(OctreeCPU &cpu)
    {
        ve_octree_info info{};
        check(ve_octree_apply_edits(cpu.handle(), &info));
        pull(cpu, info);
    }
    // Octree::findGpuSlot (octree.cpp:322-341); UINT32_MAX when absent
    uint32_t findGpuSlot(const OctreeCPU &cpu, uint32_t x, uint32_t y, uint32_t z) const
    {
        return ve_octree_find_slot(cpu.handle(), x, y, z);
    }

    uint8_t depth;
    std::vector<uint32_t> data;
    uint32_t size = 8, capacity = 8, numVoxels = 0;
    std::vector<uint32_t> changed;
    bool full = true;

private:
    void pull(OctreeCPU &cpu, const ve_octree_info &info)
    {
        const uint32_t *w = nullptr;
        uint32_t cap = 0;
        check(ve_octree_words(cpu.handle(), &w, &cap));
        data.assign(w, w + cap);
        const uint32_t *sl = nullptr;
        size_t n = 0;
        int f = 0;
        check(ve_octree_changed(cpu.handle(), &sl, &n, &f));
        changed.assign(sl, sl + n);
        full = f != 0;
        size = info.size;
        capacity = info.capacity;
        numVoxels = info.num_voxels;
    }
};

struct CameraUBO {
    float v[16];
};

inline CameraUBO camera_ubo(const float position[3], const float direction[3], float fov_deg, float aspect)
{
    CameraUBO u{};
    ve_camera_ubo(position, direction, fov_deg, aspect, u.v);
    return u;
}

inline ve_frame_config default_frame_config()
{
    ve_frame_config fc;
    ve_frame_config_defaults(&fc);
    return fc;
}

class Renderer {
public:
    Renderer(uint32_t width, uint32_t height, uint32_t max_bounces = 8, int device = 0, uint32_t hash_width = 16384)
    {
        ve_renderer_config cfg{device, width, height, hash_width, max_bounces};
        check(ve_create(&ctx_, &cfg));
        width_ = width;
        height_ = height;
    }
    ~Renderer()
    {
        if (scene_) ve_scene_release(ctx_, scene_);
        ve_destroy(ctx_);
    }
    Renderer(const Renderer &) = delete;
    Renderer &operator=(const Renderer &) = delete;

    void setOctree(const Octree &o)
    {
        if (scene_) check(ve_scene_release(ctx_, scene_), ctx_);
        scene_ = nullptr;
        check(ve_scene_upload(ctx_, o.data.data(), o.size, o.depth, &scene_), ctx_);
    }
    // After Octree::applyEdits: only the changed words travel (all of them
    // when the image was resized or rebuilt).
    void updateOctree(const Octree &o)
    {
        if (!scene_ || o.full) {
            setOctree(o);
            return;
        }
        check(ve_scene_update(ctx_, scene_, o.data.data(), o.size, o.changed.data(), o.changed.size()), ctx_);
    }
    void setCamera(const CameraUBO &u) { check(ve_set_camera(ctx_, u.v), ctx_); }
    // Cross-frame overlap of the primary pass (on by default; off = the
    // reference's strict pass order, same results).
    void setOverlap(bool on) { check(ve_set_overlap(ctx_, on ? 1 : 0), ctx_); }
    void setMaterials(const void *std140, uint32_t n) { check(ve_set_materials(ctx_, std140, n), ctx_); }
    void setSky(const uint8_t rgb[3])
    {
        const uint8_t *faces[6] = {rgb, rgb, rgb, rgb, rgb, rgb};
        check(ve_set_sky(ctx_, faces, 1), ctx_);
    }
    // Cube-map sky: six faces (+X -X +Y -Y +Z -Z) of edge x edge RGB8 texels.
    void setSkyCube(const uint8_t *const faces[6], int edge) { check(ve_set_sky(ctx_, faces, edge), ctx_); }

    // Renderer::run (renderer.cpp:654-712): one wavefront frame, then the
    // final blit's tonemap into the RGBA8 display image; returns false (and
    // keeps the message) on error, like the reference's `success` flag.
    bool run(const ve_frame_config *fc, uint32_t time)
    {
        ve_frame_inputs in{width_, height_, frameID, frameSeed(), time};
        int rc = ve_frame(ctx_, scene_, fc, &in);
        if (rc == VE_OK) rc = ve_pass(ctx_, scene_, VE_PASS_DISPLAY, fc, &in);
        frameID++;
        return rc == VE_OK;
    }

    // The displayed frame: RGBA8, row-major (what final.frag writes).
    std::vector<uint8_t> display()
    {
        std::vector<uint8_t> out((size_t)width_ * height_ * 4);
        check(ve_read(ctx_, VE_BUF_DISPLAY, out.data(), out.size()), ctx_);
        return out;
    }

    std::vector<float> output()
    {
        std::vector<float> out((size_t)width_ * height_ * 4);
        check(ve_read(ctx_, VE_BUF_OUTPUT, out.data(), out.size() * sizeof(float)), ctx_);
        return out;
    }

    const char *lastError() const { return ve_last_error(ctx_); }
    ve_ctx *handle() { return ctx_; }
    uint32_t frameID = 0;

private:
    uint32_t frameSeed() const { return frameID * 2654435761u + 1u; }
    ve_ctx *ctx_ = nullptr;
    ve_scene *scene_ = nullptr;
    uint32_t width_ = 0, height_ = 0;
};

}  // namespace ve
