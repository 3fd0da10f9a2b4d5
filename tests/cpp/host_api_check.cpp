// Host-side C++ API check (no device calls): builds a small scene through
// ve::OctreeCPU / ve::Octree exactly as the reference's demo builds one
// (src/voxelengine.cpp:129-144) and prints the node-array summary.
#include <cstdio>

#include "ve_svo.hpp"

int main()
{
    ve::OctreeCPU scene(6);
    const uint32_t L = 1u << 6;
    const float c1[3] = {L / 3.0f, L / 3.0f - 5.0f, L / 2.0f};
    scene.insertSphere(c1, L / 4.0f, 5);
    const uint32_t b0[3] = {0, 0, 0}, b1[3] = {L, 4, L};
    scene.insertBox(b0, b1, 4);
    const uint32_t f0[3] = {0, 0, 0}, f1[3] = {L, L, L};
    scene.insertFunction([](const float p[3]) { return p[1] > 50.0f && p[0] < 10.0f; }, f0, f1, 2);
    ve::Octree octree(6);
    octree.set(scene);
    uint64_t sum = 0;
    for (uint32_t i = 0; i < octree.size; i++) sum = sum * 1000003ull + octree.data[i];
    std::printf("%zu %u %u %u %llu\n", scene.inserts, octree.size, octree.capacity, octree.numVoxels,
                (unsigned long long)sum);
    const ve_frame_config fc = ve::default_frame_config();
    std::printf("%d %d\n", fc.primary_controlchecks, fc.shading_budget);
    return 0;
}
