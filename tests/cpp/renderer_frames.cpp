// ve::Renderer as the reference's main loop drives Renderer::run
// (src/voxelengine.cpp:158-290): a demo-like scene built through
// ve::OctreeCPU / ve::Octree, a material pool, a constant sky, then
// `frames` frames of run() with a moving camera.  Writes the last frame's
// RGBA32F output and RGBA8 display image, and the node words, to `out`.
//   renderer_frames <out-file> <frames>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ve_svo.hpp"

int main(int argc, char **argv)
{
    if (argc < 3) return 2;
    const int frames = std::atoi(argv[2]);
    const uint32_t depth = 6, L = 1u << depth, W = 96, H = 64;
    ve::OctreeCPU scene(depth);
    const float c1[3] = {L / 3.0f, L / 3.0f - 5.0f, L / 2.0f};
    scene.insertSphere(c1, L / 4.0f, 5);
    const uint32_t b0[3] = {0, 0, 0}, b1[3] = {L, 4, L};
    scene.insertBox(b0, b1, 4);
    const uint32_t f0[3] = {0, 0, 0}, f1[3] = {L, L, L};
    scene.insertFunction([](const float p[3]) { return p[1] > 50.0f && p[0] < 10.0f; }, f0, f1, 2);
    ve::Octree octree(depth);
    octree.set(scene);

    // std140 material entries (material.hpp:17-25): 0 unused, 2/4/5 diffuse
    struct Mat {
        float color[4], specular_color[4], roughness, specular, metallic;
        int32_t emissive;
        float intensity, pad[3];
    };
    std::vector<Mat> mats(6);
    std::memset(mats.data(), 0, mats.size() * sizeof(Mat));
    const float albedo[6][3] = {{0, 0, 0}, {.5f, .5f, .5f}, {.8f, .3f, .2f}, {.5f, .5f, .5f}, {.3f, .7f, .3f},
                                {.2f, .4f, .9f}};
    for (int i = 1; i < 6; i++) {
        for (int k = 0; k < 3; k++) mats[i].color[k] = albedo[i][k];
        for (int k = 0; k < 4; k++) mats[i].specular_color[k] = 1.0f;
        mats[i].roughness = 1.0f;
    }

    ve::Renderer r(W, H, 3);
    r.setOctree(octree);
    r.setMaterials(mats.data(), (uint32_t)mats.size());
    const uint8_t sky[3] = {140, 180, 230};
    r.setSky(sky);
    // the reference's FrameConfig with a still camera's accumulation
    // settings and one sample per visible voxel (order-independent results)
    ve_frame_config fc = ve::default_frame_config();
    fc.bounce_controlchecks = 120;
    fc.normal_precision = 1;
    fc.shading_budget = 196608;
    fc.weight_pixels = 0.0f;
    fc.weight_variance = 0.0f;
    fc.top_fraction = 0.0f;
    fc.mid_fraction = 1.0f;
    fc.sample_cap_direct = 1024;
    fc.sample_cap_indirect = 4096;
    fc.disable_stale_reset = 1;
    fc.update_time = 0xFFFFFFFFu;
    fc.firefly_floor = 1e30f;
    for (int f = 0; f < frames; f++) {
        const float eye[3] = {-20.0f + 3.0f * f, 70.0f, -30.0f}, dir[3] = {0.48f, -0.35f, 0.80f};
        r.setCamera(ve::camera_ubo(eye, dir, 70.0f, (float)W / H));
        if (!r.run(&fc, (uint32_t)(f + 1))) {
            std::fprintf(stderr, "run failed: %s\n", r.lastError());
            return 1;
        }
    }
    const std::vector<float> out = r.output();
    const std::vector<uint8_t> disp = r.display();
    FILE *fh = std::fopen(argv[1], "wb");
    if (!fh) return 1;
    const uint32_t hdr[4] = {W, H, octree.size, (uint32_t)frames};
    std::fwrite(hdr, sizeof hdr, 1, fh);
    std::fwrite(octree.data.data(), 4, octree.size, fh);
    std::fwrite(out.data(), sizeof(float), out.size(), fh);
    std::fwrite(disp.data(), 1, disp.size(), fh);
    std::fclose(fh);
    std::printf("ok %u words\n", octree.size);
    return 0;
}
