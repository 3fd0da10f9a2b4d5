"""Replay of Renderer::runWavefrontFrame on Mesa llvmpipe (test infrastructure).

Loads the reference's GLSL compute passes from the reference tree at run time
(default /root/reference/src/shd, or $VE_REF_SHADERS) and dispatches them with
the bindings, uniforms and order of src/renderer/renderer.cpp:282-649.  The
wall-clock inputs of the reference (glTime, frameSeed, frameID) are explicit
arguments.  Run with LP_NUM_THREADS=1 for a deterministic dispatch order.

Nothing here is product code and none of the reference's sources are copied:
the shaders are read from the reference checkout when a golden fixture is made.
"""
import math
import os

import numpy as np

from . import gl as G

SHADER_DIR = os.environ.get("VE_REF_SHADERS", "/root/reference/src/shd")

# renderer.hpp:90-103
LBUFFER_SLOTS = 32
NBUFFER_SLOTS = 32
RBUFFER_SLOTS = 16
RBUFFER_STRIDE = 8
RAY_CAPACITY_HOST = 49152
SHADE_BUDGET_MAX = 196608
CLAIM_WIDTH = 16384
CLAIM_PROBES = 64
EMISSIVE_CAPACITY_HOST = 16384
EMISSIVE_CLAIM_W = 8192
EMISSIVE_CLAIM_PROBES = 8
CANDIDATE_CAPACITY = 1048576
RAY_STRIDE = 96


def shaders_available():
    return os.path.exists(os.path.join(SHADER_DIR, "generate_primary.comp"))


def _prog(name, defines=None):
    return G.Program(G.load_source(os.path.join(SHADER_DIR, name), defines), name)


class GLPipeline:
    """The reference renderer's GPU state on llvmpipe for one scene."""

    def __init__(self, words, depth, W, H, max_bounces=8, materials=None, sky_rgb8=(0, 0, 0), sky_faces=None,
                 hash_width=None):
        """hash_width: the width of the l/n/rBuffer images and of the
        lBufferWidth / nBufferWidth / rBufferWidth uniforms, which the
        reference takes from GL_MAX_TEXTURE_SIZE (renderer.cpp:27-40); None =
        this driver's value.  A smaller width stands in for a driver with a
        smaller limit (every bucket then holds more voxels)."""
        self.gl = G.GL.get()
        self.W, self.H, self.depth = W, H, depth
        self.max_bounces = max_bounces
        defines = {"MAX_BOUNCES": "%du" % max_bounces}
        self.progs = {}
        for name in ("generate_primary.comp", "normal.comp", "schedule.comp", "rank.comp", "threshold.comp",
                     "emit.comp", "trace.comp", "shade.comp", "resolve.comp"):
            self.progs[name] = None
        self.defines = defines
        self.hash_w = self.gl.integer(0x0D33) if hash_width is None else int(hash_width)   # renderer.cpp:27
        self.tbo = G.TextureBuffer(np.ascontiguousarray(words, np.uint32))
        self.cam_ubo = G.Buffer(nbytes=64)
        if materials is None:
            materials = np.zeros(128 * 64, np.uint8)
        mat = np.zeros(128 * 64, np.uint8)
        mb = np.frombuffer(np.ascontiguousarray(materials).tobytes(), np.uint8)
        mat[:len(mb)] = mb
        self.mat_ubo = G.Buffer(data=mat)
        if sky_faces is None:   # constant sky: one texel per face
            face = np.array(sky_rgb8, np.uint8).reshape(1, 1, 3)
            self.sky = G.CubeMap([face] * 6, 1)
        else:                   # six (edge, edge, 3) RGB8 faces, +X -X +Y -Y +Z -Z
            faces = np.asarray(sky_faces, np.uint8)
            self.sky = G.CubeMap([faces[i] for i in range(6)], faces.shape[1])
        # screen images (framebufferEvent)
        self.gbuf = G.Image2D(W, H, G.GL_RGBA32UI)
        self.gpos = G.Image2D(W, H, G.GL_R32UI)
        self.out = G.Image2D(W, H, G.GL_RGBA32F)
        # hash images (renderer.cpp:114-161)
        self.nbuf = G.Image2D(self.hash_w, 3 * NBUFFER_SLOTS, G.GL_R32UI)
        self.lbuf = G.Image2D(self.hash_w, 12 * LBUFFER_SLOTS, G.GL_R32UI)
        self.rbuf = G.Image2D(self.hash_w, RBUFFER_STRIDE * RBUFFER_SLOTS, G.GL_R32UI)
        # wavefront buffers (initWavefrontBuffers, renderer.cpp:174-269)
        self.ring = G.Buffer(nbytes=RAY_CAPACITY_HOST * RAY_STRIDE)
        self.ring_meta = G.Buffer(nbytes=16)
        self.shade_list = G.Buffer(nbytes=SHADE_BUDGET_MAX * 8)
        self.shade_meta = G.Buffer(nbytes=16)
        self.emissives = G.Buffer(nbytes=EMISSIVE_CAPACITY_HOST * 32)
        self.emissive_meta = G.Buffer(nbytes=16)
        self.candidates = G.Buffer(nbytes=CANDIDATE_CAPACITY * 16)
        self.cand_meta = G.Buffer(nbytes=16)
        stats = np.zeros(264, np.uint32)
        stats[0] = 0x7F800000   # +inf (rank.comp:109-111 frame-0 fallback)
        stats[1] = 0xFF800000   # -inf
        self.rank_stats = G.Buffer(data=stats)
        self.claim = G.Image2D(CLAIM_WIDTH, CLAIM_PROBES, G.GL_R32UI)
        self.samples = G.Image2D(CLAIM_WIDTH, CLAIM_PROBES, G.GL_R32UI)
        self.emissive_claim = G.Image2D(EMISSIVE_CLAIM_W, EMISSIVE_CLAIM_PROBES, G.GL_R32UI)
        self.frame_id = 0
        self.stats = {}

    def prog(self, name):
        if self.progs.get(name) is None:
            self.progs[name] = _prog(name, self.defines)
            for block, binding in (("CameraUniform", 0), ("MaterialUniform", 1)):
                self.progs[name].bind_block(block, binding)
        return self.progs[name]

    def set_camera(self, ubo):
        self.cam_ubo.write(np.ascontiguousarray(ubo, np.float32))

    def _bind_common(self, p):
        self.cam_ubo.bind_ubo(0)
        self.mat_ubo.bind_ubo(1)
        self.tbo.bind(p, "octreeTexture", 0)
        p.set1ui("octreeDepth", self.depth)
        self.sky.bind(p, "skybox", 1)

    def groups2d(self):
        return math.ceil(self.W / 8.0), math.ceil(self.H / 8.0)

    # --- 1. generate_primary.comp (renderer.cpp:299-317)
    def primary(self, controlchecks):
        p = self.prog("generate_primary.comp")
        p.use()
        self._bind_common(p)
        self.gbuf.bind_image(0)
        self.gpos.bind_image(1)
        p.set2i("screenResolution", self.W, self.H)
        p.set1ui("controlchecks", controlchecks)
        p.dispatch(*self.groups2d())

    def read_gbuffer(self):
        return self.gbuf.read(), self.gpos.read()

    # --- 2. normal.comp (renderer.cpp:319-341)
    def normal(self, fc, time):
        p = self.prog("normal.comp")
        p.use()
        self.gbuf.bind_image(0)
        self.lbuf.bind_image(1)
        self.gpos.bind_image(2)
        self.nbuf.bind_image(3)
        self.tbo.bind(p, "octreeTexture", 0)
        p.set1ui("octreeDepth", self.depth)
        p.set1i("nBufferWidth", self.hash_w)
        p.set1i("nBufferSlots", NBUFFER_SLOTS)
        p.set1i("lBufferSlots", LBUFFER_SLOTS)
        p.set1i("lBufferWidth", self.hash_w)
        p.set1i("normalPrecision", fc["normal_precision"])
        p.set2i("screenResolution", self.W, self.H)
        p.set1ui("time", time)
        p.set1ui("updateTime", fc["update_time"])
        p.dispatch(*self.groups2d())

    # --- per-frame reset (renderer.cpp:343-375)
    def reset(self):
        z4 = np.zeros(4, np.uint32)
        self.ring_meta.write(z4)
        self.shade_meta.write(z4)
        self.cand_meta.write(z4)
        self.rank_stats.write(np.array([0x7F7FFFFF, 0, 0, 0], np.uint32), offset=16)
        self.rank_stats.write(np.zeros(256, np.uint32), offset=32)
        self.claim.zero()
        self.samples.zero()

    def _bind_ssbos(self):
        for i, b in enumerate((self.ring, self.ring_meta, self.shade_list, self.shade_meta, self.emissives,
                               self.emissive_meta, self.candidates, self.cand_meta, self.rank_stats)):
            b.bind_ssbo(i)

    # --- 3. schedule.comp (renderer.cpp:388-415)
    def schedule(self, frame_seed):
        self._bind_ssbos()
        p = self.prog("schedule.comp")
        p.use()
        self.gbuf.bind_image(0)
        self.claim.bind_image(1)
        self.samples.bind_image(2)
        p.set2i("screenResolution", self.W, self.H)
        p.set1i("claimWidth", CLAIM_WIDTH)
        p.set1i("claimProbes", CLAIM_PROBES)
        p.set1ui("frameSeed", frame_seed)
        p.dispatch(*self.groups2d())
        return min(int(self.cand_meta.read()[0]), CANDIDATE_CAPACITY)

    # --- 3b. rank.comp + threshold.comp (renderer.cpp:427-470)
    def rank_threshold(self, fc, cand_count, budget):
        p = self.prog("rank.comp")
        p.use()
        self.lbuf.bind_image(0)
        self.samples.bind_image(1)
        p.set1i("lBufferWidth", self.hash_w)
        p.set1i("lBufferSlots", LBUFFER_SLOTS)
        p.set1f("wPixels", fc["weight_pixels"])
        p.set1f("wVariance", fc["weight_variance"])
        p.set1f("pixelsNorm", fc["pixels_norm"])
        p.set1f("varianceNorm", fc["variance_norm"])
        p.set1i("minSamples", fc["min_samples"])
        p.set1ui("shadeBudget", budget)
        p.dispatch((cand_count + 63) // 64)
        t = self.prog("threshold.comp")
        t.use()
        t.set1f("topFraction", fc["top_fraction"])
        t.set1f("midFraction", fc["mid_fraction"])
        t.dispatch(1)
        return int(self.shade_meta.read()[0])

    # --- 4+5. emit / trace / shade chunk loop (renderer.cpp:476-606)
    def bounces(self, fc, scheduled, frame_seed, frame_idx, time, count_live=False):
        """Returns (iterations run, rays still bouncing after their trace
        dispatch, rays in all trace dispatch windows, and -- with count_live --
        rays that each trace dispatch traverses: its window's rays whose
        bounce is below MAX_BOUNCES, trace.comp:36)."""
        bounces_run = 0
        traced = 0
        window = 0
        live_in = 0
        for base in range(0, scheduled, RAY_CAPACITY_HOST):
            size = min(scheduled - base, RAY_CAPACITY_HOST)
            self.ring_meta.write(np.zeros(4, np.uint32))
            e = self.prog("emit.comp")
            e.use()
            self.gbuf.bind_image(0)
            e.set2i("screenResolution", self.W, self.H)
            e.set1ui("chunkBase", base)
            e.set1ui("chunkSize", size)
            e.set1ui("frameSeed", frame_seed)
            e.set1ui("frameIdx", frame_idx)
            e.set1ui("octreeDepth_", self.depth)
            e.dispatch((size + 63) // 64)
            for b in range(8):
                head, tail = [int(v) for v in self.ring_meta.read()[:2]]
                if head == tail:
                    break
                win = head - tail
                window += win
                if count_live:
                    ring = self.ring.read(np.uint32).reshape(-1, 24)
                    live_in += int((ring[(tail + np.arange(win)) % RAY_CAPACITY_HOST, 19] < self.max_bounces).sum())
                tr = self.prog("trace.comp")
                tr.use()
                self.tbo.bind(tr, "octreeTexture", 0)
                tr.set1ui("octreeDepth", self.depth)
                tr.set1ui("controlchecks", fc["bounce_controlchecks"])
                tr.set1ui("rayWindowBase", tail)
                tr.set1ui("rayWindowSize", win)
                tr.dispatch((win + 63) // 64)
                live = self.ring.read(np.uint32).reshape(-1, 24)
                idx = (tail + np.arange(win)) % RAY_CAPACITY_HOST
                traced += int((live[idx, 19] < self.max_bounces).sum())
                sh = self.prog("shade.comp")
                sh.use()
                self.lbuf.bind_image(0)
                self.nbuf.bind_image(1)
                self.emissive_claim.bind_image(2)
                self.rbuf.bind_image(3)
                self.tbo.bind(sh, "octreeTexture", 0)
                sh.set1ui("octreeDepth", self.depth)
                self.sky.bind(sh, "skybox", 1)
                self.mat_ubo.bind_ubo(1)
                sh.set1ui("controlchecks", fc["bounce_controlchecks"])
                sh.set1ui("rayWindowBase", tail)
                sh.set1ui("rayWindowSize", win)
                sh.set1i("lBufferWidth", self.hash_w)
                sh.set1i("lBufferSlots", LBUFFER_SLOTS)
                sh.set1i("nBufferWidth", self.hash_w)
                sh.set1i("nBufferSlots", NBUFFER_SLOTS)
                sh.set1ui("time", time)
                sh.set1ui("updateTime", fc["update_time"])
                sh.set1ui("SAMPLE_CAP_DIRECT", fc["sample_cap_direct"])
                sh.set1ui("SAMPLE_CAP_INDIRECT", fc["sample_cap_indirect"])
                sh.set1f("fireflyK", fc["firefly_k"])
                sh.set1f("fireflyFloor", fc["firefly_floor"])
                sh.set1i("disableStaleReset", fc["disable_stale_reset"])
                sh.set1i("rBufferWidth", self.hash_w)
                sh.set1i("rBufferSlots", RBUFFER_SLOTS)
                sh.set1i("restirEnabled", fc["restir_enabled"])
                sh.set1i("restirM", fc["restir_m"])
                sh.set1i("restirTemporalEnabled", fc["restir_temporal"])
                sh.set1i("restirSpatialEnabled", fc["restir_spatial"])
                sh.set1i("restirSpatialNeighbors", fc["restir_spatial_neighbors"])
                sh.set1f("restirSpatialRadius", fc["restir_spatial_radius"])
                sh.set1i("restirMaxM", fc["restir_max_m"])
                sh.dispatch((win + 63) // 64)
                self.ring_meta.write(np.array([head], np.uint32), offset=4)
                bounces_run = max(bounces_run, b + 1)
        return bounces_run, traced, window, live_in

    # --- 6. resolve.comp (renderer.cpp:612-648)
    def resolve(self, fc):
        p = self.prog("resolve.comp")
        p.use()
        self.lbuf.bind_image(0)
        self.gbuf.bind_image(1)
        self.out.bind_image(2)
        self.lbuf.bind_image(3)
        self.nbuf.bind_sampler(p, "nBuffer", 0)
        self.sky.bind(p, "skybox", 1)
        self.cam_ubo.bind_ubo(0)
        self.mat_ubo.bind_ubo(1)
        p.set1i("nBufferWidth", self.hash_w)
        p.set1i("nBufferSlots", NBUFFER_SLOTS)
        p.set1i("lBufferWidth", self.hash_w)
        p.set1i("lBufferSlots", LBUFFER_SLOTS)
        p.set2i("screenResolution", self.W, self.H)
        p.set1i("renderType", fc["render_type"])
        p.set1ui("primary_controlchecks", fc["primary_controlchecks"])
        p.set1i("reconstructionRadius", fc["reconstruction_radius"])
        p.set1f("varianceNormForView", fc["variance_norm"])
        p.dispatch(*self.groups2d())

    def frame(self, fc, frame_idx, frame_seed, time, count_live=False):
        """One Renderer::runWavefrontFrame; returns the per-frame counters."""
        budget = max(1, min(int(fc["shading_budget"]), SHADE_BUDGET_MAX))
        self.primary(fc["primary_controlchecks"])
        self.normal(fc, time)
        self.reset()
        cand = self.schedule(frame_seed)
        scheduled = 0
        if cand > 0:
            scheduled = self.rank_threshold(fc, cand, budget)
        scheduled = min(scheduled, budget)
        bounces_run, traced, window, live_in = self.bounces(fc, scheduled, frame_seed, frame_idx, time, count_live)
        self.resolve(fc)
        return dict(candidates=cand, scheduled=scheduled, bounces=bounces_run, traced=traced, window=window,
                    live_in=live_in, emissives=int(self.emissive_meta.read()[0]))

    def read_hash(self, img, height):
        """Image (x = bucket, y = field) -> bucket-major [bucket][field]."""
        return img.read().reshape(height, self.hash_w).T.copy()
