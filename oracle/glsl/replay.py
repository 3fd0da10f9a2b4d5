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

    def __init__(self, words, depth, W, H, max_bounces=8, materials=None, sky_rgb8=(0, 0, 0)):
        self.gl = G.GL.get()
        self.W, self.H, self.depth = W, H, depth
        self.max_bounces = max_bounces
        defines = {"MAX_BOUNCES": "%du" % max_bounces}
        self.progs = {}
        for name in ("generate_primary.comp", "normal.comp", "schedule.comp", "rank.comp", "threshold.comp",
                     "emit.comp", "trace.comp", "shade.comp", "resolve.comp"):
            self.progs[name] = None
        self.defines = defines
        self.hash_w = self.gl.integer(0x0D33)   # GL_MAX_TEXTURE_SIZE (renderer.cpp:27)
        self.tbo = G.TextureBuffer(np.ascontiguousarray(words, np.uint32))
        self.cam_ubo = G.Buffer(nbytes=64)
        if materials is None:
            materials = np.zeros(128 * 64, np.uint8)
        mat = np.zeros(128 * 64, np.uint8)
        mb = np.frombuffer(np.ascontiguousarray(materials).tobytes(), np.uint8)
        mat[:len(mb)] = mb
        self.mat_ubo = G.Buffer(data=mat)
        face = np.array(sky_rgb8, np.uint8).reshape(1, 1, 3)
        self.sky = G.CubeMap([face] * 6, 1)
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
