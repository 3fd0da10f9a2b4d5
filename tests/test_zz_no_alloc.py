"""ve_pass_primary_noalloc, the primary pass that allocates nothing (its jump
table lives in the library's static per-device scratch), against the
VE_PASS_PRIMARY dispatch of ve_pass on the same scene and camera: G-buffer and
positions bit-exact, on two calls in a row, the second with a smaller scene and
framebuffer, both calls through the C ABI on caller-owned device buffers and a
caller-owned stream."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """The caller's side: device buffers and a stream from the HIP runtime the
    library itself is linked against (the soname already loaded with it)."""

    def __init__(self, ve):
        ve.lib()
        self.rt = C.CDLL("libamdhip64.so.7")
        self.owned, self.streams = [], []

    def ck(self, rc, what):
        assert rc == 0, "%s failed: %d" % (what, rc)

    def alloc(self, nbytes, fill=None):
        ptr = C.c_void_p()
        self.ck(self.rt.hipMalloc(C.byref(ptr), C.c_size_t(nbytes)), "hipMalloc")
        self.owned.append(ptr)
        if fill is not None:
            self.ck(self.rt.hipMemset(ptr, C.c_int(fill), C.c_size_t(nbytes)), "hipMemset")
        return ptr

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        ptr = self.alloc(arr.nbytes)
        self.ck(self.rt.hipMemcpy(ptr, C.c_void_p(arr.ctypes.data), C.c_size_t(arr.nbytes), 1), "hipMemcpy H2D")
        return ptr

    def download(self, ptr, shape, dtype=np.uint32):
        out = np.empty(shape, dtype)
        self.ck(self.rt.hipMemcpy(C.c_void_p(out.ctypes.data), ptr, C.c_size_t(out.nbytes), 2), "hipMemcpy D2H")
        return out

    def stream(self):
        s = C.c_void_p()
        self.ck(self.rt.hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.streams.append(s)
        return s

    def sync(self, s):
        self.ck(self.rt.hipStreamSynchronize(s), "hipStreamSynchronize")

    def close(self):
        for s in self.streams:
            self.rt.hipStreamDestroy(s)
        for ptr in self.owned:
            self.rt.hipFree(ptr)
        self.owned, self.streams = [], []


@pytest.fixture
def hip(ve):
    h = _Hip(ve)
    yield h
    h.close()


def _via_context(ve, words, depth, n_words, ubo, W, H, cc):
    r = ve.Renderer(W, H, hash_width=64, max_bounces=1)
    try:
        scene = r.upload(words, depth, n_words)
        try:
            r.set_camera(ubo)
            r.run_pass(scene, "PRIMARY", ve.FrameConfig(primary_controlchecks=cc), ve.FrameInputs(W, H, 0, 1, 1))
            return r.gbuffer(W, H)
        finally:
            r.release(scene)
    finally:
        r.close()


def test_no_alloc_primary_matches_pass(ve, scenes, small_scene, hip):
    depth9, ops9 = scenes.scene_fbm(9, 5)
    words9, info9 = ve.octree_build(ops9, depth9)
    eye, d, fov = scenes.orbit_camera(1 << depth9, 3, 64, 5)
    eye_s, d_s, fov_s = scenes.camera_small()
    cases = [
        # depth 9: the scratch table (depth 6) is shallower than the scene's (depth 8)
        (words9, info9.size, depth9, ve.camera_ubo(eye, d, fov, 640 / 360), 640, 360, 4096),
        # then a smaller scene and framebuffer over the same scratch
        (small_scene["words"], small_scene["info"].size, small_scene["depth"],
         ve.camera_ubo(eye_s, d_s, fov_s, 97 / 61), 97, 61, 300),
    ]
    stream = hip.stream()
    for k, (words, n, depth, ubo, W, H, cc) in enumerate(cases):
        print("case no-alloc primary", k, "depth", depth, "words", n, "size", W, H, flush=True)
        ref_g, ref_p = _via_context(ve, words, depth, n, ubo, W, H, cc)
        words_dev = hip.upload(np.ascontiguousarray(words[:n], np.uint32))
        g_dev = hip.alloc(W * H * 16, fill=0xFF)   # every pixel must be written
        p_dev = hip.alloc(W * H * 4, fill=0xFF)
        ve.primary_noalloc(words_dev, n, depth, ubo, W, H, cc, g_dev, p_dev, stream)
        hip.sync(stream)
        g = hip.download(g_dev, (H, W, 4))
        p = hip.download(p_dev, (H, W))
        assert (ref_g[..., 3] != 0).sum() > W * H // 10   # the view hits geometry
        assert np.array_equal(g, ref_g)
        assert np.array_equal(p, ref_p)


def test_no_alloc_primary_rejects_bad_arguments(ve, hip):
    w = hip.alloc(64, fill=0)
    g = hip.alloc(256, fill=0)
    ubo = np.zeros(16, np.float32)
    with pytest.raises(ve.VeError) as e:
        ve.primary_noalloc(w, 16, 3, ubo, 16385, 1, 8, g, g)
    assert e.value.code == ve.VE_EINVAL and "16384" in str(e.value)
    with pytest.raises(ve.VeError) as e:
        ve.primary_noalloc(w, 16, 17, ubo, 4, 4, 8, g, g)
    assert e.value.code == ve.VE_EINVAL
    with pytest.raises(ve.VeError) as e:
        ve.primary_noalloc(None, 16, 3, ubo, 4, 4, 8, g, g)
    assert e.value.code == ve.VE_EINVAL
