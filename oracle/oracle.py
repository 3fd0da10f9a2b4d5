"""ctypes binding of the C oracle (oracle/libsvo_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.
Only tests/, the smoke check and the benchmark's CPU-baseline leg use it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsvo_oracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        p, u32 = C.c_void_p, C.c_uint32
        L.so_octree_build.restype = C.c_int
        L.so_octree_build.argtypes = [p, C.c_size_t, u32, u32, C.POINTER(C.POINTER(C.c_uint32))] + \
            [C.POINTER(C.c_uint32)] * 3
        L.so_free.argtypes = [p]
        L.so_camera_ubo.argtypes = [p, p, C.c_float, C.c_float, p]
        L.so_set_fp_variant.argtypes = [C.c_int]
        L.so_fp_probe.argtypes = [C.c_int, p, p, C.c_size_t]
        L.so_voxel_key.restype = u32
        L.so_voxel_key.argtypes = [p]
        L.so_generate_primary.argtypes = [p, u32, u32, p, u32, u32, u32, p, p, C.POINTER(C.c_uint64),
                                          C.POINTER(C.c_uint64), C.c_int]
        L.so_raycast.restype = SoHit
        L.so_raycast.argtypes = [p, u32, u32, p, p, p, u32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]
        L.so_retrace_f64.argtypes = [p, u32, u32, p, u32, u32, u32, p, C.c_size_t, p]
        _lib = L
    return _lib


class SoHit(C.Structure):
    _fields_ = [("hit", C.c_uint32), ("id", C.c_uint32), ("material", C.c_uint32), ("pos", C.c_uint32 * 3),
                ("size", C.c_uint32)]


def raycast(words, depth, origin, direction, controlchecks, n_words=None):
    """Raycast (internal.glsl:297-400) on node words: (hit, (x, y, z), material, checks)."""
    w = np.ascontiguousarray(words, np.uint32)
    o = np.ascontiguousarray(origin, np.float32)
    d = np.ascontiguousarray(direction, np.float32)
    inv = (np.float32(1.0) / d).astype(np.float32)
    checks = C.c_uint32()
    h = lib().so_raycast(w.ctypes.data, len(w) if n_words is None else n_words, depth, o.ctypes.data, d.ctypes.data,
                         inv.ctypes.data, controlchecks, C.byref(checks), None)
    return bool(h.hit), tuple(h.pos), h.material, checks.value


RETRACE_DTYPE = np.dtype([("key32", "<u4"), ("key64", "<u4"), ("steps32", "<u4"), ("steps64", "<u4"),
                          ("diverge_step", "<u4"), ("pad", "<u4"), ("face_dist", "<f8"), ("pos_diff", "<f8")])


def retrace_f64(words, depth, ubo, W, H, controlchecks, pixels, n_words=None):
    """FP64 twin of Raycast for the primary rays of `pixels` (y*W+x) beside
    the FP32 walk: a RETRACE_DTYPE record per pixel."""
    w = np.ascontiguousarray(words, np.uint32)
    u = np.ascontiguousarray(ubo, np.float32)
    px = np.ascontiguousarray(pixels, np.uint32)
    out = np.zeros(len(px), RETRACE_DTYPE)
    lib().so_retrace_f64(w.ctypes.data, len(w) if n_words is None else n_words, depth, u.ctypes.data, W, H,
                         controlchecks, px.ctypes.data, len(px), out.ctypes.data)
    return out


def octree_build(ops, depth, max_words=1 << 27):
    ops = np.ascontiguousarray(ops)
    out = C.POINTER(C.c_uint32)()
    cap, size, nv = C.c_uint32(), C.c_uint32(), C.c_uint32()
    rc = lib().so_octree_build(ops.ctypes.data if len(ops) else None, len(ops), depth, max_words, C.byref(out),
                               C.byref(cap), C.byref(size), C.byref(nv))
    if rc != 0:
        raise RuntimeError("oracle octree build failed (%d)" % rc)
    try:
        words = np.ctypeslib.as_array(out, shape=(cap.value,)).copy()
    finally:
        lib().so_free(out)
    return words, dict(capacity=cap.value, size=size.value, num_voxels=nv.value)


def camera_ubo(pos, direction, fov, aspect):
    pos = np.ascontiguousarray(pos, np.float32)
    d = np.ascontiguousarray(direction, np.float32)
    out = np.zeros(16, np.float32)
    lib().so_camera_ubo(pos.ctypes.data, d.ctypes.data, fov, aspect, out.ctypes.data)
    return out


def fp_probe(op, inputs):
    """Evaluate an FP primitive of the GLSL lowering (see so_fp_probe)."""
    inp = np.ascontiguousarray(inputs, np.float32).reshape(-1, 3)
    n = len(inp)
    out = np.zeros(n * (3 if op == 1 else 1), np.float32)
    lib().so_fp_probe(op, inp.ctypes.data, out.ctypes.data, n)
    return out.reshape(n, 3) if op == 1 else out


def voxel_key(x, y, z):
    p = np.array([x, y, z], np.uint32)
    return lib().so_voxel_key(p.ctypes.data)


def generate_primary(words, depth, ubo, W, H, controlchecks=70, threads=None, n_words=None):
    words = np.ascontiguousarray(words, np.uint32)
    ubo = np.ascontiguousarray(ubo, np.float32)
    gbuf = np.zeros((H, W, 4), np.uint32)
    pos = np.zeros((H, W), np.uint32)
    reads, steps = C.c_uint64(), C.c_uint64()
    if threads is None:
        threads = os.cpu_count() or 1
    lib().so_generate_primary(words.ctypes.data, len(words) if n_words is None else n_words, depth, ubo.ctypes.data,
                              W, H, controlchecks, gbuf.ctypes.data, pos.ctypes.data, C.byref(reads), C.byref(steps),
                              threads)
    return gbuf, pos, dict(node_reads=reads.value, steps=steps.value)


class FrameConfig(C.Structure):
    """so_frame_config: field for field core::FrameConfig (src/renderer/core.hpp:83-158)."""
    _fields_ = [("primary_controlchecks", C.c_int32), ("bounce_controlchecks", C.c_int32),
                ("reconstruction_radius", C.c_int32), ("normal_precision", C.c_int32),
                ("disable_stale_reset", C.c_int32), ("update_time", C.c_uint32), ("shading_budget", C.c_int32),
                ("sample_cap_direct", C.c_int32), ("sample_cap_indirect", C.c_int32),
                ("weight_pixels", C.c_float), ("weight_variance", C.c_float), ("pixels_norm", C.c_float),
                ("variance_norm", C.c_float), ("top_fraction", C.c_float), ("mid_fraction", C.c_float),
                ("min_samples", C.c_int32), ("restir_enabled", C.c_int32), ("restir_m", C.c_int32),
                ("restir_temporal", C.c_int32), ("restir_spatial", C.c_int32),
                ("restir_spatial_neighbors", C.c_int32), ("restir_spatial_radius", C.c_float),
                ("restir_max_m", C.c_int32), ("firefly_k", C.c_float), ("firefly_floor", C.c_float),
                ("render_type", C.c_int32)]

    DEFAULTS = dict(primary_controlchecks=70, bounce_controlchecks=100, reconstruction_radius=10, normal_precision=6,
                    disable_stale_reset=0, update_time=0, shading_budget=131072, sample_cap_direct=16,
                    sample_cap_indirect=64, weight_pixels=0.1, weight_variance=0.9, pixels_norm=100.0,
                    variance_norm=10000.0, top_fraction=0.30, mid_fraction=0.40, min_samples=1, restir_enabled=1,
                    restir_m=4, restir_temporal=1, restir_spatial=1, restir_spatial_neighbors=1,
                    restir_spatial_radius=2.75, restir_max_m=24, firefly_k=1.0, firefly_floor=255.0, render_type=0)

    def __init__(self, **kw):
        super().__init__()
        d = dict(self.DEFAULTS)
        d.update(kw)
        for k, v in d.items():
            setattr(self, k, v)

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


BUFFER_IDS = dict(GBUFFER=0, GBUFFER_POS=1, NBUFFER=2, LBUFFER=3, RBUFFER=4, CANDIDATES=5, SHADE_LIST=6, RAYS=7,
                  OUTPUT=8, RANK_STATS=10, SAMPLE_COUNTS=11, CLAIM_MAP=12, EMISSIVES=13)


class Frame:
    """Persistent renderer state of the C oracle (so_state) + so_frame."""

    def __init__(self, W, H, hash_width=16384, max_bounces=8):
        L = lib()
        if not getattr(L, "_frame_sigs", False):
            L.so_state_create.restype = C.c_void_p
            L.so_state_create.argtypes = [C.c_uint32] * 4
            L.so_state_destroy.argtypes = [C.c_void_p]
            L.so_state_set_materials.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
            L.so_state_set_sky.argtypes = [C.c_void_p, C.c_void_p]
            L.so_state_set_sky_cube.restype = C.c_int
            L.so_state_set_sky_cube.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
            L.so_sky_sample.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
            L.so_frame.restype = C.c_int
            L.so_frame.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                   C.POINTER(FrameConfig), C.c_uint32, C.c_uint32, C.c_uint32]
            L.so_buffer.restype = C.c_size_t
            L.so_buffer.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]
            L.so_counters.argtypes = [C.c_void_p, C.c_void_p]
            L._frame_sigs = True
        self.L = L
        self.W, self.H, self.hash_width = W, H, hash_width
        self.h = L.so_state_create(hash_width, max_bounces, W, H)
        if not self.h:
            raise MemoryError("oracle state")

    def close(self):
        if self.h:
            self.L.so_state_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_materials(self, entries):
        e = np.ascontiguousarray(entries)
        self.L.so_state_set_materials(self.h, e.ctypes.data, len(e))

    def set_sky(self, rgb):
        a = np.ascontiguousarray(rgb, np.float32)
        self.L.so_state_set_sky(self.h, a.ctypes.data)

    def set_sky_cube(self, faces):
        """Six (edge, edge, 3) RGB8 faces, +X -X +Y -Y +Z -Z (skybox.cpp:87-122)."""
        f = np.ascontiguousarray(faces, np.uint8)
        if f.ndim != 4 or f.shape[0] != 6 or f.shape[1] != f.shape[2] or f.shape[3] != 3:
            raise ValueError("faces must be (6, edge, edge, 3) uint8")
        if self.L.so_state_set_sky_cube(self.h, f.ctypes.data, f.shape[1]) != 0:
            raise MemoryError("oracle sky")

    def set_sky_rgb8(self, rgb8):
        """A constant sky given as RGB8, sampled like every cube map (one texel per face)."""
        self.set_sky_cube(np.broadcast_to(np.asarray(rgb8, np.uint8).reshape(1, 1, 1, 3), (6, 1, 1, 3)))

    def sky_sample(self, dirs):
        """texture(skybox, dir).rgb for an (n, 3) float32 array of directions."""
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.empty_like(d)
        self.L.so_sky_sample(self.h, d.ctypes.data, len(d), out.ctypes.data)
        return out

    def frame(self, words, depth, ubo, fc, frame_idx, frame_seed, time, n_words=None):
        self._words = np.ascontiguousarray(words, np.uint32)
        ubo = np.ascontiguousarray(ubo, np.float32)
        n = len(self._words) if n_words is None else n_words
        rc = self.L.so_frame(self.h, self._words.ctypes.data, n, depth, ubo.ctypes.data, C.byref(fc), frame_idx,
                             frame_seed, time)
        if rc != 0:
            raise RuntimeError("oracle frame failed")

    def resolve(self, words, depth, ubo, fc, n_words=None):
        """resolve.comp alone on the current state (render type etc. from fc)."""
        L = self.L
        L.so_resolve.restype = C.c_int
        L.so_resolve.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.POINTER(FrameConfig)]
        self._words = np.ascontiguousarray(words, np.uint32)
        ubo = np.ascontiguousarray(ubo, np.float32)
        self._fc = fc
        n = len(self._words) if n_words is None else n_words
        if L.so_resolve(self.h, self._words.ctypes.data, n, depth, ubo.ctypes.data, C.byref(fc)) != 0:
            raise RuntimeError("oracle resolve failed")

    def read(self, name, dtype=np.uint32):
        p = C.c_void_p()
        n = self.L.so_buffer(self.h, BUFFER_IDS[name], C.byref(p))
        if n == 0:
            return np.zeros(0, dtype)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n,)).copy().view(dtype)

    def counters(self):
        out = np.zeros(16, np.uint32)
        self.L.so_counters(self.h, out.ctypes.data)
        return out

    def emit_records(self, base=0, n=None):
        """emit.comp's Ray records (24 uint32 each) for shade-list entries
        [base, base + n) of the last frame."""
        L = self.L
        L.so_emit_records.restype = C.c_size_t
        L.so_emit_records.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
        if n is None:
            n = int(self.counters()[1])
        out = np.zeros((max(n, 1), 24), np.uint32)
        got = L.so_emit_records(self.h, base, n, out.ctypes.data)
        return out[:got]
