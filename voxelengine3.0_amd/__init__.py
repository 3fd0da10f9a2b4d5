"""MI355X-native sparse-voxel-octree ray pipeline (VoxelEngine3.0's hot path).

Python host binding of the C ABI in include/ve_svo.h (libve_svo.so, built
in-tree by build.py).  The classes mirror the reference's host API so a
caller can switch by name:

    OctreeCPU  src/octree_cpu.hpp:30-37   insert / insertBox / insertSphere / insertFunction
    Octree     src/renderer/octree.hpp    set(OctreeCPU)  -> flat node words
    Camera     src/renderer/camera.hpp    UBO = UpdateUBO()
    Renderer   src/renderer/renderer.hpp  run(FrameConfig) -> bool

All compute runs in the HIP library; there is no CPU fallback — loading fails
loudly when the library is missing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libve_svo.so")

VE_OK = 0
VE_EINVAL = -1
VE_ENOMEM = -2
VE_EDEVICE = -3
VE_ECAPACITY = -4
VE_EFORMAT = -5
VE_ESTATE = -6
VE_EBUDGET = -7
VE_ENOTFOUND = -8

PASS = dict(PRIMARY=0, NORMAL=1, RESET=2, SCHEDULE=3, RANK=4, THRESHOLD=5, EMIT=6, TRACE=7, SHADE=8, RESOLVE=9,
            DISPLAY=10)
PASS_NAMES = [k for k, _ in sorted(PASS.items(), key=lambda kv: kv[1])]
BUF = dict(GBUFFER=0, GBUFFER_POS=1, NBUFFER=2, LBUFFER=3, RBUFFER=4, CANDIDATES=5, SHADE_LIST=6, RAYS=7,
           OUTPUT=8, COUNTERS=9, RANK_STATS=10, SAMPLE_COUNTS=11, CLAIM_MAP=12, EMISSIVES=13, DISPLAY=14)
CNT = dict(CANDIDATES=0, SHADE=1, EMISSIVES=2, LIVE=3, RAYS_TRACED=4, SHADOW_RAYS=5, PRIMARY_RAYS=6,
           BOUNCES_RUN=7, LIVE_B=8, OVERFLOW=9, REUSED=10, TIMEOUTS=11, EM_BASE=12, EM_NEW=13, DEFERRED=14,
           TIMEOUT_TILES=15)

RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("pad0", "<f4"), ("direction", "<f4", 3), ("pad1", "<f4"),
                      ("throughput", "<f4", 3), ("primary_vid", "<u4"), ("direct", "<f4", 3), ("rng", "<u4"),
                      ("indirect", "<f4", 3), ("bounce", "<u4"), ("hit_vid", "<u4"), ("hit_material", "<u4"),
                      ("pad2", "<u4"), ("pad3", "<u4")])
assert RAY_DTYPE.itemsize == 96


class VeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("ve error %d: %s" % (code, msg))
        self.code = code


class _Insert(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("z", C.c_uint32), ("material", C.c_uint32)]


class OctreeInfo(C.Structure):
    _fields_ = [("depth", C.c_uint32), ("size", C.c_uint32), ("capacity", C.c_uint32), ("num_voxels", C.c_uint32)]


class RendererConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("max_width", C.c_uint32), ("max_height", C.c_uint32),
                ("hash_width", C.c_uint32), ("max_bounces", C.c_uint32)]


class FrameConfig(C.Structure):
    """core::FrameConfig (src/renderer/core.hpp:83-158); defaults applied on construction."""
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

    def __init__(self, **kw):
        super().__init__()
        lib().ve_frame_config_defaults(C.byref(self))
        for k, v in kw.items():
            if not hasattr(self, k):
                raise AttributeError(k)
            setattr(self, k, v)


class FrameInputs(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("frame_idx", C.c_uint32),
                ("frame_seed", C.c_uint32), ("time", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("max_sessions", C.c_uint32), ("hbm_budget", C.c_uint64),
                ("queues", C.c_uint32)]


class SessionDesc(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32), ("max_bounces", C.c_uint32),
                ("hash_width", C.c_uint32), ("materials", C.c_void_p), ("n_materials", C.c_uint32),
                ("sky_rgb8", C.c_uint8 * 3)]


class FrameDone(C.Structure):
    _fields_ = [("session", C.c_uint32), ("frame", C.c_uint32), ("tag", C.c_uint64), ("arrival_ms", C.c_double),
                ("start_ms", C.c_double), ("done_ms", C.c_double), ("rays", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("admitted", C.c_uint32), ("waiting", C.c_uint32), ("frames_in_flight", C.c_uint32),
                ("pooled", C.c_uint32), ("bytes_in_use", C.c_uint64), ("budget", C.c_uint64)]


class BatchConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("max_views", C.c_uint32), ("tile_w", C.c_uint32), ("grab", C.c_uint32),
                ("waves_per_cu", C.c_uint32), ("groups", C.c_uint32), ("whole_frame", C.c_uint32)]


_lib = None


def lib():
    """The HIP library; raises if it has not been built (no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libve_svo.so is not built: run `python voxelengine3.0_amd/build.py`")
    L = C.CDLL(LIB_PATH)
    p, u32, i32, sz = C.c_void_p, C.c_uint32, C.c_int, C.c_size_t
    sig = {
        "ve_octree_build": (i32, [p, sz, u32, u32, C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(OctreeInfo)]),
        "ve_free": (None, [p]),
        "ve_gen_caves": (i32, [u32, p, p, u32, u32, C.c_double, u32, u32, u32, p, sz, C.POINTER(sz)]),
        "ve_create": (i32, [C.POINTER(p), C.POINTER(RendererConfig)]),
        "ve_destroy": (None, [p]),
        "ve_last_error": (C.c_char_p, [p]),
        "ve_scene_upload": (i32, [p, p, u32, u32, C.POINTER(p)]),
        "ve_scene_release": (i32, [p, p]),
        "ve_set_camera": (i32, [p, p]),
        "ve_camera_ubo": (None, [p, p, C.c_float, C.c_float, p]),
        "ve_set_materials": (i32, [p, p, u32]),
        "ve_set_sky": (i32, [p, p, i32]),
        "ve_frame_config_defaults": (None, [C.POINTER(FrameConfig)]),
        "ve_frame": (i32, [p, p, C.POINTER(FrameConfig), C.POINTER(FrameInputs)]),
        "ve_pass": (i32, [p, p, i32, C.POINTER(FrameConfig), C.POINTER(FrameInputs)]),
        "ve_read": (i32, [p, i32, p, sz]),
        "ve_write": (i32, [p, i32, p, sz]),
        "ve_buffer_size": (i32, [p, i32, C.POINTER(sz)]),
        "ve_sync": (i32, [p]),
        "ve_enable_timing": (i32, [p, i32]),
        "ve_set_overlap": (i32, [p, i32]),
        "ve_timings": (i32, [p, p]),
        "ve_enable_node_stats": (i32, [p, i32]),
        "ve_node_stats": (i32, [p] + [C.POINTER(C.c_uint64)] * 4),
        "ve_node_stats_bounce": (i32, [p] + [C.POINTER(C.c_uint64)] * 2),
        "ve_timings_iter": (i32, [p, u32, p]),
        "ve_set_tuning": (i32, [p, p]),
        "ve_set_graph": (i32, [p, i32]),
        "ve_batch_create": (i32, [C.POINTER(p), C.POINTER(BatchConfig)]),
        "ve_batch_destroy": (None, [p]),
        "ve_batch_last_error": (C.c_char_p, [p]),
        "ve_batch_frame": (i32, [p, u32, p, p, p, p]),
        "ve_batch_primary_ms": (i32, [p, C.POINTER(C.c_float)]),
        "ve_batch_graph_stats": (i32, [p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "ve_graph_stats": (i32, [p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "ve_server_set_graph": (i32, [p, i32]),
        "ve_server_set_tuning": (i32, [p, p]),
        "ve_server_set_placement": (i32, [p, i32]),
        "ve_server_set_batch": (i32, [p, C.POINTER(BatchConfig)]),
        "ve_server_submit_batch": (i32, [p, u32, p, p, p, p, p, p]),
        "ve_get_tuning": (i32, [p, p]),
        "ve_ray_totals": (i32, [p] + [C.POINTER(C.c_uint64)] * 3),
        "ve_ray_totals_reused": (i32, [p, C.POINTER(C.c_uint64)]),
        "ve_reset_ray_totals": (i32, [p]),
        "ve_octree_cpu_create": (i32, [C.POINTER(p), u32, u32]),
        "ve_octree_cpu_destroy": (None, [p]),
        "ve_octree_cpu_insert": (i32, [p, p, sz]),
        "ve_octree_cpu_remove": (i32, [p, p, sz]),
        "ve_octree_cpu_lookup": (i32, [p, u32, u32, u32, C.POINTER(C.c_int), C.POINTER(u32)]),
        "ve_octree_cpu_raycast": (i32, [p, p, p, u32, C.POINTER(C.c_int), p, C.POINTER(u32)]),
        "ve_octree_set": (i32, [p, C.POINTER(OctreeInfo)]),
        "ve_octree_apply_edits": (i32, [p, C.POINTER(OctreeInfo)]),
        "ve_octree_words": (i32, [p, C.POINTER(C.POINTER(u32)), C.POINTER(u32)]),
        "ve_octree_changed": (i32, [p, C.POINTER(C.POINTER(u32)), C.POINTER(sz), C.POINTER(C.c_int)]),
        "ve_octree_find_slot": (u32, [p, u32, u32, u32]),
        "ve_scene_update": (i32, [p, p, p, u32, p, sz]),
        "ve_server_create": (i32, [C.POINTER(p), C.POINTER(ServerConfig)]),
        "ve_server_destroy": (None, [p]),
        "ve_server_last_error": (C.c_char_p, [p]),
        "ve_server_now_ms": (C.c_double, [p]),
        "ve_server_add_scene": (i32, [p, p, u32, u32, C.POINTER(u32)]),
        "ve_server_open": (i32, [p, C.POINTER(SessionDesc), C.POINTER(u32)]),
        "ve_server_submit": (i32, [p, u32, C.POINTER(FrameConfig), C.POINTER(FrameInputs), p, C.c_uint64,
                                   C.c_double]),
        "ve_server_close": (i32, [p, u32]),
        "ve_server_poll": (i32, [p, p, u32, C.POINTER(u32), i32]),
        "ve_server_read_output": (i32, [p, u32, p, sz]),
        "ve_server_get_stats": (i32, [p, C.POINTER(ServerStats)]),
        "ve_server_read": (i32, [p, u32, C.c_int, p, sz]),
        "ve_pass_primary_noalloc": (i32, [p, u32, u32, p, u32, u32, i32, p, p, p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def exported_symbols():
    """Names of the C-ABI entry points declared in include/ve_svo.h."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "ve_svo.h")
    with open(hdr) as fh:
        txt = fh.read()
    return sorted(set(re.findall(r"\b(ve_[a-z0-9_]+)\s*\(", txt)))


def _check(rc, ctx=None):
    if rc != VE_OK:
        msg = lib().ve_last_error(ctx)
        raise VeError(rc, msg.decode() if msg else "")


# ------------------------------------------------------------- octree build
def octree_build(ops, depth, max_words=1 << 27):
    """OctreeCPU inserts (an insert-dtype array in call order) -> Octree::set
    node words.  Returns (words[capacity] uint32, OctreeInfo)."""
    ops = np.ascontiguousarray(ops)
    out = C.POINTER(C.c_uint32)()
    info = OctreeInfo()
    _check(lib().ve_octree_build(ops.ctypes.data if len(ops) else None, len(ops), depth, max_words,
                                 C.byref(out), C.byref(info)))
    try:
        words = np.ctypeslib.as_array(out, shape=(info.capacity,)).copy()
    finally:
        lib().ve_free(out)
    return words, info


def gen_caves(hi, lo, band, cell, threshold, noise_seed, material_seed, threads=0):
    """Synthetic cave terrain (ve_gen_caves): the voxels of a heightfield
    ground with value-noise caves in a band under its surface, as an
    insert-dtype array (hi, lo: n x n int32 column bounds, [x][z])."""
    hi = np.ascontiguousarray(hi, np.int32)
    lo = np.ascontiguousarray(lo, np.int32)
    n = hi.shape[0]
    if hi.shape != (n, n) or lo.shape != (n, n):
        raise ValueError("hi and lo must be n x n")
    cnt = C.c_size_t()
    args = (n, hi.ctypes.data, lo.ctypes.data, band, cell, float(threshold), noise_seed & 0xFFFFFFFF,
            material_seed & 0xFFFFFFFF, threads)
    _check(lib().ve_gen_caves(*args, None, 0, C.byref(cnt)))
    from .scenes import INSERT_DTYPE
    out = np.empty(cnt.value, INSERT_DTYPE)
    _check(lib().ve_gen_caves(*args, out.ctypes.data if len(out) else None, len(out), C.byref(cnt)))
    return out


class OctreeCPU:
    """Host octree edit API (src/octree_cpu.hpp:9-60): a persistent pointer
    tree (ve_octree_cpu) with insert / remove / lookup / raycast and the
    insertBox / insertSphere / insertFunction helpers.  Octree.set() and
    Octree.applyEdits() turn it into node words.  (ve.octree_build is the
    batch path for whole scenes given as insert lists.)"""

    def __init__(self, depth, max_words=1 << 27):
        self.depth = int(depth)
        h = C.c_void_p()
        _check(lib().ve_octree_cpu_create(C.byref(h), self.depth, max_words))
        self.h = h

    def close(self):
        if self.h:
            lib().ve_octree_cpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def insert_ops(self, ops):
        ops = np.ascontiguousarray(ops, dtype=_insert_dtype())
        _check(lib().ve_octree_cpu_insert(self.h, ops.ctypes.data if len(ops) else None, len(ops)))

    def insert(self, x, y, z, material):
        self.insert_ops(np.array([(x, y, z, material)], dtype=_insert_dtype()))

    def remove(self, x, y, z):
        self.remove_many(np.array([[x, y, z]], np.uint32))

    def remove_many(self, xyz):
        xyz = np.ascontiguousarray(np.asarray(xyz, np.uint32).reshape(-1, 3))
        _check(lib().ve_octree_cpu_remove(self.h, xyz.ctypes.data if len(xyz) else None, len(xyz)))

    def lookup(self, x, y, z):
        """(present, material) of the leaf at a position (OctreeCPU::lookup)."""
        pr, m = C.c_int(), C.c_uint32()
        _check(lib().ve_octree_cpu_lookup(self.h, x, y, z, C.byref(pr), C.byref(m)))
        return bool(pr.value), m.value

    def raycast(self, origin, direction, max_steps=300):
        """OctreeCPU::raycast: (hit, (x, y, z), material)."""
        o = np.ascontiguousarray(origin, np.float32)
        d = np.ascontiguousarray(direction, np.float32)
        hit, pos, m = C.c_int(), (C.c_uint32 * 3)(), C.c_uint32()
        _check(lib().ve_octree_cpu_raycast(self.h, o.ctypes.data, d.ctypes.data, max_steps, C.byref(hit), pos,
                                           C.byref(m)))
        return bool(hit.value), tuple(pos), m.value

    def insertBox(self, lo, hi, material):
        n = 1 << self.depth
        hi = [min(int(h), n) for h in hi]
        X, Y, Z = np.meshgrid(*[np.arange(int(lo[i]), hi[i]) for i in range(3)], indexing="ij")
        ops = np.empty(X.size, _insert_dtype())
        ops["x"], ops["y"], ops["z"], ops["material"] = X.ravel(), Y.ravel(), Z.ravel(), material
        self.insert_ops(ops)

    def insertSphere(self, centre, radius, material):
        from .scenes import sphere_ops
        self.insert_ops(sphere_ops(centre, radius, material, 1 << self.depth))

    def insertFunction(self, fn, lo, hi, material):
        n = 1 << self.depth
        hi = [min(int(h), n) for h in hi]
        X, Y, Z = np.meshgrid(*[np.arange(int(lo[i]), hi[i]) for i in range(3)], indexing="ij")
        c = np.stack([X.ravel() + 0.5, Y.ravel() + 0.5, Z.ravel() + 0.5], 1).astype(np.float32)
        m = np.asarray(fn(c), bool)
        ops = np.empty(int(m.sum()), _insert_dtype())
        ops["x"], ops["y"], ops["z"], ops["material"] = X.ravel()[m], Y.ravel()[m], Z.ravel()[m], material
        self.insert_ops(ops)


def _insert_dtype():
    from .scenes import INSERT_DTYPE
    return INSERT_DTYPE


class Octree:
    """Node-word image of an OctreeCPU (src/renderer/octree.hpp:14-98):
    set() linearises the whole tree, applyEdits() syncs only what changed and
    reports the written slots (for Renderer.update)."""

    def __init__(self, depth, max_words=1 << 27):
        self.depth = int(depth)
        self.max_words = max_words
        self.data = np.zeros(8, np.uint32)
        self.size, self.capacity, self.numVoxels = 8, 8, 0
        self.changed, self.full = np.zeros(0, np.uint32), True

    def _pull(self, cpu, info):
        w, cap = C.POINTER(C.c_uint32)(), C.c_uint32()
        _check(lib().ve_octree_words(cpu.h, C.byref(w), C.byref(cap)))
        self.data = np.ctypeslib.as_array(w, shape=(cap.value,)).copy()
        sl, n, full = C.POINTER(C.c_uint32)(), C.c_size_t(), C.c_int()
        _check(lib().ve_octree_changed(cpu.h, C.byref(sl), C.byref(n), C.byref(full)))
        self.changed = np.ctypeslib.as_array(sl, shape=(n.value,)).copy() if n.value else np.zeros(0, np.uint32)
        self.full = bool(full.value)
        self.size, self.capacity, self.numVoxels = info.size, info.capacity, info.num_voxels
        return self

    def set(self, cpu):
        info = OctreeInfo()
        _check(lib().ve_octree_set(cpu.h, C.byref(info)))
        return self._pull(cpu, info)

    def applyEdits(self, cpu):
        info = OctreeInfo()
        _check(lib().ve_octree_apply_edits(cpu.h, C.byref(info)))
        return self._pull(cpu, info)

    def findGpuSlot(self, cpu, x, y, z):
        s = lib().ve_octree_find_slot(cpu.h, x, y, z)
        return None if s == 0xFFFFFFFF else s


def camera_ubo(position, direction, fov_deg, aspect):
    """Camera::UpdateUBO (src/renderer/camera.cpp:39-50) -> float32[16]."""
    pos = np.ascontiguousarray(position, np.float32)
    d = np.ascontiguousarray(direction, np.float32)
    out = np.zeros(16, np.float32)
    lib().ve_camera_ubo(pos.ctypes.data, d.ctypes.data, float(fov_deg), float(aspect), out.ctypes.data)
    return out


def primary_noalloc(words_dev, n_words, depth, ubo, width, height, controlchecks, gbuffer_dev, gbuffer_pos_dev,
                    stream=None):
    """ve_pass_primary_noalloc: the primary pass over device buffers given as
    addresses (words, G-buffer uint32 x4 per pixel, positions uint32 per
    pixel), with no context and no device allocation; enqueued on `stream`
    (a hipStream_t address, None = the null stream)."""
    ubo = np.ascontiguousarray(ubo, np.float32)
    _check(lib().ve_pass_primary_noalloc(words_dev, n_words, depth, ubo.ctypes.data, width, height, controlchecks,
                                         gbuffer_dev, gbuffer_pos_dev, stream))


class Batch:
    """ve_batch_*: frame k of many Renderers at once (one persistent primary
    launch over all views, the other passes on each view's stream; with
    whole_frame, every pass as one launch over all views)."""

    def __init__(self, device=0, max_views=256, tile_w=8, grab=4, waves_per_cu=16, groups=4, whole_frame=False):
        cfg = BatchConfig(device, max_views, tile_w, grab, waves_per_cu, groups, 1 if whole_frame else 0)
        h = C.c_void_p()
        _check(lib().ve_batch_create(C.byref(h), C.byref(cfg)))
        self.h = h
        self.config = dict(tile_w=tile_w, grab=grab, waves_per_cu=waves_per_cu, groups=groups,
                           whole_frame=bool(whole_frame))

    def close(self):
        if self.h:
            lib().ve_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def frame(self, renderers, scenes, fcs, inputs):
        n = len(renderers)
        ctxs = (C.c_void_p * n)(*[r.h.value for r in renderers])
        scs = (C.c_void_p * n)(*[s.value for s in scenes])
        fca = (FrameConfig * n)(*fcs)
        ina = (FrameInputs * n)(*inputs)
        rc = lib().ve_batch_frame(self.h, n, ctxs, scs, fca, ina)
        if rc != VE_OK:
            msg = lib().ve_batch_last_error(self.h)
            raise VeError(rc, msg.decode() if msg else "")

    def primary_ms(self):
        ms = C.c_float()
        _check(lib().ve_batch_primary_ms(self.h, C.byref(ms)))
        return ms.value

    def graph_stats(self):
        """ve_batch_graph_stats: whole-frame batch graphs captured and replayed."""
        c, r = C.c_uint64(), C.c_uint64()
        _check(lib().ve_batch_graph_stats(self.h, C.byref(c), C.byref(r)))
        return dict(captures=c.value, replays=r.value)


class Renderer:
    """One device context: Renderer(RendererConfig, ...) + run(FrameConfig)
    (src/renderer/renderer.hpp:11-13), plus per-pass dispatch and readback."""

    def __init__(self, max_width, max_height, hash_width=16384, max_bounces=8, device=0):
        cfg = RendererConfig(device, max_width, max_height, hash_width, max_bounces)
        h = C.c_void_p()
        _check(lib().ve_create(C.byref(h), C.byref(cfg)))
        self.h = h
        self.max_width, self.max_height = max_width, max_height
        self.hash_width = hash_width
        self.max_bounces = max_bounces
        self.scenes = []

    def close(self):
        if self.h:
            lib().ve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        _check(rc, self.h)

    def upload(self, words, depth, n_words=None):
        words = np.ascontiguousarray(words, np.uint32)
        n = len(words) if n_words is None else int(n_words)
        s = C.c_void_p()
        self._ck(lib().ve_scene_upload(self.h, words.ctypes.data, n, depth, C.byref(s)))
        return s

    def release(self, scene):
        self._ck(lib().ve_scene_release(self.h, scene))

    def update(self, scene, octree):
        """ve_scene_update after Octree.applyEdits: only the changed words travel."""
        words = np.ascontiguousarray(octree.data, np.uint32)
        if octree.full:
            slots = np.arange(octree.size, dtype=np.uint32)
        else:
            slots = np.ascontiguousarray(octree.changed, np.uint32)
        self._ck(lib().ve_scene_update(self.h, scene, words.ctypes.data, octree.size,
                                       slots.ctypes.data if len(slots) else None, len(slots)))

    def set_graph(self, on=True):
        """ve_set_graph: replay frames from a captured HIP graph."""
        self._ck(lib().ve_set_graph(self.h, 1 if on else 0))

    def graph_stats(self):
        a, b = C.c_uint64(), C.c_uint64()
        self._ck(lib().ve_graph_stats(self.h, C.byref(a), C.byref(b)))
        return dict(captures=a.value, replays=b.value)

    def set_tuning(self, tile_chunk=16, jump_level=0, bounce_blocks=0):
        """ve_set_tuning: launch-shape knobs (speed only, never results)."""
        t = (C.c_uint32 * 3)(tile_chunk, jump_level, bounce_blocks)
        self._ck(lib().ve_set_tuning(self.h, t))

    def tuning(self):
        t = (C.c_uint32 * 3)()
        self._ck(lib().ve_get_tuning(self.h, t))
        return dict(tile_chunk=t[0], jump_level=t[1], bounce_blocks=t[2])

    def set_camera(self, ubo):
        ubo = np.ascontiguousarray(ubo, np.float32)
        self._ck(lib().ve_set_camera(self.h, ubo.ctypes.data))

    def set_materials(self, entries):
        entries = np.ascontiguousarray(entries)
        self._ck(lib().ve_set_materials(self.h, entries.ctypes.data, len(entries)))

    def set_sky_constant(self, rgb8):
        face = np.ascontiguousarray(np.asarray(rgb8, np.uint8).reshape(3))
        arr = (C.c_void_p * 6)(*([face.ctypes.data] * 6))
        self._ck(lib().ve_set_sky(self.h, arr, 1))
        self._sky_keep = face

    def set_sky_cube(self, faces):
        """Six (edge, edge, 3) uint8 faces, +X -X +Y -Y +Z -Z (skybox.cpp:87-122)."""
        f = np.ascontiguousarray(faces, np.uint8)
        if f.ndim != 4 or f.shape[0] != 6 or f.shape[1] != f.shape[2] or f.shape[3] != 3:
            raise ValueError("faces must be (6, edge, edge, 3) uint8")
        arr = (C.c_void_p * 6)(*[f[i].ctypes.data for i in range(6)])
        self._ck(lib().ve_set_sky(self.h, arr, f.shape[1]))

    def run_pass(self, scene, name, fc, inputs):
        self._ck(lib().ve_pass(self.h, scene, PASS[name], C.byref(fc), C.byref(inputs)))

    def frame(self, scene, fc, inputs):
        self._ck(lib().ve_frame(self.h, scene, C.byref(fc), C.byref(inputs)))

    def size(self, buf):
        n = C.c_size_t()
        self._ck(lib().ve_buffer_size(self.h, BUF[buf], C.byref(n)))
        return n.value

    def read(self, buf, dtype=np.uint32, nbytes=None):
        n = self.size(buf) if nbytes is None else int(nbytes)
        out = np.empty(n, np.uint8)
        self._ck(lib().ve_read(self.h, BUF[buf], out.ctypes.data, n))
        return out.view(dtype)

    def write(self, buf, data):
        data = np.ascontiguousarray(data)
        self._ck(lib().ve_write(self.h, BUF[buf], data.ctypes.data, data.nbytes))

    def gbuffer(self, W, H):
        g = self.read("GBUFFER", np.uint32, W * H * 16).reshape(H, W, 4)
        p = self.read("GBUFFER_POS", np.uint32, W * H * 4).reshape(H, W)
        return g, p

    def counters(self):
        return self.read("COUNTERS", np.uint32)

    def sync(self):
        self._ck(lib().ve_sync(self.h))

    def enable_node_stats(self, on=True):
        self._ck(lib().ve_enable_node_stats(self.h, 1 if on else 0))

    def node_stats(self):
        v = [C.c_uint64() for _ in range(4)]
        self._ck(lib().ve_node_stats(self.h, *[C.byref(x) for x in v]))
        b = [C.c_uint64() for _ in range(2)]
        self._ck(lib().ve_node_stats_bounce(self.h, *[C.byref(x) for x in b]))
        return dict(primary_reads=v[0].value, trace_reads=v[1].value, primary_steps=v[2].value,
                    trace_steps=v[3].value, bounce_reads=b[0].value, bounce_steps=b[1].value)

    def ray_totals(self):
        v = [C.c_uint64() for _ in range(3)]
        self._ck(lib().ve_ray_totals(self.h, *[C.byref(x) for x in v]))
        reused = C.c_uint64()
        self._ck(lib().ve_ray_totals_reused(self.h, C.byref(reused)))
        # traced counts every live ray of each trace pass (the reference's
        # count); reused of them took the primary pass's hit record
        return dict(primary=v[0].value, traced=v[1].value, shadow=v[2].value, reused=reused.value)

    def reset_ray_totals(self):
        self._ck(lib().ve_reset_ray_totals(self.h))

    def set_overlap(self, on=True):
        """Cross-frame overlap of the primary pass (on by default)."""
        self._ck(lib().ve_set_overlap(self.h, 1 if on else 0))

    def enable_timing(self, on=True):
        """on: True/False, or the number of frames about to be recorded."""
        n = int(on) if not isinstance(on, bool) else (1 if on else 0)
        self._ck(lib().ve_enable_timing(self.h, n))

    def timings(self):
        out = np.zeros(len(PASS), np.float32)
        self._ck(lib().ve_timings(self.h, out.ctypes.data))
        return dict(zip(PASS_NAMES, out.tolist()))

    def timings_iter(self, it):
        """Per-pass means split by bounce iteration (read before timings())."""
        out = np.zeros(len(PASS), np.float32)
        self._ck(lib().ve_timings_iter(self.h, it, out.ctypes.data))
        return dict(zip(PASS_NAMES, out.tolist()))


class Server:
    """Session server (ve_server_*): resident octrees, per-session accumulation
    state admitted by a device-memory budget, frames of different sessions
    overlapping on the device.  queues=0: one stream per session; queues=1:
    the plain loop (every frame on one stream, in submission order)."""

    def __init__(self, device=0, max_sessions=64, hbm_budget=0, queues=0):
        cfg = ServerConfig(device, max_sessions, hbm_budget, queues)
        h = C.c_void_p()
        _check(lib().ve_server_create(C.byref(h), C.byref(cfg)))
        self.h = h
        self._keep = {}

    def close(self):
        if self.h:
            lib().ve_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        if rc != VE_OK:
            msg = lib().ve_server_last_error(self.h)
            raise VeError(rc, msg.decode() if msg else "")

    def now_ms(self):
        return lib().ve_server_now_ms(self.h)

    def add_scene(self, words, depth, n_words=None):
        words = np.ascontiguousarray(words, np.uint32)
        n = len(words) if n_words is None else int(n_words)
        sid = C.c_uint32()
        self._ck(lib().ve_server_add_scene(self.h, words.ctypes.data, n, depth, C.byref(sid)))
        return sid.value

    def open(self, scene, width, height, max_bounces, materials, sky_rgb8, hash_width=16384):
        mats = np.ascontiguousarray(materials)
        d = SessionDesc(scene, width, height, max_bounces, hash_width, mats.ctypes.data, len(mats),
                        (C.c_uint8 * 3)(*[int(v) for v in sky_rgb8]))
        sid = C.c_uint32()
        self._ck(lib().ve_server_open(self.h, C.byref(d), C.byref(sid)))
        return sid.value

    def submit(self, session, fc, inputs, ubo, tag=0, arrival_ms=-1.0):
        ubo = np.ascontiguousarray(ubo, np.float32)
        self._ck(lib().ve_server_submit(self.h, session, C.byref(fc), C.byref(inputs), ubo.ctypes.data, tag,
                                        float(arrival_ms)))

    def submit_batch(self, sessions, fcs, inputs, ubos, tags=None, arrival_ms=None):
        """ve_server_submit_batch: one frame for each of the distinct sessions
        (arrival_ms: one time per frame, None = now)."""
        n = len(sessions)
        if n == 0:
            return
        sa = (C.c_uint32 * n)(*sessions)
        fca = (FrameConfig * n)(*fcs)
        ina = (FrameInputs * n)(*inputs)
        ub = np.ascontiguousarray(np.asarray(ubos, np.float32).reshape(n, 16))
        ta = (C.c_uint64 * n)(*(tags if tags is not None else [0] * n))
        aa = (C.c_double * n)(*(arrival_ms if arrival_ms is not None else [-1.0] * n))
        self._ck(lib().ve_server_submit_batch(self.h, n, sa, fca, ina, ub.ctypes.data, ta, aa))

    def set_graph(self, on=True):
        self._ck(lib().ve_server_set_graph(self.h, 1 if on else 0))

    def set_placement(self, by_load=True):
        """ve_server_set_placement: each frame on the least-loaded shared stream."""
        self._ck(lib().ve_server_set_placement(self.h, 1 if by_load else 0))

    def set_tuning(self, tile_chunk=16, bounce_blocks=0):
        """ve_server_set_tuning: Renderer.set_tuning for every session's context."""
        t = (C.c_uint32 * 3)(tile_chunk, 0, bounce_blocks)
        self._ck(lib().ve_server_set_tuning(self.h, t))

    def set_batch(self, tile_w=8, grab=4, waves_per_cu=16, groups=4, whole_frame=False):
        cfg = BatchConfig(0, 0, tile_w, grab, waves_per_cu, groups, 1 if whole_frame else 0)
        self._ck(lib().ve_server_set_batch(self.h, C.byref(cfg)))

    def close_session(self, session):
        self._ck(lib().ve_server_close(self.h, session))

    def poll(self, wait=False, max_n=256):
        out = (FrameDone * max_n)()
        n = C.c_uint32()
        self._ck(lib().ve_server_poll(self.h, out, max_n, C.byref(n), 1 if wait else 0))
        return [dict(session=d.session, frame=d.frame, tag=d.tag, arrival_ms=d.arrival_ms, start_ms=d.start_ms,
                     done_ms=d.done_ms, rays=d.rays) for d in out[:n.value]]

    def read_output(self, session, width, height):
        out = np.empty(width * height * 4, np.float32)
        self._ck(lib().ve_server_read_output(self.h, session, out.ctypes.data, out.nbytes))
        return out

    def read(self, session, buf, dtype=np.uint32, nbytes=None):
        """Any buffer of a session's state (ve_server_read), as Renderer.read."""
        if nbytes is None:
            raise ValueError("nbytes is required")
        out = np.empty(int(nbytes), np.uint8)
        self._ck(lib().ve_server_read(self.h, session, BUF[buf], out.ctypes.data, out.nbytes))
        return out.view(dtype)

    def stats(self):
        st = ServerStats()
        self._ck(lib().ve_server_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}
