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
        _lib = L
    return _lib


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
