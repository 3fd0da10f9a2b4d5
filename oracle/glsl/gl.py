"""ctypes OpenGL 4.5 binding over the headless llvmpipe context (test infrastructure).

The oracle harness runs the reference's own GLSL compute passes on the CPU
through Mesa llvmpipe.  This module owns the context and exposes the handful
of GL objects the passes bind: compute programs with uniforms, SSBOs, UBOs,
R32UI/RGBA32UI/RGBA32F images, an R32UI texture buffer (the octree) and a
cube map (the sky).  It is never imported by the product package.
"""
import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(os.path.dirname(HERE), "_ref")
LIB_PATH = os.path.join(REF_DIR, "libdri_context.so")
DRIVER = "/usr/lib/x86_64-linux-gnu/dri/swrast_dri.so"

# GL enums used by the passes
GL_COMPUTE_SHADER = 0x91B9
GL_COMPILE_STATUS = 0x8B81
GL_LINK_STATUS = 0x8B82
GL_SHADER_STORAGE_BUFFER = 0x90D2
GL_UNIFORM_BUFFER = 0x8A11
GL_TEXTURE_BUFFER = 0x8C2A
GL_TEXTURE_2D = 0x0DE1
GL_TEXTURE_CUBE_MAP = 0x8513
GL_TEXTURE_CUBE_MAP_POSITIVE_X = 0x8515
GL_DYNAMIC_DRAW = 0x88E8
GL_R32UI = 0x8236
GL_RGBA32UI = 0x8D70
GL_RGBA32F = 0x8814
GL_RGB8 = 0x8051
GL_RGB = 0x1907
GL_RED_INTEGER = 0x8D94
GL_RGBA_INTEGER = 0x8D99
GL_RGBA = 0x1908
GL_UNSIGNED_INT = 0x1405
GL_UNSIGNED_BYTE = 0x1401
GL_FLOAT = 0x1406
GL_READ_WRITE = 0x88BA
GL_TEXTURE0 = 0x84C0
GL_TEXTURE_MIN_FILTER = 0x2801
GL_TEXTURE_MAG_FILTER = 0x2800
GL_TEXTURE_WRAP_S = 0x2802
GL_TEXTURE_WRAP_T = 0x2803
GL_TEXTURE_WRAP_R = 0x8072
GL_CLAMP_TO_EDGE = 0x812F
GL_NEAREST = 0x2600
GL_LINEAR = 0x2601
GL_ALL_BARRIER_BITS = 0xFFFFFFFF
GL_INVALID_INDEX = 0xFFFFFFFF
GL_PACK_ALIGNMENT = 0x0D05
GL_UNPACK_ALIGNMENT = 0x0CF5

_u = C.c_uint
_i = C.c_int
_f = C.c_float
_p = C.c_void_p
_sz = C.c_ssize_t

_PROTOS = {
    "glGetError": (_u, []),
    "glGetString": (C.c_char_p, [_u]),
    "glGetIntegerv": (None, [_u, C.POINTER(_i)]),
    "glCreateShader": (_u, [_u]),
    "glShaderSource": (None, [_u, _i, C.POINTER(C.c_char_p), _p]),
    "glCompileShader": (None, [_u]),
    "glGetShaderiv": (None, [_u, _u, C.POINTER(_i)]),
    "glGetShaderInfoLog": (None, [_u, _i, _p, C.c_char_p]),
    "glCreateProgram": (_u, []),
    "glAttachShader": (None, [_u, _u]),
    "glLinkProgram": (None, [_u]),
    "glGetProgramiv": (None, [_u, _u, C.POINTER(_i)]),
    "glGetProgramInfoLog": (None, [_u, _i, _p, C.c_char_p]),
    "glUseProgram": (None, [_u]),
    "glGetUniformLocation": (_i, [_u, C.c_char_p]),
    "glGetUniformBlockIndex": (_u, [_u, C.c_char_p]),
    "glUniformBlockBinding": (None, [_u, _u, _u]),
    "glUniform1i": (None, [_i, _i]),
    "glUniform1ui": (None, [_i, _u]),
    "glUniform1f": (None, [_i, _f]),
    "glUniform2i": (None, [_i, _i, _i]),
    "glGenBuffers": (None, [_i, C.POINTER(_u)]),
    "glBindBuffer": (None, [_u, _u]),
    "glBufferData": (None, [_u, _sz, _p, _u]),
    "glBufferSubData": (None, [_u, _sz, _sz, _p]),
    "glGetBufferSubData": (None, [_u, _sz, _sz, _p]),
    "glBindBufferBase": (None, [_u, _u, _u]),
    "glGenTextures": (None, [_i, C.POINTER(_u)]),
    "glBindTexture": (None, [_u, _u]),
    "glActiveTexture": (None, [_u]),
    "glTexImage2D": (None, [_u, _i, _i, _i, _i, _i, _u, _u, _p]),
    "glTexSubImage2D": (None, [_u, _i, _i, _i, _i, _i, _u, _u, _p]),
    "glTexParameteri": (None, [_u, _u, _i]),
    "glGetTexImage": (None, [_u, _i, _u, _u, _p]),
    "glTexBuffer": (None, [_u, _u, _u]),
    "glBindImageTexture": (None, [_u, _u, _i, C.c_ubyte, _i, _u, _u]),
    "glPixelStorei": (None, [_u, _i]),
    "glDispatchCompute": (None, [_u, _u, _u]),
    "glMemoryBarrier": (None, [_u]),
    "glFinish": (None, []),
}


class GL:
    """Function table bound to the current (only) llvmpipe context."""

    _inst = None

    @classmethod
    def get(cls):
        if cls._inst is None:
            cls._inst = cls()
        return cls._inst

    def __init__(self):
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("llvmpipe harness not built: run `make -C oracle ref`")
        lib = C.CDLL(LIB_PATH)
        lib.vo_gl_init.restype = C.c_int
        lib.vo_gl_init.argtypes = [C.c_char_p]
        lib.vo_gl_proc.restype = C.c_void_p
        lib.vo_gl_proc.argtypes = [C.c_char_p]
        lib.vo_gl_error.restype = C.c_char_p
        rc = lib.vo_gl_init(DRIVER.encode())
        if rc != 0:
            raise RuntimeError("llvmpipe context: %s" % lib.vo_gl_error().decode())
        self.lib = lib
        for name, (res, args) in _PROTOS.items():
            addr = lib.vo_gl_proc(name.encode())
            if not addr:
                raise RuntimeError("missing GL entry point " + name)
            setattr(self, name, C.CFUNCTYPE(res, *args)(addr))
        self.glPixelStorei(GL_PACK_ALIGNMENT, 1)
        self.glPixelStorei(GL_UNPACK_ALIGNMENT, 1)

    def integer(self, pname):
        v = _i(0)
        self.glGetIntegerv(pname, C.byref(v))
        return v.value

    def check(self, label):
        err = self.glGetError()
        if err != 0:
            raise RuntimeError("GL error 0x%x at %s" % (err, label))


# Mesa's GLSL front end is strict where the reference's original driver was
# lenient.  internal.glsl:29 applies `&&` to two bvec3 values, which GLSL only
# defines for scalar bools; the lenient reading is component-wise, so
# all(a && b) is rewritten to the identical all(a) && all(b).
_CALL1 = r"lessThanEqual\((?:[^()]|\([^()]*\))*\)"   # one level of nested parentheses
SOURCE_FIXES = [
    (re.compile(r"all\((%s)\s*&&\s*(%s)\)" % (_CALL1, _CALL1)), r"(all(\1) && all(\2))"),
]


def load_source(path, defines=None):
    """Read a reference shader and resolve one level of #include "x" from the
    shader's own directory, as the reference's compileShader does
    (src/renderer/renderer.cpp:913-942).  `defines` replaces the value of an
    existing `#define NAME value` line (compile-time constants such as
    MAX_BOUNCES); shaders that do not define the name are left unchanged."""
    with open(path) as fh:
        src = fh.read()
    base = os.path.dirname(path)

    def incl(m):
        with open(os.path.join(base, m.group(1))) as fh:
            return fh.read()

    src = re.sub(r'#include "([^"]+)"', incl, src)
    for pat, repl in SOURCE_FIXES:
        src = pat.sub(repl, src)
    for name, value in (defines or {}).items():
        pat = re.compile(r"^(\s*#define\s+%s\s+)(\S+)" % re.escape(name), re.M)
        src = pat.sub(lambda m: m.group(1) + str(value), src)
    return src


class Program:
    def __init__(self, src, label="shader"):
        gl = GL.get()
        self.gl = gl
        sh = gl.glCreateShader(GL_COMPUTE_SHADER)
        b = src.encode()
        arr = (C.c_char_p * 1)(b)
        gl.glShaderSource(sh, 1, arr, None)
        gl.glCompileShader(sh)
        st = _i(0)
        gl.glGetShaderiv(sh, GL_COMPILE_STATUS, C.byref(st))
        if not st.value:
            log = C.create_string_buffer(8192)
            gl.glGetShaderInfoLog(sh, 8192, None, log)
            raise RuntimeError("%s compile failed:\n%s" % (label, log.value.decode()))
        p = gl.glCreateProgram()
        gl.glAttachShader(p, sh)
        gl.glLinkProgram(p)
        gl.glGetProgramiv(p, GL_LINK_STATUS, C.byref(st))
        if not st.value:
            log = C.create_string_buffer(8192)
            gl.glGetProgramInfoLog(p, 8192, None, log)
            raise RuntimeError("%s link failed:\n%s" % (label, log.value.decode()))
        self.id = p
        self.label = label

    def use(self):
        self.gl.glUseProgram(self.id)

    def loc(self, name):
        return self.gl.glGetUniformLocation(self.id, name.encode())

    def set1i(self, name, v):
        self.gl.glUniform1i(self.loc(name), int(v))

    def set1ui(self, name, v):
        self.gl.glUniform1ui(self.loc(name), int(v) & 0xFFFFFFFF)

    def set1f(self, name, v):
        self.gl.glUniform1f(self.loc(name), float(v))

    def set2i(self, name, a, b):
        self.gl.glUniform2i(self.loc(name), int(a), int(b))

    def bind_block(self, name, binding):
        idx = self.gl.glGetUniformBlockIndex(self.id, name.encode())
        if idx != GL_INVALID_INDEX:
            self.gl.glUniformBlockBinding(self.id, idx, binding)

    def dispatch(self, x, y=1, z=1):
        self.gl.glDispatchCompute(int(x), int(y), int(z))
        self.gl.glMemoryBarrier(GL_ALL_BARRIER_BITS)
        self.gl.check(self.label)


class Buffer:
    """SSBO / UBO / texture-buffer storage."""

    def __init__(self, nbytes=None, data=None):
        gl = GL.get()
        self.gl = gl
        b = _u(0)
        gl.glGenBuffers(1, C.byref(b))
        self.id = b.value
        if data is not None:
            data = np.ascontiguousarray(data)
            self.nbytes = data.nbytes
            ptr = data.ctypes.data
        else:
            self.nbytes = int(nbytes)
            ptr = None
        gl.glBindBuffer(GL_SHADER_STORAGE_BUFFER, self.id)
        gl.glBufferData(GL_SHADER_STORAGE_BUFFER, self.nbytes, ptr, GL_DYNAMIC_DRAW)
        gl.glBindBuffer(GL_SHADER_STORAGE_BUFFER, 0)
        if data is None:
            self.write(np.zeros(self.nbytes, np.uint8))

    def write(self, data, offset=0):
        data = np.ascontiguousarray(data)
        self.gl.glBindBuffer(GL_SHADER_STORAGE_BUFFER, self.id)
        self.gl.glBufferSubData(GL_SHADER_STORAGE_BUFFER, offset, data.nbytes, data.ctypes.data)
        self.gl.glBindBuffer(GL_SHADER_STORAGE_BUFFER, 0)

    def read(self, dtype=np.uint32, offset=0, nbytes=None):
        n = self.nbytes - offset if nbytes is None else nbytes
        out = np.empty(n, np.uint8)
        self.gl.glBindBuffer(GL_SHADER_STORAGE_BUFFER, self.id)
        self.gl.glGetBufferSubData(GL_SHADER_STORAGE_BUFFER, offset, n, out.ctypes.data)
        self.gl.glBindBuffer(GL_SHADER_STORAGE_BUFFER, 0)
        return out.view(dtype)

    def bind_ssbo(self, binding):
        self.gl.glBindBufferBase(GL_SHADER_STORAGE_BUFFER, binding, self.id)

    def bind_ubo(self, binding):
        self.gl.glBindBufferBase(GL_UNIFORM_BUFFER, binding, self.id)


_FMT = {
    GL_R32UI: (GL_RED_INTEGER, GL_UNSIGNED_INT, 1, np.uint32),
    GL_RGBA32UI: (GL_RGBA_INTEGER, GL_UNSIGNED_INT, 4, np.uint32),
    GL_RGBA32F: (GL_RGBA, GL_FLOAT, 4, np.float32),
}


class Image2D:
    def __init__(self, w, h, fmt):
        gl = GL.get()
        self.gl = gl
        self.w, self.h, self.fmt = int(w), int(h), fmt
        t = _u(0)
        gl.glGenTextures(1, C.byref(t))
        self.id = t.value
        ext, typ, comps, dt = _FMT[fmt]
        gl.glBindTexture(GL_TEXTURE_2D, self.id)
        zeros = np.zeros(self.w * self.h * comps, dt)
        gl.glTexImage2D(GL_TEXTURE_2D, 0, fmt, self.w, self.h, 0, ext, typ, zeros.ctypes.data)
        gl.glTexParameteri(GL_TEXTURE_2D, GL_TEXTURE_MIN_FILTER, GL_NEAREST)
        gl.glTexParameteri(GL_TEXTURE_2D, GL_TEXTURE_MAG_FILTER, GL_NEAREST)
        gl.glBindTexture(GL_TEXTURE_2D, 0)
        gl.check("image alloc")

    def write(self, data):
        ext, typ, comps, dt = _FMT[self.fmt]
        data = np.ascontiguousarray(data, dt)
        assert data.size == self.w * self.h * comps
        self.gl.glBindTexture(GL_TEXTURE_2D, self.id)
        self.gl.glTexSubImage2D(GL_TEXTURE_2D, 0, 0, 0, self.w, self.h, ext, typ, data.ctypes.data)
        self.gl.glBindTexture(GL_TEXTURE_2D, 0)

    def zero(self):
        ext, typ, comps, dt = _FMT[self.fmt]
        self.write(np.zeros(self.w * self.h * comps, dt))

    def read(self):
        ext, typ, comps, dt = _FMT[self.fmt]
        out = np.empty(self.w * self.h * comps, dt)
        self.gl.glBindTexture(GL_TEXTURE_2D, self.id)
        self.gl.glGetTexImage(GL_TEXTURE_2D, 0, ext, typ, out.ctypes.data)
        self.gl.glBindTexture(GL_TEXTURE_2D, 0)
        shape = (self.h, self.w) if comps == 1 else (self.h, self.w, comps)
        return out.reshape(shape)

    def bind_image(self, unit):
        self.gl.glBindImageTexture(unit, self.id, 0, 0, 0, GL_READ_WRITE, self.fmt)

    def bind_sampler(self, prog, name, unit):
        self.gl.glActiveTexture(GL_TEXTURE0 + unit)
        self.gl.glBindTexture(GL_TEXTURE_2D, self.id)
        prog.set1i(name, unit)


class TextureBuffer:
    """R32UI texture buffer object — the octree node array binding."""

    def __init__(self, words):
        gl = GL.get()
        self.gl = gl
        words = np.ascontiguousarray(words, np.uint32)
        self.buf = Buffer(data=words)
        t = _u(0)
        gl.glGenTextures(1, C.byref(t))
        self.id = t.value
        gl.glBindTexture(GL_TEXTURE_BUFFER, self.id)
        gl.glTexBuffer(GL_TEXTURE_BUFFER, GL_R32UI, self.buf.id)
        gl.glBindTexture(GL_TEXTURE_BUFFER, 0)
        gl.check("tbo")

    def bind(self, prog, name, unit):
        self.gl.glActiveTexture(GL_TEXTURE0 + unit)
        self.gl.glBindTexture(GL_TEXTURE_BUFFER, self.id)
        prog.set1i(name, unit)


class CubeMap:
    """Six RGB8 faces with GL_LINEAR filtering, as the reference's skybox
    (src/renderer/skybox.cpp:87-122).  A constant colour makes every texel the
    same, so bilinear filtering returns exactly c/255."""

    def __init__(self, faces_rgb8, edge):
        gl = GL.get()
        self.gl = gl
        t = _u(0)
        gl.glGenTextures(1, C.byref(t))
        self.id = t.value
        gl.glBindTexture(GL_TEXTURE_CUBE_MAP, self.id)
        for i, face in enumerate(faces_rgb8):
            face = np.ascontiguousarray(face, np.uint8)
            gl.glTexImage2D(GL_TEXTURE_CUBE_MAP_POSITIVE_X + i, 0, GL_RGB8, edge, edge, 0, GL_RGB,
                            GL_UNSIGNED_BYTE, face.ctypes.data)
        for pname, val in ((GL_TEXTURE_MIN_FILTER, GL_LINEAR), (GL_TEXTURE_MAG_FILTER, GL_LINEAR),
                           (GL_TEXTURE_WRAP_S, GL_CLAMP_TO_EDGE), (GL_TEXTURE_WRAP_T, GL_CLAMP_TO_EDGE),
                           (GL_TEXTURE_WRAP_R, GL_CLAMP_TO_EDGE)):
            gl.glTexParameteri(GL_TEXTURE_CUBE_MAP, pname, val)
        gl.glBindTexture(GL_TEXTURE_CUBE_MAP, 0)
        gl.check("cubemap")

    def bind(self, prog, name, unit):
        self.gl.glActiveTexture(GL_TEXTURE0 + unit)
        self.gl.glBindTexture(GL_TEXTURE_CUBE_MAP, self.id)
        prog.set1i(name, unit)
