"""Build libve_svo.so (the C-ABI library with the gfx950 kernels) in-tree.

    python voxelengine3.0_amd/build.py

Device code is compiled for gfx950 only, with FP contraction off (the
reference's arithmetic is reproduced operation by operation) and packed-FP32
instruction selection disabled.
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libve_svo.so")

HIP_SOURCES = ["ve_kernels.hip", "ve_passes.hip"]
CXX_SOURCES = ["ve_api.cpp", "ve_server.cpp", "ve_batch.cpp", "octree_build.cpp", "octree_edit.cpp", "scene_gen.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function"]
DEVICE = ["--offload-arch=gfx950", "-ffp-contract=off", "-fno-fast-math", "-Xclang", "-target-feature",
          "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    cc = hipcc()
    tmp = tempfile.mkdtemp(prefix="ve_build_")
    objs = []
    try:
        for src in HIP_SOURCES:
            obj = os.path.join(tmp, src + ".o")
            cmd = [cc, "-c", os.path.join(CSRC, src), "-o", obj] + COMMON + DEVICE
            _run(cmd, verbose)
            objs.append(obj)
        for src in CXX_SOURCES:
            obj = os.path.join(tmp, src + ".o")
            cmd = [cc, "-x", "c++", "-c", os.path.join(CSRC, src), "-o", obj, "-D__HIP_PLATFORM_AMD__",
                   "-I/opt/rocm/include", "-ffp-contract=off"] + COMMON
            _run(cmd, verbose)
            objs.append(obj)
        out_tmp = LIB + ".tmp"
        _run([cc, "-shared", "-o", out_tmp] + objs + ["--offload-arch=gfx950", "-L/opt/rocm/lib", "-lamdhip64"],
             verbose)
        os.replace(out_tmp, LIB)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return LIB


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n%s\n%s" % (" ".join(cmd), r.stdout))
    if verbose and r.stdout.strip():
        print(r.stdout)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
