"""The C++ host classes (include/ve_svo.hpp) compile against the C ABI and
build the same node array as the Python binding for the same insert calls."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_host_api(ve, scenes, tmp_path):
    exe = tmp_path / "host_api_check"
    lib_dir = os.path.join(ROOT, "voxelengine3.0_amd")
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp",
           "host_api_check.cpp"), "-L", lib_dir, "-lve_svo", "-Wl,-rpath," + lib_dir, "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "cannot find -lve_svo" in r.stderr:
        pytest.skip("library not built")
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    n_ops, size, cap, nvox, dig = (int(v) for v in out[0].split())
    assert out[1].split() == ["70", "131072"]
    # the same calls through the Python binding
    L = 64
    cpu = ve.OctreeCPU(6)
    cpu.insertSphere((L / 3.0, L / 3.0 - 5.0, L / 2.0), L / 4.0, 5)
    cpu.insertBox((0, 0, 0), (L, 4, L), 4)
    cpu.insertFunction(lambda p: (p[:, 1] > 50.0) & (p[:, 0] < 10.0), (0, 0, 0), (L, L, L), 2)
    o = ve.Octree(6).set(cpu)
    assert (len(cpu.ops()), o.size, o.capacity, o.numVoxels) == (n_ops, size, cap, nvox)
    s = 0
    for w in o.data[: o.size].tolist():
        s = (s * 1000003 + w) & 0xFFFFFFFFFFFFFFFF
    assert s == dig
