"""The C-ABI library loads and exports every entry point of include/ve_svo.h
(no device calls)."""
import ctypes as C


def test_exports_all_declared_symbols(ve):
    L = ve.lib()
    names = ve.exported_symbols()
    assert "ve_frame" in names and "ve_octree_build" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_host_only_calls(ve):
    fc = ve.FrameConfig()
    assert fc.primary_controlchecks == 70 and fc.shading_budget == 131072 and abs(fc.weight_variance - 0.9) < 1e-7
    assert ve.lib().ve_create(None, None) == ve.VE_EINVAL
    assert b"null" in ve.lib().ve_last_error(None)
    assert ve.lib().ve_read(None, 0, None, 0) == ve.VE_EINVAL


def test_errors_are_codes_not_crashes(ve):
    L = ve.lib()
    out = C.POINTER(C.c_uint32)()
    info = ve.OctreeInfo()
    assert L.ve_octree_build(None, 0, 0, 0, C.byref(out), C.byref(info)) == ve.VE_EINVAL
    assert L.ve_octree_build(None, 0, 17, 0, C.byref(out), C.byref(info)) == ve.VE_EINVAL


def test_headline_knobs_are_the_benchmarks():
    """The GPU parity tests of the measured workload (test_gpu_headline.py)
    run the benchmark's own configuration."""
    import bench
    from frame_common import HEADLINE_KNOBS, HEADLINE_SKY
    import test_gpu_headline as t
    assert bench.knobs() == HEADLINE_KNOBS
    assert (bench.W, bench.H, bench.BOUNCES + 1, bench.ORBIT_FRAMES) == (t.W, t.H, t.MAX_BOUNCES, t.ORBIT)
    assert tuple(bench.SKY) == HEADLINE_SKY
