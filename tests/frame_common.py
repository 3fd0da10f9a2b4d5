"""Shared helpers for the frame-level parity tests (oracle and GPU)."""
import json
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# Radiance tolerance of the resolve output (AvgPass): its exp/pow/log cores
# differ from llvmpipe's by a few ulp, everything else is exact.
OUT_RTOL = 1e-4
OUT_ATOL = 1e-6


# Knobs under which the whole frame is order-independent (one sample per
# visible voxel per frame, no stale reset, caps above the sample counts, no
# firefly clamp), so every integer structure is bit-exact.
DET_KNOBS = dict(primary_controlchecks=70, bounce_controlchecks=120, normal_precision=1, shading_budget=196608,
                 weight_pixels=0.0, weight_variance=0.0, top_fraction=0.0, mid_fraction=1.0, min_samples=1,
                 sample_cap_direct=1024, sample_cap_indirect=4096, disable_stale_reset=1, update_time=0xFFFFFFFF,
                 firefly_k=1.0, firefly_floor=1e30)


def load_frame_gold():
    with open(os.path.join(GOLD, "frame_small.json")) as fh:
        meta = json.load(fh)
    return meta, np.load(os.path.join(GOLD, "frame_small.npz"))


def frame_seed(f):
    return (1 + 7919 * f) & 0xFFFFFFFF


def hash_map(bucket_major, stride, slots):
    b = np.asarray(bucket_major, np.uint32).reshape(-1, slots, stride)
    own = b[:, :, 0]
    sel = own != 0
    vids = own[sel]
    vals = b[:, :, 1:][sel]
    order = np.argsort(vids, kind="stable")
    return vids[order], vals[order]


def assert_out_close(out, ref):
    out = np.asarray(out, np.float32).reshape(-1, 4)
    ref = np.asarray(ref, np.float32).reshape(-1, 4)
    assert np.all(np.isfinite(out))
    np.testing.assert_allclose(out, ref, rtol=OUT_RTOL, atol=OUT_ATOL)


# The measured workload's knobs: a still camera's accumulation settings
# (voxelengine.cpp:201-207), central-difference normals, the full shading
# budget and one sample per visible voxel per frame.  bench.py's knobs()
# returns the same dict (checked by test_headline_knobs_are_the_benchmarks).
HEADLINE_KNOBS = dict(primary_controlchecks=70, bounce_controlchecks=120, normal_precision=1, shading_budget=196608,
                      weight_pixels=0.0, weight_variance=0.0, top_fraction=0.0, mid_fraction=1.0, min_samples=1,
                      sample_cap_direct=1024, sample_cap_indirect=4096, disable_stale_reset=1,
                      update_time=0xFFFFFFFF)
HEADLINE_SKY = (140, 180, 230)


def compare_frame(ve, rend, of, W, H, label=""):
    """Every structure of one frame, GPU (rend) against the oracle (of):
    G-buffer, packed positions, counters, candidate list, shade list,
    per-voxel pixel counts, claim map, RankStats, l/nBuffer maps bit-exact;
    the resolved image within the radiance tolerance."""
    g, p = rend.gbuffer(W, H)
    assert np.array_equal(g.reshape(-1), of.read("GBUFFER")), "G-buffer " + label
    assert np.array_equal(p.reshape(-1), of.read("GBUFFER_POS")), "packed positions " + label
    cg, co = rend.counters(), of.counters()
    assert cg[ve.CNT["OVERFLOW"]] == 0, "claim-row overflow " + label
    for i in (0, 1, 2, 4, 5, 6, 7):
        assert cg[i] == co[i], "counter %d: %d vs %d %s" % (i, cg[i], co[i], label)
    nc, ns = int(cg[0]), int(cg[1])
    assert np.array_equal(rend.read("CANDIDATES", np.uint32, nc * 16), of.read("CANDIDATES")), "candidates " + label
    assert np.array_equal(rend.read("SHADE_LIST", np.uint32, ns * 8), of.read("SHADE_LIST")), "shade list " + label
    assert np.array_equal(rend.read("SAMPLE_COUNTS"), of.read("SAMPLE_COUNTS")), "pixel counts " + label
    assert np.array_equal(rend.read("CLAIM_MAP"), of.read("CLAIM_MAP")), "claim map " + label
    assert np.array_equal(rend.read("RANK_STATS"), of.read("RANK_STATS")), "RankStats " + label
    for name, stride, slots in (("LBUFFER", 12, 32), ("NBUFFER", 3, 32)):
        gv, gval = hash_map(rend.read(name), stride, slots)
        ov, oval = hash_map(of.read(name), stride, slots)
        assert np.array_equal(gv, ov) and np.array_equal(gval, oval), name + " " + label
    assert_out_close(rend.read("OUTPUT", np.float32, W * H * 16), of.read("OUTPUT", np.float32))
    return cg
