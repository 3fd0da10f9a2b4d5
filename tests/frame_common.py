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
HEADLINE_FRAMES = 4        # orbit frames 0..3 of the llvmpipe headline fixture
HEADLINE_SAMPLES = 4096    # lBuffer entries and output pixels sampled per frame


def sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def headline_digests(gbuf, pos, cand, shade, rank_stats, nbuf_map, lbuf_map):
    """Digests of one headline frame's exact structures (the layouts ve_read
    and the llvmpipe replay share)."""
    return dict(gbuf=sha(np.asarray(gbuf, np.uint32)), pos=sha(np.asarray(pos, np.uint32)),
                candidates=sha(np.asarray(cand, np.uint32)), shade_list=sha(np.asarray(shade, np.uint32)),
                rank_stats=sha(np.asarray(rank_stats, np.uint32)),
                nbuf=sha(np.concatenate([nbuf_map[0].reshape(-1), nbuf_map[1].reshape(-1)])),
                lbuf=sha(np.concatenate([lbuf_map[0].reshape(-1), lbuf_map[1].reshape(-1)])),
                lbuf_entries=int(len(lbuf_map[0])))


def load_headline_gold():
    with open(os.path.join(GOLD, "headline_llvmpipe.json")) as fh:
        meta = json.load(fh)
    return meta, np.load(os.path.join(GOLD, "headline_llvmpipe.npz"))


def compare_headline_frame(meta, arr, f, counters, digests, lbuf_map, out):
    """One frame of the headline workload against the llvmpipe fixture.
    Counters, G-buffer, positions, candidate and shade lists, RankStats and
    nBuffer bit-exact.  lBuffer: the whole map equals llvmpipe's once the
    fixture's recorded ulp flips are put back -- entries one unit apart
    because a float deposit sat on a quantisation boundary that an ulp of
    llvmpipe's log/cos cores decides (the oracle and the GPU evaluate those
    cores in double, rounded once); each flip must hold the value recorded for
    the oracle.  Output within the radiance tolerance at the sampled pixels,
    except those the recorded flips reach through resolve."""
    fm = meta["frames"][f]
    assert [int(c) for c in counters] == fm["counters"], (counters, fm["counters"])
    for k in ("gbuf", "pos", "candidates", "shade_list", "rank_stats", "nbuf"):
        assert digests[k] == fm["digests"][k], "frame %d %s" % (f, k)
    vids, vals = lbuf_map
    assert len(vids) == fm["digests"]["lbuf_entries"]
    fv, f_llvm, f_oracle = (arr["f%d_flip_vids" % f], arr["f%d_flip_llvm" % f], arr["f%d_flip_oracle" % f])
    assert len(fv) <= LBUF_ULP_FLIPS
    at = np.searchsorted(vids, fv)
    assert np.array_equal(vids[at], fv) and np.array_equal(vals[at], f_oracle), "lBuffer flips, frame %d" % f
    put_back = vals.copy()
    put_back[at] = f_llvm
    assert sha(np.concatenate([vids.reshape(-1), put_back.reshape(-1)])) == fm["digests"]["lbuf"], \
        "lBuffer beyond the recorded flips, frame %d" % f
    px = arr["f%d_out_pixels" % f]
    got = np.asarray(out, np.float32).reshape(-1, 4)[px]
    ok = ~arr["f%d_out_flip" % f]
    np.testing.assert_allclose(got[ok], arr["f%d_out_values" % f][ok], rtol=OUT_RTOL, atol=OUT_ATOL)


# entries per frame allowed to differ from llvmpipe by one unit (observed: 1)
LBUF_ULP_FLIPS = 4
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
