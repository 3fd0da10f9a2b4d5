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
