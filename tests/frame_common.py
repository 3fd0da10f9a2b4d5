"""Shared helpers for the frame-level parity tests (oracle and GPU)."""
import json
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# Radiance tolerance of the resolve output (AvgPass): its exp/pow/log cores
# differ from llvmpipe's by a few ulp, everything else is exact.
OUT_RTOL = 1e-4
OUT_ATOL = 1e-6


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
