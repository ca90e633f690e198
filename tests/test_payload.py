"""Renderer curve payload: the plan's checks against the reference's
build_curve_batches messages, and the reference's own test values
(tests/test_main.cpp:357-415, 590-615) through oracle/_ref.  CPU only; the
GPU kernels are checked in test_gpu_payload.py against reference_payload.npz."""
import os
import sys

import numpy as np
import pytest
import torch

from nanoxgen_amd import payload

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import payload_cases  # noqa: E402


def t(a, dtype=None):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a if dtype is None else a.astype(dtype)))


def tensors(d):
    out = {k: v for k, v in d.items() if not isinstance(v, np.ndarray)}
    for k, v in d.items():
        if isinstance(v, np.ndarray):
            out[k] = t(v.view(np.int32) if v.dtype == np.uint32 else v)
    return out


def test_reference_test_values(orc):
    if orc.ref() is None:
        pytest.skip("reference core not built here")
    name, base, motion, opts = payload_cases()[0]
    r, err = orc.ref_curve_payload(base, motion, **opts)
    assert err == "" and list(r["batch_first"]) == [0, 1] and list(r["point_counts"]) == [4, 4]
    assert r["points"][0, 0] == 10.0 and r["points"][3, 1] == 1.0
    assert r["points"][0, 3] == np.float32(0.02) and r["points"][3, 3] == np.float32(0.04)
    assert r["root_uvs"][0, 0] == 0.25 and r["root_uvs"][1, 0] == 0.75
    m = r["motion_points"][:4]
    assert m[0, 0] == 10.0 and m[0, 1] == 0.5 and m[-1, 1] == 1.5


def test_64k_boundary(orc):
    if orc.ref() is None:
        pytest.skip("reference core not built here")
    n = 64 * 1024 + 1
    pts = np.zeros((2 * n, 3), np.float32)
    pts[0::2, 0] = pts[1::2, 0] = np.arange(n)
    pts[1::2, 1] = 1.0
    r, _ = orc.ref_curve_payload(dict(fixed=2, points=pts, constant_width=0.02,
                                      root_uvs=np.stack([np.arange(n), np.zeros(n)], 1).astype(np.float32)))
    assert list(r["batch_first"]) == [0, 65536]
    assert payload.plan(dict(fixed=2, points=t(pts)))[0] == 2
    assert r["root_uvs"][65536, 0] == 65536.0


def test_plan_messages_match_reference(orc):
    if orc.ref() is None:
        pytest.skip("reference core not built here")
    base = dict(counts=np.uint32([2, 3]), points=np.zeros((5, 3), np.float32), constant_width=0.01)
    cases = [
        (base, [], dict(max_strands=0)),
        (base, [], dict(resample=1)),
        (base, [], dict(radius_scale=-1.0)),
        (dict(base, constant_width=float("nan")), [], {}),
        (dict(fixed=1, points=np.zeros((4, 3), np.float32)), [], {}),
        (dict(fixed=2, points=np.zeros((5, 3), np.float32)), [], {}),
        (base, [dict(base, time=float("inf"))], {}),
        (base, [dict(counts=np.uint32([2]), points=np.zeros((2, 3), np.float32), time=0.0)], {}),
        (base, [], dict(transform=(1.0, 0, 0, float("nan"), 0, 1, 0, 0, 0, 0, 1, 0))),
    ]
    for b, m, o in cases:
        _, want = orc.ref_curve_payload(b, m, **o)
        assert want
        kw = dict(max_strands_per_batch=o.get("max_strands", 65536), resample_cv_count=o.get("resample", 0),
                  radius_scale=o.get("radius_scale", 1.0), transform=o.get("transform", payload.IDENTITY))
        with pytest.raises(ValueError) as e:
            payload.plan(tensors(b), [tensors(x) for x in m], **kw)
        assert str(e.value) == want


def test_plan_sizes():
    name, base, motion, opts = payload_cases()[3]
    b, total = payload.plan(tensors(base), [tensors(m) for m in motion], max_strands_per_batch=opts["max_strands"],
                            resample_cv_count=opts["resample"])
    assert b == 3 and total == 700 * 7
