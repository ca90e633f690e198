"""The renderer curve payload on the GPU (nxg_curve_payload_build) against
the reference's build_curve_batches outputs (tests/golden/reference_payload.npz,
inputs regenerated by make_golden.payload_cases): every batch's points,
point counts, root uvs and motion points bit for bit, and the batch table."""
import os
import sys

import numpy as np
import pytest

from nanoxgen_amd import payload

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import payload_cases  # noqa: E402


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def fixtures():
    with np.load(os.path.join(HERE, "golden", "reference_payload.npz"), allow_pickle=False) as d:
        return {k: d[k] for k in d.keys()}


def dev(torch, d):
    out = {}
    for k, v in d.items():
        if isinstance(v, np.ndarray):
            out[k] = torch.from_numpy(np.ascontiguousarray(v.view(np.int32) if v.dtype == np.uint32 else v)).cuda()
        elif v is not None:
            out[k] = v
    return out


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


@pytest.mark.parametrize("case", range(4))
def test_reference_fixtures(torch, fixtures, case):
    name, base, motion, opts = payload_cases()[case]
    print(f"payload case {name}", flush=True)
    out = payload.curve_payload(dev(torch, base), [dev(torch, m) for m in motion],
                                max_strands_per_batch=opts["max_strands"], resample_cv_count=opts["resample"],
                                radius_scale=opts.get("radius_scale", 1.0),
                                transform=opts.get("transform", payload.IDENTITY))
    assert int(out["status"].item()) == 0
    want = lambda k: fixtures[f"{name}_out_{k}"]
    n = len(want("point_counts"))
    total = len(want("points"))
    assert np.array_equal(bits(out["points"][:total].cpu().numpy()), bits(want("points")))
    assert np.array_equal(out["point_counts"][:n].cpu().numpy().view(np.uint32), want("point_counts"))
    if f"{name}_out_root_uvs" in fixtures:
        assert np.array_equal(bits(out["root_uvs"][:n].cpu().numpy()), bits(want("root_uvs")))
    if motion:
        assert np.array_equal(bits(out["motion_points"][:total * len(motion)].cpu().numpy()),
                              bits(want("motion_points")))
    table = payload.batches(out, opts["max_strands"], n)
    assert [b[0] for b in table] == list(want("batch_first"))
    counts = want("point_counts").astype(np.int64)
    assert [b[2] for b in table] == [int(counts[:f].sum()) for f in want("batch_first")]
    # a second call into the first call's buffers writes the same bytes
    points, motion_points = out["points"].clone(), out["motion_points"].clone()
    out["points"].zero_()
    out["motion_points"].zero_()
    again = payload.curve_payload(dev(torch, base), [dev(torch, m) for m in motion],
                                  max_strands_per_batch=opts["max_strands"], resample_cv_count=opts["resample"],
                                  radius_scale=opts.get("radius_scale", 1.0),
                                  transform=opts.get("transform", payload.IDENTITY), out=out)
    assert again is out and int(out["status"].item()) == 0
    assert torch.equal(out["points"], points)
    if motion:  # without motion samples the buffer is a one-row placeholder nothing writes
        assert torch.equal(out["motion_points"], motion_points)


def test_roots_give_the_uvs(torch):
    import torch as T
    rng = np.random.default_rng(2)
    n = 300
    roots = np.zeros((n, 12), np.float32)
    roots[:, 6:8] = rng.uniform(size=(n, 2))
    pts = T.from_numpy(rng.normal(size=(n * 3, 3)).astype(np.float32)).cuda()
    a = payload.curve_payload(dict(fixed=3, points=pts, constant_width=0.01,
                                   roots=T.from_numpy(roots.view(np.int32)).cuda()))
    assert np.array_equal(a["root_uvs"].cpu().numpy(), roots[:, 6:8])


def test_bad_counts_write_nothing(torch):
    import torch as T
    counts = T.tensor([2, 1, 3], dtype=T.int32, device="cuda")
    pts = T.zeros((6, 3), dtype=T.float32, device="cuda")
    out = payload.curve_payload(dict(counts=counts, points=pts, constant_width=0.01))
    assert int(out["status"].item()) == 8
