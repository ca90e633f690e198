"""The .nxc curve cache built from device-resident curves
(nxg_curve_cache_build_gpu): the same bytes as the host build of the same
arrays, which tests/test_curve_cache.py ties to the reference's
build_curve_cache byte for byte."""
import time

import numpy as np
import pytest

import nanoxgen_amd as nx
from nanoxgen_amd import abi, workloads
from nanoxgen_amd.cache import build_curve_cache, validate_curve_cache

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.parametrize("strands", [1, 777, 50000])
def test_generated_curves_cache(torch, strands):
    w = workloads.config2(strands=strands, quads=64, guides_per_side=16)
    blob = nx.build_asset(w.build)
    asset = nx.DeviceAsset(blob)
    out = nx.PackedOutputs(strands, 16)
    nx.generate_packed(asset, w.params, out, octaves=w.octaves)
    face_ids = out.roots[:, 8].contiguous()
    times = np.float32([0.0, 0.5])
    motion = torch.cat([out.points[:, :3], out.points[:, :3] + 0.25]).contiguous()
    print(f"curve cache from {strands} generated strands", flush=True)
    t0 = time.perf_counter()
    dev = build_curve_cache(out.point_counts, out.points, patch_uvs=out.root_uvs, face_ids=face_ids,
                            motion_times=times, motion_points=motion)
    t1 = time.perf_counter()
    host = build_curve_cache(out.point_counts.cpu().numpy().view(np.uint32), out.points.cpu().numpy(),
                             patch_uvs=out.root_uvs.cpu().numpy(), face_ids=face_ids.cpu().numpy().view(np.uint32),
                             motion_times=times, motion_points=motion.cpu().numpy())
    t2 = time.perf_counter()
    print(f"  gpu assembly + download + hash {1e3 * (t1 - t0):.1f} ms; host build from host copies "
          f"{1e3 * (t2 - t1):.1f} ms; {dev.nbytes} bytes", flush=True)
    assert np.array_equal(dev, host)
    assert validate_curve_cache(dev) == ""


def test_device_errors_match_host(torch):
    counts = torch.tensor([2, 1, 3], dtype=torch.int32, device="cuda")
    pts = torch.zeros((6, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError, match="fewer than two CVs"):
        build_curve_cache(counts, pts)
    counts = torch.tensor([2, 2, 3], dtype=torch.int32, device="cuda")
    with pytest.raises(ValueError, match="do not match point array"):
        build_curve_cache(counts, pts)
    counts = torch.tensor([2, 2, 2], dtype=torch.int32, device="cuda")
    pts[3, 3] = -1.0
    with pytest.raises(ValueError, match="built an invalid curve cache: curve cache contains an invalid point"):
        build_curve_cache(counts, pts)
