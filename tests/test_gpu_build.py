"""Asset build with the stencil search on the GPU (nxg_asset_build_gpu)
against the host builder, which is itself byte-identical to the reference's
build_asset: the host blob and the device copy must both equal the host
builder's blob byte for byte (stencil rows, tie order and content hash)."""
import time

import numpy as np
import pytest

import nanoxgen_amd as nx
from nanoxgen_amd import abi, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def check_same(torch, build):
    t0 = time.perf_counter()
    want = nx.build_asset(build)
    t1 = time.perf_counter()
    blob, dev = nx.build_asset_gpu(build)
    t2 = time.perf_counter()
    assert np.array_equal(blob, want)
    assert np.array_equal(dev.tensor.cpu().numpy(), want)
    return t1 - t0, t2 - t1


def test_icosphere_k4(torch):
    check_same(torch, workloads.icosphere_description())


def test_plane_k8_reduced(torch):
    check_same(torch, workloads.plane_description(quads=64, guides_per_side=16))


def test_ties_and_penalty(torch):
    """Duplicated guides give equal keys (the lower index must come first) and
    guides facing away get the 1e20 penalty; fewer guides than k leave
    invalid slots."""
    base = workloads.icosphere_description(guides=6)
    cvs = base.guide_cvs.reshape(6, -1, 3)
    nrm = base.guide_root_normals
    cvs2 = np.concatenate([cvs, cvs[:3], cvs[:1]])  # guides 6..9 repeat 0..2 and 0
    nrm2 = np.concatenate([nrm, -nrm[:3], nrm[:1]])  # 6..8 face away
    build = abi.BuildInputArrays(base.positions, base.triangles, cvs2.reshape(-1, 3),
                                 np.full(len(cvs2), cvs.shape[1]), texcoords=base.texcoords,
                                 guide_root_normals=nrm2, stencil_k=8)
    check_same(torch, build)
    few = abi.BuildInputArrays(base.positions, base.triangles, cvs[:3].reshape(-1, 3), np.full(3, cvs.shape[1]),
                               guide_root_normals=nrm[:3], stencil_k=8)
    check_same(torch, few)


def test_full_plane(torch):
    """524,288 triangles x 4,096 guides, K=8."""
    host_s, gpu_s = check_same(torch, workloads.plane_description())
    print(f"plane build: host {host_s:.3f} s, with GPU stencil {gpu_s:.3f} s")
