"""nxg_generate_packed_noalloc: the packed generation entry point that
allocates nothing and takes no workspace (the library's static device
scratch stands in for it).

It runs the same kernels as nxg_generate_packed, so its outputs must equal
that entry point's bit for bit: root records, stencil slots and weights,
point counts and every float of the points.  Two calls in a row share the
scratch, the second with a smaller request (another asset, fewer strands,
one octave) over what the first left there; a request whose workspace
exceeds the scratch is refused before any launch."""
import ctypes as C

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


def _outputs(params):
    return nx.PackedOutputs(params.strand_count, params.cvs_per_strand, stencil=True)


def _noalloc(asset, params, octaves, out):
    st = out.stencil_descriptor()
    desc = out.descriptor()
    return nx.lib().nxg_generate_packed_noalloc(
        C.byref(asset.desc), None, C.byref(params), C.byref(octaves) if octaves is not None else None,
        C.byref(desc), C.byref(st), None, nx._stream_handle(None))


def _same(a, b, label):
    for k in ("points", "roots", "root_uvs", "point_counts", "stencil_indices", "stencil_weights"):
        x, y = np.ascontiguousarray(a[k]), np.ascontiguousarray(b[k])
        assert x.shape == y.shape, f"{label}: {k} shape"
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), f"{label}: {k} differs"


def _case(torch, w, label):
    blob = nx.build_asset(w.build)
    asset = nx.DeviceAsset(blob)
    p = w.params
    need = asset.workspace_bytes(p)
    assert need <= abi.STATIC_SCRATCH_BYTES, f"{label}: fixture must fit the scratch"
    print(f"no-alloc case {label}: {p.strand_count} strands x {p.cvs_per_strand} CVs, workspace {need} B", flush=True)
    want = _outputs(p)
    nx.generate_packed(asset, p, want, octaves=w.octaves)
    got = _outputs(p)
    got.points.fill_(float("nan"))
    status = _noalloc(asset, p, w.octaves, got)
    assert status == abi.OK, nx.lib().nxg_status_message(status)
    torch.cuda.synchronize()
    _same(got.to_host(), want.to_host(), label)
    return asset


def test_noalloc_matches_workspace_entry_point(torch):
    # first call: a plane with two octaves and K = 8 (the larger workspace)
    big = workloads.config2(strands=300_000, quads=128, guides_per_side=32)
    _case(torch, big, "plane 300000x16, two octaves")
    # second call on the same scratch: another asset, fewer strands, one octave
    small = workloads.config1(strands=777)
    asset = _case(torch, small, "icosphere 777x8, one octave")
    # a request beyond the scratch is refused on the host, nothing launched
    huge = small.params.copy(strand_count=20_000_000)
    out = nx.PackedOutputs(1, small.params.cvs_per_strand)
    desc = out.descriptor()
    desc.point_capacity = huge.strand_count * huge.cvs_per_strand
    desc.root_capacity = desc.root_uv_capacity = desc.point_count_capacity = huge.strand_count
    status = nx.lib().nxg_generate_packed_noalloc(C.byref(asset.desc), None, C.byref(huge), None, C.byref(desc),
                                                 None, None, nx._stream_handle(None))
    assert status == abi.SCRATCH_LIMIT
    torch.cuda.synchronize()
