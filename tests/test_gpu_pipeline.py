"""GenerationPipeline: back-to-back generate_packed calls whose preparation
runs on a side stream, overlapping the previous call's strand kernel.

Bar: bitwise equality with serial generate_packed calls of the same
parameters (the same kernels run on the same inputs; only the scheduling
differs), for several seeds in flight at once, a shared output buffer and
every pipeline depth from 1 to 3."""
import ctypes as C

import pytest

import nanoxgen_amd as nx
from nanoxgen_amd import abi, workloads
from nanoxgen_amd.pipeline import GenerationPipeline

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def plane(torch):
    w = workloads.config2(strands=150_000, quads=96, guides_per_side=24)
    asset = nx.DeviceAsset(nx.build_asset(w.build))
    return w, asset


def with_seed(params, seed):
    p = abi.GenerationParams.from_buffer_copy(bytes(params))
    p.seed = seed
    return p


def serial(torch, w, asset, params):
    out = nx.PackedOutputs(params.strand_count, params.cvs_per_strand, roots=False, root_uvs=False,
                           point_counts=True)
    nx.generate_packed(asset, params, out, octaves=w.octaves)
    torch.cuda.synchronize()
    return out.points.clone(), out.point_counts.clone()


@pytest.mark.parametrize("depth", [1, 2, 3])
def test_pipeline_equals_serial_calls(torch, plane, depth):
    w, asset = plane
    seeds = [11, 12, 13, 14, 15]
    want = [serial(torch, w, asset, with_seed(w.params, s)) for s in seeds]
    pipe = GenerationPipeline(asset, w.params, octaves=w.octaves, depth=depth)
    outs = [nx.PackedOutputs(w.params.strand_count, w.params.cvs_per_strand, roots=False, root_uvs=False,
                             point_counts=True) for _ in seeds]
    for s, o in zip(seeds, outs):
        pipe.submit(o, with_seed(w.params, s))
    pipe.drain()
    torch.cuda.synchronize()
    for s, o, (pts, cnt) in zip(seeds, outs, want):
        assert torch.equal(o.points, pts), f"seed {s}: points differ from the serial call"
        assert torch.equal(o.point_counts, cnt), f"seed {s}: point counts differ from the serial call"
    # distinct seeds really produced distinct grooms
    assert not torch.equal(want[0][0], want[1][0])


def test_pipeline_shared_output_and_events(torch, plane):
    w, asset = plane
    pipe = GenerationPipeline(asset, w.params, octaves=w.octaves)
    out = nx.PackedOutputs(w.params.strand_count, w.params.cvs_per_strand, roots=False, root_uvs=False,
                           point_counts=True)
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(6)]
    for i in range(6):
        pipe.submit(out, with_seed(w.params, 20 + i), events=events[i])
    pipe.drain()
    torch.cuda.synchronize()
    pts, cnt = serial(torch, w, asset, with_seed(w.params, 25))
    assert torch.equal(out.points, pts) and torch.equal(out.point_counts, cnt)
    for e in events:
        assert e[0].elapsed_time(e[1]) > 0.0 and e[2].elapsed_time(e[3]) > 0.0


def test_pipeline_rejects_a_larger_call(torch, plane):
    w, asset = plane
    pipe = GenerationPipeline(asset, w.params, octaves=w.octaves)
    big = abi.GenerationParams.from_buffer_copy(bytes(w.params))
    big.strand_count = 4 * w.params.strand_count
    out = nx.PackedOutputs(big.strand_count, big.cvs_per_strand, roots=False, root_uvs=False, point_counts=True)
    with pytest.raises(ValueError, match="workspace bytes"):
        pipe.submit(out, big)
    assert C.sizeof(big) == C.sizeof(w.params)
