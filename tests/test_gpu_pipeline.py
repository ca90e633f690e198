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


def test_pipeline_deformed_frames_with_after(torch, plane):
    """Frames of a deformed patch written into one device buffer on the
    caller's stream between calls: each call's preparation follows the write
    through `after` and every frame equals a serial call on its own copy."""
    import numpy as np
    w, asset = plane
    blob = nx.build_asset(w.build)
    h = abi.header_of(blob)
    rest = blob[h.positions_offset:h.positions_offset + 12 * h.vertex_count].view(np.float32).reshape(-1, 3)
    frames = [(rest + np.float32(0.02 * (f + 1)) * np.sin(rest[:, [1, 2, 0]] * np.float32(3 + f))).astype(np.float32)
              for f in range(4)]
    src = [torch.from_numpy(np.ascontiguousarray(f)).cuda() for f in frames]

    def desc(t):
        return abi.DeformedGeometry(positions=t.data_ptr(), position_count=t.shape[0])

    want = []
    for t in src:
        out = nx.PackedOutputs(w.params.strand_count, w.params.cvs_per_strand, roots=False, root_uvs=False,
                               point_counts=True)
        nx.generate_packed(asset, w.params, out, octaves=w.octaves, deformed=desc(t))
        want.append(out.points.clone())
    torch.cuda.synchronize()
    assert not torch.equal(want[0], want[1])
    buf = torch.empty_like(src[0])
    pipe = GenerationPipeline(asset, w.params, octaves=w.octaves, deformed=desc(buf))
    outs = []
    for t in src:
        buf.copy_(t)  # on the caller's stream, after the previous strand kernel that read buf
        written = torch.cuda.Event()
        written.record(pipe.stream)
        o = nx.PackedOutputs(w.params.strand_count, w.params.cvs_per_strand, roots=False, root_uvs=False,
                             point_counts=True)
        pipe.submit(o, after=written)
        outs.append(o)
    pipe.drain()
    torch.cuda.synchronize()
    for f, (o, p) in enumerate(zip(outs, want)):
        assert torch.equal(o.points, p), f"frame {f} differs from its serial call"


@pytest.mark.parametrize("after_prepare", [False, True])
def test_pipeline_with_attribute_programs(torch, after_prepare):
    """Expression attributes evaluated per call into the pipeline's slot on
    the side stream, before or after the call's preparation: every call
    equals attributes + generate_packed run serially, for several seeds in
    flight."""
    from nanoxgen_amd import expr
    w = workloads.config3(strands=40_000)
    asset = nx.DeviceAsset(nx.build_asset(w.build))
    es, maps = expr.ExprSet(w.info["programs"]), expr.TexelMaps(w.info["maps"])
    seeds = [3, 4, 5, 6]
    want = []
    for s in seeds:
        p = with_seed(w.params, s)
        attrs = torch.empty((p.strand_count, 4), dtype=torch.float32, device="cuda")
        expr.evaluate_root_attributes(asset, p, es, maps, attrs)
        out = nx.PackedOutputs(p.strand_count, p.cvs_per_strand, roots=False, root_uvs=False, point_counts=True)
        nx.generate_packed(asset, p, out, octaves=w.octaves, attributes=attrs)
        want.append(out.points.clone())
    torch.cuda.synchronize()

    seen = []

    def attribute_fn(params, out_attrs, side, *workspace):
        seen.append(len(workspace))
        expr.evaluate_root_attributes(asset, params, es, maps, out_attrs, stream=side)

    pipe = GenerationPipeline(asset, w.params, octaves=w.octaves, attribute_fn=attribute_fn,
                              attributes_after_prepare=after_prepare)
    outs = []
    for s in seeds:
        o = nx.PackedOutputs(w.params.strand_count, w.params.cvs_per_strand, roots=False, root_uvs=False,
                             point_counts=True)
        pipe.submit(o, with_seed(w.params, s))
        outs.append(o)
    pipe.drain()
    torch.cuda.synchronize()
    assert seen == [1 if after_prepare else 0] * len(seeds)  # the workspace only after the preparation
    for s, o, p in zip(seeds, outs, want):
        assert torch.equal(o.points, p), f"seed {s}: points differ from the serial call"
