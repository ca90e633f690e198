"""Batched uneven descriptions through the C ABI (nxg_generate_batch and the
captured graph) against the single-description path and the CPU oracle.

Bars: every description's outputs from the batched launch are bit-identical
to nxg_generate_packed_attributed for that description alone, whatever the
tuning (wavefronts per block, staged CVs, tiles per claim, order), and from
every replay of the captured graph; against the oracle on a sample of each
description, root records, stencil slots, counts and radii are bit-exact
and positions within 1e-5 of the bounding-box diagonal (attributes from the
GPU expression runtime, which is itself checked against the oracle)."""
import numpy as np
import pytest

import nanoxgen_amd as nx
from nanoxgen_amd import abi, batch, expr, workloads
from test_gpu_generate import compare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def small_batch(torch):
    """Config-5 construction at 1/100 of the strands and reduced meshes."""
    ws = workloads.config5(strand_divisor=100, mesh_divisor=4)
    blobs = [nx.build_asset(w.build) for w in ws]
    assets = [nx.DeviceAsset(b) for b in blobs]
    return ws, blobs, assets


def make_batch(torch, ws, assets, stencil=True):
    b = batch.Batch()
    outs = []
    for w, a in zip(ws, assets):
        out = nx.PackedOutputs(w.params.strand_count, w.params.cvs_per_strand, stencil=stencil)
        b.add(a, w.params, out, octaves=w.octaves, exprs=expr.ExprSet(w.info["programs"]),
              maps=expr.TexelMaps(w.info["maps"]))
        outs.append(out)
    return b, outs


def single_results(torch, ws, assets, b):
    """Each description alone through the single-description entry points,
    from the attributes the batch evaluated."""
    res = []
    for w, a, e in zip(ws, assets, b.entries):
        out = nx.PackedOutputs(w.params.strand_count, w.params.cvs_per_strand, stencil=True)
        nx.generate_packed(a, w.params, out, octaves=w.octaves, attributes=e["attributes"])
        res.append(out)
    torch.cuda.synchronize()
    return res


def same(torch, x, y):
    return all(torch.equal(getattr(x, f), getattr(y, f))
               for f in ("points", "roots", "root_uvs", "point_counts", "stencil_indices", "stencil_weights"))


def test_batch_matches_single_and_oracle(torch, orc, small_batch):
    ws, blobs, assets = small_batch
    b, outs = make_batch(torch, ws, assets)
    b.generate()
    torch.cuda.synchronize()
    singles = single_results(torch, ws, assets, b)
    for w, blob, e, out, single in zip(ws, blobs, b.entries, outs, singles):
        assert same(torch, out, single), f"{w.name}: batch differs from the single-description launch"
        # attributes evaluated in the batch equal the expression runtime's
        attrs = torch.empty_like(e["attributes"])
        expr.evaluate_root_attributes(e["asset"], w.params, e["exprs"], e["maps"], attrs)
        torch.cuda.synchronize()
        assert torch.equal(attrs, e["attributes"])
        n = w.params.strand_count
        sample = np.unique(np.linspace(0, n - 1, min(n, 400)).astype(np.int64))
        # the oracle takes attributes in strand-list order
        ref = orc.generate(blob, w.params, octaves=w.octaves, attributes=e["attributes"].cpu().numpy()[sample],
                           strands=sample, stencil=True)
        gpu = out.to_host()
        c = w.params.cvs_per_strand
        idx = (sample[:, None] * c + np.arange(c)[None, :]).ravel()
        sub = dict(points=gpu["points"][idx], roots=gpu["roots"][sample], root_uvs=gpu["root_uvs"][sample],
                   point_counts=gpu["point_counts"][sample], stencil_indices=gpu["stencil_indices"][sample])
        compare(sub, ref, label=w.name)


def test_every_candidate_tuning_gives_identical_outputs(torch, small_batch):
    ws, _, assets = small_batch
    b, outs = make_batch(torch, ws, assets)
    b.generate()
    torch.cuda.synchronize()
    want = [o.points.clone() for o in outs]
    counts = [o.point_counts.clone() for o in outs]
    for t in batch.tuning_candidates() + [abi.BatchTuning(2, 8, 16, 1, 1), abi.BatchTuning(1, 4, 3, 2, 0)]:
        for o in outs:
            o.points.fill_(float("nan"))
            o.point_counts.fill_(-1)
        b.generate(tuning=t)
        torch.cuda.synchronize()
        for k, o in enumerate(outs):
            assert torch.equal(o.points, want[k]) and torch.equal(o.point_counts, counts[k]), f"{t} item {k}"


def test_graph_replays_match_eager(torch, small_batch):
    """The captured graph, replayed three times with eager work and
    synchronisation between replays, reproduces the eager batch each time."""
    ws, _, assets = small_batch
    b, outs = make_batch(torch, ws, assets)
    b.generate()
    torch.cuda.synchronize()
    want = [o.points.clone() for o in outs]
    g = b.graph()
    for replay in range(3):
        for o in outs:
            o.points.zero_()
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        for k, o in enumerate(outs):
            assert torch.equal(o.points, want[k]), f"replay {replay} item {k}"
        # unrelated eager work between replays
        (torch.arange(1 << 20, device="cuda", dtype=torch.float32) * 2.0).sum().item()
    g.close()


def test_autotune_picks_a_valid_candidate(torch, small_batch):
    ws, _, assets = small_batch
    b, outs = make_batch(torch, ws, assets, stencil=False)
    cands = batch.tuning_candidates()[:6] + [abi.BatchTuning(4, 16, 1, 0, 0)]  # the last is invalid
    best, times = b.autotune(cands, repeats=3)
    assert times[-1][1] is None and all(t is not None and t > 0 for _, t in times[:-1])
    assert best.as_tuple() in [c.as_tuple() for c, _ in times[:-1]]
    assert min(t for _, t in times[:-1]) == dict((c.as_tuple(), t) for c, t in times[:-1])[best.as_tuple()]


def test_single_item_and_tiny_descriptions(torch, orc):
    """One-strand and 65-strand descriptions (a partial last tile) next to a
    larger one, no attributes, plain and noisy."""
    w = workloads.config1(strands=1000)
    blob = nx.build_asset(w.build)
    asset = nx.DeviceAsset(blob)
    b = batch.Batch()
    params = [w.params.copy(strand_count=1), w.params.copy(strand_count=65, noise_amplitude=0.0),
              w.params.copy(strand_count=1000, cvs_per_strand=5, seed=7)]
    outs = []
    for p in params:
        outs.append(nx.PackedOutputs(p.strand_count, p.cvs_per_strand, stencil=True))
        b.add(asset, p, outs[-1])
    b.generate()
    torch.cuda.synchronize()
    for p, o in zip(params, outs):
        compare(o.to_host(), orc.generate(blob, p, stencil=True), label=f"{p.strand_count} strands")


def test_config5_full_size(torch, orc):
    """The batched uneven workload at full size: nine descriptions, 2,456,139
    strands, 104,492,896 points, one batched launch.  Per description: the
    batch's outputs equal that description's single launch bit for bit; its
    attributes equal the expression runtime's and, on every root, meet the
    oracle's attribute bar; the oracle generates a 4,096-strand sample from
    the batch's attributes to the generator's bar; counts hold everywhere."""
    from test_gpu_expr import check_attributes
    ws = workloads.config5()
    assert sum(w.params.strand_count for w in ws) == 2_456_139
    assert sum(w.params.strand_count * w.params.cvs_per_strand for w in ws) == 104_492_896
    blobs = [nx.build_asset(w.build) for w in ws]
    assets = [nx.DeviceAsset(b) for b in blobs]
    b, outs = make_batch(torch, ws, assets)
    b.generate()
    torch.cuda.synchronize()
    for w, blob, a, e, out in zip(ws, blobs, assets, b.entries, outs):
        n, c = w.params.strand_count, w.params.cvs_per_strand
        single = nx.PackedOutputs(n, c, stencil=True)
        nx.generate_packed(a, w.params, single, octaves=w.octaves, attributes=e["attributes"])
        attrs = torch.empty_like(e["attributes"])
        expr.evaluate_root_attributes(e["asset"], w.params, e["exprs"], e["maps"], attrs)
        torch.cuda.synchronize()
        assert same(torch, out, single), f"{w.name}: batch differs from the single-description launch"
        assert torch.equal(attrs, e["attributes"]), f"{w.name}: batch attributes differ"
        del single, attrs
        gpu_attrs = e["attributes"].cpu().numpy()
        ref_attrs, fails = orc.root_attributes(blob, w.params, w.info["programs"], w.info["maps"])
        check_attributes(gpu_attrs, ref_attrs, w.info["programs"])
        gpu = out.to_host()
        assert np.all(gpu["point_counts"] == c), w.name
        sample = np.unique(np.linspace(0, n - 1, min(n, 4096)).astype(np.int64))
        ref = orc.generate(blob, w.params, octaves=w.octaves, attributes=gpu_attrs[sample], strands=sample,
                           stencil=True)
        idx = (sample[:, None] * c + np.arange(c)[None, :]).ravel()
        sub = dict(points=gpu["points"][idx], roots=gpu["roots"][sample], root_uvs=gpu["root_uvs"][sample],
                   point_counts=gpu["point_counts"][sample], stencil_indices=gpu["stencil_indices"][sample],
                   stencil_weights=gpu["stencil_weights"][sample])
        print(f"{w.name}: {fails} oracle attribute errors")
        compare(sub, ref, label=w.name)
        del gpu
