"""The opt-in inverse-CDF root sampler on the GPU (nxg_cdf_sampler_build,
nxg_generate_packed_cdf) against its CPU restatement (oracle/oracle_cdf.c).

Bars: the sampler's prefix sums and the sampled roots (the whole 48-byte
RootSample) bit-exact; the strands generated from them meet the
generator's bar against the oracle generating from the same roots (root
records, stencil slots, counts and radii bit-exact, positions within 1e-5
of the bounding-box diagonal); the grouped-by-triangle processing the CDF
mode uses gives the same bits as the caller-order generate-from-roots path."""
import numpy as np
import pytest

import nanoxgen_amd as nx
from nanoxgen_amd import abi, workloads
from test_gpu_generate import compare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def tiny_description():
    """Three triangles (a T below one coarse stride and one scan block)."""
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1], [1, 0, 1], [2, 0, 0]], np.float32)
    tris = np.array([[0, 2, 1], [1, 2, 3], [1, 3, 4]], np.uint32)
    gcv = np.array([[0.2, 0, 0.2], [0.2, 0.5, 0.2], [0.8, 0, 0.5], [0.8, 0.4, 0.5]], np.float32)
    return abi.BuildInputArrays(pos, tris, gcv, np.array([2, 2], np.uint32), stencil_k=2,
                                triangle_weights=np.array([1.0, 0.0, 3.0], np.float32))


@pytest.mark.parametrize("which", ["tiny", "icosphere", "plane"])
def test_prefix_sums_match_oracle(torch, orc, which):
    build = {"tiny": tiny_description, "icosphere": lambda: workloads.config1().build,
             "plane": lambda: workloads.config2().build}[which]()
    blob = nx.build_asset(build)
    sampler = nx.CdfSampler(nx.DeviceAsset(blob))
    gpu = sampler.prefix_sums()
    ref = orc.cdf_prefix(blob)
    assert np.array_equal(gpu, ref), f"{which}: first difference at {np.flatnonzero(gpu != ref)[:4]}"


@pytest.mark.parametrize("case", [
    dict(strands=10_000),
    dict(strands=3001, seed=99, noise_amplitude=0.0),
    dict(strands=1),
])
def test_cdf_generation_matches_oracle(torch, orc, case):
    w = workloads.config1(strands=case["strands"])
    p = w.params.copy(**{k: v for k, v in case.items() if k != "strands"})
    blob = nx.build_asset(w.build)
    asset = nx.DeviceAsset(blob)
    sampler = nx.CdfSampler(asset)
    out = nx.PackedOutputs(p.strand_count, p.cvs_per_strand, stencil=True)
    nx.generate_packed(asset, p, out, sampler=sampler)
    torch.cuda.synchronize()
    gpu = out.to_host()
    roots = orc.cdf_roots(blob, p.seed, p.strand_count)
    assert np.array_equal(gpu["roots"].view(np.uint8), roots.view(np.uint8)), "CDF roots differ"
    ref = orc.generate(blob, p, roots=roots, stencil=True)
    compare(gpu, ref, label=f"cdf {case}")


def test_cdf_plane_full_size_and_order_independence(torch, orc):
    """The 2,000,000-strand plane with two octaves through the CDF mode:
    roots bit-exact on every strand, the strands bit-identical to the
    caller-order generate-from-roots path on the same roots, the oracle on
    4,096 sampled strands, and both modes' triangle histograms against the
    masses (the same distribution)."""
    w = workloads.config2()
    N, Cv = w.params.strand_count, w.params.cvs_per_strand
    blob = nx.build_asset(w.build)
    asset = nx.DeviceAsset(blob)
    sampler = nx.CdfSampler(asset)
    out = nx.PackedOutputs(N, Cv)
    nx.generate_packed(asset, w.params, out, octaves=w.octaves, sampler=sampler)
    plain = nx.PackedOutputs(N, Cv)
    nx.generate_packed(asset, w.params, plain, octaves=w.octaves, roots=out.roots)
    alias = nx.PackedOutputs(N, Cv)
    nx.generate_packed(asset, w.params, alias, octaves=w.octaves)
    torch.cuda.synchronize()
    assert torch.equal(out.points, plain.points), "processing order changed the CDF mode's strands"
    roots = out.roots.cpu().numpy().view(abi.ROOT_DTYPE).reshape(-1)
    ref_roots = orc.cdf_roots(blob, w.params.seed, N, sampler.prefix_sums())
    assert np.array_equal(roots.view(np.uint8), ref_roots.view(np.uint8)), "CDF roots differ"
    sample = np.linspace(0, N - 1, 4096).astype(np.int64)
    ref = orc.generate(blob, w.params.copy(strand_count=len(sample)), octaves=w.octaves, roots=ref_roots[sample])
    idx = (sample[:, None] * Cv + np.arange(Cv)[None, :]).ravel()
    pts = out.points.cpu().numpy()
    gpu = dict(points=pts[idx], roots=roots[sample], root_uvs=out.root_uvs.cpu().numpy()[sample],
               point_counts=out.point_counts.cpu().numpy().view(np.uint32)[sample])
    compare(gpu, ref, label="cdf plane sample")
    # same distribution: 4,096 triangle bins of each mode's counts against
    # the bins' shares of the prefix sums (the two modes' draws share their
    # hashes and are correlated, so each is tested against the masses)
    T = abi.header_of(blob).triangle_count
    pre = sampler.prefix_sums()
    bins = T // 4096
    edges = pre[bins - 1::bins].astype(np.float64)
    expected = N * np.diff(edges, prepend=0.0) / float(T << 32)
    a_tri = alias.roots.cpu().numpy().view(abi.ROOT_DTYPE).reshape(-1)["triangle_index"]
    for name, tri in (("cdf", roots["triangle_index"]), ("alias", a_tri)):
        counts = np.bincount(tri // bins, minlength=4096).astype(np.float64)
        chi2 = float(((counts - expected) ** 2 / expected).sum())
        z = (chi2 - 4095) / np.sqrt(2 * 4095)
        print(f"{name} triangle bins against the masses: chi-square {chi2:.1f} over 4095 (z = {z:.2f})")
        assert abs(z) < 5.0, name
