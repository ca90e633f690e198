"""Classic CutFX on the GPU (nxg_classic_cut_apply) against the reference's
CPU float runtime: its own outputs (tests/golden/reference_classic.npz) and
the oracle (oracle/oracle_classic.c, pinned to those outputs), bit for bit in
kept strands, point counts, roots, root uvs and positions."""
import numpy as np
import pytest

import nanoxgen_amd as nx
from nanoxgen_amd import abi, workloads
from nanoxgen_amd.classic import ClassicCutOutputs, add_endpoints, classic_cut

pytestmark = pytest.mark.gpu

CASES = [("ico5", 0.25, 0.0, 1.0), ("ico5", 0.0, 0.3, 1.0), ("ico13", 0.5, 0.05, 1.0),
         ("walk", 0.4, 0.1, 2.0), ("walk", 1.0, 0.0, 1.0), ("walk", 0.0, 0.0, 1.0)]


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def classic():
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_classic.npz")
    with np.load(path, allow_pickle=False) as d:
        return {k: d[k] for k in d.keys()}


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def dev(torch, a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    return t.cuda() if dtype is None else t.view(dtype).cuda()


def roots_dev(torch, roots):
    return torch.from_numpy(np.ascontiguousarray(roots).view(np.int32).reshape(-1, 12)).cuda()


def run(torch, pts, cvs, scale, offset, *, roots=None, uvs=None, counts=None, multiplier=None, stride=1,
        radius_scale=1.0):
    out = classic_cut(dev(torch, pts), cvs, scale, offset,
                      point_counts=None if counts is None else dev(torch, counts.view(np.int32)),
                      roots=None if roots is None else roots_dev(torch, roots),
                      root_uvs=None if uvs is None else dev(torch, uvs),
                      multiplier=None if multiplier is None else dev(torch, multiplier),
                      multiplier_stride=stride, radius_scale=radius_scale)
    return out.to_host()


def same(got, want, label):
    k = want["strand_count"]
    assert got["strand_count"] == k, f"{label}: kept {got['strand_count']} != {k}"
    assert np.array_equal(got["point_counts"], want["point_counts"]), f"{label}: point counts"
    if "roots" in got and want.get("roots") is not None:
        assert np.array_equal(bits(got["roots"]), bits(want["roots"])), f"{label}: roots"
    if "root_uvs" in got and want.get("root_uvs") is not None:
        assert np.array_equal(bits(got["root_uvs"]), bits(want["root_uvs"])), f"{label}: root uvs"
    if not np.array_equal(bits(got["points"]), bits(want["points"])):
        d = np.abs(got["points"].astype(np.float64) - want["points"])
        raise AssertionError(f"{label}: positions differ (max {d.max():.3g} at {np.unravel_index(d.argmax(), d.shape)})")


@pytest.mark.parametrize("i", range(len(CASES)))
def test_reference_fixtures(torch, classic, i):
    src, scale, offset, rscale = CASES[i]
    cvs = int(classic[f"{src}_cvs"])
    print(f"classic fixture case {i}: {src} cvs={cvs} scale={scale} offset={offset}", flush=True)
    got = run(torch, classic[f"{src}_points"], cvs, scale, offset, roots=classic[f"{src}_roots"],
              uvs=classic[f"{src}_uvs"], radius_scale=rscale)
    want = dict(strand_count=len(classic[f"case{i}_out_point_counts"]),
                points=classic[f"case{i}_out_points"], point_counts=classic[f"case{i}_out_point_counts"],
                roots=classic[f"case{i}_out_roots"], root_uvs=classic[f"case{i}_out_root_uvs"])
    assert got["status"] == 0
    same(got, want, f"fixture case {i}")


def random_strands(rng, n, cvs):
    steps = rng.normal(0, 1, size=(n, cvs, 3)) * rng.uniform(0, 0.4, size=(n, 1, 1))
    pts = np.cumsum(steps, axis=1).astype(np.float32)
    rad = rng.uniform(0.0, 0.05, size=(n, cvs, 1)).astype(np.float32)
    rad[::13] = 0.0
    pts[3::17, 1:] = pts[3::17, :1]
    return np.concatenate([pts, rad], axis=2).reshape(-1, 4)


@pytest.mark.parametrize("cvs", [2, 3, 8, 16, 24, 63, 64, 65, 100, 256])
def test_oracle_random_strands(torch, orc, cvs):
    """LDS-staged tiles up to 64 CVs, memory reads above; ragged last tile."""
    rng = np.random.default_rng(cvs)
    n = 20000 // cvs + 37
    pts = random_strands(rng, n, cvs)
    roots = np.zeros(n, abi.ROOT_DTYPE)
    roots["triangle_index"] = rng.integers(0, 1 << 20, size=n)
    roots["position"] = pts.reshape(n, cvs, 4)[:, 0, :3]
    uvs = rng.uniform(0, 1, size=(n, 2)).astype(np.float32)
    counts = rng.integers(1, 1000, size=n).astype(np.uint32)
    print(f"classic random cvs={cvs} strands={n}", flush=True)
    for scale, offset in ((0.3, 0.0), (0.0, 0.2), (0.9, 0.01), (-1.0, 0.0)):
        got = run(torch, pts, cvs, scale, offset, roots=roots, uvs=uvs, counts=counts)
        want = orc.classic_cut(pts, cvs, scale, offset, counts=counts, roots=roots, uvs=uvs)
        same(got, want, f"cvs {cvs} amount {scale} x cLength + {offset}")


def test_multiplier_column(torch, orc):
    """amount times one column of a float4-per-strand attribute buffer."""
    rng = np.random.default_rng(3)
    n, cvs = 5000, 12
    pts = random_strands(rng, n, cvs)
    attrs = rng.uniform(0.0, 2.0, size=(n, 4)).astype(np.float32)
    column = dev(torch, attrs)[:, 2]  # a strided view: strand s reads attrs[s, 2]
    got = classic_cut(dev(torch, pts), cvs, 0.3, 0.02, multiplier=column, multiplier_stride=4).to_host()
    want = orc.classic_cut(pts, cvs, 0.3, 0.02, multiplier=np.ascontiguousarray(attrs[:, 2]))
    same(got, want, "multiplier column 2")


def test_errors_write_nothing(torch, orc):
    rng = np.random.default_rng(4)
    n, cvs = 300, 8
    pts = random_strands(rng, n, cvs)
    # a non-finite amount (the reference throws): nothing written, kept 0
    assert orc.classic_cut(pts, cvs, 0.0, float("inf")) is None
    out = ClassicCutOutputs(n, cvs, roots=False, root_uvs=False)
    out.points.fill_(7.0)
    classic_cut(dev(torch, pts), cvs, 0.0, float("inf"), outputs=out)
    assert out.kept() == 0 and out.status() == abi.CLASSIC_NONFINITE_AMOUNT
    assert bool((out.points == 7.0).all())
    # a non-finite width on one strand
    bad = pts.copy()
    bad[5 * cvs, 3] = np.nan
    assert orc.classic_cut(bad, cvs, 0.25, 0.0) is None
    classic_cut(dev(torch, bad), cvs, 0.25, 0.0, outputs=out)
    assert out.kept() == 0 and out.status() == abi.CLASSIC_NONFINITE_WIDTH
    # the same outputs and workspace serve a good call afterwards
    classic_cut(dev(torch, pts), cvs, 0.25, 0.0, outputs=out)
    assert out.status() == 0
    same(out.to_host(), orc.classic_cut(pts, cvs, 0.25, 0.0), "after errors")


def test_empty_and_all_culled(torch):
    import torch as T
    out = ClassicCutOutputs(0, 4)
    classic_cut(T.zeros((0, 4), dtype=T.float32, device="cuda"), 4, 0.25, 0.0, outputs=out)
    assert out.kept() == 0 and out.status() == 0
    pts = np.zeros((100 * 4, 4), np.float32)  # zero-length strands
    got = run(torch, pts, 4, 0.25, 0.0)
    assert got["strand_count"] == 0 and got["status"] == 0


def test_generated_plane_chain(torch, orc):
    """The generator's own output (16 CVs, two octaves) through the cut,
    against the oracle's cut of the same curves."""
    w = workloads.config2(strands=50000, quads=64, guides_per_side=16)
    blob = nx.build_asset(w.build)
    asset = nx.DeviceAsset(blob)
    gen = nx.PackedOutputs(w.params.strand_count, 16)
    nx.generate_packed(asset, w.params, gen, octaves=w.octaves)
    out = classic_cut(gen.points, 16, 0.25, 0.0, point_counts=gen.point_counts, roots=gen.roots,
                      root_uvs=gen.root_uvs)
    got = out.to_host()
    host = gen.to_host()
    want = orc.classic_cut(host["points"], 16, 0.25, 0.0, counts=host["point_counts"], roots=host["roots"],
                           uvs=host["root_uvs"])
    assert 0 < want["strand_count"] <= w.params.strand_count
    same(got, want, "generated plane, 0.25 cLength")


def test_endpoints_reference_fixture(torch, classic):
    cvs = int(classic["walk_cvs"])
    got = add_endpoints(dev(torch, classic["walk_points"]), cvs).cpu().numpy()
    assert np.array_equal(bits(got), bits(classic["walk_endpoints"]))


@pytest.mark.parametrize("cvs", [2, 3, 16, 64, 256])
def test_endpoints_oracle(torch, orc, cvs):
    rng = np.random.default_rng(cvs + 100)
    n = 3000 // cvs + 65
    pts = rng.normal(size=(n * cvs, 4)).astype(np.float32)
    counts = torch.zeros(n, dtype=torch.int32, device="cuda")
    got = add_endpoints(dev(torch, pts), cvs, point_counts=counts).cpu().numpy()
    assert np.array_equal(bits(got), bits(orc.classic_endpoints(pts, cvs)))
    assert bool((counts == cvs + 2).all())


def test_cut_then_endpoints_on_device(torch, orc, classic):
    """The kept count stays on the device: endpoints read it there."""
    cvs = int(classic["walk_cvs"])
    pts = classic["walk_points"]
    n = len(pts) // cvs
    out = classic_cut(dev(torch, pts), cvs, 0.4, 0.1, radius_scale=2.0)
    sentinel = np.float32(-7.0)
    ends = torch.full((n * (cvs + 2), 4), float(sentinel), dtype=torch.float32, device="cuda")
    add_endpoints(out.points, cvs, strands=n, device_strand_count=out.words, out=ends)
    k = out.kept()
    want = classic["case3_endpoints"]
    assert k * (cvs + 2) == len(want)
    got = ends.cpu().numpy()
    assert np.array_equal(bits(got[:len(want)]), bits(want))
    assert np.all(got[len(want):] == sentinel)  # nothing past the kept strands
