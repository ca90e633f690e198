"""The Classic runtime with primitive expressions on the GPU
(nxg_classic_runtime_apply) against the reference's parser, plan compiler and
runtime (tests/golden/reference_classic_runtime.npz) and the oracle
(orc_classic_runtime, pinned to that fixture), bit for bit in kept strands,
point counts, roots and every position and radius; where the reference
throws, nothing is written and the status bits name its message."""
import os
import sys

import numpy as np
import pytest

import nanoxgen_amd as nx
from nanoxgen_amd import abi, classic, workloads

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from classic_runtime_cases import (CASES, case_clumps, clump_guides, clump_names, guide_args, has_noise,  # noqa: E402
                                   plan_sources, random_effect_cases)

pytestmark = pytest.mark.gpu

MESSAGES = {
    abi.CLASSIC_NONFINITE_AMOUNT: "Classic CutFX amount produced a non-finite value",
    abi.CLASSIC_NONFINITE_WIDTH: "Classic width expression produced a non-finite value",
    abi.CLASSIC_NEGATIVE_LENGTH: "Classic length expression produced a negative value",
    abi.CLASSIC_NONFINITE_TAPER: "Classic taper expression produced a non-finite value",
    abi.CLASSIC_NONFINITE_PROFILE: "Classic width profile produced a non-finite value",
    abi.CLASSIC_EXPRESSION_ERROR: "XGen float expression produced a non-finite value",
    abi.CLASSIC_VARIABLE_COUNTS: "Classic runtime requires fixed point counts",
    abi.CLASSIC_INVALID_GUIDE: "Classic ClumpingFX guide index is invalid",
    abi.CLASSIC_NONFINITE_CLUMP: "Classic ClumpingFX noise magnitude is non-finite",
    abi.CLASSIC_NONFINITE_NOISE_PARAMETER: "Classic NoiseFX produced a non-finite parameter",
    abi.CLASSIC_NONFINITE_NOISE: "Classic NoiseFX magnitude is non-finite",
}


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def fixture():
    with np.load(os.path.join(HERE, "golden", "reference_classic_runtime.npz"), allow_pickle=False) as d:
        return {k: d[k] for k in d.keys()}


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def dev(torch, a, view=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a)
    return torch.from_numpy(a if view is None else a.view(view)).cuda()


def make_plan(case):
    sources, ptex, _ = plan_sources(case)
    return classic.ClassicPlan(**sources, description_id=case["description_id"], description_name=case["name"],
                               ptex_count=ptex)


def run_gpu(torch, plan, pts, cvs, *, roots, uvs=None, counts=None, ids=None, prefixes=None, columns=None,
            radius_scale=1.0, strand_guides=None, guides=None, guide_uvs=None, guide_random_prefixes=None,
            clumps=None, surface_tangents=None, noise_domain_positions=None, root_relative=False):
    if clumps is None and guides is not None:
        clumps = [dict(strand_guides=strand_guides, guides=guides, guide_uvs=guide_uvs,
                       guide_random_prefixes=guide_random_prefixes)]
    bindings = [dict(strand_guides=dev(torch, b["strand_guides"], np.int32), guide_uvs=dev(torch, b["guide_uvs"]),
                     guide_random_prefixes=dev(torch, b["guide_random_prefixes"], np.int32),
                     local_axes=dev(torch, b.get("local_axes")),
                     guides={k: dev(torch, v) for k, v in b["guides"].items() if v is not None})
                for b in clumps or []]
    out = classic.classic_runtime(dev(torch, np.asarray(pts, np.float32).reshape(-1, 4)), cvs, plan,
                                  roots=dev(torch, np.ascontiguousarray(roots).view(np.int32).reshape(-1, 12)),
                                  root_uvs=dev(torch, uvs), point_counts=dev(torch, counts, np.int32),
                                  primitive_ids=dev(torch, ids, np.int32), random_prefixes=dev(torch, prefixes, np.int32),
                                  columns=dev(torch, columns), radius_scale=radius_scale, clumps=bindings,
                                  surface_tangents=dev(torch, surface_tangents),
                                  noise_domain_positions=dev(torch, noise_domain_positions),
                                  root_relative=root_relative)
    return out.to_host()


def same(got, want, label, noisy_clump=False):
    """Kept strands, counts and roots exact; positions and radii bit-exact,
    or, where a clump's noise axis is on (its frame uses the device's acosf,
    cosf and sinf), within 1e-5 of the strands' bounding-box diagonal."""
    k = want["strand_count"]
    assert got["status"] == 0, f"{label}: status {got['status']}"
    assert got["strand_count"] == k, f"{label}: kept {got['strand_count']} != {k}"
    if noisy_clump:
        g, w = got["points"], want["points"]
        diag = float(np.linalg.norm(w[:, :3].max(0) - w[:, :3].min(0))) if k else 1.0
        err = float(np.abs(g[:, :3] - w[:, :3]).max()) if k else 0.0
        print(f"{label}: max |gpu - reference| = {err:.3g} ({err / diag:.3g} of the bbox diagonal)", flush=True)
        assert err <= 1e-5 * diag, f"{label}: positions"
        assert np.allclose(g[:, 3], w[:, 3], rtol=1e-5, atol=0.0), f"{label}: radii"
    else:
        assert np.array_equal(bits(got["points"]), bits(want["points"])), f"{label}: points"
    assert np.array_equal(got["point_counts"], want["point_counts"]), f"{label}: counts"
    assert np.array_equal(bits(got["roots"]), bits(want["roots"])), f"{label}: roots"


@pytest.mark.parametrize("i", range(len(CASES)), ids=[c["name"] for c in CASES])
def test_reference_fixture(torch, fixture, i):
    g = lambda k: fixture.get(f"case{i}_in_{k}")
    inputs = {k[len(f"case{i}_in_"):]: v for k, v in fixture.items() if k.startswith(f"case{i}_in_")}
    got = run_gpu(torch, make_plan(CASES[i]), g("points"), int(g("cvs")), roots=g("roots"), uvs=g("uvs"),
                  ids=g("ids"), prefixes=g("prefixes"), columns=g("columns"), radius_scale=float(g("radius_scale")),
                  **guide_args(inputs))
    if f"case{i}_error" in fixture:
        assert got["strand_count"] == 0 and got["status"] != 0
        assert {MESSAGES[b] for b in MESSAGES if got["status"] & b} == {str(fixture[f"case{i}_error"])}
        return
    want = dict(strand_count=len(fixture[f"case{i}_out_point_counts"]), points=fixture[f"case{i}_out_points"],
                point_counts=fixture[f"case{i}_out_point_counts"], roots=fixture[f"case{i}_out_roots"])
    same(got, want, CASES[i]["name"], noisy_clump=has_noise(CASES[i]))


PLANS = [
    dict(length="1 + 0.5 * $__nxg_map_0", width="$cWidth + 0.004 + 0.01 * $__nxg_map_1", taper="0.6",
         taper_start="0.25", width_ramp="rampUI(0,1,1:0.4,0.7,2:1,0.2,3)",
         cut_amount="$cLength * (0.2 + 0.5 * rand(3)) + $cWidth"),
    dict(width_ramp="$cLength * 0.1 + 0.5", cut_amount="0.3 * $cLength"),
    dict(length="fit($__nxg_map_0, 0, 1, 0.5, 1.5)", taper="0.9"),
]


@pytest.mark.parametrize("cvs", [2, 5, 16, 33, 64, 80])
def test_random_walks_oracle(torch, orc, cvs):
    """Seeded walks at every staging form (staged tiles up to 64 CVs, memory
    reads above), ragged final tiles, every plan shape; GPU == oracle."""
    rng = np.random.default_rng(cvs)
    n = 64 * 3 + 17
    steps = rng.normal(0, 0.3, size=(n, cvs, 3)).astype(np.float32)
    pts = np.concatenate([np.cumsum(steps, 1), rng.uniform(0.0, 0.02, (n, cvs, 1)).astype(np.float32)], 2)
    pts = pts.reshape(-1, 4).astype(np.float32)
    pts[::7 * cvs, 0] = -0.0  # scale_curve turns a -0 root coordinate into +0
    roots = np.zeros(n, abi.ROOT_DTYPE)
    roots["uv"] = rng.uniform(0, 1, (n, 2))
    roots["triangle_index"] = rng.integers(0, 1000, n)
    roots["surface_face_id"] = np.where(rng.uniform(size=n) < 0.5, abi.INVALID_INDEX, rng.integers(0, 99, n))
    columns = rng.uniform(0.0, 1.0, (n, 2)).astype(np.float32)
    ids = rng.permutation(n).astype(np.uint32)
    prefixes = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    for k, sources in enumerate(PLANS):
        plan = classic.ClassicPlan(**sources, description_id=11, description_name="walks", ptex_count=2)
        want = orc.classic_runtime(pts, cvs, plan, roots=roots, primitive_ids=ids, random_prefixes=prefixes,
                                   columns=columns)
        assert want["strand_count"] is not None
        got = run_gpu(torch, plan, pts, cvs, roots=roots, ids=ids, prefixes=prefixes, columns=columns)
        same(got, want, f"{cvs} CVs, plan {k}")


CLUMP_PLANS = [
    dict(cut_amount="$cLength * 0.3 + rampUI(0,0,1:1,0.05,1)", taper="rampUI(0,0.1,1:1,0.6,1)",
         clump=dict(mask="0.2 + 0.8 * $__nxg_map_0", clump="0.5 + 0.4 * rand()", clumpScale="0.2 * $__nxg_map_1",
                    noise="0.3 * rand(2)", noiseScale="rampUI(0,0,1:1,1,1)", noiseFrequency="1 + 2 * rand()",
                    noiseCorrelation="10 + 50 * rand(9)")),
    dict(width_ramp="$cLength * 0.2",
         clump=dict(mask="$id < 60 ? 0 : 1", clump="clamp($cLength * 0.5, 0, 1)", clumpScale="-0.1",
                    noise="0.5", noiseScale="$cWidth * 20", noiseFrequency="3", noiseCorrelation="0")),
    # nothing before the clump's re-measure reads cLength, so its first measure
    # waits: strands without a guide or with mask 0 measure it afterwards
    # (scaled by the length expression), degenerate ones at once
    dict(length="0.5 + $__nxg_map_0", cut_amount="$cLength * 0.25 + 0.01",
         clump=dict(mask="$id < 60 ? 0 : 1", clump="0.6 + 0.2 * rand()", clumpScale="0.1", noise="0.2",
                    noiseScale="1", noiseFrequency="2", noiseCorrelation="0")),
]


@pytest.mark.parametrize("cvs", [2, 5, 16, 33, 64])
def test_clump_walks_oracle(torch, orc, cvs):
    """ClumpingFX (and a CutFX after it) on seeded walks at every staged CV
    count, guides with prefixes and uvs, strands without a guide, ragged
    final tiles; GPU == oracle."""
    rng = np.random.default_rng(100 + cvs)
    n = 64 * 3 + 29
    steps = rng.normal(0, 0.3, size=(n, cvs, 3)).astype(np.float32)
    pts = np.concatenate([np.cumsum(steps, 1), rng.uniform(0.001, 0.02, (n, cvs, 1)).astype(np.float32)], 2)
    pts[7:9, :, :3] = pts[7:9, :1, :3]  # two degenerate strands: every CV at the root
    pts[9, :, :3] = pts[9, :1, :3] + np.linspace(0.0, 2e-5, cvs, dtype=np.float32)[:, None]  # shorter than 1e-4
    pts = pts.reshape(-1, 4).astype(np.float32)
    pts[::5 * cvs, 2] = -0.0
    roots = np.zeros(n, abi.ROOT_DTYPE)
    roots["uv"] = rng.uniform(0, 1, (n, 2))
    columns = rng.uniform(0.0, 1.0, (n, 2)).astype(np.float32)
    ids = rng.permutation(n).astype(np.uint32)
    prefixes = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    g = clump_guides(11, cvs, rng)
    strand_guides = rng.integers(0, 11, n).astype(np.uint32)
    strand_guides[::6] = 0xFFFFFFFF
    kw = dict(strand_guides=strand_guides, guide_uvs=g["uvs"], guide_random_prefixes=g["prefixes"],
              guides=dict(axes=g["axes"], normals=g["normals"], tangents=g["tangents"], reference=g["reference"]))
    for k, sources in enumerate(CLUMP_PLANS):
        plan = classic.ClassicPlan(**sources, description_id=5, description_name="clumps", ptex_count=2)
        want = orc.classic_runtime(pts, cvs, plan, roots=roots, primitive_ids=ids, random_prefixes=prefixes,
                                   columns=columns, **kw)
        assert want["strand_count"] is not None, want["error_bits"]
        got = run_gpu(torch, plan, pts, cvs, roots=roots, ids=ids, prefixes=prefixes, columns=columns, **kw)
        same(got, want, f"{cvs} CVs, clump plan {k}", noisy_clump=True)


@pytest.mark.parametrize("cvs", [3, 16, 64])
def test_hierarchy_walks_oracle(torch, orc, cvs):
    """Hierarchical clumping: two and three ClumpingFX modules, each with its
    own guides (different counts, guideless strands, the second module's
    mask zero for some strands), and a CutFX reading t after them or
    between them; GPU == oracle."""
    rng = np.random.default_rng(300 + cvs)
    n = 64 * 2 + 37
    steps = rng.normal(0, 0.3, size=(n, cvs, 3)).astype(np.float32)
    pts = np.concatenate([np.cumsum(steps, 1), rng.uniform(0.001, 0.02, (n, cvs, 1)).astype(np.float32)], 2)
    pts = pts.reshape(-1, 4).astype(np.float32)
    roots = np.zeros(n, abi.ROOT_DTYPE)
    roots["uv"] = rng.uniform(0, 1, (n, 2))
    columns = rng.uniform(0.0, 1.0, (n, 2)).astype(np.float32)
    ids = rng.permutation(n).astype(np.uint32)
    prefixes = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    clumps = []
    for m, count in enumerate((7, 19, 4)):
        g = clump_guides(count, cvs, rng)
        sg = rng.integers(0, count, n).astype(np.uint32)
        sg[m::4 + m] = 0xFFFFFFFF
        clumps.append(dict(strand_guides=sg, guide_uvs=g["uvs"], guide_random_prefixes=g["prefixes"],
                           guides=dict(axes=g["axes"], normals=g["normals"], tangents=g["tangents"],
                                       reference=g["reference"] if m != 1 else None)))
    levels = [CLUMP_PLANS[0]["clump"], CLUMP_PLANS[1]["clump"],
              dict(mask="0.7", clump="0.5 * $__nxg_map_1", clumpScale="rampUI(0,0,1:1,0.2,1)", noise="0.1",
                   noiseScale="1", noiseFrequency="2", noiseCorrelation="20")]
    for count, after in ((2, 0), (3, 0), (2, 1), (3, 2)):
        plan = classic.ClassicPlan(clump=levels[:count], cut_amount="$cLength * 0.2 + rampUI(0,0,1:1,0.1,1)",
                                   taper="rampUI(0,0.3,1:1,0.6,1)", description_id=6, description_name="levels",
                                   ptex_count=2, clumps_after_cut=after)
        want = orc.classic_runtime(pts, cvs, plan, roots=roots, primitive_ids=ids, random_prefixes=prefixes,
                                   columns=columns, clumps=clumps[:count])
        assert want["strand_count"] is not None, want["error_bits"]
        got = run_gpu(torch, plan, pts, cvs, roots=roots, ids=ids, prefixes=prefixes, columns=columns,
                      clumps=clumps[:count])
        same(got, want, f"{cvs} CVs, {count} clump levels, {after} after the cut", noisy_clump=True)


@pytest.mark.parametrize("cvs", [2, 5, 16, 33, 64])
def test_noise_walks_oracle(torch, orc, cvs):
    """NoiseFX (after a clump, before a CutFX) on seeded walks: PTEX mask,
    rand() magnitude, per-CV magnitude scale, preserve length, zero surface
    tangents (the fallback frame), with and without domain positions;
    GPU == oracle within the noise-frame tolerance."""
    rng = np.random.default_rng(500 + cvs)
    n = 64 * 2 + 21
    steps = rng.normal(0, 0.3, size=(n, cvs, 3)).astype(np.float32)
    pts = np.concatenate([np.cumsum(steps, 1), rng.uniform(0.001, 0.02, (n, cvs, 1)).astype(np.float32)], 2)
    pts = pts.reshape(-1, 4).astype(np.float32)
    roots = np.zeros(n, abi.ROOT_DTYPE)
    roots["uv"] = rng.uniform(0, 1, (n, 2))
    roots["normal"] = rng.normal(0, 1, (n, 3))
    columns = rng.uniform(0.0, 1.0, (n, 2)).astype(np.float32)
    prefixes = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    tangents = rng.normal(0, 1, (n, 3)).astype(np.float32)
    tangents[::7] = 0.0
    domain = rng.uniform(-3, 3, (n, 3)).astype(np.float32)
    g = clump_guides(6, cvs, rng)
    sg = rng.integers(0, 6, n).astype(np.uint32)
    clumps = [dict(strand_guides=sg, guide_uvs=g["uvs"], guide_random_prefixes=g["prefixes"],
                   guides=dict(axes=g["axes"], normals=g["normals"], tangents=g["tangents"], reference=None))]
    noise = dict(mask="0.3 + 0.7 * $__nxg_map_0", magnitude="0.1 + 0.2 * rand()", magnitudeScale="rampUI(0,0.2,1:1,1,1)",
                 frequency="2 + $__nxg_map_1", correlation="30", preserveLength="40")
    for k, with_clump in enumerate((False, True)):
        plan = classic.ClassicPlan(noise=noise, cut_amount="$cLength * 0.15",
                                   clump=CLUMP_PLANS[1]["clump"] if with_clump else None,
                                   description_id=7, description_name="noise", ptex_count=2)
        kw = dict(clumps=clumps if with_clump else None, surface_tangents=tangents,
                  noise_domain_positions=domain if k else None)
        want = orc.classic_runtime(pts, cvs, plan, roots=roots, random_prefixes=prefixes, columns=columns, **kw)
        assert want["strand_count"] is not None, want["error_bits"]
        got = run_gpu(torch, plan, pts, cvs, roots=roots, prefixes=prefixes, columns=columns, **kw)
        same(got, want, f"{cvs} CVs, NoiseFX plan {k}", noisy_clump=True)


def test_generated_plane(torch, orc):
    """The generator's own curves (16 CVs, two octaves) through a full plan."""
    w = workloads.config2(strands=30000, quads=64, guides_per_side=16)
    blob = nx.build_asset(w.build)
    gen = nx.PackedOutputs(w.params.strand_count, 16)
    nx.generate_packed(nx.DeviceAsset(blob), w.params, gen, octaves=w.octaves)
    host = gen.to_host()
    plan = classic.ClassicPlan(**PLANS[0], description_id=2, description_name="plane", ptex_count=2)
    n = w.params.strand_count
    rng = np.random.default_rng(9)
    columns = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    prefixes = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    out = classic.classic_runtime(gen.points, 16, plan, roots=gen.roots, root_uvs=gen.root_uvs,
                                  point_counts=gen.point_counts, random_prefixes=dev(torch, prefixes, np.int32),
                                  columns=dev(torch, columns), radius_scale=1.0)
    got = out.to_host()
    want = orc.classic_runtime(host["points"], 16, plan, roots=host["roots"], counts=host["point_counts"],
                               uvs=host["root_uvs"], random_prefixes=prefixes, columns=columns)
    assert 0 < want["strand_count"] <= n
    same(got, want, "generated plane")
    assert np.array_equal(got["root_uvs"], want["root_uvs"])


def test_variable_counts_and_reuse(torch, orc):
    """Variable point counts raise the reference's error and write nothing;
    the same outputs and workspace then serve a valid call."""
    rng = np.random.default_rng(2)
    n, cvs = 100, 6
    pts = rng.normal(size=(n * cvs, 4)).astype(np.float32)
    pts[:, 3] = 0.01
    roots = np.zeros(n, abi.ROOT_DTYPE)
    plan = classic.ClassicPlan(taper="0.5")
    counts = np.full(n, cvs, np.uint32)
    counts[37] = cvs - 1
    T = torch
    outputs = classic.ClassicCutOutputs(n, cvs, root_uvs=False)
    ws = T.empty(classic.runtime_workspace_bytes(n), dtype=T.uint8, device="cuda")
    kw = dict(roots=dev(T, roots.view(np.int32).reshape(-1, 12)), outputs=outputs, workspace=ws)
    classic.classic_runtime(dev(T, pts), cvs, plan, point_counts=dev(T, counts, np.int32), **kw)
    assert outputs.kept() == 0 and outputs.status() == abi.CLASSIC_VARIABLE_COUNTS
    counts[37] = cvs
    classic.classic_runtime(dev(T, pts), cvs, plan, point_counts=dev(T, counts, np.int32), **kw)
    got = outputs.to_host()
    want = orc.classic_runtime(pts, cvs, plan, roots=roots, counts=counts)
    same(got, want, "after an error")


@pytest.mark.parametrize("cvs", [3, 8, 24])
def test_effect_list_walks_oracle(torch, orc, cvs):
    """Random effect orders (ClumpingFX, NoiseFX and CutFX modules, up to two
    of each, in any order) on seeded walks, root-relative with local guide
    axes in half of them; GPU == oracle (bit-exact without noise, within
    1e-5 of the diagonal with it)."""
    rng = np.random.default_rng(300 + cvs)
    n = 64 * 2 + 23
    steps = rng.normal(0, 0.25, size=(n, cvs, 3)).astype(np.float32)
    steps[:, 0] = rng.uniform(-3, 3, size=(n, 3))
    walk = np.cumsum(steps, 1)
    pts = np.concatenate([walk, rng.uniform(0.001, 0.02, (n, cvs, 1))], 2).astype(np.float32).reshape(-1, 4)
    roots = np.zeros(n, abi.ROOT_DTYPE)
    roots["uv"] = rng.uniform(0, 1, (n, 2))
    roots["position"] = walk[:, 0]
    tangents = rng.normal(0, 1, (n, 3)).astype(np.float32)
    ids = rng.permutation(n).astype(np.uint32)
    prefixes = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    for k, case in enumerate(random_effect_cases(rng, 8)):
        relative = k % 2 == 1
        sources, ptex, freqs = plan_sources(case)
        cols = rng.uniform(0.05, 1.0, (n, ptex + len(freqs))).astype(np.float32) if ptex + len(freqs) else None
        clumps = []
        for m in range(len(case_clumps(case))):
            g = clump_guides(9, cvs, rng)
            sg = rng.integers(0, 9, n).astype(np.uint32)
            sg[m::5] = 0xFFFFFFFF
            axes = g["axes"].reshape(9, cvs, 3)
            clumps.append(dict(strand_guides=sg, guide_uvs=g["uvs"], guide_random_prefixes=g["prefixes"],
                               local_axes=(axes - axes[:, :1]).reshape(-1, 3) if relative else None,
                               guides=dict(axes=g["axes"], normals=g["normals"], tangents=g["tangents"],
                                           reference=g["reference"])))
        p = pts.copy()
        if relative:
            p.reshape(n, cvs, 4)[:, :, :3] -= walk[:, :1].astype(np.float32)
        plan = classic.ClassicPlan(**sources, description_id=case["description_id"], description_name=case["name"],
                                   ptex_count=ptex)
        kw = dict(primitive_ids=ids, random_prefixes=prefixes, columns=cols, clumps=clumps,
                  surface_tangents=tangents, root_relative=relative)
        label = f"{cvs} CVs, {[e for e, _ in case['effects']]}, root_relative={relative}"
        print("case", label, flush=True)
        want = orc.classic_runtime(p, cvs, plan, roots=roots, **kw)
        assert want["strand_count"] is not None, (label, want["error_bits"])
        got = run_gpu(torch, plan, p, cvs, roots=roots, ids=ids, prefixes=prefixes, columns=cols, clumps=clumps,
                      surface_tangents=tangents, root_relative=relative)
        same(got, want, label, noisy_clump=has_noise(case))
