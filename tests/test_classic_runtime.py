"""Pins of the Classic runtime with primitive expressions: the oracle
(oracle_classic.c, orc_classic_runtime) against the reference's own test
values (tests/classic_test.cpp:198-336) and against fixtures written by the
reference's parser, plan compiler and runtime
(tests/golden/reference_classic_runtime.npz); the plan packer's and the
request validator's contracts (nxg_classic_plan_pack,
nxg_validate_classic_runtime).  CPU only."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import nanoxgen_amd as nx
from nanoxgen_amd import abi, classic

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from classic_runtime_cases import CASES, guide_args, plan_sources, random_effect_cases  # noqa: E402

FIXTURE = os.path.join(HERE, "golden", "reference_classic_runtime.npz")

# status bit -> the reference's message where it throws
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
def fixture():
    with np.load(FIXTURE, allow_pickle=False) as d:
        return {k: d[k] for k in d.keys()}


def make_plan(case):
    sources, ptex, _ = plan_sources(case)
    return classic.ClassicPlan(**sources, description_id=case["description_id"], description_name=case["name"],
                               ptex_count=ptex)


def case_args(fx, i):
    g = lambda k: fx.get(f"case{i}_in_{k}")
    return dict(points=g("points"), cvs=int(g("cvs")), roots=g("roots"), uvs=g("uvs"), ids=g("ids"),
                prefixes=g("prefixes"), columns=g("columns"), radius_scale=float(g("radius_scale")),
                guides=guide_args({k[len(f"case{i}_in_"):]: v for k, v in fx.items() if k.startswith(f"case{i}_in_")}))


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def test_reference_float_runtime_plan_pin(orc):
    # test_float_runtime_plan: length 2, width 0.2, taper 0.5 from 0.5,
    # widthRamp rampUI(0,1,1:1,0.5,1), CutFX 0.25 * $cLength on (0,0,0)..(2,0,0)
    plan = make_plan(CASES[0])
    pts = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [2, 0, 0, 0]], np.float32)
    r = orc.classic_runtime(pts, 3, plan, roots=np.zeros(1, abi.ROOT_DTYPE))
    assert r["strand_count"] == 1
    assert np.allclose(r["points"][:, 0], [0.0, 1.5, 3.0], atol=1e-6)
    assert np.allclose(r["points"][:, 3], [0.1, 0.075, 0.025], atol=1e-6)


def test_reference_id_width_pin(orc):
    # the same test: width "$id" with primitive id 7 -> radius 3.5 at every CV
    plan = classic.ClassicPlan(width="$id")
    pts = np.array([[0, 0, 0, 0], [0, 1, 0, 0]], np.float32)
    r = orc.classic_runtime(pts, 2, plan, roots=np.zeros(1, abi.ROOT_DTYPE), primitive_ids=[7])
    assert r["strand_count"] == 1 and np.all(r["points"][:, 3] == 3.5)


def test_reference_culling_pins(orc):
    # test_float_runtime_fallbacks_and_validation: width -1 culls the strand;
    # width 1e-4 is kept and 0.000099 culled (radius 0.00005)
    pts = np.array([[0, 0, 0, 0.1], [0, 1, 0, 0.1]], np.float32)
    r = orc.classic_runtime(pts, 2, classic.ClassicPlan(width="-1"), roots=np.zeros(1, abi.ROOT_DTYPE))
    assert r["strand_count"] == 0
    two = np.array([[0, 0, 0, 0.1], [0, 1, 0, 0.1], [1, 0, 0, 0.1], [1, 1, 0, 0.1]], np.float32)
    r = orc.classic_runtime(two, 2, classic.ClassicPlan(width="$id == 1 ? 0.0001 : 0.000099"),
                            roots=np.zeros(2, abi.ROOT_DTYPE), primitive_ids=[1, 2])
    assert r["strand_count"] == 1 and np.all(r["points"][:, 3] == np.float32(0.00005))


@pytest.mark.parametrize("i", range(len(CASES)), ids=[c["name"] for c in CASES])
def test_oracle_matches_reference_fixture(orc, fixture, i):
    """Kept strands, counts, roots and every position and radius bit-exact
    with the reference's runtime; where it throws, the same message."""
    a = case_args(fixture, i)
    r = orc.classic_runtime(a["points"], a["cvs"], make_plan(CASES[i]), roots=a["roots"], uvs=a["uvs"],
                            primitive_ids=a["ids"], random_prefixes=a["prefixes"], columns=a["columns"],
                            radius_scale=a["radius_scale"], **a["guides"])
    if f"case{i}_error" in fixture:
        assert r["strand_count"] is None
        assert {MESSAGES[b] for b in MESSAGES if r["error_bits"] & b} == {str(fixture[f"case{i}_error"])}
        return
    assert r["strand_count"] == len(fixture[f"case{i}_out_point_counts"])
    assert np.array_equal(bits(r["points"]), bits(fixture[f"case{i}_out_points"]))
    assert np.array_equal(r["point_counts"], fixture[f"case{i}_out_point_counts"])
    assert np.array_equal(bits(r["roots"]), bits(fixture[f"case{i}_out_roots"]))


def test_oracle_matches_reference_core(orc):
    """Beyond the fixture, where the reference core is built here: seeded
    random walks and expressions that exercise every input kind."""
    if orc.ref() is None:
        pytest.skip("reference core not built here")
    from classic_runtime_cases import xgen_text
    rng = np.random.default_rng(5)
    n, cvs = 200, 9
    steps = rng.normal(0, 0.2, size=(n, cvs, 3)).astype(np.float32)
    pts = np.concatenate([np.cumsum(steps, 1), rng.uniform(0.001, 0.02, (n, cvs, 1)).astype(np.float32)], 2)
    pts = pts.reshape(-1, 4).astype(np.float32)
    roots = np.zeros(n, abi.ROOT_DTYPE)
    roots["uv"] = rng.uniform(0, 1, (n, 2))
    for exprs in [dict(length="1 + 0.5 * map('x')", cut_amount="$cLength * map('y') * 0.8",
                       width_ramp="$cWidth * 10 + $cLength"),
                  dict(width="0.02 + 0.01 * noise($Prefg*2.5)", taper="map('z')", taper_start="0.1",
                       cut_amount="0.05"),
                  dict(length="hash($id, 3) + 0.25", taper="0.9", width_ramp="rampUI(0,1,2:1,0,2)")]:
        case = dict(name="random", description_id=4, exprs=exprs)
        sources, ptex, freqs = plan_sources(case)
        cols = rng.uniform(0.1, 1.0, (n, ptex + len(freqs))).astype(np.float32) if ptex + len(freqs) else None
        ids = rng.permutation(n).astype(np.uint32)
        want = orc.ref_classic_runtime(xgen_text(case), pts, cvs, roots=roots, primitive_ids=ids, columns=cols)
        plan = classic.ClassicPlan(**sources, description_id=4, description_name="random", ptex_count=ptex)
        got = orc.classic_runtime(pts, cvs, plan, roots=roots, primitive_ids=ids, columns=cols)
        assert got["strand_count"] == want["strand_count"], exprs
        assert np.array_equal(bits(got["points"]), bits(want["points"])), exprs


def test_effect_lists_match_reference_core(orc):
    """Random effect orders (ClumpingFX, NoiseFX and CutFX modules, up to two
    of each, any order) through the reference's parser, plan compiler and
    runtime, on seeded walks with root-relative evaluation and local guide
    axes in half of them: the oracle bit-exact."""
    if orc.ref() is None:
        pytest.skip("reference core not built here")
    from classic_runtime_cases import case_clumps, clump_guides, clump_names, xgen_text
    rng = np.random.default_rng(77)
    n, cvs = 160, 8
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
    for k, case in enumerate(random_effect_cases(rng, 12)):
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
        kw = dict(roots=roots, primitive_ids=ids, random_prefixes=prefixes, columns=cols, clumps=clumps,
                  surface_tangents=tangents, root_relative=relative)
        want = orc.ref_classic_runtime(xgen_text(case), p, cvs, clump_names=clump_names(case), **kw)
        plan = classic.ClassicPlan(**sources, description_id=case["description_id"], description_name=case["name"],
                                   ptex_count=ptex)
        got = orc.classic_runtime(p, cvs, plan, **kw)
        label = [e for e, _ in case["effects"]]
        assert want["strand_count"] is not None, (label, want.get("error"))
        assert got["strand_count"] == want["strand_count"], label
        assert np.array_equal(bits(got["points"]), bits(want["points"])), label
        assert np.array_equal(bits(got["roots"]), bits(want["roots"])), label


def test_effect_list_plan_report():
    """The packer expands the short form into the same effect list, and
    reports the trailing cut and whether the CVs are rewritten."""
    one = dict(mask="1", clump="0.5", clumpScale="0", noise="0", noiseScale="1", noiseFrequency="1",
               noiseCorrelation="0")
    noise = dict(mask="1", magnitude="0.1", magnitudeScale="1", frequency="1", correlation="0", preserveLength="0")
    short = classic.ClassicPlan(clump=[one, one], clumps_after_cut=1, cut_amount="0.1")
    listed = classic.ClassicPlan(effects=[("clump", one), ("cut", "0.1"), ("clump", one)])
    for plan in (short, listed):
        info = plan.info
        assert [(e.type, e.module) for e in info.effects[:info.effect_count]] == [(0, 0), (2, 0), (0, 1)]
        assert info.trailing_cut == 0 and info.rewrites == 1 and info.cut_count == 1
    only_cut = classic.ClassicPlan(cut_amount="0.1").info
    assert only_cut.trailing_cut == 1 and only_cut.rewrites == 0 and only_cut.effect_count == 1
    two = classic.ClassicPlan(effects=[("noise", noise), ("cut", "0.1"), ("noise", noise), ("cut", "0.2")]).info
    assert (two.noise_count, two.cut_count, two.trailing_cut, two.rewrites) == (2, 2, 1, 1)
    # module 1's programs sit in their own slots
    for key in ("noise_mask@1", "noise_preserve_length@1", "cut_amount@1"):
        assert (two.present >> abi.CLASSIC_EXPRS.index(key)) & 1
    # the programs that read $cLength
    reads = classic.ClassicPlan(length="1.5", taper="$cLength * 0.1", cut_amount="$cLength * 0.3",
                                clump=dict(one, clump="clamp($cLength, 0, 1)")).info
    readers = {k for k in abi.CLASSIC_EXPRS if (reads.length_readers >> abi.CLASSIC_EXPRS.index(k)) & 1}
    assert readers == {"taper", "cut_amount", "clump"}, readers


def test_effect_list_pack_rejects():
    L = nx.lib()
    info = abi.ClassicPlanInfo()
    amount = nx.expr.compile_expression("0.1", "amount", "CutFXModule")

    def pack(**kw):
        effects = kw.pop("effects", [])
        plan = abi.ClassicPlan(effect_count=len(effects), **kw)
        for e, (t, m) in enumerate(effects):
            plan.effects[e] = abi.ClassicEffect(t, m)
        plan.expressions[abi.CLASSIC_EXPRS.index("cut_amount")] = C.addressof(amount.struct)
        plan.expressions[abi.CLASSIC_EXPRS.index("cut_amount@1")] = C.addressof(amount.struct)
        return L.nxg_classic_plan_pack(C.byref(plan), None, 0, C.byref(info))

    cut = abi.CLASSIC_EFFECT_CUT
    assert pack(cut_count=2, effects=[(cut, 1), (cut, 0)]) == 0
    assert pack(cut_count=2, effects=[(cut, 0), (cut, 0)]) == abi.INVALID_PROGRAM  # a module twice
    assert pack(cut_count=2, effects=[(cut, 0)]) == abi.INVALID_PROGRAM            # a module missing
    assert pack(cut_count=1, effects=[(cut, 1)]) == abi.INVALID_PROGRAM            # past the modules
    assert pack(cut_count=1, effects=[(3, 0)]) == abi.INVALID_PROGRAM              # an unknown type
    assert pack(cut_count=3, effects=[(cut, 0)]) == abi.INVALID_PROGRAM            # too many modules
    assert pack(cut_count=1, has_cut=1, effects=[(cut, 0)]) == abi.INVALID_PROGRAM  # the forms mixed
    assert pack(noise_count=1, has_cut=1) == abi.INVALID_PROGRAM                     # short form with counts
    # a NoiseFX needs all six programs
    assert pack(noise_count=1, effects=[(abi.CLASSIC_EFFECT_NOISE, 0)]) == abi.INVALID_PROGRAM


# ---- host contracts --------------------------------------------------------

def test_plan_report():
    plan = make_plan(CASES[1])
    info = plan.info
    assert info.present == 0b111111 and info.has_cut == 1
    assert info.columns_required == 3 and info.needs_random_prefix == 1 and info.final_length == 0
    assert make_plan(CASES[2]).info.final_length == 1
    # programs of t alone (no input, no rand) are evaluated once per call
    assert info.uniform == 1 << 5
    assert make_plan(CASES[0]).info.uniform == 0b111011
    assert plan.name_seed == nx.lib().nxg_string_seed(b"walkAll")


def test_classic_binding():
    assert classic.classic_binding("cWidth") == abi.CLASSIC_VAR_CWIDTH
    assert classic.classic_binding("__nxg_map_2") == abi.CLASSIC_VAR_COLUMN0 + 2
    assert classic.classic_binding("__nxg_custom_0", ptex_count=3) == abi.CLASSIC_VAR_COLUMN0 + 3
    assert classic.classic_binding("__nxg_pref_noise_1", ptex_count=3, custom_count=2) == abi.CLASSIC_VAR_COLUMN0 + 6
    with pytest.raises(KeyError, match=r"not bound: \$faceid"):
        classic.classic_binding("faceid")


def test_plan_pack_rejects():
    L = nx.lib()
    info = abi.ClassicPlanInfo()
    # extension ops are not part of a Classic plan
    prog = nx.expr.compile_expression("pow($cLength, 2)", "length", "SplinePrimitive", extensions=True)
    prog.bind_indices([abi.CLASSIC_VAR_CLENGTH])
    plan = abi.ClassicPlan()
    plan.expressions[0] = C.addressof(prog.struct)
    assert L.nxg_classic_plan_pack(C.byref(plan), None, 0, C.byref(info)) == abi.INVALID_PROGRAM
    # a binding past the input row
    ok = nx.expr.compile_expression("$a + 1", "length", "SplinePrimitive").bind_indices(
        [abi.CLASSIC_VAR_COLUMN0 + abi.CLASSIC_MAX_COLUMNS])
    plan.expressions[0] = C.addressof(ok.struct)
    assert L.nxg_classic_plan_pack(C.byref(plan), None, 0, C.byref(info)) == abi.INVALID_PROGRAM
    # a CutFX without its amount
    plan = abi.ClassicPlan(has_cut=1)
    assert L.nxg_classic_plan_pack(C.byref(plan), None, 0, C.byref(info)) == abi.INVALID_PROGRAM
    assert L.nxg_classic_plan_pack(None, None, 0, C.byref(info)) == abi.STATUS.index("NULL_ARGUMENT")


def test_validate_classic_runtime():
    L = nx.lib()
    plan = make_plan(CASES[1])  # reads 3 columns and calls rand()
    n, cvs = 10, 7
    fake = 1 << 20  # never dereferenced by the host checks
    curves = abi.ClassicCurves(points=fake, roots=fake + 4096, strand_count=n, cvs_per_strand=cvs)
    out = abi.ClassicOutput(points=fake + (1 << 16), point_capacity=n * cvs, kept_count=fake + (1 << 18))
    good = dict(plan_image=fake + (1 << 19), plan=C.addressof(plan.info), random_prefixes=fake + 8192,
                columns=fake + 12288, column_count=3, radius_scale=1.0)

    def status(**kw):
        d = dict(good)
        d.update(kw)
        return L.nxg_validate_classic_runtime(C.byref(curves), C.byref(abi.ClassicRuntimeInputs(**d)), C.byref(out))

    assert status() == 0
    assert status(random_prefixes=None) == abi.STATUS.index("INVALID_CLASSIC_REQUEST")
    assert status(column_count=2) == abi.STATUS.index("INVALID_CLASSIC_REQUEST")
    assert status(columns=None) == abi.STATUS.index("INVALID_CLASSIC_REQUEST")
    assert status(radius_scale=-1.0) == abi.STATUS.index("INVALID_RADIUS_SCALE")
    assert status(radius_scale=float("inf")) == abi.STATUS.index("INVALID_RADIUS_SCALE")
    assert status(plan_image=fake + (1 << 19) + 4) == abi.STATUS.index("INVALID_CLASSIC_REQUEST")
    curves.roots = None
    assert status() == abi.STATUS.index("INVALID_CLASSIC_REQUEST")


NOISE = dict(mask="1", magnitude="0.1", magnitudeScale="1", frequency="2", correlation="0", preserveLength="0")
CLUMP = dict(mask="1", clump="0.5", clumpScale="0", noise="0", noiseScale="1", noiseFrequency="1",
             noiseCorrelation="0")


def test_effect_plan_rules():
    """Plan packing: a NoiseFX needs its six expressions and sits after the
    clumps and before the CutFX; the CutFX can precede only as many modules
    as there are, and only when it exists."""
    plan = classic.ClassicPlan(noise=NOISE, clump=[CLUMP, CLUMP], cut_amount="0.1")
    assert plan.info.has_noise == 1 and plan.info.clump_count == 2
    assert (plan.info.present >> 27) & 0x3F == 0x3F  # the six NoiseFX programs
    assert plan.info.uniform >> 27 == 0  # NoiseFX programs run per strand
    with pytest.raises(ValueError):
        classic.ClassicPlan(noise={k: v for k, v in NOISE.items() if k != "preserveLength"})
    for kw in (dict(clump=[CLUMP], cut_amount="0.1", clumps_after_cut=2),  # more than the modules
               dict(clump=[CLUMP], clumps_after_cut=1),  # no CutFX to precede them
               dict(clump=[CLUMP], cut_amount="0.1", clumps_after_cut=1, noise=NOISE)):  # NoiseFX after the cut
        with pytest.raises(nx.NxgError):
            classic.ClassicPlan(**kw)
    assert classic.ClassicPlan(clump=[CLUMP, CLUMP, CLUMP], cut_amount="0.1", clumps_after_cut=3).info.clumps_after_cut == 3


def test_noise_validation():
    """A NoiseFX needs one surface tangent per strand and at most 64 CVs."""
    L = nx.lib()
    plan = classic.ClassicPlan(noise=NOISE)
    n = 10
    fake = 1 << 20  # never dereferenced by the host checks
    out = abi.ClassicOutput(points=fake + (1 << 16), point_capacity=n * 80, kept_count=fake + (1 << 18))
    good = dict(plan_image=fake + (1 << 19), plan=C.addressof(plan.info), radius_scale=1.0,
                surface_tangents=fake + 8192)

    def status(cvs=7, **kw):
        curves = abi.ClassicCurves(points=fake, roots=fake + 4096, strand_count=n, cvs_per_strand=cvs)
        d = dict(good)
        d.update(kw)
        return L.nxg_validate_classic_runtime(C.byref(curves), C.byref(abi.ClassicRuntimeInputs(**d)), C.byref(out))

    assert status() == 0
    assert status(surface_tangents=None) == abi.STATUS.index("INVALID_CLASSIC_REQUEST")
    assert status(cvs=65) == abi.STATUS.index("INVALID_CLASSIC_REQUEST")
    assert status(noise_domain_positions=fake + 12288) == 0
