"""Cases of the Classic runtime with primitive expressions, shared by the
fixture writer (make_golden.py classic_runtime) and the tests.

Each case is a Classic description: SplinePrimitive attributes (length,
width, taper, taperStart, widthRamp), ClumpingFX modules (`clump`: one,
`clumps`: several, applied in order) and an optional CutFX amount, written
in the authoring form the reference's plan compiler reads (map('path') for
PTEX inputs, noise($Prefg*f) for reference-pose noise).  `xgen_text` renders
the description for the reference's parser; `plan_sources` rewrites the
same expressions into the bound form (map -> $__nxg_map_k in compile order,
noise($Prefg*f) -> $__nxg_pref_noise_j per distinct frequency), which is what
compile_xgen_classic_float_runtime_plan hands its compiler
(src/xgen_classic_runtime.cpp:138-430).  `case_inputs` builds the strands,
roots, ids, prefixes and input rows of a case from its seed.
"""
import re
import zlib

import numpy as np

# The order the reference compiles a plan's expressions in
# (src/xgen_classic_runtime.cpp:1211-1229, then the FX modules in object
# order: the ClumpingFX modules before the CutFX).
PRIMITIVE = (("length", "length"), ("width", "width"), ("taper", "taper"), ("taper_start", "taperStart"),
             ("width_ramp", "widthRamp"))
CLUMP = ("mask", "clump", "clumpScale", "noise", "noiseScale", "noiseFrequency", "noiseCorrelation")
NOISE = ("mask", "magnitude", "magnitudeScale", "frequency", "correlation", "preserveLength")


CASES = [
    # the reference's own test_float_runtime_plan (tests/classic_test.cpp:198-246)
    dict(name="floatRuntime", description_id=7, strands="line3",
         exprs=dict(length="2", width="0.2", taper="0.5", taper_start="0.5", width_ramp="rampUI(0,1,1:1,0.5,1)",
                    cut_amount="0.25*$cLength")),
    # every expression, PTEX inputs, rand() through SeExpr prefixes
    dict(name="ns:walkAll", description_id=12, strands="walk", ids=True, prefixes=True,
         exprs=dict(length="0.5 + rand(0.2, 1.5) * map('len.ptx')",
                    width="map('w.ptx') * 0.05 + $id * 0.00001",
                    taper="0.3 + 0.4 * map('t.ptx')", taper_start="clamp(rand(), 0.1, 0.6)",
                    width_ramp="rampUI(0,1,1:0.5,0.6,2:1,0.1,3)",
                    cut_amount="$cLength * (0.1 + 0.3 * rand(7)) + $cWidth")),
    # the generator's own strands: width from the radii, a cLength-reading ramp
    dict(name="icoRamp", description_id=3, strands="ico5", ids=True,
         exprs=dict(taper="0.8", width_ramp="$cLength > 0.3 ? 1 : 0.5", cut_amount="0.2 * $cLength")),
    # no CutFX; a $Prefg noise input after a PTEX input; radius scale 2
    dict(name="icoNoCut", description_id=5, strands="ico13", radius_scale=2.0,
         exprs=dict(length="smoothstep(0, 1, map('a.ptx'))->gamma(2) + 0.5",
                    width="0.01 * (1 + noise($Prefg*4))", taper="0.5", taper_start="0.2")),
    # the culling thresholds (tests/classic_test.cpp:316-336) by $id
    dict(name="threshold", description_id=9, strands="walk", ids=True,
         exprs=dict(length="$id < 100 ? 0 : 1", width="$id < 200 ? 0.0001 : 0.000099")),
    # a cut that removes everything, then a ramp on the collapsed strand
    dict(name="cutAll", description_id=2, strands="ico5",
         exprs=dict(cut_amount="$cLength * 2", width_ramp="$cLength + 0.5")),
    # errors: the reference's messages
    dict(name="negativeLength", description_id=1, strands="walk", exprs=dict(length="$id == 50 ? -1 : 1")),
    dict(name="nonfiniteWidth", description_id=1, strands="walk", ids=True, exprs=dict(width="1 / ($id - 5)")),
    # a widthRamp of t alone that divides by zero at t = 0.5 (CV 2 of 5)
    dict(name="uniformRampError", description_id=1, strands="ico5",
         exprs=dict(width_ramp="1 / (0.5 - rampUI(0,0,1:1,1,1))")),
    # ClumpingFX before the CutFX: expression-driven mask (PTEX input), amount
    # (strand prefix), per-CV clumpScale, noise through the guides' prefixes,
    # a per-strand noiseScale; the cut amount and taper read t, which the clump
    # leaves at 1 for the strands it touched
    dict(name="ns:clumpCut", description_id=21, strands="walk", ids=True, prefixes=True, guides=12,
         exprs=dict(width="0.01 + $cWidth", taper="rampUI(0,0.2,1:1,0.7,1)",
                    cut_amount="$cLength * 0.2 + rampUI(0,0,1:1,0.1,1)"),
         clump=dict(mask="0.3 + 0.7 * map('m.ptx')", clump="0.6 + 0.2 * rand()",
                    clumpScale="rampUI(0,0,1:1,0.3,1)", noise="0.2 * rand(5)", noiseScale="1 - $cWidth * 10",
                    noiseFrequency="2 + rand()", noiseCorrelation="30")),
    # the generator's strands, a clump reading cLength, no noise, no cut
    dict(name="icoClump", description_id=22, strands="ico5", guides=9,
         exprs=dict(width_ramp="$cLength"),
         clump=dict(mask="1", clump="clamp($cLength, 0, 1)", clumpScale="0.1 * $cLength", noise="0",
                    noiseScale="1", noiseFrequency="1", noiseCorrelation="0")),
    dict(name="clumpBadGuide", description_id=23, strands="walk", guides=4, bad_guide=True,
         clump=dict(mask="1", clump="0.5", clumpScale="0", noise="0", noiseScale="1", noiseFrequency="1",
                    noiseCorrelation="0"), exprs=dict()),
    dict(name="clumpMagnitude", description_id=24, strands="walk", guides=4,
         clump=dict(mask="0.5", clump="0.5", clumpScale="0", noise="1e30", noiseScale="1e30", noiseFrequency="1",
                    noiseCorrelation="0"), exprs=dict()),
    # hierarchical clumping: a coarse module, then a fine one with PTEX mask,
    # guide-prefix noise and a cLength-reading amount, then the cut; the cut
    # amount and taper read the t the last module with a guide leaves
    dict(name="ns:hierarchy", description_id=31, strands="walk", ids=True, prefixes=True, guides=[6, 14],
         exprs=dict(taper="rampUI(0,0.1,1:1,0.8,1)", cut_amount="$cLength * 0.15 + rampUI(0,0,1:1,0.05,1)"),
         clumps=[dict(mask="1", clump="0.5 + 0.3 * rand()", clumpScale="rampUI(0,0,1:1,0.25,1)", noise="0",
                      noiseScale="1", noiseFrequency="1", noiseCorrelation="0"),
                 dict(mask="0.2 + 0.8 * map('fine.ptx')", clump="clamp($cLength * 0.4, 0, 1)",
                      clumpScale="0.1", noise="0.05 + 0.1 * rand(3)", noiseScale="1 - 0.5 * $cWidth",
                      noiseFrequency="3 + rand()", noiseCorrelation="50")]),
    # three levels on the generator's strands; the middle one masks a third of
    # the strands out (their context's t goes back to 0)
    dict(name="threeLevels", description_id=32, strands="ico5", ids=True, guides=[5, 9, 13],
         exprs=dict(width_ramp="1 - 0.5 * $cLength", cut_amount="rampUI(0,0,1:1,0.02,1)"),
         clumps=[dict(mask="0.9", clump="0.4", clumpScale="0", noise="0", noiseScale="1", noiseFrequency="1",
                      noiseCorrelation="0"),
                 dict(mask="$id < 100 ? 0 : 1", clump="0.3 + 0.1 * $cLength", clumpScale="rampUI(0,0.1,1:1,0,1)",
                      noise="0", noiseScale="1", noiseFrequency="2", noiseCorrelation="10"),
                 dict(mask="0.5", clump="0.6", clumpScale="0.05 * $cLength", noise="0", noiseScale="1",
                      noiseFrequency="1", noiseCorrelation="0")]),
    # the CutFX between two modules: the second clumps the cut strands (a
    # strand the cut culls still runs it); the cut amount reads a PTEX input
    # numbered in object order, after the first module's
    dict(name="ns:cutBetween", description_id=33, strands="walk", ids=True, prefixes=True, guides=[8, 11],
         clumps_after_cut=1,
         exprs=dict(taper="rampUI(0,0.2,1:1,0.5,1)", cut_amount="$cLength * (0.2 + 0.6 * map('cut.ptx'))"),
         clumps=[dict(mask="map('a.ptx')", clump="0.4 + 0.4 * rand()", clumpScale="0", noise="0.05",
                      noiseScale="1", noiseFrequency="2", noiseCorrelation="30"),
                 dict(mask="1", clump="clamp($cLength, 0, 0.8)", clumpScale="rampUI(0,0,1:1,0.3,1)", noise="0",
                      noiseScale="1", noiseFrequency="1", noiseCorrelation="0")]),
    # NoiseFX: PTEX mask, rand() magnitude through the strand's prefix, a
    # per-CV magnitude scale, preserve length 50%; the cut and taper after it
    # read the t of its last CV
    dict(name="ns:noiseCut", description_id=41, strands="walk", ids=True, prefixes=True,
         exprs=dict(taper="rampUI(0,0.1,1:1,0.6,1)", cut_amount="$cLength * 0.1 + rampUI(0,0,1:1,0.05,1)"),
         noise=dict(mask="0.5 + 0.5 * map('n.ptx')", magnitude="0.2 + 0.1 * rand()",
                    magnitudeScale="rampUI(0,0,1:1,1,1)", frequency="3", correlation="20", preserveLength="50")),
    # a clump, then a NoiseFX with domain positions, masked out for some
    # strands (their t stays where the clump left it); the generator's strands
    dict(name="clumpNoise", description_id=42, strands="ico5", ids=True, guides=9, noise_domain=True,
         exprs=dict(width_ramp="1 - 0.3 * $cLength"),
         clump=dict(mask="0.8", clump="0.5", clumpScale="0", noise="0", noiseScale="1", noiseFrequency="1",
                    noiseCorrelation="0"),
         noise=dict(mask="$id < 80 ? 0 : 1", magnitude="0.05 * $cLength", magnitudeScale="1", frequency="1.5",
                    correlation="0", preserveLength="0")),
    # a length expression, then a NoiseFX with no clump before it, masked
    # out for the first 130 strands: cLength is re-measured after the module
    # whatever its mask (the scaled value gives way to the spline length),
    # and the cut amount reads it
    dict(name="noiseMaskedLength", description_id=44, strands="walk", ids=True,
         exprs=dict(length="1.37", taper="0.5", cut_amount="$cLength * 0.3"),
         noise=dict(mask="$id < 130 ? 0 : 1", magnitude="0.04", magnitudeScale="1", frequency="2",
                    correlation="0", preserveLength="0")),
    # the NoiseFX magnitude error
    dict(name="noiseMagnitude", description_id=43, strands="walk", exprs=dict(),
         noise=dict(mask="1", magnitude="1e30", magnitudeScale="1e30", frequency="1", correlation="0",
                    preserveLength="0")),
]


def case_clumps(case):
    """The case's ClumpingFX modules (attribute dicts), in object order."""
    if "clumps" in case:
        return case["clumps"]
    return [case["clump"]] if "clump" in case else []


def clump_names(case):
    """The modules' names as xgen_text writes them."""
    return ["clump" if m == 0 else f"clump{m}" for m in range(len(case_clumps(case)))]


def compile_order(case):
    """(key, attribute) in the reference's compile order; clump keys are
    "clump<m>.<attribute>"; the CutFX sits before the last
    case["clumps_after_cut"] modules."""
    count = len(case_clumps(case))
    before = count - case.get("clumps_after_cut", 0)
    clump = lambda ms: tuple((f"clump{m}.{a}", a) for m in ms for a in CLUMP)
    noise = tuple((f"noise.{a}", a) for a in NOISE) if "noise" in case else ()
    return PRIMITIVE + clump(range(before)) + noise + (("cut_amount", "amount"),) + clump(range(before, count))


def xgen_text(case):
    """The case as a Classic .xgen collection (FileVersion 18)."""
    e = case["exprs"]
    s = "FileVersion 18\nPalette\n\tname\tpalette\n\tendAttrs\n"
    s += f"Description\n\tname\t{case['name']}\n\tdescriptionId\t{case['description_id']}\n\tendAttrs\n"
    s += "SplinePrimitive\n"
    for key, attribute in PRIMITIVE:
        if key in e:
            s += f"\t{attribute}\t{e[key]}\n"
    s += "\tendAttrs\n"
    modules = list(zip(case_clumps(case), clump_names(case)))
    before = len(modules) - case.get("clumps_after_cut", 0)
    for i, module in enumerate(modules + [None]):
        if i == before and "noise" in case:
            s += "NoiseFXModule\n\tactive\ttrue\n\tname\tnoise\n\tmode\t0\n"
            s += "".join(f"\t{a}\t{case['noise'][a]}\n" for a in NOISE) + "\tendAttrs\n"
        if i == before and "cut_amount" in e:
            s += f"CutFXModule\n\tactive\ttrue\n\tname\tcut\n\tamount\t{e['cut_amount']}\n\trebuildType\t1\n\tendAttrs\n"
        if module is None:
            break
        s += f"ClumpingFXModule\n\tactive\ttrue\n\tname\t{module[1]}\n"
        s += "".join(f"\t{a}\t{module[0][a]}\n" for a in CLUMP) + "\tendAttrs\n"
    return s + "\tActive\tSplinePrimitive\n"


_MAP = re.compile(r"\bmap\s*\(\s*(['\"])(.*?)\1\s*\)")
_PREF = re.compile(r"\bnoise\s*\(\s*\$Prefg\s*\*\s*([-+]?[0-9.]+)\s*\)")


def plan_sources(case):
    """(sources in the bound form, ptex count, pref-noise frequencies); clump
    expressions under the "clump" key by attribute name (a list of such
    dicts for several modules)."""
    out, ptex, freqs = {}, 0, []
    modules = case_clumps(case)
    for key, _ in compile_order(case):
        if key.startswith("noise."):
            src = case["noise"][key[6:]]
        elif key.startswith("clump"):
            m, attribute = key[5:].split(".")
            src = modules[int(m)][attribute]
        elif key not in case["exprs"]:
            continue
        else:
            src = case["exprs"][key]

        def map_sub(m):
            nonlocal ptex
            ptex += 1
            return f"$__nxg_map_{ptex - 1}"

        src = _MAP.sub(map_sub, src)
        out[key] = src
    for key in list(out):
        def pref_sub(m):
            f = np.float32(float(m.group(1)))
            if f not in freqs:
                freqs.append(f)
            return f"$__nxg_pref_noise_{freqs.index(f)}"

        out[key] = _PREF.sub(pref_sub, out[key])
    clumps = [{a: out.pop(f"clump{m}.{a}") for a in CLUMP} for m in range(len(modules))]
    if clumps:
        out["clump"] = clumps if "clumps" in case else clumps[0]
    if case.get("clumps_after_cut"):
        out["clumps_after_cut"] = case["clumps_after_cut"]
    if "noise" in case:
        out["noise"] = {a: out.pop(f"noise.{a}") for a in NOISE}
    return out, ptex, freqs


def clump_guides(count, cvs, rng):
    """Seeded world-space clump guides (curled verticals over [-5, 5]^2), their
    uvs and SeExpr prefixes."""
    roots = np.stack([rng.uniform(-5, 5, count), np.zeros(count), rng.uniform(-5, 5, count)], 1)
    t = np.linspace(0, 1, cvs)
    length = rng.uniform(0.5, 2.0, count)
    phase = rng.uniform(0, 6.28, count)
    axes = np.zeros((count, cvs, 3))
    axes[:, :, 0] = roots[:, None, 0] + 0.2 * np.sin(phase[:, None] + 3 * t[None])
    axes[:, :, 1] = length[:, None] * t[None]
    axes[:, :, 2] = roots[:, None, 2] + 0.2 * np.cos(phase[:, None] + 3 * t[None])
    normals = np.tile([0.0, 1.0, 0.0], (count, 1)) + rng.normal(0, 0.1, (count, 3))
    tangents = np.tile([1.0, 0.0, 0.0], (count, 1))
    return dict(axes=axes.reshape(-1, 3).astype(np.float32), normals=normals.astype(np.float32),
                tangents=tangents.astype(np.float32), reference=(roots + 0.01).astype(np.float32),
                uvs=rng.uniform(0, 1, (count, 2)).astype(np.float32),
                prefixes=rng.integers(0, 2 ** 32, count, dtype=np.uint64).astype(np.uint32))


def case_inputs(case, strands_of):
    """Inputs of a case: dict(points, cvs, roots, uvs, ids, prefixes, columns).
    strands_of(name) -> (packed float4 points, cvs, roots, uvs)."""
    from nanoxgen_amd import abi
    if case["strands"] == "line3":
        pts = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [2, 0, 0, 0]], np.float32)
        cvs, roots, uvs = 3, np.zeros(1, abi.ROOT_DTYPE), None
    else:
        pts, cvs, roots, uvs = strands_of(case["strands"])
    n = len(pts) // cvs
    rng = np.random.default_rng(zlib.crc32(case["name"].encode()))
    _, ptex, freqs = plan_sources(case)
    k = ptex + len(freqs)
    columns = rng.uniform(0.05, 1.0, size=(n, k)).astype(np.float32) if k else None
    ids = (rng.permutation(n).astype(np.uint32) if case.get("ids") else None)
    prefixes = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32) if case.get("prefixes") else None
    out = dict(points=pts, cvs=cvs, roots=roots, uvs=uvs, ids=ids, prefixes=prefixes, columns=columns,
               radius_scale=case.get("radius_scale", 1.0))
    counts = case.get("guides")
    for m, count in enumerate([] if counts is None else counts if isinstance(counts, list) else [counts]):
        g = clump_guides(count, cvs, rng)
        strand_guides = rng.integers(0, count, n).astype(np.uint32)
        strand_guides[m::7 - m] = 0xFFFFFFFF
        if case.get("bad_guide"):
            strand_guides[n // 2] = count
        suffix = "" if m == 0 else f"_{m}"
        out.update({f"{k}{suffix}": v for k, v in dict(
            guide_axes=g["axes"], guide_normals=g["normals"], guide_tangents=g["tangents"],
            guide_reference=g["reference"], guide_uvs=g["uvs"], guide_prefixes=g["prefixes"],
            strand_guides=strand_guides).items()})
    if "noise" in case:  # surface tangents (every 9th zero: the fallback frame), domain positions
        tangents = rng.normal(0, 1, (n, 3)).astype(np.float32)
        tangents[::9] = 0.0
        out["surface_tangents"] = tangents
        if case.get("noise_domain"):
            out["noise_domain"] = rng.uniform(-2, 2, (n, 3)).astype(np.float32)
    return out


def guide_args(inp):
    """The `clumps` keyword argument of case_inputs' output: one binding per
    module (empty without guides)."""
    clumps = []
    while ("guide_axes" + ("" if not clumps else f"_{len(clumps)}")) in inp:
        x = lambda k: inp[k + ("" if not clumps else f"_{len(clumps)}")]
        clumps.append(dict(strand_guides=x("strand_guides"), guide_uvs=x("guide_uvs"),
                           guide_random_prefixes=x("guide_prefixes"),
                           guides=dict(axes=x("guide_axes"), normals=x("guide_normals"),
                                       tangents=x("guide_tangents"), reference=x("guide_reference"))))
    out = dict(clumps=clumps) if clumps else {}
    if "surface_tangents" in inp:
        out["surface_tangents"] = inp["surface_tangents"]
        out["noise_domain_positions"] = inp.get("noise_domain")
    return out
