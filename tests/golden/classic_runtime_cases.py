"""Cases of the Classic runtime with primitive expressions, shared by the
fixture writer (make_golden.py classic_runtime) and the tests.

Each case is a Classic description: SplinePrimitive attributes (length,
width, taper, taperStart, widthRamp) and FX modules, written in the
authoring form the reference's plan compiler reads (map('path') for PTEX
inputs, noise($Prefg*f) for reference-pose noise).  The modules are either
an `effects` list in object order (("clump", its seven attributes),
("noise", its six), ("cut", the amount)) or the short form: ClumpingFX
modules (`clump`: one, `clumps`: several), an optional `noise` module and
an optional CutFX amount (exprs["cut_amount"]) after all clumps but the last
`clumps_after_cut`.  `root_relative` runs the runtime about each strand's
root (the strands are given relative to it) and `local` lists the clump
modules whose guides carry translation-free local axes.  `xgen_text` renders
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
    # effect lists in object order: a CutFX before a NoiseFX (the noise runs
    # on the rebuilt strand; a culled strand still runs it), the NoiseFX's
    # mask and frequency read the t the cut leaves (0)
    dict(name="ns:cutNoise", description_id=51, strands="walk", ids=True, prefixes=True,
         exprs=dict(taper="rampUI(0,0.1,1:1,0.6,1)", width_ramp="1 - 0.2 * $cLength"),
         effects=[("cut", "$cLength * (0.15 + 0.2 * rand(4))"),
                  ("noise", dict(mask="0.5 + 0.5 * map('n.ptx')", magnitude="0.08 + 0.04 * rand()",
                                 magnitudeScale="rampUI(0,0,1:1,1,1)", frequency="2 + $cLength",
                                 correlation="10", preserveLength="30"))]),
    # two NoiseFX modules around a clump, then a cut whose amount reads the t
    # the second NoiseFX leaves; the second one masks a third of the strands out
    dict(name="twoNoises", description_id=52, strands="ico5", ids=True, guides=9, noise_domain=True,
         exprs=dict(width_ramp="1 - 0.3 * $cLength"),
         effects=[("noise", dict(mask="1", magnitude="0.03", magnitudeScale="1", frequency="1.5", correlation="0",
                                 preserveLength="0")),
                  ("clump", dict(mask="0.7", clump="0.4", clumpScale="0", noise="0", noiseScale="1",
                                 noiseFrequency="1", noiseCorrelation="0")),
                  ("noise", dict(mask="$id < 100 ? 0 : 1", magnitude="0.02 * $cLength",
                                 magnitudeScale="rampUI(0,0.5,1:1,1,1)", frequency="3", correlation="40",
                                 preserveLength="60")),
                  ("cut", "$cLength * 0.1 + rampUI(0,0,1:1,0.02,1)")]),
    # two CutFX modules around a clump: the first rebuilds the strand for the
    # clump (a strand it culls still runs the rest), the second reads cLength
    # and the t the clump leaves
    dict(name="ns:twoCuts", description_id=53, strands="walk", ids=True, prefixes=True, guides=10,
         exprs=dict(taper="rampUI(0,0.2,1:1,0.5,1)"),
         effects=[("cut", "$cLength * 0.2 * map('c.ptx')"),
                  ("clump", dict(mask="1", clump="0.5 + 0.3 * rand()", clumpScale="rampUI(0,0,1:1,0.2,1)",
                                 noise="0", noiseScale="1", noiseFrequency="1", noiseCorrelation="0")),
                  ("cut", "$cLength * 0.1 + rampUI(0,0,1:1,0.05,1)")]),
    # root-relative evaluation: strands given about their roots, a module
    # with local (translation-free) guide axes and one without, then a cut
    dict(name="rootRelative", description_id=54, strands="walk", ids=True, prefixes=True, guides=[8, 6],
         root_relative=True, local=[0], exprs=dict(taper="0.4"),
         effects=[("clump", dict(mask="1", clump="0.6", clumpScale="rampUI(0,0,1:1,0.3,1)", noise="0",
                                 noiseScale="1", noiseFrequency="1", noiseCorrelation="0")),
                  ("clump", dict(mask="$id < 120 ? 0 : 0.8", clump="0.3 + 0.2 * rand()", clumpScale="0",
                                 noise="0", noiseScale="1", noiseFrequency="1", noiseCorrelation="0")),
                  ("cut", "$cLength * 0.15")]),
    # root-relative with clump noise: the noise frame of the local render
    # axis, its domain at the guide's world reference position; local axes
    # given but not root-relative set only the guides' lengths
    dict(name="rootRelativeNoise", description_id=55, strands="ico5", ids=True, guides=[7, 5], root_relative=True,
         local=[0, 1], exprs=dict(width_ramp="1 - 0.2 * $cLength"),
         effects=[("clump", dict(mask="0.9", clump="0.5", clumpScale="0.05", noise="0.2", noiseScale="1",
                                 noiseFrequency="2", noiseCorrelation="20")),
                  ("clump", dict(mask="1", clump="0.3", clumpScale="0", noise="0.1",
                                 noiseScale="rampUI(0,0,1:1,1,1)", noiseFrequency="3", noiseCorrelation="0"))]),
    dict(name="localLengths", description_id=56, strands="walk", guides=6, local=[0],
         exprs=dict(cut_amount="$cLength * 0.2"),
         clump=dict(mask="1", clump="0.5", clumpScale="0", noise="0", noiseScale="1", noiseFrequency="1",
                    noiseCorrelation="0")),
]


def case_effects(case):
    """[(kind, module)] in object order: kind "clump" (attribute dict),
    "noise" (attribute dict) or "cut" (the amount source)."""
    if "effects" in case:
        return list(case["effects"])
    clumps = [case["clump"]] if "clump" in case else case.get("clumps", [])
    before = len(clumps) - case.get("clumps_after_cut", 0)
    effects = [("clump", c) for c in clumps[:before]]
    if "noise" in case:
        effects.append(("noise", case["noise"]))
    if "cut_amount" in case["exprs"]:
        effects.append(("cut", case["exprs"]["cut_amount"]))
    return effects + [("clump", c) for c in clumps[before:]]


def has_noise(case):
    """Whether the case has a NoiseFX or a clump with its noise on (positions
    then go through the device's acosf, cosf and sinf)."""
    return any(k == "noise" or (k == "clump" and m["noise"] != "0") for k, m in case_effects(case))


def case_clumps(case):
    """The case's ClumpingFX modules (attribute dicts), in object order."""
    return [m for k, m in case_effects(case) if k == "clump"]


def clump_names(case):
    """The modules' names as xgen_text writes them."""
    return ["clump" if m == 0 else f"clump{m}" for m in range(len(case_clumps(case)))]


def indexed_effects(case):
    """[(kind, index among the modules of its kind, module)] in object order."""
    seen = dict(clump=0, noise=0, cut=0)
    out = []
    for kind, module in case_effects(case):
        out.append((kind, seen[kind], module))
        seen[kind] += 1
    return out


def compile_order(case):
    """(key, attribute) in the reference's compile order (the primitive's
    attributes, then the FX modules in object order); module keys are
    "clump<m>.<attribute>", "noise<n>.<attribute>" and "cut<c>"."""
    order = PRIMITIVE
    for kind, index, _ in indexed_effects(case):
        if kind == "clump":
            order += tuple((f"clump{index}.{a}", a) for a in CLUMP)
        elif kind == "noise":
            order += tuple((f"noise{index}.{a}", a) for a in NOISE)
        else:
            order += ((f"cut{index}", "amount"),)
    return order


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
    names = clump_names(case)
    for kind, index, module in indexed_effects(case):
        if kind == "noise":
            s += f"NoiseFXModule\n\tactive\ttrue\n\tname\t{'noise' if index == 0 else f'noise{index}'}\n\tmode\t0\n"
            s += "".join(f"\t{a}\t{module[a]}\n" for a in NOISE) + "\tendAttrs\n"
        elif kind == "cut":
            s += (f"CutFXModule\n\tactive\ttrue\n\tname\t{'cut' if index == 0 else f'cut{index}'}\n"
                  f"\tamount\t{module}\n\trebuildType\t1\n\tendAttrs\n")
        else:
            s += f"ClumpingFXModule\n\tactive\ttrue\n\tname\t{names[index]}\n"
            s += "".join(f"\t{a}\t{module[a]}\n" for a in CLUMP) + "\tendAttrs\n"
    return s + "\tActive\tSplinePrimitive\n"


_MAP = re.compile(r"\bmap\s*\(\s*(['\"])(.*?)\1\s*\)")
_PREF = re.compile(r"\bnoise\s*\(\s*\$Prefg\s*\*\s*([-+]?[0-9.]+)\s*\)")


def plan_sources(case):
    """(ClassicPlan keyword arguments with the sources in the bound form, ptex
    count, pref-noise frequencies): an `effects` list for a case given as one,
    else the short form (clump expressions under "clump" by attribute name,
    a list of such dicts for several modules; "noise"; "cut_amount")."""
    out, ptex, freqs = {}, 0, []
    modules = {(kind, index): module for kind, index, module in indexed_effects(case)}
    for key, _ in compile_order(case):
        if key.startswith("noise"):
            n, attribute = key[5:].split(".")
            src = modules[("noise", int(n))][attribute]
        elif key.startswith("clump"):
            m, attribute = key[5:].split(".")
            src = modules[("clump", int(m))][attribute]
        elif key.startswith("cut"):
            src = modules[("cut", int(key[3:]))]
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
    bound = []
    for kind, index, _ in indexed_effects(case):
        if kind == "clump":
            bound.append((kind, {a: out.pop(f"clump{index}.{a}") for a in CLUMP}))
        elif kind == "noise":
            bound.append((kind, {a: out.pop(f"noise{index}.{a}") for a in NOISE}))
        else:
            bound.append((kind, out.pop(f"cut{index}")))
    if "effects" in case:
        out["effects"] = bound
        return out, ptex, freqs
    clumps = [m for k, m in bound if k == "clump"]
    if clumps:
        out["clump"] = clumps if "clumps" in case else clumps[0]
    if case.get("clumps_after_cut"):
        out["clumps_after_cut"] = case["clumps_after_cut"]
    for kind, m in bound:
        if kind == "noise":
            out["noise"] = m
        elif kind == "cut":
            out["cut_amount"] = m
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
    if case.get("root_relative"):  # the strands about their roots
        pts = pts.copy()
        pts.reshape(n, cvs, 4)[:, :, :3] -= np.asarray(roots["position"], np.float32)[:, None, :]
    rng = np.random.default_rng(zlib.crc32(case["name"].encode()))
    _, ptex, freqs = plan_sources(case)
    k = ptex + len(freqs)
    columns = rng.uniform(0.05, 1.0, size=(n, k)).astype(np.float32) if k else None
    ids = (rng.permutation(n).astype(np.uint32) if case.get("ids") else None)
    prefixes = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32) if case.get("prefixes") else None
    out = dict(points=pts, cvs=cvs, roots=roots, uvs=uvs, ids=ids, prefixes=prefixes, columns=columns,
               radius_scale=case.get("radius_scale", 1.0))
    if case.get("root_relative"):
        out["root_relative"] = np.uint32(1)
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
        if m in case.get("local", ()):  # translation-free copies: each guide about its first CV
            axes = g["axes"].reshape(count, cvs, 3)
            out[f"guide_local{suffix}"] = (axes - axes[:, :1]).reshape(-1, 3).astype(np.float32)
    if any(k == "noise" for k, _ in case_effects(case)):
        # surface tangents (every 9th zero: the fallback frame), domain positions
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
        suffix = "" if not clumps else f"_{len(clumps)}"
        x = lambda k: inp[k + suffix]
        clumps.append(dict(strand_guides=x("strand_guides"), guide_uvs=x("guide_uvs"),
                           guide_random_prefixes=x("guide_prefixes"), local_axes=inp.get("guide_local" + suffix),
                           guides=dict(axes=x("guide_axes"), normals=x("guide_normals"),
                                       tangents=x("guide_tangents"), reference=x("guide_reference"))))
    out = dict(clumps=clumps) if clumps else {}
    if "surface_tangents" in inp:
        out["surface_tangents"] = inp["surface_tangents"]
        out["noise_domain_positions"] = inp.get("noise_domain")
    if inp.get("root_relative") is not None and int(inp["root_relative"]):
        out["root_relative"] = True
    return out


EFFECT_MODULES = dict(
    clump=[dict(mask="0.3 + 0.7 * map('m')", clump="0.4 + 0.4 * rand()", clumpScale="rampUI(0,0,1:1,0.2,1)",
                noise="0.1 * rand(5)", noiseScale="1", noiseFrequency="2", noiseCorrelation="10"),
           dict(mask="$id < 50 ? 0 : 1", clump="clamp($cLength * 0.3, 0, 1)", clumpScale="0", noise="0",
                noiseScale="1", noiseFrequency="1", noiseCorrelation="0")],
    noise=[dict(mask="0.5 + 0.5 * map('n')", magnitude="0.05 + 0.05 * rand()", magnitudeScale="rampUI(0,0,1:1,1,1)",
                frequency="2", correlation="20", preserveLength="40"),
           dict(mask="$id < 80 ? 0 : 1", magnitude="0.02 * $cLength", magnitudeScale="1", frequency="1 + $cLength",
                correlation="0", preserveLength="0")],
    cut=["$cLength * (0.1 + 0.3 * map('c'))", "$cLength * 0.1 + rampUI(0,0,1:1,0.05,1)"])


def random_effect_cases(rng, count):
    """Seeded effect lists of one to six modules in random object order (at
    most two of each kind), with random primitive expressions."""
    cases = []
    for k in range(count):
        pool = [(kind, m) for kind in ("clump", "noise", "cut") for m in range(2)]
        picks = [pool[j] for j in rng.permutation(len(pool))[:rng.integers(1, 7)]]
        effects = [(kind, EFFECT_MODULES[kind][m]) for kind, m in picks]
        exprs = dict(taper="rampUI(0,0.1,1:1,0.6,1)") if rng.uniform() < 0.5 else dict(width_ramp="1 - 0.2 * $cLength")
        cases.append(dict(name=f"ns:order{k}", description_id=60 + k, strands="walk", ids=True, prefixes=True,
                          exprs=exprs, effects=effects))
    return cases
