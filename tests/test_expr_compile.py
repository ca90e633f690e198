"""The expression compiler (nxg_expr_compile): text to the float IR,
instruction for instruction the reference's compile_xgen_scalar_expression +
make_xgen_float_expression_program, with the same diagnostics; the
reference's own test values (tests/expression_test.cpp) through the oracle's
float interpreter.  CPU only."""
import numpy as np
import pytest

import nanoxgen_amd as nx
from nanoxgen_amd import expr


def same_program(a, b):
    n_ins, n_ops, n_pts, n_ramps = a.counts
    return (np.array_equal(a.ins[:n_ins].view(np.uint8), b["instructions"].view(np.uint8))
            and np.array_equal(a.ops[:n_ops], b["operands"])
            and np.array_equal(a.pts[:n_pts].view(np.uint8), b["ramp_points"].view(np.uint8))
            and np.array_equal(a.ramps_arr[:n_ramps], b["ramps"]) and a.inputs == b["inputs"]
            and a.result == b["result"] and a.expression_seed == b["expression_seed"]
            and a.random_call_count == b["random_call_count"])


REFERENCE_SOURCES = [  # tests/expression_test.cpp
    "$a=hash($foo+355) <= .3? rand(0.1,0.7):rand(0.2,0);\\n$a*$bar",
    "$a=rand(0.1,0.8);$b=rand(1,3);$a+$b",
    "$percent=30;#0,100\\n$inside=hash($id)<=($percent/100);\\n$inside ? max(2,clamp($length,0,1)) : fit(.5,0,1,4,6)",
    "$a=.5;$a->fit(0,1,4,6)\n*\n1",
    "hi=.25;$lo=.75;$hi+$lo",
    ".25->gamma(2)->contrast(.9)+smoothstep(.5,1,.34)",
    "$a/0",
    "rampUI(0,0,0:0.5,1,0:1,0,0)", "rampUI(0,0,1:0.5,1,1:1,0,1)", "rampUI(0,0,2:0.5,1,2:1,0,2)",
    "rampUI(0,0,3:0.5,1,3:1,0,3)",
    "rampUI(0,0,3:0.19692307692307692,0.3108108108108108,1:0.4523076923076923,0.3783783783783784,1:"
    "0.683076923076923,0.35135135135135137,1:1,0.24324324324324326,1)",
    "rampUI(0,-1,1:1,2,1)",
]

BAD_SOURCES = [
    "map('${DESC}/paintmaps/mask')", "noise($u)", "1 +", "$", "a & b", "a | b", "(1", "1)", "min(1)", "x", "",
    "1;;2 3", "rampUI(0,0,1:0,1,1)", "rampUI(-.1,0,1:1,1,1)", "rampUI(0,0,4:1,1,4)", "rampUI(0,0)", "'abc",
    "1 ? 2", "$a->", "$a->f(", "rand(1,2,3,4)", "hash()", "fit(1,2)", "1e999", "$a = ", "@", "true(1)",
]


@pytest.fixture(scope="module")
def ref(orc):
    if orc.ref() is None:
        pytest.skip("reference core not built here")
    return orc


@pytest.mark.parametrize("source", REFERENCE_SOURCES)
def test_reference_test_sources(ref, source):
    for name, kind in (("", ""), ("amount", "CutFXModule"), ("noise", "ClumpingFXModule")):
        got = expr.compile_expression(source, name, kind)
        want, err = ref.ref_expr_compile(source, name, kind)
        assert err == "" and same_program(got, want)


@pytest.mark.parametrize("source", BAD_SOURCES)
def test_diagnostics_match_reference(ref, source):
    _, want = ref.ref_expr_compile(source)
    assert want
    with pytest.raises(ValueError) as e:
        expr.compile_expression(source)
    assert str(e.value) == want


def test_limits_match_reference(ref):
    for source, limits in (("1234", (3, 0, 0, 0, 0)), ("1+2", (0, 0, 2, 0, 0)), ("1+2+3", (0, 4, 0, 0, 0)),
                           ("$a=1;$b=2;$a+$b", (0, 0, 0, 1, 0)), ("rampUI(0,0,1:1,1,1)", (0, 0, 0, 0, 1))):
        _, want = ref.ref_expr_compile(source, limits=limits)
        with pytest.raises(ValueError) as e:
            expr.compile_expression(source, limits=limits)
        assert str(e.value) == want


def random_source(rng, depth=0, locals_=()):
    """A random expression in the reference's language."""
    leaf = rng.random() < 0.35 + 0.15 * depth
    if leaf:
        k = rng.integers(0, 5)
        if k == 0:
            return f"{rng.uniform(-3, 3):.{rng.integers(0, 6)}f}"
        if k == 1:
            return "$" + rng.choice(["u", "v", "id", "cLength", "foo", "bar"])
        if k == 2 and locals_:
            return "$" + rng.choice(list(locals_))
        if k == 3:
            return rng.choice(["true", "false", ".5", "1e-3", "7"])
        return f"rampUI(0,{rng.uniform(0, 1):.3f},{rng.integers(0, 4)}:1,{rng.uniform(0, 1):.3f},{rng.integers(0, 4)})"
    sub = lambda: random_source(rng, depth + 1, locals_)
    k = rng.integers(0, 9)
    if k == 0:
        return f"{sub()} {rng.choice(['+', '-', '*', '/', '<', '<=', '>', '>=', '==', '!=', '&&', '||'])} {sub()}"
    if k == 1:
        return f"{sub()} ? {sub()} : {sub()}"
    if k == 2:
        return f"{rng.choice(['-', '!', '+'])}{sub()}"
    if k == 3:
        return f"({sub()})"
    if k == 4:
        fn, n = [("min", 2), ("max", 2), ("clamp", 3), ("fit", 5), ("gamma", 2), ("contrast", 2),
                 ("smoothstep", 3)][rng.integers(0, 7)]
        return f"{fn}({', '.join(sub() for _ in range(n))})"
    if k == 5:
        return f"rand({', '.join(sub() for _ in range(rng.integers(0, 4)))})"
    if k == 6:
        return f"hash({', '.join(sub() for _ in range(rng.integers(1, 4)))})"
    if k == 7:
        return f"({sub()})->{rng.choice(['gamma', 'contrast'])}({sub()})"
    return f"{sub()}\n{rng.choice(['+', '*'])}\n{sub()}"


def test_random_sources_match_reference(ref):
    rng = np.random.default_rng(12)
    for i in range(400):
        stmts, names = [], []
        for _ in range(rng.integers(0, 3)):
            name = f"l{len(names)}"
            stmts.append(f"${name} = {random_source(rng, 0, tuple(names))}")
            names.append(name)
        stmts.append(random_source(rng, 0, tuple(names)))
        sep = rng.choice([";", ";\n", "; # note\n", "\\n"])
        source = sep.join(stmts)
        want, err = ref.ref_expr_compile(source, f"n{i % 7}", "SplinePrimitive")
        if want is None:
            with pytest.raises(ValueError) as e:
                expr.compile_expression(source, f"n{i % 7}", "SplinePrimitive")
            assert str(e.value) == err
        else:
            assert same_program(expr.compile_expression(source, f"n{i % 7}", "SplinePrimitive"), want), source


def test_seeds_pin():
    import ctypes as C
    L = nx.lib()
    one = (C.c_double * 1)(1.0)
    assert abs(L.nxg_seexpr_hash(one, 1) - 0.40419923151009696) < 1e-15
    assert L.nxg_string_seed(b"widthSplinePrimitive") == 83
    assert expr.compile_expression("1", "width", "SplinePrimitive").expression_seed == 83
    assert abs(L.nxg_face_seed(0, b"rabbitShape", 7) - 0.82524268837302051) < 1e-15
    assert L.nxg_face_seed(0, b"namespace:rabbitShape", 7) == L.nxg_face_seed(0, b"rabbitShape", 7)


def ref_float_eval(R, p, inputs, u=0.0, v=0.0, face_seed=0.0, t=0.0):
    import ctypes as C
    R.ref_expr_eval.restype = C.c_int
    R.ref_expr_eval.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_float,
                                C.c_float, C.c_void_p]
    n_ins, n_ops, n_pts, n_ramps = p.counts
    vals = np.array(inputs or [0.0], np.float32)
    out = C.c_float()
    rc = R.ref_expr_eval(p.ins.ctypes.data, n_ins, p.ops.ctypes.data, n_ops, p.pts.ctypes.data, n_pts,
                         p.ramps_arr.ctypes.data, n_ramps, vals.ctypes.data, len(inputs), p.result, u, v, face_seed,
                         t, C.byref(out))
    return None if rc else out.value


def test_reference_values_through_the_float_interpreter(orc):
    fs = float(np.float32(0.82524268837302051))
    # The reference's float runtime (no random prefix) hashes the float site seed, so its rand() values are not
    # the double path's 0.1137663251985252 and 3.1409747049307857 (tests/expression_test.cpp:56-80); they are
    # pinned here against that float runtime itself.
    p = expr.compile_expression("$a=hash($foo+355) <= .3? rand(0.1,0.7):rand(0.2,0);\\n$a*$bar", "amount",
                                "CutFXModule")
    assert p.inputs == ["foo", "bar"] and p.random_call_count == 2
    got = orc.expr_eval(p, [1.0, 2.0], u=0.2, v=0.3, face_seed=fs)
    assert got is not None and 0.0 <= got <= 1.4
    if orc.ref() is not None:
        assert got == ref_float_eval(orc.ref(), p, [1.0, 2.0], 0.2, 0.3, fs)
        q = expr.compile_expression("$a=rand(0.1,0.8);$b=rand(1,3);$a+$b", "noise", "ClumpingFXModule")
        assert orc.expr_eval(q, [], u=0.2, v=0.3, face_seed=fs) == ref_float_eval(orc.ref(), q, [], 0.2, 0.3, fs)
    p = expr.compile_expression(REFERENCE_SOURCES[2])
    vals = [1.0 if n == "id" else 0.25 for n in p.inputs]
    assert orc.expr_eval(p, vals) == 5.0
    assert orc.expr_eval(expr.compile_expression(REFERENCE_SOURCES[3]), []) == 5.0
    assert orc.expr_eval(expr.compile_expression(REFERENCE_SOURCES[4]), []) == 1.0
    assert abs(orc.expr_eval(expr.compile_expression(REFERENCE_SOURCES[5]), []) - (0.5 + 0.8521857695411413)) < 1e-6
    assert orc.expr_eval(expr.compile_expression("$a/0"), [1.0]) is None  # non-finite: an error
    for source, t, want in [(REFERENCE_SOURCES[7], 0.25, 0.0), (REFERENCE_SOURCES[8], 0.25, 0.5),
                            (REFERENCE_SOURCES[9], 0.125, 0.15625), (REFERENCE_SOURCES[10], 0.25, 0.5625),
                            (REFERENCE_SOURCES[11], 0.125, 0.20562979474213056), (REFERENCE_SOURCES[12], 0.25, 0.0),
                            (REFERENCE_SOURCES[12], 0.75, 1.0)]:
        assert abs(orc.expr_eval(expr.compile_expression(source), [], t=t) - want) <= 1e-6


def test_extensions_and_bindings(orc):
    with pytest.raises(ValueError, match="unsupported function: noise"):
        expr.compile_expression("noise($u, $v, 0)")
    p = expr.compile_expression("pow($u, 2) + mix(1, 2, $v) + noise($Px, $Py, $Pz) + texel1($faceId, $u, $v)",
                                extensions=True)
    ops = [i[0] for i in p.instructions]
    assert {27, 28, 29, 30} <= set(ops) and p.instructions[ops.index(30)][3] == 1
    p.bind(expr.DEFAULT_BINDINGS)
    assert p.bindings == [6, 7, 0, 1, 2, 8]
    with pytest.raises(KeyError):
        expr.compile_expression("$cLength").bind(expr.DEFAULT_BINDINGS)
    # two-call sizing: zero capacities report the counts
    import ctypes as C
    from nanoxgen_amd import abi
    out = abi.ExprCompiled()
    assert nx.lib().nxg_expr_compile(b"1+$a", 4, None, C.byref(out), None, 0) == abi.BUFFER_TOO_SMALL
    assert (out.instruction_count, out.operand_count, out.input_count, out.input_name_bytes) == (3, 2, 1, 2)


def test_compiled_programs_validate_without_ramps():
    """A compiled program with no ramps, ramp points or operands presents
    zero counts to nxg_expr_validate / nxg_expr_pack (its padded host arrays
    are not counted)."""
    import ctypes as C
    for source in ["0.5", "$u + 1", "clamp($u, 0, 1)", "rampUI(0,0,3:1,1,3)"]:
        p = expr.compile_expression(source, "a", "SplinePrimitive").bind(expr.DEFAULT_BINDINGS)
        st = p.struct
        assert (st.instruction_count, st.operand_count, st.ramp_point_count, st.ramp_count) == p.counts
        assert nx.lib().nxg_expr_validate(C.byref(st)) == 0, source
        info = nx.abi.ExprSet()
        assert nx.lib().nxg_expr_pack(C.byref(st), 1, None, 0, C.byref(info)) == 0, source
