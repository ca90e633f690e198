"""Pins of the Classic CutFX oracle (oracle/oracle_classic.c): the reference's
own test values (tests/classic_test.cpp), fixtures written by the
reference's apply_xgen_classic_float_runtime_plan_cpu
(tests/golden/reference_classic.npz), the reference core itself where it is
built here, and closed forms of the uniform cubic B-spline.  CPU only."""
import os

import numpy as np
import pytest

from nanoxgen_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
CLASSIC = os.path.join(HERE, "golden", "reference_classic.npz")
# (strand source, amount_scale, amount_offset, radius_scale), as written by make_golden.py
CASES = [("ico5", 0.25, 0.0, 1.0), ("ico5", 0.0, 0.3, 1.0), ("ico13", 0.5, 0.05, 1.0),
         ("walk", 0.4, 0.1, 2.0), ("walk", 1.0, 0.0, 1.0), ("walk", 0.0, 0.0, 1.0)]


@pytest.fixture(scope="module")
def classic():
    with np.load(CLASSIC, allow_pickle=False) as d:
        return {k: d[k] for k in d.keys()}


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def straight(n, length, radius=0.1):
    p = np.zeros((n, 4), np.float32)
    p[:, 0] = np.linspace(0.0, length, n, dtype=np.float32)
    p[:, 3] = radius
    return p


# ---- reference test values (tests/classic_test.cpp) ----------------------

def test_spline_length_pin(orc):
    # test_float_runtime_length_and_cut: cLength of this 4-CV strand
    p = np.array([[0, 0, 0, 0], [1, 1, 0, 0], [2, -1, 0, 0], [3, 0, 0, 0]], np.float32)
    assert abs(orc.curve_spline_length(p) - 3.796864797465667) < 2e-6


def test_cut_quarter_length_pin(orc):
    # the same test: a 3-CV strand of length 4, cut by 0.25 * cLength -> x = 0, 1.5, 3
    r = orc.classic_cut(straight(3, 4.0), 3, 0.25, 0.0)
    assert r["strand_count"] == 1
    assert np.allclose(r["points"][:, 0], [0.0, 1.5, 3.0], atol=1e-5)
    assert np.all(r["points"][:, 1:3] == 0.0)


def test_cut_everything_culls_pin(orc):
    # test_float_runtime_cut_culling: cut amount 10 on a strand of length 1
    r = orc.classic_cut(straight(3, 1.0), 3, 0.0, 10.0)
    assert r["strand_count"] == 0 and r["keep"][0] == 0


# ---- reference fixtures --------------------------------------------------

@pytest.mark.parametrize("i", range(len(CASES)))
def test_oracle_matches_reference_classic(orc, classic, i):
    src, scale, offset, rscale = CASES[i]
    assert np.allclose(classic[f"case{i}_args"], [scale, offset, rscale])
    cvs = int(classic[f"{src}_cvs"])
    r = orc.classic_cut(classic[f"{src}_points"], cvs, scale, offset, roots=classic[f"{src}_roots"],
                        uvs=classic[f"{src}_uvs"], radius_scale=rscale)
    assert r["strand_count"] == len(classic[f"case{i}_out_point_counts"])
    assert np.array_equal(bits(r["points"]), bits(classic[f"case{i}_out_points"]))
    assert np.array_equal(r["point_counts"], classic[f"case{i}_out_point_counts"])
    assert np.array_equal(bits(r["roots"]), bits(classic[f"case{i}_out_roots"]))
    assert np.array_equal(bits(r["root_uvs"]), bits(classic[f"case{i}_out_root_uvs"]))


def test_fixture_cases_cover_culling(classic):
    kept = [len(classic[f"case{i}_out_point_counts"]) for i in range(len(CASES))]
    walk = len(classic["walk_points"]) // int(classic["walk_cvs"])
    assert kept[4] == 0                      # cut by the whole length: every strand goes
    assert 0 < kept[5] < walk                # no cut: only zero-width and too-short strands go
    assert 0 < kept[1] < 300                 # a fixed amount culls the short strands only


def test_oracle_matches_live_reference(orc):
    if orc.ref() is None:
        pytest.skip("reference core not built here")
    rng = np.random.default_rng(5)
    for cvs in (2, 4, 9, 32):
        n = 64
        steps = rng.normal(0, 1, size=(n, cvs, 3)) * rng.uniform(0, 0.4, size=(n, 1, 1))
        pts = np.cumsum(steps, axis=1).astype(np.float32)
        rad = rng.uniform(0.0, 0.05, size=(n, cvs, 1)).astype(np.float32)
        packed = np.concatenate([pts, rad], axis=2).reshape(-1, 4)
        for scale, offset in ((0.3, 0.0), (0.0, 0.2), (0.9, 0.01)):
            a = orc.classic_cut(packed, cvs, scale, offset)
            b = orc.ref_classic_cut(packed, cvs, scale, offset)
            assert a["strand_count"] == b["strand_count"]
            assert np.array_equal(bits(a["points"]), bits(b["points"]))
    # a non-finite amount is an error in both
    assert orc.classic_cut(straight(4, 1.0), 4, 0.0, float("inf")) is None
    assert orc.ref_classic_cut(straight(4, 1.0), 4, 0.0, float("inf")) is None


# ---- closed forms of the uniform cubic B-spline --------------------------

def test_spline_at_knots_is_the_one_four_one_average(orc):
    rng = np.random.default_rng(1)
    for n in (3, 5, 8):
        p = rng.normal(size=(n, 4)).astype(np.float32)
        out = np.zeros(3, np.float32)
        for i in range(1, n - 1):
            orc.lib().orc_curve_eval(p.ctypes.data, n, np.float32(i / (n - 1)), out.ctypes.data)
            want = (p[i - 1, :3].astype(np.float64) + 4 * p[i, :3] + p[i + 1, :3]) / 6.0
            assert np.allclose(out, want, atol=1e-5)
        orc.lib().orc_curve_eval(p.ctypes.data, n, 0.0, out.ctypes.data)
        assert np.array_equal(out, p[0, :3])
        orc.lib().orc_curve_eval(p.ctypes.data, n, 1.0, out.ctypes.data)
        assert np.array_equal(out, p[-1, :3])


@pytest.mark.parametrize("n", [2, 3, 6, 17])
def test_straight_strand_length_and_cut(orc, n):
    # evenly spaced collinear CVs: the spline is the segment, linearly parametrised
    L = 3.0
    assert abs(orc.curve_spline_length(straight(n, L)) - L) < 1e-5
    for a in (0.0, 0.5, 2.0):
        r = orc.classic_cut(straight(n, L), n, 0.0, a)
        want = np.arange(n) * (L - a) / (n - 1)
        assert np.allclose(r["points"][:, 0], want, atol=2e-5)
        assert np.all(r["points"][:, 3] == np.float32(0.1))


def test_cut_removes_the_amount_from_the_tip(orc):
    # the new tip lies on the old spline, `amount` of arc length below the old tip
    rng = np.random.default_rng(2)
    n = 12
    pts = np.cumsum(rng.normal(0, 0.3, size=(n, 3)), axis=0).astype(np.float32)
    packed = np.concatenate([pts, np.full((n, 1), 0.1, np.float32)], axis=1)
    L = orc.curve_spline_length(packed)
    ts = np.linspace(0.0, 1.0, 20001, dtype=np.float32)
    dense = np.zeros((len(ts), 3), np.float32)
    for j, t in enumerate(ts):
        orc.lib().orc_curve_eval(packed.ctypes.data, n, t, dense[j].ctypes.data)
    arc = np.concatenate([[0.0], np.cumsum(np.linalg.norm(np.diff(dense.astype(np.float64), axis=0), axis=1))])
    for frac in (0.1, 0.4, 0.7):
        r = orc.classic_cut(packed, n, frac, 0.0)
        assert np.array_equal(r["points"][0, :3], pts[0])          # the root stays
        tip = np.argmin(np.linalg.norm(dense - r["points"][-1, :3], axis=1))
        assert np.linalg.norm(dense[tip] - r["points"][-1, :3]) < 1e-3
        assert abs((arc[-1] - arc[tip]) - frac * L) < 0.03 * L     # chord-sampling error only


def test_multiplier_scales_the_amount(orc, classic):
    pts, cvs = classic["ico5_points"], int(classic["ico5_cvs"])
    n = len(pts) // cvs
    a = orc.classic_cut(pts, cvs, 0.25, 0.0)
    b = orc.classic_cut(pts, cvs, 0.25, 0.0, multiplier=np.ones(n, np.float32))
    assert np.array_equal(bits(a["points"]), bits(b["points"]))
    half = orc.classic_cut(pts, cvs, 0.5, 0.0, multiplier=np.full(n, 0.5, np.float32))
    assert np.array_equal(bits(a["points"]), bits(half["points"]))  # 0.5 * (0.5 x) == 0.25 x exactly
    zero = orc.classic_cut(pts, cvs, 0.25, 0.0, multiplier=np.zeros(n, np.float32))
    none = orc.classic_cut(pts, cvs, 0.0, 0.0)
    assert np.array_equal(bits(zero["points"]), bits(none["points"]))


def test_compaction_is_stable(orc, classic):
    r = orc.classic_cut(classic["walk_points"], int(classic["walk_cvs"]), 0.0, 0.0,
                        roots=classic["walk_roots"], uvs=classic["walk_uvs"])
    kept = np.flatnonzero(r["keep"])
    assert np.array_equal(r["roots"]["triangle_index"], classic["walk_roots"]["triangle_index"][kept])
    assert np.array_equal(r["root_uvs"], classic["walk_uvs"][kept])
    assert abi.ROOT_DTYPE.itemsize == 48


# ---- renderer endpoints (src/xgen_classic_runtime.cpp:1921-1956) ---------

def test_endpoints_pin(orc):
    # tests/classic_test.cpp:243-249: CVs at x = 0, 1.5, 3 gain -1.5 and 4.5,
    # with the neighbouring radii
    p = straight(3, 3.0)
    p[:, 3] = [0.125, 0.075, 0.025]
    e = orc.classic_endpoints(p, 3)
    assert e.shape == (5, 4)
    assert abs(e[0, 0] + 1.5) < 1e-6 and abs(e[4, 0] - 4.5) < 1e-6
    assert e[0, 3] == e[1, 3] and e[4, 3] == e[3, 3]
    assert np.array_equal(e[1:4], p)


def test_endpoints_match_reference(orc, classic):
    cvs = int(classic["walk_cvs"])
    got = orc.classic_endpoints(classic["walk_points"], cvs)
    assert np.array_equal(bits(got), bits(classic["walk_endpoints"]))
    got = orc.classic_endpoints(classic["case3_out_points"], cvs)
    assert np.array_equal(bits(got), bits(classic["case3_endpoints"]))
    if orc.ref() is not None:
        rng = np.random.default_rng(9)
        for n in (2, 5, 64):
            p = rng.normal(size=(33 * n, 4)).astype(np.float32)
            assert np.array_equal(bits(orc.classic_endpoints(p, n)), bits(orc.ref_classic_endpoints(p, n)))
