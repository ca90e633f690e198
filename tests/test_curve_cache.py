"""The .nxc curve cache (nxg_curve_cache_*): bytes identical to the
reference's build_curve_cache, its validator's messages, and the reference's
own test (tests/test_main.cpp:617-658).  CPU only."""
import numpy as np
import pytest

from nanoxgen_amd import abi
from nanoxgen_amd.cache import build_curve_cache, header_of, validate_curve_cache


def exact_case():
    # tests/test_main.cpp:617-634
    counts = np.uint32([2, 3])
    points = np.float32([[0, 0, 0, 0.01], [0, 1, 0, 0.005], [1, 0, 0, 0.02], [1, 1, 0, 0.01], [1, 2, 0, 0.002]])
    tex = np.float32([[0, 0], [0, 1], [1, 0], [1, 0.5], [1, 1]])
    patch = np.float32([[0.2, 0.3], [0.7, 0.8]])
    face = np.float32([[0.1, 0.2], [0.6, 0.7]])
    ids = np.uint32([3, 9])
    times = np.float32([0.5])
    motion = points[:, :3] + np.float32([0.25, 0, 0])
    return dict(counts=counts, points=points, texcoords=tex, patch_uvs=patch, face_uvs=face, face_ids=ids,
                motion_times=times, motion_points=motion)


def build(case):
    c = dict(case)
    return build_curve_cache(c.pop("counts"), c.pop("points"), **c)


def test_reference_exact_cache_pin():
    case = exact_case()
    blob = build(case)
    assert validate_curve_cache(blob) == ""
    h = header_of(blob)
    assert h.magic == 0x454843414347584E and h.strand_count == 2 and h.point_count == 5
    assert h.motion_sample_count == 1 and h.flags == 31
    pts = blob[h.points_offset:h.points_offset + 80].view(np.float32).reshape(5, 4)
    assert np.array_equal(pts.view(np.uint32), case["points"].view(np.uint32))
    mp = blob[h.motion_points_offset:h.motion_points_offset + 60].view(np.float32).reshape(5, 3)
    assert np.array_equal(mp, case["motion_points"])
    corrupt = abi.aligned_bytes(blob.nbytes)
    corrupt[:] = blob
    corrupt[-1] ^= 1
    assert validate_curve_cache(corrupt) == "curve-cache content hash mismatch"


def random_case(rng, strands, optional, motion):
    counts = rng.integers(2, 9, size=strands).astype(np.uint32)
    p = int(counts.sum())
    pts = rng.normal(size=(p, 4)).astype(np.float32)
    pts[:, 3] = np.abs(pts[:, 3])
    case = dict(counts=counts, points=pts)
    if "tex" in optional:
        case["texcoords"] = rng.uniform(size=(p, 2)).astype(np.float32)
    if "patch" in optional:
        case["patch_uvs"] = rng.uniform(size=(strands, 2)).astype(np.float32)
    if "face" in optional:
        case["face_uvs"] = rng.uniform(size=(strands, 2)).astype(np.float32)
    if "ids" in optional:
        case["face_ids"] = rng.integers(0, 1000, size=strands).astype(np.uint32)
    if motion:
        case["motion_times"] = np.sort(rng.uniform(size=motion)).astype(np.float32)
        case["motion_points"] = rng.normal(size=(motion * p, 3)).astype(np.float32)
    return case


@pytest.mark.parametrize("optional,motion", [((), 0), (("tex",), 0), (("patch", "ids"), 2),
                                             (("tex", "patch", "face", "ids"), 3), (("face",), 1)])
def test_bytes_match_reference(orc, optional, motion):
    if orc.ref() is None:
        pytest.skip("reference core not built here")
    rng = np.random.default_rng(len(optional) * 10 + motion)
    for strands in (1, 7, 333):
        case = random_case(rng, strands, optional, motion)
        want, err = orc.ref_build_curve_cache(**case)
        assert err == ""
        assert np.array_equal(build(case), want)


def test_build_errors_match_reference(orc):
    if orc.ref() is None:
        pytest.skip("reference core not built here")
    rng = np.random.default_rng(4)
    base = random_case(rng, 20, ("tex",), 2)
    bad = []
    c = dict(base); c["counts"] = base["counts"].copy(); c["counts"][3] = 1; bad.append(c)
    c = dict(base); c["counts"] = base["counts"].copy(); c["counts"][3] += 1; bad.append(c)
    c = dict(base); c["motion_times"] = np.float32([0.5, 0.5]); bad.append(c)
    c = dict(base); c["motion_times"] = np.float32([0.5, np.inf]); bad.append(c)
    c = dict(base); c["points"] = base["points"].copy(); c["points"][4, 1] = np.nan; bad.append(c)
    c = dict(base); c["points"] = base["points"].copy(); c["points"][4, 3] = -1.0; bad.append(c)
    for case in bad:
        _, want = orc.ref_build_curve_cache(**case)
        assert want
        with pytest.raises(ValueError) as e:
            build(case)
        assert str(e.value) == want


def test_validator_messages_match_reference(orc):
    if orc.ref() is None:
        pytest.skip("reference core not built here")
    blob = build(random_case(np.random.default_rng(5), 50, ("tex", "patch"), 1))
    h = header_of(blob)

    def mutated(fn):
        b = abi.aligned_bytes(blob.nbytes)
        b[:] = blob
        fn(b)
        return b

    cases = [
        mutated(lambda b: b.__setitem__(slice(0, 1), 0)),                           # magic
        mutated(lambda b: b.__setitem__(slice(8, 9), 7)),                           # version
        mutated(lambda b: b.__setitem__(slice(12, 13), 64)),                        # unknown flag
        mutated(lambda b: b.__setitem__(slice(h.points_offset + 4, h.points_offset + 5), 0xFF)),  # hash
        mutated(lambda b: b.__setitem__(slice(32, 36), np.uint32([0]).view(np.uint8))),  # empty
    ]
    for b in cases:
        assert validate_curve_cache(b) == orc.ref_validate_curve_cache(b) != ""
    assert validate_curve_cache(blob) == orc.ref_validate_curve_cache(blob) == ""
    short = abi.aligned_bytes(100)
    assert validate_curve_cache(short) == "blob is smaller than CurveCacheHeader"


def test_layout_from_counts():
    import ctypes as C
    import nanoxgen_amd as nx
    inp = abi.CurveCacheInput(point_counts=1, points=1, texcoords=1, strand_count=3, point_count=10,
                              motion_sample_count=2, motion_times=1, motion_points=1)
    h = abi.CurveCacheHeader()
    assert nx.lib().nxg_curve_cache_layout(C.byref(inp), C.byref(h)) == abi.OK
    assert h.point_counts_offset == 128 and h.points_offset == 144
    assert h.motion_times_offset == 144 + 160 and h.motion_points_offset == 304 + 16
    assert h.texcoords_offset == 320 + 240 and h.byte_size == 560 + 80 and h.flags == 1 | 16
    empty = abi.CurveCacheInput(strand_count=0, point_count=0)
    assert nx.lib().nxg_curve_cache_layout(C.byref(empty), C.byref(h)) == abi.INVALID_BUILD_INPUT
