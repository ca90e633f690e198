"""The C ABI on the CPU: symbol exports, the host asset builder/validator
against the oracle, and the checked device contract (validation only; no
kernel is launched).  CPU only."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, aligned_copy
import nanoxgen_amd as nx
from nanoxgen_amd import abi, batch, workloads


def header_symbols():
    text = open(os.path.join(ROOT, "include", "nanoxgen_amd.h")).read()
    return sorted(set(re.findall(r"^(?:const )?\w+\s*\**\s*(nxg_\w+)\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = nx.lib()
    declared = header_symbols()
    assert len(declared) >= 10
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(n for n, _, _ in nx.SYMBOLS) == declared


def test_build_info_names_gfx950():
    assert b"gfx950" in nx.lib().nxg_build_info()


@pytest.mark.parametrize("make", [
    lambda: workloads.icosphere_description(),
    lambda: workloads.icosphere_description(stencil_k=0),
    lambda: workloads.icosphere_description(stencil_k=1, guides=5),
    lambda: workloads.plane_description(quads=32, guides_per_side=8),
    lambda: workloads.plane_description(quads=8, guides_per_side=2, stencil_k=3),
])
def test_builder_matches_oracle_bytes(orc, make):
    b = make()
    assert np.array_equal(nx.build_asset(b), orc.build_asset(b))


def test_builder_without_normals_or_uvs(orc):
    b = workloads.icosphere_description()
    b2 = abi.BuildInputArrays(b.positions, b.triangles, b.guide_cvs, b.guide_cv_counts, stencil_k=6)
    blob = nx.build_asset(b2)
    assert np.array_equal(blob, orc.build_asset(b2))
    h = abi.header_of(blob)
    assert h.texcoords_offset == 0 and h.flags == abi.HAS_NORMALS


@pytest.mark.parametrize("mutate,message", [
    (lambda a: a.update(triangles=np.array([[0, 1, 9]])), "triangle index out of bounds"),
    (lambda a: a.update(guide_cv_counts=np.array([1]), guide_cvs=np.zeros((1, 3))), "each guide needs 2..65535 CVs"),
    (lambda a: a.update(stencil_k=9), "stencil_k must be in 0..8"),
    (lambda a: a.update(positions=np.zeros((3, 3))), "mesh has zero total area"),
])
def test_builder_rejections_match_oracle(orc, mutate, message):
    args = dict(positions=np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1]]), triangles=np.array([[0, 1, 2]]),
                guide_cvs=np.array([[0, 0, 0], [0, 1, 0]]), guide_cv_counts=np.array([2]))
    mutate(args)
    b = abi.BuildInputArrays(**args)
    with pytest.raises(ValueError, match=re.escape(message)):
        nx.build_asset(b)
    with pytest.raises(ValueError, match=re.escape(message)):
        orc.build_asset(b)


def test_validator_messages_match_oracle(orc):
    blob = nx.build_asset(workloads.icosphere_description())
    assert nx.validate_asset(blob) == "" == orc.validate_asset(blob)
    cases = []
    c = aligned_copy(blob); c[-1] ^= 1; cases.append(c)                         # hash
    c = aligned_copy(blob); c[0] ^= 1; cases.append(c)                          # magic
    c = aligned_copy(blob); c[8] = 7; cases.append(c)                           # version
    cases.append(aligned_copy(blob[:100]))                                      # short
    cases.append(aligned_copy(blob[:-4]))                                       # size
    raw = np.zeros(blob.nbytes + 64, dtype=np.uint8)
    off = (-raw.ctypes.data) % 64 + 1
    mis = raw[off:off + blob.nbytes]; mis[:] = blob; cases.append(mis)         # misaligned
    expected = ["content hash mismatch", "bad NanoXGen magic", "unsupported NanoXGen asset version",
                "blob is smaller than AssetHeader", "header byte_size does not match file size",
                "blob base is not 64-byte aligned"]
    for case, msg in zip(cases, expected):
        assert nx.validate_asset(case) == msg == orc.validate_asset(case)


# ---- checked device contract (tests/test_main.cpp:483-588) ---------------

class Request:
    """Host stand-ins for device pointers: validation never dereferences them."""

    def __init__(self):
        self.blob = nx.build_asset(workloads.icosphere_description())
        self.desc = abi.DeviceAsset()
        assert nx.lib().nxg_make_device_asset(self.blob.ctypes.data, self.blob.nbytes, self.blob.ctypes.data,
                                              self.blob.nbytes, C.byref(self.desc)) == 0
        self.params = abi.GenerationParams(strand_count=7, cvs_per_strand=5)
        n = 7 * 5
        self.points = np.zeros((n, 4), np.float32)
        self.roots = np.zeros(7, abi.ROOT_DTYPE)
        self.uvs = np.zeros((7, 2), np.float32)
        self.counts = np.zeros(7, np.uint32)
        self.out = abi.PackedOutput(points=self.points.ctypes.data, roots=self.roots.ctypes.data,
                                    root_uvs=self.uvs.ctypes.data, radius_scale=1.0,
                                    point_counts=self.counts.ctypes.data, point_capacity=n,
                                    root_capacity=7, root_uv_capacity=7, point_count_capacity=7)

    def validate(self, deformed=None, config=None, desc=None, params=None, out=None):
        return nx.lib().nxg_validate_packed_generation(
            C.byref(desc or self.desc), C.byref(deformed) if deformed else None,
            C.byref(params or self.params), C.byref(out or self.out), C.byref(config) if config else None)


def test_contract_valid_request():
    assert Request().validate() == abi.OK


def test_contract_error_codes():
    r = Request()
    out = abi.PackedOutput.from_buffer_copy(r.out)
    out.point_capacity = 34
    assert r.validate(out=out) == abi.POINT_CAPACITY_TOO_SMALL
    d = abi.DeformedGeometry(positions=r.blob.ctypes.data, position_count=r.desc.header.vertex_count - 1)
    assert r.validate(deformed=d) == abi.DEFORMED_POSITION_COUNT_MISMATCH
    assert r.validate(config=abi.LaunchConfig(block_size=0)) == abi.INVALID_BLOCK_SIZE
    assert r.validate(config=abi.LaunchConfig(block_size=1025)) == abi.INVALID_BLOCK_SIZE
    out = abi.PackedOutput.from_buffer_copy(r.out)
    out.radius_scale = float("nan")
    assert r.validate(out=out) == abi.INVALID_RADIUS_SCALE
    short = abi.DeviceAsset.from_buffer_copy(r.desc)
    short.byte_capacity = r.desc.header.byte_size - 1
    assert r.validate(desc=short) == abi.ASSET_CAPACITY_TOO_SMALL
    mis = abi.DeviceAsset.from_buffer_copy(r.desc)
    mis.data = r.desc.data + 1
    assert r.validate(desc=mis) == abi.MISALIGNED_ASSET
    bad = abi.DeviceAsset.from_buffer_copy(r.desc)
    bad.header.noise_gradients_offset = bad.header.byte_size - 4
    assert r.validate(desc=bad) == abi.INVALID_ASSET_METADATA
    null = abi.DeviceAsset.from_buffer_copy(r.desc)
    null.data = None
    assert r.validate(desc=null) == abi.NULL_ASSET
    assert r.validate(params=r.params.copy(cvs_per_strand=1)) == abi.INVALID_GENERATION_PARAMETERS
    assert r.validate(params=r.params.copy(noise_mask=1.5)) == abi.INVALID_GENERATION_PARAMETERS
    out = abi.PackedOutput.from_buffer_copy(r.out)
    out.roots = None
    assert r.validate(out=out) == abi.ROOT_CAPACITY_TOO_SMALL  # null pointer with capacity 7
    out.root_capacity = 0
    assert r.validate(out=out) == abi.OK
    out.points = None
    assert r.validate(out=out) == abi.MISSING_OUTPUT
    out = abi.PackedOutput.from_buffer_copy(r.out)
    out.point_count_capacity = 6
    assert r.validate(out=out) == abi.POINT_COUNT_CAPACITY_TOO_SMALL
    out = abi.PackedOutput.from_buffer_copy(r.out)
    out.root_uv_capacity = 6
    assert r.validate(out=out) == abi.ROOT_UV_CAPACITY_TOO_SMALL


def test_generate_rejects_before_any_device_work():
    """Every refusal happens on the host, before a HIP call."""
    r = Request()
    L = nx.lib()
    ws = np.zeros(16, np.uint8)
    st = L.nxg_generate_packed(C.byref(r.desc), None, C.byref(r.params), None, C.byref(r.out), None, None,
                               ws.ctypes.data, ws.nbytes, None)
    assert st == abi.WORKSPACE_TOO_SMALL
    oct5 = abi.NoiseOctaves(count=5)
    st = L.nxg_generate_packed(C.byref(r.desc), None, C.byref(r.params), C.byref(oct5), C.byref(r.out), None,
                               None, ws.ctypes.data, ws.nbytes, None)
    assert st == abi.INVALID_OCTAVES
    sten = abi.StencilOutput(indices=ws.ctypes.data, weights=ws.ctypes.data, capacity=8)
    st = L.nxg_generate_packed(C.byref(r.desc), None, C.byref(r.params), None, C.byref(r.out), C.byref(sten),
                               None, ws.ctypes.data, ws.nbytes, None)
    assert st == abi.STENCIL_CAPACITY_TOO_SMALL
    st = L.nxg_generate_packed_from_roots(C.byref(r.desc), C.byref(r.params), None, None, C.byref(r.out), None,
                                          ws.ctypes.data, ws.nbytes, None)
    assert st == abi.INVALID_ROOTS
    assert L.nxg_packed_generation_workspace_size(C.byref(r.desc.header), C.byref(r.params)) > 16
    # the no-alloc entry point: the contract first, then the static scratch limit
    out = abi.PackedOutput.from_buffer_copy(r.out)
    out.point_capacity = 34
    st = L.nxg_generate_packed_noalloc(C.byref(r.desc), None, C.byref(r.params), None, C.byref(out), None, None,
                                       None)
    assert st == abi.POINT_CAPACITY_TOO_SMALL
    huge = r.params.copy(strand_count=20_000_000)
    out = abi.PackedOutput.from_buffer_copy(r.out)
    out.point_capacity = huge.strand_count * huge.cvs_per_strand
    out.root_capacity = out.root_uv_capacity = out.point_count_capacity = huge.strand_count
    assert L.nxg_packed_generation_workspace_size(C.byref(r.desc.header), C.byref(huge)) > abi.STATIC_SCRATCH_BYTES
    st = L.nxg_generate_packed_noalloc(C.byref(r.desc), None, C.byref(huge), None, C.byref(out), None, None, None)
    assert st == abi.SCRATCH_LIMIT
    with open(os.path.join(ROOT, "include", "nanoxgen_amd.h")) as f:
        m = re.search(r"#define NXG_STATIC_SCRATCH_BYTES \((\d+)ull << (\d+)\)", f.read())
    assert m and int(m.group(1)) << int(m.group(2)) == abi.STATIC_SCRATCH_BYTES


def test_cdf_sampler_rejects_before_any_device_work():
    """The inverse-CDF entry points refuse a missing, misaligned or short
    sampler buffer, and output without roots, on the host."""
    r = Request()
    L = nx.lib()
    need = L.nxg_cdf_sampler_size(C.byref(r.desc.header))
    t = r.desc.header.triangle_count
    assert need >= 8 * t and need % 256 == 0
    assert L.nxg_cdf_sampler_size(None) == 0
    buf = np.zeros(need + 512, np.uint8)
    aligned = buf.ctypes.data + (-buf.ctypes.data) % 256
    assert L.nxg_cdf_sampler_build(C.byref(r.desc), None, need, None) == abi.INVALID_SAMPLER
    assert L.nxg_cdf_sampler_build(C.byref(r.desc), aligned, need - 1, None) == abi.INVALID_SAMPLER
    assert L.nxg_cdf_sampler_build(C.byref(r.desc), aligned + 64, need, None) == abi.INVALID_SAMPLER
    null = abi.DeviceAsset.from_buffer_copy(r.desc)
    null.data = None
    assert L.nxg_cdf_sampler_build(C.byref(null), aligned, need, None) == abi.NULL_ASSET
    ws = np.zeros(16, np.uint8)
    gen = lambda out, sampler, nbytes: L.nxg_generate_packed_cdf(
        C.byref(r.desc), None, C.byref(r.params), None, sampler, nbytes, C.byref(out), None, None,
        ws.ctypes.data, ws.nbytes, None)
    assert gen(r.out, None, need) == abi.INVALID_SAMPLER
    assert gen(r.out, aligned, need - 8) == abi.INVALID_SAMPLER
    out = abi.PackedOutput.from_buffer_copy(r.out)
    out.roots, out.root_capacity = None, 0
    assert gen(out, aligned, need) == abi.ROOT_CAPACITY_TOO_SMALL
    assert gen(r.out, aligned, need) == abi.WORKSPACE_TOO_SMALL


def test_status_messages():
    L = nx.lib()
    assert L.nxg_status_message(0) == b""
    assert L.nxg_status_message(6) == b"GPU block size must be in [1, 1024]"
    assert L.nxg_status_message(20) == b"device output size overflows uint64"
    assert L.nxg_status_message(999) == b"unknown device generation error"
    for s in range(1, len(abi.STATUS)):
        assert L.nxg_status_message(s) and L.nxg_status_message(s) != b"unknown device generation error"


# ---- batched descriptions (host-side validation, no device work) ---------

def _batch_items(r, count, ws_bytes=None):
    ws = np.zeros(1 << 16, np.uint8)
    r.keep = [ws]
    items = (abi.BatchItem * count)()
    for it in items:
        it.asset = C.addressof(r.desc)
        it.params = C.addressof(r.params)
        it.output = C.addressof(r.out)
        it.workspace = ws.ctypes.data
        it.workspace_bytes = ws.nbytes if ws_bytes is None else ws_bytes
    return items


def test_batch_validation():
    r = Request()
    L = nx.lib()
    items = _batch_items(r, 3)
    assert L.nxg_batch_validate(items, 3, None) == abi.OK
    assert L.nxg_batch_validate(items, 0, None) == abi.INVALID_BATCH
    many = _batch_items(r, abi.BATCH_MAX_ITEMS + 1)
    assert L.nxg_batch_validate(many, abi.BATCH_MAX_ITEMS + 1, None) == abi.INVALID_BATCH
    assert L.nxg_batch_validate(many, abi.BATCH_MAX_ITEMS, None) == abi.OK
    for bad in [(3, 8, 1, 0, 0), (4, 16, 1, 0, 0), (1, 8, 0, 0, 0), (1, 8, 17, 0, 0), (1, 8, 1, 65, 0),
                (1, 8, 1, 0, 2)]:
        assert L.nxg_batch_validate(items, 3, C.byref(abi.BatchTuning(*bad))) == abi.INVALID_BATCH, bad
    for t in batch.tuning_candidates():
        assert L.nxg_batch_validate(items, 3, C.byref(t)) == abi.OK
    # one bad item refuses the whole batch with that item's status
    small = _batch_items(r, 3)
    small[1].workspace_bytes = 16
    assert L.nxg_batch_validate(small, 3, None) == abi.WORKSPACE_TOO_SMALL
    noparams = _batch_items(r, 2)
    noparams[1].params = None
    assert L.nxg_batch_validate(noparams, 2, None) == abi.NULL_ARGUMENT
    bad_oct = abi.NoiseOctaves(count=5)
    o = _batch_items(r, 2)
    o[0].octaves = C.addressof(bad_oct)
    assert L.nxg_batch_validate(o, 2, None) == abi.INVALID_OCTAVES
    # attributes to be evaluated need a valid expression set and capacity
    ex = _batch_items(r, 1)
    es = abi.ExprSet()
    ex[0].exprs = C.addressof(es)
    attrs = np.zeros((7, 4), np.float32)
    ex[0].attributes = attrs.ctypes.data
    ex[0].attribute_capacity = attrs.size
    assert L.nxg_batch_validate(ex, 1, None) == abi.INVALID_PROGRAM
    # generation and graph capture refuse before any device work
    counter = np.zeros(1, np.uint32)
    assert L.nxg_generate_batch(small, 3, None, counter.ctypes.data, None) == abi.WORKSPACE_TOO_SMALL
    assert L.nxg_generate_batch(items, 3, None, None, None) == abi.NULL_ARGUMENT
    h = C.c_void_p()
    assert L.nxg_batch_graph_create(small, 3, None, counter.ctypes.data, C.byref(h)) == abi.WORKSPACE_TOO_SMALL
    assert not h.value
    assert L.nxg_batch_graph_destroy(None) == abi.OK


def test_batch_default_tuning_is_a_candidate():
    d = batch.default_tuning()
    assert d.as_tuple() in [c.as_tuple() for c in batch.tuning_candidates()]
    assert len(batch.tuning_candidates()) == nx.lib().nxg_batch_tuning_candidates(None, 0) == 64


# ---- motion samples (device_contract.h:256-288; tests/test_main.cpp:504-527)

def test_motion_contract():
    r = Request()
    L = nx.lib()
    n = 7 * 5
    pts = np.zeros((2 * n, 3), np.float32)
    out = abi.MotionOutput(points=pts.ctypes.data, point_capacity=2 * n)
    samples = (abi.MotionSample * 2)(abi.MotionSample(time=0.25), abi.MotionSample(time=0.75))
    v = lambda s, c, o: L.nxg_validate_motion_generation(C.byref(r.desc), s, c, C.byref(r.params), C.byref(o), None)
    assert v(samples, 2, out) == abi.OK
    same = (abi.MotionSample * 2)(abi.MotionSample(time=0.25), abi.MotionSample(time=0.25))
    assert v(same, 2, out) == abi.INVALID_MOTION_TIMES
    nan = (abi.MotionSample * 1)(abi.MotionSample(time=float("nan")))
    assert v(nan, 1, out) == abi.INVALID_MOTION_TIMES
    assert v(None, 0, out) == abi.MISSING_MOTION_SAMPLES
    assert v(samples, 0, out) == abi.MISSING_MOTION_SAMPLES
    short = abi.MotionOutput(points=pts.ctypes.data, point_capacity=2 * n - 1)
    assert v(samples, 2, short) == abi.MOTION_POINT_CAPACITY_TOO_SMALL
    bad = (abi.MotionSample * 2)(abi.MotionSample(time=0.25), abi.MotionSample(time=0.75))
    bad[1].deformation.positions = r.blob.ctypes.data
    bad[1].deformation.position_count = r.desc.header.vertex_count - 1
    assert v(bad, 2, out) == abi.DEFORMED_POSITION_COUNT_MISMATCH
    # generation refuses on the host, before any device work
    ws = np.zeros(16, np.uint8)
    assert L.nxg_generate_motion(C.byref(r.desc), samples, 2, C.byref(r.params), None, C.byref(out), None,
                                 ws.ctypes.data, ws.nbytes, None) == abi.WORKSPACE_TOO_SMALL
    assert L.nxg_generate_motion(C.byref(r.desc), same, 2, C.byref(r.params), None, C.byref(out), None,
                                 ws.ctypes.data, ws.nbytes, None) == abi.INVALID_MOTION_TIMES
    one = L.nxg_packed_generation_workspace_size(C.byref(r.desc.header), C.byref(r.params))
    assert L.nxg_motion_generation_workspace_size(C.byref(r.desc.header), C.byref(r.params), 3) == 3 * one


def test_classic_cut_contract():
    L = nx.lib()
    n, cvs = 100, 6
    pin = np.zeros((n * cvs, 4), np.float32)
    pout = np.zeros((n * cvs, 4), np.float32)
    counts = np.zeros(n, np.uint32)
    roots_in = np.zeros(n, abi.ROOT_DTYPE)
    roots_out = np.zeros(n, abi.ROOT_DTYPE)
    words = np.zeros(64, np.uint32)
    curves = abi.ClassicCurves(points=pin.ctypes.data, strand_count=n, cvs_per_strand=cvs)
    cut = abi.ClassicCutParams(amount_scale=0.25, radius_scale=1.0)
    out = abi.ClassicOutput(points=pout.ctypes.data, point_capacity=n * cvs, kept_count=words.ctypes.data)
    v = lambda c=curves, k=cut, o=out: L.nxg_validate_classic_cut(C.byref(c), C.byref(k), C.byref(o))
    assert v() == abi.OK
    assert L.nxg_validate_classic_cut(None, C.byref(cut), C.byref(out)) == abi.NULL_ARGUMENT
    for bad_cvs in (0, 1, abi.CLASSIC_MAX_CVS + 1):
        assert v(abi.ClassicCurves(points=pin.ctypes.data, strand_count=n, cvs_per_strand=bad_cvs)) \
            == abi.INVALID_CLASSIC_REQUEST
    # zero strands need no buffers
    assert v(abi.ClassicCurves(strand_count=0, cvs_per_strand=cvs),
             o=abi.ClassicOutput(kept_count=words.ctypes.data)) == abi.OK
    short = abi.ClassicOutput(points=pout.ctypes.data, point_capacity=n * cvs - 1, kept_count=words.ctypes.data)
    assert v(o=short) == abi.POINT_CAPACITY_TOO_SMALL
    no_kept = abi.ClassicOutput(points=pout.ctypes.data, point_capacity=n * cvs)
    assert v(o=no_kept) == abi.NULL_ARGUMENT
    # outputs may not overlap inputs, and must be 16-byte aligned
    alias = abi.ClassicOutput(points=pin.ctypes.data + 16 * cvs, point_capacity=n * cvs,
                              kept_count=words.ctypes.data)
    assert v(o=alias) == abi.INVALID_CLASSIC_REQUEST
    odd = abi.ClassicOutput(points=pout.ctypes.data + 4, point_capacity=n * cvs, kept_count=words.ctypes.data)
    assert v(o=odd) == abi.INVALID_CLASSIC_REQUEST
    # roots out need roots in; capacities per buffer
    with_roots = abi.ClassicOutput(points=pout.ctypes.data, point_capacity=n * cvs, roots=roots_out.ctypes.data,
                                   root_capacity=n, kept_count=words.ctypes.data)
    assert v(o=with_roots) == abi.INVALID_CLASSIC_REQUEST
    rin = abi.ClassicCurves(points=pin.ctypes.data, roots=roots_in.ctypes.data, strand_count=n,
                            cvs_per_strand=cvs)
    assert v(rin, o=with_roots) == abi.OK
    with_roots.root_capacity = n - 1
    assert v(rin, o=with_roots) == abi.ROOT_CAPACITY_TOO_SMALL
    cnt = abi.ClassicOutput(points=pout.ctypes.data, point_capacity=n * cvs, point_counts=counts.ctypes.data,
                            point_count_capacity=n - 1, kept_count=words.ctypes.data)
    assert v(o=cnt) == abi.POINT_COUNT_CAPACITY_TOO_SMALL
    assert v(k=abi.ClassicCutParams(multiplier=pin.ctypes.data, multiplier_stride=0)) == abi.INVALID_CLASSIC_REQUEST
    # workspace: cut parameters, tile counts, group totals, status and the shared sample tables
    table = 2 * (2 * abi.CLASSIC_MAX_CVS + 8) * 32
    assert L.nxg_classic_cut_workspace_size(0) == 256 + (table + 255) // 256 * 256
    assert L.nxg_classic_cut_workspace_size(2_000_000) >= 2_000_000 * 4 + 31250 * 4
    # refused on the host before any device work
    ws = np.zeros(16, np.uint8)
    assert L.nxg_classic_cut_apply(C.byref(curves), C.byref(cut), C.byref(out), ws.ctypes.data, ws.nbytes,
                                   None) == abi.WORKSPACE_TOO_SMALL
    assert L.nxg_classic_cut_apply(C.byref(curves), C.byref(cut), C.byref(alias), ws.ctypes.data, ws.nbytes,
                                   None) == abi.INVALID_CLASSIC_REQUEST



def test_classic_endpoints_contract():
    L = nx.lib()
    n, cvs = 10, 4
    pin = np.zeros((n * cvs, 4), np.float32)
    pout = np.zeros((n * (cvs + 2), 4), np.float32)
    f = lambda i=pin.ctypes.data, c=cvs, o=pout.ctypes.data, cap=n * (cvs + 2), cnt=None, ccap=0, strands=n: \
        L.nxg_classic_add_endpoints(i, strands, c, None, o, cap, cnt, ccap, None)
    assert f(c=1) == abi.INVALID_CLASSIC_REQUEST
    assert f(c=abi.CLASSIC_MAX_CVS + 1) == abi.INVALID_CLASSIC_REQUEST
    assert f(strands=0) == abi.OK                      # nothing to do, nothing launched
    assert f(i=None) == abi.NULL_ARGUMENT
    assert f(cap=n * (cvs + 2) - 1) == abi.POINT_CAPACITY_TOO_SMALL
    counts = np.zeros(n, np.uint32)
    assert f(cnt=counts.ctypes.data, ccap=n - 1) == abi.POINT_COUNT_CAPACITY_TOO_SMALL
    assert f(o=pin.ctypes.data + 16, cap=10 ** 6) == abi.INVALID_CLASSIC_REQUEST  # overlaps the input
    assert f(o=pout.ctypes.data + 8) == abi.INVALID_CLASSIC_REQUEST               # misaligned
