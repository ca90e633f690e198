"""ORACLE (test infrastructure only).

ctypes access to the C restatement of the reference's strand-generation path
(liboracle.so, built from oracle/*.c by oracle/Makefile) and, where it was
built in this container, to the reference's own CPU core (_ref/libnxg_ref.so).
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this package, and only as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from nanoxgen_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None


def build(force=False):
    """Compile liboracle.so (and _ref/ when the reference sources are present)."""
    if force or not os.path.exists(os.path.join(HERE, "liboracle.so")):
        subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    elif force is None:
        pass
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = C.CDLL(path)
        L.orc_hash32.restype = C.c_uint32
        L.orc_hash32.argtypes = [C.c_uint32]
        L.orc_random01.restype = C.c_float
        L.orc_random01.argtypes = [C.c_uint32] * 3
        L.orc_noise_hash.restype = C.c_uint32
        L.orc_noise_hash.argtypes = [C.c_int32] * 3
        L.orc_fnv1a.restype = C.c_uint64
        L.orc_fnv1a.argtypes = [C.c_void_p, C.c_uint64]
        L.orc_gradient_noise_f32.restype = C.c_float
        L.orc_gradient_noise_f32.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float]
        L.orc_gradient_noise_f64.restype = C.c_double
        L.orc_gradient_noise_f64.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.orc_generate.restype = C.c_int
        L.orc_generate.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_build_asset.restype = C.c_int
        L.orc_build_asset.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64),
                                      C.c_char_p, C.c_uint64]
        L.orc_validate_asset.restype = C.c_int
        L.orc_validate_asset.argtypes = [C.c_void_p, C.c_uint64, C.c_char_p, C.c_uint64]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_strand_state_f32.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
        L.orc_strand_state_f64.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
        L.orc_seexpr_hash_finish.restype = C.c_uint32
        L.orc_seexpr_hash_finish.argtypes = [C.c_uint32]
        L.orc_runtime_hash_component.restype = C.c_uint32
        L.orc_runtime_hash_component.argtypes = [C.c_float]
        L.orc_runtime_hash.restype = C.c_float
        L.orc_runtime_hash.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_expr_eval.restype = C.c_int
        L.orc_expr_eval.argtypes = [C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float,
                                    C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
        L.orc_sample_root_f32.restype = None
        L.orc_sample_root_f32.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
        L.orc_root_attributes.restype = C.c_uint32
        L.orc_root_attributes.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                          C.c_void_p, C.c_uint32]
        L.orc_curve_eval.restype = C.c_float
        L.orc_curve_eval.argtypes = [C.c_void_p, C.c_uint32, C.c_float, C.c_void_p]
        L.orc_curve_spline_length.restype = C.c_float
        L.orc_curve_spline_length.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_classic_endpoints.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
        L.orc_clump_prepare.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]
        L.orc_clump_apply.restype = C.c_uint32
        L.orc_clump_apply.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]
        L.orc_seexpr_gradients.restype = None
        L.orc_seexpr_gradients.argtypes = [C.c_void_p]
        L.orc_classic_runtime.restype = C.c_uint32
        L.orc_classic_runtime.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_uint32]
        L.orc_classic_cut.restype = C.c_uint32
        L.orc_classic_cut.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                      C.c_float, C.c_float, C.c_void_p, C.c_float, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]
        _LIB = L
    return _LIB


def ref():
    """The reference's own CPU core, or None where it was not built here."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libnxg_ref.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.ref_build_asset.restype = C.c_int
        L.ref_build_asset.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64),
                                      C.c_char_p, C.c_uint64]
        L.ref_free.argtypes = [C.c_void_p]
        L.ref_validate_asset.restype = C.c_int
        L.ref_validate_asset.argtypes = [C.c_void_p, C.c_uint64, C.c_char_p, C.c_uint64]
        L.ref_generate_packed.restype = C.c_int
        L.ref_generate_packed.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_float, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_char_p, C.c_uint64]
        L.ref_gradient_noise.restype = C.c_float
        L.ref_gradient_noise.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float]
        L.ref_blend.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ref_strand_state.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
        L.ref_hash32.restype = C.c_uint32
        L.ref_hash32.argtypes = [C.c_uint32]
        L.ref_random01.restype = C.c_float
        L.ref_random01.argtypes = [C.c_uint32] * 3
        L.ref_clump_apply.restype = C.c_uint32
        L.ref_clump_apply.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
        L.ref_build_curve_cache.restype = C.c_int
        L.ref_build_curve_cache.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                            C.POINTER(C.c_void_p), C.POINTER(C.c_uint64), C.c_char_p, C.c_uint64]
        L.ref_validate_curve_cache.argtypes = [C.c_void_p, C.c_uint64, C.c_char_p, C.c_uint64]
        L.ref_curve_payload.restype = C.c_uint32
        L.ref_curve_payload.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                        C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_char_p, C.c_uint64]
        L.ref_expr_compile.restype = C.c_int
        L.ref_expr_compile.argtypes = [C.c_char_p, C.c_uint64, C.c_char_p, C.c_char_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_uint32, C.c_void_p,
                                       C.c_char_p, C.c_uint64]
        L.ref_seexpr_hash.restype = C.c_double
        L.ref_seexpr_hash.argtypes = [C.c_void_p, C.c_uint32]
        L.ref_face_seed.restype = C.c_double
        L.ref_face_seed.argtypes = [C.c_uint32, C.c_char_p, C.c_uint32]
        L.ref_classic_endpoints.restype = C.c_uint32
        L.ref_classic_endpoints.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
        L.ref_classic_runtime.restype = C.c_uint32
        L.ref_classic_runtime.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                          C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_float,
                                          C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_int, C.c_char_p, C.c_uint64, C.c_void_p]
        L.ref_classic_cut.restype = C.c_uint32
        L.ref_classic_cut.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                      C.c_float, C.c_float, C.c_float, C.c_void_p]
        _REF = L
    return _REF


def _take_blob(L, free, fn, build_input):
    out = C.c_void_p()
    size = C.c_uint64()
    err = C.create_string_buffer(512)
    rc = fn(C.byref(build_input.struct), C.byref(out), C.byref(size), err, 512)
    if rc != 0:
        raise ValueError(err.value.decode())
    blob = abi.aligned_bytes(size.value)
    C.memmove(blob.ctypes.data, out.value, size.value)
    free(out)
    return blob


def build_asset(build_input):
    """Oracle build_asset (asset.cpp:229-293) -> 64-aligned uint8 blob."""
    L = lib()
    return _take_blob(L, L.orc_free, L.orc_build_asset, build_input)


def ref_build_asset(build_input):
    R = ref()
    return _take_blob(R, R.ref_free, R.ref_build_asset, build_input)


def validate_asset(blob):
    msg = C.create_string_buffer(512)
    lib().orc_validate_asset(blob.ctypes.data, blob.nbytes, msg, 512)
    return msg.value.decode()


def ref_validate_asset(blob):
    msg = C.create_string_buffer(512)
    ref().ref_validate_asset(blob.ctypes.data, blob.nbytes, msg, 512)
    return msg.value.decode()


class _Job(C.Structure):
    _fields_ = [("blob", C.c_void_p), ("deformed", C.c_void_p), ("params", C.c_void_p),
                ("octaves", C.c_void_p), ("explicit_roots", C.c_void_p), ("radius_scale", C.c_float),
                ("points", C.c_void_p), ("roots", C.c_void_p), ("root_uvs", C.c_void_p),
                ("point_counts", C.c_void_p), ("stencil_indices", C.c_void_p),
                ("stencil_weights", C.c_void_p), ("strand_list", C.c_void_p),
                ("list_count", C.c_uint32), ("attributes", C.c_void_p)]


def _deformed_struct(blob, deformed):
    if deformed is None:
        return None, []
    h = abi.header_of(blob)
    keep = []
    d = abi.DeformedGeometry()
    for name, count in (("positions", h.vertex_count), ("normals", h.vertex_count),
                        ("guide_cvs", h.guide_cv_count)):
        a = deformed.get(name)
        if a is not None:
            a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1, 3)
            keep.append(a)
            setattr(d, name, a.ctypes.data)
            setattr(d, {"positions": "position_count", "normals": "normal_count",
                        "guide_cvs": "guide_cv_count"}[name], len(a))
    keep.append(d)
    return d, keep


def generate(blob, params, *, octaves=None, deformed=None, roots=None, radius_scale=1.0,
             precision=32, workers=0, block=128, stencil=False, strands=None, attributes=None):
    """Run the oracle's packed generation.  precision 32: the float
    restatement (points float32 [N*C,4], roots ROOT_DTYPE); 64: the FP64 twin
    (points float64 [N*C,4], roots float64 [N,12]).  `strands` (optional)
    selects a subset of strand indices; outputs are then in that order."""
    L = lib()
    slist = None if strands is None else np.ascontiguousarray(strands, dtype=np.uint32)
    n = params.strand_count if slist is None else len(slist)
    cvs = params.cvs_per_strand
    dt = np.float32 if precision == 32 else np.float64
    points = np.zeros((n * cvs, 4), dtype=dt)
    root_arr = np.zeros(n, dtype=abi.ROOT_DTYPE) if precision == 32 else np.zeros((n, 12))
    uvs = np.zeros((n, 2), dtype=np.float32)
    counts = np.zeros(n, dtype=np.uint32)
    sidx = np.zeros((n, 8), dtype=np.uint32) if stencil else None
    sw = np.zeros((n, 8), dtype=np.float32) if stencil else None
    d, keep = _deformed_struct(blob, deformed)
    eroots = None if roots is None else np.ascontiguousarray(roots)
    attr = None if attributes is None else np.ascontiguousarray(attributes, dtype=np.float32).reshape(-1, 4)
    job = _Job(attributes=abi.ptr(attr), blob=blob.ctypes.data, deformed=C.addressof(d) if d is not None else None,
               params=C.addressof(params), octaves=C.addressof(octaves) if octaves is not None else None,
               explicit_roots=abi.ptr(eroots), radius_scale=radius_scale, points=points.ctypes.data,
               roots=root_arr.ctypes.data, root_uvs=uvs.ctypes.data if precision == 32 else None,
               point_counts=counts.ctypes.data, stencil_indices=abi.ptr(sidx),
               stencil_weights=abi.ptr(sw), strand_list=abi.ptr(slist),
               list_count=0 if slist is None else len(slist))
    L.orc_generate(C.byref(job), precision, workers, block)
    del keep
    out = dict(points=points, roots=root_arr, point_counts=counts)
    if precision == 32:
        out["root_uvs"] = uvs
    if stencil:
        out["stencil_indices"] = sidx
        out["stencil_weights"] = sw
    return out


def ref_generate(blob, params, *, deformed=None, roots=None, radius_scale=1.0, workers=0, block=128):
    """The reference's generate_packed_cpu / _deformed_cpu / _roots_cpu."""
    R = ref()
    n, cvs = params.strand_count, params.cvs_per_strand
    points = np.zeros((n * cvs, 4), dtype=np.float32)
    root_arr = np.zeros(n, dtype=abi.ROOT_DTYPE)
    uvs = np.zeros((n, 2), dtype=np.float32)
    counts = np.zeros(n, dtype=np.uint32)
    d, keep = _deformed_struct(blob, deformed)
    eroots = None if roots is None else np.ascontiguousarray(roots)
    err = C.create_string_buffer(512)
    rc = R.ref_generate_packed(blob.ctypes.data, blob.nbytes, C.addressof(params),
                               C.addressof(d) if d is not None else None, abi.ptr(eroots),
                               radius_scale, workers, block, points.ctypes.data, root_arr.ctypes.data,
                               uvs.ctypes.data, counts.ctypes.data, err, 512)
    del keep
    if rc != 0:
        raise ValueError(err.value.decode())
    return dict(points=points, roots=root_arr, root_uvs=uvs, point_counts=counts)


def cdf_prefix(blob):
    """The inverse-CDF sampler's prefix sums (uint64 [T]) for a blob."""
    L = lib()
    t = abi.header_of(blob).triangle_count
    out = np.zeros(t, np.uint64)
    L.orc_cdf_prefix(C.c_void_p(blob.ctypes.data), C.c_void_p(out.ctypes.data))
    return out


def cdf_roots(blob, seed, strand_count, prefix=None):
    """The inverse-CDF sampler's roots of strands 0..strand_count-1 (rest
    pose) as ROOT_DTYPE records."""
    L = lib()
    prefix = cdf_prefix(blob) if prefix is None else np.ascontiguousarray(prefix, np.uint64)
    out = np.zeros(strand_count, abi.ROOT_DTYPE)
    L.orc_cdf_roots(C.c_void_p(blob.ctypes.data), C.c_void_p(prefix.ctypes.data), C.c_uint32(seed),
                    C.c_uint32(strand_count), C.c_void_p(out.ctypes.data))
    return out


def gradients_of(blob):
    h = abi.header_of(blob)
    return blob[h.noise_gradients_offset:h.noise_gradients_offset + 256 * 12]


def gradient_noise(blob, x, y, z, precision=32):
    g = gradients_of(blob)
    if precision == 32:
        return lib().orc_gradient_noise_f32(g.ctypes.data, x, y, z)
    return lib().orc_gradient_noise_f64(g.ctypes.data, x, y, z)


def strand_state(blob, params, strand, octaves=None, precision=32):
    job = _Job(blob=blob.ctypes.data, params=C.addressof(params),
               octaves=C.addressof(octaves) if octaves is not None else None)
    out = np.zeros(8, dtype=np.float32 if precision == 32 else np.float64)
    fn = lib().orc_strand_state_f32 if precision == 32 else lib().orc_strand_state_f64
    fn(C.byref(job), strand, out.ctypes.data)
    return out


def root_attributes(blob, params, programs, maps=(), *, strands=None, workers=0):
    """Oracle of nxg_evaluate_root_attributes: float32 [N, 4] (NaN where the
    reference interpreter would raise) and the failure count.  `programs` are
    nanoxgen_amd.expr.Program objects; `maps` are (texels [F,R,R] float32)."""
    from nanoxgen_amd import expr
    L = lib()
    slist = None if strands is None else np.ascontiguousarray(strands, dtype=np.uint32)
    n = params.strand_count if slist is None else len(slist)
    out = np.zeros((n, 4), dtype=np.float32)
    structs = (abi.ExprProgram * max(1, len(programs)))(*[p.struct for p in programs])
    mkeep = [np.ascontiguousarray(m, dtype=np.float32) for m in maps]
    mstructs = (abi.TexelMap * max(1, len(mkeep)))(*[abi.TexelMap(m.ctypes.data, m.shape[0], m.shape[1]) for m in mkeep])
    job = _Job(blob=blob.ctypes.data, params=C.addressof(params), strand_list=abi.ptr(slist),
               list_count=0 if slist is None else len(slist))
    fails = L.orc_root_attributes(C.byref(job), structs, len(programs), mstructs, len(mkeep), out.ctypes.data, workers)
    return out, fails


def expr_eval(program, root_vars, u=0.0, v=0.0, face_seed=0.0, t=0.0, gradients=None, maps=()):
    """One program on one set of root variables -> float or None (error)."""
    L = lib()
    rv = np.ascontiguousarray(root_vars, dtype=np.float32)
    scratch = np.zeros(abi.EXPR_MAX_VALUES, dtype=np.float32)
    out = C.c_float()
    mkeep = [np.ascontiguousarray(m, dtype=np.float32) for m in maps]
    mstructs = (abi.TexelMap * max(1, len(mkeep)))(*[abi.TexelMap(m.ctypes.data, m.shape[0], m.shape[1]) for m in mkeep])
    g = None if gradients is None else gradients.ctypes.data
    rc = L.orc_expr_eval(C.byref(program.struct), rv.ctypes.data, u, v, face_seed, t, g, mstructs, len(mkeep),
                         scratch.ctypes.data, C.byref(out))
    return None if rc else out.value


def curve_spline_length(points):
    """curve_spline_length of one strand (float32 [C, 4] or [C, 3])."""
    p = np.zeros((len(points), 4), np.float32)
    p[:, :3] = np.asarray(points, np.float32)[:, :3]
    return lib().orc_curve_spline_length(p.ctypes.data, len(p))


def classic_cut(points, cvs, amount_scale, amount_offset, *, counts=None, roots=None, uvs=None, multiplier=None,
                radius_scale=1.0, workers=0):
    """The CutFX chain (cLength, culling, cut from tip, compaction) over
    packed float4 points [N * C, 4]: returns dict(points, point_counts,
    roots, root_uvs, keep, strand_count), or None where the reference
    would throw (a non-finite amount)."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 4)
    n = len(pts) // cvs
    counts = np.full(n, cvs, np.uint32) if counts is None else np.ascontiguousarray(counts, np.uint32)
    out = np.zeros_like(pts)
    oc = np.zeros(n, np.uint32)
    orr = np.zeros(n, abi.ROOT_DTYPE) if roots is not None else None
    ou = np.zeros((n, 2), np.float32) if uvs is not None else None
    keep = np.zeros(n, np.uint8)
    m = None if multiplier is None else np.ascontiguousarray(multiplier, np.float32)
    kept = lib().orc_classic_cut(pts.ctypes.data, counts.ctypes.data, abi.ptr(roots), abi.ptr(uvs), n, cvs,
                                 amount_scale, amount_offset, abi.ptr(m), radius_scale, out.ctypes.data,
                                 oc.ctypes.data, abi.ptr(orr), abi.ptr(ou), keep.ctypes.data, workers)
    if kept == 0xFFFFFFFF:
        return None
    return dict(points=out[:kept * cvs], point_counts=oc[:kept], roots=None if orr is None else orr[:kept],
                root_uvs=None if ou is None else ou[:kept], keep=keep, strand_count=kept)


def ref_classic_cut(points, cvs, amount_scale, amount_offset, *, counts=None, roots=None, uvs=None,
                    radius_scale=1.0):
    """The reference's apply_xgen_classic_float_runtime_plan_cpu with one
    CutFX module (amount = amount_scale * $cLength + amount_offset); the
    result's "seconds" is the wall time of the reference call alone."""
    pts = np.array(points, np.float32).reshape(-1, 4)
    n = len(pts) // cvs
    counts = np.full(n, cvs, np.uint32) if counts is None else np.array(counts, np.uint32)
    rr = np.zeros(n, abi.ROOT_DTYPE) if roots is None else np.array(roots)
    uu = None if uvs is None else np.array(uvs, np.float32)
    seconds = C.c_double()
    kept = ref().ref_classic_cut(pts.ctypes.data, counts.ctypes.data, rr.ctypes.data, abi.ptr(uu), n, cvs,
                                 amount_scale, amount_offset, radius_scale, C.byref(seconds))
    if kept == 0xFFFFFFFF:
        return None
    return dict(points=pts[:kept * cvs], point_counts=counts[:kept], roots=rr[:kept],
                root_uvs=None if uu is None else uu[:kept], strand_count=kept, seconds=seconds.value)


class _ClumpBinding(C.Structure):
    """orc_clump_binding."""
    _fields_ = [("axes3", C.c_void_p), ("render3", C.c_void_p), ("lengths", C.c_void_p), ("normals3", C.c_void_p),
                ("tangents3", C.c_void_p), ("reference3", C.c_void_p), ("guide_uvs", C.c_void_p),
                ("guide_prefixes", C.c_void_p), ("strand_guides", C.c_void_p), ("guide_count", C.c_uint32),
                ("gradients", C.c_void_p), ("local3", C.c_void_p), ("local_render3", C.c_void_p)]


_MAX_EFFECTS = 8
_EFFECT_TYPES = {"clump": 0, "noise": 1, "cut": 2}


class _ClassicPlan(C.Structure):
    """orc_classic_plan."""
    _fields_ = [("length", C.c_void_p), ("width", C.c_void_p), ("taper", C.c_void_p),
                ("taper_start", C.c_void_p), ("width_ramp", C.c_void_p),
                ("description_id", C.c_uint32), ("name_seed", C.c_uint32), ("effect_count", C.c_uint32),
                ("effect_type", C.c_uint32 * _MAX_EFFECTS), ("effect_module", C.c_uint32 * _MAX_EFFECTS),
                ("clump", (C.c_void_p * 7) * 3), ("binding", C.c_void_p * 3), ("noise", (C.c_void_p * 6) * 2),
                ("cut_amount", C.c_void_p * 2), ("root_relative", C.c_int), ("surface_tangents3", C.c_void_p),
                ("noise_domain3", C.c_void_p), ("gradients", C.c_void_p)]


def seexpr_gradients():
    """The SeExpr gradient table, float32 [256, 3]."""
    g = np.zeros((256, 3), np.float32)
    lib().orc_seexpr_gradients(g.ctypes.data)
    return g


def _u32(x, n):
    return None if x is None else np.ascontiguousarray(np.asarray(x).reshape(-1)[:n]).astype(np.uint32)


def _clump_list(clumps, strand_guides, guides, guide_uvs, guide_random_prefixes):
    if clumps is None and guides is not None:
        return [dict(strand_guides=strand_guides, guides=guides, guide_uvs=guide_uvs,
                     guide_random_prefixes=guide_random_prefixes)]
    return list(clumps or [])


def classic_runtime(points, cvs, plan, *, roots, counts=None, uvs=None, primitive_ids=None, random_prefixes=None,
                    columns=None, radius_scale=1.0, strand_guides=None, guides=None, guide_uvs=None,
                    guide_random_prefixes=None, clumps=None, surface_tangents=None, noise_domain_positions=None,
                    root_relative=False, workers=0):
    """orc_classic_runtime: the Classic runtime chain of `plan` (a
    nanoxgen_amd.classic.ClassicPlan: its compiled, bound programs and its
    effect list) over packed float4 points [N * C, 4]; each clump module of
    the plan takes a binding in `clumps` (dicts of strand_guides, guides
    (dict(axes, normals, tangents[, reference])), guide_uvs,
    guide_random_prefixes and optional local_axes), or, for one module,
    the first four as keywords.  Returns dict(points, point_counts, roots,
    root_uvs, keep, strand_count, error_bits); strand_count is None where the
    reference throws (error_bits says why)."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 4)
    n = len(pts) // cvs
    rr = np.ascontiguousarray(roots).view(abi.ROOT_DTYPE).reshape(-1)
    cc = _u32(counts, n)
    uu = None if uvs is None else np.ascontiguousarray(uvs, np.float32).reshape(-1, 2)
    ids, pre = _u32(primitive_ids, n), _u32(random_prefixes, n)
    cols = None if columns is None else np.ascontiguousarray(columns, np.float32).reshape(n, -1)
    ncol = 0 if cols is None else cols.shape[1]
    op = _ClassicPlan(description_id=plan.description_id, name_seed=plan.name_seed,
                      effect_count=len(plan.effects), root_relative=1 if root_relative else 0)
    for e, (kind, module) in enumerate(plan.effects):
        op.effect_type[e] = _EFFECT_TYPES[kind]
        op.effect_module[e] = module
    for key, prog in plan.programs.items():
        name, _, module = key.partition("@")
        module = int(module or 0)
        if name in abi.CLASSIC_NOISE_EXPRS:
            op.noise[module][abi.CLASSIC_NOISE_EXPRS.index(name)] = C.addressof(prog.struct)
        elif name in abi.CLASSIC_CLUMP_EXPRS:
            op.clump[module][abi.CLASSIC_CLUMP_EXPRS.index(name)] = C.addressof(prog.struct)
        elif name == "cut_amount":
            op.cut_amount[module] = C.addressof(prog.struct)
        else:
            setattr(op, name, C.addressof(prog.struct))
    keepalive = []
    if plan.has_noise:
        noise_arrays = [_f3(surface_tangents, n), None if noise_domain_positions is None
                        else _f3(noise_domain_positions, n), seexpr_gradients()]
        keepalive.append(noise_arrays)
        op.surface_tangents3, op.noise_domain3, op.gradients = [abi.ptr(x) for x in noise_arrays]
    bindings = _clump_list(clumps, strand_guides, guides, guide_uvs, guide_random_prefixes)
    for m in range(plan.clump_count):
        b = bindings[m]
        gs = b["guides"]
        ax = _f3(gs["axes"], -1)
        g = len(ax) // cvs
        render, lengths = clump_prepare(ax, cvs)
        local = local_render = None
        if b.get("local_axes") is not None:  # their lengths, whatever the mode (:871-887)
            local = _f3(b["local_axes"], g * cvs)
            local_render, lengths = clump_prepare(local, cvs)
        arrays = dict(axes3=ax, render3=render, lengths=lengths, local3=local, local_render3=local_render,
                      normals3=_f3(gs["normals"], g),
                      tangents3=_f3(gs["tangents"], g),
                      reference3=None if gs.get("reference") is None else _f3(gs["reference"], g),
                      guide_uvs=np.ascontiguousarray(b["guide_uvs"], np.float32).reshape(g, 2),
                      guide_prefixes=_u32(b["guide_random_prefixes"], g), strand_guides=_u32(b["strand_guides"], n),
                      gradients=seexpr_gradients())
        keepalive.append(arrays)
        binding = _ClumpBinding(guide_count=g, **{k: abi.ptr(v) for k, v in arrays.items()})
        keepalive.append(binding)
        op.binding[m] = C.addressof(binding)
    out = np.zeros_like(pts)
    oc = np.zeros(n, np.uint32)
    orr = np.zeros(n, abi.ROOT_DTYPE)
    ou = np.zeros((n, 2), np.float32)
    keep = np.zeros(n, np.uint8)
    bits = C.c_uint32()
    kept = lib().orc_classic_runtime(pts.ctypes.data, abi.ptr(cc), rr.ctypes.data, abi.ptr(uu), n, cvs,
                                     C.byref(op), abi.ptr(ids), abi.ptr(pre), abi.ptr(cols), ncol, radius_scale,
                                     out.ctypes.data, oc.ctypes.data, orr.ctypes.data, ou.ctypes.data,
                                     keep.ctypes.data, C.byref(bits), workers)
    if kept == 0xFFFFFFFF:
        return dict(strand_count=None, error_bits=bits.value, keep=keep)
    return dict(points=out[:kept * cvs], point_counts=oc[:kept], roots=orr[:kept],
                root_uvs=None if uu is None else ou[:kept], keep=keep, strand_count=kept, error_bits=0)


def ref_classic_runtime(xgen_text, points, cvs, *, roots, counts=None, uvs=None, primitive_ids=None,
                        random_prefixes=None, columns=None, radius_scale=1.0, clump_name="clump", strand_guides=None,
                        guides=None, guide_uvs=None, guide_random_prefixes=None, clumps=None, clump_names=None,
                        surface_tangents=None, noise_domain_positions=None, root_relative=False):
    """The reference's parse -> compile_xgen_classic_float_runtime_plan ->
    apply_xgen_classic_float_runtime_plan_cpu over the first description of
    `xgen_text`, with one geometry binding per ClumpingFX module (`clumps`
    and their module names `clump_names`, or one module's binding as
    keywords).  Returns dict(points, point_counts, roots, root_uvs,
    strand_count, seconds), or dict(strand_count=None, error=message) where
    the reference throws."""
    pts = np.array(points, np.float32).reshape(-1, 4)
    n = len(pts) // cvs
    cc = np.full(n, cvs, np.uint32) if counts is None else np.array(counts, np.uint32)
    rr = np.array(np.ascontiguousarray(roots).view(abi.ROOT_DTYPE).reshape(-1))
    uu = None if uvs is None else np.array(uvs, np.float32).reshape(-1, 2)
    ids, pre = _u32(primitive_ids, n), _u32(random_prefixes, n)
    cols = None if columns is None else np.ascontiguousarray(columns, np.float32).reshape(-1)
    err = C.create_string_buffer(1024)
    seconds = C.c_double()
    bindings = _clump_list(clumps, strand_guides, guides, guide_uvs, guide_random_prefixes)
    names = list(clump_names) if clump_names is not None else [clump_name] * len(bindings)
    m = len(bindings)

    def cat(f, w):
        return np.ascontiguousarray(np.concatenate([f(b) for b in bindings]).reshape(-1, w)) if m else None

    guide_counts = np.array([len(_f3(b["guides"]["axes"], -1)) // cvs for b in bindings] or [0], np.uint32)
    ga = cat(lambda b: _f3(b["guides"]["axes"], -1), 3)
    gn = cat(lambda b: _f3(b["guides"]["normals"], -1), 3)
    gt = cat(lambda b: _f3(b["guides"]["tangents"], -1), 3)
    with_reference = m and all(b["guides"].get("reference") is not None for b in bindings)
    gr = cat(lambda b: _f3(b["guides"]["reference"], -1), 3) if with_reference else None
    gu = cat(lambda b: np.asarray(b["guide_uvs"], np.float32).reshape(-1, 2), 2)
    gp = None if not m else np.concatenate([np.asarray(b["guide_random_prefixes"]).reshape(-1).astype(np.uint32)
                                            for b in bindings])
    sg = None if not m else np.concatenate([_u32(b["strand_guides"], n) for b in bindings])
    local_flags = np.array([b.get("local_axes") is not None for b in bindings] or [0], np.uint32)
    gl = cat(lambda b: _f3(b["local_axes"] if b.get("local_axes") is not None else b["guides"]["axes"], -1), 3)
    name_ptrs = (C.c_char_p * max(m, 1))(*[x.encode() for x in names])
    kept = ref().ref_classic_runtime(xgen_text.encode(), pts.ctypes.data, cc.ctypes.data, rr.ctypes.data,
                                     abi.ptr(uu), n, cvs, abi.ptr(ids), abi.ptr(pre), abi.ptr(cols),
                                     0 if cols is None else cols.size, radius_scale, m, name_ptrs,
                                     guide_counts.ctypes.data, abi.ptr(ga), abi.ptr(gn), abi.ptr(gt), abi.ptr(gr),
                                     abi.ptr(gu), abi.ptr(gp), abi.ptr(sg),
                                     abi.ptr(None if surface_tangents is None else _f3(surface_tangents, n)),
                                     abi.ptr(None if noise_domain_positions is None else _f3(noise_domain_positions, n)),
                                     local_flags.ctypes.data, abi.ptr(gl), 1 if root_relative else 0,
                                     err, 1024, C.byref(seconds))
    if kept == 0xFFFFFFFF:
        return dict(strand_count=None, error=err.value.decode())
    return dict(points=pts[:kept * cvs], point_counts=cc[:kept], roots=rr[:kept],
                root_uvs=None if uu is None else uu[:kept], strand_count=kept, seconds=seconds.value)


def classic_endpoints(points, cvs):
    """Renderer endpoints: [N * cvs, 4] -> [N * (cvs + 2), 4]."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 4)
    n = len(pts) // cvs
    out = np.zeros((n * (cvs + 2), 4), np.float32)
    lib().orc_classic_endpoints(pts.ctypes.data, n, cvs, out.ctypes.data)
    return out


def ref_classic_endpoints(points, cvs):
    """The reference's add_xgen_classic_renderer_endpoints (None where it throws)."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 4)
    n = len(pts) // cvs
    out = np.zeros((n * (cvs + 2), 4), np.float32)
    return None if ref().ref_classic_endpoints(pts.ctypes.data, n, cvs, out.ctypes.data) else out


CLUMP_KEYS = ("mask", "amount", "clump_scale", "noise", "noise_scale", "noise_frequency", "noise_correlation")


def _f3(a, rows):
    return np.ascontiguousarray(np.asarray(a, np.float32).reshape(rows, 3))


def clump_prepare(axes, cvs):
    """prepare_xgen_classic_clump_runtime_data: (render axes [G*cvs, 3], spline lengths [G])."""
    ax = _f3(axes, -1)
    g = len(ax) // cvs
    render = np.zeros_like(ax)
    lengths = np.zeros(g, np.float32)
    lib().orc_clump_prepare(ax.ctypes.data, g, cvs, render.ctypes.data, lengths.ctypes.data)
    return render, lengths


def clump_apply(points, cvs, strand_guides, guides, params, gradients, amount_multiplier=None, multiplier_stride=1,
                workers=0):
    """ClumpingFX with constant expressions (params: dict of CLUMP_KEYS) on
    packed strands; guides: dict(axes, normals, tangents[, reference_positions]);
    gradients: the 256 x 3 SeExpr table.  Returns the new points, or None
    where the reference throws."""
    pts = np.array(points, np.float32).reshape(-1, 4)
    n = len(pts) // cvs
    ax = _f3(guides["axes"], -1)
    g = len(ax) // cvs
    render, lengths = clump_prepare(ax, cvs)
    nrm, tan = _f3(guides["normals"], g), _f3(guides["tangents"], g)
    ref3 = guides.get("reference_positions")
    ref3 = None if ref3 is None else _f3(ref3, g)
    sg = np.ascontiguousarray(strand_guides, np.uint32)
    k = (C.c_float * 7)(*[params[key] for key in CLUMP_KEYS])
    grad = np.ascontiguousarray(gradients, np.float32)
    m = None if amount_multiplier is None else np.ascontiguousarray(amount_multiplier, np.float32)
    rc = lib().orc_clump_apply(pts.ctypes.data, n, cvs, sg.ctypes.data, ax.ctypes.data, render.ctypes.data,
                               lengths.ctypes.data, nrm.ctypes.data, tan.ctypes.data, abi.ptr(ref3), g,
                               grad.ctypes.data, k, abi.ptr(m), multiplier_stride, workers)
    return None if rc else pts


def ref_clump_apply(points, cvs, strand_guides, guides, params):
    """The reference's float runtime with one constant-expression ClumpingFX module."""
    pts = np.array(points, np.float32).reshape(-1, 4)
    n = len(pts) // cvs
    ax = _f3(guides["axes"], -1)
    g = len(ax) // cvs
    nrm, tan = _f3(guides["normals"], g), _f3(guides["tangents"], g)
    ref3 = guides.get("reference_positions")
    ref3 = None if ref3 is None else _f3(ref3, g)
    sg = np.ascontiguousarray(strand_guides, np.uint32)
    k = np.array([params[key] for key in CLUMP_KEYS], np.float32)
    kept = ref().ref_clump_apply(pts.ctypes.data, n, cvs, sg.ctypes.data, ax.ctypes.data, nrm.ctypes.data,
                                 tan.ctypes.data, abi.ptr(ref3), g, k.ctypes.data)
    if kept == 0xFFFFFFFF:
        return None
    assert kept == n, "strands were culled by the runtime; pick inputs with length and width"
    return pts


def ref_build_curve_cache(counts, points, *, texcoords=None, patch_uvs=None, face_uvs=None, face_ids=None,
                          motion_times=None, motion_points=None):
    """The reference's build_curve_cache: (bytes as uint8, "") or (None, message)."""
    R = ref()
    a = lambda x, dt: None if x is None else np.ascontiguousarray(x, dt)
    counts, points = a(counts, np.uint32), a(points, np.float32).reshape(-1, 4)
    tex, pu, fu, fid = a(texcoords, np.float32), a(patch_uvs, np.float32), a(face_uvs, np.float32), a(face_ids, np.uint32)
    mt, mp = a(motion_times, np.float32), a(motion_points, np.float32)
    out, size = C.c_void_p(), C.c_uint64()
    err = C.create_string_buffer(512)
    rc = R.ref_build_curve_cache(counts.ctypes.data, len(counts), points.ctypes.data, len(points), abi.ptr(tex),
                                 abi.ptr(pu), abi.ptr(fu), abi.ptr(fid), abi.ptr(mt), 0 if mt is None else len(mt),
                                 abi.ptr(mp), C.byref(out), C.byref(size), err, 512)
    if rc:
        return None, err.value.decode()
    data = np.frombuffer(C.string_at(out.value, size.value), np.uint8).copy()
    R.ref_free(out)
    return data, ""


def ref_validate_curve_cache(blob):
    """The reference's validate_curve_cache on a 64-byte aligned copy."""
    msg = C.create_string_buffer(512)
    ref().ref_validate_curve_cache(blob.ctypes.data, blob.nbytes, msg, 512)
    return msg.value.decode()


IDENTITY = (1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0)


def ref_curve_payload(base, motion=(), *, max_strands=65536, resample=0, radius_scale=1.0, transform=IDENTITY):
    """The reference's build_curve_batches.  base / each motion sample: dict(points [P, 3], counts=None|[N],
    fixed=0, widths=None|[P], constant_width=0.0[, time][, root_uvs]).  Returns dict(points, point_counts,
    root_uvs, motion_points, batch_first) with the batches concatenated, or (None, message)."""
    sets = [base] + list(motion)
    count_of = lambda s: len(s["counts"]) if s.get("counts") is not None else len(s["points"]) // s["fixed"]
    n = count_of(base)
    strand_counts = np.array([count_of(s) for s in sets], np.uint32)
    keep = []

    def arr(x, dt):
        if x is None:
            return None
        a = np.ascontiguousarray(x, dt)
        keep.append(a)
        return a

    counts = [arr(s.get("counts"), np.uint32) for s in sets]
    pts = [arr(s["points"], np.float32) for s in sets]
    wid = [arr(s.get("widths"), np.float32) for s in sets]
    P = lambda a: C.c_void_p(0) if a is None else C.c_void_p(a.ctypes.data)
    cp = (C.c_void_p * len(sets))(*[P(c) for c in counts])
    pp = (C.c_void_p * len(sets))(*[P(x) for x in pts])
    wp = (C.c_void_p * len(sets))(*[P(w) for w in wid])
    fixed = np.array([s.get("fixed", 0) for s in sets], np.uint32)
    npts = np.array([len(x) // 3 if x.ndim == 1 else len(x) for x in pts], np.uint64)
    cw = np.array([s.get("constant_width", 0.0) for s in sets], np.float32)
    times = np.array([s.get("time", 0.0) for s in motion] or [0.0], np.float32)
    uvs = arr(base.get("root_uvs"), np.float32)
    total = n * resample if resample else int(npts[0])
    out_p = np.zeros((max(total, 1), 4), np.float32)
    out_c = np.zeros(max(n, 1), np.uint32)
    out_u = np.zeros((max(n, 1), 2), np.float32)
    out_m = np.zeros((max(total * len(motion), 1), 3), np.float32)
    first = np.zeros(n + 1, np.uint32)
    err = C.create_string_buffer(512)
    tr = np.array(transform, np.float32)
    b = ref().ref_curve_payload(strand_counts.ctypes.data, len(sets), cp, fixed.ctypes.data, pp, npts.ctypes.data, wp, cw.ctypes.data,
                                times.ctypes.data, abi.ptr(uvs), max_strands, resample, radius_scale, tr.ctypes.data,
                                out_p.ctypes.data, out_c.ctypes.data, out_u.ctypes.data, out_m.ctypes.data,
                                first.ctypes.data, err, 512)
    if b == 0xFFFFFFFF:
        return None, err.value.decode()
    return dict(points=out_p[:total], point_counts=out_c[:n], root_uvs=out_u[:n] if uvs is not None else None,
                motion_points=out_m[:total * len(motion)], batch_first=first[:b]), ""


_INSTR = np.dtype([("op", "u1"), ("reserved", "u1"), ("count", "<u2"), ("offset", "<u4"), ("aux", "<u4"),
                   ("imm", "<f4")])
_RPT = np.dtype([("p", "<f4"), ("v", "<f4"), ("i", "<u4")])


def ref_expr_compile(source, name="", object_type="", limits=(0, 0, 0, 0, 0), capacity=4096):
    """The reference's compile + float conversion: dict(instructions, operands,
    ramp_points, ramps, inputs, result, expression_seed, random_call_count) or
    (None, message)."""
    src = source.encode()
    ins = np.zeros(capacity, _INSTR)
    ops = np.zeros(capacity, np.uint32)
    pts = np.zeros(capacity, _RPT)
    ramps = np.zeros((capacity, 2), np.uint32)
    names = C.create_string_buffer(65536)
    counts = np.zeros(8, np.uint32)
    lim = np.array(limits, np.uint64)
    err = C.create_string_buffer(512)
    rc = ref().ref_expr_compile(src, len(src), name.encode(), object_type.encode(), lim.ctypes.data, ins.ctypes.data,
                                ops.ctypes.data, pts.ctypes.data, ramps.ctypes.data, names, capacity,
                                counts.ctypes.data, err, 512)
    if rc:
        return None, err.value.decode()
    n_ins, n_ops, n_pts, n_ramps, n_in, result, seed, rands = map(int, counts)
    raw = names.raw
    inputs, at = [], 0
    for _ in range(n_in):
        end = raw.index(b"\0", at)
        inputs.append(raw[at:end].decode())
        at = end + 1
    return dict(instructions=ins[:n_ins], operands=ops[:n_ops], ramp_points=pts[:n_pts], ramps=ramps[:n_ramps],
                inputs=inputs, result=result, expression_seed=seed, random_call_count=rands), ""
