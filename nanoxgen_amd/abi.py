"""ctypes mirror of include/nanoxgen_amd.h.

These are the POD layouts of the C ABI (the .nxg blob types of the
reference's include/nanoxgen/types.h and the checked descriptors of
include/nanoxgen/device_contract.h).  Pure layout definitions: importing this
module loads no library.
"""
import ctypes as C

import numpy as np

MAGIC = 0x4E4547584F4E414E
GUIDE_STENCIL_SIZE = 8
NOISE_GRADIENT_COUNT = 256
INVALID_INDEX = 0xFFFFFFFF
MAX_NOISE_OCTAVES = 4
STATIC_SCRATCH_BYTES = 32 << 20  # NXG_STATIC_SCRATCH_BYTES
HAS_NORMALS = 1
HAS_TEXCOORDS = 2


class Vec2(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class Vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class AssetHeader(C.Structure):
    """types.h:77-101 (128 bytes)."""
    _fields_ = [
        ("magic", C.c_uint64),
        ("version_major", C.c_uint16),
        ("version_minor", C.c_uint16),
        ("flags", C.c_uint32),
        ("byte_size", C.c_uint64),
        ("content_hash", C.c_uint64),
        ("vertex_count", C.c_uint32),
        ("triangle_count", C.c_uint32),
        ("guide_count", C.c_uint32),
        ("guide_cv_count", C.c_uint32),
        ("guide_stencil_size", C.c_uint32),
        ("noise_gradient_count", C.c_uint32),
        ("positions_offset", C.c_uint64),
        ("normals_offset", C.c_uint64),
        ("texcoords_offset", C.c_uint64),
        ("triangles_offset", C.c_uint64),
        ("alias_table_offset", C.c_uint64),
        ("guides_offset", C.c_uint64),
        ("guide_cvs_offset", C.c_uint64),
        ("triangle_guides_offset", C.c_uint64),
        ("noise_gradients_offset", C.c_uint64),
    ]


class GenerationParams(C.Structure):
    """types.h:108-124, defaults as in the reference."""
    _fields_ = [
        ("strand_count", C.c_uint32),
        ("cvs_per_strand", C.c_uint32),
        ("seed", C.c_uint32),
        ("reserved", C.c_uint32),
        ("guide_support_scale", C.c_float),
        ("guide_weight_power", C.c_float),
        ("normal_rejection_cos", C.c_float),
        ("length_scale", C.c_float),
        ("root_width", C.c_float),
        ("tip_width", C.c_float),
        ("noise_amplitude", C.c_float),
        ("noise_frequency", C.c_float),
        ("noise_mask", C.c_float),
        ("noise_correlation", C.c_float),
        ("noise_preserve_length", C.c_float),
    ]

    def __init__(self, **kw):
        defaults = dict(strand_count=1024, cvs_per_strand=8, seed=1, reserved=0,
                        guide_support_scale=1.0, guide_weight_power=2.0,
                        normal_rejection_cos=-0.25, length_scale=1.0, root_width=0.01,
                        tip_width=0.001, noise_amplitude=0.0, noise_frequency=2.0,
                        noise_mask=1.0, noise_correlation=0.0, noise_preserve_length=0.0)
        unknown = set(kw) - set(defaults)
        if unknown:
            raise TypeError(f"unknown GenerationParams fields: {sorted(unknown)}")
        defaults.update(kw)
        super().__init__(**defaults)

    def copy(self, **kw):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d.update(kw)
        return GenerationParams(**d)


class NoiseOctaves(C.Structure):
    _fields_ = [
        ("count", C.c_uint32),
        ("amplitude", C.c_float * MAX_NOISE_OCTAVES),
        ("frequency", C.c_float * MAX_NOISE_OCTAVES),
    ]

    @classmethod
    def of(cls, pairs):
        o = cls()
        o.count = len(pairs)
        if len(pairs) > MAX_NOISE_OCTAVES:
            raise ValueError("too many octaves")
        for i, (a, f) in enumerate(pairs):
            o.amplitude[i] = a
            o.frequency[i] = f
        return o


class RootSample(C.Structure):
    """types.h:126-135 (48 bytes)."""
    _fields_ = [
        ("position", Vec3),
        ("normal", Vec3),
        ("uv", Vec2),
        ("triangle_index", C.c_uint32),
        ("barycentric", Vec2),
        ("surface_face_id", C.c_uint32),
    ]


class DeviceAsset(C.Structure):
    """device_contract.h:14-18."""
    _fields_ = [("data", C.c_void_p), ("header", AssetHeader), ("byte_capacity", C.c_uint64)]


class DeformedGeometry(C.Structure):
    _fields_ = [
        ("positions", C.c_void_p),
        ("normals", C.c_void_p),
        ("guide_cvs", C.c_void_p),
        ("position_count", C.c_uint64),
        ("normal_count", C.c_uint64),
        ("guide_cv_count", C.c_uint64),
    ]


class PackedOutput(C.Structure):
    """device_contract.h:34-40."""
    _fields_ = [
        ("points", C.c_void_p),
        ("roots", C.c_void_p),
        ("root_uvs", C.c_void_p),
        ("radius_scale", C.c_float),
        ("point_counts", C.c_void_p),
        ("point_capacity", C.c_uint64),
        ("root_capacity", C.c_uint64),
        ("root_uv_capacity", C.c_uint64),
        ("point_count_capacity", C.c_uint64),
    ]


class StencilOutput(C.Structure):
    _fields_ = [("indices", C.c_void_p), ("weights", C.c_void_p), ("capacity", C.c_uint64)]


class LaunchConfig(C.Structure):
    _fields_ = [("block_size", C.c_uint32)]


class AssetBuildInput(C.Structure):
    _fields_ = [
        ("positions", C.c_void_p),
        ("vertex_count", C.c_uint32),
        ("triangle_count", C.c_uint32),
        ("normals", C.c_void_p),
        ("texcoords", C.c_void_p),
        ("triangles", C.c_void_p),
        ("guide_count", C.c_uint32),
        ("stencil_k", C.c_uint32),
        ("guide_cv_counts", C.c_void_p),
        ("guide_cvs", C.c_void_p),
        ("guide_root_normals", C.c_void_p),
        ("guide_root_uvs", C.c_void_p),
        ("guide_triangles", C.c_void_p),
        ("guide_barycentrics", C.c_void_p),
        ("guide_support_radii", C.c_void_p),
        ("triangle_weights", C.c_void_p),
    ]


class ExprInstruction(C.Structure):
    _fields_ = [("op", C.c_uint8), ("reserved", C.c_uint8), ("operand_count", C.c_uint16),
                ("operand_offset", C.c_uint32), ("auxiliary", C.c_uint32), ("immediate", C.c_float)]


class ExprRampPoint(C.Structure):
    _fields_ = [("position", C.c_float), ("value", C.c_float), ("interpolation", C.c_uint32)]


class ExprRamp(C.Structure):
    _fields_ = [("point_offset", C.c_uint32), ("point_count", C.c_uint32)]


class ExprProgram(C.Structure):
    _fields_ = [("instructions", C.c_void_p), ("operands", C.c_void_p), ("ramp_points", C.c_void_p),
                ("ramps", C.c_void_p), ("input_bindings", C.c_void_p), ("instruction_count", C.c_uint32),
                ("operand_count", C.c_uint32), ("ramp_point_count", C.c_uint32), ("ramp_count", C.c_uint32),
                ("input_count", C.c_uint32), ("result", C.c_uint32), ("expression_seed", C.c_uint32),
                ("random_call_count", C.c_uint32)]


class ExprSet(C.Structure):
    _fields_ = [("image", C.c_void_p), ("byte_size", C.c_uint64), ("program_count", C.c_uint32),
                ("max_values", C.c_uint32), ("program_offset", C.c_uint32 * 4),
                ("texel_maps_required", C.c_uint32)]


class TexelMap(C.Structure):
    _fields_ = [("texels", C.c_void_p), ("face_count", C.c_uint32), ("resolution", C.c_uint32)]


EXPR_MAX_VALUES = 256
EXPR_MAX_PROGRAMS = 4
OPS = ["constant", "input", "negate", "logical_not", "add", "subtract", "multiply", "divide", "less",
       "less_equal", "greater", "greater_equal", "equal", "not_equal", "logical_and", "logical_or",
       "select", "hash", "random", "minimum", "maximum", "clamp", "fit", "gamma", "contrast",
       "smoothstep", "ramp", "pow", "mix", "noise3", "texel"]
OP = {name: i for i, name in enumerate(OPS)}
ROOT_VARS = ["position_x", "position_y", "position_z", "normal_x", "normal_y", "normal_z", "u", "v",
             "face", "bary1", "bary2", "id"]
ROOT_VAR = {name: i for i, name in enumerate(ROOT_VARS)}

# numpy record views of the device/host arrays
ROOT_DTYPE = np.dtype([("position", "<f4", 3), ("normal", "<f4", 3), ("uv", "<f4", 2),
                       ("triangle_index", "<u4"), ("barycentric", "<f4", 2),
                       ("surface_face_id", "<u4")])
GUIDE_DTYPE = np.dtype([("first_cv", "<u4"), ("cv_count", "<u2"), ("flags", "<u2"),
                        ("triangle_index", "<u4"), ("support_radius", "<f4"),
                        ("barycentric", "<f4", 2), ("root_uv", "<f4", 2),
                        ("root_position", "<f4", 3), ("root_normal", "<f4", 3)])

assert C.sizeof(AssetHeader) == 128
assert C.sizeof(GenerationParams) == 60
assert C.sizeof(RootSample) == 48 and ROOT_DTYPE.itemsize == 48
assert GUIDE_DTYPE.itemsize == 56
assert C.sizeof(ExprInstruction) == 16

# status codes (device_contract.h:58-80 then extensions)
class MotionSample(C.Structure):
    """nxg_motion_sample (DeviceMotionSampleDescriptor)."""
    _fields_ = [("deformation", DeformedGeometry), ("time", C.c_float)]


class MotionOutput(C.Structure):
    """nxg_motion_output (DeviceMotionOutputDescriptor): sample-major Vec3 points."""
    _fields_ = [("points", C.c_void_p), ("point_capacity", C.c_uint64)]


CLASSIC_MAX_CVS = 256
CLASSIC_NONFINITE_AMOUNT, CLASSIC_NONFINITE_WIDTH = 1, 2


class ClassicCurves(C.Structure):
    """nxg_classic_curves: packed input curves (device pointers)."""
    _fields_ = [("points", C.c_void_p), ("point_counts", C.c_void_p), ("roots", C.c_void_p),
                ("root_uvs", C.c_void_p), ("strand_count", C.c_uint32), ("cvs_per_strand", C.c_uint32)]


class ClassicCutParams(C.Structure):
    """nxg_classic_cut_params: amount = (amount_scale * cLength + amount_offset) * multiplier."""
    _fields_ = [("amount_scale", C.c_float), ("amount_offset", C.c_float), ("radius_scale", C.c_float),
                ("multiplier_stride", C.c_uint32), ("multiplier", C.c_void_p)]


class ClassicOutput(C.Structure):
    """nxg_classic_output: compacted curves plus the device kept-count and status words."""
    _fields_ = [("points", C.c_void_p), ("point_counts", C.c_void_p), ("roots", C.c_void_p),
                ("root_uvs", C.c_void_p), ("point_capacity", C.c_uint64), ("point_count_capacity", C.c_uint64),
                ("root_capacity", C.c_uint64), ("root_uv_capacity", C.c_uint64), ("kept_count", C.c_void_p),
                ("status", C.c_void_p)]


CLASSIC_VAR_ID, CLASSIC_VAR_CLENGTH, CLASSIC_VAR_CWIDTH, CLASSIC_VAR_COLUMN0 = 0, 1, 2, 3
CLASSIC_MAX_COLUMNS = 183
CLASSIC_NEGATIVE_LENGTH, CLASSIC_NONFINITE_TAPER, CLASSIC_NONFINITE_PROFILE = 8, 16, 32
CLASSIC_EXPRESSION_ERROR, CLASSIC_VARIABLE_COUNTS, CLASSIC_NONFINITE_CLUMP = 64, 128, 256
CLASSIC_CLUMP_EXPRS = ("clump_mask", "clump", "clump_scale", "clump_noise", "clump_noise_scale",
                       "clump_noise_frequency", "clump_noise_correlation")
CLASSIC_MAX_CLUMPS = 3
CLASSIC_MAX_NOISES = 2
CLASSIC_MAX_CUTS = 2
CLASSIC_MAX_EFFECTS = 8
# ClassicFloatEffectType numbering (NXG_CLASSIC_EFFECT_*)
CLASSIC_EFFECT_CLUMP, CLASSIC_EFFECT_NOISE, CLASSIC_EFFECT_CUT = 0, 1, 2
CLASSIC_EFFECT_TYPES = {"clump": CLASSIC_EFFECT_CLUMP, "noise": CLASSIC_EFFECT_NOISE, "cut": CLASSIC_EFFECT_CUT}


def clump_key(module, key):
    """The CLASSIC_EXPRS name of clump program `key` of ClumpingFX module `module`."""
    return key if module == 0 else f"{key}@{module}"


def noise_key(module, key):
    """The CLASSIC_EXPRS name of NoiseFX program `key` of NoiseFX module `module`."""
    return key if module == 0 else f"{key}@{module}"


def cut_key(module):
    """The CLASSIC_EXPRS name of CutFX module `module`'s amount."""
    return "cut_amount" if module == 0 else f"cut_amount@{module}"


# NXG_CLASSIC_EXPR_*: module m's clump programs at 6 + 7 m + k, NoiseFX module n's
# at 27 + 6 n + k, the amounts of CutFX modules 1.. at the end
CLASSIC_NOISE_EXPRS = ("noise_mask", "noise_magnitude", "noise_magnitude_scale", "noise_frequency",
                       "noise_correlation", "noise_preserve_length")
CLASSIC_EXPRS = (("length", "width", "cut_amount", "taper", "taper_start", "width_ramp")
                 + tuple(clump_key(m, k) for m in range(CLASSIC_MAX_CLUMPS) for k in CLASSIC_CLUMP_EXPRS)
                 + tuple(noise_key(n, k) for n in range(CLASSIC_MAX_NOISES) for k in CLASSIC_NOISE_EXPRS)
                 + tuple(cut_key(c) for c in range(1, CLASSIC_MAX_CUTS)))
CLASSIC_NONFINITE_NOISE_PARAMETER, CLASSIC_NONFINITE_NOISE = 512, 1024


class ClassicEffect(C.Structure):
    """nxg_classic_effect: an effect type (CLASSIC_EFFECT_*) and its module index."""
    _fields_ = [("type", C.c_uint32), ("module", C.c_uint32)]


class ClassicPlan(C.Structure):
    """nxg_classic_plan: the plan's programs (host nxg_expr_program pointers, NULL: absent)."""
    _fields_ = [("expressions", C.c_void_p * len(CLASSIC_EXPRS)), ("has_cut", C.c_uint32),
                ("clump_count", C.c_uint32), ("description_id", C.c_uint32), ("description_name_seed", C.c_uint32),
                ("clumps_after_cut", C.c_uint32), ("has_noise", C.c_uint32), ("noise_count", C.c_uint32),
                ("cut_count", C.c_uint32), ("effect_count", C.c_uint32),
                ("effects", ClassicEffect * CLASSIC_MAX_EFFECTS)]


class ClassicPlanInfo(C.Structure):
    """nxg_classic_plan_info: what nxg_classic_plan_pack reports."""
    _fields_ = [("byte_size", C.c_uint64), ("present", C.c_uint64), ("has_cut", C.c_uint32),
                ("clump_count", C.c_uint32), ("columns_required", C.c_uint32), ("needs_random_prefix", C.c_uint32),
                ("final_length", C.c_uint32), ("clump_length", C.c_uint32), ("description_id", C.c_uint32),
                ("description_name_seed", C.c_uint32), ("program_offset", C.c_uint32 * len(CLASSIC_EXPRS)),
                ("uniform", C.c_uint32), ("clumps_after_cut", C.c_uint32), ("has_noise", C.c_uint32),
                ("noise_count", C.c_uint32), ("cut_count", C.c_uint32), ("effect_count", C.c_uint32),
                ("effects", ClassicEffect * CLASSIC_MAX_EFFECTS), ("trailing_cut", C.c_uint32),
                ("rewrites", C.c_uint32), ("length_readers", C.c_uint64)]


class ClassicClumpBinding(C.Structure):
    """nxg_classic_clump_binding: one ClumpingFX module's device arrays and host guide struct."""
    _fields_ = [("strand_guides", C.c_void_p), ("guides", C.c_void_p), ("guide_uvs", C.c_void_p),
                ("guide_random_prefixes", C.c_void_p), ("local_axes", C.c_void_p)]


class ClassicRuntimeInputs(C.Structure):
    """nxg_classic_runtime_inputs (device pointers, host plan report, host clump bindings)."""
    _fields_ = [("plan_image", C.c_void_p), ("plan", C.c_void_p), ("primitive_ids", C.c_void_p),
                ("random_prefixes", C.c_void_p), ("columns", C.c_void_p), ("column_count", C.c_uint32),
                ("radius_scale", C.c_float), ("clumps", C.c_void_p), ("surface_tangents", C.c_void_p),
                ("noise_domain_positions", C.c_void_p), ("root_relative", C.c_uint32), ("reserved", C.c_uint32)]


class ExprCompileOptions(C.Structure):
    """nxg_expr_compile_options (XgenExpressionCompileOptions, xgen_expression.h:102-110)."""
    _fields_ = [("expression_name", C.c_char_p), ("object_type", C.c_char_p), ("max_source_bytes", C.c_uint64),
                ("max_tokens", C.c_uint64), ("max_instructions", C.c_uint64), ("max_locals", C.c_uint64),
                ("max_ramp_points", C.c_uint64), ("extensions", C.c_uint32), ("reserved", C.c_uint32)]


class ExprCompiled(C.Structure):
    """nxg_expr_compiled: caller buffers and the compiled counts."""
    _fields_ = [("instructions", C.c_void_p), ("operands", C.c_void_p), ("ramp_points", C.c_void_p),
                ("ramps", C.c_void_p), ("input_names", C.c_void_p), ("instruction_capacity", C.c_uint32),
                ("operand_capacity", C.c_uint32), ("ramp_point_capacity", C.c_uint32), ("ramp_capacity", C.c_uint32),
                ("input_name_capacity", C.c_uint32), ("instruction_count", C.c_uint32), ("operand_count", C.c_uint32),
                ("ramp_point_count", C.c_uint32), ("ramp_count", C.c_uint32), ("input_count", C.c_uint32),
                ("input_name_bytes", C.c_uint32), ("result", C.c_uint32), ("expression_seed", C.c_uint32),
                ("random_call_count", C.c_uint32)]


PAYLOAD_INVALID_COUNTS = 8
PAYLOAD_MAX_MOTION_SAMPLES = 16


class CurveSet(C.Structure):
    """nxg_curve_set (CurveSetView, curve_payload.h:27-38): device arrays."""
    _fields_ = [("point_counts", C.c_void_p), ("points", C.c_void_p), ("widths", C.c_void_p), ("roots", C.c_void_p),
                ("root_uvs", C.c_void_p), ("point_count", C.c_uint64), ("strand_count", C.c_uint32),
                ("fixed_cvs_per_strand", C.c_uint32), ("constant_width", C.c_float), ("time", C.c_float)]


class CurvePayloadOptions(C.Structure):
    """nxg_curve_payload_options (CurvePayloadOptions, curve_payload.h:78-85)."""
    _fields_ = [("max_strands_per_batch", C.c_uint32), ("resample_cv_count", C.c_uint32),
                ("radius_scale", C.c_float), ("transform", C.c_float * 12)]


class CurvePayloadOutput(C.Structure):
    _fields_ = [("points", C.c_void_p), ("point_counts", C.c_void_p), ("root_uvs", C.c_void_p),
                ("motion_points", C.c_void_p), ("batch_points", C.c_void_p), ("status", C.c_void_p),
                ("point_capacity", C.c_uint64), ("point_count_capacity", C.c_uint64),
                ("root_uv_capacity", C.c_uint64), ("motion_point_capacity", C.c_uint64),
                ("batch_capacity", C.c_uint64)]


class CurveCacheHeader(C.Structure):
    """nxg_curve_cache_header (CurveCacheHeader, curve_cache.h:29-47): 128 bytes."""
    _fields_ = [("magic", C.c_uint64), ("version_major", C.c_uint16), ("version_minor", C.c_uint16),
                ("flags", C.c_uint32), ("byte_size", C.c_uint64), ("content_hash", C.c_uint64),
                ("strand_count", C.c_uint32), ("point_count", C.c_uint32), ("motion_sample_count", C.c_uint32),
                ("reserved", C.c_uint32), ("point_counts_offset", C.c_uint64), ("points_offset", C.c_uint64),
                ("motion_times_offset", C.c_uint64), ("motion_points_offset", C.c_uint64),
                ("texcoords_offset", C.c_uint64), ("patch_uvs_offset", C.c_uint64),
                ("face_uvs_offset", C.c_uint64), ("face_ids_offset", C.c_uint64), ("padding", C.c_uint8 * 16)]


class CurveCacheInput(C.Structure):
    """nxg_curve_cache_input (CurveCacheBuildInput, curve_cache.h:56-67)."""
    _fields_ = [("point_counts", C.c_void_p), ("points", C.c_void_p), ("texcoords", C.c_void_p),
                ("patch_uvs", C.c_void_p), ("face_uvs", C.c_void_p), ("face_ids", C.c_void_p),
                ("motion_times", C.c_void_p), ("motion_points", C.c_void_p), ("strand_count", C.c_uint32),
                ("point_count", C.c_uint32), ("motion_sample_count", C.c_uint32), ("reserved", C.c_uint32)]


assert C.sizeof(CurveCacheHeader) == 128

CLUMP_MAX_CVS = 64
CLASSIC_INVALID_GUIDE = 4
INVALID_INDEX = 0xFFFFFFFF


class ClumpGuides(C.Structure):
    """nxg_clump_guides: world-space clump guides (device pointers)."""
    _fields_ = [("axes", C.c_void_p), ("normals", C.c_void_p), ("tangents", C.c_void_p),
                ("reference_positions", C.c_void_p), ("guide_count", C.c_uint32), ("reserved", C.c_uint32)]


class ClumpParams(C.Structure):
    """nxg_clump_params: the ClumpingFX module's constant expressions."""
    _fields_ = [("mask", C.c_float), ("amount", C.c_float), ("clump_scale", C.c_float), ("noise", C.c_float),
                ("noise_scale", C.c_float), ("noise_frequency", C.c_float), ("noise_correlation", C.c_float),
                ("amount_stride", C.c_uint32), ("amount_multiplier", C.c_void_p)]


BATCH_MAX_ITEMS = 12
BATCH_ORDER_GIVEN, BATCH_ORDER_HEAVY_FIRST = 0, 1


class BatchItem(C.Structure):
    """nxg_batch_item: one description's generation request, by pointer."""
    _fields_ = [("asset", C.c_void_p), ("deformed", C.c_void_p), ("params", C.c_void_p), ("octaves", C.c_void_p),
                ("attributes", C.c_void_p), ("attribute_capacity", C.c_uint64), ("exprs", C.c_void_p),
                ("maps", C.c_void_p), ("map_count", C.c_uint32), ("reserved", C.c_uint32), ("output", C.c_void_p),
                ("stencil", C.c_void_p), ("workspace", C.c_void_p), ("workspace_bytes", C.c_uint64),
                ("surrogate", C.c_void_p), ("surrogate_map", C.c_void_p), ("surrogate_threshold", C.c_float),
                ("surrogate_stride", C.c_uint32), ("surrogate_flags", C.c_void_p),
                ("surrogate_flag_capacity", C.c_uint64), ("surrogate_stats", C.c_void_p)]


class BatchTuning(C.Structure):
    _fields_ = [("waves_per_block", C.c_uint32), ("chunk_cvs", C.c_uint32), ("tiles_per_claim", C.c_uint32),
                ("blocks_per_cu", C.c_uint32), ("order", C.c_uint32)]

    def as_tuple(self):
        return tuple(getattr(self, n) for n, _ in self._fields_)

    def __repr__(self):
        return "BatchTuning(" + ", ".join(f"{n}={getattr(self, n)}" for n, _ in self._fields_) + ")"


STATUS = [
    "OK", "NULL_ASSET", "MISALIGNED_ASSET", "INVALID_ASSET_METADATA", "ASSET_CAPACITY_TOO_SMALL",
    "INVALID_GENERATION_PARAMETERS", "INVALID_BLOCK_SIZE", "DEFORMED_POSITION_COUNT_MISMATCH",
    "DEFORMED_NORMAL_COUNT_MISMATCH", "DEFORMED_GUIDE_CV_COUNT_MISMATCH", "MISSING_OUTPUT",
    "POINT_CAPACITY_TOO_SMALL", "WIDTH_CAPACITY_TOO_SMALL", "ROOT_CAPACITY_TOO_SMALL",
    "ROOT_UV_CAPACITY_TOO_SMALL", "POINT_COUNT_CAPACITY_TOO_SMALL", "INVALID_RADIUS_SCALE",
    "MISSING_MOTION_SAMPLES", "INVALID_MOTION_TIMES", "MOTION_POINT_CAPACITY_TOO_SMALL",
    "OUTPUT_SIZE_OVERFLOW", "HIP", "WORKSPACE_TOO_SMALL", "INVALID_OCTAVES", "NULL_ARGUMENT",
    "STENCIL_CAPACITY_TOO_SMALL", "INVALID_ROOTS", "INVALID_BUILD_INPUT", "BUFFER_TOO_SMALL",
    "INVALID_ASSET", "INVALID_PROGRAM", "ATTRIBUTE_CAPACITY_TOO_SMALL", "INVALID_MLP_REQUEST", "INVALID_BATCH",
    "INVALID_CLASSIC_REQUEST", "SCRATCH_LIMIT", "INVALID_SAMPLER",
]
globals().update({name: i for i, name in enumerate(STATUS)})


def ptr(a):
    """Address of a numpy array's data (or None for None)."""
    if a is None:
        return None
    return a.ctypes.data


class BuildInputArrays:
    """Owns contiguous numpy arrays for an AssetBuildInput and exposes the struct."""

    def __init__(self, positions, triangles, guide_cvs, guide_cv_counts, *, normals=None,
                 texcoords=None, guide_root_normals=None, guide_root_uvs=None,
                 guide_triangles=None, guide_barycentrics=None, guide_support_radii=None,
                 stencil_k=0, triangle_weights=None):
        f = lambda a, n: None if a is None else np.ascontiguousarray(a, dtype=np.float32).reshape(-1, n)
        self.positions = f(positions, 3)
        self.triangles = np.ascontiguousarray(triangles, dtype=np.uint32).reshape(-1, 3)
        self.guide_cvs = f(guide_cvs, 3)
        self.guide_cv_counts = np.ascontiguousarray(guide_cv_counts, dtype=np.uint32)
        self.normals = f(normals, 3)
        self.texcoords = f(texcoords, 2)
        self.guide_root_normals = f(guide_root_normals, 3)
        self.guide_root_uvs = f(guide_root_uvs, 2)
        self.guide_triangles = None if guide_triangles is None else np.ascontiguousarray(guide_triangles, dtype=np.uint32)
        self.guide_barycentrics = f(guide_barycentrics, 2)
        self.guide_support_radii = None if guide_support_radii is None else np.ascontiguousarray(guide_support_radii, dtype=np.float32)
        self.triangle_weights = None if triangle_weights is None else np.ascontiguousarray(triangle_weights, dtype=np.float32)
        self.struct = AssetBuildInput(
            positions=ptr(self.positions), vertex_count=len(self.positions),
            triangle_count=len(self.triangles), normals=ptr(self.normals),
            texcoords=ptr(self.texcoords), triangles=ptr(self.triangles),
            guide_count=len(self.guide_cv_counts), stencil_k=stencil_k,
            guide_cv_counts=ptr(self.guide_cv_counts), guide_cvs=ptr(self.guide_cvs),
            guide_root_normals=ptr(self.guide_root_normals), guide_root_uvs=ptr(self.guide_root_uvs),
            guide_triangles=ptr(self.guide_triangles), guide_barycentrics=ptr(self.guide_barycentrics),
            guide_support_radii=ptr(self.guide_support_radii), triangle_weights=ptr(self.triangle_weights))


def header_of(blob: np.ndarray) -> AssetHeader:
    return AssetHeader.from_buffer_copy(blob[:128].tobytes())


def aligned_bytes(size, alignment=64):
    """A zeroed uint8 numpy array of `size` bytes whose data is `alignment`-aligned."""
    raw = np.zeros(size + alignment, dtype=np.uint8)
    off = (-raw.ctypes.data) % alignment
    return raw[off:off + size]
