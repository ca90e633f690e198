"""nanoxgen_amd: MI355X (gfx950) strand generation behind a C ABI.

The compute path is libnanoxgen_amd.so (hand-written HIP kernels plus the
C++ host contract, include/nanoxgen_amd.h).  This module is the reference-
side ctypes binding of that ABI, mirroring the reference's C++ interface
(build_asset / validate_asset / make_device_asset_descriptor /
validate_device_packed_generation_request / generate_packed):
  build_asset(build_input)            -> host .nxg blob (numpy uint8)
  validate_asset(blob)                -> "" or the reference's message
  DeviceAsset(blob, device)           -> uploaded blob + checked descriptor
  generate_packed(asset, params, ...) -> device outputs, on the given stream
PyTorch provides device memory and streams only.  There is no CPU fallback:
loading fails loudly if the library is missing.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .abi import GenerationParams, NoiseOctaves  # noqa: F401  (re-export)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnanoxgen_amd.so")
_LIB = None

# every symbol of include/nanoxgen_amd.h: (name, restype, argtypes)
_P = C.c_void_p
SYMBOLS = [
    ("nxg_status_message", C.c_char_p, [C.c_uint32]),
    ("nxg_asset_build_size", C.c_uint32, [_P, _P]),
    ("nxg_asset_build", C.c_uint32, [_P, _P, C.c_uint64, _P, C.c_char_p, C.c_uint64]),
    ("nxg_asset_build_gpu", C.c_uint32, [_P, _P, C.c_uint64, _P, C.c_char_p, C.c_uint64, _P, C.c_uint64, _P]),
    ("nxg_asset_validate", C.c_uint32, [_P, C.c_uint64, C.c_char_p, C.c_uint64]),
    ("nxg_make_device_asset", C.c_uint32, [_P, C.c_uint64, _P, C.c_uint64, _P]),
    ("nxg_validate_packed_generation", C.c_uint32, [_P, _P, _P, _P, _P]),
    ("nxg_packed_generation_workspace_size", C.c_uint64, [_P, _P]),
    ("nxg_generate_packed", C.c_uint32, [_P, _P, _P, _P, _P, _P, _P, _P, C.c_uint64, _P]),
    ("nxg_generate_packed_noalloc", C.c_uint32, [_P, _P, _P, _P, _P, _P, _P, _P]),
    ("nxg_generate_packed_from_roots", C.c_uint32, [_P, _P, _P, _P, _P, _P, _P, C.c_uint64, _P]),
    ("nxg_cdf_sampler_size", C.c_uint64, [_P]),
    ("nxg_cdf_sampler_build", C.c_uint32, [_P, _P, C.c_uint64, _P]),
    ("nxg_generate_packed_cdf", C.c_uint32, [_P, _P, _P, _P, _P, C.c_uint64, _P, _P, _P, _P, C.c_uint64, _P]),
    ("nxg_prepare_packed_generation", C.c_uint32, [_P, _P, _P, _P, C.c_uint64, _P]),
    ("nxg_generate_packed_prepared", C.c_uint32, [_P, _P, _P, _P, _P, _P, _P, _P, C.c_uint64, _P]),
    ("nxg_expr_validate", C.c_uint32, [_P]),
    ("nxg_expr_compile", C.c_uint32, [C.c_char_p, C.c_uint64, _P, _P, C.c_char_p, C.c_uint64]),
    ("nxg_seexpr_hash", C.c_double, [_P, C.c_uint32]),
    ("nxg_string_seed", C.c_uint32, [C.c_char_p]),
    ("nxg_face_seed", C.c_double, [C.c_uint32, C.c_char_p, C.c_uint32]),
    ("nxg_expr_pack", C.c_uint32, [_P, C.c_uint32, _P, C.c_uint64, _P]),
    ("nxg_evaluate_root_attributes", C.c_uint32, [_P, _P, _P, _P, C.c_uint32, _P, C.c_uint64, _P, _P]),
    ("nxg_generate_packed_attributed", C.c_uint32, [_P, _P, _P, _P, _P, C.c_uint64, _P, _P, _P, _P, C.c_uint64, _P]),
    ("nxg_generate_packed_attributed_prepared", C.c_uint32,
     [_P, _P, _P, _P, _P, C.c_uint64, _P, _P, _P, _P, C.c_uint64, _P]),
    ("nxg_mlp_get_layout", None, [_P]),
    ("nxg_mlp_sync", C.c_uint32, [_P, _P]),
    ("nxg_mlp_gradients", C.c_uint32, [_P, _P, _P, C.c_uint64, C.c_uint64, C.c_uint32, _P]),
    ("nxg_mlp_train_step", C.c_uint32, [_P, _P, _P, C.c_uint64, C.c_uint64, C.c_uint32, C.c_float, C.c_uint32, _P]),
    ("nxg_mlp_forward", C.c_uint32, [_P, _P, _P, C.c_uint64, _P]),
    ("nxg_root_features", C.c_uint32, [_P, _P, _P, _P, C.c_uint64, _P]),
    ("nxg_surrogate_attributes", C.c_uint32, [_P, _P, _P, _P, _P, C.c_uint64, _P, C.c_uint64, _P]),
    ("nxg_verify_attributes", C.c_uint32, [_P, _P, _P, _P, C.c_uint32, _P, C.c_uint64, C.c_float, C.c_uint32,
                                           _P, C.c_uint64, _P, _P]),
    ("nxg_validate_motion_generation", C.c_uint32, [_P, _P, C.c_uint32, _P, _P, _P]),
    ("nxg_motion_generation_workspace_size", C.c_uint64, [_P, _P, C.c_uint32]),
    ("nxg_generate_motion", C.c_uint32, [_P, _P, C.c_uint32, _P, _P, _P, _P, _P, C.c_uint64, _P]),
    ("nxg_batch_default_tuning", None, [_P]),
    ("nxg_batch_validate", C.c_uint32, [_P, C.c_uint32, _P]),
    ("nxg_generate_batch", C.c_uint32, [_P, C.c_uint32, _P, _P, _P]),
    ("nxg_batch_graph_create", C.c_uint32, [_P, C.c_uint32, _P, _P, _P]),
    ("nxg_batch_graph_launch", C.c_uint32, [_P, _P]),
    ("nxg_batch_graph_destroy", C.c_uint32, [_P]),
    ("nxg_batch_tuning_candidates", C.c_uint32, [_P, C.c_uint32]),
    ("nxg_batch_autotune", C.c_uint32, [_P, C.c_uint32, _P, _P, _P, C.c_uint32, C.c_uint32, _P, _P]),
    ("nxg_validate_classic_cut", C.c_uint32, [_P, _P, _P]),
    ("nxg_classic_cut_workspace_size", C.c_uint64, [C.c_uint32]),
    ("nxg_classic_cut_apply", C.c_uint32, [_P, _P, _P, _P, C.c_uint64, _P]),
    ("nxg_classic_add_endpoints", C.c_uint32, [_P, C.c_uint32, C.c_uint32, _P, _P, C.c_uint64, _P, C.c_uint64, _P]),
    ("nxg_classic_clump_workspace_size", C.c_uint64, [C.c_uint32, C.c_uint32]),
    ("nxg_classic_clump_apply", C.c_uint32, [_P, C.c_uint32, C.c_uint32, _P, _P, _P, _P, C.c_uint64, _P, _P,
                                             C.c_uint64, _P]),
    ("nxg_classic_plan_pack", C.c_uint32, [_P, _P, C.c_uint64, _P]),
    ("nxg_validate_classic_runtime", C.c_uint32, [_P, _P, _P]),
    ("nxg_classic_runtime_workspace_size", C.c_uint64, [_P, C.c_uint32, C.c_uint32, C.c_uint32]),
    ("nxg_classic_runtime_apply", C.c_uint32, [_P, _P, _P, _P, C.c_uint64, _P]),
    ("nxg_curve_payload_plan", C.c_uint32, [_P, _P, C.c_uint32, _P, _P, _P, C.c_char_p, C.c_uint64]),
    ("nxg_curve_payload_workspace_size", C.c_uint64, [C.c_uint32, C.c_uint32]),
    ("nxg_curve_payload_build", C.c_uint32, [_P, _P, C.c_uint32, _P, _P, _P, C.c_uint64, _P]),
    ("nxg_curve_cache_layout", C.c_uint32, [_P, _P]),
    ("nxg_curve_cache_build", C.c_uint32, [_P, _P, C.c_uint64, _P, C.c_char_p, C.c_uint64]),
    ("nxg_curve_cache_validate", C.c_uint32, [_P, C.c_uint64, C.c_char_p, C.c_uint64]),
    ("nxg_curve_cache_build_gpu", C.c_uint32, [_P, _P, C.c_uint64, _P, C.c_uint64, _P, C.c_char_p, C.c_uint64,
                                               _P]),
    ("nxg_build_info", C.c_char_p, []),
]


class NxgError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        msg = lib().nxg_status_message(status).decode()
        super().__init__(f"{where}: status {status} ({abi.STATUS[status] if status < len(abi.STATUS) else '?'}): {msg}")


def lib():
    """The loaded libnanoxgen_amd.so (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "or `make -C nanoxgen_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SYMBOLS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(status, where):
    if status != 0:
        raise NxgError(status, where)


def build_asset(build_input):
    """build_asset (asset.cpp:229-293) -> 64-byte aligned numpy uint8 blob."""
    L = lib()
    size = C.c_uint64()
    st = L.nxg_asset_build_size(C.byref(build_input.struct), C.byref(size))
    err = C.create_string_buffer(512)
    if st == abi.INVALID_BUILD_INPUT:
        L.nxg_asset_build(C.byref(build_input.struct), C.c_void_p(64), 0, None, err, 512)
        raise ValueError(err.value.decode())
    check(st, "nxg_asset_build_size")
    blob = abi.aligned_bytes(size.value)
    st = L.nxg_asset_build(C.byref(build_input.struct), blob.ctypes.data, blob.nbytes, None, err, 512)
    if st != 0:
        raise ValueError(err.value.decode() or L.nxg_status_message(st).decode())
    return blob


def build_asset_gpu(build_input, device="cuda", stream=None):
    """build_asset with the stencil search on the GPU (nxg_asset_build_gpu):
    returns (host blob, DeviceAsset of the device copy); both are the bytes
    build_asset writes."""
    torch = _torch()
    L = lib()
    size = C.c_uint64()
    check(L.nxg_asset_build_size(C.byref(build_input.struct), C.byref(size)), "nxg_asset_build_size")
    blob = abi.aligned_bytes(size.value)
    raw = torch.empty(size.value + 64, dtype=torch.uint8, device=device)
    off = (-raw.data_ptr()) % 64
    dev = raw[off:off + size.value]
    err = C.create_string_buffer(512)
    st = L.nxg_asset_build_gpu(C.byref(build_input.struct), blob.ctypes.data, blob.nbytes, None, err, 512,
                               C.c_void_p(dev.data_ptr()), dev.numel(), _stream_handle(stream))
    if st != 0:
        raise ValueError(err.value.decode() or L.nxg_status_message(st).decode())
    return blob, DeviceAsset(blob, device=device, uploaded=(raw, dev))


def validate_asset(blob):
    """validate_asset (asset.cpp:295-338): "" when valid, else the message."""
    msg = C.create_string_buffer(512)
    lib().nxg_asset_validate(blob.ctypes.data, blob.nbytes, msg, 512)
    return msg.value.decode()


def _torch():
    import torch
    return torch


def _stream_handle(stream):
    torch = _torch()
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


class DeviceAsset:
    """A blob uploaded to device memory with its checked descriptor
    (make_device_asset_descriptor, asset.cpp:364-378)."""

    def __init__(self, blob, device="cuda", uploaded=None):
        torch = _torch()
        msg = validate_asset(blob)
        if msg:
            raise ValueError(f"invalid asset: {msg}")
        self.blob = blob
        self.header = abi.header_of(blob)
        if uploaded is not None:  # (storage, 64-byte aligned view) already holding the blob
            self.storage, self.tensor = uploaded
        else:
            # over-allocate so the base can be moved to a 64-byte boundary
            raw = torch.empty(blob.nbytes + 64, dtype=torch.uint8, device=device)
            off = (-raw.data_ptr()) % 64
            self.storage = raw
            self.tensor = raw[off:off + blob.nbytes]
            self.tensor.copy_(torch.from_numpy(blob))
        self.desc = abi.DeviceAsset()
        check(lib().nxg_make_device_asset(blob.ctypes.data, blob.nbytes, self.tensor.data_ptr(),
                                          self.tensor.numel(), C.byref(self.desc)),
              "nxg_make_device_asset")

    def workspace_bytes(self, params):
        return lib().nxg_packed_generation_workspace_size(C.byref(self.desc.header), C.byref(params))


class CdfSampler:
    """The opt-in inverse-CDF root sampler of a device asset
    (nxg_cdf_sampler_build): the alias table's distribution as u64 prefix
    sums, drawn by binary search.  Roots drawn through it differ from the
    alias draw's (generate.h:107-112) and match it in distribution only."""

    def __init__(self, asset, stream=None):
        torch = _torch()
        self.asset = asset
        self.nbytes = lib().nxg_cdf_sampler_size(C.byref(asset.desc.header))
        if self.nbytes == 0:
            raise ValueError("asset has no triangles")
        raw = torch.empty(self.nbytes + 256, dtype=torch.uint8, device=asset.tensor.device)
        off = (-raw.data_ptr()) % 256
        self.storage, self.tensor = raw, raw[off:off + self.nbytes]
        check(lib().nxg_cdf_sampler_build(C.byref(asset.desc), C.c_void_p(self.tensor.data_ptr()), self.nbytes,
                                          _stream_handle(stream)), "nxg_cdf_sampler_build")

    def prefix_sums(self):
        """The T inclusive prefix sums (numpy uint64; synchronises)."""
        t = self.asset.header.triangle_count
        return self.tensor[:8 * t].cpu().numpy().view(np.uint64)


class PackedOutputs:
    """Caller-owned device outputs of one request (torch tensors)."""

    def __init__(self, strands, cvs, device="cuda", roots=True, root_uvs=True, point_counts=True,
                 stencil=False, radius_scale=1.0):
        torch = _torch()
        self.points = torch.empty((strands * cvs, 4), dtype=torch.float32, device=device)
        self.roots = torch.empty((strands, 12), dtype=torch.int32, device=device) if roots else None
        self.root_uvs = torch.empty((strands, 2), dtype=torch.float32, device=device) if root_uvs else None
        self.point_counts = torch.empty(strands, dtype=torch.int32, device=device) if point_counts else None
        self.stencil_indices = torch.empty((strands, 8), dtype=torch.int32, device=device) if stencil else None
        self.stencil_weights = torch.empty((strands, 8), dtype=torch.float32, device=device) if stencil else None
        self.radius_scale = radius_scale
        self.strands = strands
        self.cvs = cvs

    def descriptor(self):
        t = lambda x: 0 if x is None else x.data_ptr()
        n = lambda x, k: 0 if x is None else x.numel() // k
        return abi.PackedOutput(points=t(self.points), roots=t(self.roots) or None,
                                root_uvs=t(self.root_uvs) or None, radius_scale=self.radius_scale,
                                point_counts=t(self.point_counts) or None,
                                point_capacity=n(self.points, 4), root_capacity=n(self.roots, 12),
                                root_uv_capacity=n(self.root_uvs, 2),
                                point_count_capacity=n(self.point_counts, 1))

    def stencil_descriptor(self):
        if self.stencil_indices is None:
            return None
        return abi.StencilOutput(indices=self.stencil_indices.data_ptr(),
                                 weights=self.stencil_weights.data_ptr(),
                                 capacity=self.stencil_indices.numel())

    def to_host(self):
        """Copy to numpy (synchronises)."""
        out = dict(points=self.points.cpu().numpy())
        if self.roots is not None:
            out["roots"] = self.roots.cpu().numpy().view(abi.ROOT_DTYPE).reshape(-1)
        if self.root_uvs is not None:
            out["root_uvs"] = self.root_uvs.cpu().numpy()
        if self.point_counts is not None:
            out["point_counts"] = self.point_counts.cpu().numpy().view(np.uint32)
        if self.stencil_indices is not None:
            out["stencil_indices"] = self.stencil_indices.cpu().numpy().view(np.uint32)
            out["stencil_weights"] = self.stencil_weights.cpu().numpy()
        return out


def _ref(x):
    return None if x is None else C.byref(x)


def generate_packed(asset, params, outputs, *, octaves=None, deformed=None, workspace=None,
                    stream=None, roots=None, attributes=None, sampler=None):
    """Enqueue generate_packed (generate.h:612-676) for every strand on
    `stream` (torch stream; default the current one).  `roots` (optional
    device tensor of RootSample records) selects generate-from-roots;
    `attributes` (optional float32 [N, 4] device tensor) the per-root
    attribute multipliers; `sampler` (a CdfSampler of the asset) the
    opt-in inverse-CDF root draw, whose roots land in outputs.roots."""
    torch = _torch()
    L = lib()
    if workspace is None:
        workspace = torch.empty(max(1, asset.workspace_bytes(params)), dtype=torch.uint8,
                                device=asset.tensor.device)
    out = outputs.descriptor()
    st = outputs.stencil_descriptor()
    if attributes is not None:
        status = L.nxg_generate_packed_attributed(
            C.byref(asset.desc), _ref(deformed), C.byref(params), _ref(octaves),
            C.c_void_p(attributes.data_ptr()), attributes.numel(), C.byref(out), _ref(st), None,
            C.c_void_p(workspace.data_ptr()), workspace.numel(), _stream_handle(stream))
    elif sampler is not None:
        status = L.nxg_generate_packed_cdf(
            C.byref(asset.desc), _ref(deformed), C.byref(params), _ref(octaves),
            C.c_void_p(sampler.tensor.data_ptr()), sampler.nbytes, C.byref(out), _ref(st), None,
            C.c_void_p(workspace.data_ptr()), workspace.numel(), _stream_handle(stream))
    elif roots is not None:
        status = L.nxg_generate_packed_from_roots(
            C.byref(asset.desc), C.byref(params), _ref(octaves), C.c_void_p(roots.data_ptr()),
            C.byref(out), None, C.c_void_p(workspace.data_ptr()), workspace.numel(),
            _stream_handle(stream))
    else:
        status = L.nxg_generate_packed(
            C.byref(asset.desc), _ref(deformed), C.byref(params), _ref(octaves), C.byref(out),
            _ref(st), None, C.c_void_p(workspace.data_ptr()), workspace.numel(),
            _stream_handle(stream))
    check(status, "nxg_generate_packed")
    return outputs


def generate_motion(asset, params, samples, points, *, octaves=None, workspace=None, stream=None):
    """Enqueue motion-sample generation: `samples` is a list of
    (time, DeformedGeometry); `points` a float32 [S * N * C, 3] device tensor
    receiving sample-major positions."""
    torch = _torch()
    arr = (abi.MotionSample * len(samples))(*[abi.MotionSample(deformation=d, time=t) for t, d in samples])
    out = abi.MotionOutput(points=points.data_ptr(), point_capacity=points.shape[0])
    if workspace is None:
        n = lib().nxg_motion_generation_workspace_size(C.byref(asset.desc.header), C.byref(params), len(samples))
        workspace = torch.empty(max(1, n), dtype=torch.uint8, device=asset.tensor.device)
    check(lib().nxg_generate_motion(C.byref(asset.desc), arr, len(samples), C.byref(params), _ref(octaves),
                                    C.byref(out), None, C.c_void_p(workspace.data_ptr()), workspace.numel(),
                                    _stream_handle(stream)), "nxg_generate_motion")
    return points


def prepare_packed(asset, params, workspace, *, deformed=None, stream=None):
    """First half of generate_packed: guide state + offset table into workspace."""
    check(lib().nxg_prepare_packed_generation(C.byref(asset.desc), _ref(deformed), C.byref(params),
                                              C.c_void_p(workspace.data_ptr()), workspace.numel(),
                                              _stream_handle(stream)), "nxg_prepare_packed_generation")


def generate_packed_prepared(asset, params, outputs, workspace, *, octaves=None, deformed=None, stream=None,
                             attributes=None):
    """Second half of generate_packed: the strand kernel from a prepared
    workspace (`attributes`: the per-root attribute multipliers)."""
    out = outputs.descriptor()
    st = outputs.stencil_descriptor()
    if attributes is not None:
        check(lib().nxg_generate_packed_attributed_prepared(
            C.byref(asset.desc), _ref(deformed), C.byref(params), _ref(octaves), C.c_void_p(attributes.data_ptr()),
            attributes.numel(), C.byref(out), _ref(st), None, C.c_void_p(workspace.data_ptr()), workspace.numel(),
            _stream_handle(stream)), "nxg_generate_packed_attributed_prepared")
        return outputs
    check(lib().nxg_generate_packed_prepared(C.byref(asset.desc), _ref(deformed), C.byref(params), _ref(octaves),
                                             C.byref(out), _ref(st), None, C.c_void_p(workspace.data_ptr()),
                                             workspace.numel(), _stream_handle(stream)),
          "nxg_generate_packed_prepared")
    return outputs
