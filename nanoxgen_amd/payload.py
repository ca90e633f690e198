"""Renderer curve payload on the GPU (host side of nxg_curve_payload_*).

build_curve_batches (include/nanoxgen/curve_payload.h:87-92) cuts curves
into batches of at most 64K strands for renderer APIs, transforms points,
turns widths into radii and optionally resamples curves and motion samples.
Here all batches are written back to back into device buffers; batch b is
strands [b * max, ...) and starts at point batch_points[b].
"""
import ctypes as C

import numpy as np

from . import abi, check, lib, _stream_handle, _torch

IDENTITY = (1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0)


def _set(d, strands):
    t = lambda x: None if x is None else x.data_ptr()
    pts = d["points"]
    for name in ("points", "counts", "widths", "root_uvs", "roots"):
        x = d.get(name)
        if x is not None and not x.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
    return abi.CurveSet(point_counts=t(d.get("counts")), points=pts.data_ptr(), widths=t(d.get("widths")),
                        roots=t(d.get("roots")), root_uvs=t(d.get("root_uvs")), point_count=pts.shape[0],
                        strand_count=strands, fixed_cvs_per_strand=int(d.get("fixed", 0)),
                        constant_width=float(d.get("constant_width", 0.0)), time=float(d.get("time", 0.0)))


def plan(base, motion=(), *, max_strands_per_batch=65536, resample_cv_count=0, radius_scale=1.0, transform=IDENTITY):
    """(batch count, total output points); raises ValueError with the reference's message."""
    n = _strands(base)
    b = _set(base, n)
    ms = (abi.CurveSet * max(1, len(motion)))(*[_set(m, _strands(m)) for m in motion])
    opt = abi.CurvePayloadOptions(max_strands_per_batch, resample_cv_count, radius_scale, (C.c_float * 12)(*transform))
    batches, total = C.c_uint32(), C.c_uint64()
    err = C.create_string_buffer(512)
    st = lib().nxg_curve_payload_plan(C.byref(b), ms, len(motion), C.byref(opt), C.byref(batches), C.byref(total),
                                      err, 512)
    if st != 0:
        raise ValueError(err.value.decode() or lib().nxg_status_message(st).decode())
    return batches.value, total.value


def _strands(d):
    if d.get("counts") is not None:
        return d["counts"].shape[0]
    return d["points"].shape[0] // int(d["fixed"])


def curve_payload(base, motion=(), *, max_strands_per_batch=65536, resample_cv_count=0, radius_scale=1.0,
                  transform=IDENTITY, stream=None, out=None):
    """Enqueue build_curve_batches on device curves.  base and each motion
    sample: dict of device tensors points [P, 3] float32, counts [N] int32 (or
    "fixed": CVs per curve), optional widths [P], root_uvs [N, 2], roots
    [N, 12]; "constant_width"; motion samples also "time".  Returns a dict of
    device tensors (points [T, 4], point_counts [N], root_uvs or None,
    motion_points [S * T, 3], batch_points [B + 1] int64, status [1]).
    out: a dict an earlier call of the same shapes returned, whose buffers
    (and workspace) this call reuses instead of allocating."""
    torch = _torch()
    n = _strands(base)
    batches, total = plan(base, motion, max_strands_per_batch=max_strands_per_batch,
                          resample_cv_count=resample_cv_count, radius_scale=radius_scale, transform=transform)
    dev = base["points"].device
    s = len(motion)
    has_uvs = base.get("root_uvs") is not None or base.get("roots") is not None
    if out is not None:
        if (out["points"].shape[0] != max(total, 1) or out["point_counts"].numel() != max(n, 1) or
                out["motion_points"].shape[0] != max(total * s, 1) or out["batch_points"].numel() != batches + 1 or
                (out["root_uvs"] is not None) != has_uvs):
            raise ValueError("curve_payload: out was made for other shapes")
    else:
        out = dict(points=torch.empty((max(total, 1), 4), dtype=torch.float32, device=dev),
                   point_counts=torch.empty(max(n, 1), dtype=torch.int32, device=dev),
                   root_uvs=torch.empty((max(n, 1), 2), dtype=torch.float32, device=dev) if has_uvs else None,
                   motion_points=torch.empty((max(total * s, 1), 3), dtype=torch.float32, device=dev),
                   batch_points=torch.empty(batches + 1, dtype=torch.int64, device=dev),
                   status=torch.zeros(1, dtype=torch.int32, device=dev))
    t = lambda x: None if x is None else x.data_ptr()
    o = abi.CurvePayloadOutput(points=out["points"].data_ptr(), point_counts=out["point_counts"].data_ptr(),
                               root_uvs=t(out["root_uvs"]), motion_points=t(out["motion_points"]) if s else None,
                               batch_points=out["batch_points"].data_ptr(), status=out["status"].data_ptr(),
                               point_capacity=total, point_count_capacity=n, root_uv_capacity=n if has_uvs else 0,
                               motion_point_capacity=total * s, batch_capacity=batches + 1)
    ws = out.get("_workspace")
    if ws is None:
        ws = torch.empty(int(lib().nxg_curve_payload_workspace_size(n, s)), dtype=torch.uint8, device=dev)
    b = _set(base, n)
    ms = (abi.CurveSet * max(1, s))(*[_set(m, _strands(m)) for m in motion])
    opt = abi.CurvePayloadOptions(max_strands_per_batch, resample_cv_count, radius_scale, (C.c_float * 12)(*transform))
    check(lib().nxg_curve_payload_build(C.byref(b), ms, s, C.byref(opt), C.byref(o), C.c_void_p(ws.data_ptr()),
                                        ws.numel(), _stream_handle(stream)), "nxg_curve_payload_build")
    out["_workspace"] = ws
    return out


def batches(out, max_strands_per_batch, strands):
    """Host list of (first_strand, strand_count, first_point, point_count) (synchronises)."""
    bp = out["batch_points"].cpu().numpy().astype(np.int64)
    return [(b * max_strands_per_batch, min(max_strands_per_batch, strands - b * max_strands_per_batch),
             int(bp[b]), int(bp[b + 1] - bp[b])) for b in range(len(bp) - 1)]
