"""The .nxc curve cache (host side of nxg_curve_cache_*).

The reference writes finished curves as a pointer-free, uploadable blob
(include/nanoxgen/curve_cache.h: build_curve_cache, validate_curve_cache,
save_curve_cache, load_curve_cache).  build_curve_cache here takes numpy
arrays (host build) or torch device tensors (the sections are assembled on
the GPU and leave the device in one download; the serial content hash and
the checks run on the host copy).  Both give the reference's bytes.
"""
import ctypes as C

import numpy as np

from . import abi, lib, _stream_handle, _torch


def _is_device(x):
    return x is not None and hasattr(x, "data_ptr") and getattr(x, "is_cuda", False)


def _input(counts, points, texcoords, patch_uvs, face_uvs, face_ids, motion_times, motion_points):
    keep = []

    def ptr(x, dtype):
        if x is None:
            return None
        if _is_device(x):
            if not x.is_contiguous():
                raise ValueError("device arrays must be contiguous")
            return x.data_ptr()
        a = np.ascontiguousarray(x, dtype)
        keep.append(a)
        return a.ctypes.data

    n = counts.shape[0]
    p = points.shape[0] if points.ndim == 2 else points.shape[0] // 4
    times = None if motion_times is None else np.ascontiguousarray(motion_times, np.float32)
    if times is not None:
        keep.append(times)
    inp = abi.CurveCacheInput(point_counts=ptr(counts, np.uint32), points=ptr(points, np.float32),
                              texcoords=ptr(texcoords, np.float32), patch_uvs=ptr(patch_uvs, np.float32),
                              face_uvs=ptr(face_uvs, np.float32), face_ids=ptr(face_ids, np.uint32),
                              motion_times=None if times is None else times.ctypes.data,
                              motion_points=ptr(motion_points, np.float32), strand_count=n, point_count=p,
                              motion_sample_count=0 if times is None else len(times))
    return inp, keep


def build_curve_cache(counts, points, *, texcoords=None, patch_uvs=None, face_uvs=None, face_ids=None,
                      motion_times=None, motion_points=None, stream=None):
    """build_curve_cache (src/curve_cache.cpp:104-164) -> 64-byte aligned
    uint8 numpy blob.  Device tensors (torch) take the GPU path; motion_times
    is always a host array.  Raises ValueError with the reference's message."""
    inp, keep = _input(counts, points, texcoords, patch_uvs, face_uvs, face_ids, motion_times, motion_points)
    h = abi.CurveCacheHeader()
    st = lib().nxg_curve_cache_layout(C.byref(inp), C.byref(h))
    err = C.create_string_buffer(512)
    size = C.c_uint64()
    if st != 0:
        lib().nxg_curve_cache_build(C.byref(inp), None, 0, C.byref(size), err, 512)
        raise ValueError(err.value.decode())
    blob = abi.aligned_bytes(h.byte_size)
    if _is_device(points):
        torch = _torch()
        raw = torch.empty(h.byte_size + 64, dtype=torch.uint8, device=points.device)
        off = (-raw.data_ptr()) % 64
        st = lib().nxg_curve_cache_build_gpu(C.byref(inp), C.c_void_p(raw.data_ptr() + off), h.byte_size,
                                             blob.ctypes.data, blob.nbytes, C.byref(size), err, 512,
                                             _stream_handle(stream))
    else:
        st = lib().nxg_curve_cache_build(C.byref(inp), blob.ctypes.data, blob.nbytes, C.byref(size), err, 512)
    del keep
    if st != 0:
        raise ValueError(err.value.decode() or lib().nxg_status_message(st).decode())
    return blob


def validate_curve_cache(blob):
    """validate_curve_cache (src/curve_cache.cpp:166-239): "" or the message."""
    msg = C.create_string_buffer(512)
    lib().nxg_curve_cache_validate(blob.ctypes.data, blob.nbytes, msg, 512)
    return msg.value.decode()


def header_of(blob):
    return abi.CurveCacheHeader.from_buffer_copy(blob[:128].tobytes())
