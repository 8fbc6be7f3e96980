"""TSDF volumes on disk (vf_tsdf_* in include/volumefusion.h): the native
layout and the reference's TSDFVolume::save layout (src/TSDFVolume.cpp:339-358)
with its 24-byte size field fixed."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

NATIVE, LEGACY = L.TSDF_FORMAT_NATIVE, L.TSDF_FORMAT_LEGACY


def _info(dim, size, center=(0.0, 0.0, 0.0), delta=0.01, max_weight=15.0) -> L.TsdfInfo:
    i = L.TsdfInfo()
    i.dim = L.i3(dim)
    i.size = L.f3(size)
    i.center = L.f3(center)
    i.delta = float(delta)
    i.max_weight = float(max_weight)
    return i


def _as_dict(i: L.TsdfInfo) -> dict:
    return {"dim": (i.dim.x, i.dim.y, i.dim.z), "size": (i.size.x, i.size.y, i.size.z),
            "center": (i.center.x, i.center.y, i.center.z), "delta": i.delta, "max_weight": i.max_weight,
            "format": int(i.format)}


def save_host(path: str, tsdf: np.ndarray, weights: np.ndarray, dim, size, center=(0.0, 0.0, 0.0), delta=0.01,
              max_weight=15.0, fmt: int = NATIVE) -> None:
    t = np.ascontiguousarray(tsdf, np.float32).ravel()
    w = np.ascontiguousarray(weights, np.float32).ravel()
    n = int(dim[0]) * int(dim[1]) * int(dim[2])
    if t.size != n or w.size != n:
        raise ValueError(f"volume arrays hold {t.size}/{w.size} voxels, dims say {n}")
    info = _info(dim, size, center, delta, max_weight)
    L.call("vf_tsdf_save_host", str(path).encode(), int(fmt), C.byref(info), t.ctypes.data, w.ctypes.data)


def read_info(path: str) -> dict:
    info = L.TsdfInfo()
    L.call("vf_tsdf_load_host", str(path).encode(), C.byref(info), None, None, 0)
    return _as_dict(info)


def load_host(path: str) -> tuple:
    """(info dict, tsdf float32[N], weights float32[N])."""
    info = read_info(path)
    n = info["dim"][0] * info["dim"][1] * info["dim"][2]
    t = np.empty(n, np.float32)
    w = np.empty(n, np.float32)
    i = L.TsdfInfo()
    L.call("vf_tsdf_load_host", str(path).encode(), C.byref(i), t.ctypes.data, w.ctypes.data, n)
    return _as_dict(i), t, w


def save_device(path: str, tsdf, fmt: int = NATIVE, max_weight: float = 15.0) -> None:
    """A DeviceTsdf (volumefusion_amd.ops) to `path`."""
    info = _info(tsdf.vol_dim, tsdf.vol_size, tsdf.center, tsdf.delta, max_weight)
    L.call("vf_tsdf_save", str(path).encode(), int(fmt), C.byref(info), tsdf.tsdf.ptr, tsdf.weights.ptr)


def load_device(path: str, tsdf) -> dict:
    """Into an existing DeviceTsdf of at least the file's voxel count."""
    info = L.TsdfInfo()
    L.call("vf_tsdf_load", str(path).encode(), C.byref(info), tsdf.tsdf.ptr, tsdf.weights.ptr, tsdf.voxel_count())
    return _as_dict(info)
