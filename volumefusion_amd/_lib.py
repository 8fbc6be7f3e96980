"""ctypes binding of libvolumefusion_amd.so (the C ABI in include/volumefusion.h).

The library is loaded from this package directory only; a missing library is
an error (there is no fallback path).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libvolumefusion_amd.so"


class VFError(RuntimeError):
    """A C-ABI call returned a non-zero status (the reference throws
    std::runtime_error from CUDA_CHECK, src/Helper.cu:7-20)."""

    def __init__(self, status: int, func: str, message: str):
        super().__init__(f"{func}: {message} (status {status})")
        self.status = status


class Float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Int3(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("z", C.c_int32)]


class Int2(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32)]


class Mat3(C.Structure):
    _fields_ = [("m", C.c_float * 9)]


class Mat4(C.Structure):
    _fields_ = [("m", C.c_float * 16)]


class Volume(C.Structure):
    _fields_ = [
        ("tsdf", C.c_void_p),
        ("weights", C.c_void_p),
        ("dim", Int3),
        ("voxel_size", Float3),
        ("size", Float3),
        ("center", Float3),
    ]


class Graph(C.Structure):
    _fields_ = [
        ("points", C.c_void_p),
        ("points_out", C.c_void_p),
        ("transforms", C.c_void_p),
        ("radii", C.c_void_p),
        ("count", C.c_size_t),
    ]


class Frame(C.Structure):
    _fields_ = [("depth", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32), ("intrinsics", Mat3)]


class WarpParams(C.Structure):
    _fields_ = [("k", C.c_int), ("delta", C.c_float), ("max_weight", C.c_float), ("voxels_per_node", Int3),
                ("blend", C.c_int32)]


BLEND_LBS, BLEND_DQB = 0, 1


class FrameStats(C.Structure):
    _fields_ = [
        ("supported", C.c_uint64),
        ("in_bounds", C.c_uint64),
        ("pending", C.c_uint64),
        ("collided", C.c_uint64),
        ("updated", C.c_uint64),
        ("active_bricks", C.c_uint64),
        ("overflow_bricks", C.c_uint64),
        ("screened_bricks", C.c_uint64),
        ("redo_bricks", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [("brick_shape", C.c_int32), ("wg_per_cu", C.c_int32), ("ms_per_frame", C.c_float),
                ("flags", C.c_int32)]


TUNE_NO_SCREEN = 1  # VF_TUNE_NO_SCREEN


BRICK_SHAPES = [(16, 8, 4), (16, 4, 4), (8, 8, 4), (8, 8, 8), (32, 4, 4)]  # VF_BRICK_* order
PLAN_GRAPH = 1


class TsdfInfo(C.Structure):
    _fields_ = [("dim", Int3), ("size", Float3), ("center", Float3), ("delta", C.c_float), ("max_weight", C.c_float),
                ("format", C.c_int32)]


TSDF_FORMAT_NATIVE, TSDF_FORMAT_LEGACY = 1, 2


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("k", C.c_int32), ("max_batch", C.c_int32), ("lanes", C.c_int32),
                ("device", C.c_int32)]


class StreamDesc(C.Structure):
    _fields_ = [
        ("dim", Int3),
        ("size", Float3),
        ("center", Float3),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("intrinsics", Mat3),
        ("delta", C.c_float),
        ("max_weight", C.c_float),
        ("voxels_per_node", Int3),
        ("nodes", C.c_uint32),
        ("points", C.c_void_p),
        ("radii", C.c_void_p),
        ("blend", C.c_int32),
    ]


class FrameResult(C.Structure):
    _fields_ = [
        ("ticket", C.c_uint64),
        ("stream_id", C.c_int32),
        ("status", C.c_int32),
        ("submit_s", C.c_double),
        ("issue_s", C.c_double),
        ("done_s", C.c_double),
        ("batch", C.c_uint64),
        ("batch_frames", C.c_uint32),
        ("lane", C.c_uint32),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("frames_submitted", C.c_uint64),
        ("frames_done", C.c_uint64),
        ("batches", C.c_uint64),
        ("bytes_in_use", C.c_uint64),
        ("hbm_budget", C.c_uint64),
        ("streams_open", C.c_int32),
        ("max_batch_seen", C.c_int32),
        ("totals", FrameStats),
    ]

class GNProblem(C.Structure):
    _fields_ = [
        ("corr", C.c_void_p),
        ("corr_size", C.c_uint32),
        ("canon", C.c_void_p),
        ("k", C.c_int),
        ("knn_idxs", C.c_void_p),
        ("knn_weights", C.c_void_p),
        ("normals", C.c_void_p),
        ("graph", C.c_void_p),
        ("graph_size", C.c_uint32),
        ("graph_knn", C.c_void_p),
        ("reg_k", C.c_int),
        ("weight_depth", C.c_float),
        ("weight_reg", C.c_float),
    ]


class GNReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("stopped", C.c_int), ("final_energy", C.c_double),
                ("energy", C.c_double * 32), ("cholesky", C.c_int)]


vp = C.c_void_p
i32 = C.c_int32
u32 = C.c_uint32
u64 = C.c_uint64
f32 = C.c_float
sz = C.c_size_t

# name -> argtypes (restype int unless listed in _RESTYPES)
_SIGS = {
    "vf_status_string": [C.c_int],
    "vf_last_error": [],
    "vf_version": [],
    "vf_tsdf_knn_graph": [C.c_int, vp, vp, vp, vp, vp, sz, Float3, Float3, Int3, Int3, vp],
    "vf_compute_tsdf_warp_weights": [C.c_int, vp, vp, vp, vp, vp, sz, vp],
    "vf_generate_tsdf_vertices": [vp, Float3, Float3, Int3, vp],
    "vf_warp_tsdf": [C.c_int, vp, vp, vp, vp, vp, Float3, Float3, Int3, vp],
    "vf_count_collisions_per_voxel": [vp, vp, vp, Float3, Float3, Int3, Mat3, Float3, vp],
    "vf_integrate_warped_tsdf": [vp, vp, vp, vp, vp, vp, Int2, f32, Float3, Float3, Int3, Mat3, Float3, Mat3, f32,
                                 vp],
    "vf_integrate_tsdf": [vp, vp, vp, Int2, f32, Int3, Float3, Float3, Mat3, Float3, Mat3, f32, vp],
    "vf_integrate_tsdf_camera_frame": [vp, vp, vp, Int2, f32, Int3, Float3, Float3, Mat3, Float3, Mat3, Float3, Mat3,
                                       f32, vp],
    "vf_translate": [vp, u32, Float3, vp],
    "vf_eval_warp_weight": [vp, vp, vp, sz, vp],
    "vf_compute_warp_weights": [C.c_int, vp, vp, vp, vp, sz, sz, vp],
    "vf_warp_mesh": [C.c_int, vp, vp, vp, vp, vp, sz, sz, vp],
    "vf_mesh_knn": [vp, C.c_int, vp, vp, vp, sz, vp, sz, C.c_int, vp],
    "vf_warp_points": [vp, C.c_int, C.c_int, vp, vp, sz, vp, vp, vp, sz, vp],
    "vf_warp_points_static": [C.c_int, C.c_int, vp, vp, sz, vp, vp, vp, sz, vp],
    "vf_correspondence_pairs": [vp, vp, vp, vp, sz, Mat3, vp, Int2, vp, f32, vp],
    "vf_fill_r_depth": [vp, f32, vp, u32, vp, vp, vp, vp],
    "vf_fill_r_mreg": [vp, f32, vp, vp, u32, vp, C.c_int, vp],
    "vf_fill_j_depth": [vp, f32, u32, u32, vp, u32, vp, C.c_int, vp, vp, vp, vp, vp],
    "vf_fill_j_mreg": [vp, f32, u32, u32, vp, vp, u32, vp, C.c_int, vp],
    "vf_update_transforms": [vp, vp, u32, vp],
    "vf_lu_solve": [C.c_int, vp, C.c_int, vp, vp, vp],
    "vf_gn_normal_equations": [vp, vp, vp, vp, C.POINTER(GNProblem), vp],
    "vf_gn_solver_create": [C.POINTER(vp), C.POINTER(GNProblem), vp],
    "vf_gn_solver_create_ex": [C.POINTER(vp), C.POINTER(GNProblem), C.c_int, vp],
    "vf_gn_solver_destroy": [vp],
    "vf_gn_solver_iterate": [vp, vp, vp, vp, vp, sz, vp, C.c_int, f32, C.POINTER(GNReport), vp],
    "vf_filter_depth": [vp, u32, f32, f32, vp],
    "vf_mask_depth": [vp, vp, u32, vp],
    "vf_bilateral_filter": [vp, vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, f32, vp],
    "vf_depth_to_mesh": [vp, vp, u32, u32, Mat3, vp],
    "vf_depth_normals": [vp, vp, vp, C.c_int, C.c_int, f32, f32, vp],
    "vf_remove_depth_without_normals": [vp, vp, vp, C.c_int, C.c_int, vp],
    "vf_depth_centroid": [vp, C.POINTER(Float3), C.POINTER(C.c_uint64), vp, sz, vp],
    "vf_ingest_depth": [vp, vp, C.c_int, C.c_int, Mat3, f32, C.c_int, C.c_double, C.c_double, f32, f32, f32, vp, vp,
                        C.POINTER(Float3), C.POINTER(C.c_uint64), vp],
    "vf_tsdf_save": [C.c_char_p, C.c_int, C.POINTER(TsdfInfo), vp, vp],
    "vf_tsdf_load": [C.c_char_p, C.POINTER(TsdfInfo), vp, vp, u64],
    "vf_tsdf_save_host": [C.c_char_p, C.c_int, C.POINTER(TsdfInfo), vp, vp],
    "vf_tsdf_load_host": [C.c_char_p, C.POINTER(TsdfInfo), vp, vp, u64],
    "vf_marching_cubes_mesh": [vp, vp, vp, u64, vp, vp, Int3, Float3, Float3, f32, vp],
    "vf_marching_cubes_full": [vp, vp, vp, u64, vp, vp, u64, vp, vp, Int3, Float3, Float3, f32, vp],
    "vf_workspace_create": [C.POINTER(vp), C.c_int, u64, u64, vp],
    "vf_workspace_destroy": [vp],
    "vf_workspace_bytes": [vp],
    "vf_warped_integrate": [vp, C.POINTER(Volume), C.POINTER(Graph), C.POINTER(Frame), C.POINTER(WarpParams), vp],
    "vf_warped_integrate_batch": [vp, C.c_int, C.POINTER(Volume), C.POINTER(Graph), C.POINTER(Frame),
                                  C.POINTER(WarpParams), vp],
    "vf_workspace_last_stats": [vp, C.POINTER(FrameStats), vp],
    "vf_workspace_totals": [vp, C.POINTER(FrameStats), C.c_int, vp],
    "vf_workspace_set_timing": [vp, C.c_int],
    "vf_workspace_kernel_ms": [vp, C.POINTER(f32), C.POINTER(C.c_int)],
    "vf_workspace_set_tuning": [vp, C.POINTER(Tuning)],
    "vf_workspace_set_record_budget": [vp, f32],
    "vf_plan_create": [C.POINTER(vp), C.c_int, C.POINTER(Volume), C.POINTER(Graph), C.POINTER(Frame),
                       C.POINTER(WarpParams), C.POINTER(Tuning), C.c_int, vp],
    "vf_plan_destroy": [vp],
    "vf_plan_launch": [vp, vp],
    "vf_plan_autotune": [vp, C.c_int, C.POINTER(Tuning), vp],
    "vf_plan_get_tuning": [vp, C.POINTER(Tuning)],
    "vf_plan_stats": [vp, C.POINTER(FrameStats), C.POINTER(FrameStats), C.c_int, vp],
    "vf_plan_bytes": [vp],
    "vf_server_stream_bytes": [C.POINTER(StreamDesc), C.c_int],
    "vf_server_create": [C.POINTER(vp), C.POINTER(ServerConfig)],
    "vf_server_destroy": [vp],
    "vf_server_open": [vp, C.POINTER(StreamDesc), C.c_int, C.POINTER(C.c_int32)],
    "vf_server_close": [vp, C.c_int32],
    "vf_server_submit": [vp, C.c_int32, vp, vp, C.POINTER(C.c_uint64)],
    "vf_server_submit_borrowed": [vp, C.c_int32, vp, vp, C.POINTER(C.c_uint64)],
    "vf_server_wait": [vp, C.c_uint64, C.POINTER(FrameResult)],
    "vf_server_poll": [vp, C.POINTER(FrameResult), C.c_int, C.POINTER(C.c_int)],
    "vf_server_drain": [vp],
    "vf_server_read": [vp, C.c_int32, vp, vp, vp],
    "vf_server_get_stats": [vp, C.POINTER(ServerStats)],
    "vf_server_save": [vp, C.c_int32, C.c_char_p, C.c_int],
    "vf_server_restore": [vp, C.c_int32, C.c_char_p],
    "vf_device_count": [C.POINTER(C.c_int)],
    "vf_set_device": [C.c_int],
    "vf_malloc": [C.POINTER(vp), u64],
    "vf_free": [vp],
    "vf_memcpy_h2d": [vp, vp, u64, vp],
    "vf_memcpy_d2h": [vp, vp, u64, vp],
    "vf_memcpy_d2d": [vp, vp, u64, vp],
    "vf_memset": [vp, C.c_int, u64, vp],
    "vf_fill_f32": [vp, f32, u64, vp],
    "vf_stream_create": [C.POINTER(vp)],
    "vf_stream_destroy": [vp],
    "vf_stream_synchronize": [vp],
    "vf_device_synchronize": [],
    "vf_host_alloc": [C.POINTER(vp), u64],
    "vf_host_free": [vp],
    "vf_event_create": [C.POINTER(vp)],
    "vf_event_destroy": [vp],
    "vf_event_record": [vp, vp],
    "vf_event_elapsed_ms": [C.POINTER(f32), vp, vp],
    "vf_event_synchronize": [vp],
}
_RESTYPES = {
    "vf_status_string": C.c_char_p,
    "vf_last_error": C.c_char_p,
    "vf_version": C.c_char_p,
    "vf_workspace_bytes": u64,
    "vf_plan_bytes": u64,
    "vf_server_stream_bytes": u64,
}

_lib: C.CDLL | None = None


def exported_symbols() -> list[str]:
    """Every entry point the Python binding declares (all of include/volumefusion.h)."""
    return list(_SIGS)


def load() -> C.CDLL:
    """Load the gfx950 library (built in-tree by volumefusion_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the warped-integration path has no fallback)")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | C.RTLD_GLOBAL)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)
    _lib = lib
    return lib


def check(status: int, func: str) -> None:
    if status != 0:
        msg = load().vf_last_error().decode(errors="replace") or load().vf_status_string(status).decode()
        raise VFError(status, func, msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


def f3(v) -> Float3:
    return Float3(float(v[0]), float(v[1]), float(v[2]))


def i3(v) -> Int3:
    return Int3(int(v[0]), int(v[1]), int(v[2]))


def m3(a) -> Mat3:
    m = Mat3()
    for i, x in enumerate(list(a)[:9]):
        m.m[i] = float(x)
    return m
