"""ctypes front end of the C oracle (vf_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker the HIP kernels are compared against.
Imported by the tests, by the smoke check and by bench.py's CPU-baseline leg;
never by the volumefusion_amd package.

Arrays are numpy: vectors as float32[..., 3], 4x4 transforms as
float32[G, 16] column-major (Eigen::Matrix4f storage), intrinsics as
float32[9] column-major.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libvf_oracle.so"

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


class Vol(C.Structure):
    _fields_ = [("dim", C.c_int32 * 3), ("vs", C.c_float * 3), ("size", C.c_float * 3)]


class Stats(C.Structure):
    _fields_ = [
        ("supported", C.c_int64),
        ("in_bounds", C.c_int64),
        ("collided", C.c_int64),
        ("updated", C.c_int64),
        ("candidates", C.c_int64),
        ("overflow", C.c_int64),
    ]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class FrameOut(C.Structure):
    _fields_ = [
        ("cnt", C.c_void_p),
        ("idx", C.c_void_p),
        ("dist", C.c_void_p),
        ("w", C.c_void_p),
        ("verts", C.c_void_p),
        ("col", C.c_void_p),
        ("pix", C.c_void_p),
    ]


def build() -> Path:
    """Compile the oracle with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.vfo_expw.argtypes = [C.c_float]
        L.vfo_expw.restype = C.c_float
        L.vfo_exp_libm.argtypes = [C.c_float]
        L.vfo_exp_libm.restype = C.c_float
        L.vfo_weight_tsdf.argtypes = [C.c_float, C.c_float]
        L.vfo_weight_tsdf.restype = C.c_float
        L.vfo_weight_mesh_host.argtypes = [C.c_float, C.c_float]
        L.vfo_weight_mesh_host.restype = C.c_float
        L.vfo_div_markstein.argtypes = [C.c_float, C.c_float, C.c_float]
        L.vfo_div_markstein.restype = C.c_float
        L.vfo_markstein_mismatches.argtypes = [C.c_float, C.c_uint32, C.c_uint32]
        L.vfo_markstein_mismatches.restype = C.c_int64
        L.vfo_mat3_inverse.argtypes = [_f32p, _f32p]
        L.vfo_cvt_i32_rz.argtypes = [C.c_float]
        L.vfo_cvt_i32_rz.restype = C.c_int32
        L.vfo_translate.argtypes = [_f32p, C.c_size_t, _f32p]
        L.vfo_generate_vertices.argtypes = [_f32p, C.POINTER(Vol)]
        L.vfo_knn_graph.argtypes = [C.c_int, _u32p, _f32p, _i32p, _f32p, _f32p, C.c_size_t, C.POINTER(Vol), _i32p]
        L.vfo_knn_graph.restype = C.c_int64
        L.vfo_warp_weights.argtypes = [C.c_int, _f32p, _f32p, _u32p, _f32p, _i32p, C.c_size_t]
        L.vfo_warp_tsdf.argtypes = [C.c_int, _f32p, _f32p, _u32p, _f32p, _i32p, C.c_size_t]
        L.vfo_count_collisions.argtypes = [_i32p, _f32p, _i32p, C.POINTER(Vol), _f32p, _f32p]
        L.vfo_integrate_warped.argtypes = [_f32p, _f32p, _f32p, _i32p, _i32p, _f32p, C.c_int, C.c_int, C.c_float,
                                           C.POINTER(Vol), _f32p, _f32p, _f32p, C.c_float]
        L.vfo_integrate_warped.restype = C.c_int64
        L.vfo_integrate_plain.argtypes = [_f32p, _f32p, _f32p, C.c_int, C.c_int, C.c_float, C.POINTER(Vol), _f32p,
                                          _f32p, _f32p, C.c_float]
        L.vfo_integrate_plain.restype = C.c_int64
        L.vfo_integrate_camera_frame.argtypes = [_f32p, _f32p, _f32p, C.c_int, C.c_int, C.c_float, C.POINTER(Vol),
                                                 _f32p, _f32p, _f32p, _f32p, _f32p, C.c_float]
        L.vfo_integrate_camera_frame.restype = C.c_int64
        L.vfo_tsdfvolume_integrate.argtypes = [_f32p, _f32p, _f32p, C.c_int, C.c_int, C.c_float, C.POINTER(Vol),
                                               _f32p, _f32p]
        L.vfo_tsdfvolume_integrate.restype = C.c_int64
        L.vfo_mesh_warp.argtypes = [_f32p, _f32p, C.c_size_t, _f32p, _f32p, _f32p, C.c_size_t, C.c_int, C.c_int,
                                    C.c_int]
        L.vfo_warped_frame.argtypes = [C.c_int, _f32p, _f32p, _f32p, _f32p, _f32p, C.c_size_t, C.POINTER(Vol), _i32p,
                                       _f32p, _f32p, C.c_int, C.c_int, C.c_float, _f32p, C.c_float, C.POINTER(Stats),
                                       C.POINTER(FrameOut), C.c_int]
        L.vfo_warped_frame.restype = C.c_int
        L.vfo_warped_frame_gather.argtypes = [C.c_int, _f32p, _f32p, _f32p, _f32p, _f32p, C.c_size_t,
                                              C.POINTER(Vol), _i32p, _f32p, _f32p, C.c_int, C.c_int, C.c_float,
                                              _f32p, C.c_float, C.c_int, C.c_int, C.c_int, C.POINTER(Stats), C.c_int,
                                              C.c_void_p, C.c_void_p]
        L.vfo_warped_frame_gather.restype = C.c_int
        L.vfo_warped_frame_f64.argtypes = [C.c_int, _f64p, _f64p, _f32p, _f32p, _f32p, C.c_size_t, C.POINTER(Vol),
                                           _f32p, _f32p, C.c_int, C.c_float, C.c_float, _i32p, _u32p, _i32p,
                                           C.c_int]
        L.vfo_warped_frame_f64.restype = C.c_int
        L.vfo_mesh_knn.argtypes = [C.c_int, _u32p, _f32p, _f32p, C.c_size_t, _f32p, C.c_size_t, C.c_int]
        L.vfo_compute_warp_weights.argtypes = [C.c_int, _f32p, _f32p, _u32p, _f32p, C.c_size_t, C.c_size_t]
        L.vfo_warp_mesh.argtypes = [C.c_int, _f32p, _f32p, _f32p, _u32p, _f32p, C.c_size_t, C.c_size_t]
        L.vfo_warp_points.argtypes = [C.c_int, _f32p, _f32p, C.c_size_t, _f32p, _f32p, _f32p, C.c_size_t, C.c_int]
        L.vfo_marching_cubes.argtypes = [C.c_int, _f32p, _f32p, C.POINTER(Vol), C.c_float, _f32p, C.c_uint64,
                                         _i32p, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.vfo_filter_depth.argtypes = [_f32p, C.c_size_t, C.c_float, C.c_float]
        L.vfo_mask_depth.argtypes = [_f32p, np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS"), C.c_size_t]
        L.vfo_depth_to_mesh.argtypes = [_f32p, _f32p, C.c_int, C.c_int, _f32p]
        L.vfo_depth_normals.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_float, C.c_float]
        L.vfo_remove_depth_without_normals.argtypes = [_f32p, _f32p, _f32p, C.c_int, C.c_int]
        L.vfo_depth_centroid.argtypes = [_f32p, C.c_size_t, _f32p, C.POINTER(C.c_uint64)]
        L.vfo_bilateral_tables.argtypes = [C.c_int, C.c_double, _f32p, _i32p]
        L.vfo_bilateral.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_float]
        L.vfo_mat4_to_dq.argtypes = [_f32p, _f32p]
        L.vfo_mat4_to_dq_f64.argtypes = [_f32p, _f64p]
        L.vfo_dqb_apply.argtypes = [C.POINTER(C.c_void_p), _f32p, C.c_int, _f32p, _f32p]
        L.vfo_dqb_apply_f64.argtypes = [C.POINTER(C.c_void_p), _f64p, C.c_int, _f64p, _f64p]
        L.vfo_corr_pairs.argtypes = [_i32p, _f32p, C.c_size_t, _f32p, _f32p, C.c_int, C.c_int, _f32p, C.c_float]
        L.vfo_corr_pairs.restype = C.c_int64
        L.vfo_gn_r_depth.argtypes = [_f32p, C.c_float, _i32p, C.c_size_t, _f32p, _f32p, _f32p]
        L.vfo_gn_r_depth.restype = None
        L.vfo_gn_r_mreg.argtypes = [_f32p, C.c_float, _f32p, _f32p, C.c_size_t, _u32p, C.c_int]
        L.vfo_gn_r_mreg.restype = None
        L.vfo_gn_j_depth.argtypes = [_f32p, C.c_float, C.c_size_t, C.c_size_t, _i32p, _f32p, _u32p, _f32p, C.c_int,
                                     _f32p]
        L.vfo_gn_j_depth.restype = None
        L.vfo_gn_j_mreg.argtypes = [_f32p, C.c_float, _f32p, C.c_size_t, _u32p, C.c_int]
        L.vfo_gn_j_mreg.restype = None
        L.vfo_gn_normal_eq.argtypes = [_f32p, _f32p, _f32p, _f32p, C.c_size_t, C.c_size_t, _i32p, _u32p, _f32p, C.c_int,
                                       _u32p, C.c_int]
        L.vfo_gn_normal_eq.restype = C.c_int
        L.vfo_gn_update.argtypes = [_f32p, _f32p, C.c_size_t]
        L.vfo_gn_update.restype = None
        L.vfo_lu_solve.argtypes = [C.c_int, _f32p, _f32p, _f64p]
        L.vfo_lu_solve.restype = C.c_int
        _lib = L
    return _lib


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def make_vol(dim, vs, size) -> Vol:
    v = Vol()
    for a in range(3):
        v.dim[a] = int(dim[a])
        v.vs[a] = float(np.float32(vs[a]))
        v.size[a] = float(np.float32(size[a]))
    return v


@dataclass
class FrameResult:
    stats: dict
    cnt: np.ndarray | None = None
    idx: np.ndarray | None = None
    dist: np.ndarray | None = None
    w: np.ndarray | None = None
    verts: np.ndarray | None = None
    col: np.ndarray | None = None
    pix: np.ndarray | None = None


# ---- scalar helpers -----------------------------------------------------

def expw(x: float) -> float:
    return lib().vfo_expw(float(x))


def expw_array(x: np.ndarray) -> np.ndarray:
    f = lib().vfo_expw
    return np.array([f(float(v)) for v in np.asarray(x, np.float32).ravel()], np.float32)


def mat3_inverse(m9: np.ndarray) -> np.ndarray:
    out = np.zeros(9, np.float32)
    lib().vfo_mat3_inverse(_f32(m9), out)
    return out


# ---- reference kernels one by one ----------------------------------------

def translate(pts: np.ndarray, t) -> np.ndarray:
    p = _f32(pts).copy()
    lib().vfo_translate(p, p.shape[0], _f32(t))
    return p


def generate_vertices(dim, vs, size) -> np.ndarray:
    n = int(dim[0]) * int(dim[1]) * int(dim[2])
    out = np.zeros((n, 3), np.float32)
    v = make_vol(dim, vs, size)
    lib().vfo_generate_vertices(out, C.byref(v))
    return out


def knn_graph(k, nodes, radii, dim, vs, size, cube):
    n = int(dim[0]) * int(dim[1]) * int(dim[2])
    idx = np.zeros((n, k), np.uint32)
    dist = np.zeros((n, k), np.float32)
    cnt = np.zeros(n, np.int32)
    v = make_vol(dim, vs, size)
    nodes = _f32(nodes)
    of = lib().vfo_knn_graph(k, idx, dist, cnt, nodes, _f32(radii), nodes.shape[0], C.byref(v),
                             np.ascontiguousarray(cube, np.int32))
    return idx, dist, cnt, int(of)


def warp_weights(k, radii, idx, dist, cnt) -> np.ndarray:
    n = cnt.shape[0]
    w = np.zeros((n, k), np.float32)
    lib().vfo_warp_weights(k, w, _f32(radii), np.ascontiguousarray(idx, np.uint32), _f32(dist),
                           np.ascontiguousarray(cnt, np.int32), n)
    return w


def warp_tsdf(k, verts, T, idx, w, cnt) -> np.ndarray:
    v = _f32(verts).copy()
    lib().vfo_warp_tsdf(k, v, _f32(T), np.ascontiguousarray(idx, np.uint32), _f32(w),
                        np.ascontiguousarray(cnt, np.int32), cnt.shape[0])
    return v


def count_collisions(verts, cnt, dim, vs, size, R, t) -> np.ndarray:
    n = cnt.shape[0]
    col = np.zeros(n, np.int32)
    v = make_vol(dim, vs, size)
    lib().vfo_count_collisions(col, _f32(verts), np.ascontiguousarray(cnt, np.int32), C.byref(v), _f32(R), _f32(t))
    return col


def integrate_warped(tsdf, wt, verts, col, cnt, depth, delta, dim, vs, size, R, t, Kin, maxw):
    ts = _f32(tsdf).copy()
    ww = _f32(wt).copy()
    H, W = depth.shape
    v = make_vol(dim, vs, size)
    u = lib().vfo_integrate_warped(ts, ww, _f32(verts), np.ascontiguousarray(col, np.int32),
                                   np.ascontiguousarray(cnt, np.int32), _f32(depth), W, H, float(delta), C.byref(v),
                                   _f32(R), _f32(t), _f32(Kin), float(maxw))
    return ts, ww, int(u)


def integrate_plain(tsdf, wt, depth, delta, dim, vs, size, R, t, Kin, maxw):
    ts = _f32(tsdf).copy()
    ww = _f32(wt).copy()
    H, W = depth.shape
    v = make_vol(dim, vs, size)
    u = lib().vfo_integrate_plain(ts, ww, _f32(depth), W, H, float(delta), C.byref(v), _f32(R), _f32(t), _f32(Kin),
                                  float(maxw))
    return ts, ww, int(u)


def integrate_camera_frame(tsdf, wt, depth, delta, dim, vs, size, Rw2c, tw2c, R, t, Kin, maxw):
    ts = _f32(tsdf).copy()
    ww = _f32(wt).copy()
    H, W = depth.shape
    v = make_vol(dim, vs, size)
    u = lib().vfo_integrate_camera_frame(ts, ww, _f32(depth), W, H, float(delta), C.byref(v), _f32(Rw2c),
                                         _f32(tw2c), _f32(R), _f32(t), _f32(Kin), float(maxw))
    return ts, ww, int(u)


def tsdfvolume_integrate(tsdf, wt, depth, delta, dim, vs, size, pose16, Kin):
    ts = _f32(tsdf).copy()
    ww = _f32(wt).copy()
    H, W = depth.shape
    v = make_vol(dim, vs, size)
    u = lib().vfo_tsdfvolume_integrate(ts, ww, _f32(depth), W, H, float(delta), C.byref(v), _f32(pose16), _f32(Kin))
    return ts, ww, int(u)


def mesh_warp(pts, nodes, radii, T, k, host_order=True, squared_dist=False) -> np.ndarray:
    pts = _f32(pts)
    out = np.zeros_like(pts)
    nodes = _f32(nodes)
    lib().vfo_mesh_warp(out, pts, pts.shape[0], nodes, _f32(radii), _f32(T), nodes.shape[0], int(k),
                        1 if host_order else 0, 1 if squared_dist else 0)
    return out


# ---- whole frame ----------------------------------------------------------

def mesh_knn(k, pts, nodes, exclude_closest=False):
    """K nearest nodes of each point: (idx [n, k] uint32, dist [n, k] float32)."""
    pts, nodes = _f32(pts).reshape(-1, 3), _f32(nodes).reshape(-1, 3)
    idx = np.zeros((pts.shape[0], k), np.uint32)
    dist = np.zeros((pts.shape[0], k), np.float32)
    lib().vfo_mesh_knn(int(k), idx, dist, pts, pts.shape[0], nodes, nodes.shape[0], 1 if exclude_closest else 0)
    return idx, dist


def compute_warp_weights(k, radii, idx, dist) -> np.ndarray:
    idx = np.ascontiguousarray(idx, np.uint32)
    w = np.zeros(idx.shape, np.float32)
    radii = _f32(radii)
    lib().vfo_compute_warp_weights(int(k), w, radii, idx, _f32(dist), idx.shape[0], radii.shape[0])
    return w


def warp_mesh(k, pts, T, idx, w, G) -> np.ndarray:
    pts = _f32(pts).reshape(-1, 3)
    out = np.zeros_like(pts)
    lib().vfo_warp_mesh(int(k), out, pts, _f32(T), np.ascontiguousarray(idx, np.uint32), _f32(w), pts.shape[0],
                        int(G))
    return out


def warp_points(k, pts, nodes, radii, T, blend=0) -> np.ndarray:
    pts, nodes = _f32(pts).reshape(-1, 3), _f32(nodes).reshape(-1, 3)
    out = np.zeros_like(pts)
    lib().vfo_warp_points(int(k), out, pts, pts.shape[0], nodes, _f32(radii), _f32(T), nodes.shape[0], int(blend))
    return out


def marching_cubes(tsdf, wt, dim, vs, size, iso=0.0, full=False, cap=None):
    """(vertices [n, 3], faces [m, 3] or None) in canonical order."""
    v = make_vol(dim, vs, size)
    nv, nf = C.c_uint64(0), C.c_uint64(0)
    dummy_v, dummy_f = np.zeros((1, 3), np.float32), np.zeros((1, 3), np.int32)
    t, w = _f32(tsdf), _f32(wt)
    lib().vfo_marching_cubes(1 if full else 0, t, w, C.byref(v), float(iso), dummy_v, 0, dummy_f, 0, C.byref(nv),
                             C.byref(nf))
    verts = np.zeros((max(nv.value, 1), 3), np.float32)
    faces = np.zeros((max(nf.value, 1), 3), np.int32)
    lib().vfo_marching_cubes(1 if full else 0, t, w, C.byref(v), float(iso), verts, nv.value, faces, nf.value,
                             C.byref(nv), C.byref(nf))
    return verts[:nv.value], (faces[:nf.value] if full else None)


def filter_depth(depth, threshold, fill=np.inf):
    d = _f32(depth).copy()
    lib().vfo_filter_depth(d, d.size, float(threshold), float(fill))
    return d


def mask_depth(depth, mask):
    d = _f32(depth).copy()
    lib().vfo_mask_depth(d, np.ascontiguousarray(mask, np.uint8), d.size)
    return d


def depth_to_mesh(depth, Kin):
    H, W = depth.shape
    mesh = np.zeros((H * W, 3), np.float32)
    lib().vfo_depth_to_mesh(mesh, _f32(depth), W, H, _f32(Kin))
    return mesh


def depth_normals(verts, H, W, a0, a1, normals=None):
    n = np.zeros((H * W, 3), np.float32) if normals is None else normals
    lib().vfo_depth_normals(n, _f32(verts), H, W, float(a0), float(a1))
    return n


def remove_depth_without_normals(normals, verts, depth):
    H, W = depth.shape
    v, d = _f32(verts).copy(), _f32(depth).copy()
    lib().vfo_remove_depth_without_normals(_f32(normals), v, d, H, W)
    return v, d


def depth_centroid(mesh):
    out = np.zeros(3, np.float32)
    cnt = C.c_uint64(0)
    m = _f32(mesh).reshape(-1, 3)
    lib().vfo_depth_centroid(m, m.shape[0], out, C.byref(cnt))
    return out, cnt.value


def bilateral_tables(r, sd):
    g = np.zeros(400, np.float32)
    c = np.zeros(31, np.int32)
    lib().vfo_bilateral_tables(int(r), float(sd), g, c)
    return g[:(2 * r + 1) ** 2], c


def bilateral(depth, r, sI, sS, thr):
    H, W = depth.shape
    out = np.zeros_like(_f32(depth))
    lib().vfo_bilateral(out, _f32(depth), H, W, int(r), float(sI), float(sS), float(thr))
    return out


def mat4_to_dq(T16) -> np.ndarray:
    """Dual quaternion (real xyzw, dual xyzw) of a column-major Mat4f, float32."""
    out = np.zeros(8, np.float32)
    lib().vfo_mat4_to_dq(_f32(T16), out)
    return out


def mat4_to_dq_f64(T16) -> np.ndarray:
    out = np.zeros(8, np.float64)
    lib().vfo_mat4_to_dq_f64(_f32(T16), out)
    return out


def dqb_apply(dqs, w, p, f64=False) -> np.ndarray:
    """DQB of the dual quaternions dqs [n, 8] with weights w [n] applied to p."""
    dt = np.float64 if f64 else np.float32
    rows = [np.ascontiguousarray(q, dt) for q in dqs]
    ptrs = (C.c_void_p * len(rows))(*[r.ctypes.data for r in rows])
    out = np.zeros(3, dt)
    fn = lib().vfo_dqb_apply_f64 if f64 else lib().vfo_dqb_apply
    fn(ptrs, np.ascontiguousarray(w, dt), len(rows), np.ascontiguousarray(p, dt), out)
    return out


def warped_frame(k, tsdf, wt, nodes, radii, T, dim, vs, size, cube, c, depth, delta, Kin, maxw, want=(), blend=0):
    """One frame of src/VolumeFusion.cu:699-742.  tsdf, wt and nodes are
    updated in place (nodes get the reference's round trip).  `want` names the
    per-voxel outputs to return: cnt, idx, dist, w, verts, col, pix."""
    n = int(dim[0]) * int(dim[1]) * int(dim[2])
    out = FrameOut()
    keep = {}
    shapes = {"cnt": ((n,), np.int32), "idx": ((n, k), np.uint32), "dist": ((n, k), np.float32),
              "w": ((n, k), np.float32), "verts": ((n, 3), np.float32), "col": ((n,), np.int32),
              "pix": ((n,), np.int32)}
    for name in want:
        shp, dt = shapes[name]
        arr = np.zeros(shp, dt)
        keep[name] = arr
        setattr(out, name, arr.ctypes.data)
    st = Stats()
    v = make_vol(dim, vs, size)
    H, W = depth.shape
    assert tsdf.dtype == np.float32 and wt.dtype == np.float32 and nodes.dtype == np.float32
    rc = lib().vfo_warped_frame(k, tsdf, wt, nodes, _f32(radii), _f32(T), nodes.shape[0], C.byref(v),
                                np.ascontiguousarray(cube, np.int32), _f32(c), _f32(depth), W, H, float(delta),
                                _f32(Kin), float(maxw), C.byref(st), C.byref(out), int(blend))
    if rc != 0:
        raise MemoryError("oracle frame allocation failed")
    return FrameResult(stats=st.as_dict(), **keep)


def warped_frame_gather(k, tsdf, wt, nodes, radii, T, dim, vs, size, cube, c, depth, delta, Kin, maxw,
                        nthreads=None, z0=0, z1=0, blend=0, tsdf64=None, wt64=None) -> dict:
    """Voxel-major multi-threaded restatement of the same frame (sources in
    z in [z0, z1); z1 <= 0 = whole volume).  Updates tsdf, wt, nodes in place.
    With tsdf64/wt64 (float64 arrays) it also advances the binary64 companion
    state: the same discrete decisions, continuous values in binary64 (the
    equal of warped_frame_f64)."""
    if (tsdf64 is None) != (wt64 is None):
        raise ValueError("tsdf64 and wt64 go together")
    if tsdf64 is not None:
        n = int(dim[0]) * int(dim[1]) * int(dim[2])
        assert tsdf64.dtype == np.float64 and wt64.dtype == np.float64 and tsdf64.size == n and wt64.size == n
        assert tsdf64.flags.c_contiguous and wt64.flags.c_contiguous
    if nthreads is None:
        nthreads = os.cpu_count() or 1
    st = Stats()
    v = make_vol(dim, vs, size)
    H, W = depth.shape
    assert tsdf.dtype == np.float32 and wt.dtype == np.float32 and nodes.dtype == np.float32
    rc = lib().vfo_warped_frame_gather(k, tsdf, wt, nodes, _f32(radii), _f32(T), nodes.shape[0], C.byref(v),
                                       np.ascontiguousarray(cube, np.int32), _f32(c), _f32(depth), W, H,
                                       float(delta), _f32(Kin), float(maxw), int(nthreads), int(z0), int(z1),
                                       C.byref(st), int(blend),
                                       tsdf64.ctypes.data if tsdf64 is not None else None,
                                       wt64.ctypes.data if wt64 is not None else None)
    if rc != 0:
        raise RuntimeError(f"oracle gather frame failed ({rc})")
    return st.as_dict()


def warped_frame_f64(k, tsdf64, wt64, nodes_before, radii, T, dim, vs, size, c, depth, delta, maxw, cnt, idx, pix,
                     blend=0):
    v = make_vol(dim, vs, size)
    H, W = depth.shape
    lib().vfo_warped_frame_f64(k, tsdf64, wt64, _f32(nodes_before), _f32(radii), _f32(T), nodes_before.shape[0],
                               C.byref(v), _f32(c), _f32(depth), W, float(delta), float(maxw),
                               np.ascontiguousarray(cnt, np.int32), np.ascontiguousarray(idx, np.uint32),
                               np.ascontiguousarray(pix, np.int32), int(blend))


# ---- Gauss-Newton motion optimisation (vf_gn.c) ------------------------------

def corr_pairs(warped, Kin, depth, frame_mesh, thresh) -> np.ndarray:
    """correspondencePairsKernel (src/ComputeCorrespondence.cu:46-82), pairs in
    ascending mesh index: int32 [C, 2] of (mesh index, pixel index)."""
    warped = _f32(warped)
    M = warped.shape[0]
    H, W = depth.shape
    out = np.zeros((max(M, 1), 2), np.int32)
    n = lib().vfo_corr_pairs(out, warped, M, _f32(Kin), _f32(depth), H, W, _f32(frame_mesh), float(thresh))
    return out[:n].copy()


def gn_r_depth(wD, corr, warped, dmesh, dnormals) -> np.ndarray:
    corr = np.ascontiguousarray(corr, np.int32)
    r = np.zeros(max(corr.shape[0], 1), np.float32)
    lib().vfo_gn_r_depth(r, float(wD), corr, corr.shape[0], _f32(warped), _f32(dmesh), _f32(dnormals))
    return r[:corr.shape[0]]


def gn_r_mreg(wR, graph, T, gknn) -> np.ndarray:
    gknn = np.ascontiguousarray(gknn, np.uint32)
    G, R = gknn.shape
    r = np.zeros(max(G * R * 3, 1), np.float32)
    lib().vfo_gn_r_mreg(r, float(wR), _f32(graph), _f32(T), G, gknn, R)
    return r[:G * R * 3]


def gn_j_depth(wD, corr, canon, knn, kw, dnormals, G) -> np.ndarray:
    corr = np.ascontiguousarray(corr, np.int32)
    knn = np.ascontiguousarray(knn, np.uint32)
    C = corr.shape[0]
    J = np.zeros((max(C, 1), G, 6), np.float32)
    lib().vfo_gn_j_depth(J, float(wD), C, G, corr, _f32(canon), knn, _f32(kw), knn.shape[1], _f32(dnormals))
    return J[:C]


def gn_j_mreg(wR, graph, gknn) -> np.ndarray:
    gknn = np.ascontiguousarray(gknn, np.uint32)
    G, R = gknn.shape
    J = np.zeros((max(G * R * 3, 1), G, 6), np.float32)
    lib().vfo_gn_j_mreg(J, float(wR), _f32(graph), G, gknn, R)
    return J[:G * R * 3]


def gn_normal_eq(J, r, corr, knn, kw, gknn):
    """JTJ (6G x 6G) and JTr (G x 6) from the dense J (depth rows, then the
    3 R rows of each node) and r."""
    gknn = np.ascontiguousarray(gknn, np.uint32)
    knn = np.ascontiguousarray(knn, np.uint32)
    corr = np.ascontiguousarray(corr, np.int32)
    G, R = gknn.shape
    C = corr.shape[0]
    JTJ = np.zeros((6 * G, 6 * G), np.float32)
    JTr = np.zeros((G, 6), np.float32)
    rc = lib().vfo_gn_normal_eq(JTJ, JTr, _f32(J), _f32(r), C, G, corr if C else np.zeros((1, 2), np.int32), knn,
                                _f32(kw), knn.shape[1], gknn, R)
    if rc != 0:
        raise MemoryError("oracle normal equations allocation failed")
    return JTJ, JTr


def gn_update(T, u) -> np.ndarray:
    T = _f32(T).copy()
    lib().vfo_gn_update(T, _f32(u), T.shape[0])
    return T


def lu_solve(A, b):
    """LU with partial pivoting in binary64: (x, info)."""
    A = _f32(A)
    n = A.shape[0]
    x = np.zeros(max(n, 1), np.float64)
    info = lib().vfo_lu_solve(n, A, _f32(b), x)
    return x[:n], int(info)
