"""The reference's af:: entry points on the gfx950 kernels.

Each function keeps the reference's name (snake_case), argument order and
meaning, and, like the reference wrappers (which end in cudaDeviceSynchronize,
e.g. src/TSDFWarp.cu:85), returns only when the work is done.  Errors raise
VFError (the reference throws std::runtime_error from CUDA_CHECK).

Device buffers are DeviceArray objects on the reference layouts:
  knn idx  uint32[N, K]   (Vecui<K>[N])     knn dist / weights float32[N, K]
  counts   int32[N]       tsdf / weights    float32[N] z-major
  vertices float32[N, 3]  (Vec3f[N])        transforms float32[G, 16] column-major
  depth    float32[H, W]
Small types (Vec3f, Vec3i, Mat3f) are plain sequences; Mat3f column-major.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .device import DeviceArray, Stream, default_stream

IDENTITY3 = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)


def _s(stream: Stream | None):
    return (stream or default_stream()).ptr


def _done(stream: Stream | None) -> None:
    (stream or default_stream()).synchronize()


def _k_of(arr: DeviceArray) -> int:
    if len(arr.shape) != 2:
        raise ValueError("knn buffers are [N, K]")
    return arr.shape[1]


def run_tsdf_knn_graph(knn_idxs: DeviceArray, knn_dists: DeviceArray, knn_counts: DeviceArray, graph: DeviceArray,
                       graph_radiuses: DeviceArray, graph_size: int, voxel_size, volume_size, volume_dim,
                       voxels_per_node, stream: Stream | None = None) -> None:
    """af::runTsdfKNNGraph (include/af/TSDFWarp.h:9-18)."""
    L.call("vf_tsdf_knn_graph", _k_of(knn_idxs), knn_idxs.ptr, knn_dists.ptr, knn_counts.ptr, graph.ptr,
           graph_radiuses.ptr, int(graph_size), L.f3(voxel_size), L.f3(volume_size), L.i3(volume_dim),
           L.i3(voxels_per_node), _s(stream))
    _done(stream)


def run_compute_tsdf_warp_weights_kernel(knn_weights: DeviceArray, motion_graph_radiuses: DeviceArray,
                                         knn_idxs: DeviceArray, knn_dists: DeviceArray, knn_counts: DeviceArray,
                                         grid_size: int, stream: Stream | None = None) -> None:
    """af::runComputeTsdfWarpWeightsKernel (include/af/TSDFWarp.h:20-25)."""
    L.call("vf_compute_tsdf_warp_weights", _k_of(knn_weights), knn_weights.ptr, motion_graph_radiuses.ptr,
           knn_idxs.ptr, knn_dists.ptr, knn_counts.ptr, int(grid_size), _s(stream))
    _done(stream)


def run_generate_tsdf_vertices_kernel(tsdf_vertices: DeviceArray, voxel_size, volume_size, volume_dim,
                                      stream: Stream | None = None) -> None:
    """af::runGenerateTsdfVerticesKernel (include/af/TSDFWarp.h:27)."""
    L.call("vf_generate_tsdf_vertices", tsdf_vertices.ptr, L.f3(voxel_size), L.f3(volume_size), L.i3(volume_dim),
           _s(stream))
    _done(stream)


def run_warp_tsdf_kernel(tsdf_vertices: DeviceArray, motion_graph: DeviceArray, knn_idxs: DeviceArray,
                         knn_weights: DeviceArray, knn_counts: DeviceArray, voxel_size, volume_size, volume_dim,
                         stream: Stream | None = None) -> None:
    """af::runWarpTsdfKernel (include/af/TSDFWarp.h:29-36)."""
    L.call("vf_warp_tsdf", _k_of(knn_idxs), tsdf_vertices.ptr, motion_graph.ptr, knn_idxs.ptr, knn_weights.ptr,
           knn_counts.ptr, L.f3(voxel_size), L.f3(volume_size), L.i3(volume_dim), _s(stream))
    _done(stream)


def run_count_collisions_per_voxel_kernel(collision_counts: DeviceArray, tsdf_vertices: DeviceArray,
                                          knn_counts: DeviceArray, voxel_size, volume_size, volume_dim,
                                          volume_rotate, volume_translate, stream: Stream | None = None) -> None:
    """af::runCountCollisionsPerVoxelKernel (include/af/TSDFIntegration.cuh:49-56)."""
    L.call("vf_count_collisions_per_voxel", collision_counts.ptr, tsdf_vertices.ptr, knn_counts.ptr,
           L.f3(voxel_size), L.f3(volume_size), L.i3(volume_dim), L.m3(volume_rotate), L.f3(volume_translate),
           _s(stream))
    _done(stream)


def run_integrate_warped_tsdf_from_frame_kernel(tsdf: DeviceArray, tsdf_weights: DeviceArray,
                                                tsdf_vertices: DeviceArray, collision_counts: DeviceArray,
                                                knn_counts: DeviceArray, depth_frame: DeviceArray, frame_dims,
                                                delta: float, voxel_size, volume_size, volume_dim, volume_rotate,
                                                volume_translate, depth_intrinsics, max_weight: float = 15.0,
                                                stream: Stream | None = None) -> None:
    """af::runIntegrateWarpedTSDFFromFrameKernel (include/af/TSDFIntegration.cuh:34-47)."""
    L.call("vf_integrate_warped_tsdf", tsdf.ptr, tsdf_weights.ptr, tsdf_vertices.ptr, collision_counts.ptr,
           knn_counts.ptr, depth_frame.ptr, L.Int2(int(frame_dims[0]), int(frame_dims[1])), float(delta),
           L.f3(voxel_size), L.f3(volume_size), L.i3(volume_dim), L.m3(volume_rotate), L.f3(volume_translate),
           L.m3(depth_intrinsics), float(max_weight), _s(stream))
    _done(stream)


def run_integrate_tsdf_from_frame_kernel(tsdf: DeviceArray, tsdf_weights: DeviceArray, depth_frame: DeviceArray,
                                         frame_dims, delta: float, dims, voxel_size, volume_size, volume_rotate,
                                         volume_translate, depth_intrinsics, max_weight: float = 15.0,
                                         stream: Stream | None = None) -> None:
    """af::runIntegrateTSDFFromFrameKernel (include/af/TSDFIntegration.cuh:8-18)."""
    L.call("vf_integrate_tsdf", tsdf.ptr, tsdf_weights.ptr, depth_frame.ptr,
           L.Int2(int(frame_dims[0]), int(frame_dims[1])), float(delta), L.i3(dims), L.f3(voxel_size),
           L.f3(volume_size), L.m3(volume_rotate), L.f3(volume_translate), L.m3(depth_intrinsics), float(max_weight),
           _s(stream))
    _done(stream)


def run_integrate_tsdf_from_camera_frame_kernel(tsdf: DeviceArray, tsdf_weights: DeviceArray,
                                                depth_frame: DeviceArray, frame_dims, delta: float, dims, voxel_size,
                                                volume_size, world2camera_rotate, world2camera_translate,
                                                volume_rotate, volume_translate, depth_intrinsics,
                                                max_weight: float = 15.0, stream: Stream | None = None) -> None:
    """af::runIntegrateTSDFFromCameraFrameKernel (include/af/TSDFIntegration.cuh:20-32)."""
    L.call("vf_integrate_tsdf_camera_frame", tsdf.ptr, tsdf_weights.ptr, depth_frame.ptr,
           L.Int2(int(frame_dims[0]), int(frame_dims[1])), float(delta), L.i3(dims), L.f3(voxel_size),
           L.f3(volume_size), L.m3(world2camera_rotate), L.f3(world2camera_translate), L.m3(volume_rotate),
           L.f3(volume_translate), L.m3(depth_intrinsics), float(max_weight), _s(stream))
    _done(stream)


def run_translate_kernel(mesh: DeviceArray, mesh_size: int, translation, stream: Stream | None = None) -> None:
    """af::runTranslateKernel (include/af/VertexManipulation.cuh:16)."""
    L.call("vf_translate", mesh.ptr, int(mesh_size), L.f3(translation), _s(stream))
    _done(stream)


def run_compute_warp_weights_kernel(knn_weights: DeviceArray, motion_graph_radiuses: DeviceArray,
                                    knn_idxs: DeviceArray, knn_dists: DeviceArray, mesh_size: int,
                                    graph_size: int = 0, stream: Stream | None = None) -> None:
    """af::runComputeWarpWeightsKernel (include/af/VertexManipulation.cuh:19-23).
    graph_size 0 leaves indices unchecked, as the reference does."""
    L.call("vf_compute_warp_weights", _k_of(knn_idxs), knn_weights.ptr, motion_graph_radiuses.ptr, knn_idxs.ptr,
           knn_dists.ptr, int(mesh_size), int(graph_size), _s(stream))
    _done(stream)


def run_warp_kernel(warped_mesh: DeviceArray, mesh: DeviceArray, motion_graph: DeviceArray, knn_idxs: DeviceArray,
                    knn_weights: DeviceArray, mesh_size: int, graph_size: int = 0,
                    stream: Stream | None = None) -> None:
    """af::runWarpKernel (include/af/ComputeCorrespondence.cuh:11-16)."""
    L.call("vf_warp_mesh", _k_of(knn_idxs), warped_mesh.ptr, mesh.ptr, motion_graph.ptr, knn_idxs.ptr,
           knn_weights.ptr, int(mesh_size), int(graph_size), _s(stream))
    _done(stream)


def get_knn_data(ws: "Workspace", k: int, points: DeviceArray, nodes: DeviceArray, exclude_closest: bool = False,
                 stream: Stream | None = None) -> tuple:
    """af::getKnnData<K> (include/af/GraphBuilder.h:20-32) on the device:
    (idx [n, k] uint32, dist [n, k] float32) device arrays."""
    n = points.shape[0]
    idx = DeviceArray((n, k), np.uint32, stream)
    dist = DeviceArray((n, k), np.float32, stream)
    L.call("vf_mesh_knn", ws.handle, int(k), idx.ptr, dist.ptr, points.ptr, int(n), nodes.ptr, int(nodes.shape[0]),
           1 if exclude_closest else 0, _s(stream))
    _done(stream)
    return idx, dist


def warp_points(ws: "Workspace", k: int, points: DeviceArray, graph: "DeviceMotionGraph",
                out: DeviceArray | None = None, blend: int = L.BLEND_LBS, stream: Stream | None = None) -> DeviceArray:
    """getKnnData + runComputeWarpWeightsKernel + runWarpKernel in one kernel;
    out may be points itself."""
    n = points.shape[0]
    if out is None:
        out = DeviceArray((n, 3), np.float32, stream)
    L.call("vf_warp_points", ws.handle, int(k), int(blend), out.ptr, points.ptr, int(n), graph.points.ptr,
           graph.radii.ptr, graph.transforms.ptr, int(graph.count), _s(stream))
    _done(stream)
    return out


def run_filter_depth_kernel(depth: DeviceArray, depth_size: int, threshold: float, fill_value: float = np.inf,
                            stream: Stream | None = None) -> None:
    """af::runFilterDepthKernel (include/af/VertexManipulation.cuh:11)."""
    L.call("vf_filter_depth", depth.ptr, int(depth_size), float(threshold), float(fill_value), _s(stream))
    _done(stream)


def run_mask_depth_kernel(depth: DeviceArray, mask: DeviceArray, depth_size: int,
                          stream: Stream | None = None) -> None:
    """af::runMaskDepthKernel (include/af/VertexManipulation.cuh:12); mask is uint8."""
    L.call("vf_mask_depth", depth.ptr, mask.ptr, int(depth_size), _s(stream))
    _done(stream)


def bilateral_filter_texture_opm_shared(output: DeviceArray, input: DeviceArray, height: int, width: int, r: int,
                                        sigma_i: float, sigma_s: float, threshold: float,
                                        stream: Stream | None = None) -> None:
    """af::bilateralFilterTextureOpmShared (src/BilateralFilter.cu:185-193)."""
    L.call("vf_bilateral_filter", output.ptr, input.ptr, int(height), int(width), int(r), float(sigma_i),
           float(sigma_s), float(threshold), _s(stream))
    _done(stream)


def run_depth_to_mesh_kernel(mesh: DeviceArray, depth: DeviceArray, width: int, height: int, intrinsics,
                             stream: Stream | None = None) -> None:
    """af::runDepthToMeshKernel (include/af/VertexManipulation.cuh:14)."""
    L.call("vf_depth_to_mesh", mesh.ptr, depth.ptr, int(width), int(height), L.m3(intrinsics), _s(stream))
    _done(stream)


def run_depth_vertices_normals_kernel(normals: DeviceArray, vertices: DeviceArray, depth: DeviceArray, h: int,
                                      w: int, angle_range=(1.3, 2.1), stream: Stream | None = None) -> None:
    """af::runDepthVerticesNormalsKernel (include/af/VertexManipulation.cuh:25-30)."""
    L.call("vf_depth_normals", normals.ptr, vertices.ptr, depth.ptr, int(h), int(w), float(angle_range[0]),
           float(angle_range[1]), _s(stream))
    _done(stream)


def run_remove_depth_without_normals_kernel(normals: DeviceArray, vertices: DeviceArray, depth: DeviceArray, h: int,
                                            w: int, stream: Stream | None = None) -> None:
    """af::runRemoveDepthWithoutNormalsKernel (include/af/VertexManipulation.cuh:32)."""
    L.call("vf_remove_depth_without_normals", normals.ptr, vertices.ptr, depth.ptr, int(h), int(w), _s(stream))
    _done(stream)


def depth_centroid(ws: "Workspace | None", mesh: DeviceArray, stream: Stream | None = None) -> tuple:
    """Frame centroid (src/VolumeFusion.cu:216-219): (float32[3], count)."""
    c, n = L.Float3(), C.c_uint64(0)
    L.call("vf_depth_centroid", ws.handle if ws is not None else None, C.byref(c), C.byref(n), mesh.ptr,
           int(mesh.shape[0]), _s(stream))
    return np.array([c.x, c.y, c.z], np.float32), int(n.value)


def ingest_depth(ws: "Workspace", depth: DeviceArray, intrinsics, threshold: float = 1.0, bilateral=(5, 75.0, 75.0,
                 0.025), angle_range=(1.3, 2.1), stream: Stream | None = None) -> dict:
    """The frame ingest of src/VolumeFusion.cu:190-219 in three launches;
    depth [H, W] is updated in place.  bilateral=None skips the filter."""
    H, W = depth.shape
    verts = DeviceArray((H * W, 3), np.float32, stream)
    normals = DeviceArray((H * W, 3), np.float32, stream)
    normals.fill(0.0)
    r, si, ss, bt = bilateral if bilateral is not None else (0, 1.0, 1.0, 0.0)
    c, n = L.Float3(), C.c_uint64(0)
    L.call("vf_ingest_depth", ws.handle, depth.ptr, int(W), int(H), L.m3(intrinsics), float(threshold), int(r),
           float(si), float(ss), float(bt), float(angle_range[0]), float(angle_range[1]), verts.ptr, normals.ptr,
           C.byref(c), C.byref(n), _s(stream))
    return {"vertices": verts, "normals": normals, "centroid": np.array([c.x, c.y, c.z], np.float32),
            "count": int(n.value)}


def marching_cubes(ws: "Workspace", tsdf: DeviceArray, weights: DeviceArray, dims, voxel_size, volume_size,
                   iso: float = 0.0, full: bool = False, capacity: int | None = None,
                   stream: Stream | None = None) -> tuple:
    """af::runMarchingCubesMeshKernel / runMarchingCubesFullKernel
    (include/af/MarchingCubes.cuh:10-28) into fresh buffers: (vertices
    [n, 3] float32, faces [m, 3] int32 or None) on the host, canonical order."""
    X, Y, Z = (int(d) for d in dims)
    if capacity is None:
        capacity = 15 * max(X - 2, 0) * max(Y - 2, 0) * max(Z - 2, 0) if full else \
            12 * max(X - 2, 0) * max(Y - 2, 0) * max(Z - 2, 0)
        capacity = max(min(capacity, 1 << 26), 1)
    mesh = DeviceArray((capacity, 3), np.float32, stream)
    counts = DeviceArray((2,), np.uint32, stream)
    counts.upload(np.zeros(2, np.uint32))
    if full:
        faces = DeviceArray((capacity // 3 + 1, 3), np.int32, stream)
        L.call("vf_marching_cubes_full", ws.handle, mesh.ptr, counts.ptr, int(capacity), faces.ptr, counts.ptr + 4,
               int(faces.shape[0]), tsdf.ptr, weights.ptr, L.i3(dims), L.f3(voxel_size), L.f3(volume_size),
               float(iso), _s(stream))
    else:
        L.call("vf_marching_cubes_mesh", ws.handle, mesh.ptr, counts.ptr, int(capacity), tsdf.ptr, weights.ptr,
               L.i3(dims), L.f3(voxel_size), L.f3(volume_size), float(iso), _s(stream))
    _done(stream)
    nv, nf = (int(x) for x in counts.download())
    if nv > capacity:
        raise L.VFError(6, "marching_cubes", f"{nv} vertices exceed the capacity {capacity}")
    verts = mesh.download()[:nv]
    return verts, (faces.download()[:nf] if full else None)


# ---------------------------------------------------------------------------
# Reference-layout device containers
# ---------------------------------------------------------------------------

class DeviceTsdf:
    """af::DeviceTsdf<float> (include/af/DeviceTsdf.cuh:13-84): tsdf initialised
    to -1, weights to 0, voxel size = size / dims."""

    def __init__(self, vol_size, vol_dim, center=(0.0, 0.0, 0.0), rotation=IDENTITY3, delta: float = 0.01,
                 stream: Stream | None = None):
        self.vol_dim = tuple(int(d) for d in vol_dim)
        self.vol_size = tuple(float(np.float32(s)) for s in vol_size)
        self.center = tuple(float(np.float32(c)) for c in center)
        self.rotation = tuple(float(np.float32(r)) for r in rotation)
        self.delta = float(delta)
        n = self.voxel_count()
        self.tsdf = DeviceArray((n,), np.float32, stream)
        self.weights = DeviceArray((n,), np.float32, stream)
        self.init_default()

    def voxel_count(self) -> int:
        return self.vol_dim[0] * self.vol_dim[1] * self.vol_dim[2]

    def voxel_size(self):
        """DeviceTsdf::voxelSize(): volSize.cwiseQuotient(volDim.cast<float>()) in float32."""
        return tuple(float(np.float32(s) / np.float32(d)) for s, d in zip(self.vol_size, self.vol_dim))

    def init_default(self) -> None:
        self.tsdf.fill(-1.0)
        self.weights.fill(0.0)

    def save(self, path: str, fmt: int = L.TSDF_FORMAT_NATIVE, max_weight: float = 15.0) -> None:
        """Checkpoint to disk (the reference's TSDFVolume::save layout with
        fmt=TSDF_FORMAT_LEGACY, src/TSDFVolume.cpp:339-358)."""
        from . import tsdf_io

        tsdf_io.save_device(path, self, fmt, max_weight)

    def load(self, path: str) -> dict:
        """Resume from disk; the file's dims must fit this volume."""
        from . import tsdf_io

        return tsdf_io.load_device(path, self)

    def integrate(self, depth: DeviceArray, intrinsics, max_weight: float = 15.0,
                  stream: Stream | None = None) -> None:
        """DeviceTsdf::integrate (include/af/DeviceTsdf.cuh:28-32)."""
        h, w = depth.shape
        run_integrate_tsdf_from_frame_kernel(self.tsdf, self.weights, depth, (w, h), self.delta, self.vol_dim,
                                             self.voxel_size(), self.vol_size, self.rotation, self.center,
                                             intrinsics, max_weight, stream)


class DeviceMotionGraph:
    """af::DeviceMotionGraph (include/af/DeviceMotionGraph.cuh:10-31): node
    positions Vec3f[G], transforms Mat4f[G] (column-major), radii float[G]."""

    def __init__(self, points: np.ndarray, transforms: np.ndarray, radii: np.ndarray,
                 stream: Stream | None = None):
        points = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        self.count = points.shape[0]
        self.points = DeviceArray.from_numpy(points, stream)
        self.transforms = DeviceArray.from_numpy(np.ascontiguousarray(transforms, np.float32).reshape(-1, 16), stream)
        self.radii = DeviceArray.from_numpy(np.ascontiguousarray(radii, np.float32).reshape(-1), stream)


# ---------------------------------------------------------------------------
# Fused per-frame path
# ---------------------------------------------------------------------------

class Workspace:
    """Device scratch of the fused path (vf_workspace); grows on demand."""

    def __init__(self, max_volumes: int = 1, max_total_voxels: int = 0, max_total_nodes: int = 0,
                 stream: Stream | None = None):
        self.handle = C.c_void_p(None)
        L.call("vf_workspace_create", C.byref(self.handle), int(max_volumes), int(max_total_voxels),
               int(max_total_nodes), _s(stream))

    def bytes(self) -> int:
        return int(L.load().vf_workspace_bytes(self.handle))

    def last_stats(self, stream: Stream | None = None) -> dict:
        st = L.FrameStats()
        L.call("vf_workspace_last_stats", self.handle, C.byref(st), _s(stream))
        return st.as_dict()

    def totals(self, reset: bool = False, stream: Stream | None = None) -> dict:
        st = L.FrameStats()
        L.call("vf_workspace_totals", self.handle, C.byref(st), 1 if reset else 0, _s(stream))
        return st.as_dict()

    def set_timing(self, enable, pass_a_only: bool = False) -> None:
        """Kernel timing on (all three kernel groups, or pass A alone) or off."""
        L.call("vf_workspace_set_timing", self.handle, (2 if pass_a_only else 1) if enable else 0)

    def kernel_ms(self) -> tuple:
        """(k_prep_bin, k_pass_a, k_pass_b) summed milliseconds, frames timed."""
        ms = (C.c_float * 3)()
        n = C.c_int(0)
        L.call("vf_workspace_kernel_ms", self.handle, ms, C.byref(n))
        return (float(ms[0]), float(ms[1]), float(ms[2])), int(n.value)

    def set_tuning(self, brick_shape: int = 0, wg_per_cu: int = 0, screen: bool = True) -> None:
        """Brick shape (index into _lib.BRICK_SHAPES), pass-A workgroups per
        CU (0 = occupancy limit) and the brick screen for later calls; tsdf,
        weights and node positions do not change (with the screen, the
        supported and in-bounds counts cover the swept bricks only)."""
        t = L.Tuning(int(brick_shape), int(wg_per_cu), 0.0, 0 if screen else L.TUNE_NO_SCREEN)
        L.call("vf_workspace_set_tuning", self.handle, C.byref(t))

    def set_record_budget(self, records_per_voxel: float) -> None:
        """Room for ceil(records_per_voxel x voxels) pending updates per frame
        (0: the worst case); bricks beyond it are recomputed by a further
        pass with the same results."""
        L.call("vf_workspace_set_record_budget", self.handle, float(records_per_voxel))

    def close(self) -> None:
        if self.handle:
            L.load().vf_workspace_destroy(self.handle)
            self.handle = C.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def voxels_per_node(motion_graph_radius: float, voxel_size) -> tuple:
    """Node cube of src/VolumeFusion.cu:702-705: 2 * (int(R / vs) + 1) + 1."""
    r = np.float32(motion_graph_radius)
    out = []
    for v in voxel_size:
        q = np.float32(r / np.float32(v))
        out.append(2 * (int(q) + 1) + 1)
    return tuple(out)


def make_volume(tsdf: DeviceTsdf) -> L.Volume:
    v = L.Volume()
    v.tsdf = tsdf.tsdf.ptr
    v.weights = tsdf.weights.ptr
    v.dim = L.i3(tsdf.vol_dim)
    v.voxel_size = L.f3(tsdf.voxel_size())
    v.size = L.f3(tsdf.vol_size)
    v.center = L.f3(tsdf.center)
    return v


def make_graph(g: DeviceMotionGraph, write_back: bool = True) -> L.Graph:
    gr = L.Graph()
    gr.points = g.points.ptr
    gr.points_out = g.points.ptr if write_back else None
    gr.transforms = g.transforms.ptr
    gr.radii = g.radii.ptr
    gr.count = g.count
    return gr


def make_frame(depth: DeviceArray, intrinsics) -> L.Frame:
    f = L.Frame()
    f.depth = depth.ptr
    f.height, f.width = int(depth.shape[0]), int(depth.shape[1])
    f.intrinsics = L.m3(intrinsics)
    return f


def make_params(k: int, delta: float, max_weight: float, cube, blend: int = L.BLEND_LBS) -> L.WarpParams:
    p = L.WarpParams()
    p.k = int(k)
    p.delta = float(delta)
    p.max_weight = float(max_weight)
    p.voxels_per_node = L.i3(cube)
    p.blend = int(blend)
    return p


def warped_integrate(ws: Workspace, tsdf: DeviceTsdf, graph: DeviceMotionGraph, depth: DeviceArray, intrinsics,
                     k: int, delta: float, max_weight: float, cube, write_back: bool = True,
                     stream: Stream | None = None, sync: bool = True, blend: int = L.BLEND_LBS) -> None:
    """The reference's per-frame integration block (src/VolumeFusion.cu:702-742)
    as one call: volume centre = tsdf.center, rotation I."""
    vol = make_volume(tsdf)
    gr = make_graph(graph, write_back)
    fr = make_frame(depth, intrinsics)
    p = make_params(k, delta, max_weight, cube, blend)
    L.call("vf_warped_integrate", ws.handle, C.byref(vol), C.byref(gr), C.byref(fr), C.byref(p), _s(stream))
    if sync:
        _done(stream)


def warped_integrate_batch(ws: Workspace, tsdfs, graphs, depths, intrinsics_list, k: int, delta: float,
                           max_weight: float, cubes, write_back: bool = True, stream: Stream | None = None,
                           sync: bool = True, blend: int = L.BLEND_LBS) -> None:
    """Many independent volumes, one frame each, in one set of launches."""
    n = len(tsdfs)
    vols = (L.Volume * n)(*[make_volume(t) for t in tsdfs])
    grs = (L.Graph * n)(*[make_graph(g, write_back) for g in graphs])
    frs = (L.Frame * n)(*[make_frame(d, K) for d, K in zip(depths, intrinsics_list)])
    deltas = delta if isinstance(delta, (list, tuple)) else [delta] * n
    maxws = max_weight if isinstance(max_weight, (list, tuple)) else [max_weight] * n
    blends = blend if isinstance(blend, (list, tuple)) else [blend] * n
    ps = (L.WarpParams * n)(*[make_params(k, d, m, c, b) for d, m, c, b in zip(deltas, maxws, cubes, blends)])
    L.call("vf_warped_integrate_batch", ws.handle, n, vols, grs, frs, ps, _s(stream))
    if sync:
        _done(stream)


def _tuning_dict(t) -> dict:
    return {"brick_shape": int(t.brick_shape), "brick": L.BRICK_SHAPES[t.brick_shape],
            "wg_per_cu": int(t.wg_per_cu), "ms_per_frame": float(t.ms_per_frame),
            "screen": not (int(t.flags) & L.TUNE_NO_SCREEN)}


class Plan:
    """A fixed batch of volumes replayed every frame (vf_plan): descriptors are
    uploaded once and, with graph=True, the five kernels run as one recorded
    HIP graph.  The device buffers of every argument must stay alive and in
    place; their contents may change between launches."""

    def __init__(self, tsdfs, graphs, depths, intrinsics_list, k: int, delta, max_weight, cubes,
                 write_back: bool = True, graph: bool = True, tuning: tuple | None = None,
                 transforms=None, stream: Stream | None = None, blend: int = L.BLEND_LBS):
        """transforms: optional per-volume device arrays used instead of the
        graphs' own transform buffers."""
        n = len(tsdfs)
        self.n = n
        self._keep = (list(tsdfs), list(graphs), list(depths), list(transforms or []))
        self.vols = (L.Volume * n)(*[make_volume(t) for t in tsdfs])
        grs = [make_graph(g, write_back) for g in graphs]
        if transforms is not None:
            for gr, T in zip(grs, transforms):
                gr.transforms = T.ptr
        self.graphs = (L.Graph * n)(*grs)
        self.frames = (L.Frame * n)(*[make_frame(d, K) for d, K in zip(depths, intrinsics_list)])
        deltas = delta if isinstance(delta, (list, tuple)) else [delta] * n
        maxws = max_weight if isinstance(max_weight, (list, tuple)) else [max_weight] * n
        self.params = (L.WarpParams * n)(*[make_params(k, d, m, c, blend) for d, m, c in zip(deltas, maxws, cubes)])
        self.handle = C.c_void_p(None)
        t = None
        if tuning is not None:
            screen = tuning[2] if len(tuning) > 2 else True
            t = L.Tuning(int(tuning[0]), int(tuning[1]), 0.0, 0 if screen else L.TUNE_NO_SCREEN)
        L.call("vf_plan_create", C.byref(self.handle), n, self.vols, self.graphs, self.frames, self.params,
               C.byref(t) if t is not None else None, L.PLAN_GRAPH if graph else 0, _s(stream))

    def launch(self, stream: Stream | None = None, sync: bool = False) -> None:
        L.call("vf_plan_launch", self.handle, _s(stream))
        if sync:
            _done(stream)

    def autotune(self, frames: int = 3, stream: Stream | None = None) -> dict:
        t = L.Tuning()
        L.call("vf_plan_autotune", self.handle, int(frames), C.byref(t), _s(stream))
        return _tuning_dict(t)

    def tuning(self) -> dict:
        t = L.Tuning()
        L.call("vf_plan_get_tuning", self.handle, C.byref(t))
        return _tuning_dict(t)

    def stats(self, reset: bool = False, stream: Stream | None = None) -> tuple:
        last, tot = L.FrameStats(), L.FrameStats()
        L.call("vf_plan_stats", self.handle, C.byref(last), C.byref(tot), 1 if reset else 0, _s(stream))
        return last.as_dict(), tot.as_dict()

    def bytes(self) -> int:
        return int(L.load().vf_plan_bytes(self.handle))

    def close(self) -> None:
        if self.handle:
            L.load().vf_plan_destroy(self.handle)
            self.handle = C.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
