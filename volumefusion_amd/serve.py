"""Serving: resident reconstruction streams with continuous batching
(vf_server in include/volumefusion.h).

The reference reconstructs one stream per process (af::runVolumeFusion,
src/VolumeFusion.cu:223); here one device holds many streams' volumes and warp
fields and integrates their frames in batches as they arrive.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def stream_desc(dim, size, center, width: int, height: int, intrinsics, delta: float, max_weight: float, cube,
                points: np.ndarray, radii: np.ndarray, blend: int = L.BLEND_LBS) -> tuple:
    """A vf_stream_desc and the host arrays it points at (keep both alive
    until vf_server_open returns)."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    rad = np.ascontiguousarray(radii, np.float32).reshape(-1)
    d = L.StreamDesc()
    d.dim = L.i3(dim)
    d.size = L.f3(size)
    d.center = L.f3(center)
    d.width, d.height = int(width), int(height)
    d.intrinsics = L.m3(np.asarray(intrinsics, np.float32).reshape(-1))  # column-major Mat3f
    d.delta = float(delta)
    d.max_weight = float(max_weight)
    d.voxels_per_node = L.i3(cube)
    d.nodes = pts.shape[0]
    d.points = pts.ctypes.data
    d.radii = rad.ctypes.data
    d.blend = int(blend)
    return d, (pts, rad)


def scene_desc(s, blend: int = L.BLEND_LBS) -> tuple:
    """Stream description of a synthetic scene (volumefusion_amd.scenes)."""
    return stream_desc(s.dim, s.size, s.center, s.cam.width, s.cam.height, s.intrinsics(), s.delta,
                       s.max_weight, s.cube, s.nodes, s.radii, blend)


def stream_bytes(desc: L.StreamDesc, lanes: int) -> int:
    return int(L.load().vf_server_stream_bytes(C.byref(desc), int(lanes)))


class Server:
    def __init__(self, hbm_budget: int, k: int, max_batch: int = 32, lanes: int = 2, device: int = 0):
        cfg = L.ServerConfig(int(hbm_budget), int(k), int(max_batch), int(lanes), int(device))
        self.lanes = int(lanes)
        self.handle = C.c_void_p(None)
        L.call("vf_server_create", C.byref(self.handle), C.byref(cfg))
        self._dims = {}

    def open(self, desc: L.StreamDesc, wait_ms: int = 0) -> int:
        sid = C.c_int32(-1)
        L.call("vf_server_open", self.handle, C.byref(desc), int(wait_ms), C.byref(sid))
        self._dims[sid.value] = (desc.dim.x, desc.dim.y, desc.dim.z, desc.nodes, desc.width, desc.height)
        return sid.value

    def close(self, sid: int) -> None:
        L.call("vf_server_close", self.handle, int(sid))
        self._dims.pop(sid, None)

    def submit(self, sid: int, depth: np.ndarray, transforms: np.ndarray) -> int:
        """depth float32 [H, W]; transforms float32 [G, 16], each Mat4f
        column-major (Eigen's storage)."""
        X, Y, Z, G, W, H = self._dims[sid]
        d = np.ascontiguousarray(depth, np.float32)
        if d.size != W * H:
            raise ValueError(f"depth has {d.size} pixels, the stream expects {W}x{H}")
        T = np.ascontiguousarray(transforms, np.float32).reshape(-1, 16)
        if T.shape[0] != G:
            raise ValueError(f"{T.shape[0]} transforms, the stream has {G} nodes")
        t = C.c_uint64(0)
        L.call("vf_server_submit", self.handle, int(sid), d.ctypes.data, T.ctypes.data, C.byref(t))
        return int(t.value)

    def submit_borrowed(self, sid: int, depth, transforms) -> int:
        """Zero-copy submit of PinnedArray (or device array) inputs, which must
        stay untouched until the frame is done."""
        t = C.c_uint64(0)
        L.call("vf_server_submit_borrowed", self.handle, int(sid), depth.ptr, transforms.ptr, C.byref(t))
        return int(t.value)

    def wait(self, ticket: int) -> dict:
        r = L.FrameResult()
        L.call("vf_server_wait", self.handle, int(ticket), C.byref(r))
        return _result(r)

    def poll(self, capacity: int = 4096) -> list:
        buf = (L.FrameResult * capacity)()
        n = C.c_int(0)
        L.call("vf_server_poll", self.handle, buf, int(capacity), C.byref(n))
        return [_result(buf[i]) for i in range(n.value)]

    def drain(self) -> None:
        L.call("vf_server_drain", self.handle)

    def read(self, sid: int) -> tuple:
        X, Y, Z, G, W, H = self._dims[sid]
        t = np.empty(X * Y * Z, np.float32)
        w = np.empty(X * Y * Z, np.float32)
        p = np.empty((G, 3), np.float32)
        L.call("vf_server_read", self.handle, int(sid), t.ctypes.data, w.ctypes.data, p.ctypes.data)
        return t, w, p

    def save(self, sid: int, path: str, fmt: int = L.TSDF_FORMAT_NATIVE) -> None:
        L.call("vf_server_save", self.handle, int(sid), str(path).encode(), int(fmt))

    def restore(self, sid: int, path: str) -> None:
        L.call("vf_server_restore", self.handle, int(sid), str(path).encode())

    def stats(self) -> dict:
        st = L.ServerStats()
        L.call("vf_server_get_stats", self.handle, C.byref(st))
        out = {k: int(getattr(st, k)) for k, _ in L.ServerStats._fields_ if k != "totals"}
        out["totals"] = st.totals.as_dict()
        return out

    def close_server(self) -> None:
        if self.handle:
            L.load().vf_server_destroy(self.handle)
            self.handle = C.c_void_p(None)

    def __del__(self):
        try:
            self.close_server()
        except Exception:
            pass


def _result(r: L.FrameResult) -> dict:
    return {"ticket": int(r.ticket), "stream": int(r.stream_id), "status": int(r.status), "submit_s": r.submit_s,
            "issue_s": r.issue_s, "done_s": r.done_s, "latency_ms": (r.done_s - r.submit_s) * 1e3,
            "batch": int(r.batch), "batch_frames": int(r.batch_frames), "lane": int(r.lane)}
