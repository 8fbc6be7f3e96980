"""vf_warp_points_static, the point warp that allocates nothing (its scratch
is a static device array of the library), against vf_warp_points with a
workspace: the same bits on two calls in a row, the second with fewer points
and fewer nodes (the static scratch still holds the first call's data), in
both blend modes; an oversized graph is refused on the host."""
import ctypes as C

import numpy as np
import pytest

from volumefusion_amd import _lib as L
from volumefusion_amd import ops
from volumefusion_amd.device import DeviceArray

pytestmark = pytest.mark.gpu
F = np.float32


def _graph(rng, G, n):
    nodes = rng.uniform(-0.3, 0.3, (G, 3)).astype(F)
    nodes[min(7, G - 1)] = nodes[0]  # coincident nodes: exact distance ties
    radii = rng.uniform(0.05, 0.15, G).astype(F)
    T = np.zeros((G, 16), F)
    for j in range(G):
        a = rng.normal(size=3)
        a /= np.linalg.norm(a)
        t = np.radians(rng.uniform(-20, 20))
        K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
        M = np.eye(4)
        M[:3, :3] = np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * K @ K
        M[:3, 3] = rng.uniform(-0.02, 0.02, 3)
        T[j] = M.T.reshape(16)
    pts = rng.uniform(-0.3, 0.3, (n, 3)).astype(F)
    pts[:4] = nodes[:4]
    return nodes, radii, T, pts


def _both(k, blend, nodes, radii, T, pts):
    G, n = nodes.shape[0], pts.shape[0]
    d_nodes, d_radii, d_T, d_pts = (DeviceArray.from_numpy(np.ascontiguousarray(a)) for a in (nodes, radii, T, pts))
    out_ws = DeviceArray((n, 3), F)
    out_st = DeviceArray((n, 3), F)
    out_st.upload(np.full((n, 3), np.nan, F))
    ws = ops.Workspace()
    L.call("vf_warp_points", ws.handle, k, blend, out_ws.ptr, d_pts.ptr, n, d_nodes.ptr, d_radii.ptr, d_T.ptr, G, None)
    L.call("vf_warp_points_static", k, blend, out_st.ptr, d_pts.ptr, n, d_nodes.ptr, d_radii.ptr, d_T.ptr, G, None)
    return out_ws.download(), out_st.download()


@pytest.mark.parametrize("blend", [L.BLEND_LBS, L.BLEND_DQB])
def test_static_point_warp_equals_workspace_two_calls(vf, blend):
    rng = np.random.default_rng(70 + blend)
    for G, n in ((4000, 200_000), (1500, 50_000)):
        print(f"vf_warp_points_static: blend {blend}, {G} nodes, {n} points", flush=True)
        a, b = _both(8, blend, *_graph(rng, G, n))
        assert np.isfinite(a).all()
        np.testing.assert_array_equal(b, a)


def test_static_point_warp_at_node_limit(vf):
    """65536 nodes, the documented maximum, fit the static scratch."""
    rng = np.random.default_rng(73)
    a, b = _both(6, L.BLEND_LBS, *_graph(rng, 65536, 100_000))
    assert np.isfinite(a).all()
    np.testing.assert_array_equal(b, a)


def test_static_point_warp_refuses_oversized_graph(vf):
    rng = np.random.default_rng(72)
    G, n = 65537, 1000
    nodes = rng.uniform(-1, 1, (G, 3)).astype(F)
    d_nodes = DeviceArray.from_numpy(nodes)
    d_radii = DeviceArray.from_numpy(np.full(G, 0.1, F))
    d_T = DeviceArray.from_numpy(np.tile(np.eye(4, dtype=F).reshape(16), (G, 1)))
    d_pts = DeviceArray.from_numpy(rng.uniform(-1, 1, (n, 3)).astype(F))
    out = DeviceArray((n, 3), F)
    out.upload(np.full((n, 3), 7.0, F))
    rc = L.load().vf_warp_points_static(8, L.BLEND_LBS, out.ptr, d_pts.ptr, n, d_nodes.ptr, d_radii.ptr, d_T.ptr, G,
                                        None)
    assert rc == 6  # VF_ERR_CAPACITY
    assert b"static scratch" in L.load().vf_last_error()
    assert (out.download() == 7.0).all()
