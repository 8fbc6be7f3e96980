"""GPU parity of the mesh-point path (knn, weights, warp; fused and one by
one) against the oracle, bit for bit."""
import json
from pathlib import Path

import numpy as np
import pytest

from volumefusion_amd import _lib as L
from volumefusion_amd import ops
from volumefusion_amd.device import DeviceArray

pytestmark = pytest.mark.gpu
F = np.float32
GOLDEN = Path(__file__).resolve().parent / "golden"


def dev(a):
    return DeviceArray.from_numpy(np.ascontiguousarray(a))


def random_graph(rng, G, n, spread=0.3):
    nodes = rng.uniform(-spread, spread, (G, 3)).astype(F)
    if G > 12:
        nodes[7] = nodes[2]  # coincident nodes: exact distance ties
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
    pts = rng.uniform(-spread, spread, (n, 3)).astype(F)
    m = min(10, G)
    pts[:m] = nodes[:m]  # points exactly on nodes
    return nodes, radii, T, pts


@pytest.mark.parametrize("k,exclude,G", [(1, False, 50), (4, False, 2500), (6, True, 300), (8, False, 3), (8, True, 9)])
def test_mesh_knn(vf, oracle, k, exclude, G):
    """Several LDS tiles (2500 nodes), fewer nodes than slots, exact ties."""
    rng = np.random.default_rng(k * 100 + G)
    nodes, radii, T, pts = random_graph(rng, G, 3000)
    idx, dist = ops.get_knn_data(ops.Workspace(), k, dev(pts), dev(nodes), exclude_closest=exclude)
    i_or, d_or = oracle.mesh_knn(k, pts, nodes, exclude_closest=exclude)
    np.testing.assert_array_equal(idx.download(), i_or)
    np.testing.assert_array_equal(dist.download(), d_or)


def _knn_case(oracle, k, pts, nodes, exclude=False):
    idx, dist = ops.get_knn_data(ops.Workspace(), k, dev(pts), dev(nodes), exclude_closest=exclude)
    i_or, d_or = oracle.mesh_knn(k, pts, nodes, exclude_closest=exclude)
    np.testing.assert_array_equal(idx.download(), i_or)
    np.testing.assert_array_equal(dist.download(), d_or)


@pytest.mark.parametrize("exclude", [False, True])
def test_mesh_knn_grid_geometry(vf, oracle, exclude):
    """The node grid's search against the brute force: points far outside
    the nodes' box (many rings), nodes on a plane and on a line (zero extent
    in some axes), all nodes at one position, a dense cluster beside sparse
    nodes, and points and nodes at infinity."""
    rng = np.random.default_rng(77)
    k = 8
    # far points
    nodes = rng.uniform(-0.3, 0.3, (3000, 3)).astype(F)
    far = rng.uniform(-1, 1, (500, 3)).astype(F) * F(20.0)
    _knn_case(oracle, k, np.concatenate([far, nodes[:50] + F(1e-3)]), nodes, exclude)
    # a plane and a line of nodes
    plane = nodes.copy()
    plane[:, 2] = F(0.1)
    _knn_case(oracle, k, rng.uniform(-0.4, 0.4, (2000, 3)).astype(F), plane, exclude)
    line = nodes.copy()
    line[:, 1:] = F(-0.2)
    _knn_case(oracle, k, rng.uniform(-0.4, 0.4, (2000, 3)).astype(F), line, exclude)
    # all nodes at one position: every distance ties, index order decides
    same = np.tile(np.array([[0.05, -0.1, 0.2]], F), (500, 1))
    _knn_case(oracle, k, rng.uniform(-0.3, 0.3, (500, 3)).astype(F), same, exclude)
    # a dense cluster and sparse nodes
    cluster = np.concatenate([rng.normal(0, 0.002, (4000, 3)), rng.uniform(-1, 1, (200, 3))]).astype(F)
    pts = np.concatenate([rng.normal(0, 0.01, (1500, 3)), rng.uniform(-1, 1, (1500, 3))]).astype(F)
    _knn_case(oracle, k, pts, cluster, exclude)
    # infinite coordinates: such nodes sit outside the grid, such points see every node at infinity
    inf_nodes = nodes[:600].copy()
    inf_nodes[::50, 0] = np.inf
    inf_nodes[1::50, 2] = -np.inf
    inf_pts = rng.uniform(-0.3, 0.3, (800, 3)).astype(F)
    inf_pts[::40, 1] = np.inf
    _knn_case(oracle, k, inf_pts, inf_nodes, exclude)


@pytest.mark.parametrize("k", [3, 6, 8])
def test_weights_and_warp_one_by_one(vf, oracle, k):
    rng = np.random.default_rng(k)
    G = 60
    nodes, radii, T, pts = random_graph(rng, G, 4000)
    i_or, d_or = oracle.mesh_knn(k, pts, nodes)
    w_or = oracle.compute_warp_weights(k, radii, i_or, d_or)
    out_or = oracle.warp_mesh(k, pts, T, i_or, w_or, G)
    idx, dist, rad = dev(i_or), dev(d_or), dev(radii)
    w = DeviceArray((pts.shape[0], k), F)
    ops.run_compute_warp_weights_kernel(w, rad, idx, dist, pts.shape[0])
    np.testing.assert_array_equal(w.download(), w_or)
    p = dev(pts)
    ops.run_warp_kernel(p, p, dev(T), idx, w, pts.shape[0])  # in place, as src/VolumeFusion.cu:103 calls it
    np.testing.assert_array_equal(p.download(), out_or)


@pytest.mark.parametrize("blend", [L.BLEND_LBS, L.BLEND_DQB])
@pytest.mark.parametrize("k", [1, 4, 8])
def test_warp_points_fused(vf, oracle, k, blend):
    rng = np.random.default_rng(10 + k + 7 * blend)
    G = 1500
    nodes, radii, T, pts = random_graph(rng, G, 20000)
    graph = ops.DeviceMotionGraph(nodes, T, radii)
    out = ops.warp_points(ops.Workspace(), k, dev(pts), graph, blend=blend).download()
    np.testing.assert_array_equal(out, oracle.warp_points(k, pts, nodes, radii, T, blend=blend))


def test_golden_inputs_fused(vf, oracle):
    g = json.loads((GOLDEN / "motiongraph_warp.json").read_text())
    T = np.ascontiguousarray(np.array(g["transforms_row_major"], F).transpose(0, 2, 1).reshape(-1, 16))
    pts, nodes, radii = np.array(g["mesh"], F), np.array(g["nodes"], F), np.array(g["radii"], F)
    graph = ops.DeviceMotionGraph(nodes, T, radii)
    p = dev(pts)
    ops.warp_points(ops.Workspace(), g["k"], p, graph, out=p)  # in place
    np.testing.assert_array_equal(p.download(), oracle.warp_points(g["k"], pts, nodes, radii, T))
