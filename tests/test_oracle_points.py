"""The mesh-point path (knn, weights, warp) of the oracle: the device
semantics of src/VertexManipulation.cu:105-145 and
src/ComputeCorrespondence.cu:5-43 with the knn of include/af/MotionGraph.h:50-70."""
import json
from pathlib import Path

import numpy as np
import pytest

F = np.float32
GOLDEN = Path(__file__).resolve().parent / "golden"


def _golden():
    g = json.loads((GOLDEN / "motiongraph_warp.json").read_text())
    T = np.array(g["transforms_row_major"], F)
    return g, np.ascontiguousarray(T.transpose(0, 2, 1).reshape(-1, 16))


@pytest.mark.parametrize("k,exclude", [(1, False), (3, False), (6, False), (8, False), (6, True)])
def test_mesh_knn_brute_force(oracle, k, exclude):
    """Against numpy: float32 squared distance (dx^2 + dy^2) + dz^2, sqrt,
    sorted by (distance, index); points placed on nodes and exact ties
    included."""
    rng = np.random.default_rng(k + 10 * exclude)
    nodes = rng.uniform(-1, 1, (50, 3)).astype(F)
    nodes[10] = nodes[3] + F(0.0)  # coincident nodes: an exact tie
    pts = np.concatenate([rng.uniform(-1, 1, (200, 3)).astype(F), nodes[:20]])
    idx, dist = oracle.mesh_knn(k, pts, nodes, exclude_closest=exclude)
    d = pts[:, None, :] - nodes[None, :, :]
    s2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    dd = np.sqrt(s2.astype(F)).astype(F)
    off = 1 if exclude else 0
    for i in range(pts.shape[0]):
        order = sorted(range(nodes.shape[0]), key=lambda j: (dd[i, j], j))[off:off + k]
        np.testing.assert_array_equal(idx[i], np.array(order, np.uint32))
        np.testing.assert_array_equal(dist[i], dd[i, order])


def test_mesh_knn_fills_missing_slots(oracle):
    nodes = np.array([[0, 0, 0], [1, 0, 0]], F)
    idx, dist = oracle.mesh_knn(4, np.array([[0.2, 0, 0]], F), nodes)
    assert list(idx[0]) == [0, 1, 0xFFFFFFFF, 0xFFFFFFFF]
    assert dist[0, 2] == np.finfo(F).max and dist[0, 3] == np.finfo(F).max


def test_warp_weights_closed_forms(oracle):
    """Cutoff d >= r -> 0; one valid slot -> 1; missing slots weigh 0; two
    equal distances -> 1/2 each."""
    r = np.array([0.5, 0.5, 0.25], F)
    big = np.finfo(F).max
    idx = np.array([[0, 1, 2], [0, 0xFFFFFFFF, 0xFFFFFFFF], [0, 1, 2], [2, 0, 1]], np.uint32)
    dist = np.array([[0.5, 0.6, 0.1], [0.2, big, big], [0.3, 0.3, 0.3], [0.25, 0.7, 0.8]], F)
    w = oracle.compute_warp_weights(3, r, idx, dist)
    np.testing.assert_array_equal(w[0], [0, 0, 1])
    np.testing.assert_array_equal(w[1], [1, 0, 0])
    assert w[2, 2] == 0 and w[2, 0] == w[2, 1] == F(0.5)
    np.testing.assert_array_equal(w[3], [0, 0, 0])  # every slot cut off: sum 0, left unnormalised


def test_warp_mesh_uses_the_four_term_reduction(oracle):
    """The point kernel materialises p as a Vec4f, so each row of T p is
    (a0 + a1) + (a2 + a3), unlike the voxel kernel's a0 + (a1 + a2) + a3.
    Values are chosen so the two orders round differently."""
    T = np.eye(4, dtype=F)
    T[0, :] = [1.0, 1.0, 1.0, 0.0]
    Tcm = np.ascontiguousarray(T.T.reshape(1, 16))
    h = F(2.0 ** -24)  # half an ulp of 1
    p = np.array([[1.0, h, h]], F)
    out = oracle.warp_mesh(1, p, Tcm, np.zeros((1, 1), np.uint32), np.ones((1, 1), F), 1)
    a = [F(1.0) * F(1.0), F(1.0) * h, F(1.0) * h, F(0.0) * F(1.0)]
    four = (a[0] + a[1]) + (a[2] + a[3])
    three = (a[0] + (a[1] + a[2])) + a[3]
    assert four != three
    assert out[0, 0] == four


def test_point_path_agrees_with_host_warp(oracle):
    """The device point path (float Gaussian, vfo_expw, float division) and the
    host MotionGraph::warp (binary64 Gaussian), both with the sqrt distance,
    agree to float rounding on random graphs."""
    rng = np.random.default_rng(5)
    G = 40
    nodes = rng.uniform(-0.3, 0.3, (G, 3)).astype(F)
    radii = np.full(G, 0.2, F)
    T = np.tile(np.eye(4, dtype=F).reshape(16), (G, 1))
    for j in range(G):
        M = np.eye(4)
        M[:3, 3] = rng.uniform(-0.01, 0.01, 3)
        T[j] = M.T.reshape(16)
    pts = rng.uniform(-0.3, 0.3, (500, 3)).astype(F)
    dev = oracle.warp_points(4, pts, nodes, radii, T)
    host = oracle.mesh_warp(pts, nodes, radii, T, 4, host_order=False, squared_dist=False)
    assert np.abs(dev - host).max() < 2e-6


def test_golden_inputs_through_point_path(oracle):
    """The MotionGraphTest.Warp inputs through the device point path: the
    first two points (no node within its radius) are left unchanged exactly;
    the others match the host path with sqrt distances to float rounding."""
    g, T = _golden()
    pts, nodes, radii = np.array(g["mesh"], F), np.array(g["nodes"], F), np.array(g["radii"], F)
    dev = oracle.warp_points(g["k"], pts, nodes, radii, T)
    host = oracle.mesh_warp(pts, nodes, radii, T, g["k"], host_order=False, squared_dist=False)
    np.testing.assert_array_equal(dev[:2], pts[:2])
    assert np.abs(dev - host).max() < 1e-5
