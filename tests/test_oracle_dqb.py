"""Dual-quaternion blending, the second blend mode (the reference blends Mat4f
linearly; DQB follows Kavan et al., ACM TOG 27(4), 2008).  With no reference
counterpart it is pinned by closed forms and by its binary64 version."""
import math

import numpy as np
import pytest

F = np.float32


def rot(axis, deg):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    t = math.radians(deg)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(t) * K + (1 - math.cos(t)) * K @ K


def mat4(R, t=(0, 0, 0)):
    M = np.eye(4)
    M[:3, :3] = R
    M[:3, 3] = t
    return np.ascontiguousarray(M.T.reshape(16), F)  # column-major


def test_identity_is_exact(oracle):
    dq = oracle.mat4_to_dq(mat4(np.eye(3)))
    np.testing.assert_array_equal(dq, [0, 0, 0, 1, 0, 0, 0, 0])
    rng = np.random.default_rng(0)
    for _ in range(100):
        p = rng.uniform(-2, 2, 3).astype(F)
        np.testing.assert_array_equal(oracle.dqb_apply([dq, dq], [F(0.25), F(0.75)], p), p)


@pytest.mark.parametrize("axis,deg", [((1, 2, 3), 40.0), ((1, 0, 0), 179.9), ((0, 1, 0), 179.9),
                                      ((0, 0, 1), 179.9), ((1, 1, 0), 120.0)])
def test_one_rigid_node_is_its_transform(oracle, axis, deg):
    """One node: R p + t, through every branch of the quaternion extraction
    (trace > 0, and each largest diagonal near 180 degrees)."""
    R = rot(axis, deg)
    t = np.array([0.03, -0.02, 0.01])
    M = mat4(R, t)
    Rf = M.reshape(4, 4).T[:3, :3].astype(np.float64)
    tf = M.reshape(4, 4).T[:3, 3].astype(np.float64)
    rng = np.random.default_rng(1)
    dq, dq64 = oracle.mat4_to_dq(M), oracle.mat4_to_dq_f64(M)
    for _ in range(50):
        p = rng.uniform(-1, 1, 3).astype(F)
        want = Rf @ p.astype(np.float64) + tf
        got64 = oracle.dqb_apply([dq64], [1.0], p.astype(np.float64), f64=True)
        got = oracle.dqb_apply([dq], [F(1)], p)
        assert np.abs(got64 - want).max() < 1e-6  # float32 input matrix is not exactly orthonormal
        assert np.abs(got.astype(np.float64) - got64).max() < 2e-6


def test_half_way_rotation(oracle):
    """Identity and a rotation by theta with equal weights blend to the
    rotation by theta/2 about the same axis."""
    axis = (0.3, -0.5, 0.8)
    dqs = [oracle.mat4_to_dq_f64(mat4(np.eye(3))), oracle.mat4_to_dq_f64(mat4(rot(axis, 70.0)))]
    R_half = rot(axis, 35.0)
    rng = np.random.default_rng(2)
    for _ in range(20):
        p = rng.uniform(-1, 1, 3)
        got = oracle.dqb_apply(dqs, [0.5, 0.5], p, f64=True)
        assert np.abs(got - R_half @ p).max() < 2e-7


def test_hemisphere_alignment(oracle):
    """+179 and -179 degrees about z blend to 180 degrees, not to 0: each
    quaternion is aligned with the first before summing."""
    dqs = [oracle.mat4_to_dq_f64(mat4(rot((0, 0, 1), 179.0))), oracle.mat4_to_dq_f64(mat4(rot((0, 0, 1), -179.0)))]
    got = oracle.dqb_apply(dqs, [0.5, 0.5], np.array([1.0, 0.0, 0.0]), f64=True)
    np.testing.assert_allclose(got, [-1.0, 0.0, 0.0], atol=1e-7)


def test_pure_translations_blend_linearly(oracle):
    t1, t2 = np.array([0.1, 0.0, -0.02]), np.array([-0.03, 0.05, 0.04])
    dqs = [oracle.mat4_to_dq_f64(mat4(np.eye(3), t1)), oracle.mat4_to_dq_f64(mat4(np.eye(3), t2))]
    p = np.array([0.2, -0.4, 0.9])
    for w1 in (0.0, 0.3, 0.5, 1.0):
        got = oracle.dqb_apply(dqs, [w1, 1 - w1], p, f64=True)
        np.testing.assert_allclose(got, p + w1 * F(1) * t1.astype(F) + (1 - w1) * t2.astype(F), atol=1e-7)


def test_float_path_within_bound_of_binary64(oracle):
    """Random rigid nodes and normalised weights: |float32 - binary64| stays
    below 4e-6 m for points within 1 m."""
    rng = np.random.default_rng(3)
    worst = 0.0
    for _ in range(300):
        n = int(rng.integers(1, 9))
        Ms = [mat4(rot(rng.normal(size=3), rng.uniform(-30, 30)), rng.uniform(-0.02, 0.02, 3)) for _ in range(n)]
        w = rng.uniform(0.1, 1.0, n)
        w = (w / w.sum()).astype(F)
        p = rng.uniform(-1, 1, 3).astype(F)
        a = oracle.dqb_apply([oracle.mat4_to_dq(M) for M in Ms], w, p)
        b = oracle.dqb_apply([oracle.mat4_to_dq_f64(M) for M in Ms], w.astype(np.float64), p.astype(np.float64),
                             f64=True)
        worst = max(worst, float(np.abs(a - b).max()))
    assert worst < 4e-6


@pytest.mark.parametrize("k", [4, 8])
def test_dqb_frame_node_major_equals_gather(oracle, k):
    from volumefusion_amd import scenes

    s = scenes.small_scene(seed=17, k=k, nodes=60)
    n = s.n_voxels
    ta, wa = np.full(n, -1, F), np.zeros(n, F)
    tb, wb = ta.copy(), wa.copy()
    na, nb = s.nodes.copy(), s.nodes.copy()
    for f in range(3):
        depth, T = s.depth(f), s.transforms(f)
        ra = oracle.warped_frame(k, ta, wa, na, s.radii, T, s.dim, s.voxel_size, s.size, s.cube, s.center, depth,
                                 s.delta, s.intrinsics(), s.max_weight, blend=1)
        rb = oracle.warped_frame_gather(k, tb, wb, nb, s.radii, T, s.dim, s.voxel_size, s.size, s.cube, s.center,
                                        depth, s.delta, s.intrinsics(), s.max_weight, nthreads=4, blend=1)
        assert ra.stats["updated"] == rb["updated"] and ra.stats["updated"] > 100
        np.testing.assert_array_equal(ta, tb)
        np.testing.assert_array_equal(wa, wb)


def test_dqb_frame_differs_from_lbs_and_f64_error_small(oracle):
    from volumefusion_amd import scenes

    s = scenes.small_scene(seed=18, k=4, nodes=60)
    n = s.n_voxels
    t32, w32 = np.full(n, -1, F), np.zeros(n, F)
    tl, wl = t32.copy(), w32.copy()
    t64, w64 = np.full(n, -1.0), np.zeros(n)
    nodes, nodes_l = s.nodes.copy(), s.nodes.copy()
    for f in range(3):
        before = nodes.copy()
        depth, T = s.depth(f), s.transforms(f)
        r = oracle.warped_frame(s.k, t32, w32, nodes, s.radii, T, s.dim, s.voxel_size, s.size, s.cube, s.center,
                                depth, s.delta, s.intrinsics(), s.max_weight, want=("cnt", "idx", "pix"), blend=1)
        oracle.warped_frame(s.k, tl, wl, nodes_l, s.radii, T, s.dim, s.voxel_size, s.size, s.cube, s.center, depth,
                            s.delta, s.intrinsics(), s.max_weight)
        oracle.warped_frame_f64(s.k, t64, w64, before, s.radii, T, s.dim, s.voxel_size, s.size, s.center, depth,
                                s.delta, s.max_weight, r.cnt, r.idx, r.pix, blend=1)
    assert not np.array_equal(t32, tl)
    np.testing.assert_array_equal(w32, w64.astype(F))
    m = w32 > 0
    err = np.abs(t32.astype(np.float64) - t64)[m]
    unit = float(np.spacing(F(0.6))) / s.delta
    assert err.max() <= 8 * unit
    assert math.sqrt(float(np.mean(err ** 2))) <= unit
