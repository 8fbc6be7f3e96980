"""GPU parity of the Gauss-Newton motion optimisation against the oracle:
residuals, the reference's dense Jacobians, the normal equations and the
twist update bit for bit; correspondences bit for bit; the LU solve and the
whole Gauss-Newton loop within the float LU's rounding."""
import ctypes as C

import numpy as np
import pytest

from volumefusion_amd import _lib as L
from volumefusion_amd import scenes
from volumefusion_amd.device import DeviceArray

pytestmark = pytest.mark.gpu
F = np.float32


def dev(a):
    return DeviceArray.from_numpy(np.ascontiguousarray(a))


class _Dev:
    """A GN problem's arrays on the device and its vf_gn_problem."""

    def __init__(self, p):
        self.p = p
        G = p.graph.shape[0]
        self.corr, self.canon, self.knn, self.kw = dev(p.corr), dev(p.canon), dev(p.knn), dev(p.kw)
        self.dmesh, self.normals, self.graph, self.gknn = dev(p.dmesh), dev(p.dnormals), dev(p.graph), dev(p.gknn)
        self.desc = L.GNProblem(self.corr.ptr, p.corr.shape[0], self.canon.ptr, p.knn.shape[1], self.knn.ptr,
                                self.kw.ptr, self.normals.ptr, self.graph.ptr, G, self.gknn.ptr, p.gknn.shape[1],
                                p.wD, p.wR)


def _warp(oracle, p, T):
    return oracle.warp_mesh(p.knn.shape[1], p.canon, T, p.knn, p.kw, p.graph.shape[0])


@pytest.mark.parametrize("seed", [1, 2])
def test_gn_rows_bit_exact(vf, oracle, seed):
    """Residuals, dense Jacobians (j_height x j_width of float[6]), the
    normal equations and the update: the oracle's bits."""
    p = scenes.gn_problem(seed, G=40, M=1500)
    d = _Dev(p)
    G, C_, R, K = p.graph.shape[0], p.corr.shape[0], p.gknn.shape[1], p.knn.shape[1]
    T = scenes.smooth_random_warps(p.graph, np.random.default_rng(seed), 3.0, 0.005)
    warped = _warp(oracle, p, T)
    dT, dw = dev(T), dev(warped)
    n_rows = C_ + 3 * G * R
    r = DeviceArray((n_rows,), F)
    r.upload(np.zeros(n_rows, F))
    L.call("vf_fill_r_depth", r.ptr, p.wD, d.corr.ptr, C_, dw.ptr, d.dmesh.ptr, d.normals.ptr, None)
    L.call("vf_fill_r_mreg", r.ptr + 4 * C_, p.wR, d.graph.ptr, dT.ptr, G, d.gknn.ptr, R, None)
    r_or = np.concatenate([oracle.gn_r_depth(p.wD, p.corr, warped, p.dmesh, p.dnormals),
                           oracle.gn_r_mreg(p.wR, p.graph, T, p.gknn)])
    np.testing.assert_array_equal(r.download(), r_or)
    J = DeviceArray((n_rows, G, 6), F)
    J.upload(np.zeros((n_rows, G, 6), F))
    L.call("vf_fill_j_depth", J.ptr, p.wD, n_rows, G, d.corr.ptr, C_, d.canon.ptr, K, d.knn.ptr, d.kw.ptr,
           d.dmesh.ptr, d.normals.ptr, None)
    L.call("vf_fill_j_mreg", J.ptr + 24 * C_ * G, p.wR, 3 * G * R, G, d.graph.ptr, dT.ptr, G, d.gknn.ptr, R, None)
    J_or = np.concatenate([oracle.gn_j_depth(p.wD, p.corr, p.canon, p.knn, p.kw, p.dnormals, G),
                           oracle.gn_j_mreg(p.wR, p.graph, p.gknn)])
    np.testing.assert_array_equal(J.download(), J_or)
    JTJ, JTr = DeviceArray((6 * G, 6 * G), F), DeviceArray((G, 6), F)
    L.call("vf_gn_normal_equations", None, JTJ.ptr, JTr.ptr, r.ptr, C.byref(d.desc), None)
    JTJ_or, JTr_or = oracle.gn_normal_eq(J_or, r_or, p.corr, p.knn, p.kw, p.gknn)
    np.testing.assert_array_equal(JTJ.download(), JTJ_or)
    np.testing.assert_array_equal(JTr.download(), JTr_or)
    u = (np.random.default_rng(seed + 9).normal(size=(G, 6)) * 1e-2).astype(F)
    du = dev(u)  # device buffers stay referenced until the call has returned
    L.call("vf_update_transforms", dT.ptr, du.ptr, G, None)
    np.testing.assert_array_equal(dT.download(), oracle.gn_update(T, u))


def test_gn_self_edges_and_absent_neighbours(vf, oracle):
    """Regularisation lists holding the node itself (the three pairs of a row
    coincide) and absent neighbours: normal equations bit-exact."""
    p = scenes.gn_problem(3, G=20, M=500)
    p.gknn[::3, 0] = np.arange(0, 20, 3, dtype=np.uint32)
    p.gknn[1::4, 2] = 0xFFFFFFFF
    d = _Dev(p)
    G, C_, R = p.graph.shape[0], p.corr.shape[0], p.gknn.shape[1]
    r_or = np.random.default_rng(3).normal(size=C_ + 3 * G * R).astype(F)
    J_or = np.concatenate([oracle.gn_j_depth(p.wD, p.corr, p.canon, p.knn, p.kw, p.dnormals, G),
                           oracle.gn_j_mreg(p.wR, p.graph, p.gknn)])
    JTJ_or, JTr_or = oracle.gn_normal_eq(J_or, r_or, p.corr, p.knn, p.kw, p.gknn)
    JTJ, JTr = DeviceArray((6 * G, 6 * G), F), DeviceArray((G, 6), F)
    dr = dev(r_or)
    L.call("vf_gn_normal_equations", None, JTJ.ptr, JTr.ptr, dr.ptr, C.byref(d.desc), None)
    np.testing.assert_array_equal(JTJ.download(), JTJ_or)
    np.testing.assert_array_equal(JTr.download(), JTr_or)


def test_lu_solve(vf, oracle):
    """linearSolverLUStable on a normal matrix: the float LU's solution within
    its rounding of the oracle's binary64 LU, a small normal-equation residual,
    A and b left unchanged; a singular matrix gives VF_ERR_SINGULAR."""
    p = scenes.gn_problem(4, G=30, M=1200)
    G = p.graph.shape[0]
    J = np.concatenate([oracle.gn_j_depth(p.wD, p.corr, p.canon, p.knn, p.kw, p.dnormals, G),
                        oracle.gn_j_mreg(p.wR, p.graph, p.gknn)])
    r = np.random.default_rng(4).normal(size=J.shape[0]).astype(F)
    A, b = oracle.gn_normal_eq(J, r, p.corr, p.knn, p.kw, p.gknn)
    n = 6 * G
    dA, db, dx = dev(A), dev(b.reshape(-1)), DeviceArray((n,), F)
    L.call("vf_lu_solve", n, dA.ptr, n, db.ptr, dx.ptr, None)
    x = dx.download().astype(np.float64)
    xo, info = oracle.lu_solve(A, b.reshape(-1))
    assert info == 0
    np.testing.assert_allclose(x, xo, rtol=1e-3, atol=1e-4 * np.abs(xo).max())
    res = A.astype(np.float64) @ x - b.reshape(-1)
    assert np.linalg.norm(res) <= 1e-4 * np.linalg.norm(b)
    np.testing.assert_array_equal(dA.download(), A)
    S = np.ones((4, 4), F)
    dS, d1, dx4 = dev(S), dev(np.ones(4, F)), DeviceArray((4,), F)
    rc = L.load().vf_lu_solve(4, dS.ptr, 4, d1.ptr, dx4.ptr, None)
    assert rc == 7  # VF_ERR_SINGULAR


def test_correspondence_pairs(vf, oracle):
    """Projective pairs of a warped mesh against a frame: the oracle's pairs in
    ascending mesh order, appended after the counter's value."""
    rng = np.random.default_rng(5)
    H, W = 144, 160
    K = np.array([126.0, 0, 0, 0, 126.0, 0, 80.0, 72.0, 1.0], F)
    v, u = np.mgrid[0:H, 0:W]
    depth = (0.8 + 0.1 * np.sin(u / 17.0) * np.cos(v / 13.0)).astype(F)
    depth[rng.random((H, W)) < 0.05] = 0.0
    fm = np.stack([(u - 80.0) / 126.0 * depth, (v - 72.0) / 126.0 * depth, depth], -1).astype(F).reshape(-1, 3)
    M = 20000
    src = rng.integers(0, H * W, M)
    pts = (fm[src] + rng.normal(scale=0.01, size=(M, 3))).astype(F)
    pts[::97, 2] = -0.5  # behind the camera
    want = oracle.corr_pairs(pts, K, depth, fm, 0.02)
    base = 17
    out = DeviceArray((base + M, 2), np.int32)
    out.upload(np.full((base + M, 2), -7, np.int32))
    cnt = DeviceArray((1,), np.uint32)
    cnt.upload(np.array([base], np.uint32))
    fd, dp, dd = dev(fm), dev(pts), dev(depth)
    L.call("vf_correspondence_pairs", None, out.ptr, cnt.ptr, dp.ptr, M, L.m3(K), dd.ptr, L.Int2(H, W), fd.ptr, 0.02,
           None)
    n = int(cnt.download()[0])
    assert n == base + want.shape[0] and want.shape[0] > 5000
    got = out.download()
    np.testing.assert_array_equal(got[base:n], want)
    assert (got[:base] == -7).all()


def _oracle_loop(oracle, p, T, iters, min_step, float_lu=False, float_chol=False):
    """The Gauss-Newton loop of src/VolumeFusion.cu:541-691 on the oracle,
    with its LU in binary64 (float_lu: LAPACK's single-precision LU instead,
    float_chol: its single-precision Cholesky; the yardsticks for a float
    solver's rounding)."""
    G = p.graph.shape[0]
    J = np.concatenate([oracle.gn_j_depth(p.wD, p.corr, p.canon, p.knn, p.kw, p.dnormals, G),
                        oracle.gn_j_mreg(p.wR, p.graph, p.gknn)])
    T = T.copy()
    pending = T.copy()
    warped = _warp(oracle, p, pending)
    energy, energies, steps = np.inf, [], 0
    for it in range(iters):
        r = np.concatenate([oracle.gn_r_depth(p.wD, p.corr, warped, p.dmesh, p.dnormals),
                            oracle.gn_r_mreg(p.wR, p.graph, pending, p.gknn)])
        e = float((r.astype(np.float64) ** 2).sum())
        energies.append(e)
        if energy - e < min_step:
            break
        energy = e
        if it:
            T = pending.copy()
        JTJ, JTr = oracle.gn_normal_eq(J, r, p.corr, p.knn, p.kw, p.gknn)
        if float_lu:
            import scipy.linalg as sl

            x = sl.lu_solve(sl.lu_factor(JTJ), JTr.reshape(-1)).astype(np.float64)
        elif float_chol:
            import scipy.linalg as sl

            x = sl.cho_solve(sl.cho_factor(JTJ, lower=True), JTr.reshape(-1)).astype(np.float64)
        else:
            x, info = oracle.lu_solve(JTJ, JTr.reshape(-1))
            assert info == 0
        pending = oracle.gn_update(pending, x.reshape(G, 6).astype(F))
        warped = _warp(oracle, p, pending)
        steps += 1
    return T, pending, energies, steps


@pytest.mark.parametrize("factor", [0, 1], ids=["auto", "lu"])
@pytest.mark.parametrize("G,M", [(48, 3000), (49, 3000), (50, 3000), (200, 12000)])
def test_gn_solver_loop(vf, oracle, G, M, factor):
    """The Gauss-Newton loop on the device (JTJ factorised once) against the
    oracle's loop: the same number of steps, energies within 1e-3, the energy
    falling twentyfold or more, and transforms as close to the binary64 loop
    as a single-precision LAPACK solver gets (three times the larger deviation
    of its LU and its Cholesky factorisation, the one the device uses for
    this positive definite JTJ: JTJ's condition number here is about 3e5, so
    a float factorisation moves the update by up to about 1e-4).  The sizes
    cover the factorisation's blocking: 288 unknowns (one diagonal block in
    a 256-column half and a 32-column one), 300 (a last panel of 12
    columns) and 1200 (two look-ahead steps of 512 columns and a 176-column
    last block); 294 leaves the columns of
    the factor unaligned to 16 bytes (the transposed solve's scalar loads).
    factor 1 (VF_GN_FACTOR_LU) takes the reference's LU with partial
    pivoting instead: the permutation gather, the unit lower and the upper
    substitutions."""
    p = scenes.gn_problem(6, G=G, M=M, warp_angle_deg=3.0, warp_translation=0.006)
    T0 = np.tile(np.eye(4, dtype=F).reshape(16), (G, 1))
    iters, min_step = 4, 1e-7
    T_or, pend_or, e_or, steps_or = _oracle_loop(oracle, p, T0, iters, min_step)
    T_f, pend_f, _, steps_f = _oracle_loop(oracle, p, T0, iters, min_step, float_lu=True)
    T_c, pend_c, _, steps_c = _oracle_loop(oracle, p, T0, iters, min_step, float_chol=True)
    assert steps_f == steps_or and steps_c == steps_or
    tol_T = 3.0 * max(float(np.abs(T_f - T_or).max()), float(np.abs(T_c - T_or).max())) + 1e-6
    tol_P = 3.0 * max(float(np.abs(pend_f - pend_or).max()), float(np.abs(pend_c - pend_or).max())) + 1e-6
    d = _Dev(p)
    solver = C.c_void_p()
    L.call("vf_gn_solver_create_ex", C.byref(solver), C.byref(d.desc), factor, None)
    try:
        dT, dP = dev(T0), DeviceArray((G, 16), F)
        warped = dev(_warp(oracle, p, T0))
        rep = L.GNReport()
        L.call("vf_gn_solver_iterate", solver, dT.ptr, dP.ptr, warped.ptr, d.canon.ptr, p.canon.shape[0],
               d.dmesh.ptr, iters, min_step, C.byref(rep), None)
    finally:
        L.load().vf_gn_solver_destroy(solver)
    e_dev = [rep.energy[i] for i in range(len(e_or))]
    print(f"energies oracle {e_or}, device {e_dev}, steps {steps_or}/{rep.iterations}", flush=True)
    assert rep.iterations == steps_or
    if factor == 0:
        assert rep.cholesky == 1, "a positive definite JTJ is factorised by Cholesky"
    else:
        assert rep.cholesky == 0, "VF_GN_FACTOR_LU takes the LU"
    np.testing.assert_allclose(e_dev, e_or, rtol=1e-3, atol=1e-9)
    assert e_or[-1] < 0.05 * e_or[0]
    print(f"transform deviation from the binary64 loop: device {np.abs(dT.download() - T_or).max():.3g}, "
          f"float LAPACK LU or Cholesky {(tol_T - 1e-6) / 3:.3g}", flush=True)
    np.testing.assert_allclose(dT.download(), T_or, rtol=0, atol=tol_T)
    np.testing.assert_allclose(dP.download(), pend_or, rtol=0, atol=tol_P)


def test_gn_solver_singular_normal_matrix(vf):
    """A normal matrix that is not positive definite: with zero normals and a
    zero regularisation weight every node's rows have zero Jacobians, so JTJ
    is zero wherever a node has rows.  The Cholesky factorisation meets a
    zero pivot in its first diagonal block and leaves the later ones alone,
    the solver falls back to the LU with partial pivoting, which finds a zero
    pivot too: solver creation reports VF_ERR_SINGULAR (the reference's LU
    would fail on the same matrix).  522 unknowns: three diagonal blocks."""
    import dataclasses
    G = 87
    p = scenes.gn_problem(6, G=G, M=3000, warp_angle_deg=3.0, warp_translation=0.006)
    p = dataclasses.replace(p, dnormals=np.zeros_like(p.dnormals), wR=0.0)
    d = _Dev(p)
    solver = C.c_void_p()
    rc = L.load().vf_gn_solver_create(C.byref(solver), C.byref(d.desc), None)
    if rc == 0:
        L.load().vf_gn_solver_destroy(solver)
    assert rc == 7  # VF_ERR_SINGULAR
