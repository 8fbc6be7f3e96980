"""Depth-frame ingest in the oracle (src/VertexManipulation.cu:6-71,
147-232; src/BilateralFilter.cu; src/VolumeFusion.cu:190-219): closed forms."""
import math

import numpy as np
import pytest

F = np.float32
KIN = np.array([500.0, 0, 0, 0, 500.0, 0, 32.0, 24.0, 1.0], F)  # column-major


def test_filter_and_mask(oracle):
    d = np.array([0.5, 1.5, 5e-5, np.nan, 0.9999, 1.0], F)
    out = oracle.filter_depth(d, 1.0)
    assert out[0] == F(0.5) and np.isinf(out[1]) and np.isinf(out[2]) and np.isinf(out[3])
    assert out[4] == F(0.9999) and out[5] == F(1.0)
    np.testing.assert_array_equal(oracle.filter_depth(d, 1.0, fill=0.0)[[1, 2, 3]], [0, 0, 0])
    m = oracle.mask_depth(np.ones(4, F), np.array([1, 0, 1, 0], np.uint8))
    assert list(np.isinf(m)) == [False, True, False, True]


def test_depth_to_mesh_closed_form(oracle):
    H, W = 48, 64
    d = np.full((H, W), 2.0, F)
    d[5, 7] = np.inf
    d[6, 7] = 0.0
    mesh = oracle.depth_to_mesh(d, KIN).reshape(H, W, 3)
    np.testing.assert_array_equal(mesh[24, 32], [0.0, 0.0, 2.0])  # principal point
    x0 = (F(40) - F(32)) * (F(1) / F(500))
    y0 = (F(10) - F(24)) * (F(1) / F(500))
    np.testing.assert_array_equal(mesh[10, 40], [x0 * F(2), y0 * F(2), 2.0])
    np.testing.assert_array_equal(mesh[5, 7], 0)
    np.testing.assert_array_equal(mesh[6, 7], 0)


def plane_depth(H, W, n, d0):
    """Depth of the plane n . p = d0 seen through KIN."""
    ys, xs = np.mgrid[0:H, 0:W]
    rx = (xs - 32.0) / 500.0
    ry = (ys - 24.0) / 500.0
    z = d0 / (n[0] * rx + n[1] * ry + n[2])
    return z.astype(F)


def test_normals_of_planes(oracle):
    """A fronto-parallel plane gets (0, 0, -1) exactly (gY x gX of the unit
    image axes); a tilted plane gets its normal up to sign and rounding; a
    plane whose normal makes an angle in (1.3, 2.1) rad with +z is dropped."""
    H, W = 48, 64
    v = oracle.depth_to_mesh(np.full((H, W), 1.5, F), KIN)
    nrm = oracle.depth_normals(v, H, W, 1.3, 2.1).reshape(H, W, 3)
    np.testing.assert_array_equal(nrm[1:-1, 1:-1], np.broadcast_to([0, 0, -1], (H - 2, W - 2, 3)))
    assert (nrm[0] == 0).all() and (nrm[:, 0] == 0).all()  # border untouched (zero-initialised here)
    n = np.array([0.3, -0.2, 1.0])
    n /= np.linalg.norm(n)
    v = oracle.depth_to_mesh(plane_depth(H, W, n, 1.5), KIN)
    nrm = oracle.depth_normals(v, H, W, 1.3, 2.1).reshape(H, W, 3)[1:-1, 1:-1].reshape(-1, 3)
    err = np.minimum(np.abs(nrm - n).max(1), np.abs(nrm + n).max(1))
    assert err.max() < 1e-3
    # angle(acos(-n_z)) ~ 1.75 rad for a plane almost edge-on: excluded
    t = np.array([0.95, 0.0, 0.3122])
    t /= np.linalg.norm(t)
    v = oracle.depth_to_mesh(plane_depth(H, W, t, 0.5), KIN)
    nrm = oracle.depth_normals(v, H, W, 1.3, 2.1).reshape(H, W, 3)[1:-1, 1:-1]
    ok = np.isfinite(v.reshape(H, W, 3)[1:-1, 1:-1]).all(-1)
    assert (nrm[ok] == 0).all()


def test_remove_depth_without_normals(oracle):
    H, W = 6, 5
    nrm = np.zeros((H * W, 3), F)
    nrm[7] = [0, 0, 1]
    nrm[8] = [1e-5, 0, 0]  # within 1e-4 of zero: removed
    v = np.ones((H * W, 3), F)
    d = np.ones((H, W), F)
    v2, d2 = oracle.remove_depth_without_normals(nrm, v, d)
    inner = [y * W + x for y in range(1, H - 1) for x in range(1, W - 1)]
    for i in inner:
        if i == 7:
            assert (v2[i] == 1).all() and d2.ravel()[i] == 1
        else:
            assert (v2[i] == 0).all() and np.isinf(d2.ravel()[i])
    assert (v2[0] == 1).all()  # border untouched


def test_centroid(oracle):
    rng = np.random.default_rng(1)
    m = rng.uniform(-1, 1, (1000, 3)).astype(F)
    m[:, 2] = np.abs(m[:, 2])
    m[::7] = 0
    c, cnt = oracle.depth_centroid(m)
    assert cnt == int((m[:, 2] > F(1e-4)).sum())
    want = (m.astype(np.float64).sum(0)).astype(F) / F(cnt)
    assert np.abs(c - want).max() <= np.abs(want).max() * 2 ** -22


def test_bilateral_tables(oracle):
    g, c = oracle.bilateral_tables(5, 75.0)
    assert g.shape == (121,) and g[60] == F(1.0)  # centre weight expf(0)^2
    np.testing.assert_array_equal(g.reshape(11, 11), g.reshape(11, 11).T)
    assert c[0] == int(F(1.0) / F(math.exp(-4.5))) == 90 and c[30] == 1
    assert all(c[i] >= c[i + 1] for i in range(30))


def test_bilateral_closed_forms(oracle):
    """Constant regions stay constant (to float rounding of the 121-term
    weighted sum: 2e-6 relative, ~16 ulps); a step larger than the
    threshold is not blurred; infinite pixels stay infinite and do not leak."""
    H, W = 24, 30
    d = np.full((H, W), 1.25, F)
    d[:, 15:] = 1.40  # step of 0.15 > thr
    d[3, 4] = np.inf
    out = oracle.bilateral(d, 5, 75.0, 75.0, 0.025)
    assert np.isinf(out[3, 4])
    fin = np.isfinite(out)
    np.testing.assert_allclose(out[:, :15][fin[:, :15]], F(1.25), rtol=2e-6)
    np.testing.assert_allclose(out[:, 15:], F(1.40), rtol=2e-6)
