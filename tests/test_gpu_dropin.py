"""GPU parity of the per-entry-point kernels (the reference's af::run*
functions) against the oracle, bit for bit.  Each stage gets the oracle's
inputs for that stage, so a mismatch names the stage."""
import numpy as np
import pytest

from volumefusion_amd import ops, scenes
from volumefusion_amd.device import DeviceArray

pytestmark = pytest.mark.gpu
F = np.float32
I3 = np.eye(3, dtype=F).reshape(9)


def dev(a):
    return DeviceArray.from_numpy(np.ascontiguousarray(a))


def _setup(k, seed=3, nodes=60, dim=(40, 36, 44)):
    s = scenes.small_scene(seed=seed, k=k, nodes=nodes, dim=dim)
    return s


def test_weight_function_bit_exact(vf, oracle):
    rng = np.random.default_rng(0)
    n = 200000
    r = rng.uniform(0.005, 0.05, n).astype(F)
    d = (r * rng.uniform(0.0, 2.0, n)).astype(F)
    out = DeviceArray((n,), F)
    g_r, g_d = dev(r), dev(d)  # keep the allocations alive across the call
    vf.call("vf_eval_warp_weight", g_r.ptr, g_d.ptr, out.ptr, n, None)
    got = out.download()
    L = oracle.lib()
    exp = np.array([L.vfo_weight_tsdf(float(a), float(b)) for a, b in zip(r, d)], F)
    np.testing.assert_array_equal(got, exp)


def test_translate_and_generate_vertices(vf, oracle):
    s = _setup(4)
    c = np.array(s.center, F)
    pts = dev(s.nodes)
    ops.run_translate_kernel(pts, s.nodes.shape[0], -c)
    np.testing.assert_array_equal(pts.download(), oracle.translate(s.nodes, -c))
    verts = DeviceArray((s.n_voxels, 3), F)
    ops.run_generate_tsdf_vertices_kernel(verts, s.voxel_size, s.size, s.dim)
    np.testing.assert_array_equal(verts.download(), oracle.generate_vertices(s.dim, s.voxel_size, s.size))


@pytest.mark.parametrize("k", [1, 4, 6, 8])
def test_knn_graph(vf, oracle, k):
    s = _setup(k, nodes=80)
    nodes_t = oracle.translate(s.nodes, -np.array(s.center, F))
    idx, dist, cnt, over = oracle.knn_graph(k, nodes_t, s.radii, s.dim, s.voxel_size, s.size, s.cube)
    n = s.n_voxels
    gi, gd, gc = DeviceArray((n, k), np.uint32), DeviceArray((n, k), F), DeviceArray((n,), np.int32)
    gi.fill(0), gd.fill(0), gc.fill(0)
    ops.run_tsdf_knn_graph(gi, gd, gc, dev(nodes_t), dev(s.radii), s.nodes.shape[0], s.voxel_size, s.size, s.dim,
                           s.cube)
    c = gc.download()
    np.testing.assert_array_equal(c, cnt)
    mask = np.arange(k)[None, :] < cnt[:, None]
    np.testing.assert_array_equal(gi.download()[mask], idx[mask])
    np.testing.assert_array_equal(gd.download()[mask], dist[mask])
    assert (cnt > 0).sum() > 1000
    if k <= 4:
        assert over > 0  # some voxels had more than k candidates


def test_knn_graph_spilled_brick_lists(vf, oracle):
    """More than 128 nodes overlapping one brick: the kernel walks every node
    in chunks instead of the brick list; same result."""
    rng = np.random.default_rng(9)
    dim, vs = (24, 20, 16), (0.005,) * 3
    size = tuple(float(F(d * v)) for d, v in zip(dim, vs))
    G = 300
    nodes = rng.uniform(-0.02, 0.02, (G, 3)).astype(F)
    radii = np.full(G, 0.03, F)
    cube = scenes.voxels_per_node(0.03, tuple(float(F(s) / F(d)) for s, d in zip(size, dim)))
    k = 8
    vsz = tuple(float(F(s) / F(d)) for s, d in zip(size, dim))
    idx, dist, cnt, over = oracle.knn_graph(k, nodes, radii, dim, vsz, size, cube)
    n = int(np.prod(dim))
    gi, gd, gc = DeviceArray((n, k), np.uint32), DeviceArray((n, k), F), DeviceArray((n,), np.int32)
    gi.fill(0), gd.fill(0), gc.fill(0)
    ops.run_tsdf_knn_graph(gi, gd, gc, dev(nodes), dev(radii), G, vsz, size, dim, cube)
    np.testing.assert_array_equal(gc.download(), cnt)
    mask = np.arange(k)[None, :] < cnt[:, None]
    np.testing.assert_array_equal(gi.download()[mask], idx[mask])
    np.testing.assert_array_equal(gd.download()[mask], dist[mask])
    assert over > 100


def test_knn_graph_nodes_on_voxels(vf, oracle):
    """Nodes exactly on voxel positions and a hair away from them: squared
    distances of 0, below 2^-96 and subnormal take the square root's slow
    path; the distances and the order they give match bit for bit."""
    dim, k = (24, 20, 16), 8
    vsz = (0.005,) * 3
    size = tuple(float(F(d) * F(v)) for d, v in zip(dim, vsz))
    half = np.array(size, F) * F(0.5)
    # voxel (12, 10, 8) sits at the origin; (5, 3, 9) at an exact float position
    on = (np.array([5, 3, 9], F) * np.array(vsz, F)).astype(F) - half
    nodes = np.array([[0, 0, 0], [1e-17, 0, 0], [0, -1e-20, 0], [0, 0, 3e-30], on,
                      on + np.array([0, 0, 1e-19], F), [1e-10, 1e-10, 1e-10]], F)
    rng = np.random.default_rng(4)
    nodes = np.concatenate([nodes, rng.uniform(-0.04, 0.04, (20, 3)).astype(F)])
    G = nodes.shape[0]
    radii = np.full(G, 0.02, F)
    cube = scenes.voxels_per_node(0.02, vsz)
    idx, dist, cnt, _ = oracle.knn_graph(k, nodes, radii, dim, vsz, size, cube)
    n = int(np.prod(dim))
    gi, gd, gc = DeviceArray((n, k), np.uint32), DeviceArray((n, k), F), DeviceArray((n,), np.int32)
    gi.fill(0), gd.fill(0), gc.fill(0)
    g_nodes, g_radii = dev(nodes), dev(radii)
    ops.run_tsdf_knn_graph(gi, gd, gc, g_nodes, g_radii, G, vsz, size, dim, cube)
    np.testing.assert_array_equal(gc.download(), cnt)
    mask = np.arange(k)[None, :] < cnt[:, None]
    np.testing.assert_array_equal(gi.download()[mask], idx[mask])
    np.testing.assert_array_equal(gd.download()[mask], dist[mask])
    d = dist[mask]
    assert (d == 0).any() and ((d > 0) & (d < 2.0 ** -48)).any()


def test_knn_graph_empty_and_outside(vf, oracle):
    s = _setup(4)
    n = s.n_voxels
    gi, gd, gc = DeviceArray((n, 4), np.uint32), DeviceArray((n, 4), F), DeviceArray((n,), np.int32)
    gc.fill(0)
    ops.run_tsdf_knn_graph(gi, gd, gc, dev(s.nodes), dev(s.radii), 0, s.voxel_size, s.size, s.dim, s.cube)
    assert (gc.download() == 0).all()
    far = (s.nodes + F(10.0)).astype(F)
    ops.run_tsdf_knn_graph(gi, gd, gc, dev(far), dev(s.radii), far.shape[0], s.voxel_size, s.size, s.dim, s.cube)
    assert (gc.download() == 0).all()


@pytest.mark.parametrize("k", [4, 8])
def test_weights_warp_collisions_integrate(vf, oracle, k):
    s = _setup(k, nodes=70)
    c = np.array(s.center, F)
    n = s.n_voxels
    nodes_t = oracle.translate(s.nodes, -c)
    idx, dist, cnt, _ = oracle.knn_graph(k, nodes_t, s.radii, s.dim, s.voxel_size, s.size, s.cube)
    T = s.transforms(0)
    # weights
    w = oracle.warp_weights(k, s.radii, idx, dist, cnt)
    gw = DeviceArray((n, k), F)
    gw.fill(0)
    g_idx, g_dist, g_cnt = dev(idx), dev(dist), dev(cnt)
    ops.run_compute_tsdf_warp_weights_kernel(gw, dev(s.radii), g_idx, g_dist, g_cnt, n)
    np.testing.assert_array_equal(gw.download(), w)
    # vertices + c, warp
    verts = oracle.translate(oracle.generate_vertices(s.dim, s.voxel_size, s.size), c)
    warped = oracle.warp_tsdf(k, verts, T, idx, w, cnt)
    gv = dev(verts)
    ops.run_warp_tsdf_kernel(gv, dev(T), g_idx, dev(w), g_cnt, s.voxel_size, s.size, s.dim)
    np.testing.assert_array_equal(gv.download(), warped)
    # collisions (exact int counts)
    col = oracle.count_collisions(warped, cnt, s.dim, s.voxel_size, s.size, I3, c)
    gcol = DeviceArray((n,), np.int32)
    gcol.fill(0)
    g_warped = dev(warped)
    ops.run_count_collisions_per_voxel_kernel(gcol, g_warped, g_cnt, s.voxel_size, s.size, s.dim, I3, c)
    np.testing.assert_array_equal(gcol.download(), col)
    assert (col > 1).sum() > 0
    # integrate warped, with some inf and zero depth pixels
    depth = s.depth(0)
    depth[::7, ::5] = np.inf
    tsdf0 = np.full(n, -1, F)
    wt0 = np.zeros(n, F)
    t, wt, upd = oracle.integrate_warped(tsdf0, wt0, warped, col, cnt, depth, s.delta, s.dim, s.voxel_size, s.size,
                                         I3, c, s.intrinsics(), s.max_weight)
    gt, gwt = dev(tsdf0), dev(wt0)
    H, W = depth.shape
    ops.run_integrate_warped_tsdf_from_frame_kernel(gt, gwt, g_warped, dev(col), g_cnt, dev(depth), (W, H), s.delta,
                                                    s.voxel_size, s.size, s.dim, I3, c, s.intrinsics(), s.max_weight)
    np.testing.assert_array_equal(gt.download(), t)
    np.testing.assert_array_equal(gwt.download(), wt)
    assert upd > 500


def test_collisions_and_integration_special_vertices(vf, oracle):
    """Warped vertices at NaN, infinities, beyond the int32 range and just
    outside the volume: the float -> int conversions saturate (NaN gives 0)
    in the target voxel and in the pixel, bit for bit with the oracle."""
    s = _setup(4, nodes=40)
    c = np.array(s.center, F)
    n = s.n_voxels
    rng = np.random.default_rng(8)
    verts = oracle.translate(oracle.generate_vertices(s.dim, s.voxel_size, s.size), c)
    cnt = np.ones(n, np.int32)
    cnt[rng.random(n) < 0.1] = 0
    half = np.array(s.size, F) * F(0.5)
    special = np.array([np.nan, np.inf, -np.inf, 3e9, -3e9, 2.0 ** 31, -(2.0 ** 31), 1e30, -1e30, 0.0, -0.0],
                       np.float64).astype(F)
    m = 4000
    idx = rng.choice(n, m, replace=False)
    for j, i in enumerate(idx):
        a = j % 3
        if j % 4 == 0:
            verts[i, a] = special[(j // 4) % special.shape[0]]
        elif j % 4 == 1:  # just below the volume's low face: truncation toward zero keeps voxel 0
            verts[i, a] = c[a] - half[a] - F(s.voxel_size[a]) * F(0.5)
        elif j % 4 == 2:
            verts[i, a] = c[a] + half[a] + F(1e-4)
        else:
            verts[i, :] = special[(j // 4) % special.shape[0]]
    col = oracle.count_collisions(verts, cnt, s.dim, s.voxel_size, s.size, I3, c)
    g_verts, g_cnt = dev(verts), dev(cnt)
    gcol = DeviceArray((n,), np.int32)
    gcol.fill(0)
    ops.run_count_collisions_per_voxel_kernel(gcol, g_verts, g_cnt, s.voxel_size, s.size, s.dim, I3, c)
    np.testing.assert_array_equal(gcol.download(), col)
    assert col[0] > 1  # the NaN vertices all land on voxel 0
    depth = s.depth(0)
    tsdf0 = np.full(n, -1, F)
    wt0 = np.zeros(n, F)
    t, wt, upd = oracle.integrate_warped(tsdf0, wt0, verts, col, cnt, depth, s.delta, s.dim, s.voxel_size, s.size,
                                         I3, c, s.intrinsics(), s.max_weight)
    gt, gwt, g_col, g_depth = dev(tsdf0), dev(wt0), dev(col), dev(depth)
    H, W = depth.shape
    ops.run_integrate_warped_tsdf_from_frame_kernel(gt, gwt, g_verts, g_col, g_cnt, g_depth, (W, H), s.delta,
                                                    s.voxel_size, s.size, s.dim, I3, c, s.intrinsics(), s.max_weight)
    np.testing.assert_array_equal(gt.download(), t)
    np.testing.assert_array_equal(gwt.download(), wt)
    assert upd > 100


def test_reference_sequence_chained(vf, oracle):
    """src/VolumeFusion.cu:708-741 through the drop-in entry points, GPU buffers
    only, equals the oracle frame (tsdf, weights, node round trip)."""
    k = 6
    s = _setup(k, nodes=70)
    c = np.array(s.center, F)
    n = s.n_voxels
    T = s.transforms(1)
    depth = s.depth(1)
    H, W = depth.shape
    t_or, w_or, nodes_or = np.full(n, -1, F), np.zeros(n, F), s.nodes.copy()
    oracle.warped_frame(k, t_or, w_or, nodes_or, s.radii, T, s.dim, s.voxel_size, s.size, s.cube, c, depth, s.delta,
                        s.intrinsics(), 15.0)

    tsdf = ops.DeviceTsdf(s.size, s.dim, center=c, delta=s.delta)
    g_nodes, g_T, g_r, g_depth = dev(s.nodes), dev(T), dev(s.radii), dev(depth)
    g_idx, g_dist, g_w = DeviceArray((n, k), np.uint32), DeviceArray((n, k), F), DeviceArray((n, k), F)
    g_cnt, g_col = DeviceArray((n,), np.int32), DeviceArray((n,), np.int32)
    g_verts = DeviceArray((n, 3), F)
    g_cnt.fill(0), g_col.fill(0)
    G = s.nodes.shape[0]
    vs = tsdf.voxel_size()
    ops.run_translate_kernel(g_nodes, G, -c)
    ops.run_tsdf_knn_graph(g_idx, g_dist, g_cnt, g_nodes, g_r, G, vs, s.size, s.dim, s.cube)
    ops.run_compute_tsdf_warp_weights_kernel(g_w, g_r, g_idx, g_dist, g_cnt, n)
    ops.run_generate_tsdf_vertices_kernel(g_verts, vs, s.size, s.dim)
    ops.run_translate_kernel(g_nodes, G, c)
    ops.run_translate_kernel(g_verts, n, c)
    ops.run_warp_tsdf_kernel(g_verts, g_T, g_idx, g_w, g_cnt, vs, s.size, s.dim)
    ops.run_count_collisions_per_voxel_kernel(g_col, g_verts, g_cnt, vs, s.size, s.dim, I3, c)
    ops.run_integrate_warped_tsdf_from_frame_kernel(tsdf.tsdf, tsdf.weights, g_verts, g_col, g_cnt, g_depth, (W, H),
                                                    s.delta, vs, s.size, s.dim, I3, c, s.intrinsics(), 15.0)
    np.testing.assert_array_equal(tsdf.tsdf.download(), t_or)
    np.testing.assert_array_equal(tsdf.weights.download(), w_or)
    np.testing.assert_array_equal(g_nodes.download(), nodes_or)


def test_plain_device_vs_cpu_tsdfvolume_path(vf, oracle):
    """The plain-method workload (sphere128: one static-camera frame into a
    128^3 volume) on the device kernel against the reference's CPU path,
    TSDFVolume::integrate (src/TSDFVolume.cpp:55-171, restated in the oracle),
    with the reference's own assertion for this pair
    (test/src/TSDFIntegrationTest.cu:76-79): every tsdf within 4 ulps
    (ASSERT_FLOAT_EQ), every weight equal."""
    s = scenes.sphere128()
    depth = s.depth(0)
    n = s.n_voxels
    c = np.array(s.center, F)
    pose = np.eye(4, dtype=F)
    pose[:3, 3] = c
    tc, wc, uc = oracle.tsdfvolume_integrate(np.full(n, -1, F), np.zeros(n, F), depth, s.delta, s.dim, s.voxel_size,
                                             s.size, np.ascontiguousarray(pose.T.reshape(16)), s.intrinsics())
    tsdf = ops.DeviceTsdf(s.size, s.dim, center=c, delta=s.delta)
    tsdf.integrate(dev(depth), s.intrinsics(), 15.0)
    td, wd = tsdf.tsdf.download(), tsdf.weights.download()
    np.testing.assert_array_equal(wd, wc)
    ulps = np.abs(td.view(np.int32).astype(np.int64) - tc.view(np.int32))
    assert ulps.max() <= 4
    assert uc > 10000 and int((wd > 0).sum()) == uc


def test_integrate_plain_and_camera_frame(vf, oracle):
    s = scenes.sphere128()
    depth = s.depth(0)
    H, W = depth.shape
    n = s.n_voxels
    c = np.array(s.center, F)
    t, w, upd = oracle.integrate_plain(np.full(n, -1, F), np.zeros(n, F), depth, s.delta, s.dim, s.voxel_size, s.size,
                                       I3, c, s.intrinsics(), 15.0)
    tsdf = ops.DeviceTsdf(s.size, s.dim, center=c, delta=s.delta)
    tsdf.integrate(dev(depth), s.intrinsics(), 15.0)
    np.testing.assert_array_equal(tsdf.tsdf.download(), t)
    np.testing.assert_array_equal(tsdf.weights.download(), w)
    assert upd > 10000
    # camera-frame variant with a small rotation
    th = 0.05
    Rw = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]], F)
    Rw_cm = np.ascontiguousarray(Rw.T.reshape(9))
    tw = np.array([0.01, -0.02, 0.03], F)
    t2, w2, _ = oracle.integrate_camera_frame(t, w, depth, s.delta, s.dim, s.voxel_size, s.size, Rw_cm, tw, I3, c,
                                              s.intrinsics(), 15.0)
    ops.run_integrate_tsdf_from_camera_frame_kernel(tsdf.tsdf, tsdf.weights, dev(depth), (W, H), s.delta, s.dim,
                                                    tsdf.voxel_size(), s.size, Rw_cm, tw, I3, c, s.intrinsics(), 15.0)
    np.testing.assert_array_equal(tsdf.tsdf.download(), t2)
    np.testing.assert_array_equal(tsdf.weights.download(), w2)
