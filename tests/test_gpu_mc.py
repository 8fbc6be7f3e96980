"""GPU marching cubes against the oracle, bit for bit in canonical order."""
import ctypes as C

import numpy as np
import pytest

from test_gpu_fused import dev
from volumefusion_amd import _lib as L
from volumefusion_amd import ops, scenes
from volumefusion_amd.device import DeviceArray

pytestmark = pytest.mark.gpu
F = np.float32


def sphere(n, vs, R, center=(0.0, 0.0, 0.0), trunc=0.1):
    size = n * vs
    c = np.arange(n, dtype=np.float64) * vs - size / 2
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt((x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2)
    tsdf = np.clip((r - R) / trunc, -1, 1).astype(F).ravel()
    wt = np.where(r < R + 2 * trunc, 1.0, 0.0).astype(F).ravel()
    return tsdf, wt, (n, n, n), (F(size / n),) * 3, (F(size),) * 3


@pytest.mark.parametrize("full", [False, True])
@pytest.mark.parametrize("n", [20, 70])
def test_sphere(vf, oracle, n, full):
    tsdf, wt, dims, vs, size = sphere(n, 0.8 / n, 0.27)
    v, f = ops.marching_cubes(ops.Workspace(), dev(tsdf), dev(wt), dims, vs, size, full=full)
    vo, fo = oracle.marching_cubes(tsdf, wt, dims, vs, size, full=full)
    assert v.shape[0] > 100
    np.testing.assert_array_equal(v, vo)
    if full:
        np.testing.assert_array_equal(f, fo)


@pytest.mark.parametrize("full", [False, True])
def test_special_values_and_near_iso_corners(vf, oracle, full):
    """NaN, +-inf and near-iso values (within 1e-7, and tiny next to the
    other corners) on a sphere: the count sweep's screen must send these
    cubes to the exact count, which the emit sweep must agree with."""
    tsdf, wt, dims, vs, size = sphere(48, 0.8 / 48, 0.27)
    rng = np.random.default_rng(7)
    idx = rng.choice(tsdf.size, 3000, replace=False)
    vals = np.array([np.nan, np.inf, -np.inf, 5e-8, -5e-8, 1e-6, -3e-7, 0.0], F)
    tsdf[idx] = vals[rng.integers(0, vals.size, idx.size)]
    v, f = ops.marching_cubes(ops.Workspace(), dev(tsdf), dev(wt), dims, vs, size, full=full)
    vo, fo = oracle.marching_cubes(tsdf, wt, dims, vs, size, full=full)
    assert v.shape[0] > 100
    np.testing.assert_array_equal(v, vo)
    if full:
        np.testing.assert_array_equal(f, fo)


@pytest.mark.parametrize("iso", [0.0, 0.5])
def test_degenerate_triangles(vf, oracle, iso):
    """TSDF values on a coarse grid with the iso value among them: edge points
    snap to cube corners, so triangles with two equal points occur and are
    dropped, exactly as the oracle drops them."""
    rng = np.random.default_rng(21)
    n = 24
    tsdf = rng.choice(np.array([-1.0, -0.5, 0.0, 0.5, 1.0], F), n ** 3).astype(F)
    wt = np.ones(n ** 3, F)
    dims, vs, size = (n, n, n), (F(0.01),) * 3, (F(0.24),) * 3
    v, f = ops.marching_cubes(ops.Workspace(), dev(tsdf), dev(wt), dims, vs, size, iso=iso, full=True)
    vo, fo = oracle.marching_cubes(tsdf, wt, dims, vs, size, iso=iso, full=True)
    np.testing.assert_array_equal(v, vo)
    np.testing.assert_array_equal(f, fo)
    # the cases' triangles outnumber the kept ones: degenerate ones were dropped
    from test_oracle_mc import tri_table
    ntri = np.array([len(t) for t in tri_table()])
    g = tsdf.reshape(n, n, n)  # [z, y, x]
    corner = [(1, 1, 0), (1, 0, 0), (0, 0, 0), (0, 1, 0), (1, 1, 1), (1, 0, 1), (0, 0, 1), (0, 1, 1)]
    case = np.zeros((n - 2, n - 2, n - 2), np.int64)
    for b, (dx, dy, dz) in enumerate(corner):
        case |= (g[dz:dz + n - 2, dy:dy + n - 2, dx:dx + n - 2] > F(iso)).astype(np.int64) << b
    assert fo.shape[0] > 1000 and fo.shape[0] < ntri[case].sum()


def test_coinciding_corner_positions(vf, oracle):
    """A volume size far larger than dims x voxel size puts every corner at
    the same binary32 position: the exact point comparison decides."""
    tsdf, wt, dims, vs, size = sphere(20, 0.04, 0.27)
    size = (F(1e9),) * 3
    v, f = ops.marching_cubes(ops.Workspace(), dev(tsdf), dev(wt), dims, vs, size, full=True)
    vo, fo = oracle.marching_cubes(tsdf, wt, dims, vs, size, full=True)
    np.testing.assert_array_equal(v, vo)
    np.testing.assert_array_equal(f, fo)
    vm, _ = ops.marching_cubes(ops.Workspace(), dev(tsdf), dev(wt), dims, vs, size, full=False)
    vmo, _ = oracle.marching_cubes(tsdf, wt, dims, vs, size, full=False)
    np.testing.assert_array_equal(vm, vmo)
    assert vmo.shape[0] > 100


@pytest.mark.parametrize("full", [False, True])
def test_integrated_volume(vf, oracle, full):
    """The TSDF the fused integration leaves after two frames (weight-0 voxels
    outside the observed band), at a non-zero iso value."""
    s = scenes.small_scene(seed=30, k=4, nodes=70)
    tsdf = ops.DeviceTsdf(s.size, s.dim, center=s.center, delta=s.delta)
    graph = ops.DeviceMotionGraph(s.nodes, s.transforms(0), s.radii)
    ws = ops.Workspace()
    for f in range(2):
        graph.transforms.upload(s.transforms(f))
        ops.warped_integrate(ws, tsdf, graph, dev(s.depth(f)), s.intrinsics(), s.k, s.delta, s.max_weight, s.cube)
    t, w = tsdf.tsdf.download(), tsdf.weights.download()
    assert (w == 0).any() and (w > 0).sum() > 1000
    v, fc = ops.marching_cubes(ws, tsdf.tsdf, tsdf.weights, s.dim, s.voxel_size, s.size, iso=0.1, full=full)
    vo, fo = oracle.marching_cubes(t, w, s.dim, s.voxel_size, s.size, iso=0.1, full=full)
    assert v.shape[0] > 50
    np.testing.assert_array_equal(v, vo)
    if full:
        np.testing.assert_array_equal(fc, fo)


def test_append_after_count_and_capacity(vf, oracle):
    """Vertices go after the caller's count (the reference's counter
    semantics); the counters advance by the full count even past the
    capacity, and nothing is written there."""
    tsdf, wt, dims, vs, size = sphere(24, 0.8 / 24, 0.27)
    vo, fo = oracle.marching_cubes(tsdf, wt, dims, vs, size, full=True)
    cap = 5 + vo.shape[0] // 2
    mesh = DeviceArray((cap, 3), F)
    mesh.fill(7.0)
    faces = DeviceArray((cap, 3), np.int32)
    counts = DeviceArray((2,), np.uint32)
    counts.upload(np.array([5, 2], np.uint32))
    ws, t_d, w_d = ops.Workspace(), dev(tsdf), dev(wt)  # alive across the call
    L.call("vf_marching_cubes_full", ws.handle, mesh.ptr, counts.ptr, cap, faces.ptr, counts.ptr + 4, cap, t_d.ptr,
           w_d.ptr, L.i3(dims), L.f3(vs), L.f3(size), 0.0, None)
    L.call("vf_device_synchronize")
    nv, nf = counts.download()
    assert nv == 5 + vo.shape[0] and nf == 2 + fo.shape[0]
    m = mesh.download()
    assert (m[:5] == 7.0).all()
    np.testing.assert_array_equal(m[5:], vo[:cap - 5])
    fc = faces.download()
    k = min(cap - 2, fo.shape[0])
    np.testing.assert_array_equal(fc[2:2 + k], fo[:k] + 5)


def test_default_workspace_entry(vf, oracle):
    """A NULL workspace (the af:: shims' path) uses the device default."""
    tsdf, wt, dims, vs, size = sphere(16, 0.8 / 16, 0.27)
    vo, _ = oracle.marching_cubes(tsdf, wt, dims, vs, size)
    mesh = DeviceArray((vo.shape[0] + 10, 3), F)
    counts = DeviceArray((1,), np.uint32)
    counts.upload(np.zeros(1, np.uint32))
    t_d, w_d = dev(tsdf), dev(wt)
    L.call("vf_marching_cubes_mesh", None, mesh.ptr, counts.ptr, mesh.shape[0], t_d.ptr, w_d.ptr, L.i3(dims),
           L.f3(vs), L.f3(size), 0.0, None)
    L.call("vf_device_synchronize")
    assert counts.download()[0] == vo.shape[0]
    np.testing.assert_array_equal(mesh.download()[:vo.shape[0]], vo)
