"""GPU depth ingest against the oracle, bit for bit: each reference step as a
drop-in entry point, and the fused three-launch ingest."""
import numpy as np
import pytest

from test_gpu_fused import dev
from volumefusion_amd import ops, scenes
from volumefusion_amd.device import DeviceArray

pytestmark = pytest.mark.gpu
F = np.float32


def frame(seed=0):
    s = scenes.sphere128()
    d = s.depth(seed).copy()  # 640x576, sphere with noise, zeros for no hit
    d[100:110, 200:215] = np.nan
    d[300, 300:310] = 2.5  # beyond the threshold
    return d, s.intrinsics()


def test_steps_one_by_one(vf, oracle):
    d, K = frame()
    H, W = d.shape
    # filter
    dd = dev(d)
    ops.run_filter_depth_kernel(dd, d.size, 1.3)
    f_or = oracle.filter_depth(d, 1.3)
    np.testing.assert_array_equal(dd.download(), f_or)
    # bilateral
    out = DeviceArray((H, W), F)
    ops.bilateral_filter_texture_opm_shared(out, dd, H, W, 5, 75.0, 75.0, 0.025)
    b_or = oracle.bilateral(f_or, 5, 75.0, 75.0, 0.025)
    np.testing.assert_array_equal(out.download(), b_or)
    # back-projection, normals, removal
    mesh = DeviceArray((H * W, 3), F)
    ops.run_depth_to_mesh_kernel(mesh, out, W, H, K)
    m_or = oracle.depth_to_mesh(b_or, K)
    np.testing.assert_array_equal(mesh.download(), m_or)
    normals = DeviceArray((H * W, 3), F)
    normals.fill(0.0)
    ops.run_depth_vertices_normals_kernel(normals, mesh, out, H, W, (1.3, 2.1))
    n_or = oracle.depth_normals(m_or, H, W, 1.3, 2.1)
    np.testing.assert_array_equal(normals.download(), n_or)
    ops.run_remove_depth_without_normals_kernel(normals, mesh, out, H, W)
    v_or, d_or = oracle.remove_depth_without_normals(n_or, m_or, b_or)
    np.testing.assert_array_equal(mesh.download(), v_or)
    np.testing.assert_array_equal(out.download(), d_or)
    c, cnt = ops.depth_centroid(ops.Workspace(), mesh)
    c_or, cnt_or = oracle.depth_centroid(v_or)
    assert cnt == cnt_or > 1000
    np.testing.assert_array_equal(c, c_or)


def test_mask(vf, oracle):
    d, _ = frame()
    mask = (np.arange(d.size) % 3 != 0).astype(np.uint8)
    dd = dev(d)
    ops.run_mask_depth_kernel(dd, dev(mask), d.size)
    np.testing.assert_array_equal(dd.download(), oracle.mask_depth(d, mask))


@pytest.mark.parametrize("bilateral", [(5, 75.0, 75.0, 0.025), (2, 30.0, 2.0, 0.01), None])
def test_fused_ingest_equals_sequence(vf, oracle, bilateral):
    d, K = frame(1)
    H, W = d.shape
    dd = dev(d)
    res = ops.ingest_depth(ops.Workspace(), dd, K, threshold=1.3, bilateral=bilateral)
    f_or = oracle.filter_depth(d, 1.3)
    b_or = oracle.bilateral(f_or, *bilateral) if bilateral else f_or
    m_or = oracle.depth_to_mesh(b_or, K)
    n_or = oracle.depth_normals(m_or, H, W, 1.3, 2.1)
    v_or, d_or = oracle.remove_depth_without_normals(n_or, m_or, b_or)
    c_or, cnt_or = oracle.depth_centroid(v_or)
    np.testing.assert_array_equal(dd.download(), d_or)
    np.testing.assert_array_equal(res["vertices"].download(), v_or)
    np.testing.assert_array_equal(res["normals"].download(), n_or)
    assert res["count"] == cnt_or
    np.testing.assert_array_equal(res["centroid"], c_or)
