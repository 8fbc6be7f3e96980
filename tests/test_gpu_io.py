"""TSDF files from and into device volumes; server checkpoint and resume."""
import numpy as np
import pytest

from test_gpu_fused import dev
from volumefusion_amd import ops, scenes, serve, tsdf_io

pytestmark = pytest.mark.gpu
F = np.float32


@pytest.mark.parametrize("fmt", [tsdf_io.NATIVE, tsdf_io.LEGACY])
def test_device_round_trip(vf, tmp_path, fmt):
    s = scenes.small_scene(seed=70, k=4, nodes=60)
    tsdf = ops.DeviceTsdf(s.size, s.dim, center=s.center, delta=s.delta)
    graph = ops.DeviceMotionGraph(s.nodes, s.transforms(0), s.radii)
    ops.warped_integrate(ops.Workspace(), tsdf, graph, dev(s.depth(0)), s.intrinsics(), s.k, s.delta, 64.0, s.cube)
    p = tmp_path / "v.bin"
    tsdf.save(p, fmt, max_weight=64.0)
    info, t, w = tsdf_io.load_host(p)
    np.testing.assert_array_equal(t, tsdf.tsdf.download())
    np.testing.assert_array_equal(w, tsdf.weights.download())
    other = ops.DeviceTsdf(s.size, s.dim, center=s.center, delta=s.delta)
    info2 = other.load(p)
    assert info2["dim"] == tuple(s.dim)
    np.testing.assert_array_equal(other.tsdf.download(), t)
    np.testing.assert_array_equal(other.weights.download(), w)


def test_server_checkpoint_and_resume(vf, tmp_path):
    """Frames 0-1, checkpoint, a fresh server restores and takes frame 2: the
    same bits as frames 0-2 on one server (node positions carried over by the
    caller, as the file holds only the volume)."""
    s = scenes.small_scene(seed=71, k=8, nodes=50)
    desc, keep = serve.scene_desc(s)
    a = serve.Server(1 << 31, k=8)
    sa = a.open(desc)
    for f in range(3):
        a.submit(sa, s.depth(f), s.transforms(f))
    a.drain()
    t_all, w_all, _ = a.read(sa)
    a.close_server()

    b = serve.Server(1 << 31, k=8)
    sb = b.open(desc)
    for f in range(2):
        b.submit(sb, s.depth(f), s.transforms(f))
    b.drain()
    _, _, nodes_after = b.read(sb)
    b.save(sb, tmp_path / "ckpt.vft")
    b.close_server()

    desc2, keep2 = serve.stream_desc(s.dim, s.size, s.center, s.cam.width, s.cam.height, s.intrinsics(), s.delta,
                                     s.max_weight, s.cube, nodes_after, s.radii)
    c = serve.Server(1 << 31, k=8)
    sc = c.open(desc2)
    c.restore(sc, tmp_path / "ckpt.vft")
    c.submit(sc, s.depth(2), s.transforms(2))
    c.drain()
    t_res, w_res, _ = c.read(sc)
    c.close_server()
    np.testing.assert_array_equal(t_res, t_all)
    np.testing.assert_array_equal(w_res, w_all)
