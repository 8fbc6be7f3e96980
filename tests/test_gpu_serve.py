"""The serving loop: results are the same alone and in a mix, in frame
order, admission follows the device-memory budget."""
import numpy as np
import pytest

from volumefusion_amd import _lib as L
from volumefusion_amd import scenes, serve

pytestmark = pytest.mark.gpu
F = np.float32
GiB = 1 << 30


def _scenes():
    return [scenes.small_scene(seed=60 + i, k=8, nodes=40 + 15 * i, dim=(28 + 6 * i, 24 + 8 * i, 32 + 4 * i))
            for i in range(5)]


def _oracle_frames(oracle, s, frames):
    t, w, nodes = np.full(s.n_voxels, -1, F), np.zeros(s.n_voxels, F), s.nodes.copy()
    for f in frames:
        oracle.warped_frame(s.k, t, w, nodes, s.radii, s.transforms(f), s.dim, s.voxel_size, s.size, s.cube,
                            s.center, s.depth(f), s.delta, s.intrinsics(), s.max_weight)
    return t, w, nodes


def test_stream_same_alone_and_in_mix(vf, oracle):
    sc = _scenes()
    frames = [0, 1, 2]
    # alone
    srv = serve.Server(4 * GiB, k=8, max_batch=8, lanes=2)
    desc, keep = serve.scene_desc(sc[0])
    a = srv.open(desc)
    for f in frames:
        srv.submit(a, sc[0].depth(f), sc[0].transforms(f))
    srv.drain()
    alone = srv.read(a)
    srv.close_server()
    # in a mix, frames interleaved across five streams
    srv = serve.Server(4 * GiB, k=8, max_batch=3, lanes=2)
    ids, keeps = [], []
    for s in sc:
        d, kp = serve.scene_desc(s)
        ids.append(srv.open(d))
        keeps.append(kp)
    tickets = []
    for f in frames:
        for sid, s in zip(ids, sc):
            tickets.append(srv.submit(sid, s.depth(f), s.transforms(f)))
    results = [srv.wait(t) for t in tickets]
    assert all(r["status"] == 0 for r in results)
    assert max(r["batch_frames"] for r in results) <= 3
    mixed = [srv.read(sid) for sid in ids]
    st = srv.stats()
    assert st["frames_done"] == len(tickets) and st["streams_open"] == len(sc)
    # these scenes have pending updates on 7-16 % of their voxels, above the
    # lanes' budget of 1/24 per voxel: bricks went through the redo pass
    assert st["totals"]["redo_bricks"] > 0
    srv.close_server()
    for x, y in zip(alone, mixed[0]):
        np.testing.assert_array_equal(x, y)
    # every stream against the oracle
    for s, (t, w, p) in zip(sc, mixed):
        t_or, w_or, n_or = _oracle_frames(oracle, s, frames)
        np.testing.assert_array_equal(w, w_or)
        np.testing.assert_array_equal(t, t_or)
        np.testing.assert_array_equal(p, n_or)


def test_frames_of_a_stream_stay_in_order(vf, oracle):
    """Many frames of one stream queued at once integrate in submission order."""
    s = _scenes()[1]
    srv = serve.Server(2 * GiB, k=8, max_batch=4, lanes=3)
    d, keep = serve.scene_desc(s)
    sid = srv.open(d)
    tickets = [srv.submit(sid, s.depth(f), s.transforms(f)) for f in range(5)]
    res = [srv.wait(t) for t in tickets]
    done = [r["done_s"] for r in res]
    assert done == sorted(done)
    assert all(r["batch_frames"] == 1 for r in res)
    t, w, p = srv.read(sid)
    srv.close_server()
    t_or, w_or, n_or = _oracle_frames(oracle, s, range(5))
    np.testing.assert_array_equal(w, w_or)
    np.testing.assert_array_equal(t, t_or)


def test_admission_by_budget(vf):
    sc = _scenes()
    d0, k0 = serve.scene_desc(sc[0])
    d1, k1 = serve.scene_desc(sc[1])
    need0, need1 = serve.stream_bytes(d0, 2), serve.stream_bytes(d1, 2)
    srv = serve.Server(need0 + need1 // 2, k=8, max_batch=4, lanes=2)
    a = srv.open(d0)
    with pytest.raises(L.VFError) as e:
        srv.open(d1)
    assert e.value.status == 6  # VF_ERR_CAPACITY
    srv.close(a)
    b = srv.open(d1)
    assert srv.stats()["bytes_in_use"] == need1
    srv.submit(b, sc[1].depth(0), sc[1].transforms(0))
    srv.drain()
    srv.close_server()


def test_admission_charges_lane_workspaces_once(vf):
    """The charge is every open stream's buffers plus `lanes` workspaces sized
    for the max_batch largest workspace shares, not a share of every lane per
    stream: with max_batch 2 and four streams, the two largest shares count."""
    sc = _scenes()[:4]
    descs = [serve.scene_desc(s) for s in sc]
    ws = [serve.stream_bytes(d, 2) - serve.stream_bytes(d, 1) for d, _ in descs]
    own = [serve.stream_bytes(d, 1) - w for (d, _), w in zip(descs, ws)]
    lanes, max_batch = 3, 2
    srv = serve.Server(1 << 40, k=8, max_batch=max_batch, lanes=lanes)
    ids = [srv.open(d) for d, _ in descs]
    top = sorted(ws, reverse=True)[:max_batch]
    assert srv.stats()["bytes_in_use"] == sum(own) + lanes * sum(top)
    assert srv.stats()["bytes_in_use"] < sum(serve.stream_bytes(d, lanes) for d, _ in descs)
    big = int(np.argmax(ws))
    srv.close(ids[big])
    rest = [w for i, w in enumerate(ws) if i != big]
    assert srv.stats()["bytes_in_use"] == sum(own) - own[big] + lanes * sum(sorted(rest, reverse=True)[:max_batch])
    srv.close_server()


def test_borrowed_submit_equals_copied(vf):
    """Frames submitted from page-locked memory without a copy give the bits of
    copied submissions."""
    from volumefusion_amd.device import PinnedArray

    s = _scenes()[2]
    out = []
    for borrowed in (False, True):
        srv = serve.Server(2 * GiB, k=8, max_batch=4, lanes=2)
        d, keep = serve.scene_desc(s)
        sid = srv.open(d)
        pins = []
        for f in range(3):
            if borrowed:
                pins.append((PinnedArray(s.depth(f)), PinnedArray(s.transforms(f))))
                srv.submit_borrowed(sid, *pins[-1])
            else:
                srv.submit(sid, s.depth(f), s.transforms(f))
        srv.drain()
        out.append(srv.read(sid))
        srv.close_server()
        for a, b in pins:
            a.free()
            b.free()
    for x, y in zip(*out):
        np.testing.assert_array_equal(x, y)


def test_streams32_full_size_one_frame_each(vf, oracle):
    """The 32-stream serving workload at its full size (129^3..491^3,
    624..3721 nodes, k=8, 640x576 frames) through the serving loop with its
    benchmark settings (eight lanes of at most eight frames): every stream's
    volume and node positions bit for bit against the gather oracle on 16
    host threads."""
    vols = scenes.streams32()
    frames = [(s.depth(0), s.transforms(0)) for s in vols]
    srv = serve.Server(200 * GiB, k=8, max_batch=8, lanes=8)
    ids, keeps, tickets = [], [], []
    for s in vols:
        d, kp = serve.scene_desc(s)
        ids.append(srv.open(d))
        keeps.append(kp)
    for sid, (depth, T) in zip(ids, frames):
        tickets.append(srv.submit(sid, depth, T))
    results = [srv.wait(t) for t in tickets]
    assert all(r["status"] == 0 for r in results)
    for i, (sid, s, (depth, T)) in enumerate(zip(ids, vols, frames)):
        t, w, p = srv.read(sid)
        t_or, w_or, n_or = np.full(s.n_voxels, -1, F), np.zeros(s.n_voxels, F), s.nodes.copy()
        oracle.warped_frame_gather(s.k, t_or, w_or, n_or, s.radii, T, s.dim, s.voxel_size, s.size, s.cube,
                                   s.center, depth, s.delta, s.intrinsics(), s.max_weight, nthreads=16)
        np.testing.assert_array_equal(w, w_or, err_msg=f"weights of stream {i}")
        np.testing.assert_array_equal(t, t_or, err_msg=f"tsdf of stream {i}")
        np.testing.assert_array_equal(p, n_or, err_msg=f"nodes of stream {i}")
    srv.close_server()
