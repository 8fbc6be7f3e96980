"""GPU parity of the fused five-kernel frame (and its batched form) against
the oracle's frame, bit for bit, over several frames."""
import numpy as np
import pytest

from volumefusion_amd import ops, scenes
from volumefusion_amd.device import DeviceArray

pytestmark = pytest.mark.gpu
F = np.float32


def dev(a):
    return DeviceArray.from_numpy(np.ascontiguousarray(a))


def run_frames(vf, oracle, s, frames, gather=False, threads=16, check_stats=True, tuning=None, blend=0,
               screen=True, depth_fn=None, transforms_fn=None, f64=False, budget=None, per_frame=None):
    """Frames through the fused call against the oracle, bit for bit after
    every frame.  With the brick screen off the supported and in-bounds
    counts equal the oracle's too; with it they cover the swept bricks only.
    With f64 the gather oracle also carries the binary64 companion state, and
    the device tsdf's max and RMS error against it (in units of the
    truncation distance: the stored tsdf is sdf / delta) is checked after
    every frame and reported.  With budget, the workspace gets that
    pending-record budget (records per voxel).  per_frame, a list, receives
    each frame's device statistics.  Returns the last frame's (oracle stats,
    device stats)."""
    n = s.n_voxels
    t_or, w_or, nodes_or = np.full(n, -1, F), np.zeros(n, F), s.nodes.copy()
    t64 = w64 = None
    if f64:
        gather = True
        t64, w64 = np.full(n, -1.0), np.zeros(n)
    errs = []
    tsdf = ops.DeviceTsdf(s.size, s.dim, center=s.center, delta=s.delta)
    graph = ops.DeviceMotionGraph(s.nodes, transforms_fn(frames[0]) if transforms_fn else s.transforms(frames[0]),
                                  s.radii)
    ws = ops.Workspace()
    tuning = tuple(tuning) if tuning is not None else (0, 0)
    ws.set_tuning(tuning[0], tuning[1], screen)
    if budget is not None:
        ws.set_record_budget(budget)
    gs = None
    for f in frames:
        depth = depth_fn(f) if depth_fn else s.depth(f)
        T = transforms_fn(f) if transforms_fn else s.transforms(f)
        if gather:
            st = oracle.warped_frame_gather(s.k, t_or, w_or, nodes_or, s.radii, T, s.dim, s.voxel_size, s.size,
                                            s.cube, s.center, depth, s.delta, s.intrinsics(), s.max_weight,
                                            nthreads=threads, blend=blend, tsdf64=t64, wt64=w64)
        else:
            st = oracle.warped_frame(s.k, t_or, w_or, nodes_or, s.radii, T, s.dim, s.voxel_size, s.size, s.cube,
                                     s.center, depth, s.delta, s.intrinsics(), s.max_weight, blend=blend).stats
        graph.transforms.upload(T)
        ops.warped_integrate(ws, tsdf, graph, dev(depth), s.intrinsics(), s.k, s.delta, s.max_weight, s.cube,
                             blend=blend)
        gs = ws.last_stats()
        if per_frame is not None:
            per_frame.append(gs)
        if check_stats:
            if not screen:
                assert gs["supported"] == st["supported"]
                assert gs["in_bounds"] == st["in_bounds"]
                assert gs["screened_bricks"] == 0
            assert gs["updated"] == st["updated"]
        np.testing.assert_array_equal(tsdf.weights.download(), w_or, err_msg=f"weights, frame {f}")
        np.testing.assert_array_equal(tsdf.tsdf.download(), t_or, err_msg=f"tsdf, frame {f}")
        np.testing.assert_array_equal(graph.points.download(), nodes_or, err_msg=f"nodes, frame {f}")
        if f64:
            td = tsdf.tsdf.download()
            np.testing.assert_array_equal(w64.astype(F), w_or, err_msg=f"binary64 weights, frame {f}")
            m = w_or > 0
            e = np.abs(td[m].astype(np.float64) - t64[m])
            emax, erms = float(e.max()), float(np.sqrt(np.mean(e * e)))
            errs.append((emax, erms))
            assert erms <= 1e-5, f"frame {f}: RMS |tsdf - tsdf64| = {erms:.3g} above 1e-5 of delta"
    if f64:
        from fp64_report import record

        record(s.name, len(frames), max(x[0] for x in errs), errs[-1][1], max(x[1] for x in errs),
               int((w_or > 0).sum()))
    return st, gs


@pytest.mark.parametrize("screen", [True, False])
@pytest.mark.parametrize("k", [1, 4, 6, 8])
def test_fused_small_multi_frame(vf, oracle, k, screen):
    s = scenes.small_scene(seed=11, k=k, nodes=70)
    st, gs = run_frames(vf, oracle, s, [0, 1, 2, 3], screen=screen)
    assert st["updated"] > 500 and st["collided"] > 0
    if screen:
        assert gs["screened_bricks"] > 0, "the far side of the sphere is screened"


def test_fused_spill_path(vf, oracle):
    """Dense node cluster: brick lists above 128 entries take the all-nodes
    path inside the kernel."""
    s = scenes.small_scene(seed=12, k=8, nodes=400, radius=0.03)
    run_frames(vf, oracle, s, [0, 1])
    run_frames(vf, oracle, s, [0, 1], screen=False)


@pytest.mark.parametrize("screen", [True, False])
@pytest.mark.parametrize("budget", [1e-7, 0.02, 1.0])
def test_fused_record_budget(vf, oracle, budget, screen):
    """A pending-record budget: with room for one record every brick with a
    pending voxel goes to the redo pass; with room for about 1,300 (a few
    records per workgroup arena, the rest in the overflow arena) some bricks
    are stored, some overflow reservations are left as null records by
    bricks that then go to the redo pass, and the rest are redone; a budget
    of one record per voxel holds everything.  Every frame's bits equal the
    oracle's, and every statistic the default workspace's."""
    s = scenes.small_scene(seed=11, k=8, nodes=70)
    frames = [0, 1, 2]
    ref, got = [], []
    run_frames(vf, oracle, s, frames, screen=screen, per_frame=ref)
    run_frames(vf, oracle, s, frames, screen=screen, budget=budget, per_frame=got)
    for f, (a, b) in enumerate(zip(ref, got)):
        assert a["redo_bricks"] == 0
        assert {k: v for k, v in a.items() if k != "redo_bricks"} == \
               {k: v for k, v in b.items() if k != "redo_bricks"}, f"frame {f}"
    redo = [g["redo_bricks"] for g in got]
    if budget < 1e-6:
        assert min(redo) > 10, redo
    elif budget < 0.5:
        assert min(redo) > 0 and max(redo) < max(g["active_bricks"] for g in got), redo
    else:
        assert max(redo) == 0, redo


def test_batch_record_budget_with_dqb_and_spill(vf, oracle):
    """A batch on a workspace with a small record budget: volumes of uneven
    size, LBS and DQB, one with a node cluster whose brick lists spill; the
    redo pass finds each brick's volume and list.  Bits equal the oracle's."""
    vols = [scenes.small_scene(seed=20 + i, k=8, nodes=30 + 25 * i, dim=(24 + 8 * i, 32, 20 + 12 * i))
            for i in range(3)]
    vols.append(scenes.small_scene(seed=12, k=8, nodes=400, radius=0.03))
    blends = [0, 1, 0, 1]
    tsdfs = [ops.DeviceTsdf(s.size, s.dim, center=s.center, delta=s.delta) for s in vols]
    graphs = [ops.DeviceMotionGraph(s.nodes, s.transforms(0), s.radii) for s in vols]
    depths = [dev(s.depth(0)) for s in vols]
    ws = ops.Workspace()
    ws.set_record_budget(1e-3)
    ops.warped_integrate_batch(ws, tsdfs, graphs, depths, [s.intrinsics() for s in vols], 8, 0.012, 64.0,
                               [s.cube for s in vols], blend=blends)
    st = ws.last_stats()
    assert st["redo_bricks"] > 0 and st["overflow_bricks"] > 0, st
    for s, t, g, b in zip(vols, tsdfs, graphs, blends):
        n = s.n_voxels
        t_or, w_or, nodes_or = np.full(n, -1, F), np.zeros(n, F), s.nodes.copy()
        oracle.warped_frame(s.k, t_or, w_or, nodes_or, s.radii, s.transforms(0), s.dim, s.voxel_size, s.size,
                            s.cube, s.center, s.depth(0), s.delta, s.intrinsics(), s.max_weight, blend=b)
        np.testing.assert_array_equal(t.tsdf.download(), t_or)
        np.testing.assert_array_equal(t.weights.download(), w_or)
        np.testing.assert_array_equal(g.points.download(), nodes_or)


def _special_depth(s, f, nan):
    """The frame with +-inf, zero, negative and too small pixels sprinkled
    over it (all fail the projective test), and NaN pixels (which pass it:
    their sdf compares false both ways): none, a block in the middle, or
    sprinkled everywhere as well."""
    d = s.depth(f).copy()
    rng = np.random.default_rng([s.seed, 77, f])
    H, W = d.shape
    specials = ((np.inf, 0.01), (-np.inf, 0.01), (0.0, 0.01), (-0.5, 0.01), (5e-5, 0.01))
    if nan == "everywhere":
        specials = ((np.nan, 0.02),) + specials
    for value, frac in specials:
        m = rng.random(d.shape) < frac
        d[m] = value
    if nan != "none":
        d[H // 2 - 6:H // 2 + 6, W // 2 - 6:W // 2 + 6] = np.nan
    return d


@pytest.mark.parametrize("nan", ["none", "block", "everywhere"])
def test_fused_screen_special_depth_pixels(vf, oracle, nan):
    """Brick screen with NaN, infinite, zero and negative depth pixels:
    bit-exact against the oracle over four frames with the screen on.  The
    depth-tile maxima must keep every brick a NaN pixel can make pending (a
    NaN block in the middle keeps the bricks behind the sphere's centre and
    whatever can share their targets); with NaN pixels all over the frame
    every brick's footprint holds one, so the screen keeps everything; with
    no NaN the screen drops bricks."""
    s = scenes.small_scene(seed=15, k=8, nodes=80)
    st, gs = run_frames(vf, oracle, s, [0, 1, 2, 3], depth_fn=lambda f: _special_depth(s, f, nan))
    assert st["updated"] > 500
    if nan == "everywhere":
        assert gs["screened_bricks"] == 0
    elif nan == "none":
        assert gs["screened_bricks"] > 0


@pytest.mark.parametrize("bad", ["nan_transform", "inf_transform", "zero_radius", "tiny_radius", "nan_node",
                                 "inf_node"])
def test_fused_screen_unbounded_node(vf, oracle, bad):
    """A node whose displacement the screen cannot bound (non-finite
    transform or position, a radius whose weights can be NaN): the screen
    keeps every brick of the volume, and the frame stays bit-exact.  A node
    at infinity has its cube outside the volume and binds nothing."""
    s = scenes.small_scene(seed=16, k=8, nodes=60)
    clean = scenes.small_scene(seed=16, k=8, nodes=60)  # warps are made from finite positions
    j = 7
    if bad == "zero_radius":
        s.radii = s.radii.copy()
        s.radii[j] = 0.0
    elif bad == "tiny_radius":
        s.radii = s.radii.copy()
        s.radii[j] = 1e-30
    elif bad in ("nan_node", "inf_node"):
        s.nodes = s.nodes.copy()
        s.nodes[j, 0] = np.nan if bad == "nan_node" else np.inf

    def transforms(f):
        T = clean.transforms(f).copy()
        if bad == "nan_transform":
            T[j, 12] = np.nan
        elif bad == "inf_transform":
            T[j, 0] = np.inf
        return T

    st, gs = run_frames(vf, oracle, s, [0, 1], transforms_fn=transforms)
    if bad == "inf_node":
        assert gs["screened_bricks"] > 0
    else:
        assert gs["screened_bricks"] == 0


@pytest.mark.parametrize("radius", [1e20, 1e19, 1e-11])
def test_fused_extreme_radius(vf, oracle, radius):
    """A node radius whose 2r^2 is infinite, huge or tiny: the brick's weights
    take IEEE divisions (exp(-d^2 / inf) = 1 where a precomputed reciprocal
    would give NaN), bit-exact against the oracle; the screen keeps the whole
    volume."""
    s = scenes.small_scene(seed=19, k=8, nodes=60)
    s.radii = s.radii.copy()
    s.radii[5] = radius
    st, gs = run_frames(vf, oracle, s, [0, 1])
    assert st["updated"] > 500 and gs["screened_bricks"] == 0


def test_fused_screen_volume_across_camera_plane(vf, oracle):
    """A volume reaching behind the camera (z from -0.03 to 0.17 m) around a
    small sphere in front of it: bricks near z = 0 have unbounded pixel
    footprints and are kept, bricks behind the camera are dropped."""
    import dataclasses

    rng = np.random.default_rng([17, 9])
    sph = scenes.make_surface(rng, (0.0, 0.0, 0.11), (0.045, 0.045, 0.045), eps=0.0, rot=np.eye(3))
    nodes = scenes.poisson_disk([sph], 60, rng)
    base = scenes.small_scene(seed=17, k=8, nodes=8)
    dim = (40, 40, 40)
    s = dataclasses.replace(base, dim=dim, size=(0.2, 0.2, 0.2), center=(0.0, 0.0, 0.07), nodes=nodes,
                            radii=np.full(60, 0.015, F), graph_radius=0.015, surfaces=[sph],
                            cam=scenes.Camera(160, 144, 126.0, 126.0, 80.0, 72.0))
    st, gs = run_frames(vf, oracle, s, [0, 1, 2])
    assert st["updated"] > 100


def test_fused_empty_graph_leaves_volume(vf, oracle):
    s = scenes.small_scene(seed=13, k=4, nodes=8)
    s.nodes = s.nodes[:0]
    s.radii = s.radii[:0]
    tsdf = ops.DeviceTsdf(s.size, s.dim, center=s.center, delta=s.delta)
    graph = ops.DeviceMotionGraph(np.zeros((0, 3), F), np.zeros((0, 16), F), np.zeros(0, F))
    ws = ops.Workspace()
    ops.warped_integrate(ws, tsdf, graph, dev(s.depth(0)), s.intrinsics(), 4, s.delta, 64.0, s.cube)
    assert (tsdf.tsdf.download() == -1).all() and (tsdf.weights.download() == 0).all()


def test_fused_sphere128(vf, oracle):
    """128^3, 256 nodes, identity transforms, k=4 (the CPU-runnable config):
    its one frame with the screen and the binary64 error report, then two
    frames without the screen (frame counters equal the oracle's)."""
    s = scenes.sphere128()
    st, gs = run_frames(vf, oracle, s, [0], f64=True)
    assert st["updated"] > 10000 and gs["screened_bricks"] > 0
    run_frames(vf, oracle, s, [0, 1], screen=False)


@pytest.mark.slow
def test_fused_sphere256_warped_frames(vf, oracle):
    """256^3, 2048 nodes with random warps, k=8, all 30 frames of the workload
    in the benchmark's configuration: every voxel and node position checked
    bit for bit after every frame against the gather oracle (itself equal to
    the node-major oracle), and the binary64 error after every frame."""
    s = scenes.sphere256_warped()
    run_frames(vf, oracle, s, list(range(30)), tuning=(2, 5), f64=True)


@pytest.mark.parametrize("tuning,screen", [((0, 0), False), ((2, 5), True)])
def test_fused_ellipsoids512_full_size(vf, oracle, tuning, screen):
    """The headline workload at its full size (512^3 at 2 mm, 4096 nodes, k=8,
    1024x1024 depth with noise and dropout), two frames, every voxel and node
    position bit for bit against the gather oracle on 16 host threads: the
    default 16x8x4 bricks without the brick screen (frame counters equal
    too), and the configuration the benchmark times (8x8x4 bricks, five
    workgroups per CU, screen on)."""
    s = scenes.ellipsoids512()
    st, gs = run_frames(vf, oracle, s, [0, 1], gather=True, threads=16, tuning=tuning, screen=screen, f64=screen)
    assert st["updated"] > 1_000_000 and st["supported"] > 8_000_000
    if screen:
        assert gs["screened_bricks"] > 10_000


def test_batched_equals_sequential(vf, oracle):
    """Volumes of uneven size and node count in one batched call give the same
    bits as one call per volume, and as the oracle."""
    vols = [scenes.small_scene(seed=20 + i, k=8, nodes=30 + 25 * i, dim=(24 + 8 * i, 32, 20 + 12 * i))
            for i in range(4)]
    tsdfs = [ops.DeviceTsdf(s.size, s.dim, center=s.center, delta=s.delta) for s in vols]
    graphs = [ops.DeviceMotionGraph(s.nodes, s.transforms(0), s.radii) for s in vols]
    depths = [dev(s.depth(0)) for s in vols]
    ws = ops.Workspace()
    ops.warped_integrate_batch(ws, tsdfs, graphs, depths, [s.intrinsics() for s in vols], 8, 0.012, 64.0,
                               [s.cube for s in vols])
    for s, t, g in zip(vols, tsdfs, graphs):
        n = s.n_voxels
        t_or, w_or, nodes_or = np.full(n, -1, F), np.zeros(n, F), s.nodes.copy()
        oracle.warped_frame(s.k, t_or, w_or, nodes_or, s.radii, s.transforms(0), s.dim, s.voxel_size, s.size,
                            s.cube, s.center, s.depth(0), s.delta, s.intrinsics(), s.max_weight)
        np.testing.assert_array_equal(t.tsdf.download(), t_or)
        np.testing.assert_array_equal(t.weights.download(), w_or)
        np.testing.assert_array_equal(g.points.download(), nodes_or)


def test_batch_of_many_volumes_some_without_nodes(vf, oracle):
    """Twenty small volumes in one call, every fourth with an empty motion
    graph: an empty graph owns no (node, brick) items, so its first item is
    the next volume's, and the kernels must still find each item's and each
    brick's volume.  Every volume gets the oracle's bits; the empty ones stay
    untouched."""
    vols = []
    for i in range(20):
        s = scenes.small_scene(seed=60 + i, k=8, nodes=20 + 3 * i, dim=(16 + 2 * (i % 5), 20, 18 + 4 * (i % 3)))
        if i % 4 == 1:
            s.nodes = s.nodes[:0]
            s.radii = s.radii[:0]
        vols.append(s)
    tsdfs = [ops.DeviceTsdf(s.size, s.dim, center=s.center, delta=s.delta) for s in vols]
    graphs = [ops.DeviceMotionGraph(s.nodes, s.transforms(0).reshape(-1, 16), s.radii) for s in vols]
    depths = [dev(s.depth(0)) for s in vols]
    ws = ops.Workspace(max_volumes=len(vols))
    ops.warped_integrate_batch(ws, tsdfs, graphs, depths, [s.intrinsics() for s in vols], 8, 0.012, 64.0,
                               [s.cube for s in vols])
    updated = 0
    for i, (s, t, g) in enumerate(zip(vols, tsdfs, graphs)):
        n = s.n_voxels
        t_or, w_or, nodes_or = np.full(n, -1, F), np.zeros(n, F), s.nodes.copy()
        if s.nodes.shape[0]:
            oracle.warped_frame(s.k, t_or, w_or, nodes_or, s.radii, s.transforms(0), s.dim, s.voxel_size, s.size,
                                s.cube, s.center, s.depth(0), s.delta, s.intrinsics(), s.max_weight)
        np.testing.assert_array_equal(t.tsdf.download(), t_or, err_msg=f"tsdf of volume {i}")
        np.testing.assert_array_equal(t.weights.download(), w_or, err_msg=f"weights of volume {i}")
        updated += int((w_or > 0).sum())
    assert updated > 1000


def test_batch_at_the_volume_limit(vf, oracle):
    """256 volumes, the most one call takes, every seventh with an empty motion
    graph: the per-volume tables in LDS are full and the searches over them
    take every step.  Each volume gets the oracle's bits; a 257th volume is
    rejected on the host with VF_ERR_INVALID_ARGUMENT before anything runs.
    The scenes share two node sets (dart throwing is slow) and differ in
    extent, node count, warps and noise."""
    import dataclasses

    from volumefusion_amd import _lib as L

    bases = [scenes.small_scene(seed=300 + b, k=8, nodes=14, dim=(12, 12, 12)) for b in range(2)]
    vols = []
    for i in range(256):
        b = bases[i % 2]
        dim = (9 + i % 5, 10 + i % 3, 8 + i % 4)
        g = 0 if i % 7 == 3 else 4 + i % 11
        vs = b.voxel_size[0]
        vols.append(dataclasses.replace(b, dim=dim, size=tuple(float(np.float32(d * vs)) for d in dim),
                                        nodes=b.nodes[:g].copy(), radii=b.radii[:g].copy(), seed=400 + i))
    tsdfs = [ops.DeviceTsdf(s.size, s.dim, center=s.center, delta=s.delta) for s in vols]
    graphs = [ops.DeviceMotionGraph(s.nodes, s.transforms(0).reshape(-1, 16), s.radii) for s in vols]
    depths = [dev(s.depth(0)) for s in vols]
    ws = ops.Workspace(max_volumes=257)
    with pytest.raises(L.VFError) as e:
        ops.warped_integrate_batch(ws, tsdfs + tsdfs[:1], graphs + graphs[:1], depths + depths[:1],
                                   [s.intrinsics() for s in vols + vols[:1]], 8, 0.012, 64.0,
                                   [s.cube for s in vols + vols[:1]])
    assert e.value.status == 1  # VF_ERR_INVALID_ARGUMENT
    ops.warped_integrate_batch(ws, tsdfs, graphs, depths, [s.intrinsics() for s in vols], 8, 0.012, 64.0,
                               [s.cube for s in vols])
    updated = 0
    for i, (s, t, g) in enumerate(zip(vols, tsdfs, graphs)):
        n = s.n_voxels
        t_or, w_or, nodes_or = np.full(n, -1, F), np.zeros(n, F), s.nodes.copy()
        if s.nodes.shape[0]:
            oracle.warped_frame(s.k, t_or, w_or, nodes_or, s.radii, s.transforms(0), s.dim, s.voxel_size, s.size,
                                s.cube, s.center, s.depth(0), s.delta, s.intrinsics(), s.max_weight)
        np.testing.assert_array_equal(t.tsdf.download(), t_or, err_msg=f"tsdf of volume {i}")
        np.testing.assert_array_equal(t.weights.download(), w_or, err_msg=f"weights of volume {i}")
        np.testing.assert_array_equal(g.points.download(), nodes_or, err_msg=f"nodes of volume {i}")
        updated += int((w_or > 0).sum())
    assert updated > 10000


def test_fused_deterministic_repeat(vf, oracle):
    """Same inputs twice (fresh volumes): identical bits (no race in the knn)."""
    s = scenes.small_scene(seed=14, k=8, nodes=90)
    out = []
    for _ in range(2):
        tsdf = ops.DeviceTsdf(s.size, s.dim, center=s.center, delta=s.delta)
        graph = ops.DeviceMotionGraph(s.nodes, s.transforms(0), s.radii)
        ops.warped_integrate(ops.Workspace(), tsdf, graph, dev(s.depth(0)), s.intrinsics(), 8, s.delta, 64.0, s.cube)
        out.append((tsdf.tsdf.download(), tsdf.weights.download()))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("k", [1, 4, 8])
def test_fused_dqb_multi_frame(vf, oracle, k):
    """Dual-quaternion blending: the device frame equals the oracle's DQB
    frame bit for bit."""
    s = scenes.small_scene(seed=21, k=k, nodes=70)
    st, _ = run_frames(vf, oracle, s, [0, 1, 2], blend=1)
    assert st["updated"] > 500


def test_fused_dqb_spill_path(vf, oracle):
    s = scenes.small_scene(seed=22, k=6, nodes=400, radius=0.03)
    run_frames(vf, oracle, s, [0, 1], blend=1)


def test_batch_mixes_blend_modes(vf, oracle):
    """Volumes with either blend mode in one batched call."""
    vols = [scenes.small_scene(seed=23 + i, k=8, nodes=40 + 10 * i, dim=(24 + 8 * i, 28, 24)) for i in range(4)]
    blends = [0, 1, 1, 0]
    tsdfs = [ops.DeviceTsdf(s.size, s.dim, center=s.center, delta=s.delta) for s in vols]
    graphs = [ops.DeviceMotionGraph(s.nodes, s.transforms(0), s.radii) for s in vols]
    ops.warped_integrate_batch(ops.Workspace(), tsdfs, graphs, [dev(s.depth(0)) for s in vols],
                               [s.intrinsics() for s in vols], 8, 0.012, 64.0, [s.cube for s in vols], blend=blends)
    for s, t, b in zip(vols, tsdfs, blends):
        t_or, w_or, nodes_or = np.full(s.n_voxels, -1, F), np.zeros(s.n_voxels, F), s.nodes.copy()
        oracle.warped_frame(s.k, t_or, w_or, nodes_or, s.radii, s.transforms(0), s.dim, s.voxel_size, s.size,
                            s.cube, s.center, s.depth(0), s.delta, s.intrinsics(), s.max_weight, blend=b)
        np.testing.assert_array_equal(t.tsdf.download(), t_or)
        np.testing.assert_array_equal(t.weights.download(), w_or)
