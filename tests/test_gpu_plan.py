"""Brick shapes, frame plans, recorded graphs and the autotuner: every launch
configuration gives the oracle's bits."""
import numpy as np
import pytest

from test_gpu_fused import dev, run_frames
from volumefusion_amd import _lib as L
from volumefusion_amd import ops, scenes
from volumefusion_amd.device import DeviceArray, Stream

pytestmark = pytest.mark.gpu
F = np.float32


@pytest.mark.parametrize("shape", range(len(L.BRICK_SHAPES)))
def test_every_brick_shape_matches_oracle(vf, oracle, shape):
    print(f"brick shape {L.BRICK_SHAPES[shape]}", flush=True)
    s = scenes.small_scene(seed=31, k=8, nodes=80)
    st, _ = run_frames(vf, oracle, s, [0, 1, 2], tuning=(shape, 0))
    assert st["updated"] > 500


@pytest.mark.parametrize("shape", range(len(L.BRICK_SHAPES)))
def test_node_cubes_wider_than_bricks(vf, oracle, shape):
    """Node cubes 41-43 voxels wide (radius 0.1 m at about 5 mm voxels):
    wider than every brick shape and than the 40x36x44 volume along x and y,
    so each brick's range of a cube is clipped on both sides, and voxels have
    more than k = 8 candidates among the 12 nodes."""
    print(f"brick shape {L.BRICK_SHAPES[shape]}", flush=True)
    s = scenes.small_scene(seed=33, k=8, nodes=12, radius=0.1)
    assert min(s.cube) > 33 and s.cube[0] > s.dim[0] and s.cube[1] > s.dim[1]
    st, _ = run_frames(vf, oracle, s, [0, 1], tuning=(shape, 0))
    assert st["updated"] > 500


def test_one_workgroup_per_cu_and_spill(vf, oracle):
    """One resident workgroup per CU (long static schedules) and spilled brick
    lists, on the smallest brick."""
    s = scenes.small_scene(seed=32, k=6, nodes=400, radius=0.03)
    run_frames(vf, oracle, s, [0, 1], tuning=(2, 1))


class _Batch:
    """Uneven small volumes with their oracle state."""

    def __init__(self, oracle, count=4, seed=40):
        self.oracle = oracle
        self.scenes = [scenes.small_scene(seed=seed + i, k=8, nodes=30 + 20 * i, dim=(24 + 8 * i, 28 + 4 * i, 20 + 8 * i))
                       for i in range(count)]
        self.tsdfs = [ops.DeviceTsdf(s.size, s.dim, center=s.center, delta=s.delta) for s in self.scenes]
        self.graphs = [ops.DeviceMotionGraph(s.nodes, s.transforms(0), s.radii) for s in self.scenes]
        self.depths = [dev(s.depth(0)) for s in self.scenes]
        self.ref = [(np.full(s.n_voxels, -1, F), np.zeros(s.n_voxels, F), s.nodes.copy()) for s in self.scenes]

    def plan(self, **kw):
        return ops.Plan(self.tsdfs, self.graphs, self.depths, [s.intrinsics() for s in self.scenes], 8, 0.012, 64.0,
                        [s.cube for s in self.scenes], **kw)

    def load_frame(self, f):
        for s, g, d in zip(self.scenes, self.graphs, self.depths):
            g.transforms.upload(s.transforms(f))
            d.upload(s.depth(f))

    def oracle_frame(self, f):
        for s, (t, w, nodes) in zip(self.scenes, self.ref):
            self.oracle.warped_frame(s.k, t, w, nodes, s.radii, s.transforms(f), s.dim, s.voxel_size, s.size, s.cube,
                                     s.center, s.depth(f), s.delta, s.intrinsics(), s.max_weight)

    def check(self, what):
        for i, (tsdf, g, (t, w, nodes)) in enumerate(zip(self.tsdfs, self.graphs, self.ref)):
            np.testing.assert_array_equal(tsdf.weights.download(), w, err_msg=f"{what}: weights of volume {i}")
            np.testing.assert_array_equal(tsdf.tsdf.download(), t, err_msg=f"{what}: tsdf of volume {i}")
            np.testing.assert_array_equal(g.points.download(), nodes, err_msg=f"{what}: nodes of volume {i}")


@pytest.mark.parametrize("graph", [True, False])
def test_plan_replays_match_oracle(vf, oracle, graph):
    """A plan launched frame after frame, its inputs rewritten in place between
    launches and unrelated eager work in between; every replay is checked."""
    b = _Batch(oracle)
    stream = Stream()
    plan = b.plan(graph=graph, stream=stream)
    other = scenes.small_scene(seed=49, k=4, nodes=40)
    other_tsdf = ops.DeviceTsdf(other.size, other.dim, center=other.center, delta=other.delta)
    other_graph = ops.DeviceMotionGraph(other.nodes, other.transforms(0), other.radii)
    other_ws = ops.Workspace()
    for f in range(4):
        print(f"plan frame {f} graph={graph}", flush=True)
        b.load_frame(f)
        plan.launch(stream, sync=True)
        b.oracle_frame(f)
        b.check(f"replay {f}")
        last, _ = plan.stats()
        assert last["updated"] > 0
        # eager work between replays: a plain call on another volume and a fill
        ops.warped_integrate(other_ws, other_tsdf, other_graph, dev(other.depth(f)), other.intrinsics(), 4,
                             other.delta, 64.0, other.cube)
        scratch = DeviceArray((1 << 16,), F)
        scratch.fill(3.0)
    plan.close()


def test_plan_autotune_leaves_volumes_and_keeps_results(vf, oracle):
    b = _Batch(oracle, count=3, seed=50)
    stream = Stream()
    plan = b.plan(graph=True, stream=stream)
    b.load_frame(0)
    plan.launch(stream, sync=True)
    b.oracle_frame(0)
    b.check("before tuning")
    chosen = plan.autotune(frames=1, stream=stream)
    print("autotune chose", chosen, flush=True)
    assert 0 <= chosen["brick_shape"] < len(L.BRICK_SHAPES) and chosen["wg_per_cu"] >= 1
    assert chosen["ms_per_frame"] > 0
    b.check("after tuning")  # scratch copies only; nodes not written back
    assert plan.tuning()["brick_shape"] == chosen["brick_shape"]
    for f in (1, 2):
        b.load_frame(f)
        plan.launch(stream, sync=True)
        b.oracle_frame(f)
        b.check(f"tuned frame {f}")
    # a new plan of the same size class starts from the cached choice
    plan2 = b.plan(graph=False, stream=stream)
    assert plan2.tuning()["brick_shape"] == chosen["brick_shape"]
    assert plan2.tuning()["wg_per_cu"] == chosen["wg_per_cu"]
    plan.close()
    plan2.close()


def test_batch16_full_size_graph_replays(vf, oracle):
    """The 16-volume batching workload at its full size (64^3..346^3, 274..3044
    nodes, k=8, one 640x576 frame each) as one plan replayed as a HIP graph,
    two frames: every voxel and node position of every volume bit for bit
    against the gather oracle on 16 host threads."""
    vols = scenes.batch16()
    tsdfs = [ops.DeviceTsdf(s.size, s.dim, center=s.center, delta=s.delta) for s in vols]
    graphs = [ops.DeviceMotionGraph(s.nodes, s.transforms(0), s.radii) for s in vols]
    depths = [dev(s.depth(0)) for s in vols]
    ref = [(np.full(s.n_voxels, -1, F), np.zeros(s.n_voxels, F), s.nodes.copy()) for s in vols]
    stream = Stream()
    plan = ops.Plan(tsdfs, graphs, depths, [s.intrinsics() for s in vols], 8, [s.delta for s in vols],
                    [s.max_weight for s in vols], [s.cube for s in vols], graph=True, stream=stream)
    for f in range(2):
        print(f"batch16 frame {f}", flush=True)
        frames = [(s.depth(f), s.transforms(f)) for s in vols]
        for g, d, (depth, T) in zip(graphs, depths, frames):
            g.transforms.upload(T)
            d.upload(depth)
        plan.launch(stream, sync=True)
        for s, (t, w, nodes), (depth, T) in zip(vols, ref, frames):
            oracle.warped_frame_gather(s.k, t, w, nodes, s.radii, T, s.dim, s.voxel_size, s.size, s.cube, s.center,
                                       depth, s.delta, s.intrinsics(), s.max_weight, nthreads=16)
        for i, (tsdf, g, (t, w, nodes)) in enumerate(zip(tsdfs, graphs, ref)):
            np.testing.assert_array_equal(tsdf.weights.download(), w, err_msg=f"frame {f}: weights of volume {i}")
            np.testing.assert_array_equal(tsdf.tsdf.download(), t, err_msg=f"frame {f}: tsdf of volume {i}")
            np.testing.assert_array_equal(g.points.download(), nodes, err_msg=f"frame {f}: nodes of volume {i}")
        last, _ = plan.stats()
        assert last["updated"] > 1_000_000
    plan.close()
