#!/usr/bin/env python3
"""Warped TSDF integration benchmark on one MI355X.

A step is one frame of warped integration (the reference's per-frame block,
src/VolumeFusion.cu:702-742) of the workload's volume, on inputs already
resident in HBM: k_prep_bin, k_pass_a, k_pass_b on one stream.  Prints one
JSON line:

  value      warped voxel updates per second (voxels whose tsdf/weight changed)
  roofline   dominant kernel (k_pass_a): algorithmic bytes per launch / its
             average duration from HIP events recorded on its stream, against
             the 8 TB/s HBM peak.  Algorithmic bytes per frame (whole path):
             16 U + 4 W H + 80 G (tsdf+weight read and write of updated voxels,
             one depth read, node position+radius+transform); k_pass_a's share
             is 8 U + 4 W H + 80 G, k_pass_b's is 8 U.
  cpu_baseline  the C oracle's voxel-major frame on the host cores, on a bounded
             z-slab sample of the same workload.

Usage: python bench.py [--steps K] [--warmup W] [--workload config3]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

METRIC = "warped voxel updates/s + % HBM roofline; p50/p99 frame latency @32 streams"
HBM_PEAK = 8.0e12


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="config3", choices=["config1", "config2", "config3"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target length of the CPU sample")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a device copy of known size (HBM counter calibration under rocprofv3)")
    return ap.parse_args()


def workload(name: str):
    from volumefusion_amd import scenes

    return {"config1": scenes.config1, "config2": scenes.config2, "config3": scenes.config3}[name]()


def algorithmic_bytes(U: float, W: int, H: int, G: int) -> dict:
    return {"frame": 16.0 * U + 4.0 * W * H + 80.0 * G, "k_pass_a": 8.0 * U + 4.0 * W * H + 80.0 * G,
            "k_pass_b": 8.0 * U}


def cpu_baseline(s, depth, T, seconds: float) -> dict:
    """Time the oracle's voxel-major frame on a z-slab sample around the
    volume centre, sized to about `seconds` of work on the host cores."""
    from oracle import oracle as O

    threads = min(16, os.cpu_count() or 1)
    F = np.float32
    n = s.n_voxels
    Z = s.dim[2]

    def run(z0, z1, nthreads):
        t = np.full(n, -1, F)
        w = np.zeros(n, F)
        nodes = s.nodes.copy()
        t0 = time.perf_counter()
        st = O.warped_frame_gather(s.k, t, w, nodes, s.radii, T, s.dim, s.voxel_size, s.size, s.cube, s.center,
                                   depth, s.delta, s.intrinsics(), s.max_weight, nthreads=nthreads, z0=z0, z1=z1)
        return time.perf_counter() - t0, st

    # whole frames on all cores, repeated up to about `seconds`
    total_t, total_u, frames = 0.0, 0, 0
    while total_t < seconds and frames < 50:
        dt, st = run(0, Z, threads)
        total_t += dt
        total_u += st["updated"]
        frames += 1
    out = {"value": total_u / total_t, "unit": "warped voxel updates/s", "cores": threads, "kind": "port",
           "sample": f"{s.name}: {frames} whole frames on {threads} threads ({total_u} updates in {total_t:.2f} s; "
                     "oracle vfo_warped_frame_gather, voxel-major, z-slabs over threads)"}
    # one core on a z-slab through the centre sized to about seconds/3
    mid = Z // 2
    dt, st = run(mid - 4, mid + 4, 1)
    half = int(min(Z // 2, max(4, (seconds / 3) / max(dt / 8.0, 1e-4) / 2)))
    dt1, st1 = run(mid - half, mid + half, 1)
    out["single_core"] = {"value": st1["updated"] / dt1, "cores": 1,
                          "sample": f"one frame, source voxels z in [{mid - half}, {mid + half}) of {Z}, {dt1:.2f} s"}
    return out


def main() -> int:
    args = parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE={world}); the warped "
              "integration path has no multi-GPU mode in this build", file=sys.stderr)
        return 2

    from volumefusion_amd import _lib as L
    from volumefusion_amd import device, ops
    from volumefusion_amd.device import DeviceArray, Event, Stream

    L.load()
    if device.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1

    s = workload(args.workload)
    depths = [s.depth(f) for f in range(2)]
    Ts = [s.transforms(f) for f in range(2)]
    H, W = depths[0].shape
    G = s.nodes.shape[0]

    stream = Stream()
    tsdf = ops.DeviceTsdf(s.size, s.dim, center=s.center, delta=s.delta, stream=stream)
    graph = ops.DeviceMotionGraph(s.nodes, Ts[0], s.radii, stream=stream)
    T_dev = [DeviceArray.from_numpy(T, stream) for T in Ts]
    d_dev = [DeviceArray.from_numpy(d, stream) for d in depths]
    ws = ops.Workspace(stream=stream)
    vol = ops.make_volume(tsdf)
    frames = [ops.make_frame(d, s.intrinsics()) for d in d_dev]
    graphs = []
    for Td in T_dev:
        g = ops.make_graph(graph, write_back=True)
        g.transforms = Td.ptr
        graphs.append(g)
    params = ops.make_params(s.k, s.delta, s.max_weight, s.cube)
    lib = L.load()
    fn = lib.vf_warped_integrate

    def step(i):
        L.check(fn(ws.handle, C.byref(vol), C.byref(graphs[i % 2]), C.byref(frames[i % 2]), C.byref(params),
                   stream.ptr), "vf_warped_integrate")

    for i in range(args.warmup):
        step(i)
    stream.synchronize()
    ws.totals(reset=True, stream=stream)
    ws.set_timing(True)
    e0, e1 = Event(), Event()
    stream.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for i in range(args.steps):
        step(args.warmup + i)
    e1.record(stream)
    stream.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_ms(e1)
    kms, nk = ws.kernel_ms()
    ws.set_timing(False)
    tot = ws.totals(stream=stream)

    U = tot["updated"] / args.steps
    ab = algorithmic_bytes(U, W, H, G)
    t_a = kms[1] / nk * 1e-3
    t_b = kms[2] / nk * 1e-3
    t_frame = sum(kms) / nk * 1e-3
    achieved = ab["k_pass_a"] / t_a
    traffic = None
    tfile = REPO / "profiles" / f"traffic_{args.workload}.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get("k_pass_a", {}).get("hbm_bytes_per_launch")
        except (ValueError, AttributeError):
            traffic = None

    if args.copy_calibration:
        # a plain device-to-device copy of known size, for scaling HBM counters
        nbytes = 1 << 30
        a, b = DeviceArray((nbytes // 4,), np.float32, stream), DeviceArray((nbytes // 4,), np.float32, stream)
        a.fill(1.0)
        for _ in range(3):
            b.copy_from(a)
        stream.synchronize()

    result = {
        "metric": METRIC,
        "value": tot["updated"] / wall,
        "unit": "warped voxel updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{s.name}: {s.dim[0]}x{s.dim[1]}x{s.dim[2]} volume at {s.voxel_size[0] * 1e3:.1f} mm, "
                        f"{G} nodes, k={s.k}, {W}x{H} depth frames",
            "volume": list(s.dim), "nodes": G, "k": s.k, "frame": [W, H], "node_cube": list(s.cube),
            "updates_per_frame": U, "supported_per_frame": tot["supported"] / args.steps,
            "active_bricks_per_frame": tot["active_bricks"] / args.steps,
            "collided_per_frame": tot["collided"] / args.steps,
        },
        "roofline": {
            "bound": "hbm", "kernel": "k_pass_a", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
            "algorithmic_bytes_per_launch": ab["k_pass_a"], "avg_launch_ms": t_a * 1e3,
        },
        "frame_roofline": {
            "achieved_GBps": ab["frame"] / t_frame / 1e9, "frac": ab["frame"] / t_frame / HBM_PEAK,
            "algorithmic_bytes": ab["frame"], "kernel_ms": {"k_prep_bin": kms[0] / nk, "k_pass_a": kms[1] / nk,
                                                            "k_pass_b": kms[2] / nk},
            "gpu_ms_per_step": gpu_ms / args.steps,
        },
        "dense_equivalent_voxels_per_s": s.n_voxels * args.steps / wall,
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(s, depths[0], Ts[0], args.cpu_seconds)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
