#!/usr/bin/env python3
"""Warped TSDF integration benchmark on one MI355X.

A step is one frame of warped integration (the reference's per-frame block,
src/VolumeFusion.cu:702-742) of the workload's volume, on inputs already
resident in HBM: k_prep_bin, k_compact, k_pass_a, k_pass_b on one stream.
Prints one JSON line:

  value      warped voxel updates per second (voxels whose tsdf/weight changed)
  roofline   dominant kernel (k_pass_a): algorithmic bytes per launch / its
             average duration from HIP events carried by its launches in the
             timed frames (on its stream), against the 8 TB/s HBM peak.  Algorithmic bytes per frame (whole path):
             16 U + 4 W H + 80 G (tsdf+weight read and write of updated voxels,
             one depth read, node position+radius+transform); k_pass_a's share
             is 8 U + 4 W H + 80 G, k_pass_b's is 8 U.
  cpu_baseline  the C oracle's voxel-major frame on the host cores, on a bounded
             z-slab sample of the same workload.

  batching   batch16: 16 volumes of uneven size, one frame each: 16 plain
             per-volume calls per step against one autotuned frame plan
             replayed as a HIP graph (the same work, results identical).
  extras     on the ellipsoids512 state: marching cubes of the integrated volume,
             the fused depth ingest of a frame, the fused point warp.
  serving    streams32: 32 resident streams, Poisson frame arrivals at 30 fps
             each, continuous batching; p50/p99 frame latency (submission to
             completion) and aggregate updates/s; then the same streams
             saturated against plain per-volume calls.

Usage: python bench.py [--steps K] [--warmup W] [--workload ellipsoids512]
                       [--no-cpu] [--no-batching] [--no-serving]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

METRIC = "warped voxel updates/s + % HBM roofline; p50/p99 frame latency @32 streams"
HBM_PEAK = 8.0e12
FP32_PEAK = 157.3e12  # dense fp32 matrix peak of MI355X
SIMDS = 1024          # MI355X: 256 CUs x 4 SIMDs
CLOCK_HZ = 2.4e9      # MI355X peak engine clock
VALU_CYCLES = 4       # a wave64 VALU instruction on a 16-lane SIMD


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="ellipsoids512", choices=["sphere128", "sphere256_warped", "ellipsoids512"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target length of the CPU sample")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a device copy of known size (HBM counter calibration under rocprofv3)")
    ap.add_argument("--no-batching", action="store_true", help="skip the batch16 batching measurement")
    ap.add_argument("--no-serving", action="store_true", help="skip the streams32 serving measurement")
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="length of the Poisson arrival run")
    ap.add_argument("--no-tune", action="store_true", help="default brick shape instead of the autotuned one")
    ap.add_argument("--tuning", default=None, metavar="SHAPE,WG",
                    help="fixed brick shape index and workgroups per CU instead of autotuning (e.g. 2,0)")
    ap.add_argument("--no-extras", action="store_true", help="skip marching cubes, ingest and point warp")
    ap.add_argument("--no-gn", action="store_true", help="skip the Gauss-Newton (gn4096) measurement")
    ap.add_argument("--no-fp64", action="store_true",
                    help="skip the binary64 error report of sphere128, sphere256_warped and ellipsoids512")
    return ap.parse_args()


def workload(name: str):
    from volumefusion_amd import scenes

    return {"sphere128": scenes.sphere128, "sphere256_warped": scenes.sphere256_warped,
            "ellipsoids512": scenes.ellipsoids512}[name]()


# The sources the fused frame's kernels are built from: counter profiles are
# valid only for the sources they were taken on.
KERNEL_SOURCES = ("volumefusion_amd/csrc/vf_fused.hip", "volumefusion_amd/csrc/vf_device.hpp",
                  "volumefusion_amd/csrc/vf_host.hpp")


def kernel_sources_sha256() -> str:
    import hashlib

    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        h.update(rel.encode())
        h.update((REPO / rel).read_bytes())
    return h.hexdigest()


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
    out["host_cpus"] = os.cpu_count()
    out["threads_note"] = ("threads = min(16, os.cpu_count()): one GPU's share of the host is 16 cores on the "
                           "MI355X boxes this runs on")
    out["plain_sphere128"] = cpu_plain_sphere128(seconds / 6)
    return out


def cpu_plain_sphere128(seconds: float) -> dict:
    """The plain method on the CPU path (TSDFVolume::integrate,
    src/TSDFVolume.cpp:55-171, restated in the oracle, single-threaded like the
    reference's): one static-camera frame of sphere128 into a fresh 128^3
    volume, repeated for about `seconds`."""
    from oracle import oracle as O
    from volumefusion_amd import scenes

    s = scenes.sphere128()
    depth = s.depth(0)
    n = s.n_voxels
    pose = np.eye(4, dtype=np.float32)
    pose[:3, 3] = np.asarray(s.center, np.float32)
    pose16 = np.ascontiguousarray(pose.T.reshape(16))
    vol = O.make_vol(s.dim, s.voxel_size, s.size)
    H, W = depth.shape
    K = np.ascontiguousarray(s.intrinsics(), np.float32)
    fn = O.lib().vfo_tsdfvolume_integrate
    total, frames, upd = 0.0, 0, 0
    while total < seconds or frames < 2:
        t, w = np.full(n, -1, np.float32), np.zeros(n, np.float32)
        t0 = time.perf_counter()
        u = fn(t, w, depth, W, H, float(s.delta), O.C.byref(vol), pose16, K)
        total += time.perf_counter() - t0
        frames += 1
        upd += int(u)
    return {"value": upd / total, "unit": "voxel updates/s (plain method)", "cores": 1, "kind": "port",
            "voxels_per_s": n * frames / total, "ms_per_frame": total / frames * 1e3,
            "sample": f"sphere128: {frames} frames of TSDFVolume::integrate on one core, {upd // frames} updates "
                      "per frame"}


def fp64_error(tuning, threads: int) -> dict:
    """Max and RMS |tsdf - tsdf64| of the device frames against the oracle's
    binary64 companion (same discrete decisions, continuous values in
    binary64), in units of the truncation distance (the stored tsdf is
    sdf / delta), over the voxels with weight > 0: sphere128 (its one frame),
    sphere256_warped (all 30 frames, checked after every frame),
    ellipsoids512 (two frames).  The device frames are also checked bit for
    bit against the float oracle.  Outside the timed region."""
    from oracle import oracle as O
    from volumefusion_amd import ops, scenes
    from volumefusion_amd.device import DeviceArray

    F = np.float32
    out = {}
    for name, fn, frames in (("sphere128", scenes.sphere128, 1), ("sphere256_warped", scenes.sphere256_warped, 30),
                             ("ellipsoids512", scenes.ellipsoids512, 2)):
        s = fn()
        n = s.n_voxels
        t_or, w_or, nodes_or = np.full(n, -1, F), np.zeros(n, F), s.nodes.copy()
        t64, w64 = np.full(n, -1.0), np.zeros(n)
        tsdf = ops.DeviceTsdf(s.size, s.dim, center=s.center, delta=s.delta)
        graph = ops.DeviceMotionGraph(s.nodes, s.transforms(0), s.radii)
        ws = ops.Workspace()
        ws.set_tuning(*tuning)
        per_frame = []
        exact = True
        t0 = time.perf_counter()
        for f in range(frames):
            depth, T = s.depth(f), s.transforms(f)
            O.warped_frame_gather(s.k, t_or, w_or, nodes_or, s.radii, T, s.dim, s.voxel_size, s.size, s.cube,
                                  s.center, depth, s.delta, s.intrinsics(), s.max_weight, nthreads=threads,
                                  tsdf64=t64, wt64=w64)
            graph.transforms.upload(T)
            ops.warped_integrate(ws, tsdf, graph, DeviceArray.from_numpy(depth), s.intrinsics(), s.k, s.delta,
                                 s.max_weight, s.cube)
            td, wd = tsdf.tsdf.download(), tsdf.weights.download()
            exact = exact and np.array_equal(td, t_or) and np.array_equal(wd, w_or)
            m = w_or > 0
            e = np.abs(td[m].astype(np.float64) - t64[m])
            per_frame.append((float(e.max()) if e.size else 0.0, float(np.sqrt(np.mean(e * e))) if e.size else 0.0))
        out[name] = {"frames": frames, "voxels_weighted": int((w_or > 0).sum()),
                     "max": max(x[0] for x in per_frame), "rms_last_frame": per_frame[-1][1],
                     "rms_max_over_frames": max(x[1] for x in per_frame),
                     "bit_exact_vs_float_oracle": bool(exact), "seconds": time.perf_counter() - t0}
        del tsdf, graph, ws
    out["unit"] = "truncation distance (tsdf = sdf / delta)"
    out["tolerance_rms"] = 1e-5
    return out


def _pct(values, q):
    v = sorted(values)
    if not v:
        return None
    i = min(len(v) - 1, max(0, int(math.ceil(q / 100.0 * len(v))) - 1))
    return v[i]


def bench_extras(s, tsdf, graph, depth_np, stream) -> dict:
    """The rows beside the path on the headline workload's state: marching
    cubes over the integrated 512^3 volume, the fused depth ingest of a frame,
    and the fused point warp of the marching-cubes mesh through the motion
    graph."""
    from volumefusion_amd import _lib as L
    from volumefusion_amd import ops
    from volumefusion_amd.device import DeviceArray, Event

    out = {}
    ws = ops.Workspace(stream=stream)
    e0, e1 = Event(), Event()
    reps = 3

    # marching cubes (full: vertices and faces)
    X, Y, Z = s.dim
    cap = 64 << 20
    mesh = DeviceArray((cap, 3), np.float32, stream)
    faces = DeviceArray((cap // 3, 3), np.int32, stream)
    counts = DeviceArray((2,), np.uint32, stream)

    def mc():
        counts.upload(np.zeros(2, np.uint32))
        L.call("vf_marching_cubes_full", ws.handle, mesh.ptr, counts.ptr, cap, faces.ptr, counts.ptr + 4, cap // 3,
               tsdf.tsdf.ptr, tsdf.weights.ptr, L.i3(s.dim), L.f3(s.voxel_size), L.f3(s.size), 0.0, stream.ptr)

    mc()
    stream.synchronize()
    ms = 0.0
    for _ in range(reps):
        counts.upload(np.zeros(2, np.uint32))
        e0.record(stream)
        L.call("vf_marching_cubes_full", ws.handle, mesh.ptr, counts.ptr, cap, faces.ptr, counts.ptr + 4, cap // 3,
               tsdf.tsdf.ptr, tsdf.weights.ptr, L.i3(s.dim), L.f3(s.voxel_size), L.f3(s.size), 0.0, stream.ptr)
        e1.record(stream)
        stream.synchronize()
        ms += e0.elapsed_ms(e1)
    ms /= reps
    nv, nf = (int(x) for x in counts.download())
    alg = 8.0 * X * Y * Z + 12.0 * nv + 12.0 * nf
    out["marching_cubes"] = {"vertices": nv, "faces": nf, "ms": ms, "algorithmic_bytes": alg,
                             "achieved_GBps": alg / (ms * 1e-3) / 1e9, "frac_hbm": alg / (ms * 1e-3) / HBM_PEAK}
    del faces

    # depth ingest (filter, bilateral r=5, back-projection, normals, removal, centroid)
    H, W = depth_np.shape
    d = DeviceArray((H, W), np.float32, stream)
    verts = DeviceArray((H * W, 3), np.float32, stream)
    normals = DeviceArray((H * W, 3), np.float32, stream)
    cen, cnt = L.Float3(), C.c_uint64(0)
    K = L.m3(s.intrinsics())
    ms = 0.0
    for r in range(reps + 1):
        d.upload(depth_np)
        t0 = time.perf_counter()
        L.call("vf_ingest_depth", ws.handle, d.ptr, W, H, K, 2.0, 5, 75.0, 75.0, 0.025, 1.3, 2.1, verts.ptr,
               normals.ptr, C.byref(cen), C.byref(cnt), stream.ptr)
        if r:
            ms += (time.perf_counter() - t0) * 1e3
    ms /= reps
    alg = 32.0 * H * W
    out["depth_ingest"] = {"pixels": H * W, "points": int(cnt.value), "ms_wall_incl_centroid_readback": ms,
                           "algorithmic_bytes": alg, "achieved_GBps": alg / (ms * 1e-3) / 1e9}

    # point warp: the marching-cubes mesh through the 4096-node graph, k=8 (the
    # reference warps its extracted mesh, src/VolumeFusion.cu:410-423)
    pts = mesh
    npts = nv
    L.call("vf_translate", pts.ptr, npts, L.f3(s.center), stream.ptr)  # volume frame -> world, as the nodes
    warped = DeviceArray((npts, 3), np.float32, stream)
    L.call("vf_warp_points", ws.handle, s.k, 0, warped.ptr, pts.ptr, npts, graph.points.ptr, graph.radii.ptr,
           graph.transforms.ptr, graph.count, stream.ptr)
    stream.synchronize()
    e0.record(stream)
    for _ in range(reps):
        L.call("vf_warp_points", ws.handle, s.k, 0, warped.ptr, pts.ptr, npts, graph.points.ptr, graph.radii.ptr,
               graph.transforms.ptr, graph.count, stream.ptr)
    e1.record(stream)
    stream.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    out["point_warp"] = {"points": npts, "source": "marching-cubes vertices", "nodes": int(graph.count), "k": s.k,
                         "ms": ms,
                         "points_per_s": npts / (ms * 1e-3),
                         "node_tests_per_s": npts * graph.count / (ms * 1e-3)}
    return out


def bench_plain_sphere128(stream) -> dict:
    """The plain method on the device (the reference's
    integrateTSDFFromFrameKernel, src/TSDFIntegration.cu:69-96): one
    static-camera frame of sphere128 into a 128^3 volume, timed with HIP
    events over repeated frames (the volume keeps integrating the same
    frame)."""
    from volumefusion_amd import ops, scenes
    from volumefusion_amd.device import DeviceArray, Event

    s = scenes.sphere128()
    tsdf = ops.DeviceTsdf(s.size, s.dim, center=s.center, delta=s.delta, stream=stream)
    d = DeviceArray.from_numpy(s.depth(0), stream)
    K = s.intrinsics()
    for _ in range(3):
        tsdf.integrate(d, K, 15.0, stream=stream)
    reps = 50
    e0, e1 = Event(), Event()
    e0.record(stream)
    for _ in range(reps):
        tsdf.integrate(d, K, 15.0, stream=stream)
    e1.record(stream)
    stream.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    n = s.n_voxels
    return {"workload": "sphere128: 128^3 at 4 mm, one 640x576 frame, plain method", "ms_per_frame": ms,
            "voxels_per_s": n / (ms * 1e-3), "hbm_bytes_per_frame": 16.0 * n,
            "achieved_GBps": 16.0 * n / (ms * 1e-3) / 1e9}


def bench_gauss_newton(stream, G: int = 4096, M: int = 300_000, iters: int = 5) -> dict:
    """One ICP iteration's Gauss-Newton solve (src/VolumeFusion.cu:434-692) on
    the gn4096 workload: 4096 nodes, 300k mesh points with 6 knn slots, 4
    regularisation neighbours, targets made by a smooth rigid motion of every
    node.  vf_gn_solver_create builds the normal equations from the sparse
    rows and factorises them (a blocked Cholesky factorisation: one-workgroup
    diagonal blocks, rocBLAS trsm/syrk/gemm updates, look-ahead on a second
    stream; rocSOLVER getrf where a pivot is not positive); vf_gn_solver_iterate then runs `iters` Gauss-Newton
    iterations (residuals, energy, JTr, the two triangular solves, update,
    mesh warp).  The reference rebuilds and refactorises JTJ on every
    iteration, i.e. one create per iteration."""
    from volumefusion_amd import _lib as L
    from volumefusion_amd import scenes
    from volumefusion_amd.device import DeviceArray

    p = scenes.gn_problem(7, G=G, M=M)
    K, R = p.knn.shape[1], p.gknn.shape[1]
    dev = lambda a: DeviceArray.from_numpy(np.ascontiguousarray(a), stream)
    corr, canon, knn, kw = dev(p.corr), dev(p.canon), dev(p.knn), dev(p.kw)
    dmesh, normals, graph, gknn = dev(p.dmesh), dev(p.dnormals), dev(p.graph), dev(p.gknn)
    desc = L.GNProblem(corr.ptr, p.corr.shape[0], canon.ptr, K, knn.ptr, kw.ptr, normals.ptr, graph.ptr, G,
                       gknn.ptr, R, p.wD, p.wR)
    n = 6 * G
    JTJ = DeviceArray((n, n), np.float32, stream)
    T0 = np.tile(np.eye(4, dtype=np.float32).reshape(16), (G, 1))
    dT, dP, dw = dev(T0), DeviceArray((G, 16), np.float32, stream), DeviceArray((M, 3), np.float32, stream)

    def reset():
        dT.upload(T0)
        L.call("vf_warp_mesh", K, dw.ptr, canon.ptr, dT.ptr, knn.ptr, kw.ptr, M, G, stream.ptr)
        stream.synchronize()

    def timed(fn):
        t0 = time.perf_counter()
        fn()
        stream.synchronize()
        return (time.perf_counter() - t0) * 1e3

    ne = [timed(lambda: L.call("vf_gn_normal_equations", None, JTJ.ptr, None, None, C.byref(desc), stream.ptr))
          for _ in range(3)]
    del JTJ
    solver = C.c_void_p()
    create, per_iter, rep = [], [], None
    for r in range(3):
        reset()
        create.append(timed(lambda: L.call("vf_gn_solver_create", C.byref(solver), C.byref(desc), stream.ptr)))
        rep = L.GNReport()
        ms = timed(lambda: L.call("vf_gn_solver_iterate", solver, dT.ptr, dP.ptr, dw.ptr, canon.ptr, M, dmesh.ptr,
                                  iters, -1e30, C.byref(rep), stream.ptr))
        per_iter.append(ms / iters)
        L.call("vf_gn_solver_destroy", solver)
    c_ms, ne_ms, it_ms = min(create), min(ne), min(per_iter)
    lu_ms = max(c_ms - ne_ms, 1e-6)
    chol = bool(rep.cholesky)
    flops = (1.0 if chol else 2.0) / 3.0 * float(n) ** 3  # Cholesky n^3/3, LU 2n^3/3
    out = {
        "workload": f"gn4096: {G} nodes, {M} mesh points ({p.corr.shape[0]} correspondences), k={K}, "
                    f"{R} regularisation neighbours, {n}x{n} normal matrix",
        "rows": int(p.corr.shape[0] + 3 * G * R),
        "normal_equations_ms": ne_ms,
        "solver_create_ms": c_ms,
        "factorisation": "cholesky (blocked, look-ahead)" if chol else "lu (getrf)",
        "factor_ms_est": lu_ms,
        "factor_tflops_est": flops / (lu_ms * 1e-3) / 1e12,
        "factor_frac_fp32_peak": flops / (lu_ms * 1e-3) / FP32_PEAK,
        "gn_iteration_ms": it_ms,
        "iterations": iters,
        "icp_iteration_ms": c_ms + iters * it_ms,
        "refactorised_every_iteration_ms": iters * (c_ms + it_ms),
        "energy_first_last": [rep.energy[0], rep.energy[iters - 1]],
    }
    out["speedup_vs_refactorising"] = out["refactorised_every_iteration_ms"] / out["icp_iteration_ms"]
    return out


def bench_batching(vols, rendered, steps: int, warmup: int) -> dict:
    """16 plain per-volume calls per step against one frame plan (autotuned,
    replayed as a HIP graph), each on its own fresh copy of the volumes."""
    from volumefusion_amd import _lib as L
    from volumefusion_amd import ops
    from volumefusion_amd.device import DeviceArray, Event, Stream

    lib = L.load()
    stream = Stream()
    n = len(vols)
    inputs = []  # per volume: ([depth0, depth1], [T0, T1]) on the device
    for i, v in enumerate(vols):
        (d0, T0), (d1, T1) = rendered[2 * i], rendered[2 * i + 1]
        inputs.append(([DeviceArray.from_numpy(d0, stream), DeviceArray.from_numpy(d1, stream)],
                       [DeviceArray.from_numpy(T0, stream), DeviceArray.from_numpy(T1, stream)]))

    def volume_set():
        tsdfs = [ops.DeviceTsdf(v.size, v.dim, center=v.center, delta=v.delta, stream=stream) for v in vols]
        graphs = [ops.DeviceMotionGraph(v.nodes, rendered[2 * i][1], v.radii, stream=stream)
                  for i, v in enumerate(vols)]
        return tsdfs, graphs

    params = [ops.make_params(v.k, v.delta, v.max_weight, v.cube) for v in vols]
    e0, e1 = Event(), Event()

    # plain: one call per volume
    tsdfs, graphs = volume_set()
    ws = ops.Workspace(stream=stream)
    vol_s = [ops.make_volume(t) for t in tsdfs]
    gr_s, fr_s = [], []
    for i in range(n):
        pair_g, pair_f = [], []
        for f in range(2):
            g = ops.make_graph(graphs[i], write_back=True)
            g.transforms = inputs[i][1][f].ptr
            pair_g.append(g)
            pair_f.append(ops.make_frame(inputs[i][0][f], vols[i].intrinsics()))
        gr_s.append(pair_g)
        fr_s.append(pair_f)
    fn = lib.vf_warped_integrate

    def plain_step(k):
        for i in range(n):
            L.check(fn(ws.handle, C.byref(vol_s[i]), C.byref(gr_s[i][k % 2]), C.byref(fr_s[i][k % 2]),
                       C.byref(params[i]), stream.ptr), "vf_warped_integrate")

    for k in range(warmup):
        plain_step(k)
    stream.synchronize()
    ws.totals(reset=True, stream=stream)
    e0.record(stream)
    for k in range(steps):
        plain_step(warmup + k)
    e1.record(stream)
    stream.synchronize()
    plain_ms = e0.elapsed_ms(e1) / steps
    plain_updates = ws.totals(stream=stream)["updated"] / steps
    del tsdfs, graphs, ws

    # batched: one plan per input parity over a second fresh set; the first
    # is autotuned, the second starts from the tuning cache
    tsdfs, graphs = volume_set()
    plans = []
    tune = None
    for f in range(2):
        plans.append(ops.Plan(tsdfs, graphs, [inputs[i][0][f] for i in range(n)], [v.intrinsics() for v in vols],
                              vols[0].k, [v.delta for v in vols], [v.max_weight for v in vols],
                              [v.cube for v in vols], transforms=[inputs[i][1][f] for i in range(n)], graph=True,
                              stream=stream))
        if f == 0:
            tune = plans[0].autotune(frames=2, stream=stream)
    assert plans[1].tuning()["brick_shape"] == tune["brick_shape"]
    for k in range(warmup):
        plans[k % 2].launch(stream)
    stream.synchronize()
    for p_ in plans:
        p_.stats(reset=True, stream=stream)
    e0.record(stream)
    for k in range(steps):
        plans[(warmup + k) % 2].launch(stream)
    e1.record(stream)
    stream.synchronize()
    batched_ms = e0.elapsed_ms(e1) / steps
    batched_updates = sum(p_.stats(stream=stream)[1]["updated"] for p_ in plans) / steps
    for p_ in plans:
        p_.close()
    return {
        "workload": f"batch16: {n} volumes, {min(v.dim[0] for v in vols)}^3..{max(v.dim[0] for v in vols)}^3, "
                    f"{min(v.nodes.shape[0] for v in vols)}..{max(v.nodes.shape[0] for v in vols)} nodes, k=8, "
                    "640x576 frames",
        "voxels": int(sum(v.n_voxels for v in vols)), "nodes": int(sum(v.nodes.shape[0] for v in vols)),
        "plain_ms_per_step": plain_ms, "batched_ms_per_step": batched_ms, "speedup": plain_ms / batched_ms,
        "plain_updates_per_s": plain_updates / (plain_ms * 1e-3),
        "batched_updates_per_s": batched_updates / (batched_ms * 1e-3),
        "updates_per_step": [plain_updates, batched_updates],
        "tuning": {"brick": list(tune["brick"]), "wg_per_cu": tune["wg_per_cu"],
                   "tuned_ms_per_frame": tune["ms_per_frame"]},
        "steps": steps,
    }


def bench_serving(vols, rendered, seconds: float, fps: float = 30.0, seed: int = 5, max_batch: int = 8,
                  lanes: int = 8, rounds: int = 32,
                  ramp_rates=(120.0, 240.0, 300.0, 360.0, 400.0, 440.0, 480.0, 500.0, 520.0, 540.0, 560.0)) -> dict:
    """Poisson arrivals into the serving loop at 30 fps per stream and at
    rising rates (one second each; the steps narrow where the p99 latency
    approaches 33 ms), the streams saturated, and plain per-volume calls
    (each frame uploaded, then integrated) on resident copies of the same
    streams.
    The saturated server and the plain calls both run `rounds` frames per
    stream (32: the server's fill and drain are then a small part of the
    run; with 8 rounds they cost it about an eighth)."""
    from volumefusion_amd import _lib as L
    from volumefusion_amd import ops, serve
    from volumefusion_amd.device import DeviceArray, Event, PinnedArray, Stream

    n = len(vols)
    # eight lanes of at most eight frames: a lane uploads its batch while the
    # others compute (measured 1.4x the saturated rate of two lanes of 32)
    srv = serve.Server(200 << 30, k=8, max_batch=max_batch, lanes=lanes)
    ids, keep = [], []
    for v in vols:
        d, kp = serve.scene_desc(v)
        keep.append(kp)
        ids.append(srv.open(d))
    # every stream's two frames in page-locked memory, submitted without copies
    pinned = [[(PinnedArray(rendered[2 * i + f][0]), PinnedArray(rendered[2 * i + f][1])) for f in range(2)]
              for i in range(n)]
    # warm-up: batches of every size class grow the lane workspaces
    for r in range(3):
        for i, sid in enumerate(ids):
            srv.submit_borrowed(sid, *pinned[i][r % 2])
        srv.drain()
    srv.poll(1 << 20)

    def poisson_run(rate, length, rseed):
        rng = np.random.default_rng([seed, rseed])
        arrivals = []
        for i in range(n):
            t = rng.exponential(1.0 / rate)
            f = 1
            while t < length:
                arrivals.append((t, i, f))
                t += rng.exponential(1.0 / rate)
                f += 1
        arrivals.sort()
        upd0 = srv.stats()["totals"]["updated"]
        late = []
        t0 = time.perf_counter()
        for t, i, f in arrivals:
            target = t0 + t
            while True:
                now = time.perf_counter()
                if now >= target:
                    break
                if target - now > 2e-3:
                    time.sleep(target - now - 1e-3)
            late.append(now - target)
            srv.submit_borrowed(ids[i], *pinned[i][f % 2])
        srv.drain()
        wall = time.perf_counter() - t0
        res = srv.poll(1 << 22)
        lat = [r["latency_ms"] for r in res]
        return {"fps_per_stream": rate, "offered_fps": len(arrivals) / length, "frames": len(res),
                "achieved_fps": len(res) / wall, "p50_ms": _pct(lat, 50), "p99_ms": _pct(lat, 99),
                "max_ms": max(lat) if lat else None,
                "mean_batch_frames": float(np.mean([r["batch_frames"] for r in res])) if res else 0.0,
                "updates_per_s": (srv.stats()["totals"]["updated"] - upd0) / wall,
                "submit_lateness_p99_ms": _pct([x * 1e3 for x in late], 99)}

    poisson = poisson_run(fps, seconds, 0)
    ramp = [poisson_run(rate, 1.0, r + 1) for r, rate in enumerate(ramp_rates)]
    within = [x for x in [poisson] + ramp if x["p99_ms"] is not None and x["p99_ms"] < 33.0]
    capacity = max(within, key=lambda x: x["achieved_fps"]) if within else None

    # saturated: every stream has frames queued
    upd1 = srv.stats()["totals"]["updated"]
    t0 = time.perf_counter()
    for r in range(rounds):
        for i, sid in enumerate(ids):
            srv.submit_borrowed(sid, *pinned[i][r % 2])
    srv.drain()
    sat_wall = time.perf_counter() - t0
    res = srv.poll(1 << 20)
    st = srv.stats()
    sat_lat = [r_["latency_ms"] for r_ in res]
    saturated = {"frames": rounds * n, "seconds": sat_wall, "fps": rounds * n / sat_wall,
                 "updates_per_s": (st["totals"]["updated"] - upd1) / sat_wall,
                 "mean_batch_frames": float(np.mean([r_["batch_frames"] for r_ in res])),
                 "p50_ms": _pct(sat_lat, 50), "p99_ms": _pct(sat_lat, 99)}
    resident = sum(8 * v.n_voxels for v in vols)  # the streams' tsdf and weights
    hbm = {"bytes_admitted": st["bytes_in_use"], "budget": st["hbm_budget"], "streams": st["streams_open"],
           "tsdf_weight_bytes": resident, "admitted_vs_tsdf_weight": st["bytes_in_use"] / resident,
           "redo_bricks": st["totals"]["redo_bricks"]}
    srv.close_server()

    # plain per-volume calls: upload the frame, integrate, next volume
    lib = L.load()
    stream = Stream()
    ws = ops.Workspace(stream=stream)
    items = []
    for i, v in enumerate(vols):
        t = ops.DeviceTsdf(v.size, v.dim, center=v.center, delta=v.delta, stream=stream)
        g = ops.DeviceMotionGraph(v.nodes, rendered[2 * i][1], v.radii, stream=stream)
        d = DeviceArray.from_numpy(rendered[2 * i][0], stream)
        items.append((t, g, d, ops.make_volume(t), ops.make_graph(g, write_back=True),
                      ops.make_frame(d, v.intrinsics()), ops.make_params(v.k, v.delta, v.max_weight, v.cube)))
    fn = lib.vf_warped_integrate
    h2d = lib.vf_memcpy_h2d

    def plain_round(r, upload):
        for i, it in enumerate(items):
            if upload:
                dp, Tp = pinned[i][r % 2]
                L.check(h2d(it[2].ptr, dp.ptr, dp.nbytes, stream.ptr), "vf_memcpy_h2d")
                L.check(h2d(it[1].transforms.ptr, Tp.ptr, Tp.nbytes, stream.ptr), "vf_memcpy_h2d")
            L.check(fn(ws.handle, C.byref(it[3]), C.byref(it[4]), C.byref(it[5]), C.byref(it[6]), stream.ptr),
                    "vf_warped_integrate")

    plain = {}
    for upload in (True, False):
        plain_round(0, upload)
        stream.synchronize()
        ws.totals(reset=True, stream=stream)
        t0 = time.perf_counter()
        for r in range(rounds):
            plain_round(r + 1, upload)
        stream.synchronize()
        wall = time.perf_counter() - t0
        plain["with_upload" if upload else "resident"] = {
            "frames": rounds * n, "seconds": wall, "fps": rounds * n / wall,
            "updates_per_s": ws.totals(stream=stream)["updated"] / wall}
    for p_ in pinned:
        for a, b in p_:
            a.free()
            b.free()
    return {
        "workload": f"streams32: {n} streams, {min(v.dim[0] for v in vols)}^3..{max(v.dim[0] for v in vols)}^3, "
                    f"{min(v.nodes.shape[0] for v in vols)}..{max(v.nodes.shape[0] for v in vols)} nodes, k=8, "
                    f"640x576 frames from page-locked host memory, Poisson arrivals",
        "poisson": poisson, "ramp": ramp,
        "capacity_within_p99_33ms_fps": capacity["achieved_fps"] if capacity else None,
        "server": {"max_batch": max_batch, "lanes": lanes, "rounds_per_stream": rounds},
        "saturated": saturated, "plain": plain,
        "saturated_vs_plain": saturated["fps"] / plain["with_upload"]["fps"], "hbm": hbm,
    }


def main() -> int:
    args = parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE={world}); the warped "
              "integration path has no multi-GPU mode in this build", file=sys.stderr)
        return 2

    from volumefusion_amd import scenes

    # host-side inputs first: rendering forks worker processes, which must
    # happen before this process touches the device
    s = workload(args.workload)
    c4 = c5 = None
    jobs = [(s, 0), (s, 1)]
    if not args.no_batching:
        c4 = scenes.batch16()
        jobs += [(v, f) for v in c4 for f in (0, 1)]
    if not args.no_serving:
        c5 = scenes.streams32()
        jobs += [(v, f) for v in c5 for f in (0, 1)]
    rendered = scenes.render_frames(jobs)
    print(f"inputs ready ({len(jobs)} frames)", file=sys.stderr, flush=True)

    from volumefusion_amd import _lib as L
    from volumefusion_amd import device, ops
    from volumefusion_amd.device import DeviceArray, Event, Stream

    L.load()
    if device.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1

    depths = [rendered[0][0], rendered[1][0]]
    Ts = [rendered[0][1], rendered[1][1]]
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

    # brick shape and workgroups per CU from the autotuner (scratch copies of
    # the volume; results do not depend on the choice)
    tuning = {"brick": [16, 8, 4], "wg_per_cu": 0, "tuned_ms_per_frame": None}
    if args.tuning:
        shape, wg = (int(x) for x in args.tuning.split(","))
        ws.set_tuning(shape, wg)
        tuning = {"brick": list(L.BRICK_SHAPES[shape]), "wg_per_cu": wg, "tuned_ms_per_frame": None}
    elif not args.no_tune:
        plan = ops.Plan([tsdf], [graph], [d_dev[0]], [s.intrinsics()], s.k, s.delta, s.max_weight, [s.cube],
                        transforms=[T_dev[0]], graph=False, stream=stream)
        t = plan.autotune(frames=3, stream=stream)
        plan.close()
        ws.set_tuning(t["brick_shape"], t["wg_per_cu"])
        tuning = {"brick": list(t["brick"]), "wg_per_cu": t["wg_per_cu"], "tuned_ms_per_frame": t["ms_per_frame"]}
        print(f"autotuned: brick {t['brick']}, {t['wg_per_cu']} workgroups per CU", file=sys.stderr, flush=True)

    for i in range(args.warmup):
        step(i)
    stream.synchronize()
    ws.totals(reset=True, stream=stream)
    # the timed frames carry HIP events on pass A's launches (its duration for
    # the roofline); the prep and pass-B durations come from a short run after
    ws.set_timing(True, pass_a_only=True)
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
    tot = ws.totals(stream=stream)
    nb = min(args.steps, 20)
    ws.set_timing(True)
    for i in range(nb):
        step(args.warmup + args.steps + i)
    kb, nkb = ws.kernel_ms()
    ws.set_timing(False)
    kms = [kb[0] / nkb * nk, kms[1], kb[2] / nkb * nk]
    print(f"{s.name}: {wall / args.steps * 1e3:.3f} ms per frame", file=sys.stderr, flush=True)

    U = tot["updated"] / args.steps
    ab = algorithmic_bytes(U, W, H, G)
    t_a = kms[1] / nk * 1e-3
    t_frame = sum(kms) / nk * 1e-3
    achieved = ab["k_pass_a"] / t_a
    # HBM bytes and VALU instructions per k_pass_a launch, from the committed
    # rocprofv3 counter runs of this workload: valid only for the brick shape
    # and the kernel sources they were taken with (else reported as stale)
    traffic = None
    valu = None
    counters_note = None
    tfile = REPO / "profiles" / f"traffic_{args.workload}.json"
    if tfile.exists():
        try:
            prof = json.loads(tfile.read_text())
            ka = prof.get("k_pass_a", {})
            if prof.get("kernel_sources_sha256") != kernel_sources_sha256():
                counters_note = (f"stale: {tfile.name} was measured on other kernel sources; traffic and "
                                 f"valu_roofline omitted")
            elif prof.get("brick") != tuning["brick"]:
                counters_note = f"stale: {tfile.name} was measured with brick {prof.get('brick')}"
            else:
                traffic = ka.get("hbm_bytes_per_launch")
                if ka.get("valu_insts_per_launch"):
                    issue = ka["valu_insts_per_launch"] * VALU_CYCLES / (SIMDS * CLOCK_HZ)
                    valu = {"bound": "valu", "kernel": "k_pass_a", "insts_per_launch": ka["valu_insts_per_launch"],
                            "issue_ms": issue * 1e3, "avg_launch_ms": t_a * 1e3, "frac": issue / t_a,
                            "model": "wave64 VALU instruction = 4 cycles on a 16-lane SIMD, 1024 SIMDs at 2.4 GHz"}
        except (ValueError, AttributeError, TypeError):
            traffic = valu = None

    if args.copy_calibration:
        # a plain device-to-device copy of known size, for scaling HBM counters
        nbytes = 1 << 30
        a, b = DeviceArray((nbytes // 4,), np.float32, stream), DeviceArray((nbytes // 4,), np.float32, stream)
        a.fill(1.0)
        for _ in range(3):
            b.copy_from(a)
        stream.synchronize()
        del a, b

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
            "tuning": tuning,
        },
        "roofline": {
            "bound": "hbm", "kernel": "k_pass_a", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
            "algorithmic_bytes_per_launch": ab["k_pass_a"], "avg_launch_ms": t_a * 1e3,
            **({"counters": counters_note} if counters_note else {}),
        },
        "valu_roofline": valu,
        "frame_roofline": {
            "achieved_GBps": ab["frame"] / t_frame / 1e9, "frac": ab["frame"] / t_frame / HBM_PEAK,
            "algorithmic_bytes": ab["frame"], "kernel_ms": {"k_prep_bin": kms[0] / nk, "k_pass_a": kms[1] / nk,
                                                            "k_pass_b": kms[2] / nk},
            "gpu_ms_per_step": gpu_ms / args.steps,
        },
        "dense_equivalent_voxels_per_s": s.n_voxels * args.steps / wall,
    }
    if not args.no_extras:
        result["extras"] = bench_extras(s, tsdf, graph, depths[0], stream)
        ex = result["extras"]
        print(f"extras: marching cubes {ex['marching_cubes']['ms']:.3f} ms, ingest "
              f"{ex['depth_ingest']['ms_wall_incl_centroid_readback']:.3f} ms, point warp "
              f"{ex['point_warp']['ms']:.3f} ms", file=sys.stderr, flush=True)
    result["plain_sphere128_device"] = bench_plain_sphere128(stream)
    del tsdf, graph, ws
    if c4 is not None:
        off = 2
        result["batching"] = bench_batching(c4, rendered[off:off + 2 * len(c4)], steps=20, warmup=3)
        b = result["batching"]
        print(f"batching: plain {b['plain_ms_per_step']:.3f} ms, batched {b['batched_ms_per_step']:.3f} ms "
              f"per step ({b['speedup']:.2f}x)", file=sys.stderr, flush=True)
    if c5 is not None:
        off = 2 + (2 * len(c4) if c4 is not None else 0)
        result["serving"] = bench_serving(c5, rendered[off:off + 2 * len(c5)], args.serve_seconds)
        sv = result["serving"]
        print(f"serving: p50 {sv['poisson']['p50_ms']:.2f} ms, p99 {sv['poisson']['p99_ms']:.2f} ms, "
              f"saturated {sv['saturated']['fps']:.0f} fps vs plain {sv['plain']['with_upload']['fps']:.0f} fps",
              file=sys.stderr, flush=True)
    if not args.no_gn:
        result["gauss_newton"] = bench_gauss_newton(stream)
        g = result["gauss_newton"]
        print(f"gauss-newton: normal equations {g['normal_equations_ms']:.2f} ms, solver create "
              f"{g['solver_create_ms']:.2f} ms ({g['factorisation']} ~{g['factor_tflops_est']:.1f} TFLOP/s), "
              f"{g['gn_iteration_ms']:.3f} ms per iteration", file=sys.stderr, flush=True)
    if not args.no_fp64:
        result["fp64_error"] = fp64_error((L.BRICK_SHAPES.index(tuple(tuning["brick"])), tuning["wg_per_cu"]),
                                          min(16, os.cpu_count() or 1))
        fe = result["fp64_error"]
        print("fp64 error: " + ", ".join(f"{k} max {v['max']:.3g} rms {v['rms_last_frame']:.3g}"
                                         for k, v in fe.items() if isinstance(v, dict)), file=sys.stderr, flush=True)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(s, depths[0], Ts[0], args.cpu_seconds)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
