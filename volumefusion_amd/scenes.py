"""Seeded synthetic workloads for the warped-integration path.

The reference reads recorded depth frames and builds its motion graph from
them (src/VolumeFusion.cu:367-404); with no dataset available here, frames are
ray-cast from analytic star-shaped surfaces (spheres, ellipsoids with a
low-frequency radial displacement), nodes are Poisson-disk samples on those
surfaces, and node transforms are smooth random rigid motions.  Everything is
a deterministic function of the seed.

Conventions follow the reference: camera frame = world frame, the volume is
centred on `center` (the reference's frame centroid), voxel size is
size / float(dim) in float32 (include/af/DeviceTsdf.cuh:64), transforms are
column-major 4x4 (Eigen::Matrix4f storage), intrinsics column-major 3x3.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

F32 = np.float32


@dataclass
class Camera:
    width: int
    height: int
    fx: float
    fy: float
    cx: float
    cy: float

    def intrinsics(self) -> np.ndarray:
        """Column-major Mat3f: [fx 0 cx; 0 fy cy; 0 0 1]."""
        K = np.array([[self.fx, 0, self.cx], [0, self.fy, self.cy], [0, 0, 1]], np.float32)
        return np.ascontiguousarray(K.T.reshape(9))


@dataclass
class Surface:
    """Star-shaped surface: x(u) = center + R @ (axes * u) * (1 + eps * f(u))
    for unit directions u, f(u) = sum_i a_i sin(k_i . u + phi_i), |f| <= 1."""
    center: np.ndarray
    axes: np.ndarray
    rot: np.ndarray
    eps: float = 0.0
    waves_k: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    waves_a: np.ndarray = field(default_factory=lambda: np.zeros(0))
    waves_phi: np.ndarray = field(default_factory=lambda: np.zeros(0))

    def f(self, u: np.ndarray) -> np.ndarray:
        if self.eps == 0.0 or len(self.waves_a) == 0:
            return np.zeros(u.shape[:-1])
        ph = u @ self.waves_k.T + self.waves_phi
        return np.sin(ph) @ self.waves_a

    def point(self, u: np.ndarray) -> np.ndarray:
        s = 1.0 + self.eps * self.f(u)
        return self.center + ((self.axes * u) * s[..., None]) @ self.rot.T

    def area(self) -> float:
        a, b, c = self.axes
        p = 1.6075
        return 4 * math.pi * (((a * b) ** p + (a * c) ** p + (b * c) ** p) / 3) ** (1 / p)


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def make_surface(rng, center, axes, eps=0.0, rot=None, nwaves=4) -> Surface:
    rot = random_rotation(rng) if rot is None else rot
    k = rng.normal(size=(nwaves, 3))
    k *= (rng.uniform(1.5, 3.0, size=nwaves) / np.linalg.norm(k, axis=1))[:, None]
    a = rng.uniform(-1, 1, size=nwaves)
    a /= max(np.abs(a).sum(), 1e-9)
    phi = rng.uniform(0, 2 * math.pi, size=nwaves)
    return Surface(np.asarray(center, float), np.asarray(axes, float), rot, eps, k, a, phi)


def render_depth(cam: Camera, surfaces: list[Surface], rng: np.random.Generator, noise_sigma: float = 0.0,
                 dropout: float = 0.0, steps: int = 48, bisect: int = 30) -> np.ndarray:
    """Depth (z in metres) of the nearest surface hit per pixel centre, 0 where
    no surface is hit; Gaussian noise on valid pixels, then dropout to 0."""
    H, W = cam.height, cam.width
    u = (np.arange(W) - cam.cx) / cam.fx
    v = (np.arange(H) - cam.cy) / cam.fy
    dx, dy = np.meshgrid(u, v)
    d = np.stack([dx, dy, np.ones_like(dx)], -1).reshape(-1, 3)  # ray x(t) = t d, t = depth
    best = np.full(d.shape[0], np.inf)
    for s in surfaces:
        inv_axes = 1.0 / s.axes
        # normalised coordinates y(t) = (R^T (t d - c)) / axes = t a - b
        a = (d @ s.rot) * inv_axes
        b = (s.rot.T @ s.center) * inv_axes
        rmax = 1.0 + abs(s.eps) + 1e-6
        A = np.einsum("ij,ij->i", a, a)
        B = -2 * (a @ b)
        Cc = b @ b - rmax * rmax
        disc = B * B - 4 * A * Cc
        hit = disc > 0
        sq = np.sqrt(np.where(hit, disc, 0))
        t0 = (-B - sq) / (2 * A)
        t1 = (-B + sq) / (2 * A)
        t0 = np.maximum(t0, 1e-3)
        hit &= t1 > t0
        idx = np.nonzero(hit)[0]
        if idx.size == 0:
            continue
        aa, ta, tb = a[idx], t0[idx], t1[idx]

        def g(t):
            y = t[:, None] * aa - b
            n = np.linalg.norm(y, axis=1)
            un = y / np.maximum(n, 1e-12)[:, None]
            return n - (1.0 + s.eps * s.f(un))

        prev_t = ta.copy()
        prev_g = g(prev_t)
        found = np.zeros(idx.size, bool)
        lo = np.zeros(idx.size)
        hi = np.zeros(idx.size)
        for i in range(1, steps + 1):
            t = ta + (tb - ta) * (i / steps)
            gv = g(t)
            cross = (~found) & (prev_g > 0) & (gv <= 0)
            lo[cross] = prev_t[cross]
            hi[cross] = t[cross]
            found |= cross
            prev_t, prev_g = t, gv
        sel = np.nonzero(found)[0]
        if sel.size == 0:
            continue
        lo, hi = lo[sel], hi[sel]
        aa_s = aa[sel]
        for _ in range(bisect):
            mid = 0.5 * (lo + hi)
            y = mid[:, None] * aa_s - b
            n = np.linalg.norm(y, axis=1)
            gm = n - (1.0 + s.eps * s.f(y / np.maximum(n, 1e-12)[:, None]))
            inside = gm <= 0
            hi = np.where(inside, mid, hi)
            lo = np.where(inside, lo, mid)
        tt = 0.5 * (lo + hi)
        gi = idx[sel]
        best[gi] = np.minimum(best[gi], tt)
    depth = np.where(np.isfinite(best), best, 0.0)
    valid = depth > 0
    if noise_sigma > 0:
        depth = np.where(valid, depth + rng.normal(0.0, noise_sigma, size=depth.shape), 0.0)
    if dropout > 0:
        depth = np.where(rng.uniform(size=depth.shape) < dropout, 0.0, depth)
    return depth.reshape(H, W).astype(np.float32)


def poisson_disk(surfaces: list[Surface], n: int, rng: np.random.Generator, rmin: float | None = None) -> np.ndarray:
    """n points on the surfaces, pairwise >= rmin apart (dart throwing in
    batches of 4096 candidates, each accepted in order when no accepted point
    lies closer than rmin; rmin shrinks by 10% whenever 30 batches add
    nothing)."""
    from scipy.spatial import cKDTree

    area = sum(s.area() for s in surfaces)
    if rmin is None:
        rmin = 0.8 * math.sqrt(area / max(n, 1))
    weights = np.array([s.area() for s in surfaces]) / area
    pts = np.empty((0, 3))
    stale = 0
    while len(pts) < n:
        which = rng.choice(len(surfaces), size=4096, p=weights)
        u = rng.normal(size=(4096, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        cand = np.empty((4096, 3))
        for i, s in enumerate(surfaces):
            m = which == i
            if m.any():
                cand[m] = s.point(u[m])
        # against the points accepted so far
        if len(pts):
            d, _ = cKDTree(pts).query(cand, k=1, distance_upper_bound=rmin)
            cand = cand[~(d < rmin)]
        # within the batch, in candidate order
        accepted = []
        if len(cand):
            near = cKDTree(cand).query_ball_point(cand, rmin * (1.0 - 1e-12))
            taken = np.zeros(len(cand), bool)
            for i in range(len(cand)):
                if len(pts) + len(accepted) >= n:
                    break
                if not any(taken[j] for j in near[i] if j < i):
                    taken[i] = True
                    accepted.append(i)
        if accepted:
            pts = np.concatenate([pts, cand[accepted]])
            stale = 0
        else:
            stale += 1
            if stale >= 30:
                rmin *= 0.9
                stale = 0
    return np.array(pts[:n], np.float32)


def rodrigues(w: np.ndarray) -> np.ndarray:
    """Rotation matrices [G,3,3] from rotation vectors [G,3]."""
    th = np.linalg.norm(w, axis=1)
    k = w / np.maximum(th, 1e-300)[:, None]
    K = np.zeros((w.shape[0], 3, 3))
    K[:, 0, 1], K[:, 0, 2] = -k[:, 2], k[:, 1]
    K[:, 1, 0], K[:, 1, 2] = k[:, 2], -k[:, 0]
    K[:, 2, 0], K[:, 2, 1] = -k[:, 1], k[:, 0]
    s, c = np.sin(th)[:, None, None], np.cos(th)[:, None, None]
    return np.eye(3)[None] + s * K + (1 - c) * (K @ K)


def identity_transforms(G: int) -> np.ndarray:
    return np.tile(np.eye(4, dtype=np.float32).reshape(16), (G, 1))


def smooth_random_warps(nodes: np.ndarray, rng: np.random.Generator, max_angle_deg: float = 5.0,
                        max_translation: float = 0.010, smooth_radius: float | None = None) -> np.ndarray:
    """Per-node rigid motions about the node (T n = n + t): rotation about a
    random axis by U[0, max angle], translation uniform in a ball, both averaged
    over graph neighbours within smooth_radius.  Column-major float32 [G, 16]."""
    G = nodes.shape[0]
    if G == 0:
        return np.zeros((0, 16), np.float32)
    axis = rng.normal(size=(G, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    ang = np.deg2rad(rng.uniform(0, max_angle_deg, size=G))
    w = axis * ang[:, None]
    t = rng.normal(size=(G, 3))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    t *= (max_translation * rng.uniform(0, 1, size=G) ** (1 / 3))[:, None]
    n64 = nodes.astype(np.float64)
    if smooth_radius is None:
        from scipy.spatial import cKDTree
        dd, _ = cKDTree(n64).query(n64, k=min(2, G))
        spacing = float(np.median(dd[:, -1])) if G > 1 else 0.0
        smooth_radius = 2.5 * spacing
    if smooth_radius > 0 and G > 1:
        from scipy.spatial import cKDTree
        tree = cKDTree(n64)
        nb = tree.query_ball_point(n64, smooth_radius)
        w = np.stack([w[j].mean(axis=0) for j in nb])
        t = np.stack([t[j].mean(axis=0) for j in nb])
    R = rodrigues(w)
    T = np.zeros((G, 4, 4))
    T[:, :3, :3] = R
    T[:, :3, 3] = n64 + t - np.einsum("gij,gj->gi", R, n64)
    T[:, 3, 3] = 1.0
    return np.ascontiguousarray(T.transpose(0, 2, 1).reshape(G, 16)).astype(np.float32)


def voxels_per_node(radius: float, voxel_size) -> tuple:
    """2 * (int(R / vs) + 1) + 1 per axis, src/VolumeFusion.cu:702-705 (float32)."""
    r = F32(radius)
    return tuple(2 * (int(F32(r / F32(v))) + 1) + 1 for v in voxel_size)


@dataclass
class VolumeScene:
    """One volume, its motion graph and its frames."""
    name: str
    dim: tuple
    size: tuple
    center: tuple
    cam: Camera
    nodes: np.ndarray          # [G,3] float32
    radii: np.ndarray          # [G] float32
    graph_radius: float        # settings.motionGraphRadius (node cube)
    k: int
    delta: float
    max_weight: float
    surfaces: list
    seed: int
    noise_sigma: float = 0.0
    dropout: float = 0.0
    warp_angle_deg: float = 0.0
    warp_translation: float = 0.0

    @property
    def voxel_size(self) -> tuple:
        return tuple(float(F32(s) / F32(d)) for s, d in zip(self.size, self.dim))

    @property
    def n_voxels(self) -> int:
        return int(self.dim[0]) * int(self.dim[1]) * int(self.dim[2])

    @property
    def cube(self) -> tuple:
        return voxels_per_node(self.graph_radius, self.voxel_size)

    def intrinsics(self) -> np.ndarray:
        return self.cam.intrinsics()

    def depth(self, frame: int = 0) -> np.ndarray:
        rng = np.random.default_rng([self.seed, 1, frame])
        return render_depth(self.cam, self.surfaces, rng, self.noise_sigma, self.dropout)

    def transforms(self, frame: int = 0) -> np.ndarray:
        if self.warp_angle_deg == 0.0 and self.warp_translation == 0.0:
            return identity_transforms(self.nodes.shape[0])
        rng = np.random.default_rng([self.seed, 2, frame])
        return smooth_random_warps(self.nodes, rng, self.warp_angle_deg, self.warp_translation)

    def summary(self) -> dict:
        return {"name": self.name, "dim": list(self.dim), "voxel_size": list(self.voxel_size),
                "nodes": int(self.nodes.shape[0]), "k": self.k, "cube": list(self.cube),
                "frame": [self.cam.width, self.cam.height]}


CAM_640 = Camera(640, 576, 504.0, 504.0, 320.0, 288.0)
CAM_1024 = Camera(1024, 1024, 900.0, 900.0, 512.0, 512.0)


def sphere128(seed: int = 0) -> VolumeScene:
    """128^3 at 4 mm centred at z=1 m; one 640x576 frame of a sphere (r=0.3 m
    at 1 m, sigma 1 mm); 256 Poisson-disk nodes, identity transforms, k=4,
    r=25 mm, delta 12 mm, max weight 64."""
    rng = np.random.default_rng([seed, 0])
    sph = make_surface(rng, (0.0, 0.0, 1.0), (0.3, 0.3, 0.3), eps=0.0, rot=np.eye(3))
    nodes = poisson_disk([sph], 256, rng)
    return VolumeScene("sphere128", (128, 128, 128), (0.512, 0.512, 0.512), (0.0, 0.0, 1.0), CAM_640, nodes,
                       np.full(256, 0.025, np.float32), 0.025, 4, 0.012, 64.0, [sph], seed, noise_sigma=0.001)


def sphere256_warped(seed: int = 0) -> VolumeScene:
    """256^3 at 3 mm; 640x576 frames of a sphere; 2048 nodes with smooth random
    warps (<= 5 deg, <= 10 mm); k=8, r=20 mm."""
    rng = np.random.default_rng([seed, 0])
    sph = make_surface(rng, (0.0, 0.0, 1.0), (0.3, 0.3, 0.3), eps=0.0, rot=np.eye(3))
    nodes = poisson_disk([sph], 2048, rng)
    return VolumeScene("sphere256_warped", (256, 256, 256), (0.768, 0.768, 0.768), (0.0, 0.0, 1.0), CAM_640, nodes,
                       np.full(2048, 0.020, np.float32), 0.020, 8, 0.012, 64.0, [sph], seed, noise_sigma=0.001,
                       warp_angle_deg=5.0, warp_translation=0.010)


def ellipsoids512(seed: int = 0) -> VolumeScene:
    """512^3 at 2 mm; 1024x1024 frame (fx=fy=900) of two intersecting deformed
    ellipsoids, sigma 1.5 mm, 5% dropout; 4096 nodes, k=8, r=20 mm."""
    rng = np.random.default_rng([seed, 0])
    e1 = make_surface(rng, (-0.09, 0.0, 1.0), (0.30, 0.22, 0.18), eps=0.06)
    e2 = make_surface(rng, (0.12, 0.05, 1.05), (0.22, 0.28, 0.20), eps=0.06)
    nodes = poisson_disk([e1, e2], 4096, rng)
    return VolumeScene("ellipsoids512", (512, 512, 512), (1.024, 1.024, 1.024), (0.0, 0.0, 1.0), CAM_1024, nodes,
                       np.full(4096, 0.020, np.float32), 0.020, 8, 0.012, 64.0, [e1, e2], seed, noise_sigma=0.0015,
                       dropout=0.05, warp_angle_deg=5.0, warp_translation=0.010)


def random_volume(name: str, rng_seed, dim: int, nodes: int, side: float = 0.8) -> VolumeScene:
    """A randomly posed deformed sphere in a side^3 volume of dim^3 voxels."""
    rng = np.random.default_rng(rng_seed)
    r = rng.uniform(0.18, 0.28)
    c = (rng.uniform(-0.08, 0.08), rng.uniform(-0.08, 0.08), 1.0 + rng.uniform(-0.08, 0.08))
    s = make_surface(rng, c, (r, r, r), eps=0.05)
    pts = poisson_disk([s], nodes, rng)
    spacing = math.sqrt(s.area() / nodes)
    radius = float(np.float32(max(0.020, 0.5 * spacing)))
    return VolumeScene(name, (dim, dim, dim), (side, side, side), (0.0, 0.0, 1.0), CAM_640, pts,
                       np.full(nodes, radius, np.float32), radius, 8, 0.012, 64.0, [s], int(rng_seed[-1]),
                       noise_sigma=0.001, warp_angle_deg=5.0, warp_translation=0.010)


def batch16(seed: int = 0, count: int = 16) -> list[VolumeScene]:
    """16 volumes, resolution log-uniform in [64, 384]^3, nodes log-uniform in
    [128, 4096], k=8, each with its own 640x576 frame."""
    rng = np.random.default_rng([seed, 4])
    out = []
    for i in range(count):
        dim = int(round(math.exp(rng.uniform(math.log(64), math.log(384)))))
        g = int(round(math.exp(rng.uniform(math.log(128), math.log(4096)))))
        out.append(random_volume(f"batch16_{i}", [seed, 40, i], dim, g))
    return out


def streams32(seed: int = 0, count: int = 32) -> list[VolumeScene]:
    """32 streams, volumes 128^3..512^3, 512..4096 nodes, k=8."""
    rng = np.random.default_rng([seed, 5])
    out = []
    for i in range(count):
        dim = int(round(math.exp(rng.uniform(math.log(128), math.log(512)))))
        g = int(round(math.exp(rng.uniform(math.log(512), math.log(4096)))))
        out.append(random_volume(f"streams32_{i}", [seed, 50, i], dim, g))
    return out


def small_scene(seed: int = 0, dim=(40, 36, 44), nodes: int = 48, k: int = 4, radius: float = 0.02,
                warps: bool = True, vs: float = 0.005, cam: Camera | None = None) -> VolumeScene:
    """A few-second-oracle workload for parity tests: uneven dims (partial
    bricks), a sphere cutting through the volume, optional random warps."""
    rng = np.random.default_rng([seed, 9])
    size = tuple(float(np.float32(d * vs)) for d in dim)
    r = 0.35 * min(size)
    c = (0.0, 0.0, 0.6)
    s = make_surface(rng, (c[0] + 0.01, c[1] - 0.01, c[2] + 0.02), (r, r, r), eps=0.0, rot=np.eye(3))
    pts = poisson_disk([s], nodes, rng)
    cam = cam or Camera(160, 144, 126.0, 126.0, 80.0, 72.0)
    return VolumeScene("small", tuple(dim), size, c, cam, pts, np.full(nodes, radius, np.float32), radius, k, 0.012,
                       64.0, [s], seed, noise_sigma=0.001, warp_angle_deg=5.0 if warps else 0.0,
                       warp_translation=0.01 if warps else 0.0)


def _render_job(job):
    s, frame = job
    return s.depth(frame), s.transforms(frame)


def render_frames(jobs: list, workers: int | None = None) -> list:
    """(depth, transforms) for each (scene, frame) in `jobs`, rendered on a
    pool of forked CPU processes.  Call it before any device work: the pool
    forks this process."""
    import concurrent.futures
    import multiprocessing
    import os

    workers = workers or min(16, os.cpu_count() or 1)
    if workers <= 1 or len(jobs) <= 1:
        return [_render_job(j) for j in jobs]
    ctx = multiprocessing.get_context("fork")
    with concurrent.futures.ProcessPoolExecutor(max_workers=workers, mp_context=ctx) as ex:
        return list(ex.map(_render_job, jobs, chunksize=1))


@dataclass
class GNProblem:
    """One linearisation of the reference's motion optimisation
    (src/VolumeFusion.cu:434-692): a motion graph with its regularisation
    neighbours, a canonical mesh bound to the graph by k nearest nodes and
    normalised Gaussian weights, and correspondences of mesh points to target
    points with unit normals."""
    graph: np.ndarray    # [G, 3] float32 node positions
    T: np.ndarray        # [G, 16] float32 column-major transforms
    gknn: np.ndarray     # [G, R] uint32 regularisation neighbours (>= G: none)
    canon: np.ndarray    # [M, 3] float32 canonical mesh
    knn: np.ndarray      # [M, K] uint32 nearest nodes of each mesh point
    kw: np.ndarray       # [M, K] float32 normalised weights (0 beyond the radius)
    corr: np.ndarray     # [C, 2] int32 (mesh index, target index)
    dmesh: np.ndarray    # [D, 3] float32 target points
    dnormals: np.ndarray # [D, 3] float32 unit normals
    wD: float = 1.0
    wR: float = 1.0


def gn_problem(seed: int, G: int = 24, M: int = 600, K: int = 6, R: int = 4, radius: float = 0.05,
               motion=None, warp_angle_deg: float = 2.0, warp_translation: float = 0.004,
               wD: float = 1.0, wR: float = 0.5, drop_gknn: float = 0.1) -> GNProblem:
    """Nodes on a sphere patch, mesh points near the surface, targets on the
    mesh points moved by `motion` (a twist u [G, 6] applied as T - [u], or a
    random rigid motion of each node) with noise-free normals."""
    rng = np.random.default_rng([seed, 60])
    K = min(K, G)
    dirs = rng.normal(size=(G, 3))
    dirs[:, 2] = np.abs(dirs[:, 2]) + 1.0
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    graph = (0.3 * dirs + np.array([0.0, 0.0, 0.8])).astype(F32)
    # regularisation neighbours: the R nearest other nodes, some dropped
    d2 = ((graph[:, None, :].astype(np.float64) - graph[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    gknn = np.full((G, R), 0xFFFFFFFF, np.uint32)  # fewer than R other nodes: the rest absent
    gknn[:, :min(R, G - 1)] = np.argsort(d2, axis=1)[:, :min(R, G - 1)]
    gknn[rng.random(gknn.shape) < drop_gknn] = np.uint32(0xFFFFFFFF)
    # mesh points around the nodes
    from scipy.spatial import cKDTree

    src = rng.integers(0, G, M)
    canon = (graph[src] + rng.normal(scale=0.4 * radius, size=(M, 3))).astype(F32)
    dk, knn = cKDTree(graph.astype(np.float64)).query(canon.astype(np.float64), k=K)
    dk, knn = dk.reshape(M, K), knn.reshape(M, K).astype(np.uint32)
    w = np.where(dk < radius, np.exp(-(dk ** 2) / (2 * radius ** 2)), 0.0)
    s = w.sum(1, keepdims=True)
    w = np.where(s > 0, w / np.where(s > 0, s, 1.0), 0.0)
    kw = w.astype(F32)
    T = np.tile(np.eye(4, dtype=F32).reshape(16), (G, 1))
    if motion is None:
        motion = smooth_random_warps(graph, rng, warp_angle_deg, warp_translation)
        T_true = motion
    else:
        T_true = T.copy()
        for g in range(G):
            x = motion[g]
            U = np.array([[0, -x[2], x[1], x[3]], [x[2], 0, -x[0], x[4]], [-x[1], x[0], 0, x[5]], [0, 0, 0, 0]])
            T_true[g] = (T[g].reshape(4, 4).T - U).T.reshape(16)
    # targets: every mesh point warped by the true transforms (binary64,
    # slots in order)
    Mt = T_true.reshape(G, 4, 4).astype(np.float64)  # Mt[g, j, a] = T_g(a, j)
    p64 = canon.astype(np.float64)
    tgt = np.zeros((M, 3))
    for s_ in range(K):
        Ts = Mt[knn[:, s_]]
        moved = (Ts[:, 0, :3] * p64[:, 0:1] + Ts[:, 1, :3] * p64[:, 1:2]) + (Ts[:, 2, :3] * p64[:, 2:3] + Ts[:, 3, :3])
        tgt += np.where(kw[:, s_:s_ + 1] > 0, kw[:, s_:s_ + 1].astype(np.float64) * moved, 0.0)
    normals = rng.normal(size=(M, 3))
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    has_w = kw.sum(1) > 0
    idx = np.nonzero(has_w)[0].astype(np.int32)
    corr = np.stack([idx, idx], 1).astype(np.int32)
    return GNProblem(graph, T, gknn, canon, knn, kw, corr, tgt.astype(F32), normals.astype(F32), wD, wR)
