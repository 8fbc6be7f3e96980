"""Marching cubes (src/MarchingCubes.cu) in the oracle: the packed table, and
closed forms on analytic fields."""
import re
from pathlib import Path

import numpy as np
import pytest

F = np.float32
TABLES = Path(__file__).resolve().parent.parent / "volumefusion_amd" / "csrc" / "vf_mc_tables.h"
EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]


def tri_table():
    words = [int(x, 16) for x in re.findall(r"0x([0-9a-f]{16})ull", TABLES.read_text())]
    assert len(words) == 256
    out = []
    for w in words:
        t = w & 15
        out.append([[(w >> (4 + 4 * (3 * q + v))) & 15 for v in range(3)] for q in range(t)])
    return out


def crossing_edges(c):
    return {e for e, (a, b) in enumerate(EDGES) if ((c >> a) ^ (c >> b)) & 1}


def test_table_uses_exactly_the_crossed_edges():
    """Every case's triangles use each edge its corner signs cross, and no
    other; complementary cases cross the same edges."""
    tab = tri_table()
    assert tab[0] == [] and tab[255] == []
    for c in range(256):
        used = {e for tri in tab[c] for e in tri}
        assert used == crossing_edges(c), c
        assert all(len(set(t)) == 3 for t in tab[c]), c
        assert len(tab[c]) <= 5
        assert crossing_edges(c) == crossing_edges(255 - c)


def test_table_triangles_close_each_component():
    """On each cube face the crossed edges come in pairs joined by one triangle
    edge each: every triangle edge that lies on a cube face joins two crossed
    edges of that face, and each crossed edge is used by exactly that many
    face-boundary segments (one per face it borders)."""
    faces = [{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 4, 8, 9}, {1, 5, 9, 10}, {2, 6, 10, 11}, {3, 7, 11, 8}]
    tab = tri_table()
    for c in range(1, 255):
        count = {}
        for tri in tab[c]:
            for a, b in ((tri[0], tri[1]), (tri[1], tri[2]), (tri[2], tri[0])):
                key = frozenset((a, b))
                count[key] = count.get(key, 0) + 1
        # segments used once lie on the cube surface: both edges on a common face
        boundary = [k for k, n in count.items() if n == 1]
        for seg in boundary:
            assert any(seg <= f for f in faces), (c, seg)
        deg = {}
        for seg in boundary:
            for e in seg:
                deg[e] = deg.get(e, 0) + 1
        assert all(d == 2 for d in deg.values()), c


def plane_field(dim, vs, size, z0, delta=0.05):
    X, Y, Z = dim
    z = np.arange(Z, dtype=F) * F(vs) - F(size) * F(0.5)
    t = ((z - F(z0)) / F(delta)).astype(F)
    tsdf = np.broadcast_to(t[:, None, None], (Z, Y, X)).astype(F).ravel()
    return tsdf, np.ones_like(tsdf)


@pytest.mark.parametrize("full", [False, True])
def test_plane_closed_form(oracle, full):
    """tsdf linear in z crossing 0 between two layers: one vertex per vertical
    edge of that cube layer at height z0 (4 per cube, shared edges repeated),
    or two triangles per cube."""
    dim, vs, size = (12, 10, 9), 0.1, 0.9
    z0 = 0.0123
    tsdf, wt = plane_field(dim, vs, size, z0)
    verts, faces = oracle.marching_cubes(tsdf, wt, dim, (vs,) * 3, (size,) * 3, 0.0, full=full)
    ncubes = (dim[0] - 2) * (dim[1] - 2)
    if full:
        assert faces.shape[0] == 2 * ncubes and verts.shape[0] == 6 * ncubes
        np.testing.assert_array_equal(faces.ravel(), np.arange(verts.shape[0]))
    else:
        assert verts.shape[0] == 4 * ncubes
    np.testing.assert_allclose(verts[:, 2], z0, atol=1e-6)
    xs = np.unique(verts[:, 0])
    assert xs.size == dim[0] - 1  # x from 0 to X-2 (cubes i < X-2)


def test_weight_zero_corner_and_last_layer(oracle):
    """A cube with a weight-0 corner emits nothing (src/MarchingCubes.cu:352-357);
    crossings in the last cube layer (i = X-2) are not visited."""
    dim, vs, size = (8, 8, 8), 0.1, 0.8
    tsdf, wt = plane_field(dim, vs, size, 0.0)
    v_all, _ = oracle.marching_cubes(tsdf, wt, dim, (vs,) * 3, (size,) * 3)
    wt2 = wt.copy().reshape(8, 8, 8)
    wt2[:, :, 0] = 0  # x = 0 column of voxels
    v_cut, _ = oracle.marching_cubes(tsdf, wt2.ravel(), dim, (vs,) * 3, (size,) * 3)
    assert v_cut.shape[0] == v_all.shape[0] - 4 * (dim[1] - 2)
    # a field crossing only between x = X-2 and X-1
    X = dim[0]
    x = np.arange(X)
    t = np.where(x >= X - 1, 1.0, -1.0).astype(F)
    tsdf3 = np.broadcast_to(t[None, None, :], (8, 8, 8)).astype(F).ravel()
    v3, _ = oracle.marching_cubes(tsdf3, wt, dim, (vs,) * 3, (size,) * 3)
    assert v3.shape[0] == 0


def test_sphere_vertices_on_surface(oracle):
    n, vs = 40, 0.05
    size = n * vs
    c = np.arange(n) * vs - size / 2
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(x * x + y * y + z * z)
    R = 0.7
    tsdf = np.clip((r - R) / 0.1, -1, 1).astype(F).ravel()
    wt = np.ones_like(tsdf)
    verts, faces = oracle.marching_cubes(tsdf, wt, (n,) * 3, (vs,) * 3, (size,) * 3, full=True)
    rr = np.linalg.norm(verts.astype(np.float64), axis=1)
    assert faces.shape[0] > 1000
    assert np.abs(rr - R).max() < 0.1 * vs
