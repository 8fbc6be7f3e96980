"""TSDF files (host side, no device needed): the native layout and the
reference's TSDFVolume::save layout (src/TSDFVolume.cpp:315-358) with its
24-byte size field fixed."""
import struct

import numpy as np
import pytest

from volumefusion_amd import _lib as L
from volumefusion_amd import tsdf_io

F = np.float32


def volume(dim, seed=0):
    rng = np.random.default_rng(seed)
    n = dim[0] * dim[1] * dim[2]
    return rng.uniform(-1, 1, n).astype(F), rng.integers(0, 16, n).astype(F)


def test_native_round_trip_and_header(tmp_path):
    dim = (7, 5, 3)
    t, w = volume(dim)
    p = tmp_path / "v.vft"
    tsdf_io.save_host(p, t, w, dim, (0.35, 0.25, 0.15), center=(0.1, -0.2, 1.0), delta=0.012, max_weight=64.0)
    raw = p.read_bytes()
    assert len(raw) == 64 + 8 * t.size
    assert raw[:8] == b"VFTSDF\x00\x01"
    hb, dx, dy, dz = struct.unpack_from("<I3i", raw, 8)
    assert (hb, dx, dy, dz) == (64, 7, 5, 3)
    size = struct.unpack_from("<3f", raw, 24)
    center = struct.unpack_from("<3f", raw, 36)
    delta, maxw, flags = struct.unpack_from("<ffI", raw, 48)
    assert size == tuple(F(x) for x in (0.35, 0.25, 0.15)) and center == tuple(F(x) for x in (0.1, -0.2, 1.0))
    assert delta == F(0.012) and maxw == 64.0 and flags == 1
    np.testing.assert_array_equal(np.frombuffer(raw, F, t.size, 64), t)
    np.testing.assert_array_equal(np.frombuffer(raw, F, t.size, 64 + 4 * t.size), w)
    info, t2, w2 = tsdf_io.load_host(p)
    assert info["dim"] == dim and info["format"] == tsdf_io.NATIVE and info["max_weight"] == 64.0
    np.testing.assert_array_equal(t2, t)
    np.testing.assert_array_equal(w2, w)


def test_legacy_layout(tmp_path):
    """Byte offsets of TSDFVolume::save: dims 0, size 12 (24 bytes: the size
    and the voxel size that follows it in the object), delta 36, tsdf 40,
    weights 40 + 4N, then 3N colour bytes and 4N colour weights."""
    dim = (4, 3, 2)
    t, w = volume(dim, 1)
    n = t.size
    p = tmp_path / "v.tsdf"
    tsdf_io.save_host(p, t, w, dim, (0.4, 0.3, 0.2), delta=0.01, fmt=tsdf_io.LEGACY)
    raw = p.read_bytes()
    assert len(raw) == 12 + 24 + 4 + 4 * n + 4 * n + 3 * n + 4 * n
    assert struct.unpack_from("<3i", raw, 0) == dim
    assert struct.unpack_from("<3f", raw, 12) == tuple(F(x) for x in (0.4, 0.3, 0.2))
    # m_voxelSize, which TSDFVolume::load takes from the second half of the
    # field: size / float(dim) in float
    assert struct.unpack_from("<3f", raw, 24) == tuple(F(x) / F(d) for x, d in zip((0.4, 0.3, 0.2), dim))
    assert struct.unpack_from("<f", raw, 36)[0] == F(0.01)
    np.testing.assert_array_equal(np.frombuffer(raw, F, n, 40), t)
    np.testing.assert_array_equal(np.frombuffer(raw, F, n, 40 + 4 * n), w)
    assert raw[40 + 8 * n:] == bytes(7 * n)


def test_legacy_file_with_garbage_in_the_size_field(tmp_path):
    """A file as the reference writes it: 12 bytes of whatever followed the
    Vec3f size in the object.  Only the first 12 bytes are read."""
    dim = (3, 3, 3)
    t, w = volume(dim, 2)
    n = t.size
    raw = struct.pack("<3i", *dim) + struct.pack("<3f", 0.3, 0.3, 0.3) + bytes(range(200, 212)) + \
        struct.pack("<f", 0.02) + t.tobytes() + w.tobytes() + bytes(range(3 * n)) + np.ones(n, F).tobytes()
    p = tmp_path / "ref.tsdf"
    p.write_bytes(raw)
    info, t2, w2 = tsdf_io.load_host(p)
    assert info["format"] == tsdf_io.LEGACY and info["dim"] == dim
    assert info["size"] == (F(0.3),) * 3 and info["delta"] == F(0.02)
    np.testing.assert_array_equal(t2, t)
    np.testing.assert_array_equal(w2, w)


def test_errors(tmp_path):
    dim = (4, 4, 4)
    t, w = volume(dim)
    p = tmp_path / "v.vft"
    tsdf_io.save_host(p, t, w, dim, (1, 1, 1))
    raw = p.read_bytes()
    (tmp_path / "short.vft").write_bytes(raw[:-10])
    with pytest.raises(L.VFError):
        tsdf_io.load_host(tmp_path / "short.vft")
    bad = bytearray(raw)
    bad[7] = 9
    (tmp_path / "ver.vft").write_bytes(bytes(bad))
    with pytest.raises(L.VFError, match="version"):
        tsdf_io.load_host(tmp_path / "ver.vft")
    with pytest.raises(L.VFError, match="cannot open"):
        tsdf_io.load_host(tmp_path / "missing.vft")
    # capacity: a 64-voxel file into 10-voxel buffers
    import ctypes as C

    info = L.TsdfInfo()
    small = np.zeros(10, F)
    st = L.load().vf_tsdf_load_host(str(p).encode(), C.byref(info), small.ctypes.data, small.ctypes.data, 10)
    assert st == 6  # VF_ERR_CAPACITY
