"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: this runs without a GPU)."""
import re
import subprocess
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (REPO / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[\w\s\*]*?\b(vf_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_reference_entry_points():
    names = declared_symbols()
    for ref in ("vf_tsdf_knn_graph", "vf_compute_tsdf_warp_weights", "vf_generate_tsdf_vertices", "vf_warp_tsdf",
                "vf_count_collisions_per_voxel", "vf_integrate_warped_tsdf", "vf_integrate_tsdf",
                "vf_integrate_tsdf_camera_frame", "vf_translate", "vf_warped_integrate",
                "vf_warped_integrate_batch"):
        assert ref in names


def test_library_exports_every_declared_symbol():
    import volumefusion_amd._lib as L

    lib_path = L.LIB_PATH
    assert lib_path.exists(), "build the library first (__graft_entry__.build())"
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib_path)], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in declared_symbols() if n not in exported]
    assert not missing, missing
    # the ctypes binding declares exactly the header's entry points
    assert sorted(L.exported_symbols()) == declared_symbols()


def test_library_loads_without_gpu_and_reports_status():
    import volumefusion_amd._lib as L

    lib = L.load()
    assert lib.vf_version().startswith(b"volumefusion_amd")
    assert lib.vf_status_string(0) == b"ok"
    assert lib.vf_status_string(2) == b"unsupported knn slot count"


def test_invalid_arguments_rejected_on_host():
    """Argument checks run on the host before any device work, as errors
    (the reference throws std::runtime_error)."""
    import volumefusion_amd._lib as L

    with pytest.raises(L.VFError) as e:
        L.call("vf_compute_tsdf_warp_weights", 9, None, None, None, None, None, 10, None)
    assert e.value.status == 2
    with pytest.raises(L.VFError) as e:
        L.call("vf_generate_tsdf_vertices", None, L.f3((0.01,) * 3), L.f3((1,) * 3), L.i3((0, 4, 4)), None)
    assert e.value.status == 1
    with pytest.raises(L.VFError):
        L.call("vf_integrate_tsdf", None, None, None, L.Int2(4, 4), 0.01, L.i3((4, 4, 4)), L.f3((0.01,) * 3),
               L.f3((0.04,) * 3), L.m3([1, 0, 0, 0, 1, 0, 0, 0, 1]), L.f3((0, 0, 0)), L.m3([1] * 9), 15.0, None)
    # the Gauss-Newton solver's factorisation choice is checked first
    import ctypes as C

    solver = C.c_void_p()
    with pytest.raises(L.VFError) as e:
        L.call("vf_gn_solver_create_ex", C.byref(solver), None, 7, None)
    assert e.value.status == 1 and "factorisation 7" in str(e.value)
    with pytest.raises(L.VFError) as e:
        L.call("vf_gn_solver_create_ex", C.byref(solver), None, 1, None)
    assert e.value.status == 1 and "NULL problem" in str(e.value)
    assert not solver.value


def test_voxels_per_node_matches_reference_formula():
    """src/VolumeFusion.cu:702-705: 2 * (int(R / vs) + 1) + 1 in float32.
    float32(0.02) / float32(0.002) = 9.999999, which truncates to 9: the 2 mm
    volume gets a 21^3 cube, not 23^3."""
    from volumefusion_amd import scenes

    assert scenes.voxels_per_node(0.025, (0.004,) * 3) == (15, 15, 15)
    assert scenes.voxels_per_node(0.020, (0.002,) * 3) == (21, 21, 21)
    assert scenes.voxels_per_node(0.020, (0.003,) * 3) == (15, 15, 15)


def test_server_admission_bytes_formula():
    """The admission charge grows with the volume, the graph and the lanes
    (pure host arithmetic, no device needed)."""
    import numpy as np

    from volumefusion_amd import serve

    def charge(dim, G, lanes):
        d, keep = serve.stream_desc((dim,) * 3, (0.5,) * 3, (0, 0, 1), 640, 576, np.eye(3).reshape(-1), 0.01, 64.0,
                                    (11, 11, 11), np.zeros((G, 3), np.float32), np.full(G, 0.02, np.float32))
        return serve.stream_bytes(d, lanes)

    base = charge(128, 512, 1)
    assert base >= 8 * 128 ** 3 + 512 * 80 + 640 * 576 * 4
    assert charge(256, 512, 1) > base
    assert charge(128, 1024, 1) > base
    assert charge(128, 512, 2) > base
