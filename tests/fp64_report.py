"""Collects the binary64 error of the device tsdf per workload (filled by the
GPU parity tests, printed at the end of the session by conftest.py)."""

RESULTS = []


def record(workload: str, frames: int, max_err: float, rms_last: float, rms_max: float, voxels: int) -> None:
    RESULTS.append({"workload": workload, "frames": frames, "max": max_err, "rms_last": rms_last,
                    "rms_max": rms_max, "voxels": voxels})


def lines() -> list:
    return [f"{r['workload']}: {r['frames']} frame(s), {r['voxels']} voxels with weight > 0: "
            f"max |tsdf - tsdf64| = {r['max']:.3g}, RMS = {r['rms_last']:.3g} (largest RMS over frames "
            f"{r['rms_max']:.3g}), in units of the truncation distance" for r in RESULTS]
