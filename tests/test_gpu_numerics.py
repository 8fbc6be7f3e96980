"""Exhaustive device checks of the library's rounding helpers.

tests/cpp/numerics_check.hip compiles the helpers of
volumefusion_amd/csrc/vf_device.hpp into two kernels and compares them with
the operations they stand for, on every input of their range:

- rcp_rn(x), the reciprocal used for 1/z in the projection (the reference's
  zInv = 1.0f / z, src/TSDFIntegration.cu:21) and for the weight sum
  (src/TSDFWarp.cu:112-113), against the IEEE quotient 1.0f / x for every
  positive float from 2^-126 through +inf;
- sqrt_rn(s), the distance square root of the knn search (Eigen's norm,
  src/TSDFWarp.cu:54), against the correctly rounded __builtin_sqrtf for
  every non-negative float and +inf.

The binary is built by build() (make -C tests/cpp numerics).
"""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "_build" / "numerics_check"


def test_numerics_check_is_built():
    assert BIN.exists(), "tests/cpp/_build/numerics_check missing: run __graft_entry__.build()"


@pytest.mark.gpu
def test_rounding_helpers_exhaustive():
    proc = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=100)
    out = proc.stdout
    lines = {ln.split()[0]: ln for ln in out.splitlines() if " checked " in ln}
    assert set(lines) == {"rcp", "sqrt"}, out + proc.stderr
    for name, ln in lines.items():
        words = ln.split()
        checked, bad = int(words[2]), int(words[4])
        assert checked > 2_000_000_000, ln
        assert bad == 0, ln
    assert proc.returncode == 0, out + proc.stderr
