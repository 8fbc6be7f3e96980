import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: larger CPU workloads")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.build()
    return O


@pytest.fixture(scope="session")
def vf():
    """The HIP library; on a GPU run it must load and see a device."""
    import volumefusion_amd._lib as L
    from volumefusion_amd import device

    L.load()
    n = device.device_count()
    assert n >= 1, "no HIP device visible: GPU tests need an MI355X"
    return L


def pytest_terminal_summary(terminalreporter):
    """The binary64 error report of the GPU parity tests (device tsdf against
    the oracle's binary64 companion, same discrete decisions)."""
    try:
        import fp64_report
    except ImportError:
        return
    if fp64_report.RESULTS:
        terminalreporter.section("binary64 error of the device tsdf")
        for line in fp64_report.lines():
            terminalreporter.write_line(line)
