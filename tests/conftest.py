import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device discovery only)
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


# Measurement lines the GPU tests print (errors against the f32 and FP64 oracles, mass drift)
# are repeated in the terminal summary, so that a quiet run (-q, output captured) still states
# them next to the pass/fail record.
_MEASURED = []


def pytest_runtest_logreport(report):
    if report.when == "call" and report.capstdout:
        lines = [l.strip() for l in report.capstdout.splitlines()
                 if any(k in l for k in ("max|", "drift", "err", "relative"))]
        if lines:
            _MEASURED.append((report.nodeid, lines))


def pytest_terminal_summary(terminalreporter):
    if not _MEASURED:
        return
    terminalreporter.section("measured errors")
    for nodeid, lines in _MEASURED:
        terminalreporter.write_line(nodeid)
        for l in lines:
            terminalreporter.write_line("    " + l)
