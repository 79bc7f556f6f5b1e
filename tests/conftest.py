import json
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: test needs a ROCm GPU and the built HIP library")


def pytest_collection_modifyitems(config, items):
    import torch

    if torch.cuda.is_available():
        return
    skip = pytest.mark.skip(reason="no ROCm GPU visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


_PARITY = {}


class _Parity:
    """Records the worst error of a parity test in units of its scale."""

    def __init__(self, nodeid):
        self.nodeid = nodeid

    def record(self, rel_err, bar):
        worst, _ = _PARITY.get(self.nodeid, (0.0, bar))
        _PARITY[self.nodeid] = (max(worst, float(rel_err)), bar)


@pytest.fixture
def parity(request):
    return _Parity(request.node.nodeid)


def pytest_terminal_summary(terminalreporter):
    if not _PARITY:
        return
    terminalreporter.section("parity: worst error / scale per test (bar)")
    for nodeid, (worst, bar) in sorted(_PARITY.items()):
        terminalreporter.write_line(f"{worst:10.3e}  (bar {bar:.0e})  {nodeid}")


@pytest.fixture(scope="session")
def ref_vectors():
    return json.loads((GOLDEN / "reference_vectors.json").read_text())


@pytest.fixture(scope="session")
def configs():
    return json.loads((GOLDEN / "configs.json").read_text())


@pytest.fixture(scope="session")
def dev():
    import torch

    return torch.device("cuda", 0)
