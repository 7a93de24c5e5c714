import os
import sys

import pytest

# Measured errors and figures that tests report; printed in the terminal
# summary so that they appear in a quiet (-q) run's log as well.
_REPORTED = []

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X); run with -m gpu")
    config.addinivalue_line("markers", "slow: longer CPU test")


def pytest_collection_modifyitems(config, items):
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:  # pragma: no cover
        has_gpu = False
    if has_gpu:
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture
def report():
    """report(line): add a measured figure to the run's terminal summary."""
    def add(line: str) -> None:
        print(line)
        _REPORTED.append(line)
    return add


def pytest_terminal_summary(terminalreporter):
    if _REPORTED:
        terminalreporter.section("measured errors and figures")
        for line in _REPORTED:
            terminalreporter.write_line(line)
