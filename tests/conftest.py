"""Test configuration: import paths and the ``gpu`` marker.

``-m "not gpu"`` runs the oracle pins, the host logic (reader, noise, planner, sampling) and the
C-ABI load/validation checks on any machine; ``-m gpu`` runs the parity tests of the gfx950
kernels against the oracle through the C ABI, on the MI355X.
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / "lucivant-sim_amd", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config: pytest.Config) -> None:
    config.addinivalue_line("markers", "gpu: needs the MI355X and the built liblsim.so")


@pytest.fixture(scope="session")
def golden_runs() -> dict:
    return json.loads((GOLDEN / "runs.json").read_text())["runs"]


@pytest.fixture(scope="session")
def golden_programs() -> dict:
    return json.loads((GOLDEN / "programs.json").read_text())


@pytest.fixture(scope="session")
def golden_host() -> dict:
    return json.loads((GOLDEN / "host.json").read_text())
