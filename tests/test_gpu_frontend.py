"""Batches and request streams read by the front-end worker pool (with wide noisy count requests
run by its trajectory worker processes) give the same Results (counts, fingerprints,
expectations) as the same work in the launching process alone. The pool has to be forked before
a process first uses the GPU, so the check runs in a child process."""

from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu


def test_pool_results_equal_in_process_results() -> None:
    child = Path(__file__).with_name("frontend_child.py")
    out = subprocess.run([sys.executable, str(child)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-3000:]
    rec = json.loads(out.stdout.strip().splitlines()[-1])
    assert rec == {"workers": 3, "trajectory_workers": 2, "batch_same": True, "noisy_same": True, "error_kept": True, "serve_same": True}, rec
