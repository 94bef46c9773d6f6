"""Build-time guard on the gfx950 kernels: no kernel may use scratch (private) memory. A kernel
that spills runs at lower occupancy, and the runtime must allocate its scratch per queue on first
launch, which has failed when a second stream launched a spilling kernel. Every csrc/*.hip file
is compiled for the device alone with the compiler's resource-usage remarks, and each kernel's
ScratchSize must be 0."""

from __future__ import annotations

import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "lucivant-sim_amd" / "csrc"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _usage(src: Path, tmp: Path) -> dict[str, int]:
    cmd = [HIPCC, *FLAGS, "--cuda-device-only", "-c", str(src), "-o", str(tmp / (src.stem + ".o")), "-Rpass-analysis=kernel-resource-usage"]
    proc = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert proc.returncode == 0, proc.stderr[-2000:]
    kernels: dict[str, int] = {}
    name = None
    for line in proc.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            continue
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name is not None:
            kernels[name] = int(m.group(1))
            name = None
    return kernels


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc is not installed")
@pytest.mark.parametrize("src", sorted(CSRC.glob("*.hip")), ids=lambda p: p.name)
def test_no_kernel_uses_scratch(src: Path, tmp_path: Path) -> None:
    kernels = _usage(src, tmp_path)
    spilling = {k: v for k, v in kernels.items() if v}
    assert not spilling, spilling
    if src.name == "lsim_program.hip":
        assert any("k_program" in k for k in kernels)
    if src.name == "lsim_sv.hip":
        assert any("k_dense_mfma" in k for k in kernels) and any("k_dense" in k for k in kernels)
