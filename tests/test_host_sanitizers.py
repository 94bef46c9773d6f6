"""The C ABI's host code under AddressSanitizer and UndefinedBehaviorSanitizer (no GPU needed):
every csrc/*.hip file is compiled with the sanitizers on the host side only (device code is
unchanged), linked into a separate library, and tests/sanitize/abi_driver.c calls each entry
point with bad arguments (every one must return LSIM_EINVAL with nothing reported) and the host
cumulative sum with good ones."""

from __future__ import annotations

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "lucivant-sim_amd" / "csrc"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CLANG = "/opt/rocm/llvm/bin/clang"
HOST_SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]


@pytest.mark.skipif(not (Path(HIPCC).exists() and Path(CLANG).exists()), reason="ROCm's hipcc and clang are needed")
def test_abi_validation_is_clean_under_host_sanitizers(tmp_path: Path) -> None:
    flags = ["--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

    def compile_one(src: Path) -> Path:
        obj = tmp_path / (src.stem + ".o")
        proc = subprocess.run([HIPCC, *flags, *HOST_SAN, "-c", str(src), "-o", str(obj)], capture_output=True, text=True, timeout=900)
        assert proc.returncode == 0, proc.stderr[-2000:]
        return obj

    with ThreadPoolExecutor(max_workers=4) as pool:
        objects = list(pool.map(compile_one, sorted(CSRC.glob("*.hip"))))
    lib = tmp_path / "liblsim_san.so"
    proc = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *HOST_SAN, "-o", str(lib), *map(str, objects)], capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr[-2000:]
    driver = tmp_path / "abi_driver"
    proc = subprocess.run(
        [CLANG, "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "sanitize" / "abi_driver.c"),
         "-o", str(driver), f"-L{tmp_path}", "-llsim_san", f"-Wl,-rpath,{tmp_path}", "-Wl,-rpath,/opt/rocm/lib"],
        capture_output=True, text=True,
    )  # fmt: skip
    assert proc.returncode == 0, proc.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    run = subprocess.run([str(driver)], capture_output=True, text=True, timeout=300, env=env)
    assert run.returncode == 0, run.stdout[-3000:] + run.stderr[-3000:]
    assert "ok: 0 failures" in run.stdout
    assert "runtime error" not in run.stderr and "AddressSanitizer" not in run.stderr
