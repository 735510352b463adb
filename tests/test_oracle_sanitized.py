"""The CPU oracle under AddressSanitizer, UndefinedBehaviorSanitizer and
bounds-checked index macros.

The oracle's own CPU tests (restatements against each other, closed forms,
golden fixtures, whole model steps) run in a child process against the checked
build of the same sources (oracle/Makefile target `sanitize`:
-fsanitize=address,undefined, -DHT3D_ORACLE_BOUNDS), with the ASan runtime
preloaded.  An index outside an array's extents, a heap overflow or an
undefined operation aborts the child and fails this test.
"""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SUITES = ["tests/test_oracle.py", "tests/test_oracle_pins.py", "tests/test_oracle_dim.py", "tests/test_golden.py"]


def _asan_runtime() -> str | None:
    gcc = shutil.which("gcc")
    if not gcc:
        return None
    path = subprocess.run([gcc, "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return path if os.path.isabs(path) and Path(path).exists() else None


def test_oracle_suite_under_sanitizers():
    asan = _asan_runtime()
    if asan is None:
        pytest.skip("gcc's AddressSanitizer runtime is not installed")
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "sanitize"], check=True)
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan + (":" + env["LD_PRELOAD"] if env.get("LD_PRELOAD") else "")
    env.update(ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               HT3D_ORACLE_LIB=str(ROOT / "oracle" / "libht3d_oracle_san.so"),
               OMP_NUM_THREADS="4", PYTHONPATH=str(ROOT))
    r = subprocess.run([sys.executable, "-m", "pytest", *SUITES, "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-6000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-6000:]
    assert " passed" in out
