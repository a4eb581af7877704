"""AddressSanitizer + UndefinedBehaviorSanitizer over the host code and the oracle (CPU only): the mesh
generator, deformation with tangents and the binary mesh file (with a truncated file), and the oracle's
FP32 / FP64 residuals, serial and threaded drivers, frozen limiter and tangent, built with g++ from
tests/sanitize/ and run in a subprocess; any sanitizer report aborts it."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent / "sanitize"


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_code_under_asan_ubsan(tmp_path):
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4", SAN_TMP=str(tmp_path))
    out = subprocess.run([str(HERE / "_build" / "sanitize_check")], capture_output=True, text=True, timeout=600,
                         env=env)
    assert out.returncode == 0, out.stdout[-3000:] + out.stderr[-3000:]
    assert "sanitize: ok" in out.stdout
    assert "runtime error" not in out.stderr and "AddressSanitizer" not in out.stderr
