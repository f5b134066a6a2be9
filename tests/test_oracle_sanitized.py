"""The least-squares extrapolation oracle (oracle/lsq.c, the bit-exact
checker of the HIP extrapolation) built with AddressSanitizer and
UndefinedBehaviorSanitizer (gcc, host code only) and run standalone
(oracle/lsq_sanitize_main.c): a soft sphere's extrapolation band on the
doubled grid, filled from a linear field that a first-order fit must
reproduce.  The run must end cleanly (no sanitizer report, which would stop
it with a nonzero status), take at least one sweep, fill every target cell
and hold the linear field to 1e-12.  CPU only."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not installed")
def test_lsq_oracle_clean_under_asan_and_ubsan(tmp_path):
    exe = tmp_path / "lsq_sanitized"
    cmd = ["gcc", "-std=c11", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           os.path.join(ROOT, "oracle", "lsq.c"),
           os.path.join(ROOT, "oracle", "lsq_sanitize_main.c"), "-o", str(exe), "-lm"]
    subprocess.run(cmd, check=True, capture_output=True, timeout=120)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120, env=env)
    report = run.stdout + run.stderr
    assert run.returncode == 0, report
    assert "Sanitizer" not in report and "runtime error" not in report, report
    sweeps = int(report.split()[1])
    assert sweeps >= 1
