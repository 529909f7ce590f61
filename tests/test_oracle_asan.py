"""Host sanitizer build of the CPU oracle: oracle/orpic.c and a driver that
calls every oracle entry point on exactly-sized buffers
(tests/asan/oracle_asan.c), compiled with AddressSanitizer and
UndefinedBehaviorSanitizer (gcc, host code only) and run; any access
outside a buffer, use after free or undefined arithmetic aborts the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOURCES = [os.path.join(ROOT, "oracle", "orpic.c"), os.path.join(ROOT, "tests", "asan", "oracle_asan.c")]
FLAGS = ["-std=c11", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off", "-fopenmp",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]


def test_oracle_under_address_and_ub_sanitizers(tmp_path):
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("gcc not available")
    exe = str(tmp_path / "oracle_asan")
    build = subprocess.run([cc] + FLAGS + ["-I" + os.path.join(ROOT, "oracle")] + SOURCES +
                           ["-lm", "-o", exe], capture_output=True, text=True)
    assert build.returncode == 0, build.stderr
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               OMP_NUM_THREADS="2")
    run = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "oracle asan run: ok" in run.stdout
    assert "runtime error" not in run.stderr, run.stderr
