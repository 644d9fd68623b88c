"""AddressSanitizer and UndefinedBehaviorSanitizer builds on the CPU
(tests/sanitize/Makefile): the FP64 oracle over every scheme, noise mode,
adaptive dt and the parameter.h switches; and the product library's
host-side C-ABI code (parameter validation, footprints, step counts, the MT
stream and the reference's random start, disc noise, the initial-state
generators, the startpos reader, the job server's failed-admission path and
its ownership of failed jobs).  Each must finish without a sanitizer
report."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
OUT = os.path.join(SAN, "_build")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


@pytest.fixture(scope="module")
def built():
    objs = [os.path.join(ROOT, "bd-simulations_amd", "_build", n) for n in ("bdk_kernels.hip.o", "bdk_noalloc.hip.o")]
    if not all(os.path.exists(o) for o in objs):
        subprocess.check_call([sys.executable, os.path.join(ROOT, "bd-simulations_amd", "build.py")])
    out = subprocess.run(["make", "-s", "-C", SAN], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr


@pytest.mark.parametrize("name", ["oracle_check", "host_check"])
def test_sanitized_run_is_clean(built, name):
    out = subprocess.run([os.path.join(OUT, name)], capture_output=True, text=True, timeout=300, env=ENV)
    report = out.stdout + out.stderr
    assert out.returncode == 0, report
    assert "runtime error" not in report and "AddressSanitizer" not in report, report
    assert "%s: OK" % name in out.stdout
