"""CPU builds of the host-side table builders (tprf_amd/csrc/host_tables.cpp:
tail-fit operator, D_q tails, nu = 0 weights, the Matsubara bubble tables)
and of the oracle under the address and undefined-behaviour sanitizers
(tests/sanitize/Makefile, g++/gcc only).  Each driver exercises its code on a
set of meshes and checks what the mathematics fixes; a sanitizer report
aborts the driver with a non-zero status."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def built():
    if shutil.which("g++") is None or shutil.which("gcc") is None:
        pytest.fail("host compilers missing")
    d = tempfile.mkdtemp()
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "sanitize"), "OUT=" + d],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    yield d
    shutil.rmtree(d, ignore_errors=True)


@pytest.mark.parametrize("prog", ["sanitize_tables", "sanitize_oracle"])
def test_sanitized_run(built, prog):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(built, prog)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "ok: 0 failure(s)" in r.stdout
