"""Host sanitizer runs (CPU): the oracle's C restatement and the `main`
driver's host utilities, built with AddressSanitizer and
UndefinedBehaviorSanitizer (tests/sanitize/Makefile), run their self-tests
(tests/sanitize/selftest_*.c*) with halt-on-error, so an out-of-bounds access
or undefined operation fails the test with the sanitizer's report."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    if not shutil.which("gcc") or not shutil.which("g++"):
        pytest.skip("gcc / g++ not available")
    out = tmp_path_factory.mktemp("san")
    r = subprocess.run(["make", "-s", "-C", HERE, f"OUT={out}"], capture_output=True, text=True,
                       timeout=300)
    if r.returncode != 0 and "asan" in (r.stderr + r.stdout).lower():
        pytest.skip("sanitizer runtime not available: " + r.stderr[-300:])
    assert r.returncode == 0, r.stderr[-2000:]
    return out


@pytest.mark.parametrize("prog", ["selftest_oracle", "selftest_host"])
def test_sanitized_selftest(built, prog, tmp_path):
    # verify_asan_link_order=0: a library preloaded by the environment may
    # come before the ASan runtime in the link order
    env = dict(os.environ,
               ASAN_OPTIONS="halt_on_error=1:detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="2",
               SANITIZE_TMP=str(tmp_path))
    r = subprocess.run([os.path.join(built, prog)], capture_output=True, text=True, timeout=300,
                       env=env)
    assert r.returncode == 0, r.stdout[-1000:] + r.stderr[-3000:]
    assert "ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
