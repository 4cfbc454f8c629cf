"""Host sanitizer builds (no GPU): the CPU oracle driven through every entry point, and the bake server's
host logic (admission, queues, worker threads, resident-mesh cache and eviction, fetch and release) driven by
four client threads over a fake device layer, each under AddressSanitizer + UndefinedBehaviorSanitizer and
under ThreadSanitizer (tests/sanitize/)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")
BINARIES = ["oracle_asan", "oracle_tsan", "server_asan", "server_tsan"]


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    if not shutil.which(os.environ.get("CXX", "g++")) or not os.path.exists("/opt/rocm/llvm/bin/clang++"):
        pytest.skip("host compilers for the sanitizer builds are missing")
    out = str(tmp_path_factory.mktemp("sanitize"))
    subprocess.run(["make", "-s", "-C", HERE, "-j4", f"OUT={out}"], check=True, timeout=600)
    return out


@pytest.mark.parametrize("name", BINARIES)
def test_sanitized_run_is_clean(built, name):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1:exitcode=24",
               TSAN_OPTIONS="halt_on_error=1:exitcode=25")
    r = subprocess.run([os.path.join(built, name)], env=env, capture_output=True, text=True, timeout=600)
    report = r.stdout + r.stderr
    assert r.returncode == 0, report[-4000:]
    assert "ERROR: AddressSanitizer" not in report and "WARNING: ThreadSanitizer" not in report
    assert "runtime error" not in report, report[-4000:]
    assert ": ok" in r.stdout
