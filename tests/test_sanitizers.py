"""The host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only): libsfft.so
built with host-side instrumentation (tests/sanitize/Makefile; device code untouched) runs the
host checks of tests/sanitize/host_checks.c, and the instrumented oracle runs the oracle's own
test files.  A planted out-of-bounds read on the oracle shows that the instrumentation is live.
Builds go to a temporary directory."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_TESTS = ["test_oracle_stages.py", "test_golden.py", "test_oracle_fft.py",
                "test_oracle_e2e.py", "test_oracle_adjoint.py", "test_oracle_fp32.py",
                "test_oracle_params.py", "test_oracle_filter.py"]


@pytest.fixture(scope="module")
def san_build(tmp_path_factory):
    out = tmp_path_factory.mktemp("sanitize")
    jobs = str(max(1, min(8, len(os.sched_getaffinity(0)))))
    subprocess.run(["make", "-s", "-j", jobs, "-C", os.path.join(HERE, "sanitize"),
                    f"OUT={out}"], check=True, capture_output=True, text=True, timeout=900)
    return str(out)


def _asan_env(san_build):
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                          text=True, check=True).stdout.strip()
    env = dict(os.environ)
    env.update(LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0",
               SFFT_ORACLE_LIBRARY=os.path.join(san_build, "liboracle.so"))
    return env


def test_host_checks_under_sanitizers(san_build):
    r = subprocess.run([os.path.join(san_build, "host_checks")], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "host_checks: ok" in r.stdout


def test_oracle_tests_under_sanitizers(san_build):
    env = _asan_env(san_build)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider"] +
                       [os.path.join(HERE, f) for f in ORACLE_TESTS],
                       capture_output=True, text=True, env=env, timeout=1200)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert " passed" in r.stdout


def test_instrumentation_is_live(san_build):
    """A read past the end of a numpy buffer inside the instrumented oracle must abort."""
    code = (
        "import os, sys, ctypes as C, numpy as np\n"
        "sys.path.insert(0, os.path.join(%r, '..', 'oracle'))\n"
        "import oracle as O\n"
        "assert O._LIB_PATH == os.environ['SFFT_ORACLE_LIBRARY']\n"
        "X = np.zeros(64, dtype=np.complex128); s = np.zeros(128)\n"
        "O.lib().orc_samples64(X.ctypes.data_as(C.c_void_p), 128, s.ctypes.data_as(C.c_void_p))\n"
        "print('not caught')\n" % HERE)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       env=_asan_env(san_build), timeout=300)
    assert r.returncode != 0 and "not caught" not in r.stdout
    assert "heap-buffer-overflow" in r.stderr, r.stderr[-3000:]
