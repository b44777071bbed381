"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

The C host programs (hackforce, hackcode1, csf and the snapshot codec) and
the C oracle are rebuilt with -fsanitize=address,undefined (nemo_amd/san.mk,
oracle/san.mk) and the CPU tests that exercise them run again on those
builds: codec round trips of the reference's own snapshots, truncated and
garbage files, keyword errors, typed and coerced reads, and every oracle
test (tree build, buckets, walk, leapfrog, treecode1).  Any report aborts
the run.  This is the counterpart of the reference's valgrind runs of its
test suite (src/scripts/testsuite:62-68)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SANFLAGS = "-fsanitize=address,undefined -fno-sanitize-recover=undefined -fno-omit-frame-pointer -g"


def _libasan():
    p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.fixture(scope="module")
def san_env():
    if _libasan() is None:
        pytest.skip("gcc's libasan is not installed")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "nemo_amd")], check=True)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "nemo_amd"), "-f", "san.mk"], check=True)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-f", "san.mk"], check=True)
    env = dict(os.environ)
    env.update(ASAN_OPTIONS="abort_on_error=0:halt_on_error=1:exitcode=99",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1:exitcode=99")
    return env


def _pytest(env, *args):
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu"] +
                          list(args), cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)


def test_cli_and_codec_under_sanitizers(san_env):
    env = dict(san_env, NEMO_BIN=os.path.join(ROOT, "nemo_amd", "bin-san"), NEMO_SAN_CFLAGS=SANFLAGS,
               ASAN_OPTIONS=san_env["ASAN_OPTIONS"] + ":detect_leaks=1")
    r = _pytest(env, "tests/test_cli.py", "-k", "not plugin")
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "passed" in r.stdout and "ERROR: AddressSanitizer" not in r.stdout + r.stderr
    # the acceleration(5) plugin is loaded into the (uninstrumented) Python
    # process: the runtime is preloaded and leaks are not checked there
    env.update(LD_PRELOAD=_libasan(), ASAN_OPTIONS=san_env["ASAN_OPTIONS"] + ":detect_leaks=0")
    r = _pytest(env, "tests/test_cli.py", "-k", "plugin")
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "passed" in r.stdout


def test_oracle_under_sanitizers(san_env):
    # the Python interpreter itself is not instrumented: the runtime is
    # preloaded, and leak checking is left to the C programs above
    env = dict(san_env, NEMO_ORACLE_LIB=os.path.join(ROOT, "oracle", "_san", "liboracle.so"), LD_PRELOAD=_libasan(),
               ASAN_OPTIONS=san_env["ASAN_OPTIONS"] + ":detect_leaks=0")
    r = _pytest(env, "tests/test_oracle.py", "tests/test_tc_oracle.py", "tests/test_mlp_oracle.py",
                "tests/test_lc_oracle.py", "tests/test_le_oracle.py")
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "passed" in r.stdout and "runtime error" not in r.stderr
