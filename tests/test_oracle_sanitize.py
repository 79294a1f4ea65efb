"""The CPU oracle (oracle/clim_oracle.c, scene_oracle.c) built with
AddressSanitizer and UndefinedBehaviorSanitizer (`make -C oracle sanitize`)
runs the whole path on small frames of every CFA layout and level filter; a
report from either sanitizer fails the test.  Host code only."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _gcc_lib(name):
    r = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True)
    p = r.stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_oracle_under_asan_ubsan():
    asan, ubsan = _gcc_lib("libasan.so"), _gcc_lib("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("gcc sanitizer runtimes not available")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    env = dict(os.environ, LD_PRELOAD=f"{asan}:{ubsan}", ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2",
               CLIM_ORACLE_LIB=os.path.join(ROOT, "oracle", "_build", "liboracle_san.so"))
    r = subprocess.run([sys.executable, os.path.join(HERE, "sanitize", "oracle_run.py")], capture_output=True,
                       text=True, env=env, timeout=600)
    assert r.returncode == 0 and "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-4000:]
    assert "ok" in r.stdout
