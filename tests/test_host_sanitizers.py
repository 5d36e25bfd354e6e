"""The host file-format layer (pm_io.cpp) under AddressSanitizer and
UndefinedBehaviorSanitizer: tests/sanitize/host_io_check.cpp drives the .inp
parser, the raw and log writers, the raw reader and the rank map over the
reference's input files, malformed variants and edge sizes.  g++ only (no HIP
code in this build)."""
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_io_under_asan_and_ubsan(tmp_path):
    exe = tmp_path / "host_io_check"
    cmd = ["g++", "-std=c++17", "-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", os.path.join(ROOT, "tests", "sanitize", "host_io_check.cpp"),
           os.path.join(ROOT, "poremaps_amd", "csrc", "pm_io.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    scratch = tmp_path / "run"
    scratch.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(exe), GOLDEN, str(scratch)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "host io checks: ok" in r.stdout
