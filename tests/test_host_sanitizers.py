"""The host C++ the drop-in runs before any device work (equation parser,
evaluator, postfix lowering, grid and twiddle builders) under AddressSanitizer
and UndefinedBehaviorSanitizer, built with g++ (no HIP code involved)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "gpue_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_code_is_sanitizer_clean(tmp_path):
    exe = str(tmp_path / "sanitize_main")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-I" + os.path.join(os.path.dirname(HERE), "include"),
           os.path.join(HERE, "host", "sanitize_main.cpp"), os.path.join(CSRC, "eqn.cpp"),
           os.path.join(CSRC, "host_util.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
    assert r.returncode == 0, r.stderr[-2000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-2000:]
    assert r.stdout.startswith("parsed ")
