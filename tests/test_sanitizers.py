"""CPU sanitizer runs (address + undefined behaviour) of the FP64 oracle that
every parity test trusts, and of the host build of the device linear
algebra header.  Host code only: no GPU code is built with a sanitizer."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]


def _env():
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1"
    # the sanitizer runtime must come first in the process: a preloaded
    # library of the environment would be ahead of it
    env.pop("LD_PRELOAD", None)
    return env


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no gcc")
def test_oracle_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    r = subprocess.run([os.path.join(ROOT, "oracle", "_asan", "oracle_asan")], capture_output=True, text=True,
                       env=_env(), timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_linalg_header_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "linalg_asan")
    host = os.path.join(ROOT, "tests", "host")
    subprocess.run(["g++", *SAN, "-std=c++17", "-o", exe, os.path.join(host, "linalg_harness.cpp"),
                    os.path.join(host, "linalg_asan_main.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, env=_env(), timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok" in r.stdout
