"""The host code of the library (case setup: units, face codes, initial populations,
neiSolid, outlet sponges, findPoints and the near-wall distances; output: VTK in its three
forms, restart write/read and its refusals, run statistics, FP64 coefficients, running
means) built with AddressSanitizer and UndefinedBehaviorSanitizer and driven through every
entry point by host_sanitize_driver.cpp, on the CPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOURCES = [os.path.join(ROOT, "tests", "host_sanitize_driver.cpp"),
           os.path.join(ROOT, "lbffs_amd", "csrc", "lbffs_case.cpp"),
           os.path.join(ROOT, "lbffs_amd", "csrc", "lbffs_output.cpp")]


def test_host_code_under_asan_and_ubsan(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = str(tmp_path / "driver")
    build = subprocess.run([gxx, "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                            "-fno-sanitize-recover=all", "-I" + os.path.join(ROOT, "include"),
                            "-I" + os.path.join(ROOT, "lbffs_amd", "csrc"), *SOURCES, "-o", exe],
                           capture_output=True, text=True, timeout=300)
    if build.returncode != 0 and "asan" in (build.stderr or "").lower():
        pytest.skip("the sanitizer runtimes are not installed: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr[-3000:]
    # verify_asan_link_order=0: the sanitizer runtime need not come first among the loaded
    # libraries (a library preloaded into every process would otherwise stop the run)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    run = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-4000:])
    assert "0 failed expectations" in run.stdout
    assert "runtime error" not in run.stderr
