"""AddressSanitizer / UndefinedBehaviorSanitizer runs on the CPU (no GPU).

* the FP64 oracle (oracle/joseph_oracle.c) driven by tests/sanitize/oracle_checks.c:
  FP / BP in SET and ADD mode on parallel and cone geometries with rays along
  every axis, a tilted detector and detector supersampling, checking the adjoint
  identity;
* the host half of libtsx (tomosipo_amd/csrc/tsx.hip, sanitizers on the host
  pass only; the device code is built as usual and never run) driven by
  tests/sanitize/host_checks.cpp: geometry validation and its error paths, host
  plans of every kernel kind, plan queries, workspace sizes, the argument checks
  of every compute entry point, plan destruction (LeakSanitizer).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.join(ROOT, "tests", "sanitize")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
SAN = ["-fsanitize=address", "-fsanitize=undefined", "-fno-omit-frame-pointer"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def _run(exe):
    res = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=300)
    out = res.stdout + res.stderr
    assert res.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "ERROR: AddressSanitizer" not in out, out[-4000:]
    return out


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "oracle_checks")
    cmd = ["gcc", "-O1", "-g", *SAN, "-o", exe, os.path.join(HERE, "oracle_checks.c"),
           os.path.join(ROOT, "oracle", "joseph_oracle.c"), "-lm"]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    out = _run(exe)
    assert "oracle checks passed" in out


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_library_host_code_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "host_checks")
    # the sanitizers instrument the host pass only (each flag behind -Xarch_host)
    host_san = [x for f in SAN for x in ("-Xarch_host", f)]
    cmd = [HIPCC, "-x", "hip", "--offload-arch=gfx950", "-std=c++17", "-O1", "-g",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", *host_san,
           "-I", os.path.join(ROOT, "include"), "-o", exe, os.path.join(HERE, "host_checks.cpp"),
           os.path.join(ROOT, "tomosipo_amd", "csrc", "tsx.hip"),
           f"-L{ROCM}/lib", "-lrocprofiler-sdk-roctx", f"-Wl,-rpath,{ROCM}/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stderr[-4000:]
    out = _run(exe)
    assert "host checks passed" in out
