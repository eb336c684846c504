"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only, every report fatal).

* The oracle restatement (oracle/hydro_oracle.cpp, `make -C oracle sanitize`) runs the oracle's own test files
  (closed forms, golden fixtures, Flu_Close / fix-up / interpolation pins, the two-level composition) with GCC's
  ASan runtime preloaded.
* The C-ABI library's host code (argument checks, error paths, buffer bookkeeping;
  gamer_amd.build.build_host_sanitized) runs tests/test_abi.py with the ROCm compiler's ASan runtime preloaded.
  Without a GPU those tests stop at the argument checks, which is what they exercise.
"""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "_build")
ORACLE_TESTS = ["tests/test_oracle_props.py", "tests/test_oracle_golden.py", "tests/test_oracle_sync.py",
                "tests/test_oracle_interp.py", "tests/test_amr_two_level.py"]


def _run(env_extra, args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="1", **env_extra)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu", *args],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=1200)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    return out


def test_oracle_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize", f"SAN_OUT={BUILD}"], check=True)
    rt = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(rt):
        pytest.skip("GCC's ASan runtime is not installed")
    out = _run({"LD_PRELOAD": rt, "ORACLE_LIB": os.path.join(BUILD, "liboracle_san.so")}, ORACLE_TESTS)
    assert " passed" in out


def test_c_abi_host_code_under_asan_ubsan():
    rts = glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
    if not rts:
        pytest.skip("the ROCm compiler's ASan runtime is not installed")
    sys.path.insert(0, ROOT)
    from gamer_amd import build
    lib = build.build_host_sanitized(os.path.join(BUILD, "san"))
    # the C99 link test links a program against the library; with the sanitizer runtime that needs the
    # compiler driver's link line, so it runs only against the plain build
    out = _run({"LD_PRELOAD": rts[0], "GAMER_AMD_HYDRO_LIB": lib}, ["tests/test_abi.py", "-k", "not c99_and_links"])
    assert " passed" in out
