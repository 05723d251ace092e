"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

* The oracle (oracle/par2_oracle*.c) built with -fsanitize=address,undefined
  (``make -C oracle asan``) runs its whole pin suite (tests/test_oracle.py,
  tests/test_golden.py) in a subprocess with the sanitizer runtimes
  preloaded; a canary read past a buffer first shows that the instrumented
  build is the one loaded and that a fault is reported.
* The product's host-only planners (par2rust_amd/csrc/p2g_plan.cpp) and
  volume split (par2_scheme.cpp) are built with g++ and the same sanitizers
  into tests/host_plan_check.cpp's harness, which sweeps shapes and schemes.
"""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ASAN_ORACLE = os.path.join(ROOT, "oracle", "_asan", "libpar2_oracle.so")


def _runtimes():
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    libs = []
    for name in ("libasan.so", "libubsan.so"):
        p = subprocess.run([gcc, f"-print-file-name={name}"], capture_output=True,
                           text=True).stdout.strip()
        if not os.path.isabs(p) or not os.path.exists(p):
            pytest.skip(f"{name} not available")
        libs.append(p)
    return libs


def _env():
    env = dict(os.environ)
    env["LD_PRELOAD"] = " ".join(_runtimes())
    env["ASAN_OPTIONS"] = "detect_leaks=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1"
    env["PAR2_ORACLE_LIB"] = ASAN_ORACLE
    return env


@pytest.fixture(scope="module")
def asan_oracle():
    _runtimes()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    return ASAN_ORACLE


CANARY = """
import ctypes
from oracle import oracle as O
buf = ctypes.create_string_buffer(1 << 20)
O.L.or_crc32(buf, (1 << 20) + 4096)   # reads past the end of the buffer
print("not detected")
"""


def test_canary_is_detected(asan_oracle):
    r = subprocess.run([sys.executable, "-c", CANARY], cwd=ROOT, env=_env(), capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "AddressSanitizer" in r.stderr, r.stderr[-2000:]
    assert "not detected" not in r.stdout


def test_oracle_pins_under_sanitizers(asan_oracle):
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(HERE, "test_oracle.py"), os.path.join(HERE, "test_golden.py")],
                       cwd=ROOT, env=_env(), capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-3000:]
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-3000:]


def test_planners_and_schemes_under_sanitizers(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("g++ or the ROCm headers not available")
    csrc = os.path.join(ROOT, "par2rust_amd", "csrc")
    exe = str(tmp_path / "host_plan_check")
    subprocess.run([cxx, "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
                    "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I" + os.path.join(ROOT, "include"), "-I" + csrc,
                    os.path.join(HERE, "host_plan_check.cpp"), os.path.join(csrc, "p2g_plan.cpp"),
                    os.path.join(csrc, "par2_scheme.cpp"), "-o", exe], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    assert r.stdout.startswith("ok ")
