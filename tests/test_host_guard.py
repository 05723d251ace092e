"""The C ABI's exception barrier (p2g::guarded in
par2rust_amd/csrc/p2g_context.hpp), built with g++ into a temporary harness:
a body's own status passes through, a failed host allocation becomes
P2G_ERR_OUT_OF_MEMORY and any other C++ exception P2G_ERR_IO, so none
crosses the C boundary.  CPU only."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_exceptions_become_status_codes(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("g++ or the ROCm headers not available")
    exe = str(tmp_path / "guard_check")
    subprocess.run([cxx, "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "par2rust_amd", "csrc"),
                    os.path.join(HERE, "guard_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, check=True).stdout.decode().split()
    from par2rust_amd import ERRORS

    assert [ERRORS.get(int(x), x) for x in out[1:]] == ["OutOfMemory", "Io", "Io"]
    assert out[0] == "7"
