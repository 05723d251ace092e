"""The product's host MD5 (par2rust_amd/csrc/host_md5.hpp: hash_full, hash16k,
file ids, critical packets) against RFC 1321 vectors and hashlib, built
with g++ into a temporary harness.  CPU only."""
import hashlib
import os
import random
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("md5") / "host_md5_check")
    subprocess.run([cxx, "-O2", "-std=c++17", os.path.join(HERE, "host_md5_check.cpp"), "-o", exe],
                   check=True)
    return exe


def run(exe, data, piece):
    return subprocess.run([exe, str(piece)], input=data, capture_output=True,
                          check=True).stdout.decode().strip()


def test_rfc1321_vectors(harness):
    from tests.test_oracle import RFC1321

    for msg, hexd in RFC1321.items():
        assert run(harness, msg, 4096) == hexd


def test_against_hashlib_with_split_updates(harness):
    rnd = random.Random(9)
    for n in [0, 1, 55, 56, 57, 63, 64, 65, 119, 120, 1000, 16384, 100003]:
        d = bytes(rnd.randrange(256) for _ in range(n))
        for piece in (1, 3, 64, 100, 65536):
            assert run(harness, d, piece) == hashlib.md5(d).hexdigest(), (n, piece)
