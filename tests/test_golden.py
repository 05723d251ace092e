"""Golden fixtures (tests/golden/) and structural checks of create output.

CPU tests: the oracle still reproduces every committed fixture, and its
output sets pass the independent structural checks of par2_check.py.  The
GPU side compares the product against the same fixtures in
test_gpu_create.py.
"""
import hashlib
import json
import os

from oracle import oracle as O
from tests.par2_check import check_set

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


def test_config0_manifest(tmp_path):
    m = load("config0.json")
    data = O.xorshift64_bytes(m["input"]["bytes"], m["input"]["seed"])
    src = tmp_path / "input.bin"
    src.write_bytes(data)
    files = O.run_create(str(tmp_path / "recovery.par2"), m["slice_size"], m["recovery"],
                         [(str(src), b"input.bin")], O.SCHEME_EXPONENTIAL)
    got = [{"name": os.path.basename(p), "size": os.path.getsize(p),
            "sha256": hashlib.sha256(open(p, "rb").read()).hexdigest()} for p in files]
    assert got == m["files"]
    blobs = {os.path.basename(p): open(p, "rb").read() for p in files}
    check_set(blobs, [(b"input.bin", data)], m["slice_size"], m["recovery"])


def test_tiny_case(tmp_path):
    t = load("tiny.json")
    src = tmp_path / "a.bin"
    src.write_bytes(t["input"].encode())
    files = O.run_create(str(tmp_path / "recovery.par2"), 4, 2, [(str(src), b"a.bin")],
                         O.SCHEME_SINGLE)
    assert [os.path.basename(p) for p in files] == [f["name"] for f in t["files"]]
    for p, f in zip(files, t["files"]):
        assert open(p, "rb").read().hex() == f["hex"]
    blobs = {f["name"]: bytes.fromhex(f["hex"]) for f in t["files"]}
    check_set(blobs, [(b"a.bin", t["input"].encode())], 4, 2)


def test_vectors():
    v = load("vectors.json")
    for a, b, p in v["gf_mul"]:
        assert O.gf_mul(a, b) == p
    assert O.input_bases(8) == v["bases_first_8"]
    e = v["encode"]
    files = [bytes.fromhex(x) for x in e["files_hex"]]
    assert O.encode(files, e["slice_size"], e["recovery"]).hex() == e["recovery_hex"]
    c = v["slice_checksums"]
    md5s, crcs = O.slice_checksums(c["data"].encode(), c["slice_size"])
    assert [m.hex() for m in md5s] == c["md5"] and crcs == c["crc32"]
    p = v["packet"]
    pkt = O.recovery_packet(bytes.fromhex(p["set_id"]), p["exponent"], bytes.fromhex(p["data_hex"]))
    assert pkt[16:32].hex() == p["md5"]


def test_multi_file_set_structure(tmp_path):
    srcs = []
    for i, n in enumerate([5000, 4096, 1, 12289]):
        d = bytes(n) if i == 1 else O.xorshift64_bytes(n, 77 + i)
        (tmp_path / f"f{i}").write_bytes(d)
        srcs.append((f"name{i}.dat".encode(), d))
    files = O.run_create(str(tmp_path / "set.par2"), 1024, 9,
                         [(str(tmp_path / f"f{i}"), nm) for i, (nm, _) in enumerate(srcs)],
                         O.SCHEME_UNIFORM, uniform_count=4)
    blobs = {os.path.basename(p): open(p, "rb").read() for p in files}
    assert list(blobs)[1:] == ["set.vol0+2.par2", "set.vol2+2.par2", "set.vol4+2.par2",
                               "set.vol6+3.par2"]
    check_set(blobs, srcs, 1024, 9)
