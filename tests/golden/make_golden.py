"""Writes the golden fixtures of tests/golden/ from the CPU oracle.

    python tests/golden/make_golden.py

config0.json  SHA-256 and size of every file of the config-0 create: one
              4 MiB file of xorshift64 bytes (seed 1) named "input.bin",
              slice size 4096 (1024 slices), 10 recovery blocks, exponential
              volume split.
tiny.json     the complete bytes (hex) of the create of src/creator.rs:1210-
              1233: "hello par2 world!!!!" as a.bin, slice size 4, 2 recovery
              blocks, single volume.
vectors.json  small known answers for the kernels: GF products, RS
              coefficients, one 64-byte encode, slice checksums and a packet
              hash.
config2.json  the many-files create at full size: the 64 files of
              par2rust_amd.synth.config2_files (304 MB; 4 zero-length, whose
              create fails with EmptyFile, and 8 all-zero), slice size 65536,
              1000 recovery blocks, uniform split into 15 volumes: SHA-256
              and size of every file the create of the 60 other files writes,
              and of recovery blocks 0, 1, 2, 499, 998 and 999 alone (the
              scalar encode of those exponents, a check of the set's SIMD
              encode).
config1.json  the headline create at full size: one 538,218,411-byte file of
              xorshift64 bytes (seed 1), slice size 524288 (1027 slices),
              200 recovery blocks, exponential split: SHA-256 and size of
              every file written.

    python tests/golden/make_golden.py [config1|config2]   # only that file

The oracle is pinned independently (tests/test_oracle.py); these files freeze
its output so that a GPU box can check the product without rebuilding it and
any later change of either side shows up as a diff.
"""
import hashlib
import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def create(tmp, data, name, S, R, kind):
    src = os.path.join(tmp, name)
    with open(src, "wb") as f:
        f.write(data)
    out = os.path.join(tmp, "recovery.par2")
    return O.run_create(out, S, R, [(src, name.encode())], kind)


CONFIG2_ROWS = [0, 1, 2, 499, 998, 999]


def config2():
    from par2rust_amd import synth

    entries = synth.config2_files()
    S, R, vols = 65536, 1000, 15
    with tempfile.TemporaryDirectory() as tmp:
        inputs, datas, err = [], [], None
        for e in entries:
            path = os.path.join(tmp, e[0])
            with open(path, "wb") as f:
                f.write(synth.config2_bytes(e))
            inputs.append((path, e[0].encode()))
        try:
            O.run_create(os.path.join(tmp, "all.par2"), S, 1, inputs, O.SCHEME_SINGLE)
        except O.OracleError as x:
            err = x.code
        kept = [(p, nm) for (p, nm), e in zip(inputs, entries) if e[2] != "empty"]
        datas = [open(p, "rb").read() for p, _ in kept]
        out = O.run_create(os.path.join(tmp, "recovery.par2"), S, R, kept, O.SCHEME_UNIFORM, vols)
        manifest = {
            "files": [{"name": e[0], "size": e[1], "kind": e[2], "seed": e[3]} for e in entries],
            "slice_size": S, "recovery": R, "scheme": f"uniform {vols}", "uniform_count": vols,
            "all_files_error_code": err,
            "slices": sum((len(d) + S - 1) // S for d in datas),
            "outputs": [{"name": os.path.basename(p), "size": os.path.getsize(p),
                         "sha256": hashlib.sha256(open(p, "rb").read()).hexdigest()}
                        for p in out],
            "rows": {str(r): hashlib.sha256(O.encode_rows(datas, S, r, 1)).hexdigest()
                     for r in CONFIG2_ROWS},
        }
    with open(os.path.join(HERE, "config2.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    print("wrote config2.json")


def config1():
    from par2rust_amd import synth

    nbytes, S, R = 538_218_411, 524288, 200
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "input.bin")
        with open(path, "wb") as f:
            f.write(synth.xorshift64_bytes(nbytes, 1))
        out = O.run_create(os.path.join(tmp, "recovery.par2"), S, R, [(path, b"input.bin")],
                           O.SCHEME_EXPONENTIAL)
        manifest = {
            "input": {"generator": "xorshift64", "seed": 1, "bytes": nbytes, "name": "input.bin"},
            "slice_size": S, "recovery": R, "scheme": "exponential",
            "outputs": [{"name": os.path.basename(p), "size": os.path.getsize(p),
                         "sha256": hashlib.sha256(open(p, "rb").read()).hexdigest()}
                        for p in out],
        }
    with open(os.path.join(HERE, "config1.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    print("wrote config1.json")


def main():
    only = {"config1": config1, "config2": config2}
    if sys.argv[1:] and sys.argv[1] in only:
        only[sys.argv[1]]()
        return
    config1()
    config2()
    with tempfile.TemporaryDirectory() as tmp:
        files = create(tmp, O.xorshift64_bytes(4 << 20, 1), "input.bin", 4096, 10,
                       O.SCHEME_EXPONENTIAL)
        manifest = {
            "input": {"generator": "xorshift64", "seed": 1, "bytes": 4 << 20,
                      "name": "input.bin"},
            "slice_size": 4096, "recovery": 10, "scheme": "exponential",
            "files": [{"name": os.path.basename(p), "size": os.path.getsize(p),
                       "sha256": hashlib.sha256(open(p, "rb").read()).hexdigest()}
                      for p in files],
        }
    with open(os.path.join(HERE, "config0.json"), "w") as f:
        json.dump(manifest, f, indent=1)

    with tempfile.TemporaryDirectory() as tmp:
        files = create(tmp, b"hello par2 world!!!!", "a.bin", 4, 2, O.SCHEME_SINGLE)
        tiny = {"input": "hello par2 world!!!!", "name": "a.bin", "slice_size": 4,
                "recovery": 2, "scheme": "single",
                "files": [{"name": os.path.basename(p), "hex": open(p, "rb").read().hex()}
                          for p in files]}
    with open(os.path.join(HERE, "tiny.json"), "w") as f:
        json.dump(tiny, f, indent=1)

    enc_in = [O.xorshift64_bytes(150, 7), O.xorshift64_bytes(70, 8)]
    vec = {
        "gf_mul": [[a, b, O.gf_mul(a, b)] for a, b in
                   [(2, 0x8000), (0x1234, 0x5678), (0xFFFF, 0xFFFF), (3, 7), (0x100B, 2)]],
        "bases_first_8": O.input_bases(8),
        "coef_r3_i0_7": [O.gf_pow(b, 3) for b in O.input_bases(8)],
        "encode": {"files_hex": [x.hex() for x in enc_in], "slice_size": 64, "recovery": 3,
                   "recovery_hex": O.encode(enc_in, 64, 3).hex()},
        "slice_checksums": {"data": "ABCDE", "slice_size": 4,
                            "md5": [m.hex() for m in O.slice_checksums(b"ABCDE", 4)[0]],
                            "crc32": O.slice_checksums(b"ABCDE", 4)[1]},
        "packet": {"set_id": "00112233445566778899aabbccddeeff", "exponent": 5,
                   "data_hex": bytes(range(64)).hex(),
                   "md5": O.recovery_packet(bytes.fromhex("00112233445566778899aabbccddeeff"),
                                            5, bytes(range(64)))[16:32].hex()},
    }
    with open(os.path.join(HERE, "vectors.json"), "w") as f:
        json.dump(vec, f, indent=1)
    print("wrote config0.json, tiny.json, vectors.json")


if __name__ == "__main__":
    main()
