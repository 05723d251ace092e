"""Full-size creates end to end on the GPU engine (p2g_run_create), every
written file byte-identical to the oracle's (committed SHA-256 manifests),
and each set passing the independent structure check:

* the headline: one 538,218,411-byte file, slice size 524288 (1027 slices),
  200 recovery blocks, exponential split;
* many files: 64 files (sizes log-uniform in [1 KiB, 64 MiB], 4 zero-length,
  8 all-zero), slice size 65536, 1000 recovery blocks, uniform split into 15
  volumes.  With the zero-length files the create fails with EmptyFile, as
  the reference's SourceFile::scan does (src/source.rs:73-75)."""
import hashlib
import json
import os

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_config2_full_size(tmp_path, ctx):
    import par2rust_amd as p2
    from par2rust_amd import synth
    from tests.par2_check import check_set

    m = json.load(open(os.path.join(HERE, "golden", "config2.json")))
    entries = synth.config2_files()
    assert [{"name": e[0], "size": e[1], "kind": e[2], "seed": e[3]} for e in entries] == m["files"]
    inputs, sources = [], []
    for e in entries:
        path = tmp_path / e[0]
        data = synth.config2_bytes(e)
        path.write_bytes(data)
        inputs.append((str(path), e[0].encode()))
        if e[2] != "empty":
            sources.append((e[0].encode(), data))
    S, R = m["slice_size"], m["recovery"]
    scheme = p2.VolumeScheme.uniform(m["uniform_count"])
    out_dir = tmp_path / "out"
    out_dir.mkdir()
    with pytest.raises(p2.Par2Error) as err:
        ctx.run_create(p2.CreateOptions(str(out_dir / "all.par2"), S, R, scheme), inputs)
    assert err.value.kind == "EmptyFile" and err.value.code == m["all_files_error_code"]
    kept = [(p, nm) for (p, nm), e in zip(inputs, entries) if e[2] != "empty"]
    written = ctx.run_create(p2.CreateOptions(str(out_dir / "recovery.par2"), S, R, scheme), kept)
    got = [{"name": os.path.basename(p), "size": os.path.getsize(p),
            "sha256": hashlib.sha256(open(p, "rb").read()).hexdigest()} for p in written]
    assert got == m["outputs"]
    seen = check_set({os.path.basename(p): open(p, "rb").read() for p in written}, sources, S, R)
    for r, h in m["rows"].items():
        assert hashlib.sha256(seen[int(r)]).hexdigest() == h, r


def test_config1_full_size(tmp_path, ctx):
    import par2rust_amd as p2
    from par2rust_amd import synth
    from tests.par2_check import check_set

    m = json.load(open(os.path.join(HERE, "golden", "config1.json")))
    data = synth.xorshift64_bytes(m["input"]["bytes"], m["input"]["seed"])
    src = tmp_path / "input.bin"
    src.write_bytes(data)
    out_dir = tmp_path / "out"
    out_dir.mkdir()
    written = ctx.run_create(p2.CreateOptions(str(out_dir / "recovery.par2"), m["slice_size"],
                                              m["recovery"], p2.VolumeScheme.exponential()),
                             [(str(src), b"input.bin")])
    got = [{"name": os.path.basename(p), "size": os.path.getsize(p),
            "sha256": hashlib.sha256(open(p, "rb").read()).hexdigest()} for p in written]
    assert got == m["outputs"]
    check_set({os.path.basename(p): open(p, "rb").read() for p in written}, [(b"input.bin", data)],
              m["slice_size"], m["recovery"])
