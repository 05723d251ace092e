"""Full-size creates end to end on the GPU engine (p2g_run_create), every
written file byte-identical to the oracle's (committed SHA-256 manifests),
and each set passing the independent structure check:

* the headline: one 538,218,411-byte file, slice size 524288 (1027 slices),
  200 recovery blocks, exponential split;
* many files: 64 files (sizes log-uniform in [1 KiB, 64 MiB], 4 zero-length,
  8 all-zero), slice size 65536, 1000 recovery blocks, uniform split into 15
  volumes.  With the zero-length files the create fails with EmptyFile, as
  the reference's SourceFile::scan does (src/source.rs:73-75).

The serving workload's 32 GiB job is checked on the device engine alone (no
files): served and alone equal, every slice checksum against the oracle, and
the recovery blocks through properties that hold at any size."""
import hashlib
import json
import os
import struct
import time

import pytest

from oracle import oracle as O

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


def test_serving_large_job_full_size(ctx):
    """The serving workload's large job at its full size: one 32 GiB file
    (xorshift64, seed 4000, written on the device by the server), slice size
    1 MiB (32768 slices, the PAR2 maximum), 3000 recovery blocks.  The server
    runs it across 10 ms steps: each 16-row block (~27 ms of device time) in
    three pieces of its column groups, the packet hashes carrying their MD5
    state from piece to piece; p2g_dev_create_pass then runs it alone on the
    same device bytes (a one-job batched launch).  Checks:
    * both give the same recovery blocks, slice checksums and packet MD5s;
    * every slice MD5 and CRC-32 equals the oracle's, the source streamed to
      the host in 1 GiB chunks;
    * recovery block 0 equals the XOR of all 32768 slices
      (src/reedsolomon.rs:261-276: exponent 0 makes every coefficient 1);
    * three 16-byte column windows of all 3000 blocks equal the oracle's
      encode of the same windows of every slice (the multiply is
      position-wise, so a window of the inputs encodes to that window of the
      outputs);
    * every packet MD5 equals the oracle's over the downloaded blocks."""
    import ctypes

    import numpy as np

    import par2rust_amd as p2
    from par2rust_amd import hipmem

    S, R, N = 1 << 20, 3000, 32768
    size = N * S
    sid = bytes(range(100, 116))
    srv = ctx.server(48 << 30, step_ms=10.0)
    try:
        rid = srv.submit([size], S, R, sid, seeds=[4000])
        t0, done = time.time(), []
        while not done:
            done = srv.poll(1000.0)
            assert time.time() - t0 < 300, "the server did not finish the job"
        assert done == [rid]
        job, _ = srv.result(rid)
        df = ctypes.cast(job.files, ctypes.POINTER(p2._DevFile))
        assert job.nfiles == 1 and df[0].len == size
        src = df[0].data
        served = srv.download(rid)
        outs = [hipmem.DeviceBuffer(R * S), hipmem.DeviceBuffer(16 * N), hipmem.DeviceBuffer(4 * N),
                hipmem.DeviceBuffer(16 * R)]
        ctx.dev_create_pass([(src, size)], S, R, sid, *outs)
        ctx.synchronize()
        alone = (outs[0].download(), outs[1].download(),
                 list(struct.unpack(f"<{N}I", outs[2].download())), outs[3].download())
        del outs
        assert served[0] == alone[0], "recovery blocks: served != alone"
        assert served[1:] == alone[1:], "slice checksums or packet MD5s: served != alone"
        del alone
        rec, md5s, crcs, pkt = served

        offs = np.concatenate([np.arange(o, o + 16) for o in (0, 524288 + 496, S - 16)])
        row0 = np.zeros(S // 8, dtype=np.uint64)
        windows = []
        chunk = 1024  # slices per 1 GiB chunk
        for c0 in range(0, N, chunk):
            buf = hipmem.download_ptr(src + c0 * S, chunk * S)
            m, cr = O.slice_checksums_mt(buf, S, 16)
            assert b"".join(m) == md5s[16 * c0:16 * (c0 + chunk)], f"slice MD5 in slices {c0}+"
            assert cr == crcs[c0:c0 + chunk], f"slice CRC-32 in slices {c0}+"
            row0 ^= np.bitwise_xor.reduce(np.frombuffer(buf, dtype=np.uint64).reshape(chunk, S // 8),
                                          axis=0)
            windows.append(np.frombuffer(buf, dtype=np.uint8).reshape(chunk, S)[:, offs].copy())
            del buf
        recv = np.frombuffer(rec, dtype=np.uint8).reshape(R, S)
        assert row0.tobytes() == rec[:S], "recovery block 0 != XOR of the slices"
        want = O.encode([np.concatenate(windows).tobytes()], len(offs), R)
        assert recv[:, offs].tobytes() == want, "column windows != oracle"
        assert b"".join(O.recovery_md5_mt(rec, R, S, 0, sid, 16)) == pkt, "packet MD5s != oracle"
        srv.release(rid)
    finally:
        srv.close()
