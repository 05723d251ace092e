"""GPU tests of the autotuner: a tuned choice is cached, saved and loaded;
online tuning (P2G_AUTOTUNE=1) returns the same bytes as the oracle on the
tuning call and after; a cached bit-sliced or narrow-band choice drives the
plain pass and the batched planner without changing any output."""
import os
import struct

import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def data(n: int, seed: int) -> bytes:
    return O.xorshift64_bytes(n, seed)


def expected(files, S, R, set_id):
    rec = O.encode(files, S, R)
    md5s, crcs = [], []
    for f in files:
        m, cr = O.slice_checksums(f, S)
        md5s += m
        crcs += cr
    pkt = b"".join(O.recovery_packet(set_id, r, rec[r * S:(r + 1) * S])[16:32] for r in range(R))
    return rec, b"".join(md5s), crcs, pkt


def run_pass(c, files, S, R, sid):
    from par2rust_amd.hipmem import DeviceBuffer

    dev = [DeviceBuffer.from_bytes(f) for f in files]
    n = sum((len(f) + S - 1) // S for f in files)
    rec, md5, crc, pkt = (DeviceBuffer(R * S), DeviceBuffer(16 * n), DeviceBuffer(4 * n),
                          DeviceBuffer(16 * R))
    c.dev_create_pass([(d.addr, len(f)) for d, f in zip(dev, files)], S, R, sid, rec, md5, crc, pkt)
    c.synchronize()
    return (rec.download(), md5.download(), list(struct.unpack(f"<{n}I", crc.download())),
            pkt.download())


@pytest.fixture
def env_ctx():
    import par2rust_amd as p2

    made = []

    def open_with(**env):
        old = {k: os.environ.get(k) for k in env}
        os.environ.update({k: str(v) for k, v in env.items()})
        try:
            c = p2.Context(0)
        finally:
            for k, v in old.items():
                if v is None:
                    del os.environ[k]
                else:
                    os.environ[k] = v
        made.append(c)
        return c

    yield open_with
    for c in made:
        c.close()


def test_autotune_caches_saves_loads(env_ctx, tmp_path):
    c = env_ctx()
    got = c.autotune(600, 40, 16384)
    assert got["variant"] in (0, 1, 2) and got["ms"] > 0
    enc = c.tune_get_encode(600, 40, 16384)
    assert enc["variant"] in (0, 1, 2) and enc["ms"] > 0
    assert c.tune_get(600, 40, 16384) == {"variant": got["variant"], "band_cols": got["band_cols"]}
    # the same bucket: shapes within the same bit lengths
    assert c.tune_get(1000, 60, 32000) == c.tune_get(600, 40, 16384)
    path = str(tmp_path / "tune.txt")
    c.tune_save(path)
    d = env_ctx()
    d.tune_load(path)
    assert d.tune_get(600, 40, 16384) == c.tune_get(600, 40, 16384)
    a, b = d.tune_get_encode(600, 40, 16384), c.tune_get_encode(600, 40, 16384)
    assert a["variant"] == b["variant"] and abs(a["ms"] - b["ms"]) < 1e-3  # 4 decimals in the file


def test_online_tuning_keeps_results(env_ctx, tmp_path):
    cache = tmp_path / "cache.txt"
    c = env_ctx(P2G_AUTOTUNE=1, P2G_TUNE_CACHE=str(cache))
    files, S, R = [data(300000, 11), data(5000, 12)], 8192, 21
    sid = data(16, 13)
    want = expected(files, S, R, sid)
    assert run_pass(c, files, S, R, sid) == want  # the tuning call
    assert run_pass(c, files, S, R, sid) == want  # the cached choice
    import par2rust_amd as p2

    n = sum((len(f) + S - 1) // S for f in files)
    lines = cache.read_text().split()
    # bucket, variant, band, ms, encode kernel (-1: not timed online), its ms
    assert str(p2.tune_bucket(n, R, S)) == lines[0] and len(lines) == 6 and lines[4] == "-1"


@pytest.mark.parametrize("variant,band", [(1, 0), (2, 0), (0, 1), (0, 3)])
def test_cached_choice_drives_pass_and_batch(env_ctx, tmp_path, variant, band):
    import par2rust_amd as p2
    from par2rust_amd.hipmem import DeviceBuffer

    files, S, R = [data(1 << 20, 21)], 65536, 18
    sid = data(16, 22)
    n = 16
    cache = tmp_path / "forced.txt"
    cache.write_text(f"{p2.tune_bucket(n, R, S)} {variant} {band} 1.0\n")
    c = env_ctx(P2G_TUNE_CACHE=str(cache))
    assert c.tune_get(n, R, S) == {"variant": variant, "band_cols": band}
    want = expected(files, S, R, sid)
    assert run_pass(c, files, S, R, sid) == want
    dev = DeviceBuffer.from_bytes(files[0])
    outs = [DeviceBuffer(R * S), DeviceBuffer(16 * n), DeviceBuffer(4 * n), DeviceBuffer(16 * R)]
    c.dev_create_batch([p2.BatchJob([(dev.addr, len(files[0]))], S, R, sid, *outs)])
    c.synchronize()
    assert (outs[0].download(), outs[1].download(),
            list(struct.unpack(f"<{n}I", outs[2].download())), outs[3].download()) == want
