"""GPU parity of the batched create (p2g_batch_plan / launch, p2g_dev_create_batch):
every job of a batch gets exactly what the oracle computes for it alone
(recovery blocks, slice MD5 / CRC-32, packet MD5s), across mixed slice
sizes and recovery counts, jobs without recovery blocks, jobs past one K
chunk of 4096 inputs, row pieces and several launch groups, graph replays;
and the device xorshift64 generator against the documented stream."""
import os
import struct

import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def data(n: int, seed: int) -> bytes:
    return O.xorshift64_bytes(n, seed)


def expected(files, S, R, set_id):
    big = sum(len(f) for f in files) * R >= (4 << 20)
    rec = (O.encode_simd(files, S, R) if big else O.encode(files, S, R)) if R else b""
    md5s, crcs = [], []
    for f in files:
        m, cr = O.slice_checksums(f, S)
        md5s += m
        crcs += cr
    pkt = b"".join(O.recovery_packet(set_id, r, rec[r * S:(r + 1) * S])[16:32] for r in range(R))
    return rec, b"".join(md5s), crcs, pkt


class Job:
    """One batch job with its device buffers and expected outputs."""

    def __init__(self, lens, S, R, seed):
        import par2rust_amd as p2
        from par2rust_amd.hipmem import DeviceBuffer

        self.files = [data(n, seed + i) for i, n in enumerate(lens)]
        self.S, self.R = S, R
        self.sid = data(16, seed + 99)
        self.n = sum((len(f) + S - 1) // S for f in self.files)
        self.dev = [DeviceBuffer.from_bytes(f) if f else DeviceBuffer(16) for f in self.files]
        self.rec = DeviceBuffer(max(16, R * S))
        self.md5 = DeviceBuffer(16 * max(1, self.n))
        self.crc = DeviceBuffer(4 * max(1, self.n))
        self.pkt = DeviceBuffer(16 * max(1, R))
        self.job = p2.BatchJob([(d.addr, len(f)) for d, f in zip(self.dev, self.files)], S, R,
                               self.sid, self.rec, self.md5, self.crc, self.pkt)

    def clear(self):
        for b in (self.rec, self.md5, self.crc, self.pkt):
            b.zero()

    def got(self):
        n, R, S = self.n, self.R, self.S
        return (self.rec.download(R * S) if R else b"", self.md5.download(16 * n),
                list(struct.unpack(f"<{n}I", self.crc.download(4 * n))), self.pkt.download(16 * R))

    def want(self):
        return expected(self.files, self.S, self.R, self.sid)


MIX = [
    # (file lengths, S, R)
    ([70000, 4100], 4096, 7),
    ([777], 12, 4),                  # S % 16 != 0
    ([1000], 100, 3),
    ([1 << 20], 16384, 33),          # three row blocks
    ([300000], 65536, 70),           # two packet tasks
    ([4 * 600, 8], 4, 5),            # 602 slices, one-byte-ish partial
    ([5000], 64, 0),                 # no recovery blocks: checksums only
    ([16 * 5000 + 9], 16, 20),       # 5001 inputs: two K chunks
    ([150, 64, 3], 64, 1),
]


def make_jobs(spec, seed0=1000):
    return [Job(lens, S, R, seed0 + 10 * k) for k, (lens, S, R) in enumerate(spec)]


@pytest.fixture
def env_ctx():
    """Contexts opened with environment overrides (restored afterwards)."""
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


def test_batch_matches_oracle(ctx):
    jobs = make_jobs(MIX)
    b = ctx.batch_plan([j.job for j in jobs])
    b.launch()
    b.wait()
    for k, j in enumerate(jobs):
        assert j.got() == j.want(), f"job {k}"
    ms = b.job_ms()
    assert len(ms) == len(jobs) and all(m >= 0 for m in ms)
    b.free()


def test_batch_replays_and_direct_launches(ctx):
    # graph replays and direct launches give the same bytes every time
    jobs = make_jobs(MIX[:4] + MIX[7:8], seed0=2000)
    want = [j.want() for j in jobs]
    for graph in (True, False):
        b = ctx.batch_plan([j.job for j in jobs], use_graph=graph)
        for rep in range(3):
            for j in jobs:
                j.clear()
            b.launch()
            b.wait()
            assert [j.got() for j in jobs] == want, (graph, rep)
        b.free()


def test_batch_row_pieces_and_groups(env_ctx, monkeypatch):
    # a tiny fragment budget: every job splits into 16-row pieces and every
    # piece is a launch group of its own
    monkeypatch.setenv("P2G_BATCH_AFRAG_BYTES", "1")
    c = env_ctx()
    jobs = make_jobs([([1 << 18], 8192, 50), ([40000, 11], 2048, 37), ([999], 36, 0),
                      ([16 * 2100], 16, 17)], seed0=3000)
    b = c.batch_plan([j.job for j in jobs])
    b.launch()
    b.wait()
    for k, j in enumerate(jobs):
        assert j.got() == j.want(), f"job {k}"
    b.free()


def test_batch_narrow_bands(env_ctx):
    c = env_ctx(P2G_BAND_COLS=1)
    jobs = make_jobs([([1 << 20], 65536, 20), ([23000, 7], 5000, 9)], seed0=4000)
    c.dev_create_batch([j.job for j in jobs])
    c.synchronize()
    for k, j in enumerate(jobs):
        assert j.got() == j.want(), f"job {k}"


def test_batch_equals_plain_pass(ctx):
    # the batched form and p2g_dev_create_pass give the same bytes
    from par2rust_amd.hipmem import DeviceBuffer

    jobs = make_jobs([([200000], 8192, 12), ([50000, 3], 4096, 5)], seed0=5000)
    ctx.dev_create_batch([j.job for j in jobs])
    ctx.synchronize()
    for j in jobs:
        rec = DeviceBuffer(j.R * j.S)
        md5, crc, pkt = DeviceBuffer(16 * j.n), DeviceBuffer(4 * j.n), DeviceBuffer(16 * j.R)
        ctx.dev_create_pass([(d.addr, len(f)) for d, f in zip(j.dev, j.files)], j.S, j.R, j.sid,
                            rec, md5, crc, pkt)
        ctx.synchronize()
        assert j.got() == (rec.download(), md5.download(),
                           list(struct.unpack(f"<{j.n}I", crc.download())), pkt.download())


def test_batch_errors(ctx):
    import par2rust_amd as p2

    with pytest.raises(p2.Par2Error) as e:
        ctx.batch_plan([])
    assert e.value.kind == "InvalidArgument"
    j = Job([100], 4, 2, 6000)
    bad = p2.BatchJob([], 4, 2, j.sid, j.rec, j.md5, j.crc, j.pkt)  # no slices
    with pytest.raises(p2.Par2Error) as e:
        ctx.batch_plan([j.job, bad])
    assert e.value.kind == "InvalidArgument"
    bad = p2.BatchJob(j.job.files, 6, 2, j.sid, j.rec, j.md5, j.crc, j.pkt)
    with pytest.raises(p2.Par2Error) as e:
        ctx.batch_plan([bad])
    assert e.value.kind == "InvalidSliceSize"


@pytest.mark.parametrize("n,seed", [(1, 1), (7, 2), (8, 3), (16, 4), (128 * 8 + 5, 8),
                                    (3 * 128 * 8 - 1, 9), (2048 * 8 + 5, 5),
                                    (3 * (1 << 20) + 3, 6), (40 << 20, 7)])
def test_fill_xorshift64(ctx, n, seed):
    from par2rust_amd import synth
    from par2rust_amd.hipmem import DeviceBuffer

    buf = DeviceBuffer(n + 16)
    buf.zero()
    ctx.dev_fill_xorshift64(buf, n, seed)
    ctx.synchronize()
    got = buf.download(n + 16)
    assert got[n:] == bytes(16)  # nothing written past n
    want = synth.xorshift64_bytes(n, seed) if n > 4096 else O.xorshift64_bytes(n, seed)
    assert got[:n] == want
