"""Failure paths of the persistent launches, forced from the host.

A packet-hash lane that waits for the encode's bands gives up when nothing
moves for a stall limit and flags the launch.  P2G_STALL_TICKS (read at each
launch) sets that limit; at 0 every wait that finds its band unfinished gives
up at once, while the encode and the slice hashes still run to their normal
end, so every kernel exits normally.  The jobs below are sized so that the
first band takes far longer than the waiters need to reach their first
check.  Checked: an error is reported by the next wait even when a later
launch ran after it (the error word is sticky), and the serving loop reports
the requests of a failed step as failed (not done) without losing their ids.
"""
import os
import time

import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

# 2000 slices of 4 KiB, 64 recovery blocks: every unit runs 500 K steps
# (~0.2 ms) before the first band can complete
N_SLICES, S, R = 2000, 4096, 64


class stall_limit:
    def __init__(self, ticks):
        self.ticks = ticks

    def __enter__(self):
        os.environ["P2G_STALL_TICKS"] = str(self.ticks)

    def __exit__(self, *exc):
        os.environ.pop("P2G_STALL_TICKS", None)


def _buffers(n, S, R):
    from par2rust_amd.hipmem import DeviceBuffer

    return [DeviceBuffer(R * S), DeviceBuffer(16 * n), DeviceBuffer(4 * n), DeviceBuffer(16 * R)]


def test_pass_error_is_kept_across_a_later_pass(ctx):
    import par2rust_amd as p2
    from par2rust_amd.hipmem import DeviceBuffer

    src = DeviceBuffer.from_bytes(O.xorshift64_bytes(N_SLICES * S, 31))
    files = [(src.addr, N_SLICES * S)]
    outs = _buffers(N_SLICES, S, R)
    ctx.synchronize()
    with stall_limit(0):
        ctx.dev_create_pass(files, S, R, bytes(16), *outs)  # its packet lanes give up
    ctx.dev_create_pass(files, S, R, bytes(16), *outs)      # queued behind it, completes
    with pytest.raises(p2.Par2Error) as e:
        ctx.synchronize()
    assert e.value.kind == "Device"
    # reported once: the next wait is clean, and a pass then gives the oracle's bytes
    ctx.synchronize()
    ctx.dev_create_pass(files, S, R, bytes(16), *outs)
    ctx.synchronize()
    data = src.download()
    rec = O.encode_simd([data], S, R)
    assert outs[0].download() == rec
    assert outs[3].download() == b"".join(
        O.recovery_packet(bytes(16), r, rec[r * S:(r + 1) * S])[16:32] for r in range(R))


def test_batch_error_is_kept_across_a_later_launch(ctx):
    import par2rust_amd as p2
    from par2rust_amd.hipmem import DeviceBuffer

    src = DeviceBuffer.from_bytes(O.xorshift64_bytes(N_SLICES * S, 32))
    outs = _buffers(N_SLICES, S, R)
    plan = ctx.batch_plan([p2.BatchJob([(src.addr, N_SLICES * S)], S, R, bytes(16), *outs)],
                          use_graph=False)
    with stall_limit(0):
        plan.launch()
    plan.launch()  # runs of one plan are serialized; this one completes
    with pytest.raises(p2.Par2Error) as e:
        plan.wait()
    assert e.value.kind == "Device"
    plan.launch()
    plan.wait()  # clean again
    plan.free()


def test_server_failed_step_reports_its_requests_failed(ctx):
    import par2rust_amd as p2

    srv = ctx.server(1 << 30)
    lens = [N_SLICES * S]
    with stall_limit(0):
        bad = srv.submit(lens, S, R, bytes(16), seeds=[33])
        t, raised = time.time(), None
        while raised is None:
            try:
                assert not srv.poll(20.0), "a failed step's request was reported as done"
            except p2.Par2Error as err:
                raised = err
            assert time.time() - t < 60, "the step never ended"
    assert raised.kind == "Device"
    # the id is not lost: the next poll reports it, and result says it failed
    ids = []
    t = time.time()
    while bad not in ids:
        ids += srv.poll(20.0)
        assert time.time() - t < 60
    with pytest.raises(p2.Par2Error) as e:
        srv.result(bad)
    assert e.value.kind == "Device"
    srv.release(bad)
    assert srv.stats() == {"queued": 0, "active": 0, "hbm_used": 0}
    # the server keeps serving: a request after the failure is correct
    good = srv.submit([70000], 4096, 5, bytes(16), seeds=[34])
    while good not in srv.poll(50.0):
        pass
    rec, md5, crc, pkt = srv.download(good)
    want = O.xorshift64_bytes(70000, 34)
    assert rec == O.encode([want], 4096, 5)
    assert (md5, crc) == (b"".join(O.slice_checksums(want, 4096)[0]), O.slice_checksums(want, 4096)[1])
    srv.close()
