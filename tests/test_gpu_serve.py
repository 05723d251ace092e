"""GPU tests of the serving loop (p2g_server_*): every request's outputs are
the same as its job computed alone (the oracle), whether it runs whole in a
step, in row or column pieces across steps, or waits in the queue for HBM;
and the loop's error and accounting paths."""
import time

import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def data(n: int, seed: int) -> bytes:
    return O.xorshift64_bytes(n, seed)


def alone(files, S, R, set_id):
    big = sum(len(f) for f in files) * R >= (4 << 20)
    rec = (O.encode_simd(files, S, R) if big else O.encode(files, S, R)) if R else b""
    md5s, crcs = [], []
    for f in files:
        m, cr = O.slice_checksums(f, S)
        md5s += m
        crcs += cr
    pkt = b"".join(O.recovery_packet(set_id, r, rec[r * S:(r + 1) * S])[16:32] for r in range(R))
    return rec, b"".join(md5s), crcs, pkt


REQS = [
    # (file lengths, S, R)
    ([70000, 4100], 4096, 7),
    ([1 << 20], 16384, 40),
    ([1000], 100, 3),
    ([5000], 64, 0),
    ([16 * 3000 + 9], 16, 50),
    ([300000], 65536, 70),
    ([4 * 7000], 4, 3),              # 7000 slices: checksums over several steps at a tiny target
]


def run_all(srv, reqs, stagger=True):
    """Submits the requests (host bytes), half of them while the first ones
    run, polls until all are done; returns {id: (files, S, R, sid)}."""
    ids, done = {}, set()
    for k, (lens, S, R) in enumerate(reqs):
        files = [data(n, 700 + 10 * k + i) for i, n in enumerate(lens)]
        sid = data(16, 900 + k)
        ids[srv.submit(lens, S, R, sid, host_data=files)] = (files, S, R, sid)
        if stagger and k == len(reqs) // 2:
            done.update(srv.poll(1.0))  # a poll reports each finished id once
    t = time.time()
    while len(done) < len(ids):
        done.update(srv.poll(50.0))
        assert time.time() - t < 60, "server made no progress"
    return ids


def test_server_results_equal_alone(ctx):
    srv = ctx.server(1 << 30)
    ids = run_all(srv, REQS)
    for rid, (files, S, R, sid) in ids.items():
        assert srv.download(rid) == alone(files, S, R, sid), rid
        srv.release(rid)
    assert srv.stats() == {"queued": 0, "active": 0, "hbm_used": 0}
    srv.close()


def test_server_row_pieces_across_steps(ctx):
    # a tiny step target: every job with more than 16 rows runs 16-row pieces
    # over several steps, mixed with the other requests, and the 64 KiB-slice
    # job's row blocks (128 column groups) run as column pieces of 32 groups
    # whose packet hashes continue from the saved MD5 state
    srv = ctx.server(1 << 30, step_ms=1e-6)
    ids = run_all(srv, REQS)
    for rid, (files, S, R, sid) in ids.items():
        assert srv.download(rid) == alone(files, S, R, sid), rid
        _, (t_sub, t_adm, t_done) = srv.result(rid)
        assert t_sub <= t_adm <= t_done
    srv.close()


ADMIT_REQS = [([500000], 4096, 20), ([450000, 3], 4096, 12), ([400000], 2048, 9)]


def test_server_admission_waits_for_memory(ctx):
    # room for about one request at a time: the others wait in the queue,
    # in arrival order (no overtaking)
    reqs = ADMIT_REQS
    srv = ctx.server(1 << 20)
    srv.set_bypass(0.0)
    ids = {}
    for k, (lens, S, R) in enumerate(reqs):
        files = [data(n, 1700 + 10 * k + i) for i, n in enumerate(lens)]
        sid = data(16, 1900 + k)
        ids[srv.submit(lens, S, R, sid, host_data=files)] = (files, S, R, sid)
    finished = srv.poll(0.0)
    st = srv.stats()
    assert st["queued"] >= 1 and 0 < st["hbm_used"] <= 1 << 20
    done, t = [], time.time()
    while len(done) < len(ids):
        for rid in finished:
            files, S, R, sid = ids[rid]
            assert srv.download(rid) == alone(files, S, R, sid), rid
            srv.release(rid)  # frees room for the next request
            done.append(rid)
        assert time.time() - t < 60, "server made no progress"
        finished = srv.poll(50.0)
    assert done == sorted(ids)  # admitted and finished in arrival order
    srv.close()


def submit_admit_reqs(srv, ids, ks):
    for k in ks:
        lens, S, R = ADMIT_REQS[k]
        files = [data(n, 1700 + 10 * k + i) for i, n in enumerate(lens)]
        sid = data(16, 1900 + k)
        ids[srv.submit(lens, S, R, sid, host_data=files)] = (files, S, R, sid)


def finish_all(srv, ids, finished=()):
    """Polls until every request is done (`finished`: ids earlier polls
    reported), checks each against the oracle, releases it; returns
    {id: admit time}."""
    admitted, t = {}, time.time()
    finished = list(finished)
    while len(admitted) < len(ids):
        for rid in finished:
            files, S, R, sid = ids[rid]
            assert srv.download(rid) == alone(files, S, R, sid), rid
            admitted[rid] = srv.result(rid)[1][1]
            srv.release(rid)
        assert time.time() - t < 60, "server made no progress"
        finished = srv.poll(50.0)
    return admitted


def test_server_admission_bypass(ctx):
    # A (582 KB) fits the 1 MiB budget, B (500 KB) does not beside it, C
    # (418 KB) does: C goes ahead of the waiting B, which is admitted once A
    # and C are released
    srv = ctx.server(1 << 20)
    srv.set_bypass(60000.0)
    ids = {}
    submit_admit_reqs(srv, ids, [0, 1, 2])
    finished = srv.poll(0.0)
    st = srv.stats()
    assert st["queued"] == 1 and st["active"] + len(finished) == 2
    a, b, c = sorted(ids)
    t = finish_all(srv, ids, finished)
    assert t[a] <= t[c] < t[b]
    srv.close()


def test_server_admission_bypass_window_ends(ctx):
    # once B has waited its window, C queues behind it although it would fit
    srv = ctx.server(1 << 20)
    srv.set_bypass(50.0)
    ids = {}
    submit_admit_reqs(srv, ids, [0, 1])
    finished = srv.poll(0.0)
    assert srv.stats()["queued"] == 1
    time.sleep(0.2)
    submit_admit_reqs(srv, ids, [2])
    finished += srv.poll(0.0)  # A may be done by now: not released yet
    assert srv.stats()["queued"] == 2
    a, b, c = sorted(ids)
    t = finish_all(srv, ids, finished)
    assert t[a] < t[b] <= t[c]
    srv.close()


def test_server_device_filled_requests(ctx):
    # files filled on the device from seeds equal the documented stream
    srv = ctx.server(1 << 30)
    lens, S, R = [123457, 4096], 4096, 9
    rid = srv.submit(lens, S, R, bytes(16), seeds=[5, 6])
    while rid not in srv.poll(50.0):
        pass
    files = [data(lens[0], 5), data(lens[1], 6)]
    assert srv.download(rid) == alone(files, S, R, bytes(16))
    srv.close()


def test_server_device_filled_many(ctx):
    # many seeded requests admitted together: their 60 files are filled by
    # shared launches (up to 48 files each), across request boundaries, with
    # sizes from a few bytes to tens of KiB and not multiples of 8
    srv = ctx.server(1 << 30)
    want = {}
    for k in range(30):
        lens = [1 + (k * 7919) % 50000, 16 * k + 3]
        S, R = 4096, 1 + k % 5
        seeds = [100 + 2 * k, 101 + 2 * k]
        rid = srv.submit(lens, S, R, bytes(16), seeds=seeds)
        want[rid] = ([data(n, sd) for n, sd in zip(lens, seeds)], S, R)
    done, t = set(), time.time()
    while len(done) < len(want):
        done.update(srv.poll(50.0))
        assert time.time() - t < 60, "server made no progress"
    for rid, (files, S, R) in want.items():
        assert srv.download(rid) == alone(files, S, R, bytes(16)), rid
    srv.close()


def test_server_errors(ctx):
    import par2rust_amd as p2

    srv = ctx.server(1 << 20)
    with pytest.raises(p2.Par2Error) as e:
        srv.submit([4 << 20], 4096, 4, bytes(16), seeds=[1])  # never fits the budget
    assert e.value.kind == "OutOfMemory"
    with pytest.raises(p2.Par2Error) as e:
        srv.submit([], 4096, 4, bytes(16))
    assert e.value.kind == "InvalidArgument"
    with pytest.raises(p2.Par2Error) as e:
        srv.submit([100], 6, 4, bytes(16), seeds=[1])
    assert e.value.kind == "InvalidSliceSize"
    rid = srv.submit([100], 4, 2, bytes(16), seeds=[1])
    with pytest.raises(p2.Par2Error):
        srv.release(rid)  # not done yet
    while rid not in srv.poll(50.0):
        pass
    srv.release(rid)
    srv.close()
