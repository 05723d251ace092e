"""End-to-end parity: p2g_run_create (scan + run_create on the GPU engine)
writes files byte-identical to the oracle's restatement of run_create, and
reports the reference's errors in the reference's order."""
import os

import pytest

import par2rust_amd as p2
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["resident", "stream"])
def create_path(request, monkeypatch):
    """p2g_run_create's two pipelines: every file read once into HBM (the
    default when the set fits), or scan then a second read for the encode."""
    if request.param == "stream":
        monkeypatch.setenv("P2G_CREATE_STREAM", "1")
    else:
        monkeypatch.delenv("P2G_CREATE_STREAM", raising=False)
    return request.param


def write(path, b):
    with open(path, "wb") as f:
        f.write(b)
    return path


def both(tmp_path, inputs, S, R, scheme: p2.VolumeScheme, ctx, out_name="recovery.par2"):
    da, db = tmp_path / "gpu", tmp_path / "cpu"
    da.mkdir()
    db.mkdir()
    got = ctx.run_create(p2.CreateOptions(str(da / out_name), S, R, scheme), inputs)
    want = O.run_create(str(db / out_name), S, R, inputs, scheme.kind, scheme.uniform_count,
                        scheme.limited, scheme.limit_cap, scheme.explicit_sizes, scheme.outer_caps)
    assert [os.path.basename(p) for p in got] == [os.path.basename(p) for p in want]
    for g, w in zip(got, want):
        assert open(g, "rb").read() == open(w, "rb").read(), os.path.basename(g)
    return got


def test_config0_exponential(tmp_path, ctx, create_path):
    src = write(tmp_path / "input.bin", O.xorshift64_bytes(4 << 20, 1))
    got = both(tmp_path, [(str(src), b"input.bin")], 4096, 10, p2.VolumeScheme.exponential(), ctx)
    assert [os.path.basename(p) for p in got] == [
        "recovery.par2", "recovery.vol0+1.par2", "recovery.vol1+1.par2", "recovery.vol2+2.par2",
        "recovery.vol4+4.par2", "recovery.vol8+2.par2"]


def test_small_single_volume(tmp_path, ctx):
    # src/creator.rs:1210-1233
    src = write(tmp_path / "a.bin", b"hello par2 world!!!!")
    got = both(tmp_path, [(str(src), b"a.bin")], 4, 2, p2.VolumeScheme.single(), ctx)
    assert [os.path.basename(p) for p in got] == ["recovery.par2", "recovery.vol0+2.par2"]


def test_index_only(tmp_path, ctx, create_path):
    src = write(tmp_path / "a.bin", b"hello par2 world")
    got = both(tmp_path, [(str(src), b"a.bin")], 4, 0, p2.VolumeScheme.single(), ctx)
    assert len(got) == 1


def test_multi_file_uniform_with_zero_files(tmp_path, ctx, create_path):
    names = []
    for i, n in enumerate([70000, 65536, 1, 200000, 65537]):
        b = bytes(n) if i == 1 else O.xorshift64_bytes(n, 1000 + i)
        names.append((str(write(tmp_path / f"f{i}.dat", b)), f"f{i}.dat".encode()))
    both(tmp_path, names, 65536, 20, p2.VolumeScheme.uniform(3), ctx)


def test_explicit_and_limited(tmp_path, ctx, create_path):
    src = write(tmp_path / "x.bin", O.xorshift64_bytes(30000, 5))
    got = both(tmp_path, [(str(src), b"x.bin")], 1024, 6, p2.VolumeScheme.explicit([3, 3]), ctx)
    assert [os.path.basename(p) for p in got][1:] == ["recovery.vol0+3.par2", "recovery.vol3+3.par2"]
    (tmp_path / "lim").mkdir()
    both(tmp_path / "lim", [(str(src), b"x.bin")], 1024, 20,
         p2.VolumeScheme.exponential().capped(3), ctx, out_name="lim.par2")
    # Limited nested in Limited (src/creator.rs:131-134): caps 6 then 4
    (tmp_path / "nest").mkdir()
    got = both(tmp_path / "nest", [(str(src), b"x.bin")], 1024, 10,
               p2.VolumeScheme.explicit([10]).capped(6).capped(4), ctx, out_name="n.par2")
    assert [os.path.basename(p) for p in got][1:] == ["n.vol0+4.par2", "n.vol4+2.par2",
                                                       "n.vol6+4.par2"]


def test_output_name_without_extension(tmp_path, ctx):
    src = write(tmp_path / "x.bin", O.xorshift64_bytes(3000, 6))
    got = both(tmp_path, [(str(src), b"x.bin")], 512, 3, p2.VolumeScheme.single(), ctx,
               out_name="noext")
    assert os.path.basename(got[1]) == "noext.vol0+3.par2"


def expect_error(fn, kind):
    with pytest.raises(p2.Par2Error) as e:
        fn()
    assert e.value.kind == kind
    return e.value


def test_errors(tmp_path, ctx, create_path):
    out = str(tmp_path / "r.par2")
    empty = write(tmp_path / "empty.bin", b"")
    ok = write(tmp_path / "ok.bin", b"hi there")
    opts = lambda S=4, R=1, sch=None: p2.CreateOptions(out, S, R, sch or p2.VolumeScheme.single())
    expect_error(lambda: ctx.run_create(opts(), [(str(empty), b"empty.bin")]), "EmptyFile")
    expect_error(lambda: ctx.run_create(opts(), [(str(ok), b"bad\0name")]), "InvalidFileName")
    expect_error(lambda: ctx.run_create(opts(), [(str(ok), b"")]), "InvalidFileName")
    expect_error(lambda: ctx.run_create(opts(S=5), [(str(ok), b"ok.bin")]), "InvalidSliceSize")
    expect_error(lambda: ctx.run_create(opts(S=0), [(str(ok), b"ok.bin")]), "InvalidSliceSize")
    expect_error(lambda: ctx.run_create(opts(), []), "NoInputFiles")
    expect_error(lambda: ctx.run_create(opts(R=65536), [(str(ok), b"ok.bin")]),
                 "TooManyRecoveryBlocks")
    expect_error(lambda: ctx.run_create(opts(), [(str(tmp_path / "missing"), b"m")]), "Io")
    big = write(tmp_path / "big.bin", bytes(4 * 32769))
    e = expect_error(lambda: ctx.run_create(opts(), [(str(big), b"big.bin")]), "TooManyInputBlocks")
    assert "slice_size >= 8" in str(e)  # ceil(131076 / 32768) = 5 -> 8
    # a bad scheme fails after the index file is written (src/creator.rs:237 then :244)
    expect_error(lambda: ctx.run_create(opts(R=3, sch=p2.VolumeScheme.uniform(4)),
                                        [(str(ok), b"ok.bin")]), "InvalidVolumeScheme")
    assert os.path.exists(out)


def test_config0_against_golden_manifest(tmp_path, ctx):
    import hashlib
    import json

    from tests.par2_check import check_set

    m = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "config0.json")))
    data = O.xorshift64_bytes(m["input"]["bytes"], m["input"]["seed"])
    src = write(tmp_path / "input.bin", data)
    files = ctx.run_create(p2.CreateOptions(str(tmp_path / "recovery.par2"), m["slice_size"],
                                            m["recovery"], p2.VolumeScheme.exponential()),
                           [(str(src), b"input.bin")])
    got = [{"name": os.path.basename(p), "size": os.path.getsize(p),
            "sha256": hashlib.sha256(open(p, "rb").read()).hexdigest()} for p in files]
    assert got == m["files"]
    check_set({os.path.basename(p): open(p, "rb").read() for p in files},
              [(b"input.bin", data)], m["slice_size"], m["recovery"])


def test_tiny_against_golden_bytes(tmp_path, ctx):
    import json

    t = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "tiny.json")))
    src = write(tmp_path / "a.bin", t["input"].encode())
    files = ctx.run_create(p2.CreateOptions(str(tmp_path / "recovery.par2"), 4, 2,
                                            p2.VolumeScheme.single()), [(str(src), b"a.bin")])
    assert [open(p, "rb").read().hex() for p in files] == [f["hex"] for f in t["files"]]


def test_failed_volume_error_and_the_other_volumes(tmp_path, ctx, create_path):
    # src/creator.rs:296-316 writes the volumes as a parallel map and returns
    # the first failure in volume order: when one volume cannot be written
    # (here its name is taken by a directory, so its rename fails) the create
    # reports an I/O error, and the index and every other volume are on disk
    src = write(tmp_path / "input.bin", O.xorshift64_bytes(4 << 20, 1))
    out = tmp_path / "out"
    out.mkdir()
    (out / "recovery.vol2+2.par2").mkdir()
    with pytest.raises(p2.Par2Error) as e:
        ctx.run_create(p2.CreateOptions(str(out / "recovery.par2"), 4096, 10,
                                        p2.VolumeScheme.exponential()), [(str(src), b"input.bin")])
    assert e.value.kind == "Io"
    names = set(os.listdir(out))
    assert {"recovery.par2", "recovery.vol0+1.par2", "recovery.vol1+1.par2",
            "recovery.vol4+4.par2", "recovery.vol8+2.par2"} <= names
    assert (out / "recovery.vol2+2.par2").is_dir()
