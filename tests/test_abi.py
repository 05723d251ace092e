"""The C-ABI library loads and exports every entry point include/par2gpu.h
declares.  No compute call is made: these run without a GPU."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "par2gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(p2g_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("p2g_open", "p2g_encode_begin", "p2g_encode_push", "p2g_encode_finish",
                 "p2g_slice_checksums", "p2g_recovery_packet_md5", "p2g_dev_encode",
                 "p2g_run_create"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import par2rust_amd as p2

    lib = ctypes.CDLL(p2.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(p2.EXPORTED_SYMBOLS) == declared_functions()


def test_version_string():
    import par2rust_amd as p2

    assert p2.version().startswith("par2gpu ")


def test_status_codes_match_header():
    import par2rust_amd as p2

    text = open(HEADER).read()
    codes = dict((int(v), k) for k, v in re.findall(r"(P2G_ERR_[A-Z_]+)\s*=\s*(\d+)", text))
    assert set(codes) == set(p2.ERRORS)
    assert codes[1] == "P2G_ERR_IO" and p2.ERRORS[1] == "Io"
    assert codes[8] == "P2G_ERR_TOO_MANY_INPUT_BLOCKS" and p2.ERRORS[8] == "TooManyInputBlocks"


def test_tune_bucket_is_bit_lengths():
    import par2rust_amd as p2

    # the headline shape: 1027 slices (11 bits), 200 blocks (8), 524288 bytes (20)
    assert p2.tune_bucket(1027, 200, 524288) == 11 | 8 << 8 | 20 << 16
    assert p2.tune_bucket(1024, 128, 1 << 19) == p2.tune_bucket(2047, 255, (1 << 20) - 4)
    assert p2.tune_bucket(2048, 200, 524288) != p2.tune_bucket(2047, 200, 524288)
