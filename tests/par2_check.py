"""Independent structural check of a PAR2 create output set (test helper).

Parses packets with the layout of src/format.rs:33-54 and checks every field
that can be recomputed with library routines (hashlib, zlib) from the
source files: packet hashes, set id, file ids, whole-file hashes, per-slice
MD5/CRC-32, and recovery row 0 (exponent 0 has every coefficient 1, so it
is the XOR of all zero-padded input slices).
"""
import hashlib
import struct
import zlib

MAGIC = b"PAR2\0PKT"
T_MAIN = b"PAR 2.0\0Main\0\0\0\0"
T_DESC = b"PAR 2.0\0FileDesc"
T_IFSC = b"PAR 2.0\0IFSC\0\0\0\0"
T_RECV = b"PAR 2.0\0RecvSlic"
T_CREATOR = b"PAR 2.0\0Creator\0"


def packets(blob: bytes):
    out, i = [], 0
    while i < len(blob):
        assert blob[i:i + 8] == MAGIC, f"bad magic at {i}"
        (length,) = struct.unpack_from("<Q", blob, i + 8)
        assert length >= 64 and length % 4 == 0 and i + length <= len(blob)
        pkt_md5, set_id, ptype = blob[i + 16:i + 32], blob[i + 32:i + 48], blob[i + 48:i + 64]
        body = blob[i + 64:i + length]
        assert hashlib.md5(set_id + ptype + body).digest() == pkt_md5, f"packet hash at {i}"
        out.append((ptype, set_id, body))
        i += length
    return out


def check_set(files: dict, sources: list, S: int, R: int):
    """files: name -> bytes of the written set (index first);
    sources: [(display name, bytes)] in caller order."""
    names = list(files)
    index = packets(files[names[0]])
    main = index[0]
    assert main[0] == T_MAIN
    set_id = main[1]
    assert hashlib.md5(main[2]).digest() == set_id
    slice_size, nfiles = struct.unpack_from("<QI", main[2], 0)
    assert slice_size == S and nfiles == len(sources)
    ids = []
    for k, (name, data) in enumerate(sources):
        desc, ifsc = index[1 + 2 * k], index[2 + 2 * k]
        assert desc[0] == T_DESC and ifsc[0] == T_IFSC
        h16 = hashlib.md5(data[:16384]).digest()
        fid = hashlib.md5(h16 + struct.pack("<Q", len(data)) + name).digest()
        b = desc[2]
        assert b[0:16] == fid and b[16:32] == hashlib.md5(data).digest() and b[32:48] == h16
        assert struct.unpack_from("<Q", b, 48)[0] == len(data)
        assert b[56:56 + len(name)] == name and set(b[56 + len(name):]) <= {0}
        assert len(b) == 56 + ((len(name) + 3) & ~3)
        ib = ifsc[2]
        assert ib[:16] == fid
        nsl = (len(data) + S - 1) // S
        assert len(ib) == 16 + 20 * nsl
        for s in range(nsl):
            sl = data[s * S:(s + 1) * S].ljust(S, b"\0")
            assert ib[16 + 20 * s:32 + 20 * s] == hashlib.md5(sl).digest()
            assert struct.unpack_from("<I", ib, 32 + 20 * s)[0] == zlib.crc32(sl)
        ids.append(fid)
    assert main[2][12:] == b"".join(sorted(ids))
    assert index[-1][0] == T_CREATOR and index[-1][2].startswith(b"Created by par2rust v0.1.0")
    assert len(index) == 2 + 2 * len(sources)
    critical = files[names[0]]
    seen = {}
    for vname in names[1:]:
        v = files[vname]
        assert v.startswith(critical) and v.endswith(critical)
        mid = packets(v[len(critical):len(v) - len(critical)])
        exps = []
        for ptype, sid, body in mid:
            assert ptype == T_RECV and sid == set_id and len(body) == 4 + S
            e = struct.unpack_from("<I", body, 0)[0]
            exps.append(e)
            seen[e] = body[4:]
        first, count = map(int, vname.rsplit(".vol", 1)[1].split(".")[0].split("+"))
        assert exps == list(range(first, first + count))
    if R:
        import numpy as np

        assert sorted(seen) == list(range(R))
        parity = np.zeros(S, dtype=np.uint8)
        for _, data in sources:
            n = (len(data) + S - 1) // S
            if n:
                a = np.zeros(n * S, dtype=np.uint8)
                a[:len(data)] = np.frombuffer(data, dtype=np.uint8)
                parity ^= np.bitwise_xor.reduce(a.reshape(n, S), axis=0)
        assert seen[0] == parity.tobytes()
    return seen
