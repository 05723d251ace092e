// par2_create.cpp -- SourceFile::scan and run_create on the GPU engine.
//
// Control flow and byte layout follow src/source.rs:59-136 and
// src/creator.rs:188-321, 1025-1087; the heavy loops go through the C ABI:
// p2g_slice_checksums (IFSC MD5+CRC-32), p2g_encode_* (recovery blocks and
// their packet hashes).  hash_full / hash16k and the small critical-packet
// hashes stay on the host.
#include "par2_create.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cerrno>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fcntl.h>
#include <mutex>
#include <sys/stat.h>
#include <thread>
#include <unistd.h>

#include "host_md5.hpp"
#include "p2g_context.hpp"

namespace par2 {
namespace {

using p2g::HostMd5;

const uint8_t kMagic[8] = {'P', 'A', 'R', '2', 0, 'P', 'K', 'T'};
const uint8_t kTypeMain[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0, 'M', 'a', 'i', 'n', 0, 0, 0, 0};
const uint8_t kTypeFileDesc[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0,
                                   'F', 'i', 'l', 'e', 'D', 'e', 's', 'c'};
const uint8_t kTypeIfsc[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0, 'I', 'F', 'S', 'C', 0, 0, 0, 0};
const uint8_t kTypeRecv[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0,
                               'R', 'e', 'c', 'v', 'S', 'l', 'i', 'c'};
const uint8_t kTypeCreator[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0,
                                  'C', 'r', 'e', 'a', 't', 'o', 'r', 0};
const char kCreatorString[] = "Created by par2rust v0.1.0";  // src/packet/creator.rs:3

constexpr uint64_t kChunkBytes = 64ull << 20;  // bytes read per GPU hand-off

int set_err(Error &err, int code, const std::string &msg) {
    err.code = code;
    err.message = msg;
    return code;
}

void put_le64(std::vector<uint8_t> &v, uint64_t x) {
    for (int i = 0; i < 8; i++) v.push_back((uint8_t)(x >> (8 * i)));
}
void put_le32(std::vector<uint8_t> &v, uint32_t x) {
    for (int i = 0; i < 4; i++) v.push_back((uint8_t)(x >> (8 * i)));
}

// build_packet, src/format.rs:33-54.
void append_packet(std::vector<uint8_t> &out, const uint8_t set_id[16], const uint8_t type[16],
                   const std::vector<uint8_t> &body) {
    uint8_t hash[16];
    HostMd5 m;
    m.update(set_id, 16);
    m.update(type, 16);
    m.update(body.data(), body.size());
    m.finish(hash);
    out.insert(out.end(), kMagic, kMagic + 8);
    put_le64(out, 64 + body.size());
    out.insert(out.end(), hash, hash + 16);
    out.insert(out.end(), set_id, set_id + 16);
    out.insert(out.end(), type, type + 16);
    out.insert(out.end(), body.begin(), body.end());
}

// write_atomic, src/creator.rs:1070-1087.
int write_atomic(const std::string &path, const std::vector<const std::vector<uint8_t> *> &parts,
                 const std::vector<std::pair<const uint8_t *, size_t>> &raw, Error &err) {
    const std::string tmp = path + ".par2rust.tmp";
    int fd = ::open(tmp.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) return set_err(err, P2G_ERR_IO, "cannot create " + tmp + ": " + strerror(errno));
    auto write_all = [&](const uint8_t *p, size_t n) {
        while (n) {
            ssize_t w = ::write(fd, p, n);
            if (w < 0) {
                if (errno == EINTR) continue;
                return false;
            }
            p += w;
            n -= (size_t)w;
        }
        return true;
    };
    bool ok = true;
    for (auto *v : parts) ok = ok && write_all(v->data(), v->size());
    for (auto &r : raw) ok = ok && write_all(r.first, r.second);
    ok = ok && ::fsync(fd) == 0;
    ::close(fd);
    if (!ok) return set_err(err, P2G_ERR_IO, "write failed for " + tmp + ": " + strerror(errno));
    if (::rename(tmp.c_str(), path.c_str()) != 0)
        return set_err(err, P2G_ERR_IO, "rename failed for " + path + ": " + strerror(errno));
    return P2G_OK;
}


// Pinned host buffer for file reads.
struct Pinned {
    uint8_t *p = nullptr;
    size_t cap = 0;
    ~Pinned() {
        if (p) hipHostFree(p);
    }
    bool reserve(size_t n) {
        if (n <= cap) return true;
        if (p) hipHostFree(p);
        p = nullptr;
        cap = 0;
        if (hipHostMalloc(reinterpret_cast<void **>(&p), n, hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            return false;
        }
        cap = n;
        return true;
    }
};

bool read_full(int fd, uint8_t *buf, size_t n, size_t &got) {
    got = 0;
    while (got < n) {
        ssize_t r = ::read(fd, buf + got, n - got);
        if (r < 0) {
            if (errno == EINTR) continue;
            return false;
        }
        if (r == 0) break;
        got += (size_t)r;
    }
    return true;
}

uint64_t chunk_bytes_for(uint64_t S) {
    const uint64_t slices = std::max<uint64_t>(1, kChunkBytes / S);
    return slices * S;
}

}  // namespace

int scan(p2g_ctx *ctx, const std::string &path, const std::vector<uint8_t> &name,
         uint64_t slice_size, SourceFile &out, Error &err) {
    // validation order, src/source.rs:65-75
    if (name.empty() || std::find(name.begin(), name.end(), 0) != name.end())
        return set_err(err, P2G_ERR_INVALID_FILE_NAME, "filename " + path + " is empty or contains a NUL byte");
    if (slice_size == 0 || slice_size % 4 != 0)
        return set_err(err, P2G_ERR_INVALID_SLICE_SIZE,
                       "invalid slice size " + std::to_string(slice_size));
    struct stat st;
    if (::stat(path.c_str(), &st) != 0)
        return set_err(err, P2G_ERR_IO, "cannot stat " + path + ": " + strerror(errno));
    const uint64_t length = (uint64_t)st.st_size;
    if (length == 0) return set_err(err, P2G_ERR_EMPTY_FILE, "input file " + path + " is empty");
    const uint64_t total_slices = (length + slice_size - 1) / slice_size;

    out = SourceFile();
    out.name = name;
    out.path = path;
    out.length = length;
    out.slice_checksums.resize(total_slices);

    int fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) return set_err(err, P2G_ERR_IO, "cannot open " + path + ": " + strerror(errno));
    const uint64_t chunk = chunk_bytes_for(slice_size);
    Pinned buf;
    if (!buf.reserve((size_t)std::min<uint64_t>(chunk, length))) {
        ::close(fd);
        return set_err(err, P2G_ERR_OUT_OF_MEMORY, "pinned read buffer");
    }
    std::vector<uint8_t> md5s(16 * total_slices);
    std::vector<uint32_t> crcs(total_slices);
    HostMd5 full, first16k;
    uint64_t done = 0, slice_at = 0;
    int rc = P2G_OK;
    while (done < length) {
        const size_t want = (size_t)std::min<uint64_t>(chunk, length - done);
        size_t got = 0;
        if (!read_full(fd, buf.p, want, got) || got != want) {
            rc = set_err(err, P2G_ERR_IO, "short read on " + path);
            break;
        }
        // hash_full / hash16k stream on the host while the GPU hashes slices.
        const uint64_t at = done;
        std::thread host_hash([&, at, got]() {
            full.update(buf.p, got);
            if (at < 16384) first16k.update(buf.p, (size_t)std::min<uint64_t>(got, 16384 - at));
        });
        const uint64_t n = (got + slice_size - 1) / slice_size;
        rc = p2g_slice_checksums(ctx, buf.p, got, slice_size, md5s.data() + 16 * slice_at,
                                 crcs.data() + slice_at);
        host_hash.join();
        if (rc != P2G_OK) {
            set_err(err, rc, std::string("slice checksums: ") + p2g_last_error(ctx));
            break;
        }
        done += got;
        slice_at += n;
    }
    ::close(fd);
    if (rc != P2G_OK) return rc;
    full.finish(out.hash_full);
    first16k.finish(out.hash16k);
    for (uint64_t s = 0; s < total_slices; s++) {
        memcpy(out.slice_checksums[s].md5, md5s.data() + 16 * s, 16);
        out.slice_checksums[s].crc32 = crcs[s];
    }
    // compute_file_id, src/source.rs:303-309
    HostMd5 id;
    id.update(out.hash16k, 16);
    uint8_t l[8];
    for (int i = 0; i < 8; i++) l[i] = (uint8_t)(length >> (8 * i));
    id.update(l, 8);
    id.update(name.data(), name.size());
    id.finish(out.file_id);
    return P2G_OK;
}

namespace {

// run_create's validation, src/creator.rs:193-217 (after the scans).
int check_create(const CreateOptions &opts, const std::vector<SourceFile> &sources, Error &err) {
    if (sources.empty()) return set_err(err, P2G_ERR_NO_INPUT_FILES, "no input files supplied");
    if (sources.size() > P2G_MAX_FILES)
        return set_err(err, P2G_ERR_TOO_MANY_FILES, "too many input files");
    if (opts.recovery_block_count > P2G_MAX_RECOVERY_BLOCKS)
        return set_err(err, P2G_ERR_TOO_MANY_RECOVERY_BLOCKS,
                       "recovery block count exceeds PAR2 limit of 65535");
    const uint64_t S = opts.slice_size;
    if (S == 0 || S % 4 != 0)
        return set_err(err, P2G_ERR_INVALID_SLICE_SIZE, "invalid slice size " + std::to_string(S));
    uint64_t total_blocks = 0, total_bytes = 0;
    for (auto &s : sources) {
        total_blocks += (s.length + S - 1) / S;
        total_bytes += s.length;
    }
    if (total_blocks > P2G_MAX_INPUT_BLOCKS) {
        const uint64_t raw = (total_bytes + P2G_MAX_INPUT_BLOCKS - 1) / P2G_MAX_INPUT_BLOCKS;
        const uint64_t suggested = (raw + 3) / 4 * 4;
        return set_err(err, P2G_ERR_TOO_MANY_INPUT_BLOCKS,
                       "too many input blocks (" + std::to_string(total_blocks) + ") at slice_size " +
                           std::to_string(S) + "; retry with slice_size >= " +
                           std::to_string(suggested));
    }
    return P2G_OK;
}

// set_id: MD5 of the main packet body, src/packet/main_packet.rs:25-42 (file
// ids sorted).
void main_body(uint64_t S, const std::vector<SourceFile> &sources, std::vector<uint8_t> &body,
               uint8_t set_id[16]) {
    std::vector<std::array<uint8_t, 16>> ids(sources.size());
    for (size_t f = 0; f < sources.size(); f++) memcpy(ids[f].data(), sources[f].file_id, 16);
    std::sort(ids.begin(), ids.end());
    body.clear();
    put_le64(body, S);
    put_le32(body, (uint32_t)sources.size());
    for (auto &id : ids) body.insert(body.end(), id.begin(), id.end());
    HostMd5::digest(body.data(), body.size(), set_id);
}

// The critical packets: Main, FileDesc + IFSC per source in caller order
// (src/creator.rs:227-230), Creator.
void make_critical(uint64_t S, const std::vector<SourceFile> &sources, uint8_t set_id[16],
                   std::vector<uint8_t> &critical) {
    std::vector<uint8_t> body;
    main_body(S, sources, body, set_id);
    critical.clear();
    append_packet(critical, set_id, kTypeMain, body);
    for (auto &s : sources) {
        body.clear();  // src/packet/file_desc.rs:17-28
        body.insert(body.end(), s.file_id, s.file_id + 16);
        body.insert(body.end(), s.hash_full, s.hash_full + 16);
        body.insert(body.end(), s.hash16k, s.hash16k + 16);
        put_le64(body, s.length);
        body.insert(body.end(), s.name.begin(), s.name.end());
        body.resize(56 + ((s.name.size() + 3) & ~size_t(3)), 0);
        append_packet(critical, set_id, kTypeFileDesc, body);
        body.clear();  // src/packet/file_verify.rs:12-23
        body.insert(body.end(), s.file_id, s.file_id + 16);
        for (auto &c : s.slice_checksums) {
            body.insert(body.end(), c.md5, c.md5 + 16);
            put_le32(body, c.crc32);
        }
        append_packet(critical, set_id, kTypeIfsc, body);
    }
    body.assign(kCreatorString, kCreatorString + sizeof(kCreatorString) - 1);
    body.resize((body.size() + 3) & ~size_t(3), 0);  // src/packet/creator.rs:8-14
    append_packet(critical, set_id, kTypeCreator, body);
}

// Volumes, src/creator.rs:279-317 and build_volume_packet :1025-1042:
// critical || RecvSlic packets in exponent order || critical.  Packet hashes
// come from the GPU; the headers are assembled here.
int write_volumes(const CreateOptions &opts, const std::vector<uint32_t> &sizes,
                  const std::vector<uint8_t> &critical, const uint8_t set_id[16],
                  const uint8_t *rec, const uint8_t *pkt_md5, std::vector<std::string> &written,
                  Error &err) {
    const uint64_t S = opts.slice_size;
    // Each volume is its own file (written, fsynced and renamed by
    // write_atomic): up to 8 are written at once, and every volume is
    // attempted.  As in src/creator.rs:296-316 (a parallel map over the
    // volumes, then the results in volume order), the first failed volume's
    // error is returned and the volumes written beside it stay on disk.
    const size_t nv = sizes.size();
    std::vector<uint32_t> first(nv);
    for (size_t v = 0, e = 0; v < nv; e += sizes[v], v++) first[v] = (uint32_t)e;
    std::vector<std::string> names(nv);
    std::vector<Error> errs(nv);
    std::vector<int> rcs(nv, P2G_OK);
    auto write_one = [&](size_t v) {
        const uint32_t count = sizes[v], first_exp = first[v];
        std::vector<std::vector<uint8_t>> hdrs(count);
        for (uint32_t k = 0; k < count; k++) {
            const uint32_t exp = first_exp + k;
            auto &h = hdrs[k];
            h.insert(h.end(), kMagic, kMagic + 8);
            put_le64(h, 64 + 4 + S);
            h.insert(h.end(), &pkt_md5[16 * (size_t)exp], &pkt_md5[16 * (size_t)exp] + 16);
            h.insert(h.end(), set_id, set_id + 16);
            h.insert(h.end(), kTypeRecv, kTypeRecv + 16);
            put_le32(h, exp);
        }
        std::vector<std::pair<const uint8_t *, size_t>> parts;
        for (uint32_t k = 0; k < count; k++) {
            parts.push_back({hdrs[k].data(), hdrs[k].size()});
            parts.push_back({rec + (size_t)(first_exp + k) * S, (size_t)S});
        }
        parts.push_back({critical.data(), critical.size()});
        names[v] = volume_name(opts.output, first_exp, count);
        rcs[v] = write_atomic(names[v], {&critical}, parts, errs[v]);
    };
    const size_t nthreads = std::min<size_t>(nv, 8);
    std::vector<std::thread> pool;
    std::mutex mu;
    size_t next = 0;
    for (size_t t = 0; t < nthreads; t++)
        pool.emplace_back([&]() {
            for (;;) {
                size_t v;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    if (next == nv) return;
                    v = next++;
                }
                write_one(v);
            }
        });
    for (auto &th : pool) th.join();
    for (size_t v = 0; v < nv; v++) {
        if (rcs[v] != P2G_OK) {
            err = errs[v];
            return rcs[v];
        }
        written.push_back(names[v]);
    }
    return P2G_OK;
}

}  // namespace

int run_create(p2g_ctx *ctx, const CreateOptions &opts, const std::vector<SourceFile> &sources,
               std::vector<std::string> &written, Error &err) {
    written.clear();
    int rc = check_create(opts, sources, err);
    if (rc != P2G_OK) return rc;
    const uint64_t S = opts.slice_size;
    uint64_t total_blocks = 0;
    for (auto &s : sources) total_blocks += s.slice_checksums.size();
    uint8_t set_id[16];
    std::vector<uint8_t> critical;
    make_critical(S, sources, set_id, critical);

    rc = write_atomic(opts.output, {&critical}, {}, err);
    if (rc != P2G_OK) return rc;
    written.push_back(opts.output);
    const uint32_t R = opts.recovery_block_count;
    if (R == 0) return P2G_OK;

    std::vector<uint32_t> sizes;
    rc = resolve_volume_sizes(opts.volume_scheme, R, sizes, err);
    if (rc != P2G_OK) return rc;  // after the index is written, as src/creator.rs:237 then :244

    // Encode, src/creator.rs:883-1018: every file read again, its slices
    // pushed at their running global input index.
    p2g_job *job = nullptr;
    rc = p2g_encode_begin(ctx, S, (uint32_t)total_blocks, R, &job);
    if (rc != P2G_OK) return set_err(err, rc, std::string("encode: ") + p2g_last_error(ctx));
    const uint64_t chunk = chunk_bytes_for(S);
    Pinned buf;
    uint32_t input_idx = 0;
    for (auto &s : sources) {
        int fd = ::open(s.path.c_str(), O_RDONLY);
        if (fd < 0) {
            p2g_encode_free(job);
            return set_err(err, P2G_ERR_IO, "cannot open " + s.path + ": " + strerror(errno));
        }
        uint64_t done = 0;
        while (done < s.length) {
            const size_t want = (size_t)std::min<uint64_t>(chunk, s.length - done);
            size_t got = 0;
            if (!buf.reserve(want) || !read_full(fd, buf.p, want, got) || got != want) {
                ::close(fd);
                p2g_encode_free(job);
                return set_err(err, P2G_ERR_IO, "short read on " + s.path);
            }
            rc = p2g_encode_push(job, input_idx, buf.p, got);
            if (rc != P2G_OK) {
                ::close(fd);
                p2g_encode_free(job);
                return set_err(err, rc, std::string("encode: ") + p2g_last_error(ctx));
            }
            input_idx += (uint32_t)((got + S - 1) / S);
            done += got;
        }
        ::close(fd);
    }
    Pinned rec;
    if (!rec.reserve((size_t)R * S)) {
        p2g_encode_free(job);
        return set_err(err, P2G_ERR_OUT_OF_MEMORY, "pinned recovery buffer");
    }
    std::vector<uint8_t> pkt_md5((size_t)R * 16);
    rc = p2g_encode_finish(job, rec.p, set_id, pkt_md5.data());
    p2g_encode_free(job);
    if (rc != P2G_OK) return set_err(err, rc, std::string("encode: ") + p2g_last_error(ctx));
    return write_volumes(opts, sizes, critical, set_id, rec.p, pkt_md5.data(), written, err);
}

// ---- single-read create ----------------------------------------------
//
// The fused shape of src/creator.rs:447-700: each file is read once, in
// 64 MiB chunks through a ring of pinned buffers; every chunk is copied to
// the set's HBM arena on a copy stream while a host thread streams it
// through hash_full / hash16k (src/source.rs:223-228).  Once every file is
// resident, one p2g_dev_create_pass computes the slice checksums, the
// recovery blocks and the packet hashes; the index and volumes are written
// as run_create writes them.  The per-file checks run first, in caller
// order and in scan's order (name, slice size, stat, empty), so a set fails
// with the error the scans would report (a read that fails part way is the
// one case reported after later files' checks rather than before).
int create_resident(p2g_ctx *ctx, const CreateOptions &opts, const std::vector<std::string> &paths,
                    const std::vector<std::vector<uint8_t>> &names, std::vector<std::string> &written,
                    Error &err, bool *declined) {
    written.clear();
    *declined = false;
    // P2G_CREATE_VERBOSE: phase times on stderr (read + hash, device pass,
    // read-back, index and volume writes)
    const bool verbose = getenv("P2G_CREATE_VERBOSE") != nullptr;
    const auto t_start = std::chrono::steady_clock::now();
    auto ms_since = [](std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
    };
    const uint64_t S = opts.slice_size;
    const size_t nf = paths.size();
    std::vector<SourceFile> sources(nf);
    for (size_t f = 0; f < nf; f++) {
        const std::vector<uint8_t> &name = names[f];
        if (name.empty() || std::find(name.begin(), name.end(), 0) != name.end())
            return set_err(err, P2G_ERR_INVALID_FILE_NAME,
                           "filename " + paths[f] + " is empty or contains a NUL byte");
        if (S == 0 || S % 4 != 0)
            return set_err(err, P2G_ERR_INVALID_SLICE_SIZE, "invalid slice size " + std::to_string(S));
        struct stat st;
        if (::stat(paths[f].c_str(), &st) != 0)
            return set_err(err, P2G_ERR_IO, "cannot stat " + paths[f] + ": " + strerror(errno));
        if (st.st_size == 0)
            return set_err(err, P2G_ERR_EMPTY_FILE, "input file " + paths[f] + " is empty");
        sources[f].name = name;
        sources[f].path = paths[f];
        sources[f].length = (uint64_t)st.st_size;
    }
    int rc = check_create(opts, sources, err);
    if (rc != P2G_OK) return rc;
    const uint32_t R = opts.recovery_block_count;
    std::vector<uint32_t> sizes;
    Error scheme_err;
    const int scheme_rc = R ? resolve_volume_sizes(opts.volume_scheme, R, sizes, scheme_err) : P2G_OK;

    // HBM for the whole set: files (256-byte aligned), recovery blocks,
    // digests; declined (the caller streams instead) when it would not
    // leave room for the pass's own buffers.
    uint64_t n = 0, arena = 0;
    std::vector<uint64_t> off(nf);
    for (size_t f = 0; f < nf; f++) {
        off[f] = arena;
        arena += (sources[f].length + 255) / 256 * 256;
        n += (sources[f].length + S - 1) / S;
    }
    const uint64_t out_bytes = (uint64_t)R * S + 20 * n + 16ull * R + 1024;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess ||
        arena + out_bytes + (8ull << 30) > (uint64_t)free_b) {
        *declined = true;
        return P2G_OK;
    }
    uint8_t *dev = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&dev), arena + out_bytes) != hipSuccess) {
        *declined = true;
        return P2G_OK;
    }
    struct DevFree {
        uint8_t *p;
        ~DevFree() { hipFree(p); }
    } dev_free{dev};
    uint8_t *d_rec = dev + arena;
    uint8_t *d_md5 = d_rec + (uint64_t)R * S;
    uint32_t *d_crc = reinterpret_cast<uint32_t *>(d_md5 + 16 * n);
    uint8_t *d_pkt = reinterpret_cast<uint8_t *>(d_crc + n);

    // Read once: a ring of pinned buffers -> HBM, and the whole-file hashes
    // on host threads.  A file's chunks are hashed in order by one thread
    // (the one with the fewest bytes pending when the file's first chunk is
    // read), so different files hash in parallel; a buffer is refilled once
    // its copy and its hashing are done.
    const int H = (int)std::min<size_t>(nf, 6);  // hashing threads
    const int kSlots = H > 1 ? H + 2 : 4;        // pinned 64 MiB buffers
    std::vector<Pinned> slot(kSlots);
    std::vector<hipEvent_t> copied(kSlots, nullptr);
    std::vector<char> busy(kSlots, 0);  // filled and not yet hashed
    hipStream_t copy = nullptr;
    hipError_t he = hipStreamCreateWithFlags(&copy, hipStreamNonBlocking);
    for (int k = 0; k < kSlots && he == hipSuccess; k++)
        he = hipEventCreateWithFlags(&copied[k], hipEventDisableTiming);
    struct Cleanup {
        hipStream_t &s;
        std::vector<hipEvent_t> &e;
        ~Cleanup() {
            if (s) hipStreamSynchronize(s);
            for (hipEvent_t ev : e)
                if (ev) hipEventDestroy(ev);
            if (s) hipStreamDestroy(s);
        }
    } cleanup{copy, copied};
    if (he != hipSuccess) return set_err(err, P2G_ERR_DEVICE, "copy stream");
    struct Item {
        int slot;
        size_t f, got;
        uint64_t at;
    };
    std::vector<HostMd5> full(nf), first16k(nf);
    std::mutex mu;
    std::condition_variable cv;
    std::vector<std::deque<Item>> queue(H);
    std::vector<uint64_t> pending(H, 0);  // bytes queued or being hashed per thread
    bool stop = false;
    std::vector<std::thread> hashers;
    for (int t = 0; t < H; t++)
        hashers.emplace_back([&, t]() {
            for (;;) {
                Item it;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return stop || !queue[t].empty(); });
                    if (queue[t].empty()) return;
                    it = queue[t].front();
                    queue[t].pop_front();
                }
                const uint8_t *p = slot[it.slot].p;
                full[it.f].update(p, it.got);
                if (it.at < 16384)
                    first16k[it.f].update(p, (size_t)std::min<uint64_t>(it.got, 16384 - it.at));
                {
                    std::lock_guard<std::mutex> lk(mu);
                    busy[it.slot] = 0;
                    pending[t] -= it.got;
                }
                cv.notify_all();
            }
        });
    auto finish_hashers = [&]() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_all();
        for (auto &th : hashers) th.join();
    };
    const uint64_t chunk = kChunkBytes;
    int next = 0;  // slot to try first
    for (size_t f = 0; f < nf && rc == P2G_OK; f++) {
        int fd = ::open(paths[f].c_str(), O_RDONLY);
        if (fd < 0) {
            rc = set_err(err, P2G_ERR_IO, "cannot open " + paths[f] + ": " + strerror(errno));
            break;
        }
        int owner = 0;  // the hashing thread of this file
        {
            std::lock_guard<std::mutex> lk(mu);
            for (int t = 1; t < H; t++)
                if (pending[t] < pending[owner]) owner = t;
        }
        for (uint64_t at = 0; at < sources[f].length; at += chunk) {
            int k = -1;
            {
                // a buffer whose previous chunk is hashed (its copy is waited below)
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] {
                    for (int j = 0; j < kSlots; j++)
                        if (!busy[(next + j) % kSlots]) {
                            k = (next + j) % kSlots;
                            return true;
                        }
                    return false;
                });
            }
            next = (k + 1) % kSlots;
            if (hipEventSynchronize(copied[k]) != hipSuccess) {
                rc = set_err(err, P2G_ERR_DEVICE, "chunk copy");
                break;
            }
            const size_t want = (size_t)std::min<uint64_t>(chunk, sources[f].length - at);
            size_t got = 0;
            if (!slot[k].reserve((size_t)chunk) || !read_full(fd, slot[k].p, want, got) ||
                got != want) {
                rc = set_err(err, P2G_ERR_IO, "short read on " + paths[f]);
                break;
            }
            if (hipMemcpyAsync(dev + off[f] + at, slot[k].p, got, hipMemcpyHostToDevice, copy) !=
                    hipSuccess ||
                hipEventRecord(copied[k], copy) != hipSuccess) {
                rc = set_err(err, P2G_ERR_DEVICE, "chunk copy");
                break;
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                busy[k] = 1;
                pending[owner] += got;
                queue[owner].push_back({k, f, got, at});
            }
            cv.notify_all();
        }
        ::close(fd);
    }
    finish_hashers();
    if (rc != P2G_OK) return rc;
    const double ms_read = ms_since(t_start);
    if (hipStreamSynchronize(copy) != hipSuccess) return set_err(err, P2G_ERR_DEVICE, "chunk copies");
    for (size_t f = 0; f < nf; f++) {
        SourceFile &s = sources[f];
        full[f].finish(s.hash_full);
        first16k[f].finish(s.hash16k);
        HostMd5 id;  // compute_file_id, src/source.rs:303-309
        id.update(s.hash16k, 16);
        uint8_t l[8];
        for (int i = 0; i < 8; i++) l[i] = (uint8_t)(s.length >> (8 * i));
        id.update(l, 8);
        id.update(s.name.data(), s.name.size());
        id.finish(s.file_id);
    }
    uint8_t set_id[16];
    std::vector<uint8_t> body;
    main_body(S, sources, body, set_id);

    // one device pass: slice checksums, recovery blocks, packet hashes
    std::vector<p2g_dev_file> files(nf);
    for (size_t f = 0; f < nf; f++) files[f] = {dev + off[f], sources[f].length};
    const auto t_pass = std::chrono::steady_clock::now();
    rc = p2g_dev_create_pass(ctx, files.data(), (uint32_t)nf, S, R, set_id, d_rec, d_md5, d_crc,
                             d_pkt, nullptr);
    if (rc == P2G_OK) rc = p2g_synchronize(ctx);
    if (rc != P2G_OK) return set_err(err, rc, std::string("create pass: ") + p2g_last_error(ctx));
    std::vector<uint8_t> md5s(16 * n), pkt_md5(16ull * R);
    std::vector<uint32_t> crcs(n);
    Pinned rec;
    if (R && !rec.reserve((size_t)R * S)) return set_err(err, P2G_ERR_OUT_OF_MEMORY, "pinned recovery buffer");
    he = hipMemcpy(md5s.data(), d_md5, 16 * n, hipMemcpyDeviceToHost);
    if (he == hipSuccess) he = hipMemcpy(crcs.data(), d_crc, 4 * n, hipMemcpyDeviceToHost);
    if (he == hipSuccess && R) he = hipMemcpy(pkt_md5.data(), d_pkt, 16ull * R, hipMemcpyDeviceToHost);
    if (he == hipSuccess && R) he = hipMemcpy(rec.p, d_rec, (size_t)R * S, hipMemcpyDeviceToHost);
    if (he != hipSuccess) return set_err(err, P2G_ERR_DEVICE, "read back the create pass");
    const double ms_pass = ms_since(t_pass);  // pass and read-back
    const auto t_write = std::chrono::steady_clock::now();
    for (size_t f = 0, sl = 0; f < nf; f++) {
        SourceFile &s = sources[f];
        s.slice_checksums.resize((s.length + S - 1) / S);
        for (auto &c : s.slice_checksums) {
            memcpy(c.md5, &md5s[16 * sl], 16);
            c.crc32 = crcs[sl++];
        }
    }
    std::vector<uint8_t> critical;
    make_critical(S, sources, set_id, critical);
    rc = write_atomic(opts.output, {&critical}, {}, err);
    if (rc != P2G_OK) return rc;
    written.push_back(opts.output);
    if (R == 0) return P2G_OK;
    if (scheme_rc != P2G_OK) {  // after the index is written, as src/creator.rs:237 then :244
        err = scheme_err;
        return scheme_rc;
    }
    rc = write_volumes(opts, sizes, critical, set_id, rec.p, pkt_md5.data(), written, err);
    if (verbose)
        fprintf(stderr, "create: read+hash %.1f ms, pass+read-back %.1f ms, write %.1f ms, total %.1f ms\n",
                ms_read, ms_pass, ms_since(t_write), ms_since(t_start));
    return rc;
}

}  // namespace par2

// =====================================================================
// extern "C" p2g_run_create
// =====================================================================

extern "C" int p2g_run_create(p2g_ctx *ctx, const char *output, uint64_t slice_size,
                              uint32_t recovery_count, const p2g_volume_scheme *scheme,
                              uint32_t nfiles, const char *const *paths,
                              const uint8_t *const *names, const uint32_t *name_lens,
                              char *names_out, size_t names_cap) {
    return p2g::guarded(reinterpret_cast<p2g::Ctx *>(ctx), [&]() -> int {
        if (!ctx || !output || !scheme || (nfiles && (!paths || !names || !name_lens)))
            return P2G_ERR_INVALID_ARGUMENT;
        if (names_out && names_cap) names_out[0] = 0;
        par2::Error err;
        par2::CreateOptions opts;
        opts.output = output;
        opts.slice_size = slice_size;
        opts.recovery_block_count = recovery_count;
        opts.volume_scheme.kind = scheme->kind;
        opts.volume_scheme.uniform_count = scheme->uniform_count;
        if (scheme->limited) opts.volume_scheme.caps.push_back(scheme->limit_cap);
        if (scheme->n_outer_caps) {
            if (!scheme->limited || !scheme->outer_caps) return P2G_ERR_INVALID_ARGUMENT;
            opts.volume_scheme.caps.insert(opts.volume_scheme.caps.end(), scheme->outer_caps,
                                           scheme->outer_caps + scheme->n_outer_caps);
        }
        if (scheme->explicit_sizes)
            opts.volume_scheme.explicit_sizes.assign(scheme->explicit_sizes,
                                                     scheme->explicit_sizes + scheme->n_explicit);
        std::vector<std::string> written;
        std::vector<std::string> path_v(paths, paths + nfiles);
        std::vector<std::vector<uint8_t>> name_v(nfiles);
        for (uint32_t f = 0; f < nfiles; f++) name_v[f].assign(names[f], names[f] + name_lens[f]);
        // single read with the set resident in HBM; the streaming scan + encode
        // when the set does not fit (or P2G_CREATE_STREAM=1)
        bool declined = true;
        int rc = P2G_OK;
        const char *stream_env = getenv("P2G_CREATE_STREAM");
        if (!(stream_env && atoi(stream_env)))
            rc = par2::create_resident(ctx, opts, path_v, name_v, written, err, &declined);
        if (declined) {
            std::vector<par2::SourceFile> sources(nfiles);
            for (uint32_t f = 0; f < nfiles && rc == P2G_OK; f++)
                rc = par2::scan(ctx, paths[f], name_v[f], slice_size, sources[f], err);
            if (rc == P2G_OK) rc = par2::run_create(ctx, opts, sources, written, err);
        }
        size_t used = 0;
        for (auto &w : written) {
            if (names_out && used + w.size() + 1 <= names_cap) {
                memcpy(names_out + used, w.c_str(), w.size() + 1);
                used += w.size() + 1;
            }
        }
        if (rc != P2G_OK) reinterpret_cast<p2g::Ctx *>(ctx)->err = err.message;
        return rc;
    });
}
