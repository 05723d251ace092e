// p2g_abi.hip -- context, device memory and the extern "C" entry points of
// include/par2gpu.h (p2g_run_create lives in par2_create.cpp).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "p2g_context.hpp"
#include "p2g_internal.hpp"

using namespace p2g;

namespace p2g {

// ---------------------------------------------------------------------
// GF(2^16) tables on the host: build_tables, src/galois.rs:33-53.
// ---------------------------------------------------------------------
static void build_gf_tables(std::vector<uint16_t> &log, std::vector<uint16_t> &antilog) {
    log.assign(65536, 0);
    antilog.assign(65536, 0);
    uint32_t b = 1;
    for (uint32_t l = 0; l < kFieldLimit; l++) {
        log[b] = (uint16_t)l;
        antilog[l] = (uint16_t)b;
        b <<= 1;
        if (b & 0x10000u) b ^= kFieldGenerator;
    }
    log[0] = (uint16_t)kFieldLimit;
    antilog[kFieldLimit] = 0;
}

// Logs of the RS input bases: the non-negative integers coprime to 65535 in
// increasing order, generate_input_bases, src/reedsolomon.rs:130-149
// (gcd(65535, 0) = 65535, so 0 is not one of them).
static std::vector<uint16_t> input_base_logs() {
    std::vector<uint16_t> out;
    out.reserve(kMaxInputs);
    for (uint32_t l = 0; l < kFieldLimit && out.size() < kMaxInputs; l++) {
        uint32_t a = kFieldLimit, b = l;
        while (b) { uint32_t t = a % b; a = b; b = t; }
        if (a == 1) out.push_back((uint16_t)l);
    }
    return out;
}

int Ctx::fail(int code, const char *what, hipError_t e) {
    char buf[512];
    if (e != hipSuccess)
        snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    else
        snprintf(buf, sizeof buf, "%s", what);
    err = buf;
    return code;
}

int DevBuf::reserve(size_t n) {
    if (n <= cap) return P2G_OK;
    // Work queued earlier may still read the old allocation.
    if (ptr) { hipDeviceSynchronize(); hipFree(ptr); }
    ptr = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(n, 256);
    if (hipMalloc(&ptr, want) != hipSuccess) { ptr = nullptr; return P2G_ERR_OUT_OF_MEMORY; }
    cap = want;
    return P2G_OK;
}

void DevBuf::release() {
    if (ptr) hipFree(ptr);
    ptr = nullptr;
    cap = 0;
}

int PinBuf::reserve(size_t n) {
    if (n <= cap) return P2G_OK;
    // An asynchronous copy queued earlier may still read the old allocation.
    if (ptr) { hipDeviceSynchronize(); hipHostFree(ptr); }
    ptr = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(n, 256);
    if (hipHostMalloc(&ptr, want, hipHostMallocDefault) != hipSuccess) {
        ptr = nullptr;
        return P2G_ERR_OUT_OF_MEMORY;
    }
    cap = want;
    return P2G_OK;
}

void PinBuf::release() {
    if (ptr) hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
}

int Ctx::open(int dev) {
    device = dev;
    hipError_t e = hipSetDevice(dev);
    if (e != hipSuccess) return fail(P2G_ERR_DEVICE, "hipSetDevice", e);
    e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(P2G_ERR_DEVICE, "hipStreamCreate", e);
    e = hipEventCreateWithFlags(&enc_descs.free, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sum_descs.free, hipEventDisableTiming);
    for (hipEvent_t &ev : scratch_done)
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&noalloc_done, hipEventDisableTiming);
    if (e != hipSuccess) return fail(P2G_ERR_DEVICE, "hipEventCreate", e);
    if (noalloc_scratch_init(dev, &noalloc_scratch) != P2G_OK)
        return fail(P2G_ERR_DEVICE, "address of the allocation-free scratch array");
    build_gf_tables(h_log, h_antilog);
    base_logs = input_base_logs();
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess || cus <= 0) return fail(P2G_ERR_DEVICE, "compute unit count", e);
    num_cus = (uint32_t)cus;
    {
        const int po = pass_occupancy(), bo = batch_occupancy();
        if (po <= 0 || bo <= 0)
            return fail(P2G_ERR_DEVICE, "occupancy of the persistent kernels could not be computed");
        pass_resident = (uint32_t)po * num_cus;
        batch_resident = (uint32_t)bo * num_cus;
    }
    if (const char *v = getenv("P2G_BAND_COLS")) band_cols = (uint32_t)strtoul(v, nullptr, 10);
    if (const char *v = getenv("P2G_AUTOTUNE")) autotune = atoi(v) != 0;
    if (const char *v = getenv("P2G_TUNE_CACHE")) {
        tune_path = v;
        if (FILE *f = fopen(v, "r")) {  // a missing file is created at the first tuning
            fclose(f);
            int rc = tune_load(this, v);
            if (rc != P2G_OK) return rc;
        }
    }
    if (const char *v = getenv("P2G_ENCODE")) {
        if (!strcmp(v, "bitslice")) encode_variant = kEncodeBitslice;
        else if (!strcmp(v, "nibble")) encode_variant = kEncodeNibble;
        else if (!strcmp(v, "mfma")) encode_variant = kEncodeMfma;
    }
    if (d_antilog.reserve(65536 * 2) != P2G_OK) return fail(P2G_ERR_OUT_OF_MEMORY, "antilog table");
    e = hipMemcpy(d_antilog.ptr, h_antilog.data(), 65536 * 2, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(P2G_ERR_DEVICE, "upload antilog", e);
    if (d_base_logs.reserve(base_logs.size() * 2) != P2G_OK)
        return fail(P2G_ERR_OUT_OF_MEMORY, "base logs");
    e = hipMemcpy(d_base_logs.ptr, base_logs.data(), base_logs.size() * 2, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(P2G_ERR_DEVICE, "upload base logs", e);
    if (d_err.reserve(256) != P2G_OK) return fail(P2G_ERR_OUT_OF_MEMORY, "error word");
    e = hipMemset(d_err.ptr, 0, 256);
    if (e != hipSuccess) return fail(P2G_ERR_DEVICE, "clear error word", e);
    return P2G_OK;
}

Ctx::~Ctx() {
    hipSetDevice(device);
    if (stream) hipStreamSynchronize(stream);
    if (side) hipStreamSynchronize(side);
    for (BatchPlan *b : pending_batches) free_batch_plan(b);
    d_antilog.release();
    d_base_logs.release();
    d_mats.release();
    d_md5.release();
    d_crc.release();
    d_stage.release();
    d_sync.release();
    d_err.release();
    d_partials.release();
    d_xs_jumps.release();
    for (hipEvent_t ev : scratch_done)
        if (ev) hipEventDestroy(ev);
    if (noalloc_done) hipEventDestroy(noalloc_done);
    for (DescSlot *sl : {&enc_descs, &sum_descs}) {
        sl->h.release();
        sl->d.release();
        sl->tb.release();
        if (sl->free) hipEventDestroy(sl->free);
    }
    if (stream) hipStreamDestroy(stream);
    if (side) hipStreamDestroy(side);
}

int Ctx::side_stream(hipStream_t *out) {
    // Created on first use: hardware queues are handed out to streams in
    // creation order, and a context that never streams copies should not
    // take one from the caller's streams.
    if (!side) {
        hipError_t e = hipStreamCreateWithFlags(&side, hipStreamNonBlocking);
        if (e != hipSuccess) {
            side = nullptr;
            return fail(P2G_ERR_DEVICE, "hipStreamCreate", e);
        }
    }
    *out = side;
    return P2G_OK;
}

int Ctx::scratch_acquire(Scratch which, hipStream_t st) {
    hipError_t e = hipStreamWaitEvent(st, scratch_done[which], 0);
    return e == hipSuccess ? P2G_OK : fail(P2G_ERR_DEVICE, "stream wait", e);
}

int Ctx::scratch_release(Scratch which, hipStream_t st) {
    hipError_t e = hipEventRecord(scratch_done[which], st);
    return e == hipSuccess ? P2G_OK : fail(P2G_ERR_DEVICE, "event record", e);
}

int check_slice_size(Ctx *c, uint64_t S) {
    if (S == 0 || S % 4 != 0)
        return c->fail(P2G_ERR_INVALID_SLICE_SIZE, "slice size must be > 0 and a multiple of 4");
    if (S > kMaxSliceBytes)
        return c->fail(P2G_ERR_INVALID_ARGUMENT, "slice size above the 2 GiB kernel limit");
    return P2G_OK;
}

int append_descs(Ctx *c, const p2g_dev_file *files, uint32_t nfiles, uint64_t S,
                 std::vector<SliceDesc> &out) {
    for (uint32_t f = 0; f < nfiles; f++) {
        if (files[f].len && !files[f].data)
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "null device file pointer");
        if (reinterpret_cast<uintptr_t>(files[f].data) % 16 != 0)
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "device file pointers must be 16-byte aligned");
        const uint64_t n = (files[f].len + S - 1) / S;
        if (out.size() + n > 0x7FFFFFFFull)
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "too many slices in one call");
        for (uint64_t s = 0; s < n; s++) {
            const uint64_t have = files[f].len - s * S;
            out.push_back({files[f].data + s * S, (uint32_t)(have < S ? have : S), 0});
        }
    }
    return P2G_OK;
}

int upload_descs(Ctx *c, DescSlot &slot, const std::vector<SliceDesc> &descs, hipStream_t st) {
    const size_t bytes = descs.size() * sizeof(SliceDesc);
    if (bytes == 0) return P2G_OK;
    // The previous upload out of slot.h must have been consumed first.
    hipError_t e = hipEventSynchronize(slot.free);
    if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "descriptor staging wait", e);
    if (slot.h.reserve(bytes) != P2G_OK || slot.d.reserve(bytes) != P2G_OK)
        return c->fail(P2G_ERR_OUT_OF_MEMORY, "descriptor buffers");
    memcpy(slot.h.ptr, descs.data(), bytes);
    e = hipMemcpyAsync(slot.d.ptr, slot.h.ptr, bytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(slot.free, st);
    if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "upload slice descriptors", e);
    return P2G_OK;
}

int dev_encode(Ctx *c, const SliceDesc *d_descs, uint32_t n, uint64_t S, uint32_t first_input,
               uint32_t total_inputs, uint32_t R, uint8_t *d_rec, int accumulate,
               hipStream_t st, int variant) {
    if (variant < 0) {
        // the bucket's tuned encode kernel, else the context's
        TuneChoice ch;
        variant = tune_lookup(c, total_inputs, R, S, &ch) && ch.enc_variant >= 0 ? ch.enc_variant
                                                                                 : c->encode_variant;
    }
    if (R > P2G_MAX_RECOVERY_BLOCKS)
        return c->fail(P2G_ERR_TOO_MANY_RECOVERY_BLOCKS, "recovery count above 65535");
    if (total_inputs > kMaxInputs)
        return c->fail(P2G_ERR_TOO_MANY_INPUT_BLOCKS, "more than 32768 input slices");
    if ((uint64_t)first_input + n > total_inputs)
        return c->fail(P2G_ERR_INVALID_ARGUMENT, "input index range exceeds total_inputs");
    if (R == 0) return P2G_OK;
    hipError_t e;
    if (n == 0) {
        if (!accumulate) {
            e = hipMemsetAsync(d_rec, 0, (size_t)R * S, st);
            if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "hipMemsetAsync", e);
        }
        return P2G_OK;
    }
    const uint16_t *dlog = static_cast<const uint16_t *>(c->d_base_logs.ptr) + first_input;
    const uint16_t *dalog = static_cast<const uint16_t *>(c->d_antilog.ptr);
    if (variant == kEncodeBitslice) {
        // Inputs go in chunks so the coefficient matrices (32 B per (r, i))
        // stay within kMatsBudget.
        uint64_t ch = kMatsBudget / ((uint64_t)R * 32);
        ch = std::max<uint64_t>(1, std::min<uint64_t>(ch, n));
        if (c->d_mats.reserve((size_t)ch * R * 32) != P2G_OK)
            return c->fail(P2G_ERR_OUT_OF_MEMORY, "coefficient matrices");
        for (uint64_t c0 = 0; c0 < n; c0 += ch) {
            const uint32_t cn = (uint32_t)std::min<uint64_t>(ch, n - c0);
            e = launch_coeff_mats(dalog, dlog + c0, cn, R, static_cast<uint16_t *>(c->d_mats.ptr), st);
            if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "coefficient kernel", e);
            e = launch_encode(d_descs + c0, cn, static_cast<const uint16_t *>(c->d_mats.ptr), R,
                              (uint32_t)S, d_rec, (c0 > 0 || accumulate) ? 1 : 0, st);
            if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "encode kernel", e);
        }
        return P2G_OK;
    }
    if (variant == kEncodeNibble) {
        // Inputs go in chunks so the byte tables (96 B per (r, i)) stay
        // within kMatsBudget.
        uint64_t ch = kMatsBudget / ((uint64_t)R * kNibbleTableBytes);
        ch = std::max<uint64_t>(1, std::min<uint64_t>(ch, n));
        if (c->d_mats.reserve((size_t)ch * R * kNibbleTableBytes) != P2G_OK)
            return c->fail(P2G_ERR_OUT_OF_MEMORY, "split-nibble tables");
        for (uint64_t c0 = 0; c0 < n; c0 += ch) {
            const uint32_t cn = (uint32_t)std::min<uint64_t>(ch, n - c0);
            uint32_t *tables = static_cast<uint32_t *>(c->d_mats.ptr);
            e = launch_nibble_tables(dalog, dlog + c0, cn, R, tables, st);
            if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "split-nibble table kernel", e);
            e = launch_encode_nibble(d_descs + c0, cn, tables, R, (uint32_t)S, d_rec,
                                     (c0 > 0 || accumulate) ? 1 : 0, st);
            if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "encode kernel", e);
        }
        return P2G_OK;
    }
    e = launch_desc_tails(const_cast<SliceDesc *>(d_descs), n, nullptr, st);
    if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "descriptor tail kernel", e);
    // Matrix-core form: A fragments (1 KiB per 32 rows x 4 inputs) for a chunk
    // of inputs, a multiple of 4, within kMatsBudget.
    const uint64_t per_quad = (uint64_t)mfma_row_tiles(R) * 1024;
    uint64_t ch = std::max<uint64_t>(1, kMatsBudget / per_quad) * 4;
    ch = std::min<uint64_t>(std::min<uint64_t>(ch, kMfmaMaxInputsPerLaunch), n);
    if (c->d_mats.reserve(mfma_afrag_bytes((uint32_t)ch, R)) != P2G_OK)
        return c->fail(P2G_ERR_OUT_OF_MEMORY, "matrix-core coefficient fragments");
    for (uint64_t c0 = 0; c0 < n; c0 += ch) {
        const uint32_t cn = (uint32_t)std::min<uint64_t>(ch, n - c0);
        uint4 *afrag = static_cast<uint4 *>(c->d_mats.ptr);
        e = launch_mfma_afrag(dalog, dlog + c0, cn, R, afrag, st);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "coefficient fragment kernel", e);
        e = launch_encode_mfma(d_descs + c0, cn, afrag, R, (uint32_t)S, d_rec,
                               (c0 > 0 || accumulate) ? 1 : 0, st);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "encode kernel", e);
    }
    return P2G_OK;
}

int dev_encode_files(Ctx *c, const p2g_dev_file *files, uint32_t nfiles, uint64_t slice_size,
                     uint32_t first_input_index, uint32_t total_inputs, uint32_t recovery_count,
                     uint8_t *rec, int accumulate, hipStream_t st, int variant) {
    if ((nfiles && !files) || (recovery_count && !rec))
        return c->fail(P2G_ERR_INVALID_ARGUMENT, "null argument");
    if (reinterpret_cast<uintptr_t>(rec) % 4 != 0)
        return c->fail(P2G_ERR_INVALID_ARGUMENT, "recovery buffer must be 4-byte aligned");
    int rc = check_slice_size(c, slice_size);
    if (rc != P2G_OK) return rc;
    hipSetDevice(c->device);
    std::vector<SliceDesc> descs;
    rc = append_descs(c, files, nfiles, slice_size, descs);
    if (rc != P2G_OK) return rc;
    if (descs.size() > kMaxInputs)
        return c->fail(P2G_ERR_TOO_MANY_INPUT_BLOCKS, "more than 32768 input slices");
    rc = c->scratch_acquire(Ctx::kScratchEnc, st);
    if (rc == P2G_OK) rc = upload_descs(c, c->enc_descs, descs, st);
    if (rc == P2G_OK)
        rc = dev_encode(c, static_cast<const SliceDesc *>(c->enc_descs.d.ptr), (uint32_t)descs.size(),
                        slice_size, first_input_index, total_inputs, recovery_count, rec, accumulate,
                        st, variant);
    if (rc != P2G_OK) return rc;
    return c->scratch_release(Ctx::kScratchEnc, st);
}

// GF(2) 64x64 matrices as 64 column images.
static uint64_t xs_apply_host(const uint64_t *m, uint64_t v) {
    uint64_t out = 0;
    for (int b = 0; b < 64; b++)
        if ((v >> b) & 1) out ^= m[b];
    return out;
}

// The jump tables of the synthetic-input fill (built and uploaded once per
// context): c->d_xs_jumps.
int xs_jump_tables(Ctx *c) {
    if (!c->d_xs_jumps.ptr) {
        // step^(W 2^b) for b = 0..31, W = kXorshiftLaneWords
        std::vector<uint64_t> step(64), m(64), t(64), jumps(64 * 32);
        for (int b = 0; b < 64; b++) {
            uint64_t x = 1ull << b;
            x ^= x << 13;
            x ^= x >> 7;
            x ^= x << 17;
            step[b] = x;
        }
        m = step;  // step^(2^k), squared until 2^k = W
        for (uint32_t w = 1; w < kXorshiftLaneWords; w *= 2) {
            for (int b = 0; b < 64; b++) t[b] = xs_apply_host(m.data(), m[b]);
            m = t;
        }
        for (int k = 0; k < 32; k++) {
            std::copy(m.begin(), m.end(), jumps.begin() + 64 * k);
            for (int b = 0; b < 64; b++) t[b] = xs_apply_host(m.data(), m[b]);
            m = t;
        }
        // each jump as 8 byte tables: entry v of table q of jump k is the
        // XOR of the columns 8 q + i of jumps[k] for the set bits i of v
        std::vector<uint64_t> tables((size_t)32 * kXsJumpTableWords);
        for (int k = 0; k < 32; k++)
            for (int q = 0; q < 8; q++)
                for (int v = 0; v < 256; v++) {
                    uint64_t x = 0;
                    for (int i = 0; i < 8; i++)
                        if ((v >> i) & 1) x ^= jumps[64 * k + 8 * q + i];
                    tables[(size_t)k * kXsJumpTableWords + 256 * q + v] = x;
                }
        if (c->d_xs_jumps.reserve(tables.size() * 8) != P2G_OK)
            return c->fail(P2G_ERR_OUT_OF_MEMORY, "xorshift jump table");
        hipError_t e = hipMemcpy(c->d_xs_jumps.ptr, tables.data(), tables.size() * 8,
                                 hipMemcpyHostToDevice);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "upload xorshift jumps", e);
    }
    return P2G_OK;
}

}  // namespace p2g

// =====================================================================
// extern "C"
// =====================================================================

extern "C" {

const char *p2g_version(void) { return "par2gpu 0.1.0 gfx950"; }

int p2g_open(int device, p2g_ctx **out) {
    return guarded(nullptr, [&]() -> int {
        if (!out) return P2G_ERR_INVALID_ARGUMENT;
        *out = nullptr;
        Ctx *c = new (std::nothrow) Ctx();
        if (!c) return P2G_ERR_OUT_OF_MEMORY;
        int rc = c->open(device);
        if (rc != P2G_OK) {
            fprintf(stderr, "p2g_open: %s\n", c->err.c_str());
            delete c;
            return rc;
        }
        *out = reinterpret_cast<p2g_ctx *>(c);
        return P2G_OK;
    });
}

void p2g_close(p2g_ctx *ctx) { delete reinterpret_cast<Ctx *>(ctx); }

const char *p2g_last_error(const p2g_ctx *ctx) {
    return ctx ? reinterpret_cast<const Ctx *>(ctx)->err.c_str() : "null context";
}

int p2g_dev_fill_xorshift64(p2g_ctx *ctx, uint8_t *dst, uint64_t nbytes, uint64_t seed,
                            void *stream) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c) return P2G_ERR_INVALID_ARGUMENT;
        if (nbytes == 0) return P2G_OK;
        if (!dst || reinterpret_cast<uintptr_t>(dst) % 16 != 0)
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "destination must be 16-byte aligned");
        hipSetDevice(c->device);
        const int rc = xs_jump_tables(c);
        if (rc != P2G_OK) return rc;
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
        hipError_t e = launch_xorshift_fill(dst, nbytes, seed,
                                            static_cast<const uint64_t *>(c->d_xs_jumps.ptr), st);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "xorshift fill kernel", e);
        return P2G_OK;
    });
}

int p2g_synchronize(p2g_ctx *ctx) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c) return P2G_ERR_INVALID_ARGUMENT;
        hipSetDevice(c->device);
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess && c->side) e = hipStreamSynchronize(c->side);
        for (hipEvent_t ev : c->scratch_done)  // work queued on callers' streams
            if (e == hipSuccess) e = hipEventSynchronize(ev);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "synchronize", e);
        int brc = P2G_OK;
        for (BatchPlan *b : c->pending_batches) {
            const int rc = wait_plan(b);
            if (brc == P2G_OK) brc = rc;
            free_batch_plan(b);
        }
        c->pending_batches.clear();
        if (brc != P2G_OK) return brc;
        if (c->pass_pending) {
            // every pass queued since the last wait ORs its errors into the
            // sticky word; it is cleared only here, after all of them ended
            c->pass_pending = false;
            uint32_t flags = 0;
            e = hipMemcpy(&flags, c->d_err.ptr, 4, hipMemcpyDeviceToHost);
            if (e == hipSuccess && flags) e = hipMemset(c->d_err.ptr, 0, 4);
            if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "read create-pass flags", e);
            if (flags & kPassErrWait)
                return c->fail(P2G_ERR_DEVICE,
                               "create pass: a packet hash stopped waiting for the encode (no progress for 2 s)");
        }
        return P2G_OK;
    });
}

int p2g_dev_encode(p2g_ctx *ctx, const p2g_dev_file *files, uint32_t nfiles, uint64_t slice_size,
                   uint32_t first_input_index, uint32_t total_inputs, uint32_t recovery_count,
                   uint8_t *rec, int accumulate, void *stream) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c) return P2G_ERR_INVALID_ARGUMENT;
        return dev_encode_files(c, files, nfiles, slice_size, first_input_index, total_inputs,
                                recovery_count, rec, accumulate,
                                stream ? static_cast<hipStream_t>(stream) : c->stream, -1);
    });
}

int p2g_dev_slice_checksums(p2g_ctx *ctx, const p2g_dev_file *files, uint32_t nfiles,
                            uint64_t slice_size, uint8_t *md5_out, uint32_t *crc_out,
                            void *stream) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c) return P2G_ERR_INVALID_ARGUMENT;
        if ((nfiles && !files) || !md5_out || !crc_out)
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "null argument");
        if (reinterpret_cast<uintptr_t>(md5_out) % 4 != 0 || reinterpret_cast<uintptr_t>(crc_out) % 4 != 0)
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "checksum outputs must be 4-byte aligned");
        int rc = check_slice_size(c, slice_size);
        if (rc != P2G_OK) return rc;
        hipSetDevice(c->device);
        std::vector<SliceDesc> descs;
        rc = append_descs(c, files, nfiles, slice_size, descs);
        if (rc != P2G_OK) return rc;
        if (descs.empty()) return P2G_OK;
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
        rc = c->scratch_acquire(Ctx::kScratchSum, st);
        if (rc == P2G_OK) rc = upload_descs(c, c->sum_descs, descs, st);
        if (rc != P2G_OK) return rc;
        SliceDesc *dd = static_cast<SliceDesc *>(c->sum_descs.d.ptr);
        const uint32_t n = (uint32_t)descs.size();
        if (c->sum_descs.tb.reserve((size_t)n * 64) != P2G_OK)
            return c->fail(P2G_ERR_OUT_OF_MEMORY, "slice tail blocks");
        uint4 *tb = static_cast<uint4 *>(c->sum_descs.tb.ptr);
        hipError_t e = launch_desc_tails(dd, n, tb, st);
        if (e == hipSuccess) e = launch_slice_md5(dd, tb, n, (uint32_t)slice_size, md5_out, st);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "slice md5 kernel", e);
        e = launch_slice_crc32(dd, n, (uint32_t)slice_size, crc_out, st);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "slice crc kernel", e);
        return c->scratch_release(Ctx::kScratchSum, st);
    });
}

int p2g_dev_recovery_packet_md5(p2g_ctx *ctx, const uint8_t *rec, uint32_t count,
                                uint64_t slice_size, uint32_t first_exponent,
                                const uint8_t set_id[16], uint8_t *pkt_md5_out, void *stream) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c) return P2G_ERR_INVALID_ARGUMENT;
        if (count && (!rec || !set_id || !pkt_md5_out))
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "null argument");
        if (reinterpret_cast<uintptr_t>(rec) % 4 != 0 || reinterpret_cast<uintptr_t>(pkt_md5_out) % 4 != 0)
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "recovery blocks and digests must be 4-byte aligned");
        int rc = check_slice_size(c, slice_size);
        if (rc != P2G_OK) return rc;
        if ((uint64_t)first_exponent + count > 65536)
            return c->fail(P2G_ERR_TOO_MANY_RECOVERY_BLOCKS, "recovery exponent above 65535");
        hipSetDevice(c->device);
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
        hipError_t e = launch_recovery_md5(rec, count, (uint32_t)slice_size, first_exponent, set_id,
                                           pkt_md5_out, st);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "recovery md5 kernel", e);
        return P2G_OK;
    });
}

int p2g_pass_timing(p2g_ctx *ctx, double *ms_out) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c || !ms_out) return P2G_ERR_INVALID_ARGUMENT;
        if (!c->last_pass_fused || !c->d_sync.ptr) {  // separate kernels, batched, or none yet
            for (uint32_t k = 0; k < 2 * kPassRoles; k++) ms_out[k] = 0.0;
            return P2G_OK;
        }
        hipSetDevice(c->device);
        hipError_t e = hipEventSynchronize(c->scratch_done[Ctx::kScratchEnc]);
        unsigned long long t[2 * kPassRoles];
        if (e == hipSuccess)
            e = hipMemcpy(t, static_cast<uint32_t *>(c->d_sync.ptr) + kPassTimes, sizeof t,
                          hipMemcpyDeviceToHost);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "read create-pass stamps", e);
        unsigned long long t0 = ~0ull;
        for (uint32_t r = 0; r < kPassRoles; r++)
            if (t[2 * r + 1]) t0 = std::min(t0, t[2 * r]);
        for (uint32_t k = 0; k < 2 * kPassRoles; k++)  // 100 MHz counter: 1e5 ticks per ms
            ms_out[k] = t[k | 1] && t[k] >= t0 ? (double)(t[k] - t0) / 1e5 : 0.0;
        return P2G_OK;
    });
}

}  // extern "C"

namespace p2g {

// p2g_dev_create_pass with an explicit encode variant and band width.
int create_pass_impl(Ctx *c, const p2g_dev_file *files, uint32_t nfiles, uint64_t slice_size,
                     uint32_t recovery_count, const uint8_t set_id[16], uint8_t *rec,
                     uint8_t *slice_md5_out, uint32_t *slice_crc_out, uint8_t *pkt_md5_out,
                     hipStream_t st, int variant, uint32_t band_cols) {
    const uint32_t R = recovery_count;
    if ((nfiles && !files) || !set_id || !slice_md5_out || !slice_crc_out ||
        (R && (!rec || !pkt_md5_out)))
        return c->fail(P2G_ERR_INVALID_ARGUMENT, "null argument");
    for (const void *p : {(const void *)rec, (const void *)slice_md5_out,
                          (const void *)slice_crc_out, (const void *)pkt_md5_out})
        if (reinterpret_cast<uintptr_t>(p) % 4 != 0)
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "outputs must be 4-byte aligned");
    if (R > P2G_MAX_RECOVERY_BLOCKS)
        return c->fail(P2G_ERR_TOO_MANY_RECOVERY_BLOCKS, "recovery count above 65535");
    int rc = check_slice_size(c, slice_size);
    if (rc != P2G_OK) return rc;
    hipSetDevice(c->device);
    std::vector<SliceDesc> descs;
    rc = append_descs(c, files, nfiles, slice_size, descs);
    if (rc != P2G_OK) return rc;
    if (descs.size() > kMaxInputs)
        return c->fail(P2G_ERR_TOO_MANY_INPUT_BLOCKS, "more than 32768 input slices");
    const uint32_t n = (uint32_t)descs.size();
    const uint32_t S = (uint32_t)slice_size;
    PassPlan plan;
    const bool fits_pass = plan_create_pass(n, R, S, c->num_cus, c->pass_resident, band_cols, &plan) &&
                           mfma_afrag_bytes(n, R) <= kPassAfragBudget;
    c->last_pass_fused = variant == kEncodeMfma && fits_pass;
    if (variant == kEncodeMfma && !fits_pass && n && R) {
        // More slices than the pass's descriptor table holds (or fragments
        // above its budget): the batched pass runs the set as one job.  Its
        // units take the slices in 4096-slice chunks and its pieces split
        // the rows by a fragment budget, so the encode stays on the LDS-DMA
        // loop and the hashes stay in the same launches.
        p2g_batch_job job;
        job.files = files;
        job.slice_size = slice_size;
        job.rec = rec;
        job.slice_md5 = slice_md5_out;
        job.slice_crc = slice_crc_out;
        job.pkt_md5 = pkt_md5_out;
        job.nfiles = nfiles;
        job.recovery_count = R;
        memcpy(job.set_id, set_id, 16);
        return create_pass_batched(c, job, st);
    }
    rc = c->scratch_acquire(Ctx::kScratchEnc, st);
    if (rc != P2G_OK) return rc;
    // Partial slices (the last one of each file) are staged into zero-padded
    // copies when they fit kStageBudget, so that every descriptor is a whole
    // slice and the encode takes the full-slice loads.  The copies go after
    // the n descriptors in the same table: {source, valid, staging slot}.
    std::vector<SliceDesc> parts;
    for (uint32_t i = 0; i < n; i++)
        if (descs[i].valid < S) parts.push_back({descs[i].ptr, descs[i].valid, (uint32_t)parts.size()});
    uint64_t stage_budget = kStageBudget;
    if (const char *v = getenv("P2G_STAGE_BYTES")) stage_budget = strtoull(v, nullptr, 10);
    const bool stage = !parts.empty() && (uint64_t)parts.size() * S <= stage_budget;
    if (stage) {
        if (c->d_partials.reserve(parts.size() * (size_t)S) != P2G_OK)
            return c->fail(P2G_ERR_OUT_OF_MEMORY, "partial slice staging");
        uint8_t *staging = static_cast<uint8_t *>(c->d_partials.ptr);
        for (uint32_t i = 0, k = 0; i < n; i++)
            if (descs[i].valid < S) descs[i] = {staging + (size_t)(k++) * S, S, 0};
        descs.insert(descs.end(), parts.begin(), parts.end());
    }
    const bool all_full = parts.empty() || stage;
    rc = upload_descs(c, c->enc_descs, descs, st);
    if (rc != P2G_OK) return rc;
    SliceDesc *dd = static_cast<SliceDesc *>(c->enc_descs.d.ptr);
    if (stage) {
        hipError_t es = launch_stage_partials(dd + n, (uint32_t)parts.size(), S,
                                              static_cast<uint8_t *>(c->d_partials.ptr), st);
        if (es != hipSuccess) return c->fail(P2G_ERR_DEVICE, "partial slice staging kernel", es);
    }
    if (c->enc_descs.tb.reserve((size_t)std::max<uint32_t>(n, 1) * 64) != P2G_OK)
        return c->fail(P2G_ERR_OUT_OF_MEMORY, "slice tail blocks");
    uint4 *tb = static_cast<uint4 *>(c->enc_descs.tb.ptr);
    hipError_t e;

    if (!c->last_pass_fused) {
        // The same results from the separate kernels, in stream order.
        rc = dev_encode(c, dd, n, slice_size, 0, n, R, rec, 0, st, variant);
        if (rc != P2G_OK) return rc;
        if (n) {
            e = launch_desc_tails(dd, n, tb, st);
            if (e == hipSuccess) e = launch_slice_md5(dd, tb, n, S, slice_md5_out, st);
            if (e == hipSuccess) e = launch_slice_crc32(dd, n, S, slice_crc_out, st);
            if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "slice checksum kernels", e);
        }
        e = launch_recovery_md5(rec, R, S, 0, set_id, pkt_md5_out, st);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "recovery md5 kernel", e);
        return c->scratch_release(Ctx::kScratchEnc, st);
    }
    if (c->d_mats.reserve(mfma_afrag_bytes(n, R)) != P2G_OK)
        return c->fail(P2G_ERR_OUT_OF_MEMORY, "matrix-core coefficient fragments");
    if (c->d_sync.reserve((size_t)pass_sync_words(plan) * 4) != P2G_OK)
        return c->fail(P2G_ERR_OUT_OF_MEMORY, "create-pass sync words");
    uint4 *afrag = static_cast<uint4 *>(c->d_mats.ptr);
    e = launch_desc_tails(dd, n, tb, st);
    if (e == hipSuccess)
        e = launch_mfma_afrag(static_cast<const uint16_t *>(c->d_antilog.ptr),
                              static_cast<const uint16_t *>(c->d_base_logs.ptr), n, R, afrag, st);
    if (e == hipSuccess)
        e = launch_create_pass(plan, dd, tb, n, afrag, R, S, rec, slice_md5_out, slice_crc_out, set_id,
                               pkt_md5_out, static_cast<uint32_t *>(c->d_sync.ptr),
                               static_cast<uint32_t *>(c->d_err.ptr), all_full, st);
    if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "create pass kernels", e);
    c->pass_pending = true;
    return c->scratch_release(Ctx::kScratchEnc, st);
}

}  // namespace p2g

extern "C" {

int p2g_dev_create_pass(p2g_ctx *ctx, const p2g_dev_file *files, uint32_t nfiles,
                        uint64_t slice_size, uint32_t recovery_count, const uint8_t set_id[16],
                        uint8_t *rec, uint8_t *slice_md5_out, uint32_t *slice_crc_out,
                        uint8_t *pkt_md5_out, void *stream) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c) return P2G_ERR_INVALID_ARGUMENT;
        if ((nfiles && !files) || slice_size == 0)
            return create_pass_impl(c, files, nfiles, slice_size, recovery_count, set_id, rec,
                                    slice_md5_out, slice_crc_out, pkt_md5_out, c->stream,
                                    c->encode_variant, c->band_cols);  // reports the argument error
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
        uint64_t n = 0;
        for (uint32_t f = 0; f < nfiles; f++) n += (files[f].len + slice_size - 1) / slice_size;
        auto run = [&](int variant, uint32_t band_cols) {
            return create_pass_impl(c, files, nfiles, slice_size, recovery_count, set_id, rec,
                                    slice_md5_out, slice_crc_out, pkt_md5_out, st, variant, band_cols);
        };
        TuneChoice ch;
        if (tune_lookup(c, n, recovery_count, slice_size, &ch)) return run(ch.variant, ch.band_cols);
        if (c->autotune && n && recovery_count)
            return tune_and_run(c, n, recovery_count, slice_size, st, run, nullptr);
        return run(c->encode_variant, c->band_cols);
    });
}

// ---- host-buffer entry points ----------------------------------------

int p2g_slice_checksums(p2g_ctx *ctx, const uint8_t *data, uint64_t len, uint64_t slice_size,
                        uint8_t *md5_out, uint32_t *crc_out) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c) return P2G_ERR_INVALID_ARGUMENT;
        if (len && (!data || !md5_out || !crc_out))
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "null argument");
        int rc = check_slice_size(c, slice_size);
        if (rc != P2G_OK) return rc;
        if (len == 0) return P2G_OK;
        hipSetDevice(c->device);
        const uint64_t n = (len + slice_size - 1) / slice_size;
        if (c->d_stage.reserve(len) != P2G_OK || c->d_md5.reserve(n * 16) != P2G_OK ||
            c->d_crc.reserve(n * 4) != P2G_OK)
            return c->fail(P2G_ERR_OUT_OF_MEMORY, "checksum buffers");
        hipError_t e = hipMemcpyAsync(c->d_stage.ptr, data, len, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "upload file bytes", e);
        p2g_dev_file f{static_cast<const uint8_t *>(c->d_stage.ptr), len};
        rc = p2g_dev_slice_checksums(ctx, &f, 1, slice_size, static_cast<uint8_t *>(c->d_md5.ptr),
                                     static_cast<uint32_t *>(c->d_crc.ptr), c->stream);
        if (rc != P2G_OK) return rc;
        e = hipMemcpyAsync(md5_out, c->d_md5.ptr, n * 16, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(crc_out, c->d_crc.ptr, n * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "slice checksums", e);
        return P2G_OK;
    });
}

int p2g_recovery_packet_md5(p2g_ctx *ctx, const uint8_t *rec, uint32_t count, uint64_t slice_size,
                            uint32_t first_exponent, const uint8_t set_id[16],
                            uint8_t *pkt_md5_out) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c) return P2G_ERR_INVALID_ARGUMENT;
        if (count && (!rec || !set_id || !pkt_md5_out))
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "null argument");
        int rc = check_slice_size(c, slice_size);
        if (rc != P2G_OK) return rc;
        if (count == 0) return P2G_OK;
        hipSetDevice(c->device);
        const size_t bytes = (size_t)count * slice_size;
        if (c->d_stage.reserve(bytes) != P2G_OK || c->d_md5.reserve((size_t)count * 16) != P2G_OK)
            return c->fail(P2G_ERR_OUT_OF_MEMORY, "packet md5 buffers");
        hipError_t e = hipMemcpyAsync(c->d_stage.ptr, rec, bytes, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "upload recovery blocks", e);
        rc = p2g_dev_recovery_packet_md5(ctx, static_cast<const uint8_t *>(c->d_stage.ptr), count,
                                         slice_size, first_exponent, set_id,
                                         static_cast<uint8_t *>(c->d_md5.ptr), c->stream);
        if (rc != P2G_OK) return rc;
        e = hipMemcpyAsync(pkt_md5_out, c->d_md5.ptr, (size_t)count * 16, hipMemcpyDeviceToHost,
                           c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "recovery packet md5", e);
        return P2G_OK;
    });
}

// ---- streaming encode job ----------------------------------------------

int p2g_encode_begin(p2g_ctx *ctx, uint64_t slice_size, uint32_t total_inputs,
                     uint32_t recovery_count, p2g_job **out) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c || !out) return P2G_ERR_INVALID_ARGUMENT;
        *out = nullptr;
        int rc = check_slice_size(c, slice_size);
        if (rc != P2G_OK) return rc;
        if (recovery_count > P2G_MAX_RECOVERY_BLOCKS)
            return c->fail(P2G_ERR_TOO_MANY_RECOVERY_BLOCKS, "recovery count above 65535");
        if (total_inputs > kMaxInputs)
            return c->fail(P2G_ERR_TOO_MANY_INPUT_BLOCKS, "more than 32768 input slices");
        hipSetDevice(c->device);
        Job *j = new (std::nothrow) Job();
        if (!j) return P2G_ERR_OUT_OF_MEMORY;
        j->ctx = c;
        j->S = slice_size;
        j->N = total_inputs;
        j->R = recovery_count;
        const size_t rec_bytes = (size_t)recovery_count * slice_size;
        const size_t desc_bytes = (size_t)std::max<uint32_t>(total_inputs, 1) * sizeof(SliceDesc);
        if ((rec_bytes && j->d_rec.reserve(rec_bytes) != P2G_OK) ||
            j->d_descs.reserve(desc_bytes) != P2G_OK || j->h_descs.reserve(desc_bytes) != P2G_OK) {
            p2g_encode_free(reinterpret_cast<p2g_job *>(j));
            return c->fail(P2G_ERR_OUT_OF_MEMORY, "encode job buffers");
        }
        if (rec_bytes) {
            hipError_t e = hipMemsetAsync(j->d_rec.ptr, 0, rec_bytes, c->stream);
            if (e != hipSuccess) {
                p2g_encode_free(reinterpret_cast<p2g_job *>(j));
                return c->fail(P2G_ERR_DEVICE, "clear accumulator", e);
            }
        }
        *out = reinterpret_cast<p2g_job *>(j);
        return P2G_OK;
    });
}

int p2g_encode_push(p2g_job *job, uint32_t first_input_index, const uint8_t *data, uint64_t len) {
    return guarded(job ? reinterpret_cast<Job *>(job)->ctx : nullptr, [&]() -> int {
        Job *j = reinterpret_cast<Job *>(job);
        if (!j) return P2G_ERR_INVALID_ARGUMENT;
        Ctx *c = j->ctx;
        if (len && !data) return c->fail(P2G_ERR_INVALID_ARGUMENT, "null argument");
        if (len == 0 || j->R == 0) return P2G_OK;
        hipSetDevice(c->device);
        const uint64_t n = (len + j->S - 1) / j->S;
        if ((uint64_t)first_input_index + n > j->N)
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "pushed slices exceed total_inputs");
        // Double-buffered staging: this push's copy runs on the side stream while
        // the previous push's encode runs on the main stream.  A staging buffer
        // is reused only after the encode that read it has finished, and the
        // call returns once `data` has been copied, so the caller may reuse it.
        Job::Stage &sg = j->stage[j->next_stage];
        j->next_stage ^= 1;
        hipError_t e;
        if (sg.done) {
            e = hipEventSynchronize(sg.done);
            if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "staging wait", e);
        } else {
            e = hipEventCreateWithFlags(&sg.done, hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&sg.copied, hipEventDisableTiming);
            if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "event create", e);
        }
        if (sg.buf.reserve(len) != P2G_OK) return c->fail(P2G_ERR_OUT_OF_MEMORY, "staging buffer");
        hipStream_t side;
        int rc = c->side_stream(&side);
        if (rc != P2G_OK) return rc;
        e = hipMemcpyAsync(sg.buf.ptr, data, len, hipMemcpyHostToDevice, side);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "upload slices", e);
        // Descriptors of this push go to their own range of the job's tables,
        // which no other push touches.
        SliceDesc *hd = static_cast<SliceDesc *>(j->h_descs.ptr) + first_input_index;
        const uint8_t *base = static_cast<const uint8_t *>(sg.buf.ptr);
        for (uint64_t s = 0; s < n; s++) {
            const uint64_t have = len - s * j->S;
            hd[s] = {base + s * j->S, (uint32_t)(have < j->S ? have : j->S), 0};
        }
        SliceDesc *dd = static_cast<SliceDesc *>(j->d_descs.ptr) + first_input_index;
        e = hipMemcpyAsync(dd, hd, n * sizeof(SliceDesc), hipMemcpyHostToDevice, side);
        if (e == hipSuccess) e = hipEventRecord(sg.copied, side);
        if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, sg.copied, 0);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "upload slice descriptors", e);
        rc = c->scratch_acquire(Ctx::kScratchEnc, c->stream);
        if (rc == P2G_OK)
            rc = dev_encode(c, dd, (uint32_t)n, j->S, first_input_index, j->N, j->R,
                            static_cast<uint8_t *>(j->d_rec.ptr), 1, c->stream);
        if (rc == P2G_OK) rc = c->scratch_release(Ctx::kScratchEnc, c->stream);
        if (rc != P2G_OK) return rc;
        e = hipEventRecord(sg.done, c->stream);
        if (e == hipSuccess) e = hipEventSynchronize(sg.copied);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "event record", e);
        return P2G_OK;
    });
}

int p2g_encode_finish(p2g_job *job, uint8_t *rec_out, const uint8_t *set_id, uint8_t *pkt_md5_out) {
    return guarded(job ? reinterpret_cast<Job *>(job)->ctx : nullptr, [&]() -> int {
        Job *j = reinterpret_cast<Job *>(job);
        if (!j) return P2G_ERR_INVALID_ARGUMENT;
        Ctx *c = j->ctx;
        if (j->R == 0) return P2G_OK;
        if (!rec_out || (set_id && !pkt_md5_out))
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "null argument");
        hipSetDevice(c->device);
        const size_t bytes = (size_t)j->R * j->S;
        hipError_t e;
        if (set_id) {
            if (c->d_md5.reserve((size_t)j->R * 16) != P2G_OK)
                return c->fail(P2G_ERR_OUT_OF_MEMORY, "packet md5 buffer");
            e = launch_recovery_md5(static_cast<const uint8_t *>(j->d_rec.ptr), j->R, (uint32_t)j->S,
                                    0, set_id, static_cast<uint8_t *>(c->d_md5.ptr), c->stream);
            if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "recovery md5 kernel", e);
        }
        e = hipMemcpyAsync(rec_out, j->d_rec.ptr, bytes, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && set_id)
            e = hipMemcpyAsync(pkt_md5_out, c->d_md5.ptr, (size_t)j->R * 16, hipMemcpyDeviceToHost,
                               c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "encode finish", e);
        return P2G_OK;
    });
}

void p2g_encode_free(p2g_job *job) {
    Job *j = reinterpret_cast<Job *>(job);
    if (!j) return;
    hipSetDevice(j->ctx->device);
    hipStreamSynchronize(j->ctx->stream);
    for (auto &sg : j->stage) {
        if (sg.done) hipEventDestroy(sg.done);
        if (sg.copied) hipEventDestroy(sg.copied);
        sg.buf.release();
    }
    j->d_rec.release();
    j->d_descs.release();
    j->h_descs.release();
    delete j;
}

}  // extern "C"
