// p2g_tune.cpp -- the autotuner: per (N, R, slice size) bucket, the create
// pass variant (matrix-core fused pass, or separate launches with the
// bit-sliced or the split-nibble VALU encode) and the column band width of
// the fused pass that ran fastest, plus the fastest encode kernel alone (for
// p2g_dev_encode and p2g_encode_push), cached in the context and optionally
// in a text file.
//
// Buckets are the bit lengths of N (slices), R (recovery blocks) and S
// (slice bytes), so one measurement covers shapes within a factor of two.
// A bucket is tuned on demand (p2g_autotune, on synthetic input of the
// given shape) or, with P2G_AUTOTUNE=1, on the first p2g_dev_create_pass
// call that falls in it, using that call's own data: every candidate
// computes the same outputs, so the call's results are those of the last
// (fastest) run.  The batched planner takes the band width of a job's
// bucket from the same cache.
//
// Cache file (P2G_TUNE_CACHE, p2g_tune_save / p2g_tune_load): one line per
// bucket, "bucket variant band_cols ms enc_variant enc_ms", bucket =
// nbits(N) | nbits(R) << 8 | nbits(S) << 16, enc_variant -1 when the
// encode kernels were not timed (lines of the four first fields load too).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <vector>

#include "p2g_context.hpp"

namespace p2g {

static uint32_t nbits(uint64_t x) {
    uint32_t b = 0;
    while (x) {
        b++;
        x >>= 1;
    }
    return b;
}

uint32_t tune_bucket(uint64_t n, uint32_t R, uint64_t S) {
    return nbits(n) | nbits(R) << 8 | nbits(S) << 16;
}

bool tune_lookup(Ctx *c, uint64_t n, uint32_t R, uint64_t S, TuneChoice *out) {
    auto it = c->tune.find(tune_bucket(n, R, S));
    if (it == c->tune.end()) return false;
    *out = it->second;
    return true;
}

int tune_and_run(Ctx *c, uint64_t n, uint32_t R, uint64_t S, hipStream_t st,
                 const std::function<int(int, uint32_t)> &run, TuneChoice *out) {
    // candidates: the fused pass at several band widths (when it applies),
    // then the bit-sliced kernels
    std::vector<std::pair<int, uint32_t>> cand;
    PassPlan plan;
    const uint32_t col_groups = (uint32_t)((S + 511) / 512);
    if (n <= kMaxInputs && plan_create_pass((uint32_t)n, R, (uint32_t)S, c->num_cus, c->pass_resident, 0, &plan)) {
        for (uint32_t w : {16u, 32u, 64u, 128u})
            if (w <= std::max<uint32_t>(col_groups, 16)) cand.push_back({kEncodeMfma, w});
    } else {
        cand.push_back({kEncodeMfma, 0});
    }
    cand.push_back({kEncodeBitslice, 0});
    cand.push_back({kEncodeNibble, 0});
    double best = 1e300;
    size_t best_k = 0;
    for (size_t k = 0; k < cand.size(); k++) {
        double t_best = 1e300;
        for (int rep = 0; rep < 2; rep++) {
            hipError_t e = hipStreamSynchronize(st);
            if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "autotune sync", e);
            const auto t0 = std::chrono::steady_clock::now();
            int rc = run(cand[k].first, cand[k].second);
            if (rc != P2G_OK) return rc;
            e = hipStreamSynchronize(st);
            if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "autotune sync", e);
            t_best = std::min(
                t_best,
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
            if (t_best > 2 * best) break;  // clearly slower: no second measurement
        }
        if (t_best < best) {
            best = t_best;
            best_k = k;
        }
    }
    if (best_k + 1 != cand.size()) {  // leave the outputs of the winner
        int rc = run(cand[best_k].first, cand[best_k].second);
        if (rc != P2G_OK) return rc;
    }
    TuneChoice &ch = c->tune[tune_bucket(n, R, S)];  // keeps an encode-kernel choice
    ch.variant = cand[best_k].first;
    ch.band_cols = cand[best_k].second;
    ch.ms = best;
    if (out) *out = ch;
    if (!c->tune_path.empty()) return tune_save(c, c->tune_path.c_str());
    return P2G_OK;
}

int tune_encode(Ctx *c, uint64_t n, uint32_t R, uint64_t S, hipStream_t st,
                const std::function<int(int)> &run_encode) {
    double best = 1e300;
    int best_v = kEncodeMfma;
    for (int v = 0; v < kEncodeVariants; v++) {
        double t_best = 1e300;
        for (int rep = 0; rep < 3; rep++) {  // the first run also warms the kernel up
            hipError_t e = hipStreamSynchronize(st);
            if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "autotune sync", e);
            const auto t0 = std::chrono::steady_clock::now();
            int rc = run_encode(v);
            if (rc != P2G_OK) return rc;
            e = hipStreamSynchronize(st);
            if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "autotune sync", e);
            t_best = std::min(
                t_best,
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
            if (t_best > 4 * best) break;  // clearly slower
        }
        if (t_best < best) {
            best = t_best;
            best_v = v;
        }
    }
    TuneChoice &ch = c->tune[tune_bucket(n, R, S)];
    ch.enc_variant = best_v;
    ch.enc_ms = best;
    return P2G_OK;
}

// One bucket per line: "bucket variant band_cols ms [enc_variant enc_ms]".
int tune_load(Ctx *c, const char *path) {
    FILE *f = fopen(path, "r");
    if (!f) return c->fail(P2G_ERR_IO, "cannot open the tuning cache");
    char line[256];
    while (fgets(line, sizeof line, f)) {
        unsigned bucket, band;
        int variant, enc = -1;
        double ms, enc_ms = 0;
        const int got = sscanf(line, "%u %d %u %lf %d %lf", &bucket, &variant, &band, &ms, &enc, &enc_ms);
        if (got < 4 || variant < 0 || variant >= kEncodeVariants) continue;
        TuneChoice ch;
        ch.variant = variant;
        ch.band_cols = band;
        ch.ms = ms;
        if (got == 6 && enc >= 0 && enc < kEncodeVariants) {
            ch.enc_variant = enc;
            ch.enc_ms = enc_ms;
        }
        c->tune[bucket] = ch;
    }
    fclose(f);
    return P2G_OK;
}

int tune_save(Ctx *c, const char *path) {
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "w");
    if (!f) return c->fail(P2G_ERR_IO, "cannot write the tuning cache");
    for (const auto &kv : c->tune)
        fprintf(f, "%u %d %u %.4f %d %.4f\n", kv.first, kv.second.variant, kv.second.band_cols,
                kv.second.ms, kv.second.enc_variant, kv.second.enc_ms);
    const bool ok = fclose(f) == 0 && rename(tmp.c_str(), path) == 0;
    return ok ? P2G_OK : c->fail(P2G_ERR_IO, "cannot write the tuning cache");
}

}  // namespace p2g

using namespace p2g;

extern "C" {

uint32_t p2g_tune_bucket(uint64_t slices, uint32_t recovery_count, uint64_t slice_size) {
    return tune_bucket(slices, recovery_count, slice_size);
}

int p2g_autotune(p2g_ctx *ctx, uint32_t slices, uint32_t recovery_count, uint64_t slice_size,
                 int *variant_out, uint32_t *band_cols_out, double *ms_out) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c) return P2G_ERR_INVALID_ARGUMENT;
        if (slices == 0 || slices > kMaxInputs || recovery_count == 0 ||
            recovery_count > P2G_MAX_RECOVERY_BLOCKS)
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "autotune shape out of range");
        int rc = check_slice_size(c, slice_size);
        if (rc != P2G_OK) return rc;
        hipSetDevice(c->device);
        // synthetic input of the shape: one file of slices * slice_size bytes
        const uint64_t bytes = (uint64_t)slices * slice_size;
        const uint64_t rec_bytes = (uint64_t)recovery_count * slice_size;
        DevBuf src, out;
        if (src.reserve(bytes) != P2G_OK ||
            out.reserve(rec_bytes + 20ull * slices + 16ull * recovery_count + 1024) != P2G_OK) {
            src.release();
            out.release();
            return c->fail(P2G_ERR_OUT_OF_MEMORY, "autotune buffers");
        }
        uint8_t *o = static_cast<uint8_t *>(out.ptr);
        uint8_t *rec = o, *md5 = o + (rec_bytes + 255) / 256 * 256;
        uint32_t *crc = reinterpret_cast<uint32_t *>(md5 + (16ull * slices + 255) / 256 * 256);
        uint8_t *pkt = reinterpret_cast<uint8_t *>(crc) + (4ull * slices + 255) / 256 * 256;
        static const uint8_t sid[16] = {};
        rc = p2g_dev_fill_xorshift64(ctx, static_cast<uint8_t *>(src.ptr), bytes, 1, c->stream);
        p2g_dev_file file{static_cast<const uint8_t *>(src.ptr), bytes};
        TuneChoice ch;
        // the encode kernels alone (p2g_dev_encode, p2g_encode_push), then the
        // create pass candidates
        if (rc == P2G_OK)
            rc = tune_encode(c, slices, recovery_count, slice_size, c->stream, [&](int variant) {
                return dev_encode_files(c, &file, 1, slice_size, 0, slices, recovery_count, rec, 0,
                                        c->stream, variant);
            });
        if (rc == P2G_OK)
            rc = tune_and_run(c, slices, recovery_count, slice_size, c->stream,
                              [&](int variant, uint32_t band) {
                                  return create_pass_impl(c, &file, 1, slice_size, recovery_count, sid,
                                                          rec, md5, crc, pkt, c->stream, variant, band);
                              },
                              &ch);
        if (rc == P2G_OK) rc = p2g_synchronize(ctx);
        src.release();
        out.release();
        if (rc != P2G_OK) return rc;
        if (variant_out) *variant_out = ch.variant;
        if (band_cols_out) *band_cols_out = ch.band_cols;
        if (ms_out) *ms_out = ch.ms;
        return P2G_OK;
    });
}

int p2g_tune_get(p2g_ctx *ctx, uint32_t slices, uint32_t recovery_count, uint64_t slice_size,
                 int *variant_out, uint32_t *band_cols_out) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c) return P2G_ERR_INVALID_ARGUMENT;
        TuneChoice ch;
        if (!tune_lookup(c, slices, recovery_count, slice_size, &ch)) {
            ch.variant = c->encode_variant;
            ch.band_cols = c->band_cols;
        }
        if (variant_out) *variant_out = ch.variant;
        if (band_cols_out) *band_cols_out = ch.band_cols;
        return P2G_OK;
    });
}

int p2g_tune_get_encode(p2g_ctx *ctx, uint32_t slices, uint32_t recovery_count, uint64_t slice_size,
                        int *variant_out, double *ms_out) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c) return P2G_ERR_INVALID_ARGUMENT;
        TuneChoice ch;
        const bool tuned = tune_lookup(c, slices, recovery_count, slice_size, &ch) && ch.enc_variant >= 0;
        if (variant_out) *variant_out = tuned ? ch.enc_variant : c->encode_variant;
        if (ms_out) *ms_out = tuned ? ch.enc_ms : 0.0;
        return P2G_OK;
    });
}

int p2g_tune_save(p2g_ctx *ctx, const char *path) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c || !path) return P2G_ERR_INVALID_ARGUMENT;
        return tune_save(c, path);
    });
}

int p2g_tune_load(p2g_ctx *ctx, const char *path) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c || !path) return P2G_ERR_INVALID_ARGUMENT;
        return tune_load(c, path);
    });
}

}  // extern "C"
