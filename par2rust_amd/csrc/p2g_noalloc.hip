// p2g_noalloc.hip -- p2g_dev_create_pass_noalloc: the fused create pass with
// every buffer beyond the caller's inputs and outputs in one static device
// array, for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops).  The call makes no hipMalloc,
// hipHostMalloc or hipFree and builds its slice descriptors on the device
// from kernel arguments, so nothing is staged through host memory either.
//
// The array `scratch` is file scope and per device: every GPU the library
// runs on holds kNoAllocScratchBytes of it until the process ends, and every
// call on that GPU shares it (calls that overlap on one GPU race; calls of
// one context are ordered by an event across its streams).  In host code the
// name `scratch` is only a stand-in: its device address, one per device,
// comes from hipGetSymbolAddress at p2g_open and is the only form host code
// passes to kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "p2g_context.hpp"
#include "p2g_internal.hpp"

namespace p2g {

constexpr uint64_t kNoAllocScratchBytes = 64ull << 20;
constexpr uint32_t kNoAllocMaxFiles = 64;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

// Device address of `scratch` per device (nullptr: not looked up yet).
static std::mutex g_scratch_mu;
static uint8_t *g_scratch_ptr[kNoAllocMaxDevices] = {};

// Looks the array up once per device, with that device current (p2g_open).
int noalloc_scratch_init(int device, uint8_t **out) {
    *out = nullptr;
    if (device < 0 || device >= kNoAllocMaxDevices) return P2G_ERR_DEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return P2G_ERR_DEVICE;
        g_scratch_ptr[device] = static_cast<uint8_t *>(p);
    }
    *out = g_scratch_ptr[device];
    return P2G_OK;
}

// The set's files as kernel arguments: file f holds slices
// [first[f], first[f + 1]); part[f] is the staging slot of its partial last
// slice, or -1 (no partial slice, or not staged).
struct NoAllocFiles {
    const uint8_t *ptr[kNoAllocMaxFiles];
    uint64_t len[kNoAllocMaxFiles];
    uint32_t first[kNoAllocMaxFiles + 1];
    int32_t part[kNoAllocMaxFiles];
    uint32_t nfiles, S, n;
};

// Thread per slice: descriptor i of the set (file order, slice order), and
// for a staged partial slice also its staging entry {source, valid, slot}
// after the n descriptors (the layout k_stage_partials reads).
__global__ __launch_bounds__(256) void k_noalloc_descs(NoAllocFiles a, SliceDesc *__restrict__ descs,
                                                       uint8_t *__restrict__ staging) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    uint32_t f = 0;
    while (i >= a.first[f + 1]) f++;
    const uint64_t s = i - a.first[f];
    const uint64_t have = a.len[f] - s * a.S;
    const uint32_t valid = (uint32_t)(have < a.S ? have : a.S);
    if (valid < a.S && a.part[f] >= 0) {
        const uint32_t slot = (uint32_t)a.part[f];
        descs[i] = {staging + (size_t)slot * a.S, a.S, 0};
        descs[a.n + slot] = {a.ptr[f] + s * a.S, valid, slot};
    } else {
        descs[i] = {a.ptr[f] + s * a.S, valid, 0};
    }
}

static uint64_t a256(uint64_t x) { return (x + 255) / 256 * 256; }

}  // namespace p2g

using namespace p2g;

extern "C" int p2g_dev_create_pass_noalloc(p2g_ctx *ctx, const p2g_dev_file *files, uint32_t nfiles,
                                           uint64_t slice_size, uint32_t recovery_count,
                                           const uint8_t set_id[16], uint8_t *rec,
                                           uint8_t *slice_md5_out, uint32_t *slice_crc_out,
                                           uint8_t *pkt_md5_out, void *stream) {
    return guarded(reinterpret_cast<Ctx *>(ctx), [&]() -> int {
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        if (!c) return P2G_ERR_INVALID_ARGUMENT;
        const uint32_t R = recovery_count;
        if ((nfiles && !files) || !set_id || !slice_md5_out || !slice_crc_out || !rec || !pkt_md5_out)
            return c->fail(P2G_ERR_INVALID_ARGUMENT, "null argument");
        for (const void *p : {(const void *)rec, (const void *)slice_md5_out,
                              (const void *)slice_crc_out, (const void *)pkt_md5_out})
            if (reinterpret_cast<uintptr_t>(p) % 4 != 0)
                return c->fail(P2G_ERR_INVALID_ARGUMENT, "outputs must be 4-byte aligned");
        if (R > P2G_MAX_RECOVERY_BLOCKS)
            return c->fail(P2G_ERR_TOO_MANY_RECOVERY_BLOCKS, "recovery count above 65535");
        int rc = check_slice_size(c, slice_size);
        if (rc != P2G_OK) return rc;
        if (nfiles > kNoAllocMaxFiles)
            return c->fail(P2G_ERR_INVALID_ARGUMENT,
                           "p2g_dev_create_pass_noalloc takes at most 64 files per call");
        if (!c->noalloc_scratch)
            return c->fail(P2G_ERR_DEVICE, "the allocation-free scratch array was not found at open");
        const uint32_t S = (uint32_t)slice_size;
        NoAllocFiles a;
        memset(&a, 0, sizeof a);
        a.nfiles = nfiles;
        a.S = S;
        uint64_t n = 0;
        uint32_t parts = 0;
        for (uint32_t f = 0; f < nfiles; f++) {
            if (files[f].len && !files[f].data)
                return c->fail(P2G_ERR_INVALID_ARGUMENT, "null device file pointer");
            if (reinterpret_cast<uintptr_t>(files[f].data) % 16 != 0)
                return c->fail(P2G_ERR_INVALID_ARGUMENT, "device file pointers must be 16-byte aligned");
            a.ptr[f] = files[f].data;
            a.len[f] = files[f].len;
            a.first[f] = (uint32_t)std::min<uint64_t>(n, 0xFFFFFFFFull);
            n += (files[f].len + S - 1) / S;
            a.part[f] = files[f].len % S ? (int32_t)parts++ : -1;
        }
        if (n > kMaxInputs) return c->fail(P2G_ERR_TOO_MANY_INPUT_BLOCKS, "more than 32768 input slices");
        if (n == 0 || R == 0)
            return c->fail(P2G_ERR_INVALID_ARGUMENT,
                           "p2g_dev_create_pass_noalloc needs at least one slice and one recovery block");
        a.first[nfiles] = (uint32_t)n;
        a.n = (uint32_t)n;
        TuneChoice ch;
        const uint32_t band = tune_lookup(c, n, R, S, &ch) && ch.variant == kEncodeMfma && ch.band_cols
                                  ? ch.band_cols
                                  : c->band_cols;
        PassPlan plan;
        if (!plan_create_pass(a.n, R, S, c->num_cus, c->pass_resident, band, &plan))
            return c->fail(P2G_ERR_INVALID_ARGUMENT,
                           "p2g_dev_create_pass_noalloc takes the sets one fused pass holds "
                           "(at most 2048 slices)");
        // the scratch layout: sync words, descriptors (n, then the staged
        // partial slices), tail blocks, staging slots, A fragments
        const uint64_t sync_b = a256((uint64_t)pass_sync_words(plan) * 4);
        const uint64_t desc_b = a256((n + parts) * sizeof(SliceDesc));
        const uint64_t tb_b = a256(n * 64);
        const uint64_t afrag_b = a256(mfma_afrag_bytes(a.n, R));
        const uint64_t fixed = sync_b + desc_b + tb_b + afrag_b;
        if (fixed > kNoAllocScratchBytes)
            return c->fail(P2G_ERR_INVALID_ARGUMENT,
                           "set too large for p2g_dev_create_pass_noalloc: its scratch exceeds the "
                           "64 MiB limit (A fragments take 128 bytes per recovery block per slice, "
                           "rounded up to 16 recovery blocks x 16 slices)");
        // partial slices are staged (whole-slice loads) when their copies fit
        // the rest of the array, else read in place
        const uint64_t stage_b = a256((uint64_t)parts * S);
        const bool stage = parts > 0 && fixed + stage_b <= kNoAllocScratchBytes;
        if (!stage)
            for (uint32_t f = 0; f < nfiles; f++) a.part[f] = -1;
        uint8_t *base = c->noalloc_scratch;
        uint32_t *d_sync = reinterpret_cast<uint32_t *>(base);
        SliceDesc *d_desc = reinterpret_cast<SliceDesc *>(base + sync_b);
        uint4 *d_tb = reinterpret_cast<uint4 *>(base + sync_b + desc_b);
        uint4 *d_afrag = reinterpret_cast<uint4 *>(base + sync_b + desc_b + tb_b);
        uint8_t *d_stage = base + fixed;

        hipSetDevice(c->device);
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
        // this context's earlier calls are done with the array
        hipError_t e = hipStreamWaitEvent(st, c->noalloc_done, 0);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_noalloc_descs, dim3((a.n + 255) / 256), dim3(256), 0, st, a, d_desc,
                               d_stage);
            e = hipGetLastError();
        }
        if (e == hipSuccess && stage) e = launch_stage_partials(d_desc + n, parts, S, d_stage, st);
        if (e == hipSuccess) e = launch_desc_tails(d_desc, a.n, d_tb, st);
        if (e == hipSuccess)
            e = launch_mfma_afrag(static_cast<const uint16_t *>(c->d_antilog.ptr),
                                  static_cast<const uint16_t *>(c->d_base_logs.ptr), a.n, R, d_afrag, st);
        if (e == hipSuccess)
            e = launch_create_pass(plan, d_desc, d_tb, a.n, d_afrag, R, S, rec, slice_md5_out,
                                   slice_crc_out, set_id, pkt_md5_out, d_sync,
                                   static_cast<uint32_t *>(c->d_err.ptr), stage || parts == 0, st);
        if (e == hipSuccess) e = hipEventRecord(c->noalloc_done, st);
        if (e != hipSuccess) return c->fail(P2G_ERR_DEVICE, "allocation-free create pass kernels", e);
        c->pass_pending = true;       // p2g_synchronize reads the sticky error word
        c->last_pass_fused = false;   // p2g_pass_timing reads the pass's own stamps only
        return P2G_OK;
    });
}
