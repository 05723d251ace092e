// p2g_context.hpp -- the state behind p2g_ctx and p2g_job.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <exception>
#include <functional>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "p2g_internal.hpp"

namespace p2g {

constexpr uint32_t kMaxInputs = 32768;            // MAX_INPUT_BLOCKS, src/creator.rs:106
constexpr uint64_t kMaxSliceBytes = P2G_MAX_SLICE_BYTES;  // kernels index slices with u32
constexpr uint64_t kMatsBudget = 256ull << 20;    // coefficient matrices per launch chunk
constexpr uint64_t kPassAfragBudget = 4ull << 30; // A fragments of one fused create pass
constexpr uint64_t kStageBudget = 256ull << 20;   // zero-padded copies of partial slices (P2G_STAGE_BYTES overrides)

// Grow-only device allocation.
struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    int reserve(size_t n);
    void release();
};

// Grow-only pinned host allocation.
struct PinBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    int reserve(size_t n);
    void release();
};

// Slice descriptors staged through pinned memory into a device table.  One
// slot per kind of launch, so launches on different streams never share one.
struct DescSlot {
    PinBuf h;
    DevBuf d;
    DevBuf tb;  // 64-byte tail block per descriptor (slice MD5)
    hipEvent_t free = nullptr;  // `h` may be rewritten once this fires
};

// Encode kernel variants: the matrix-core GF(2) contraction (default), the
// bit-sliced VALU kernel and the split-nibble byte-permute kernel.
// P2G_ENCODE=mfma|bitslice|nibble picks one at p2g_open.
enum { kEncodeMfma = 0, kEncodeBitslice = 1, kEncodeNibble = 2, kEncodeVariants = 3 };

struct BatchPlan;  // p2g_batch.cpp

// Autotuner choice for one (N, R, slice size) bucket (p2g_tune.cpp).
struct TuneChoice {
    int variant = kEncodeMfma;
    uint32_t band_cols = 0;  // 0: the default band width
    double ms = 0;           // the winning time when it was measured
    // the fastest encode kernel alone (p2g_dev_encode, p2g_encode_push), -1
    // when not measured, and its time
    int enc_variant = -1;
    double enc_ms = 0;
};

struct Ctx {
    int device = 0;
    int encode_variant = kEncodeMfma;
    uint32_t num_cus = 0;          // compute units of the device
    // workgroups of each persistent kernel the device holds at once
    // (occupancy x CUs): a persistent launch is never larger
    uint32_t pass_resident = 0, batch_resident = 0;
    uint32_t band_cols = 0;        // fused-pass band width in 512-byte column groups (0: default)
    hipStream_t stream = nullptr;  // kernels and copies of this context
    hipStream_t side = nullptr;    // copies overlapped with `stream`; see side_stream()
    std::string err;

    std::vector<uint16_t> h_log, h_antilog;
    std::vector<uint16_t> base_logs;  // L_i of every possible input (32768)

    DevBuf d_antilog, d_base_logs;
    DevBuf d_mats, d_md5, d_crc, d_stage;
    DevBuf d_sync;                 // fused-pass stamps and band counts
    DevBuf d_err;                  // fused-pass sticky error word (read and cleared by p2g_synchronize)
    // p2g_dev_create_pass_noalloc: the device address of its static scratch
    // array on this device, and the end of this context's last use of it
    uint8_t *noalloc_scratch = nullptr;
    hipEvent_t noalloc_done = nullptr;
    DevBuf d_partials;             // zero-padded copies of partial slices (fused pass)
    DevBuf d_xs_jumps;             // xorshift64 jumps as byte tables (synthetic input)
    bool pass_pending = false;     // a fused pass may have set error flags not yet read
    bool last_pass_fused = false;  // the last p2g_dev_create_pass ran as one k_create_pass
    // Shared device buffers and the event recorded after the last kernel that
    // reads them; an entry point that rewrites them first makes its stream
    // wait for that event, so calls on different streams never overlap on
    // them.  Encode side: enc_descs, d_mats, d_sync.  Checksum side: sum_descs.
    enum Scratch { kScratchEnc = 0, kScratchSum = 1 };
    hipEvent_t scratch_done[2] = {nullptr, nullptr};
    DescSlot enc_descs;  // encode launches
    DescSlot sum_descs;  // slice checksum launches
    std::vector<BatchPlan *> pending_batches;  // p2g_dev_create_batch, until p2g_synchronize
    CostRates rates;                           // planners' cost estimates, updated from runs
    std::map<uint32_t, TuneChoice> tune;       // autotuner cache by bucket
    bool autotune = false;                     // P2G_AUTOTUNE=1: tune a bucket on its first call
    std::string tune_path;                     // P2G_TUNE_CACHE: cache file, loaded at open

    int open(int dev);
    int side_stream(hipStream_t *out);
    int scratch_acquire(Scratch which, hipStream_t st);  // st waits for earlier users
    int scratch_release(Scratch which, hipStream_t st);  // marks the end of this use on st
    ~Ctx();
    int fail(int code, const char *what, hipError_t e = hipSuccess);
};

struct Job {
    Ctx *ctx = nullptr;
    uint64_t S = 0;
    uint32_t N = 0, R = 0;
    DevBuf d_rec;
    DevBuf d_descs;   // one descriptor per global input index
    PinBuf h_descs;
    struct Stage {
        DevBuf buf;
        hipEvent_t done = nullptr;    // the encode that read `buf` has finished
        hipEvent_t copied = nullptr;  // host bytes and descriptors are on the device
    } stage[2];
    int next_stage = 0;
};

int check_slice_size(Ctx *c, uint64_t S);

// The device address of p2g_dev_create_pass_noalloc's static scratch array
// on `device` (the current device), looked up once per device
// (p2g_noalloc.hip).
int noalloc_scratch_init(int device, uint8_t **out);

// Batched create plans (p2g_batch.cpp).  An item is recovery rows [r0, r1)
// of a job and the checksums of its slices [s0, s1) (s1 past the job's
// slice count means all of them).
struct PlanItem {
    const p2g_batch_job *job;
    uint32_t r0, r1;
    uint32_t s0, s1;
    // column groups [c0, c1) of the rows (c1 past the row: to its end); a
    // range short of a whole row continues or saves each row's packet hash
    // in md5_state (the job's R chaining states, 16 bytes each)
    uint32_t c0 = 0, c1 = 0xFFFFFFFFu;
    void *md5_state = nullptr;
};
int check_batch_job(Ctx *c, const p2g_batch_job &J, std::vector<SliceDesc> &descs);
int build_plan(Ctx *c, const std::vector<PlanItem> &items, bool use_graph, BatchPlan *b);
// Builds and uploads the synthetic-input jump tables (c->d_xs_jumps) once.
int xs_jump_tables(Ctx *c);
int launch_plan(BatchPlan *b, hipStream_t st);
int wait_plan(BatchPlan *b);  // waits for the last launch, reads its flags and stamps
void free_batch_plan(BatchPlan *b);
// One set as a one-job batched plan launched on st (p2g_dev_create_pass for
// sets the fused pass cannot hold); the plan is freed by p2g_synchronize, or
// here once more than two such plans are pending.
int create_pass_batched(Ctx *c, const p2g_batch_job &job, hipStream_t st);

// Appends the descriptors of every slice of `files` (file order, slice order).
int append_descs(Ctx *c, const p2g_dev_file *files, uint32_t nfiles, uint64_t S,
                 std::vector<SliceDesc> &out);

// Copies descriptors into a slot's device table through its pinned staging
// buffer, in stream order.
int upload_descs(Ctx *c, DescSlot &slot, const std::vector<SliceDesc> &descs, hipStream_t st);

// rec (^)= encode of the n slices described by d_descs (device), whose global
// input indices start at first_input.
// variant: kEncodeMfma / kEncodeBitslice / kEncodeNibble, or -1 for the
// bucket's tuned encode kernel (else the context's).
int dev_encode(Ctx *c, const SliceDesc *d_descs, uint32_t n, uint64_t S, uint32_t first_input,
               uint32_t total_inputs, uint32_t R, uint8_t *d_rec, int accumulate,
               hipStream_t st, int variant = -1);

// p2g_dev_encode with an explicit encode kernel (-1: the bucket's tuned one,
// else the context's).
int dev_encode_files(Ctx *c, const p2g_dev_file *files, uint32_t nfiles, uint64_t slice_size,
                     uint32_t first_input_index, uint32_t total_inputs, uint32_t recovery_count,
                     uint8_t *rec, int accumulate, hipStream_t st, int variant);

// p2g_dev_create_pass with an explicit encode variant and band width
// (band_cols 0: the default).
int create_pass_impl(Ctx *c, const p2g_dev_file *files, uint32_t nfiles, uint64_t slice_size,
                     uint32_t recovery_count, const uint8_t set_id[16], uint8_t *rec,
                     uint8_t *slice_md5_out, uint32_t *slice_crc_out, uint8_t *pkt_md5_out,
                     hipStream_t st, int variant, uint32_t band_cols);

// Autotuner (p2g_tune.cpp).  Buckets: bit lengths of N, R and S.
uint32_t tune_bucket(uint64_t n, uint32_t R, uint64_t S);
bool tune_lookup(Ctx *c, uint64_t n, uint32_t R, uint64_t S, TuneChoice *out);
// Runs every candidate (variant x band width) through `run` on stream st,
// timing each, stores the fastest for the bucket (and the cache file), and
// leaves the outputs of a run of the fastest.
int tune_and_run(Ctx *c, uint64_t n, uint32_t R, uint64_t S, hipStream_t st,
                 const std::function<int(int, uint32_t)> &run, TuneChoice *out);
// Times each encode kernel alone (run_encode(variant) queues one encode on
// st) and records the fastest as the bucket's encode kernel.
int tune_encode(Ctx *c, uint64_t n, uint32_t R, uint64_t S, hipStream_t st,
                const std::function<int(int)> &run_encode);
int tune_load(Ctx *c, const char *path);
int tune_save(Ctx *c, const char *path);

// Runs the body of a C entry point so that no C++ exception crosses the C
// ABI: a failed host allocation (std::bad_alloc) returns
// P2G_ERR_OUT_OF_MEMORY, any other exception (a thread that could not start,
// a failed system call reported by the standard library) P2G_ERR_IO, as the
// reference reports OS failures as Par2Error::Io (src/error.rs).  The text
// goes to c's last error when there is a context.
template <typename F>
int guarded(Ctx *c, F &&body) noexcept {
    const char *what = nullptr;
    int rc = P2G_ERR_IO;
    try {
        return body();
    } catch (const std::bad_alloc &) {
        what = "host memory allocation failed";
        rc = P2G_ERR_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        what = e.what();
    } catch (...) {
        what = "unexpected host failure";
    }
    if (c) {
        try {
            c->err = what;
        } catch (...) {
        }
    }
    return rc;
}

}  // namespace p2g
