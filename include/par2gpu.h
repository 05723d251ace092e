/*
 * par2gpu.h -- C ABI of the MI355X PAR2 create engine (libpar2gpu.so).
 *
 * The reference (javi11/par2rust) has no FFI on this path: its encode is a
 * set of internal Rust calls.  Each entry point below replaces one of those
 * internal call sites, cited as file:line of the reference tree.  Plain
 * pointers and sizes only; every call returns a p2g status (0 = success);
 * p2g_last_error() gives the message of the last failure on a context.
 * No C++ exception crosses this ABI: a host allocation that fails inside a
 * call returns P2G_ERR_OUT_OF_MEMORY, any other host failure P2G_ERR_IO.
 *
 * Threading: a context is used by one host thread at a time.  All device
 * work of a context runs on that context's HIP streams; the host-buffer
 * entry points return after their results are in host memory.
 */
#ifndef PAR2GPU_H
#define PAR2GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  1..9 map one-to-one onto the Par2Error variants of
 * src/error.rs:7-41; 10.. are conditions the reference does not have. */
enum {
    P2G_OK = 0,
    P2G_ERR_IO = 1,                      /* Par2Error::Io */
    P2G_ERR_INVALID_SLICE_SIZE = 2,      /* Par2Error::InvalidSliceSize */
    P2G_ERR_INVALID_FILE_NAME = 3,       /* Par2Error::InvalidFileName */
    P2G_ERR_NO_INPUT_FILES = 4,          /* Par2Error::NoInputFiles */
    P2G_ERR_EMPTY_FILE = 5,              /* Par2Error::EmptyFile */
    P2G_ERR_TOO_MANY_FILES = 6,          /* Par2Error::TooManyFiles */
    P2G_ERR_TOO_MANY_RECOVERY_BLOCKS = 7,/* Par2Error::TooManyRecoveryBlocks */
    P2G_ERR_TOO_MANY_INPUT_BLOCKS = 8,   /* Par2Error::TooManyInputBlocks */
    P2G_ERR_INVALID_VOLUME_SCHEME = 9,   /* Par2Error::InvalidVolumeScheme */
    P2G_ERR_INVALID_ARGUMENT = 10,       /* bad pointer / size at this ABI */
    P2G_ERR_DEVICE = 11,                 /* HIP runtime or kernel failure */
    P2G_ERR_OUT_OF_MEMORY = 12,          /* device or pinned allocation failed */
};

/* PAR2 limits, src/creator.rs:102-108. */
#define P2G_MAX_FILES 32768u
#define P2G_MAX_INPUT_BLOCKS 32768u
#define P2G_MAX_RECOVERY_BLOCKS 65535u
/* Largest slice size the kernels take (they index a slice with 32 bits);
 * the reference accepts any multiple of 4.  A larger slice size is refused
 * with P2G_ERR_INVALID_ARGUMENT (a limit of this engine, not a PAR2 error).
 * Such a set has recovery blocks of over 2 GiB each. */
#define P2G_MAX_SLICE_BYTES 2147483648ull

typedef struct p2g_ctx p2g_ctx;
typedef struct p2g_job p2g_job;

/* Opens `device`, builds the GF(2^16) log/antilog tables on it and creates
 * the context's streams.  No reference counterpart: the reference's
 * equivalent state is gf16_tables() (src/galois.rs:28-31) and the global
 * rayon pool. */
int p2g_open(int device, p2g_ctx **out);
void p2g_close(p2g_ctx *ctx);
const char *p2g_last_error(const p2g_ctx *ctx);
/* Version string of the library, e.g. "par2gpu 0.1.0 gfx950". */
const char *p2g_version(void);

/* ---------------------------------------------------------------------
 * Host-buffer entry points (the drop-in cut points)
 * ------------------------------------------------------------------- */

/* Per-slice MD5 and CRC-32 of one file's bytes, the last slice zero padded
 * to slice_size: ceil(len / slice_size) entries into md5_out (16 bytes each)
 * and crc_out.  Replaces the per-slice closure of scan_via_mmap,
 * src/source.rs:165-220 (called from scan_with_progress, :100-109). */
int p2g_slice_checksums(p2g_ctx *ctx, const uint8_t *data, uint64_t len,
                        uint64_t slice_size, uint8_t *md5_out, uint32_t *crc_out);

/* Streaming recovery encode.  Replaces encode_all_recovery_blocks,
 * src/creator.rs:883-1018 (called at :252-257), and, through push with a
 * running first_input_index, encode_file_slices_into_buffers, :728-843.
 *   begin:  total_inputs = N of the whole set (fixes the RS input bases),
 *           recovery_count = R, first exponent 0 (src/creator.rs:915).
 *   push:   `len` bytes holding consecutive input slices starting at global
 *           input index first_input_index; the last slice may be partial
 *           and is zero padded.  Pushes may come in any order.
 *   finish: recovery block r (exponent r) into rec_out + r * slice_size.
 *           With set_id non-NULL, pkt_md5_out[16 r ..] also receives the
 *           packet hash of recovery packet r, MD5(set_id || "PAR 2.0\0RecvSlic"
 *           || le32(r) || block r), as build_packet computes it
 *           (src/format.rs:40-44 via src/packet/recovery.rs:11-19). */
int p2g_encode_begin(p2g_ctx *ctx, uint64_t slice_size, uint32_t total_inputs,
                     uint32_t recovery_count, p2g_job **out);
int p2g_encode_push(p2g_job *job, uint32_t first_input_index, const uint8_t *data,
                    uint64_t len);
int p2g_encode_finish(p2g_job *job, uint8_t *rec_out, const uint8_t *set_id,
                      uint8_t *pkt_md5_out);
void p2g_encode_free(p2g_job *job);

/* Packet hashes of R recovery packets whose blocks are in host memory
 * (block r at rec + r * slice_size, exponent first_exponent + r).  Replaces
 * the R calls of build_packet made by build_volume_packet,
 * src/creator.rs:1031-1041. */
int p2g_recovery_packet_md5(p2g_ctx *ctx, const uint8_t *rec, uint32_t count,
                            uint64_t slice_size, uint32_t first_exponent,
                            const uint8_t set_id[16], uint8_t *pkt_md5_out);

/* ---------------------------------------------------------------------
 * Device-buffer entry points (inputs already resident in HBM).  The same
 * kernels as above, without host copies; `stream` is a hipStream_t or NULL
 * for the context's own stream.  They return once the work is queued;
 * p2g_synchronize waits for it.
 * ------------------------------------------------------------------- */

/* One input file resident on the device. */
typedef struct {
    const uint8_t *data; /* device pointer, 16-byte aligned */
    uint64_t len;        /* bytes; ceil(len / slice_size) slices */
} p2g_dev_file;

/* Encode every slice of `files` (file order, slice order) into rec
 * (device, recovery_count * slice_size bytes, exponent-major).
 * first_input_index is the global index of the first slice (0 for a whole
 * set) and total_inputs the N that fixes the input bases.  accumulate = 0
 * overwrites rec, 1 XORs into it. */
int p2g_dev_encode(p2g_ctx *ctx, const p2g_dev_file *files, uint32_t nfiles,
                   uint64_t slice_size, uint32_t first_input_index, uint32_t total_inputs,
                   uint32_t recovery_count, uint8_t *rec, int accumulate, void *stream);

/* Per-slice MD5 (16 B each) and CRC-32 of every slice of `files`, written to
 * device arrays in file order then slice order. */
int p2g_dev_slice_checksums(p2g_ctx *ctx, const p2g_dev_file *files, uint32_t nfiles,
                            uint64_t slice_size, uint8_t *md5_out, uint32_t *crc_out,
                            void *stream);

/* Packet MD5s of `count` recovery packets whose blocks are on the device. */
int p2g_dev_recovery_packet_md5(p2g_ctx *ctx, const uint8_t *rec, uint32_t count,
                                uint64_t slice_size, uint32_t first_exponent,
                                const uint8_t set_id[16], uint8_t *pkt_md5_out,
                                void *stream);

/* One device pass of the create hot path over a whole set resident in HBM:
 * the recovery encode (as p2g_dev_encode with first_input_index 0,
 * total_inputs = the slice count, accumulate 0), the per-slice MD5 and
 * CRC-32 (as p2g_dev_slice_checksums) and the packet hashes of the R
 * recovery packets (as p2g_dev_recovery_packet_md5 with first_exponent 0),
 * with the packet hashes following the encode band by band inside one
 * launch.  Replaces, for one set, the scan checksums (src/source.rs:165-220),
 * encode_all_recovery_blocks (src/creator.rs:883-1018) and the build_packet
 * hashes of build_volume_packet (src/creator.rs:1031-1041) together.
 * Outputs: rec (R * slice_size), slice_md5_out (16 B per slice),
 * slice_crc_out (one per slice), pkt_md5_out (16 B per packet), all device
 * memory.  A set of more than 2048 slices (or whose coefficient fragments
 * exceed the pass's budget) runs as a one-job batched launch
 * (p2g_dev_create_batch), with the same results.  p2g_synchronize reports a
 * pass that could not complete. */
int p2g_dev_create_pass(p2g_ctx *ctx, const p2g_dev_file *files, uint32_t nfiles,
                        uint64_t slice_size, uint32_t recovery_count, const uint8_t set_id[16],
                        uint8_t *rec, uint8_t *slice_md5_out, uint32_t *slice_crc_out,
                        uint8_t *pkt_md5_out, void *stream);

/* p2g_dev_create_pass without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  Same arguments and results as
 * p2g_dev_create_pass for the sets one fused pass holds; everything beyond
 * the caller's inputs and outputs (slice descriptors built on the device,
 * tail blocks, zero-padded copies of partial slices, coefficient fragments,
 * sync words) lives in one static 64 MiB device array per GPU, kept until
 * the process ends.  Limits: at most 64 files, 1..2048 slices, R >= 1, and
 * a scratch need within 64 MiB (coefficient fragments take
 * ceil(R / 16) * ceil(slices / 16) * 8 KiB; the 1027-slice, R = 200
 * headline set takes 28 MB); larger sets fail with P2G_ERR_INVALID_ARGUMENT
 * and a message naming the limit.  The call makes no hipMalloc, hipFree or
 * host staging copy.  Calls on one GPU share the array: calls of one
 * context are ordered across its streams, calls of different contexts (or
 * processes' libraries) that overlap on one GPU race.  p2g_synchronize
 * reports a pass that could not complete. */
int p2g_dev_create_pass_noalloc(p2g_ctx *ctx, const p2g_dev_file *files, uint32_t nfiles,
                                uint64_t slice_size, uint32_t recovery_count,
                                const uint8_t set_id[16], uint8_t *rec, uint8_t *slice_md5_out,
                                uint32_t *slice_crc_out, uint8_t *pkt_md5_out, void *stream);

/* Diagnostics, no reference counterpart: after the last p2g_dev_create_pass
 * of this context has finished (it waits for that), the first start and the
 * last end of its packet-MD5, slice-MD5, encode and CRC work, in ms from the
 * earliest start: ms_out[2k] start and ms_out[2k + 1] end of role k, 0 for a
 * role with no work (all 0 when the last pass ran as separate kernels or as
 * a batched launch, or when none has run). */
int p2g_pass_timing(p2g_ctx *ctx, double *ms_out);

/* Diagnostics, no reference counterpart; host only (needs no device).  The
 * role of every workgroup of the persistent launch the planner makes for one
 * set of `slices` slices of slice_size bytes and recovery_count blocks, on
 * a device of num_cus CUs that holds max_resident workgroups of the kernel
 * at once: kind 0 the fused pass of p2g_dev_create_pass, kind 1 a one-job
 * batched plan.  roles_out[b] (up to cap entries) = 0 packet hash, 1 slice
 * hash, 2 encode; grid_out = workgroups of the launch.  The encode
 * workgroups come first, so that a partly resident grid runs its producers
 * before the roles that wait on them.  P2G_ERR_INVALID_ARGUMENT for a shape
 * the planner does not take (kind 0: a set the fused pass does not hold). */
int p2g_plan_roles(int kind, uint32_t num_cus, uint32_t max_resident, uint32_t slices,
                   uint32_t recovery_count, uint64_t slice_size, uint32_t *roles_out, uint32_t cap,
                   uint32_t *grid_out);

/* ---------------------------------------------------------------------
 * Batched create (no reference counterpart: the reference runs one
 * run_create at a time, src/creator.rs:188-321).  Many independent sets,
 * each exactly the work of one p2g_dev_create_pass, share persistent
 * launches: one table of slices, A fragments and work items (recovery row
 * block x 512-byte column group) over all jobs, so small jobs fill the
 * device together and their serial hashes overlap.  For callers that
 * create many sets at once (a posting service, a backup tool packaging
 * many directories).
 * ------------------------------------------------------------------- */

/* One create job of a batch: what p2g_dev_create_pass takes. */
typedef struct {
    const p2g_dev_file *files; /* the set's files, device resident, 16-byte aligned */
    uint64_t slice_size;
    uint8_t *rec;              /* recovery_count * slice_size bytes, device */
    uint8_t *slice_md5;        /* 16 B per slice, device */
    uint32_t *slice_crc;       /* one per slice, device */
    uint8_t *pkt_md5;          /* 16 B per recovery packet, device */
    uint32_t nfiles;
    uint32_t recovery_count;
    uint8_t set_id[16];
} p2g_batch_job;

typedef struct p2g_batch p2g_batch;

/* Plans a batch: validates every job (as p2g_dev_create_pass does, plus:
 * a job needs at least one slice), builds and uploads the descriptor, piece
 * and work tables and, with use_graph != 0, records the launch sequence as
 * a HIP graph.  The files and outputs must stay allocated while the plan
 * is launched; their contents may change between launches. */
int p2g_batch_plan(p2g_ctx *ctx, const p2g_batch_job *jobs, uint32_t njobs, int use_graph,
                   p2g_batch **out);
/* Queues one run of the plan on `stream` (NULL: the context's stream).
 * Runs of one plan are serialized. */
int p2g_batch_launch(p2g_batch *b, void *stream);
/* Waits for the last run; reports a run that could not complete. */
int p2g_batch_wait(p2g_batch *b);
/* After p2g_batch_wait: each job's completion time in ms from the start
 * of the run (device clock), in the order of the jobs given to the plan. */
int p2g_batch_job_ms(p2g_batch *b, double *ms_out);
void p2g_batch_free(p2g_batch *b);
/* Plan, queue and, at the next p2g_synchronize, free: one batched run. */
int p2g_dev_create_batch(p2g_ctx *ctx, const p2g_batch_job *jobs, uint32_t njobs, void *stream);

/* ---------------------------------------------------------------------
 * Serving (no reference counterpart): a request loop for a service that
 * creates PAR2 sets on demand.  Requests are admitted in arrival order
 * while their footprint (files, recovery blocks, digests) fits an HBM
 * budget taken at open (a request that does not fit lets later ones that
 * fit go ahead of it for a bounded time, p2g_server_set_bypass);
 * admitted requests stay resident and run in
 * continuous batching steps (batched launches over every admitted job with
 * work left; a job longer than the step target runs a piece of its
 * recovery rows per step), so requests that arrive while the device is
 * busy join the next step.  Each job's results are those of
 * p2g_dev_create_pass on the same input.
 * ------------------------------------------------------------------- */

typedef struct p2g_server p2g_server;

/* One create request.  File f is host_data[f] (file_lens[f] bytes, kept
 * valid until the request is done) or, with host_data NULL, the xorshift64
 * stream of fill_seeds[f] (p2g_dev_fill_xorshift64) written on the device. */
typedef struct {
    const uint64_t *file_lens;
    const uint8_t *const *host_data;
    const uint64_t *fill_seeds;
    uint64_t slice_size;
    uint32_t nfiles;
    uint32_t recovery_count;
    uint8_t set_id[16];
} p2g_request;

/* hbm_budget: bytes of device memory for admitted requests (allocated at
 * open, with the steps' 16 GiB A-fragment buffer, or half the free memory
 * if that is less); step_ms: target device time of one step (0: 35 ms). */
int p2g_server_open(p2g_ctx *ctx, uint64_t hbm_budget, double step_ms, p2g_server **out);
/* How long a queued request that finds no memory may be overtaken by later
 * requests that fit, from when it first found none (default: 4 step
 * targets, 140 ms at the default target; 0: strict arrival order).  After
 * that it holds back every later request until it is admitted. */
int p2g_server_set_bypass(p2g_server *s, double bypass_ms);
/* Queues a request; id_out identifies it.  Fails at once for a request
 * that could never be admitted. */
int p2g_server_submit(p2g_server *s, const p2g_request *req, uint64_t *id_out);
/* Drives the loop: retires finished steps, admits queued requests, launches
 * the next steps; returns after wait_ms or as soon as requests are done,
 * with up to cap of their ids in done_ids. */
int p2g_server_poll(p2g_server *s, double wait_ms, uint64_t *done_ids, uint32_t cap,
                    uint32_t *n_done);
/* A done request's outputs (device pointers, as a p2g_batch_job) and its
 * submit, admit and done times in ms on the server's clock. */
int p2g_server_result(p2g_server *s, uint64_t id, p2g_batch_job *out, double times_ms[3]);
/* Returns a done request's memory to the budget. */
int p2g_server_release(p2g_server *s, uint64_t id);
int p2g_server_stats(p2g_server *s, uint32_t *queued, uint32_t *active, uint64_t *hbm_used);
void p2g_server_close(p2g_server *s);

/* ---------------------------------------------------------------------
 * Autotuner (no reference counterpart; the reference picks its SIMD kernel
 * by CPU feature, src/galois_simd.rs:605-636).  Per (N, R, slice size)
 * bucket -- the bit lengths of the three -- the create-pass variant
 * (P2G_VARIANT_MFMA: matrix-core fused pass; P2G_VARIANT_BITSLICE /
 * P2G_VARIANT_NIBBLE: separate launches with the bit-sliced or the
 * split-nibble byte-permute VALU encode) and the column band width of the
 * fused pass that ran fastest, and the fastest encode kernel alone (mfma,
 * bitslice or nibble), which p2g_dev_encode and p2g_encode_push use.
 * p2g_dev_create_pass and the batched planner use a tuned bucket's choice;
 * with P2G_AUTOTUNE=1 a bucket is tuned on its first p2g_dev_create_pass
 * call (on that call's data; the create-pass choice only); P2G_TUNE_CACHE
 * names a file the cache is loaded from at p2g_open and saved to after each
 * tuning.
 * ------------------------------------------------------------------- */
enum { P2G_VARIANT_MFMA = 0, P2G_VARIANT_BITSLICE = 1, P2G_VARIANT_NIBBLE = 2 };

/* The bucket id of a shape: nbits(slices) | nbits(R) << 8 | nbits(S) << 16. */
uint32_t p2g_tune_bucket(uint64_t slices, uint32_t recovery_count, uint64_t slice_size);
/* Tunes the bucket of (slices, recovery_count, slice_size) on synthetic
 * input of that shape -- the encode kernels alone, then the create-pass
 * candidates -- and caches both choices; the create-pass winner's variant,
 * band width (0: default) and time in ms are returned when the pointers are
 * set (p2g_tune_get_encode gives the encode kernel's). */
int p2g_autotune(p2g_ctx *ctx, uint32_t slices, uint32_t recovery_count, uint64_t slice_size,
                 int *variant_out, uint32_t *band_cols_out, double *ms_out);
/* The choice p2g_dev_create_pass would use for the shape now. */
int p2g_tune_get(p2g_ctx *ctx, uint32_t slices, uint32_t recovery_count, uint64_t slice_size,
                 int *variant_out, uint32_t *band_cols_out);
/* The encode kernel (P2G_VARIANT_*) p2g_dev_encode and p2g_encode_push use
 * for the shape now, and its tuned time in ms (0 when not tuned). */
int p2g_tune_get_encode(p2g_ctx *ctx, uint32_t slices, uint32_t recovery_count,
                        uint64_t slice_size, int *variant_out, double *ms_out);
int p2g_tune_save(p2g_ctx *ctx, const char *path);
int p2g_tune_load(p2g_ctx *ctx, const char *path);

/* Synthetic input for benchmarks and tests (no reference counterpart):
 * fills nbytes of dst (device, 16-byte aligned) with the xorshift64 stream
 * from `seed` (x ^= x << 13; x ^= x >> 7; x ^= x << 17; each new state
 * emitted as 8 little-endian bytes), the documented workload generator. */
int p2g_dev_fill_xorshift64(p2g_ctx *ctx, uint8_t *dst, uint64_t nbytes, uint64_t seed,
                            void *stream);

/* Waits for all work queued by this context, on its own streams and on the
 * streams callers passed in. */
int p2g_synchronize(p2g_ctx *ctx);

/* ---------------------------------------------------------------------
 * Whole-pipeline entry point: SourceFile::scan for every input, then
 * run_create.  Replaces the library caller of examples/create_from_lib.rs:
 * 12-41 (scan, src/source.rs:53, then run_create, src/creator.rs:182).
 * ------------------------------------------------------------------- */

enum { P2G_SCHEME_SINGLE = 0, P2G_SCHEME_EXPONENTIAL = 1, P2G_SCHEME_UNIFORM = 2,
       P2G_SCHEME_EXPLICIT = 3 };

/* VolumeScheme, src/creator.rs:110-139.  `kind` is the base scheme; with
 * limited != 0 it is wrapped in Limited { max_blocks_per_volume = limit_cap,
 * inner }, and that in n_outer_caps further Limited wrappers whose caps are
 * outer_caps[0] (the next one out) .. outer_caps[n_outer_caps - 1] (the
 * outermost), as Limited nests through its boxed inner scheme. */
typedef struct {
    int kind;                        /* P2G_SCHEME_*: the innermost scheme */
    uint32_t uniform_count;          /* Uniform { count } */
    int limited;                     /* Limited { max_blocks_per_volume, inner } */
    uint32_t limit_cap;              /* max_blocks_per_volume */
    const uint32_t *explicit_sizes;  /* Explicit(sizes) */
    uint32_t n_explicit;
    const uint32_t *outer_caps;      /* further Limited wrappers, innermost first */
    uint32_t n_outer_caps;           /* 0: none (needs limited != 0 otherwise) */
} p2g_volume_scheme;

/* Writes the index file `output` and its volume files.  names_out receives
 * the NUL-separated paths written, index first (capacity names_cap). */
int p2g_run_create(p2g_ctx *ctx, const char *output, uint64_t slice_size,
                   uint32_t recovery_count, const p2g_volume_scheme *scheme,
                   uint32_t nfiles, const char *const *paths,
                   const uint8_t *const *names, const uint32_t *name_lens,
                   char *names_out, size_t names_cap);

#ifdef __cplusplus
}
#endif
#endif
