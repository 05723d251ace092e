// p2g_internal.hpp -- types shared by the kernels, the C ABI and the host
// pipeline of libpar2gpu.so.  Not part of the public ABI (include/par2gpu.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/par2gpu.h"

namespace p2g {

// One input slice as the encode and checksum kernels see it: `ptr` points at
// the slice's first byte on the device, bytes [valid, slice_size) read as
// zero (the zero padding of a file's trailing partial slice,
// src/creator.rs:79-85 and src/source.rs:174-183).
struct SliceDesc {
    const uint8_t *ptr;
    uint32_t valid;
    uint32_t tail;  // bytes [valid & ~3, valid) as a zero-padded word (set on the device)
};
static_assert(sizeof(SliceDesc) == 16, "SliceDesc must stay 16 bytes");

// GF(2^16) field constants, src/galois.rs:15-18.
constexpr uint32_t kFieldLimit = 65535u;
constexpr uint32_t kFieldGenerator = 0x1100Bu;

// Launch wrappers implemented in p2g_kernels.hip.  All are asynchronous on
// `stream` and return hipSuccess or the launch error.

// mats[(i * R + r) * 16 + j] = row j of the 16x16 GF(2) matrix of
// multiplication by coef(r, first_input + i) = base^r, base = antilog[L].
hipError_t launch_coeff_mats(const uint16_t *d_antilog, const uint16_t *d_logbase,
                             uint32_t n_inputs, uint32_t R, uint16_t *d_mats,
                             hipStream_t stream);

// rec[r][*] (^)= XOR_i coef(r, i) * slice_i over the n_inputs descriptors.
hipError_t launch_encode(const SliceDesc *d_inputs, uint32_t n_inputs, const uint16_t *d_mats,
                         uint32_t R, uint32_t S, uint8_t *d_rec, int accumulate,
                         hipStream_t stream);

// Split-nibble form: 96 bytes of byte-permute tables per coefficient
// (tables[(i * R + r) * 24 ..]) for inputs first_input + i, rows 0..R-1, then
// rec[r][*] (^)= XOR_i coef(r, i) * slice_i from them.
constexpr uint32_t kNibbleTableBytes = 96;
hipError_t launch_nibble_tables(const uint16_t *d_antilog, const uint16_t *d_logbase,
                                uint32_t n_inputs, uint32_t R, uint32_t *d_tables,
                                hipStream_t stream);
hipError_t launch_encode_nibble(const SliceDesc *d_inputs, uint32_t n_inputs, const uint32_t *d_tables,
                                uint32_t R, uint32_t S, uint8_t *d_rec, int accumulate,
                                hipStream_t stream);

// Fills SliceDesc::tail of n descriptors in place.
// With d_tail_blocks (64 bytes per descriptor, may be null), also writes each
// slice's zero-padded 64-byte block holding byte `valid` (for slice MD5).
hipError_t launch_desc_tails(SliceDesc *d_descs, uint32_t n, uint4 *d_tail_blocks,
                             hipStream_t stream);

// Encode unit geometry of the matrix-core kernels: a workgroup of
// kMfThreads threads (4 waves) computes kMfTiles 32-row MFMA tiles (16
// recovery rows) over 512 bytes of every slice (256 symbols); the fused
// pass holds at most kMfMaxInputs slice descriptors in LDS.
constexpr int kMfTiles = 8;
constexpr int kMfThreads = 256;
constexpr int kMfMaxInputs = 2048;

// Matrix-core encode.  A fragments for `n` inputs (first input's L at
// d_logbase[0]) and R recovery rows: mfma_row_tiles(R) * mfma_steps(n) * 1 KiB.
// A K step covers four inputs; the step count is padded to whole groups of
// four steps with zero fragments (zero A contributes nothing).
__host__ __device__ inline uint32_t mfma_row_tiles(uint32_t R) { return (R + 15) / 16 * 8; }
__host__ __device__ inline uint32_t mfma_steps(uint32_t n) { return ((n + 3) / 4 + 3) / 4 * 4; }
constexpr uint32_t kMfmaMaxInputsPerLaunch = 2048;  // LDS descriptor table of k_encode_mfma
inline size_t mfma_afrag_bytes(uint32_t n, uint32_t R) {
    return (size_t)mfma_row_tiles(R) * mfma_steps(n) * 1024;
}
hipError_t launch_mfma_afrag(const uint16_t *d_antilog, const uint16_t *d_logbase, uint32_t n,
                             uint32_t R, uint4 *d_afrag, hipStream_t stream);
constexpr uint32_t kMfmaChunkInputs = 4096;  // inputs per K chunk of a batched unit (64 KiB of LDS)
hipError_t launch_encode_mfma(const SliceDesc *d_inputs, uint32_t n_inputs, const uint4 *d_afrag,
                              uint32_t R, uint32_t S, uint8_t *d_rec, int accumulate,
                              hipStream_t stream);

// Fused create pass (encode + slice MD5 + packet MD5, one launch).  The
// sync words (cleared by each launch): [kPassFlags] unused; from kPassTimes, 2 kPassRoles u64
// real-time counter stamps (100 MHz): the first start and the last end of
// each role's work (packet MD5, slice MD5, encode, CRC); the per-XCD work
// queue counters from kPassQueues; [kPassBands + b]
// finished units of column band b; [kPassBands + n_bands + b] the first
// column group of band b (n_bands + 1 entries).  Errors go to a separate
// sticky word (kPassErrWait bits) that no launch clears: the host reads and
// clears it when it waits, so an error of an earlier pass queued before the
// wait is not lost to a later pass's reset.
constexpr uint32_t kPassFlags = 0, kPassTimes = 2, kPassRoles = 4;
// [kPassQueues + x]: units taken so far from the work queue of XCD x (8)
constexpr uint32_t kPassQueues = kPassTimes + 4 * kPassRoles, kPassXcds = 8;
constexpr uint32_t kPassBands = kPassQueues + kPassXcds;
enum { kPassPacket = 0, kPassSlice = 1, kPassEncode = 2, kPassCrc = 3 };
constexpr uint32_t kPassErrWait = 1u;  // a packet lane stopped waiting for a band
// Width (in 512-byte column groups) of the next band after `done` groups,
// with `remaining` left: growing from one group (1, 1, 2, 4, ... up to W) so
// that the packet hashes can start early, then W, then halving toward the
// end so that the last band, which they can only start once it is complete,
// is one group again.
__host__ __device__ inline uint32_t next_band_width(uint32_t done, uint32_t remaining, uint32_t W) {
    const uint32_t ramp = done == 0 ? 1u : (done < W ? done : W);
    const uint32_t tail = remaining > 2 * W ? W : remaining > 1 ? (remaining + 1) / 2 : remaining;
    return ramp < tail ? ramp : tail;
}

// Unit v of a band of w column groups x rows row blocks, in tiles of
// kPassTileRows row blocks x tc column groups (the last tiles of each
// direction ragged), tile after tile and row-block-major inside a tile:
// its row block *y and column group *x within the band.
constexpr uint32_t kPassTileRows = 4;
__host__ __device__ inline void band_tile_unit(uint32_t v, uint32_t w, uint32_t rows, uint32_t tc,
                                               uint32_t *y, uint32_t *x) {
    const uint32_t rt = v / (kPassTileRows * w), rows_left = rows - rt * kPassTileRows;
    const uint32_t rows_in = rows_left < kPassTileRows ? rows_left : kPassTileRows;
    const uint32_t v1 = v - rt * kPassTileRows * w;
    const uint32_t ct = v1 / (tc * rows_in), cols_left = w - ct * tc;
    const uint32_t cols_in = cols_left < tc ? cols_left : tc;
    const uint32_t v2 = v1 - ct * tc * rows_in;
    *y = rt * kPassTileRows + v2 / cols_in;
    *x = ct * tc + v2 % cols_in;
}

// Band widths of a batched piece: one band when W covers every column
// group (a short row, whose packet hashes gain nothing from early bands),
// else next_band_width.
__host__ __device__ inline uint32_t batch_band_width(uint32_t done, uint32_t remaining, uint32_t W) {
    return done == 0 && W >= remaining ? remaining : next_band_width(done, remaining, W);
}

// Role of workgroup `blk` of a persistent create launch (fused or batched)
// of `grid` workgroups, h2 of them hashing slices and h3 hashing recovery
// packets.  The encode workgroups, the producers, take the lowest indices
// [0, grid - h2 - h3); the slice hashes, which wait for nothing, follow; the
// packet hashes, which wait for the encode's bands, take the highest.  The
// dispatcher hands out a grid's workgroups in index order, so when fewer
// CUs are free than the grid has workgroups (another process on the GPU),
// every producer is dispatched before any waiter and the waiters never hold
// the CUs the producers need.  `index` is the workgroup's rank in its role.
struct PassRole {
    uint32_t role, index;
};
__host__ __device__ inline PassRole pass_role(uint32_t blk, uint32_t grid, uint32_t h2,
                                              uint32_t h3) {
    const uint32_t e = grid - h2 - h3;
    if (blk < e) return {(uint32_t)kPassEncode, blk};
    if (blk < e + h2) return {(uint32_t)kPassSlice, blk - e};
    return {(uint32_t)kPassPacket, blk - e - h2};
}

struct PassPlan {
    uint32_t grid = 0, h2_blocks = 0, h3_blocks = 0;
    uint32_t row_blocks = 0, col_groups = 0, band_cols = 0, n_bands = 0, n_units = 0;
};
// Default band width of the fused pass, in 512-byte column groups: the
// packet lanes hash a band once it is encoded, so the last full band is
// hashed after the encode ends; 32 column groups (16 KiB of each packet,
// 256 MD5 blocks) keep that tail short at the 512 KiB-slice headline shape,
// where 24-32 measured 5.17-5.20 ms per pass against 5.31-5.35 at 64.
constexpr uint32_t kPassDefaultBand = 32;
// False when the fused pass does not apply (no inputs or recovery blocks,
// more inputs than one encode launch takes, too few CUs).  band_cols = 0
// picks the default band width.  max_resident: how many workgroups of the
// pass the device holds at once (occupancy x CUs); the grid is at most that.
bool plan_create_pass(uint32_t n, uint32_t R, uint32_t S, uint32_t num_cus, uint32_t max_resident,
                      uint32_t band_cols, PassPlan *p);
// Workgroups per CU of the persistent kernels as compiled (their LDS and
// registers allow), from the HIP occupancy calculator; 0 when it fails.
int pass_occupancy();
int batch_occupancy();
uint32_t pass_sync_words(const PassPlan &p);
// Zeroes the sync words, then runs the pass.  d_inputs must have their tail
// words (launch_desc_tails) and d_afrag the fragments of all n inputs.
hipError_t launch_create_pass(const PassPlan &p, const SliceDesc *d_inputs,
                              const uint4 *d_tail_blocks, uint32_t n, const uint4 *d_afrag,
                              uint32_t R, uint32_t S, uint8_t *d_rec,
                              uint8_t *d_slice_md5, uint32_t *d_slice_crc, const uint8_t set_id[16],
                              uint8_t *d_pkt_md5, uint32_t *d_sync, uint32_t *d_err,
                              bool all_full, hipStream_t stream);
// ---- batched create pass ------------------------------------------------
//
// Many create jobs in one persistent launch per group of jobs.  A job's
// recovery rows may be split into pieces (16-row multiples) so that the A
// fragments of a group stay within a budget; every piece has its own
// fragments, units, band counters and packet tasks.  Tables live in device
// memory and are read by the kernel.

// One piece: rows [first_exp, first_exp + R) of a job.  Units of the group
// [unit0, unit_end); packet wave tasks [task0, task_end) (64 packets each);
// band counters at sync[band0 + b], band starts at sync[bstart0 + b].
struct BatchPiece {
    const SliceDesc *inputs;  // the job's n descriptors, all whole slices
    const void *afrag;        // fragments of these rows (uint4), mfma_steps(n) steps
    uint8_t *rec;             // block of row 0 of the piece
    uint8_t *pkt_md5;         // digest of row 0 of the piece
    uint32_t pre[8];          // set id and packet type words of the packet hash
    uint32_t n, nq, R, S;
    uint32_t first_exp, job, row_blocks, col_groups;
    uint32_t band_cols, n_bands, unit0, unit_end;
    uint32_t task0, task_end, band0, bstart0;
    // Column groups [cg0, cg0 + col_groups) of the rows (all of them unless
    // a serving step split the rows' columns over several launches).  A
    // piece that starts past column group 0 continues each row's packet hash
    // from md5_state; one that ends before the last group saves it there.
    uint4 *md5_state;         // chaining state of row 0 of the piece (or null)
    uint32_t cg0, cg_total;   // first column group; column groups of a whole row
};
static_assert(sizeof(BatchPiece) == 144, "BatchPiece layout");

// One job's slices: global slice indices [slice0, slice_end) of the batch
// descriptor table.
struct BatchSlices {
    uint8_t *md5;
    uint32_t *crc;
    uint32_t slice0, slice_end, S, job;
};
static_assert(sizeof(BatchSlices) == 32, "BatchSlices layout");

// A partial slice copied into a zero-padded S-byte slot.
struct StagePart {
    const uint8_t *src;
    uint8_t *dst;
    uint32_t valid, S;
    uint64_t pad;
};

struct BatchGroupLaunch {
    const BatchPiece *pieces;   // the group's pieces (device)
    uint32_t n_pieces;
    const BatchSlices *sjobs;   // the group's slice jobs (device)
    uint32_t n_sjobs;
    uint32_t slice_lo, slice_hi;  // global slice range of the group
    uint32_t n_units, n_tasks;
    uint32_t h2_blocks, h3_blocks, grid;
    uint32_t *sync;             // group sync words
    uint32_t *err;              // the plan's sticky error word (not cleared by launches)
    uint32_t sync_words;        // flags, times, counters (zeroed by the reset)
    uint32_t n_jobs;            // whole batch
    unsigned long long *job_done;  // n_jobs completion stamps, then the batch start
    bool first;                 // resets job_done too
    uint64_t afrag_entries;     // uint4 entries of the group's fragments
};

// Zero-padded copies of partial slices (one launch for all of them).
hipError_t launch_stage_parts(const StagePart *d_parts, uint32_t nparts, uint32_t max_S,
                              hipStream_t stream);
// A fragments of every piece of a group (pieces[k].afrag), exponent first_exp + r.
hipError_t launch_batch_afrag(const uint16_t *d_antilog, const uint16_t *d_logbase,
                              const BatchGroupLaunch &g, hipStream_t stream);
// Sync reset, band starts, then the persistent launch of one group.
hipError_t launch_create_batch(const BatchGroupLaunch &g, const SliceDesc *d_descs,
                               const uint4 *d_tail_blocks, hipStream_t stream);
// The work of one launch group, for sizing its roles: encode unit K steps
// (units x steps of their job), the serial MD5 work of the slice lanes in
// 64-byte blocks and in slices, the count of the largest slices, and per
// piece in launch order its unit-steps, its packet tasks and the blocks of
// one task's hash.
struct BatchGroupWork {
    double enc_steps = 0;
    double slice_blocks = 0, slices = 0;
    double big_slices = 0, big_blocks = 0;  // the slices of the largest size and their blocks
    double small_blocks = 0;                // blocks of a slice of the smallest size
    std::vector<double> piece_steps, piece_tasks, piece_task_blocks;
};
// What the planners estimate times from (µs): an encode unit's K step on
// one workgroup; a 64-byte block of a slice lane (MD5 and CRC-32) and of a
// packet lane; the fixed cost of a slice and of a packet task.  A context
// starts from values measured on MI355X at the batched mix and updates them
// from the role stamps of every batched run it finishes (update_rates).
struct CostRates {
    double step_us = 0.38, slice_block_us = 1.31, slice_us = 36;
    double packet_block_us = 0.65, task_us = 40;
    uint32_t updates = 0;  // runs that updated the rates
};
// Folds one finished group's measured role times (µs from its first start
// to its last end; 0 when the role had no work) into the rates.
void update_rates(CostRates &r, const BatchGroupWork &w, uint32_t grid, uint32_t h2, uint32_t h3,
                  double enc_us, double slice_us);
// Group sizing: MD5 workgroups for n_tasks packet tasks and n_slices slices,
// as few as are estimated to finish within the encode (the rest encode).
void plan_batch_group(uint32_t num_cus, uint32_t max_resident, uint32_t n_tasks, uint32_t n_slices,
                      uint32_t n_units, const BatchGroupWork &w, const CostRates &rates,
                      BatchGroupLaunch *g);
constexpr uint32_t kBatchProgress = 1;  // sync word: units published so far
// Sync word of a batched group: slice tasks (64 consecutive slices) taken so
// far by the slice-hash waves (the fused pass leaves this word unused).
constexpr uint32_t kBatchSliceQueue = kPassFlags;
// MD5 + CRC-32 waves of a batched slice-hash workgroup (64 lanes each, one
// slice per lane; each wave keeps 16 KiB of CRC tables in the workgroup's LDS).
constexpr uint32_t kBatchSliceWaves = 4;
constexpr uint32_t kBatchSliceLanes = 64 * kBatchSliceWaves;

// Synthetic input: the xorshift64 stream from `seed` into d_dst.  d_jumps
// holds the 32 GF(2) matrices step^(kXorshiftLaneWords 2^b) as byte tables
// (kXsJumpTableWords each); the destination must be 16-byte aligned.  A lane writes kXorshiftLaneWords
// words (1 KiB) after jumping to its start, so a file of a few MB is already
// thousands of lanes (with 16 KiB per lane a 13 MB file ran on four
// workgroups at ~100 GB/s).
constexpr uint32_t kXorshiftLaneWords = 128;
// A jump step^(W 2^b) as 8 tables of 256 words: table q maps byte q of a
// state to that byte's share of the jumped state.
constexpr uint32_t kXsJumpTableWords = 8 * 256;
hipError_t launch_xorshift_fill(uint8_t *d_dst, uint64_t nbytes, uint64_t seed,
                                const uint64_t *d_jumps, hipStream_t stream);
// Up to kXsFillFiles files in one launch, passed by value (kernel
// arguments): file k is nbytes[k] bytes of the stream from seed[k] at dst[k]
// (16-byte aligned), its lanes numbered from lane0[k]; lane0[n] is the total.
constexpr uint32_t kXsFillFiles = 48;
struct XsFillFiles {
    uint8_t *dst[kXsFillFiles];
    uint64_t nbytes[kXsFillFiles];
    uint64_t seed[kXsFillFiles];
    uint64_t lane0[kXsFillFiles + 1];
    uint32_t n;
};
hipError_t launch_xorshift_fill_files(const XsFillFiles &f, const uint64_t *d_jumps,
                                      hipStream_t stream);

// Copies the nparts partial slices described by d_parts (ptr, valid, and in
// `tail` the staging slot) into zero-padded S-byte slots of d_staging.
hipError_t launch_stage_partials(const SliceDesc *d_parts, uint32_t nparts, uint32_t S,
                                 uint8_t *d_staging, hipStream_t stream);

// MD5 of each zero-padded S-byte slice (16 bytes per slice into d_md5).
// d_tail_blocks: the 64-byte tail blocks of launch_desc_tails.
hipError_t launch_slice_md5(const SliceDesc *d_slices, const uint4 *d_tail_blocks, uint32_t n,
                            uint32_t S, uint8_t *d_md5, hipStream_t stream);

// CRC-32 of each zero-padded S-byte slice.
hipError_t launch_slice_crc32(const SliceDesc *d_slices, uint32_t n, uint32_t S,
                              uint32_t *d_crc, hipStream_t stream);

// Packet MD5 of recovery packets r = 0..count-1: MD5(set_id || "PAR 2.0\0RecvSlic"
// || le32(first_exp + r) || rec[r * S .. r * S + S)).
hipError_t launch_recovery_md5(const uint8_t *d_rec, uint32_t count, uint32_t S,
                               uint32_t first_exp, const uint8_t set_id[16],
                               uint8_t *d_md5, hipStream_t stream);

}  // namespace p2g
