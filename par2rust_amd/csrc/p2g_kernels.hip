// p2g_kernels.hip -- gfx950 kernels of the PAR2 create path.
//
//   k_coeff_mats     RS coefficients base_i^r from the log/antilog tables
//                    (src/reedsolomon.rs:56-60, src/galois.rs:74-87) expanded
//                    to the 16x16 GF(2) matrix of multiplication by each one
//   k_encode_bits    GF(2^16) recovery accumulator (src/creator.rs:934-995)
//                    on bit-sliced symbols
//   k_slice_md5      IFSC slice MD5, one lane per slice (src/source.rs:168-183)
//   k_slice_crc32    IFSC slice CRC-32, one wave per slice
//   k_recovery_md5   recovery packet hash, one lane per packet
//                    (src/format.rs:40-44 via src/packet/recovery.rs:11-19)
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <type_traits>

#include "p2g_internal.hpp"

namespace p2g {

// ======================================================================
// Coefficient matrices
// ======================================================================

// Thread per (input i, recovery r).  coef = antilog[(L_i * r) mod 65535]
// equals gf_pow(base_i, r) for base_i = antilog[L_i] (r = 0 gives
// antilog[0] = 1).  Column k of the matrix is coef * x^k; row j collects
// bit j of every column, so out_bit_j = parity(row_j & in_symbol).
__global__ __launch_bounds__(256) void k_coeff_mats(const uint16_t *__restrict__ antilog,
                                                    const uint16_t *__restrict__ logbase,
                                                    uint32_t n, uint32_t R,
                                                    uint16_t *__restrict__ mats) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (uint64_t)n * R) return;
    const uint32_t i = (uint32_t)(idx / R), r = (uint32_t)(idx % R);
    const uint32_t e = (uint32_t)(((uint64_t)logbase[i] * r) % kFieldLimit);
    uint32_t v = antilog[e];
    uint32_t col[16];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        col[k] = v;
        v = ((v << 1) & 0xFFFFu) ^ ((v & 0x8000u) ? (kFieldGenerator & 0xFFFFu) : 0u);
    }
    uint16_t *out = mats + idx * 16;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        uint32_t row = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) row |= ((col[k] >> j) & 1u) << k;
        out[j] = (uint16_t)row;
    }
}

hipError_t launch_coeff_mats(const uint16_t *d_antilog, const uint16_t *d_logbase,
                             uint32_t n_inputs, uint32_t R, uint16_t *d_mats,
                             hipStream_t stream) {
    const uint64_t total = (uint64_t)n_inputs * R;
    if (total == 0) return hipSuccess;
    const uint64_t blocks = (total + 255) / 256;
    hipLaunchKernelGGL(k_coeff_mats, dim3((uint32_t)blocks), dim3(256), 0, stream, d_antilog,
                       d_logbase, n_inputs, R, d_mats);
    return hipGetLastError();
}

// ======================================================================
// Bounded loads: bytes at or past `valid` read as zero.
// ======================================================================

// Slice pointers arrive in descriptors as generic pointers, and a load
// through a generic pointer is a flat_load, which the wait counters track
// as both a vector-memory and an LDS operation: every LDS wait after it
// would also wait for the load.  All source data is global memory, so loads
// go through the global address space (global_load, vmcnt only).
template <typename T>
__device__ __forceinline__ T gload(const T *p) {
    return *(const __attribute__((address_space(1))) T *)p;
}
// uint4's copy constructor takes a generic reference, so it is loaded as a
// native four-lane vector instead.
__device__ __forceinline__ uint4 gload(const uint4 *p) {
    typedef uint32_t v4u_t __attribute__((ext_vector_type(4)));
    const v4u_t v = *(const __attribute__((address_space(1))) v4u_t *)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ uint32_t load_word(const uint8_t *p, uint32_t off, uint32_t valid) {
    if (off + 4 <= valid) return gload(reinterpret_cast<const uint32_t *>(p + off));
    uint32_t w = 0;
    for (uint32_t b = 0; b < 4; b++)
        if (off + b < valid) w |= (uint32_t)gload(p + off + b) << (8 * b);
    return w;
}

// Tail word of each descriptor: the bytes of a trailing partial dword, so
// that word loads never touch memory past `valid`.  With tail_blocks, also
// the 64-byte block of the slice that holds byte `valid` (block valid / 64),
// zero padded: MD5 lanes read it in place of the slice, so every lane of a
// wave runs the same number of whole-block steps.
__global__ __launch_bounds__(256) void k_desc_tails(SliceDesc *__restrict__ d, uint32_t n,
                                                    uint4 *__restrict__ tail_blocks) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = d[i].valid;
    d[i].tail = (v & 3) ? load_word(d[i].ptr, v & ~3u, v) : 0u;
    if (tail_blocks) {
        const uint32_t base = v / 64 * 64;
#pragma unroll
        for (int q = 0; q < 4; q++)
            tail_blocks[4 * (size_t)i + q] = make_uint4(
                load_word(d[i].ptr, base + 16 * q, v), load_word(d[i].ptr, base + 16 * q + 4, v),
                load_word(d[i].ptr, base + 16 * q + 8, v), load_word(d[i].ptr, base + 16 * q + 12, v));
    }
}

hipError_t launch_desc_tails(SliceDesc *d_descs, uint32_t n, uint4 *d_tail_blocks,
                             hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_desc_tails, dim3((n + 255) / 256), dim3(256), 0, stream, d_descs, n,
                       d_tail_blocks);
    return hipGetLastError();
}

// Partial slices copied into zero-padded S-byte staging slots: thread per
// staging word, parts[k] = {source, valid bytes, slot}.
__global__ __launch_bounds__(256) void k_stage_partials(const SliceDesc *__restrict__ parts,
                                                        uint32_t nparts, uint32_t S,
                                                        uint8_t *__restrict__ staging) {
    const uint32_t words = S / 4;
    const uint64_t total = (uint64_t)nparts * words;
    for (uint64_t idx = (uint64_t)blockIdx.x * 256 + threadIdx.x; idx < total;
         idx += (uint64_t)gridDim.x * 256) {
        const uint32_t k = (uint32_t)(idx / words), w = (uint32_t)(idx % words);
        const SliceDesc d = parts[k];
        reinterpret_cast<uint32_t *>(staging + (size_t)d.tail * S)[w] = load_word(d.ptr, 4 * w, d.valid);
    }
}

hipError_t launch_stage_partials(const SliceDesc *d_parts, uint32_t nparts, uint32_t S,
                                 uint8_t *d_staging, hipStream_t stream) {
    if (nparts == 0) return hipSuccess;
    const uint64_t total = (uint64_t)nparts * (S / 4);
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_stage_partials, dim3(blocks), dim3(256), 0, stream, d_parts, nparts, S,
                       d_staging);
    return hipGetLastError();
}

// A word that is always zero: the target of masked-off word loads.
__device__ const uint32_t kZeroWord = 0;

// ======================================================================
// K1: GF(2^16) accumulator on bit-sliced symbols
// ======================================================================
//
// A lane owns 32 consecutive symbols (64 bytes) of the slice column range.
// Its 16 input words are transposed into 16 bit planes (plane b = bit b of
// the 32 symbols), so multiplication by a coefficient becomes 16 output
// planes, each the XOR of the input planes selected by one matrix row.
// RT recovery rows stay in registers while the inputs stream past.

// Transposes the two 16x16 bit matrices held in the low and high halves of
// 16 words (row q = word q, column b = bit b).  Its own inverse.
__device__ __forceinline__ void transpose16(uint32_t w[16]) {
#pragma unroll
    for (int st = 0; st < 4; st++) {
        const int k = 8 >> st;
        const uint32_t m = (k == 8) ? 0x00FF00FFu : (k == 4) ? 0x0F0F0F0Fu
                         : (k == 2) ? 0x33333333u : 0x55555555u;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            if ((q & k) == 0) {
                const uint32_t t = ((w[q] >> k) ^ w[q + k]) & m;
                w[q + k] ^= t;
                w[q] ^= t << k;
            }
        }
    }
}

// acc ^ (x & m) in one v_bitop3_b32: truth-table bit (s0 << 2 | s1 << 1 | s2)
// of f(s0, s1, s2) = s0 ^ (s1 & s2) gives 0x78.
__device__ __forceinline__ uint32_t bitop3_xor_and(uint32_t acc, uint32_t x, uint32_t m) {
    return __builtin_amdgcn_bitop3_b32(acc, x, m, 0x78);
}

constexpr int kEncThreads = 256;
constexpr int kEncRT = 8;

template <bool VEC>
__global__ __launch_bounds__(kEncThreads) void k_encode_bits(
    const SliceDesc *__restrict__ inputs, uint32_t n_inputs, const uint16_t *__restrict__ mats,
    uint32_t R, uint32_t S, uint8_t *__restrict__ rec, int accumulate) {
    const uint32_t col = (blockIdx.x * kEncThreads + threadIdx.x) * 64u;
    if (col >= S) return;
    const uint32_t r0 = blockIdx.y * kEncRT;
    const uint32_t rt = (R - r0) < (uint32_t)kEncRT ? (R - r0) : (uint32_t)kEncRT;

    uint32_t acc[kEncRT][16];
#pragma unroll
    for (int rr = 0; rr < kEncRT; rr++)
#pragma unroll
        for (int j = 0; j < 16; j++) acc[rr][j] = 0;

    for (uint32_t i = 0; i < n_inputs; i++) {
        const SliceDesc d = inputs[i];
        uint32_t x[16];
        if (VEC && col + 64 <= d.valid) {
            const uint4 *p = reinterpret_cast<const uint4 *>(d.ptr + col);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint4 v = gload(p + q);
                x[4 * q] = v.x; x[4 * q + 1] = v.y; x[4 * q + 2] = v.z; x[4 * q + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const uint32_t off = col + 4u * q;
                x[q] = (off < S) ? load_word(d.ptr, off, d.valid) : 0u;
            }
        }
        transpose16(x);
        const uint16_t *m = mats + (size_t)i * R * 16 + (size_t)r0 * 16;
#pragma unroll
        for (int rr = 0; rr < kEncRT; rr++) {
            if ((uint32_t)rr < rt) {
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    const uint32_t row = m[rr * 16 + j];
                    uint32_t a = acc[rr][j];
#pragma unroll
                    for (int k = 0; k < 16; k++)
                        a = bitop3_xor_and(a, x[k], 0u - ((row >> k) & 1u));
                    acc[rr][j] = a;
                }
            }
        }
    }

#pragma unroll
    for (int rr = 0; rr < kEncRT; rr++) {
        if ((uint32_t)rr >= rt) continue;
        transpose16(acc[rr]);
        uint8_t *out = rec + (size_t)(r0 + rr) * S + col;
        if (VEC && col + 64 <= S) {
            uint4 *o = reinterpret_cast<uint4 *>(out);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                uint4 v = make_uint4(acc[rr][4 * q], acc[rr][4 * q + 1], acc[rr][4 * q + 2],
                                     acc[rr][4 * q + 3]);
                if (accumulate) {
                    const uint4 old = o[q];
                    v.x ^= old.x; v.y ^= old.y; v.z ^= old.z; v.w ^= old.w;
                }
                o[q] = v;
            }
        } else {
            uint32_t *o = reinterpret_cast<uint32_t *>(out);
#pragma unroll
            for (int q = 0; q < 16; q++) {
                if (col + 4u * q < S) o[q] = accumulate ? (o[q] ^ acc[rr][q]) : acc[rr][q];
            }
        }
    }
}

hipError_t launch_encode(const SliceDesc *d_inputs, uint32_t n_inputs, const uint16_t *d_mats,
                         uint32_t R, uint32_t S, uint8_t *d_rec, int accumulate,
                         hipStream_t stream) {
    if (R == 0) return hipSuccess;
    const uint32_t per_block = kEncThreads * 64u;
    dim3 grid((S + per_block - 1) / per_block, (R + kEncRT - 1) / kEncRT);
    // 128-bit loads and stores need 16-byte aligned slices and output rows;
    // slice pointers are checked by the caller, the output base here.
    if (S % 16 == 0 && reinterpret_cast<uintptr_t>(d_rec) % 16 == 0)
        hipLaunchKernelGGL(k_encode_bits<true>, grid, dim3(kEncThreads), 0, stream, d_inputs,
                           n_inputs, d_mats, R, S, d_rec, accumulate);
    else
        hipLaunchKernelGGL(k_encode_bits<false>, grid, dim3(kEncThreads), 0, stream, d_inputs,
                           n_inputs, d_mats, R, S, d_rec, accumulate);
    return hipGetLastError();
}

// ======================================================================
// K1 (split-nibble form): byte-permute table lookups
// ======================================================================
//
// The GF(2^16) product c * x is linear in the bits of x, so it splits over
// chunks of x: c * x = XOR_q T_q[chunk_q(x)], T_q[v] = c * (v << shift_q).
// The chunks here are three bits wide (bits 0-2, 3-5, 6-7 of each byte of
// the symbol: six chunks), so each table has at most 8 entries and one
// v_perm_b32 looks up four of them at once: its selector bytes index the 8
// bytes of a register pair.  A table of 16-bit products is two byte tables
// (low and high byte of the product), so a coefficient has 12 byte tables
// of 8 entries, 96 bytes (the reference's SSSE3 / AVX2 kernels split by
// 4-bit nibbles for their 16-entry byte shuffle, src/galois_simd.rs:109-137,
// 432-539; v_perm_b32 indexes 8 bytes, hence 3-bit chunks).
constexpr int kNibChunks = 6;
__host__ __device__ constexpr uint32_t nib_shift(int q) { return q < 3 ? 3 * q : 8 + 3 * (q - 3); }
__host__ __device__ constexpr uint32_t nib_width(int q) { return (q % 3) == 2 ? 2 : 3; }

// tables[(i * R + r) * 24 + 4 q + 2 o + h]: byte table o (0: low byte of the
// product, 1: high byte) of chunk q of coefficient coef(r, first_input + i),
// entries 4 h .. 4 h + 3 (h = 0, 1), entry v in byte v % 4.  Thread per
// (input, row).
__global__ __launch_bounds__(256) void k_nibble_tables(const uint16_t *__restrict__ antilog,
                                                       const uint16_t *__restrict__ logbase,
                                                       uint32_t n, uint32_t R,
                                                       uint32_t *__restrict__ tables) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (uint64_t)n * R) return;
    const uint32_t i = (uint32_t)(idx / R), r = (uint32_t)(idx % R);
    uint32_t v = antilog[(uint32_t)(((uint64_t)logbase[i] * r) % kFieldLimit)];
    uint32_t pw[16];  // coef * x^m
#pragma unroll
    for (int m = 0; m < 16; m++) {
        pw[m] = v;
        v = ((v << 1) & 0xFFFFu) ^ ((v & 0x8000u) ? (kFieldGenerator & 0xFFFFu) : 0u);
    }
    uint32_t *out = tables + idx * 24;
#pragma unroll
    for (int q = 0; q < kNibChunks; q++) {
        uint32_t t[2][2] = {{0, 0}, {0, 0}};  // [byte o][half h]
#pragma unroll
        for (int e = 0; e < 8; e++) {
            uint32_t prod = 0;
#pragma unroll
            for (int j = 0; j < 3; j++)
                if (j < (int)nib_width(q) && ((e >> j) & 1)) prod ^= pw[nib_shift(q) + j];
            // entries past the chunk's width are never selected (selectors
            // are masked to the width); they stay the products of their bits
            t[0][e >> 2] |= (prod & 0xFFu) << (8 * (e & 3));
            t[1][e >> 2] |= (prod >> 8) << (8 * (e & 3));
        }
        out[4 * q + 0] = t[0][0];
        out[4 * q + 1] = t[0][1];
        out[4 * q + 2] = t[1][0];
        out[4 * q + 3] = t[1][1];
    }
}

hipError_t launch_nibble_tables(const uint16_t *d_antilog, const uint16_t *d_logbase,
                                uint32_t n_inputs, uint32_t R, uint32_t *d_tables,
                                hipStream_t stream) {
    const uint64_t total = (uint64_t)n_inputs * R;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_nibble_tables, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, stream,
                       d_antilog, d_logbase, n_inputs, R, d_tables);
    return hipGetLastError();
}

// XOR of three words in one v_bitop3_b32 (truth table 0x96).
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

constexpr int kNibThreads = 256;
constexpr int kNibRT = 8;   // recovery rows per workgroup
constexpr int kNibSym = 8;  // symbols per lane (16 bytes)

// A lane owns 8 consecutive symbols (16 bytes) of the slice column range and
// kNibRT recovery rows, over the inputs of its workgroup's input chunk
// (blockIdx.z: chunks of in_chunk inputs, so that small slices still fill the
// GPU; with more than one chunk the partial blocks are XORed into rec with
// atomics, XOR being the addition of the field).  Per input it gathers the low bytes of its 8 symbols
// into two words and the high bytes into two more (v_perm), splits them into
// the six chunk selector planes once, then per row looks every chunk up in
// the row's byte tables (one v_perm per 4 symbols, per output byte) and
// XORs the results into byte-plane accumulators; the accumulators are
// interleaved back into symbols at the end.  The tables of a (row, input)
// are the same for the whole workgroup, so they are read as scalars.
template <bool VEC>
__global__ __launch_bounds__(kNibThreads) void k_encode_nibble(
    const SliceDesc *__restrict__ inputs, uint32_t n_inputs, const uint32_t *__restrict__ tables,
    uint32_t R, uint32_t S, uint8_t *__restrict__ rec, int accumulate, uint32_t in_chunk) {
    const uint32_t col = (blockIdx.x * kNibThreads + threadIdx.x) * (2 * kNibSym);
    if (col >= S) return;
    const uint32_t i_lo = blockIdx.z * in_chunk, i_hi = min(i_lo + in_chunk, n_inputs);
    const bool split = gridDim.z > 1;
    const uint32_t r0 = blockIdx.y * kNibRT;
    const uint32_t rt = min(R - r0, (uint32_t)kNibRT);
    uint32_t acc[kNibRT][2][2];  // [row][byte o][symbols 4 j .. 4 j + 3]
#pragma unroll
    for (int rr = 0; rr < kNibRT; rr++)
#pragma unroll
        for (int o = 0; o < 2; o++) acc[rr][o][0] = acc[rr][o][1] = 0;

    for (uint32_t i = i_lo; i < i_hi; i++) {
        const SliceDesc d = inputs[i];
        uint32_t x[4];
        if (VEC && col + 16 <= d.valid) {
            const uint4 v = gload(reinterpret_cast<const uint4 *>(d.ptr + col));
            x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
        } else {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t off = col + 4u * q;
                x[q] = (off < S) ? load_word(d.ptr, off, d.valid) : 0u;
            }
        }
        // low bytes of symbols 0-3 / 4-7, high bytes likewise
        uint32_t sel[kNibChunks][2];
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const uint32_t lo = __builtin_amdgcn_perm(x[2 * j + 1], x[2 * j], 0x06040200u);
            const uint32_t hi = __builtin_amdgcn_perm(x[2 * j + 1], x[2 * j], 0x07050301u);
            sel[0][j] = lo & 0x07070707u;
            sel[1][j] = (lo >> 3) & 0x07070707u;
            sel[2][j] = (lo >> 6) & 0x03030303u;
            sel[3][j] = hi & 0x07070707u;
            sel[4][j] = (hi >> 3) & 0x07070707u;
            sel[5][j] = (hi >> 6) & 0x03030303u;
        }
        const uint32_t *tb = tables + ((size_t)i * R + r0) * 24;
#pragma unroll
        for (int rr = 0; rr < kNibRT; rr++) {
            if ((uint32_t)rr >= rt) break;
            const uint32_t *t = tb + rr * 24;
#pragma unroll
            for (int o = 0; o < 2; o++) {
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    uint32_t p[kNibChunks];
#pragma unroll
                    for (int q = 0; q < kNibChunks; q++)
                        p[q] = __builtin_amdgcn_perm(t[4 * q + 2 * o + 1], t[4 * q + 2 * o], sel[q][j]);
                    uint32_t a = xor3(acc[rr][o][j], p[0], p[1]);
                    a = xor3(a, p[2], p[3]);
                    acc[rr][o][j] = xor3(a, p[4], p[5]);
                }
            }
        }
    }

#pragma unroll
    for (int rr = 0; rr < kNibRT; rr++) {
        if ((uint32_t)rr >= rt) break;
        uint32_t y[4];
#pragma unroll
        for (int j = 0; j < 2; j++) {
            // symbol s: low byte acc[.][0][j] byte s, high byte acc[.][1][j] byte s
            y[2 * j] = __builtin_amdgcn_perm(acc[rr][1][j], acc[rr][0][j], 0x05010400u);
            y[2 * j + 1] = __builtin_amdgcn_perm(acc[rr][1][j], acc[rr][0][j], 0x07030602u);
        }
        uint8_t *out = rec + (size_t)(r0 + rr) * S + col;
        if (split) {
            uint32_t *o = reinterpret_cast<uint32_t *>(out);
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (col + 4u * q < S) atomicXor(o + q, y[q]);
        } else if (VEC && col + 16 <= S) {
            uint4 *o4 = reinterpret_cast<uint4 *>(out);
            uint4 v = make_uint4(y[0], y[1], y[2], y[3]);
            if (accumulate) {
                const uint4 old = *o4;
                v.x ^= old.x; v.y ^= old.y; v.z ^= old.z; v.w ^= old.w;
            }
            *o4 = v;
        } else {
            uint32_t *o = reinterpret_cast<uint32_t *>(out);
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (col + 4u * q < S) o[q] = accumulate ? (o[q] ^ y[q]) : y[q];
        }
    }
}

hipError_t launch_encode_nibble(const SliceDesc *d_inputs, uint32_t n_inputs, const uint32_t *d_tables,
                                uint32_t R, uint32_t S, uint8_t *d_rec, int accumulate,
                                hipStream_t stream) {
    if (R == 0) return hipSuccess;
    const uint32_t per_block = kNibThreads * 2 * kNibSym;
    dim3 grid((S + per_block - 1) / per_block, (R + kNibRT - 1) / kNibRT);
    // input chunks of at least 16 inputs until the grid has ~2048 workgroups
    const uint64_t xy = (uint64_t)grid.x * grid.y;
    const uint32_t want = (uint32_t)std::max<uint64_t>(1, (2048 + xy - 1) / xy);
    const uint32_t chunks = std::max<uint32_t>(1, std::min<uint32_t>(want, (n_inputs + 15) / 16));
    const uint32_t in_chunk = (n_inputs + chunks - 1) / chunks;
    grid.z = (n_inputs + in_chunk - 1) / in_chunk;
    if (grid.z > 1 && !accumulate) {
        // the chunks' partials are XORed into a zeroed block
        const hipError_t e = hipMemsetAsync(d_rec, 0, (size_t)R * S, stream);
        if (e != hipSuccess) return e;
    }
    if (S % 16 == 0 && reinterpret_cast<uintptr_t>(d_rec) % 16 == 0)
        hipLaunchKernelGGL(k_encode_nibble<true>, grid, dim3(kNibThreads), 0, stream, d_inputs,
                           n_inputs, d_tables, R, S, d_rec, accumulate, in_chunk);
    else
        hipLaunchKernelGGL(k_encode_nibble<false>, grid, dim3(kNibThreads), 0, stream, d_inputs,
                           n_inputs, d_tables, R, S, d_rec, accumulate, in_chunk);
    return hipGetLastError();
}

// ======================================================================
// K1 (matrix-core form): GF(2) contraction on fp4 MFMA
// ======================================================================
//
// Over GF(2) the whole encode is one bit-matrix product:
//   out[16 r + j][p] = parity( sum_{i,b} A[16 r + j][16 i + b] * in[16 i + b][p] )
// with A[16 r + j][16 i + b] = bit j of coef(r, i) * x^b and in[16 i + b][p]
// = bit b of symbol p of slice i.  Bits become fp4 (e2m1) 1.0 / 0.0, the
// v_mfma_f32_32x32x64_f8f6f4 sums them exactly in fp32 (at most 16 N =
// 524288 < 2^24 terms), and the parity of each sum is the output bit.
//
// One MFMA covers 32 rows (recovery rows 2t, 2t + 1), 64 k (inputs
// 4g .. 4g + 3) and 32 columns (symbols).  A lane in half h (lane / 32)
// holds k values of inputs i0 = 4g + 2h and i1 = 4g + 2h + 1, 16 bits each,
// as four dwords of eight nibbles, the same order in A and in B, so each
// (A lane, nibble) meets the B nibble of the same k.
//
// Nibble n of dword d carries bit 4 (n % 4) + d of input i0 (n < 4) or i1
// (n >= 4).  Any order shared by A and B gives the same sums; this one makes
// the B operand of a symbol one byte permute (the symbol's 16 bits of both
// inputs side by side) and four masks of the result, with the set bit left
// where it lies in its nibble: fp4 0.5 (bit 0), 1.0 (bit 1), 2.0 (bit 2);
// bit 3 (the fp4 sign) is shifted down to bit 2 (2.0).  The A nibble of the
// same k is the reciprocal (2.0, 1.0, 0.5, 0.5 for d = 0..3), so every
// product is exactly 1.0 and the sums stay integers.

typedef int v8i_t __attribute__((ext_vector_type(8)));
typedef float v16f_t __attribute__((ext_vector_type(16)));

// B dwords of one symbol: x = the symbol's 16 bits of input i0 (low half)
// and of input i1 (high half).
__device__ __forceinline__ void mf_b_dwords(uint32_t x, uint32_t out[4]) {
    out[0] = x & 0x11111111u;         // bits 4q     -> 0.5
    out[1] = x & 0x22222222u;         // bits 4q + 1 -> 1.0
    out[2] = x & 0x44444444u;         // bits 4q + 2 -> 2.0
    out[3] = (x >> 1) & 0x44444444u;  // bits 4q + 3 -> 2.0
}

// A dword d for the matrix rows m0 (input i0) and m1 (input i1) of one
// output bit (bit b = coefficient of symbol bit b): nibble n holds bit
// 4 (n % 4) + d of its row as the fp4 reciprocal of the B value of dword d.
__host__ __device__ __forceinline__ uint32_t mf_a_dword(uint32_t m0, uint32_t m1, int d) {
    const int sh = d == 0 ? 2 : d == 1 ? 1 : 0;  // 2.0 (0x4), 1.0 (0x2), 0.5 (0x1)
    const uint32_t lo = ((m0 >> d) & 0x1111u) << sh, hi = ((m1 >> d) & 0x1111u) << sh;
    return lo | (hi << 16);
}

// E8M0 block scale of 1.0 (biased exponent 127) for both operands; a scale
// operand of 0 would mean 2^-127.
constexpr int kE8M0One = 127;


// A fragments: frag (t, g) is 64 lanes x 16 bytes at afrag[(t * nq + g) * 64 + lane].
// Entry idx = (t * nq + g) * 64 + lane.  Row r of the fragments is recovery
// exponent first_exp + r; rows 2t + rr >= R and inputs >= n are zero.
// Row j of the GF(2) matrix of multiplication by x^k, k, j < 16: bit b is
// bit j of x^(k + b) mod the field polynomial.  The matrix of multiplication
// by v is the sum of those of the x^k for the bits k of v, so a row of it is
// the XOR of 16 such rows selected by v's bits.
struct MulXRows {
    uint16_t row[16][16];  // [j][k]
};
constexpr MulXRows make_mulx_rows() {
    MulXRows t{};
    uint32_t pw[31] = {};
    uint32_t v = 1;
    for (int e = 0; e < 31; e++) {
        pw[e] = v;
        v = (v << 1) ^ ((v & 0x8000u) ? kFieldGenerator : 0u);
    }
    for (int j = 0; j < 16; j++)
        for (int k = 0; k < 16; k++) {
            uint32_t m = 0;
            for (int b = 0; b < 16; b++) m |= ((pw[k + b] >> j) & 1u) << b;
            t.row[j][k] = (uint16_t)m;
        }
    return t;
}
__constant__ MulXRows kMulXRows = make_mulx_rows();

// x mod 65535 for x < 2^32 (65536 = 1 mod 65535): two folds of the halves.
__host__ __device__ __forceinline__ uint32_t mod_field_limit(uint32_t x) {
    x = (x & 0xFFFFu) + (x >> 16);
    x = (x & 0xFFFFu) + (x >> 16);
    return x == kFieldLimit ? 0u : x;
}

// The rows of the x^k matrices a fragment lane uses: lane l holds matrix row
// j = l % 16 of its coefficients in every fragment.
struct AfragLane {
    uint32_t rows[16];
    __device__ __forceinline__ void load(uint32_t lane) {
        const uint32_t j = lane & 15;
#pragma unroll
        for (int k = 0; k < 16; k++) rows[k] = kMulXRows.row[j][k];
    }
};

// Lane `lane` of fragment (t, g): matrix row j = lane % 16 of the two
// coefficients of recovery row r = 2 t + lane / 16 % 2 and inputs
// 4 g + 2 h, 4 g + 2 h + 1 (h = lane / 32).
__device__ __forceinline__ uint4 mfma_afrag_entry(const AfragLane &L,
                                                  const uint16_t *__restrict__ antilog,
                                                  const uint16_t *__restrict__ logbase, uint32_t n,
                                                  uint32_t R, uint32_t first_exp, uint32_t t,
                                                  uint32_t g, uint32_t lane) {
    const uint32_t row = lane & 31, h = lane >> 5;
    const uint32_t r = 2 * t + (row >> 4);
    uint32_t m[2] = {0, 0};
#pragma unroll
    for (uint32_t e = 0; e < 2; e++) {
        const uint32_t i = 4 * g + 2 * h + e;
        if (r >= R || i >= n) continue;
        // exponents and logs are below 65535, so the product fits 32 bits
        const uint32_t v = antilog[mod_field_limit((uint32_t)logbase[i] * (first_exp + r))];
        uint32_t me = 0;  // row j of the matrix of multiplication by v
#pragma unroll
        for (int k = 0; k < 16; k++) {
            // all ones when bit k of v is set; then me ^= rows[k] & mask as one
            // v_bitop3 (truth-table bit (me << 2 | row << 1 | mask): 0x78)
            const uint32_t mask = (uint32_t)__builtin_amdgcn_sbfe((int)v, k, 1);
            me = __builtin_amdgcn_bitop3_b32(me, L.rows[k], mask, 0x78);
        }
        m[e] = me;
    }
    return make_uint4(mf_a_dword(m[0], m[1], 0), mf_a_dword(m[0], m[1], 1),
                      mf_a_dword(m[0], m[1], 2), mf_a_dword(m[0], m[1], 3));
}

// Every fragment (t, g) of tiles x nq, entry (t * nq + g) * 64 + lane: a
// thread keeps its lane and walks fragments frag0, frag0 + F, ... (F = the
// grid's waves), stepping (t, g) without a division.
__device__ __forceinline__ void afrag_fill(const AfragLane &L, const uint16_t *__restrict__ antilog,
                                           const uint16_t *__restrict__ logbase, uint32_t n,
                                           uint32_t R, uint32_t first_exp, uint32_t nq,
                                           uint32_t tiles, uint4 *__restrict__ out) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t frag0 = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t F = (gridDim.x * blockDim.x) >> 6;
    const uint32_t dt = F / nq, dg = F % nq;
    uint32_t t = frag0 / nq, g = frag0 % nq;
    while (t < tiles) {
        out[((size_t)t * nq + g) * 64 + lane] = mfma_afrag_entry(L, antilog, logbase, n, R, first_exp, t, g, lane);
        t += dt;
        g += dg;
        if (g >= nq) {
            g -= nq;
            t++;
        }
    }
}

// Grid-stride over the fragments; the block size is a multiple of 64.
__global__ __launch_bounds__(256) void k_mfma_afrag(const uint16_t *__restrict__ antilog,
                                                    const uint16_t *__restrict__ logbase,
                                                    uint32_t n, uint32_t R, uint32_t nq,
                                                    uint32_t tiles, uint4 *__restrict__ afrag) {
    AfragLane L;
    L.load(threadIdx.x & 63);
    afrag_fill(L, antilog, logbase, n, R, 0, nq, tiles, afrag);
}

hipError_t launch_mfma_afrag(const uint16_t *d_antilog, const uint16_t *d_logbase, uint32_t n,
                             uint32_t R, uint4 *d_afrag, hipStream_t stream) {
    const uint32_t nq = mfma_steps(n);
    const uint32_t tiles = mfma_row_tiles(R);
    const uint64_t total = (uint64_t)tiles * nq * 64;
    if (total == 0) return hipSuccess;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((total + 255) / 256, 16384);
    hipLaunchKernelGGL(k_mfma_afrag, dim3(blocks), dim3(256), 0, stream, d_antilog, d_logbase, n,
                       R, nq, tiles, d_afrag);
    return hipGetLastError();
}

// LDS: input descriptors of the launch (at most kMfMaxInputs) and a double
// buffer of the workgroup's A fragments for one K step (8 tiles x 1 KiB).
constexpr int kMfAhead = 4;  // K steps between issuing a load and consuming it

// Branch-free word loads of one K step: whole words come from the slice,
// the partial last word from the descriptor's tail field in global memory,
// anything else reads kZeroWord.  Only the address is selected, so the
// loaded word is final and nothing waits for it before it is consumed.
// s_desc has 4 nq entries, those past the inputs with valid = 0, and
// valid <= S, so inputs past n and columns past S read zero without tests
// of their own (g_desc is dereferenced only for a partial word, i < n).
// The descriptor is read as one 16-byte LDS load {ptr lo, ptr hi, valid,
// tail} and all three addresses are formed unconditionally, so the select
// stays two conditional moves per address instead of a branch.
//
// FULL: every descriptor is a whole slice (valid = S; partial slices were
// staged into zero-padded copies), entries past the inputs repeat input 0
// (their A fragments are zero) and `off` is clamped into the slice, so the
// word is just ptr + off.
template <bool FULL>
__device__ __forceinline__ void mf_load_words(const uint4 *s_desc, const SliceDesc *__restrict__ g_desc,
                                              uint32_t g, uint32_t h, uint32_t off, uint32_t &w0,
                                              uint32_t &w1) {
    static_assert(sizeof(SliceDesc) == 16 && offsetof(SliceDesc, valid) == 8 &&
                      offsetof(SliceDesc, tail) == 12,
                  "descriptor layout read as uint4");
    if (FULL) {
        const uint2 p0 = reinterpret_cast<const uint2 *>(s_desc + 4 * g + 2 * h)[0];
        const uint2 p1 = reinterpret_cast<const uint2 *>(s_desc + 4 * g + 2 * h + 1)[0];
        w0 = gload(reinterpret_cast<const uint32_t *>(((uint64_t)p0.y << 32 | p0.x) + off));
        w1 = gload(reinterpret_cast<const uint32_t *>(((uint64_t)p1.y << 32 | p1.x) + off));
        return;
    }
    const uint64_t zero = reinterpret_cast<uint64_t>(&kZeroWord);
    uint32_t wd[2];
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const uint32_t i = 4 * g + 2 * h + e;
        uint4 d = s_desc[i];
        // Opaque to the optimizer: keeps the 16-byte read whole and the
        // address select a select (it would otherwise sink the pointer read
        // into a branch on `valid`, serializing two LDS round trips).
        asm("" : "+v"(d.x), "+v"(d.y), "+v"(d.z));
        const uint64_t data = ((uint64_t)d.y << 32 | d.x) + off;
        const uint64_t tail = reinterpret_cast<uint64_t>(&g_desc[i].tail);
        // full word: data; partial word: tail; past the end: zero.  As masks
        // (full implies partial) so that no branch is formed.
        uint32_t mf = 0u - (uint32_t)(off + 4 <= d.z), mp = 0u - (uint32_t)(off < d.z);
        asm("" : "+v"(mf), "+v"(mp));
        const uint64_t f64 = (uint64_t)mf << 32 | mf, p64 = (uint64_t)mp << 32 | mp;
        const uint64_t a = (data & f64) | (tail & p64 & ~f64) | (zero & ~p64);
        wd[e] = gload(reinterpret_cast<const uint32_t *>(a));
    }
    w0 = wd[0];
    w1 = wd[1];
}

// B operands of the two column tiles (cb = symbol 0 / 1 of each word) from
// the words of inputs 4g + 2h (w0) and 4g + 2h + 1 (w1): 12 VALU.
__device__ __forceinline__ void mf_make_b(uint32_t w0, uint32_t w1, v8i_t B[2]) {
    uint32_t p[4], q[4];
    mf_b_dwords(__builtin_amdgcn_perm(w1, w0, 0x05040100u), p);  // symbol 0 of both
    mf_b_dwords(__builtin_amdgcn_perm(w1, w0, 0x07060302u), q);  // symbol 1 of both
    B[0] = v8i_t{(int)p[0], (int)p[1], (int)p[2], (int)p[3], 0, 0, 0, 0};
    B[1] = v8i_t{(int)q[0], (int)q[1], (int)q[2], (int)q[3], 0, 0, 0, 0};
}

// LDS of the encode: the padded descriptor table and the A double buffer.
struct MfLds {
    uint4 desc[kMfMaxInputs];     // SliceDesc as {ptr lo, ptr hi, valid, tail}
    uint4 a[2][kMfTiles * 64];
};

// Fills lds.desc with the n descriptors and 4 nq - n more: empty ones
// (valid = 0), or, FULL, copies of input 0 (read, never counted).  The caller
// synchronizes before the first unit.
template <bool FULL>
__device__ __forceinline__ void mf_load_desc(uint4 *desc, const SliceDesc *__restrict__ inputs,
                                             uint32_t n, uint32_t nq) {
    for (uint32_t i = threadIdx.x; i < 4 * nq; i += kMfThreads)
        desc[i] = i < n  ? gload(reinterpret_cast<const uint4 *>(inputs) + i)
                      : FULL ? gload(reinterpret_cast<const uint4 *>(inputs))
                             : make_uint4(0, 0, 0, 0);
}

// One encode unit: row tiles tile0 .. tile0 + 7 (recovery rows 2 tile0 ..
// 2 tile0 + 15) over the 256 symbols at byte colbase of every slice, nq K
// steps (the descriptors in lds.desc), stored to rec (XORed in when
// accumulate).  Fragment (t, g) is afrag[(t * frag_stride + g) * 64 + lane]:
// frag_stride = nq for a whole launch, the job's step count when this is
// one K chunk of it.  All 256 threads call it; it leaves lds.a free for the
// next unit.
template <bool FULL>
__device__ __forceinline__ void mf_unit(MfLds &lds, const SliceDesc *__restrict__ inputs,
                                        const uint4 *__restrict__ afrag, uint32_t nq,
                                        uint32_t frag_stride, uint32_t R, uint32_t S,
                                        uint8_t *__restrict__ rec, int accumulate, uint32_t tile0,
                                        uint32_t colbase) {
    const uint4 *s_desc = lds.desc;
    uint4 (*s_a)[kMfTiles * 64] = lds.a;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63, w = tid >> 6;
    const uint32_t c = lane & 31, h = lane >> 5;
    // this lane's dword: symbols base + 2c (column tile 0) and base + 2c + 1 (tile 1)
    const uint32_t off = colbase + w * 128 + 4 * c;
    const bool col_ok = off < S;
    // FULL loads stay inside the slice; columns past S are computed from any
    // bytes and never stored.
    const uint32_t off_ld = FULL ? min(off, S - 4) : off;

    // Thread tid stages A entries tid and tid + 256 of each K step: tile
    // (e / 64), lane (e % 64) of fragment (tile0 + tile, g).
    const uint4 *ga0 = afrag + ((size_t)(tile0 + (tid >> 6)) * frag_stride) * 64 + lane;
    const uint4 *ga1 = afrag + ((size_t)(tile0 + 4 + (tid >> 6)) * frag_stride) * 64 + lane;

    v16f_t acc[kMfTiles][2];
#pragma unroll
    for (int a = 0; a < kMfTiles; a++)
#pragma unroll
        for (int cb = 0; cb < 2; cb++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[a][cb][v] = 0.0f;

    // Software pipeline, kMfAhead steps deep.  Ring slot p holds the staged
    // A entries and source words of a step gg with gg % kMfAhead == p.  Step
    // gg: issue the loads of step gg + kMfAhead into its own (consumed) slot,
    // run the MFMAs of step gg from LDS buffer gg & 1 and B(gg) while B(gg + 1)
    // is built, store A(gg + 1) into the other LDS buffer, barrier.  The slot
    // index of each step is a literal (MF_STEP), so the ring stays in registers.
    // slot p: staged A entries a0_p, a1_p and source words w0_p, w1_p
    uint4 a0_0, a1_0, a0_1, a1_1, a0_2, a1_2, a0_3, a1_3;
    uint32_t w0_0, w1_0, w0_1, w1_1, w0_2, w1_2, w0_3, w1_3;
#define MF_FILL(P)                                                                            \
    do {                                                                                      \
        const uint32_t gp = min((uint32_t)(P), nq - 1);                                       \
        a0_##P = gload(ga0 + (size_t)gp * 64);                                                \
        a1_##P = gload(ga1 + (size_t)gp * 64);                                                \
        mf_load_words<FULL>(s_desc, inputs, gp, h, off_ld, w0_##P, w1_##P);                         \
    } while (0)
    MF_FILL(0);
    MF_FILL(1);
    MF_FILL(2);
    MF_FILL(3);
#undef MF_FILL
    s_a[0][tid] = a0_0;
    s_a[0][tid + 256] = a1_0;
    v8i_t Bc[2];
    mf_make_b(w0_0, w1_0, Bc);
    __syncthreads();

#define MF_STEP(GG, P, PN)                                                                    \
    do {                                                                                      \
        const uint32_t gf = min((GG) + kMfAhead, nq - 1);                                     \
        a0_##P = gload(ga0 + (size_t)gf * 64);                                                \
        a1_##P = gload(ga1 + (size_t)gf * 64);                                                \
        mf_load_words<FULL>(s_desc, inputs, gf, h, off_ld, w0_##P, w1_##P);                         \
        const uint4 *sa = s_a[(GG) & 1];                                                      \
        v8i_t Bn[2];                                                                          \
        mf_make_b(w0_##PN, w1_##PN, Bn);                                                      \
        uint4 Af[kMfTiles];                                                                   \
        _Pragma("unroll") for (int a = 0; a < kMfTiles; a++) Af[a] = sa[a * 64 + lane];       \
        _Pragma("unroll") for (int a = 0; a < kMfTiles; a++) {                                \
            const uint4 A = Af[a];                                                            \
            const v8i_t av = v8i_t{(int)A.x, (int)A.y, (int)A.z, (int)A.w, 0, 0, 0, 0};       \
            _Pragma("unroll") for (int cb = 0; cb < 2; cb++) acc[a][cb] =                     \
                __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, Bc[cb], acc[a][cb], 4, 4, \
                                                                0, kE8M0One, 0, kE8M0One);    \
        }                                                                                     \
        _Pragma("unroll") for (int k = 0; k < 2 * kMfTiles; k++) {                            \
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                                \
            __builtin_amdgcn_sched_group_barrier(0x002, 6, 0);                                \
        }                                                                                     \
        uint4 *da = s_a[((GG) + 1) & 1];                                                      \
        da[tid] = a0_##PN;                                                                    \
        da[tid + 256] = a1_##PN;                                                              \
        __syncthreads();                                                                      \
        Bc[0] = Bn[0];                                                                        \
        Bc[1] = Bn[1];                                                                        \
    } while (0)

    // nq is a multiple of 4 (mfma_steps), so every iteration runs all four
    // steps and the wait counts are the same on every trip.
    static_assert(kMfAhead == 4, "the step sequence below is written out for 4 slots");
    for (uint32_t g = 0; g < nq; g += kMfAhead) {
        MF_STEP(g, 0, 1);
        MF_STEP(g + 1, 1, 2);
        MF_STEP(g + 2, 2, 3);
        MF_STEP(g + 3, 3, 0);
    }
#undef MF_STEP

    // Epilogue.  Accumulator v of a lane holds row 8 (v / 4) + 4 h + v % 4 of
    // the tile at column c: rows 0..15 are bits of recovery row 2t, rows
    // 16..31 bits of 2t + 1.  Each lane half owns 8 of a symbol's 16 bits;
    // one exchange across halves gives half h the whole symbol of row 2t + h.
#pragma unroll
    for (int a = 0; a < kMfTiles; a++) {
        uint32_t sym[2];
#pragma unroll
        for (int cb = 0; cb < 2; cb++) {
            uint32_t lo = 0, hi = 0;  // bits of rows 2t and 2t + 1 held here
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const uint32_t bit = ((uint32_t)acc[a][cb][v]) & 1u;
                const int pos = (v & 3) + ((v >> 2) & 1) * 8;
                if (v < 8) lo |= bit << pos;
                else hi |= bit << pos;
            }
            lo <<= 4 * h;
            hi <<= 4 * h;
            const uint32_t give = h ? lo : hi;
            const uint32_t got = (uint32_t)__shfl_xor((int)give, 32, 64);
            sym[cb] = (h ? hi : lo) | got;
        }
        const uint32_t r = 2 * (tile0 + a) + h;
        if (r < R && col_ok) {
            uint32_t *o = reinterpret_cast<uint32_t *>(rec + (size_t)r * S + off);
            const uint32_t v = sym[0] | (sym[1] << 16);
            *o = accumulate ? (*o ^ v) : v;
        }
    }
}

// The same unit for whole-slice descriptors with every operand loaded
// straight into LDS (global_load_lds): the A fragments (dwordx4) and the
// source words (dword) of step g land in ring buffer g % 4 three steps
// ahead, take no registers and cannot be moved late by the scheduler.  A
// ring buffer is kMfRing uint4: the 512 A entries, then 8 runs of 64 words
// (wave w, word j at run 2 w + j, lane l's word at position l).  The four
// buffers are distinct __shared__ objects, so the compiler does not make
// LDS reads wait for the DMA into the other buffers; the waits are explicit
// instead.  Every thread issues the same loads in the same order on every
// step (clamped past nq; a compiler fence between steps' groups): per step
// A(g + 3) x 2 then W(g + 3) x 2, after a prologue A0 W0 A1 W1 A2 W2.  At
// the top of step g, group g + 1 (A and W of step g + 1) is complete once
// at most the 4 loads of step g - 1 are in flight: vmcnt(4) (with
// lgkmcnt(0): this wave's LDS reads of step g - 1 are done), then one
// s_barrier (no fence: its release would wait for every load in flight)
// publishes every wave's share.  Step g then reads A(g + 1) and W(g + 1)
// from LDS while its MFMAs run on A(g) and B(g), read and built in step
// g - 1, so no MFMA waits for LDS.  The buffer that step g's DMA overwrites,
// (g + 3) % 4, was last read in step g - 2.  Scheduling barriers keep each
// step's instructions between its barriers (MFMAs have no memory operands,
// so the compiler fences alone would let them sink past later barriers).
// The dense schedule runs the same ring one step deeper: step g's DMA is for
// step g + 4 into buffer g % 4 (read in step g - 1, before the barrier that
// opens step g), and the wait at the top of step g is vmcnt(8), so a load has
// three steps instead of two to land (the batched mix 0.409-0.412 ->
// 0.405 s per replay; A fragments there often come from HBM).
constexpr int kMfRing = kMfTiles * 64 + 128;  // uint4 per ring buffer: A, then 512 words
constexpr int kWaitVmcnt4 = 0x0074;  // s_waitcnt vmcnt(4) lgkmcnt(0) (this wave's LDS reads done)
constexpr int kWaitVmcnt8 = 0x0F78;
constexpr int kWaitVmcnt12 = 0x0F7C;
constexpr int kWaitVmcnt8Lds = 0x0078;  // vmcnt(8) lgkmcnt(0)
constexpr int kWaitVmcnt0 = 0x0070;  // vmcnt(0) lgkmcnt(0)
constexpr int kWaitLgkm0 = 0xC07F;   // lgkmcnt(0) alone (vmcnt 63, expcnt 7)

// Addresses of the two source words of K step g of this lane (whole slices).
__device__ __forceinline__ void mf_word_addrs(const uint4 *s_desc, uint32_t g, uint32_t h,
                                              uint32_t off, const uint32_t *&p0,
                                              const uint32_t *&p1) {
    const uint2 d0 = reinterpret_cast<const uint2 *>(s_desc + 4 * g + 2 * h)[0];
    const uint2 d1 = reinterpret_cast<const uint2 *>(s_desc + 4 * g + 2 * h + 1)[0];
    p0 = reinterpret_cast<const uint32_t *>(((uint64_t)d0.y << 32 | d0.x) + off);
    p1 = reinterpret_cast<const uint32_t *>(((uint64_t)d1.y << 32 | d1.x) + off);
}

// DENSE selects the step schedule.  The dense one keeps the matrix pipe
// busiest, and the GPU then runs at a lower clock under its power limit
// (2.12-2.17 against 2.27-2.32 GHz measured in the fused pass), which
// lengthens every serial MD5 chain by the same factor.  The batched kernel,
// bound by its encode, takes the dense schedule; the fused pass, bound by
// its packet-MD5 chain once the encode is done, takes the light one.
template <bool DENSE>
__device__ __forceinline__ void mf_unit_dma(const uint4 *s_desc, uint4 *A0, uint4 *A1, uint4 *A2,
                                            uint4 *A3, const SliceDesc *__restrict__ inputs,
                                            const uint4 *__restrict__ afrag, uint32_t nq,
                                            uint32_t frag_stride, uint32_t R, uint32_t S,
                                            uint8_t *__restrict__ rec, int accumulate,
                                            uint32_t tile0, uint32_t colbase) {
    (void)inputs;
    const uint32_t lane = threadIdx.x & 63;
    // the wave index as a scalar, so the DMA's LDS destinations (M0) are
    // scalar arithmetic
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t c = lane & 31, h = lane >> 5;
    // DMA share of wave w: column quarter w (128 bytes) of the step's four
    // inputs and the A fragments of tiles w and w + 4
    const uint32_t off_ld = min(colbase + w * 128 + 4 * c, S - 4);
    // wave-uniform fragment bases (scalar registers); the lane offset is
    // added at each load, so the DMA takes a scalar base and a lane offset
    const uint4 *ga0 = afrag + ((size_t)(tile0 + w) * frag_stride) * 64;
    const uint4 *ga1 = afrag + ((size_t)(tile0 + 4 + w) * frag_stride) * 64;
    // Compute share of wave w = 2 wr + wc: row tiles 4 wr .. 4 wr + 3 over
    // column quarters 2 wc and 2 wc + 1.  Column tile 2 p + s is symbol s of
    // the lane's word in quarter 2 wc + p; the word of input 4g + 2h + j
    // there is run 2 (2 wc + p) + j, position lane, of the ring buffer.
    const uint32_t wr = w >> 1, wc = w & 1;
    const uint32_t wrd = kMfTiles * 64 * 4 + (4 * wc) * 64 + lane;
    const uint32_t ard = (4 * wr) * 64 + lane;

    v16f_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int ct = 0; ct < 4; ct++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[a][ct][v] = 0.0f;

    // word addresses of the next DMA, read from the descriptors one step ahead
    const uint32_t *pw0, *pw1;
#define MF_DMA(G, BUF)                                                                        \
    do {                                                                                      \
        const uint32_t gq = min((uint32_t)(G), nq - 1);                                       \
        __builtin_amdgcn_global_load_lds(ga0 + (size_t)gq * 64 + lane, (BUF) + 64 * w, 16, 0, 0); \
        __builtin_amdgcn_global_load_lds(ga1 + (size_t)gq * 64 + lane, (BUF) + 256 + 64 * w, 16, 0, 0); \
        __builtin_amdgcn_global_load_lds(pw0, reinterpret_cast<uint32_t *>(BUF) + kMfTiles * 64 * 4 + 2 * w * 64, 4, 0, 0); \
        __builtin_amdgcn_global_load_lds(pw1, reinterpret_cast<uint32_t *>(BUF) + kMfTiles * 64 * 4 + (2 * w + 1) * 64, 4, 0, 0); \
        asm volatile("" ::: "memory");                                                        \
        mf_word_addrs(s_desc, min((uint32_t)(G) + 1, nq - 1), h, off_ld, pw0, pw1);           \
    } while (0)
    // B operands of the four column tiles from the words in BUF: the words
    // are read first, so the B build waits only for them
#define MF_B(BUF, OUT)                                                                        \
    do {                                                                                      \
        const uint32_t *wb = reinterpret_cast<const uint32_t *>(BUF) + wrd;                   \
        const uint32_t x00 = wb[0], x01 = wb[64], x10 = wb[128], x11 = wb[192];               \
        mf_make_b(x00, x01, OUT);                                                             \
        mf_make_b(x10, x11, OUT + 2);                                                         \
    } while (0)
    mf_word_addrs(s_desc, 0, h, off_ld, pw0, pw1);
    MF_DMA(0, A0);
    MF_DMA(1, A1);
    MF_DMA(2, A2);
    if constexpr (DENSE) MF_DMA(3, A3);
    v8i_t Bc[4];
    uint4 Ac[4];
    // group 0 of this wave
    if constexpr (DENSE) __builtin_amdgcn_s_waitcnt(kWaitVmcnt12);
    else __builtin_amdgcn_s_waitcnt(kWaitVmcnt8);
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    MF_B(A0, Bc);
#pragma unroll
    for (int a = 0; a < 4; a++) Ac[a] = A0[ard + a * 64];

    // step GG (both schedules): MFMAs on Ac = A(GG), Bc = B(GG); DMA(GG + 3)
    // (light) or DMA(GG + 4) (dense) into BUFW; B(GG + 1) and A(GG + 1) from
    // BUFN.
    //
    // Light schedule: scheduling-group hints only, which leave the DMA issue
    // and most of the B build in runs between few MFMAs.
#define MF_STEP_LIGHT(GG, BUFN, BUFW)                                                           \
    do {                                                                                      \
        __builtin_amdgcn_sched_barrier(0);                                                    \
        /* the step's first MFMA ahead of its wait and barrier: the pipe */                  \
        /* runs it while the wave waits for the others */                                    \
        {                                                                                     \
            const uint4 A = Ac[0];                                                            \
            const v8i_t av = v8i_t{(int)A.x, (int)A.y, (int)A.z, (int)A.w, 0, 0, 0, 0};       \
            acc[0][0] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, Bc[0], acc[0][0], \
                                                   4, 4, 0, kE8M0One, 0, kE8M0One);           \
        }                                                                                     \
        __builtin_amdgcn_sched_barrier(0);                                                    \
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt4);                                              \
        asm volatile("" ::: "memory");                                                        \
        __builtin_amdgcn_s_barrier();                                                         \
        asm volatile("" ::: "memory");                                                        \
        __builtin_amdgcn_sched_barrier(0);                                                    \
        MF_DMA((GG) + 3, BUFW);                                                               \
        v8i_t Bn[4];                                                                          \
        MF_B(BUFN, Bn);                                                                       \
        uint4 An[4];                                                                          \
        _Pragma("unroll") for (int a = 0; a < 4; a++) An[a] = (BUFN)[ard + a * 64];           \
        _Pragma("unroll") for (int a = 0; a < 4; a++) {                                       \
            const uint4 A = Ac[a];                                                            \
            const v8i_t av = v8i_t{(int)A.x, (int)A.y, (int)A.z, (int)A.w, 0, 0, 0, 0};       \
            _Pragma("unroll") for (int ct = 0; ct < 4; ct++) if (a > 0 || ct > 0) acc[a][ct] = \
                __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, Bc[ct], acc[a][ct], 4, 4, \
                                                                0, kE8M0One, 0, kE8M0One);    \
        }                                                                                     \
        /* Three MFMAs straight after the barrier (operands in registers) */                 \
        /* while the DMA is issued, then the LDS reads two per MFMA, then */                 \
        /* the B build's VALU three per MFMA (an MFMA issue holds the wave */                \
        /* until the pipe takes it, so only work placed between MFMAs */                     \
        /* overlaps them). */                                                                 \
        __builtin_amdgcn_sched_group_barrier(0x008, 3, 0);                                    \
        __builtin_amdgcn_sched_group_barrier(0x010, 4, 0);                                    \
        _Pragma("unroll") for (int k = 0; k < 3; k++) {                                       \
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                                \
            __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);                                \
        }                                                                                     \
        _Pragma("unroll") for (int k = 0; k < 9; k++) {                                       \
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                                \
            __builtin_amdgcn_sched_group_barrier(0x002, 3, 0);                                \
        }                                                                                     \
        _Pragma("unroll") for (int ct = 0; ct < 4; ct++) Bc[ct] = Bn[ct];                    \
        _Pragma("unroll") for (int a = 0; a < 4; a++) Ac[a] = An[a];                          \
    } while (0)

    // Dense schedule: the step is written out in the order it issues, each
    // group of other instructions fenced between two
    // MFMAs (an MFMA issue holds the wave until the pipe takes it, so only
    // work placed between MFMAs overlaps them): the first MFMA ahead of the
    // wait and barrier, then one DMA per MFMA, the LDS reads (words first,
    // so the B build waits only for them), and the B build three VALU per
    // MFMA.
#define MF_MM(A_, CT_)                                                                        \
    acc[A_][CT_] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(                           \
        v8i_t{(int)Ac[A_].x, (int)Ac[A_].y, (int)Ac[A_].z, (int)Ac[A_].w, 0, 0, 0, 0}, Bc[CT_], \
        acc[A_][CT_], 4, 4, 0, kE8M0One, 0, kE8M0One)
#define MF_SB __builtin_amdgcn_sched_barrier(0)
    // half of the B operand of one symbol: x from a byte permute, then two masks
#define MF_BH0(X, W0, W1, SEL, OUT)                                                           \
    do {                                                                                      \
        X = __builtin_amdgcn_perm(W1, W0, SEL);                                               \
        OUT[0] = (int)(X & 0x11111111u);                                                      \
        OUT[1] = (int)(X & 0x22222222u);                                                      \
    } while (0)
#define MF_BH1(X, OUT)                                                                        \
    do {                                                                                      \
        OUT[2] = (int)(X & 0x44444444u);                                                      \
        OUT[3] = (int)((X >> 1) & 0x44444444u);                                               \
        OUT[4] = OUT[5] = OUT[6] = OUT[7] = 0;                                                \
    } while (0)
#define MF_STEP_DENSE(GG, BUFN, BUFW)                                                         \
    do {                                                                                      \
        const uint32_t gq = min((uint32_t)(GG) + 4, nq - 1);                                  \
        const uint32_t *wb = reinterpret_cast<const uint32_t *>(BUFN) + wrd;                  \
        uint32_t x00, x01, x10, x11, xs;                                                      \
        uint2 d0, d1;                                                                         \
        v8i_t Bn[4];                                                                          \
        uint4 An[4];                                                                          \
        MF_SB;                                                                                \
        MF_MM(0, 0);                                                                          \
        MF_SB;                                                                                \
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt8Lds);                                           \
        asm volatile("" ::: "memory");                                                        \
        __builtin_amdgcn_s_barrier();                                                         \
        asm volatile("" ::: "memory");                                                        \
        MF_SB;                                                                                \
        MF_MM(0, 1);                                                                          \
        __builtin_amdgcn_global_load_lds(ga0 + (size_t)gq * 64 + lane, (BUFW) + 64 * w, 16, 0, 0); \
        x00 = wb[0];                                                                          \
        x01 = wb[64];                                                                         \
        MF_SB;                                                                                \
        MF_MM(0, 2);                                                                          \
        __builtin_amdgcn_global_load_lds(ga1 + (size_t)gq * 64 + lane, (BUFW) + 256 + 64 * w, 16, 0, 0); \
        x10 = wb[128];                                                                        \
        x11 = wb[192];                                                                        \
        {                                                                                     \
            const uint32_t gd = min((uint32_t)(GG) + 5, nq - 1);                              \
            d0 = reinterpret_cast<const uint2 *>(s_desc + 4 * gd + 2 * h)[0];                 \
            d1 = reinterpret_cast<const uint2 *>(s_desc + 4 * gd + 2 * h + 1)[0];             \
        }                                                                                     \
        MF_SB;                                                                                \
        MF_MM(0, 3);                                                                          \
        __builtin_amdgcn_global_load_lds(pw0, reinterpret_cast<uint32_t *>(BUFW) + kMfTiles * 64 * 4 + 2 * w * 64, 4, 0, 0); \
        MF_SB;                                                                                \
        MF_MM(1, 0);                                                                          \
        __builtin_amdgcn_global_load_lds(pw1, reinterpret_cast<uint32_t *>(BUFW) + kMfTiles * 64 * 4 + (2 * w + 1) * 64, 4, 0, 0); \
        asm volatile("" ::: "memory");                                                        \
        MF_SB;                                                                                \
        MF_MM(1, 1);                                                                          \
        MF_BH0(xs, x00, x01, 0x05040100u, Bn[0]);                                             \
        MF_SB;                                                                                \
        MF_MM(1, 2);                                                                          \
        MF_BH1(xs, Bn[0]);                                                                    \
        An[0] = (BUFN)[ard];                                                                  \
        MF_SB;                                                                                \
        MF_MM(1, 3);                                                                          \
        MF_BH0(xs, x00, x01, 0x07060302u, Bn[1]);                                             \
        An[1] = (BUFN)[ard + 64];                                                             \
        MF_SB;                                                                                \
        MF_MM(2, 0);                                                                          \
        MF_BH1(xs, Bn[1]);                                                                    \
        An[2] = (BUFN)[ard + 128];                                                            \
        MF_SB;                                                                                \
        MF_MM(2, 1);                                                                          \
        MF_BH0(xs, x10, x11, 0x05040100u, Bn[2]);                                             \
        An[3] = (BUFN)[ard + 192];                                                            \
        MF_SB;                                                                                \
        MF_MM(2, 2);                                                                          \
        MF_BH1(xs, Bn[2]);                                                                    \
        MF_SB;                                                                                \
        MF_MM(2, 3);                                                                          \
        MF_BH0(xs, x10, x11, 0x07060302u, Bn[3]);                                             \
        MF_SB;                                                                                \
        MF_MM(3, 0);                                                                          \
        MF_BH1(xs, Bn[3]);                                                                    \
        MF_SB;                                                                                \
        MF_MM(3, 1);                                                                          \
        pw0 = reinterpret_cast<const uint32_t *>(((uint64_t)d0.y << 32 | d0.x) + off_ld);     \
        pw1 = reinterpret_cast<const uint32_t *>(((uint64_t)d1.y << 32 | d1.x) + off_ld);     \
        MF_SB;                                                                                \
        MF_MM(3, 2);                                                                          \
        MF_SB;                                                                                \
        MF_MM(3, 3);                                                                          \
        MF_SB;                                                                                \
        _Pragma("unroll") for (int ct = 0; ct < 4; ct++) Bc[ct] = Bn[ct];                    \
        _Pragma("unroll") for (int a = 0; a < 4; a++) Ac[a] = An[a];                          \
    } while (0)

    if constexpr (DENSE) {
        for (uint32_t g = 0; g < nq; g += 4) {
            MF_STEP_DENSE(g, A1, A0);
            MF_STEP_DENSE(g + 1, A2, A1);
            MF_STEP_DENSE(g + 2, A3, A2);
            MF_STEP_DENSE(g + 3, A0, A3);
        }
    } else {
        for (uint32_t g = 0; g < nq; g += 4) {
            MF_STEP_LIGHT(g, A1, A3);
            MF_STEP_LIGHT(g + 1, A2, A0);
            MF_STEP_LIGHT(g + 2, A3, A1);
            MF_STEP_LIGHT(g + 3, A0, A2);
        }
    }
#undef MF_STEP_DENSE
#undef MF_STEP_LIGHT
#undef MF_BH0
#undef MF_BH1
#undef MF_SB
#undef MF_MM
#undef MF_B
#undef MF_DMA
    // the clamped DMAs of steps past nq, and this wave's last LDS reads, are
    // done before anything reuses the ring
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");

    // Epilogue, as in mf_unit: half h of the lanes ends with the symbols of
    // recovery row 2 t + h of row tile t.
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const uint32_t r = 2 * (tile0 + 4 * wr + a) + h;
#pragma unroll
        for (int p = 0; p < 2; p++) {
            uint32_t sym[2];
#pragma unroll
            for (int s = 0; s < 2; s++) {
                uint32_t lo = 0, hi = 0;
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const uint32_t bit = ((uint32_t)acc[a][2 * p + s][v]) & 1u;
                    const int pos = (v & 3) + ((v >> 2) & 1) * 8;
                    if (v < 8) lo |= bit << pos;
                    else hi |= bit << pos;
                }
                lo <<= 4 * h;
                hi <<= 4 * h;
                const uint32_t give = h ? lo : hi;
                const uint32_t got = (uint32_t)__shfl_xor((int)give, 32, 64);
                sym[s] = (h ? hi : lo) | got;
            }
            const uint32_t off = colbase + wc * 256 + p * 128 + 4 * c;
            if (r < R && off < S) {
                uint32_t *o = reinterpret_cast<uint32_t *>(rec + (size_t)r * S + off);
                const uint32_t v = sym[0] | (sym[1] << 16);
                *o = accumulate ? (*o ^ v) : v;
            }
        }
    }
}

__global__ __launch_bounds__(kMfThreads, 1) void k_encode_mfma(
    const SliceDesc *__restrict__ inputs, uint32_t n, const uint4 *__restrict__ afrag,
    uint32_t nq, uint32_t R, uint32_t S, uint8_t *__restrict__ rec, int accumulate) {
    __shared__ MfLds lds;
    mf_load_desc<false>(lds.desc, inputs, n, nq);
    __syncthreads();
    mf_unit<false>(lds, inputs, afrag, nq, nq, R, S, rec, accumulate, blockIdx.y * kMfTiles,
                   blockIdx.x * (kMfThreads * 2));
}

hipError_t launch_encode_mfma(const SliceDesc *d_inputs, uint32_t n_inputs, const uint4 *d_afrag,
                              uint32_t R, uint32_t S, uint8_t *d_rec, int accumulate,
                              hipStream_t stream) {
    if (R == 0) return hipSuccess;
    const uint32_t nq = mfma_steps(n_inputs);
    if (4 * nq > (uint32_t)kMfMaxInputs) return hipErrorInvalidValue;  // caller chunks
    const uint32_t per_block = kMfThreads * 2;  // bytes of symbol columns per workgroup
    dim3 grid((S + per_block - 1) / per_block, mfma_row_tiles(R) / kMfTiles);
    hipLaunchKernelGGL(k_encode_mfma, grid, dim3(kMfThreads), 0, stream, d_inputs, n_inputs,
                       d_afrag, nq, R, S, d_rec, accumulate);
    return hipGetLastError();
}

// ======================================================================
// MD5 (RFC 1321)
// ======================================================================

// constexpr (not __constant__): indexed by unrolled round numbers, the
// constants become instruction literals instead of 64 scalar registers.
constexpr uint32_t kMd5K[64] = {
    0xd76aa478u, 0xe8c7b756u, 0x242070dbu, 0xc1bdceeeu, 0xf57c0fafu, 0x4787c62au, 0xa8304613u, 0xfd469501u,
    0x698098d8u, 0x8b44f7afu, 0xffff5bb1u, 0x895cd7beu, 0x6b901122u, 0xfd987193u, 0xa679438eu, 0x49b40821u,
    0xf61e2562u, 0xc040b340u, 0x265e5a51u, 0xe9b6c7aau, 0xd62f105du, 0x02441453u, 0xd8a1e681u, 0xe7d3fbc8u,
    0x21e1cde6u, 0xc33707d6u, 0xf4d50d87u, 0x455a14edu, 0xa9e3e905u, 0xfcefa3f8u, 0x676f02d9u, 0x8d2a4c8au,
    0xfffa3942u, 0x8771f681u, 0x6d9d6122u, 0xfde5380cu, 0xa4beea44u, 0x4bdecfa9u, 0xf6bb4b60u, 0xbebfbc70u,
    0x289b7ec6u, 0xeaa127fau, 0xd4ef3085u, 0x04881d05u, 0xd9d4d039u, 0xe6db99e5u, 0x1fa27cf8u, 0xc4ac5665u,
    0xf4292244u, 0x432aff97u, 0xab9423a7u, 0xfc93a039u, 0x655b59c3u, 0x8f0ccc92u, 0xffeff47du, 0x85845dd1u,
    0x6fa87e4fu, 0xfe2ce6e0u, 0xa3014314u, 0x4e0811a1u, 0xf7537e82u, 0xbd3af235u, 0x2ad7d2bbu, 0xeb86d391u};

// Rotate left by s (1..31) in one v_alignbit_b32: (x:x) >> (32 - s).
__device__ __forceinline__ uint32_t rotl32(uint32_t x, int s) {
    return __builtin_amdgcn_alignbit(x, x, 32 - s);
}

__device__ __forceinline__ void md5_compress(uint32_t h[4], const uint32_t x[16]) {
    constexpr int sh[16] = {7, 12, 17, 22, 5, 9, 14, 20, 4, 11, 16, 23, 6, 10, 15, 21};
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3];
#pragma unroll
    for (int i = 0; i < 64; i++) {
        // The four round functions as one v_bitop3 each (truth-table bit
        // (b << 2 | c << 1 | d)): F = b ? c : d (0xCA), G = d ? b : c (0xE4),
        // H = b ^ c ^ d (0x96), I = c ^ (b | ~d) (0x39).
        uint32_t f;
        int g;
        if (i < 16) { f = __builtin_amdgcn_bitop3_b32(b, c, d, 0xCA); g = i; }
        else if (i < 32) { f = __builtin_amdgcn_bitop3_b32(b, c, d, 0xE4); g = (5 * i + 1) & 15; }
        else if (i < 48) { f = __builtin_amdgcn_bitop3_b32(b, c, d, 0x96); g = (3 * i + 5) & 15; }
        else { f = __builtin_amdgcn_bitop3_b32(b, c, d, 0x39); g = (7 * i) & 15; }
        const uint32_t t = d;
        d = c;
        c = b;
        b = b + rotl32(a + f + (x[g] + kMd5K[i]), sh[(i >> 4) * 4 + (i & 3)]);
        a = t;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d;
}

// The message schedule of MD5: round i reads word md5_word(i) of the block.
__host__ __device__ constexpr int md5_word(int i) {
    return i < 16 ? i : i < 32 ? (5 * i + 1) & 15 : i < 48 ? (3 * i + 5) & 15 : (7 * i) & 15;
}

// One block from its 64 round inputs w[i] = x[md5_word(i)] + K[i], made
// beforehand by another wave: 4 VALU per round (v_bitop3_b32, v_add3_u32,
// v_alignbit_b32, v_add_u32) instead of md5_compress's 5, and no loads.
__device__ __forceinline__ void md5_compress_w(uint32_t h[4], const uint32_t w[64]) {
    constexpr int sh[16] = {7, 12, 17, 22, 5, 9, 14, 20, 4, 11, 16, 23, 6, 10, 15, 21};
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3];
#pragma unroll
    for (int i = 0; i < 64; i++) {
        uint32_t f;
        if (i < 16) f = __builtin_amdgcn_bitop3_b32(b, c, d, 0xCA);
        else if (i < 32) f = __builtin_amdgcn_bitop3_b32(b, c, d, 0xE4);
        else if (i < 48) f = __builtin_amdgcn_bitop3_b32(b, c, d, 0x96);
        else f = __builtin_amdgcn_bitop3_b32(b, c, d, 0x39);
        const uint32_t t = d;
        d = c;
        c = b;
        b = b + rotl32(a + f + w[i], sh[(i >> 4) * 4 + (i & 3)]);
        a = t;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d;
}

__device__ __forceinline__ void md5_init(uint32_t h[4]) {
    h[0] = 0x67452301u; h[1] = 0xefcdab89u; h[2] = 0x98badcfeu; h[3] = 0x10325476u;
}

__device__ __forceinline__ void store_digest(uint8_t *out, const uint32_t h[4]) {
    uint32_t *o = reinterpret_cast<uint32_t *>(out);  // 16-byte aligned slot
#pragma unroll
    for (int i = 0; i < 4; i++) o[i] = h[i];
}

// Number of 64-byte MD5 blocks for a message of `len` bytes (len % 4 == 0).
__device__ __forceinline__ uint32_t md5_blocks(uint64_t len) { return (uint32_t)((len + 8) / 64 + 1); }

// Padding word at byte offset o >= len of a len-byte message (len % 4 == 0)
// inside block b of nb.
__device__ __forceinline__ uint32_t md5_pad_word(uint64_t o, uint64_t len, uint32_t w,
                                                 uint32_t b, uint32_t nb) {
    if (o == len) return 0x80u;
    if (b == nb - 1 && w == 14) return (uint32_t)(len * 8);
    if (b == nb - 1 && w == 15) return (uint32_t)((len * 8) >> 32);
    return 0u;
}

// Bulk blocks: nblk whole 64-byte blocks at p, loaded kMd5Ahead blocks ahead of
// the block being compressed so that the loads of one lane stream behind
// its (serial) compression.  W is uint4 when p is 16-byte aligned, uint32_t
// otherwise.
constexpr int kMd5Ahead = 8;

template <typename W>
__device__ __forceinline__ void md5_load_block(uint32_t x[16], const W *p) {
    constexpr int n = 16 / (sizeof(W) / 4);
#pragma unroll
    for (int k = 0; k < n; k++) {
        const W v = gload(p + k);
        memcpy(&x[k * (sizeof(W) / 4)], &v, sizeof(W));
    }
}

// nblk blocks, block b loaded from addr(b) (a const W *).  Block j lives in
// ring slot j % kMd5Ring; the load of block j + kMd5Ahead goes into the slot
// compressed one step earlier, so the compression reads the loaded registers
// without copies.  The main loop's loads are unconditional (their wait counts
// are static); the tail loads only what exists.
constexpr int kMd5Ring = kMd5Ahead + 1;
struct Md5NoVisit {
    __device__ void operator()(const uint32_t *) const {}
};
// visit(x) sees each block's 16 words before it is compressed.
template <typename W, typename Addr, typename Visit = Md5NoVisit>
__device__ __forceinline__ void md5_blocks_at(uint32_t h[4], Addr addr, uint32_t nblk,
                                              Visit visit = Visit()) {
    // Fewer blocks than one turn of the ring plus the look-ahead: one load
    // and compression after the other.  Longer messages fill the ring with
    // unconditional loads in block order, so that the compiler's wait
    // counts at the main loop's top stay 8 blocks deep (after loads that
    // may have been skipped or were reordered, it waits there until every
    // load but the newest is done, which drains the ring once a turn).
    if (nblk < (uint32_t)(kMd5Ring + kMd5Ahead)) {
        for (uint32_t b = 0; b < nblk; b++) {
            uint32_t x[16];
            md5_load_block(x, addr(b));
            visit(x);
            md5_compress(h, x);
        }
        return;
    }
    uint32_t buf[kMd5Ring][16];
    // in block order: block 0 must be the oldest load when it is compressed
#pragma unroll
    for (int k = 0; k < kMd5Ahead; k++) {
        md5_load_block(buf[k], addr((uint32_t)k));
        __builtin_amdgcn_sched_barrier(0);
    }
    uint32_t b = 0;
    for (; b + kMd5Ring + kMd5Ahead <= nblk; b += kMd5Ring) {
#pragma unroll
        for (int k = 0; k < kMd5Ring; k++) {
            md5_load_block(buf[(k + kMd5Ahead) % kMd5Ring], addr(b + k + kMd5Ahead));
            // keeps each load at the top of its block: left free, the
            // scheduler gathers a whole ring's loads at the loop's end and
            // the loop top then waits a full memory latency for them
            __builtin_amdgcn_sched_barrier(0);
            visit(buf[k]);
            md5_compress(h, buf[k]);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    for (; b < nblk; b += kMd5Ring) {
#pragma unroll
        for (int k = 0; k < kMd5Ring; k++) {
            if (b + k < nblk) {
                const uint32_t nxt = b + k + kMd5Ahead;
                if (nxt < nblk) md5_load_block(buf[(k + kMd5Ahead) % kMd5Ring], addr(nxt));
                visit(buf[k]);
                md5_compress(h, buf[k]);
            }
        }
    }
}

// nblk consecutive blocks at p: md5_blocks_at's ring, but the loads of a
// ring turn take their addresses as constant offsets (at most 572 bytes)
// from one pointer advanced once per turn.  A single MD5 wave issues one
// instruction per 4 cycles whatever its kind, so per-block address
// arithmetic (64-bit shifts and adds, ~5 instructions) costs as much as a
// VALU round step.
// step(y) consumes each block's 16 loaded words in order.
template <typename W, typename Step>
__device__ __forceinline__ void md5_stream_steps(const W *p, uint32_t nblk, Step step) {
    constexpr int WB = 64 / sizeof(W);  // W per block
    if (nblk < (uint32_t)(kMd5Ring + kMd5Ahead)) {
        for (uint32_t b = 0; b < nblk; b++) {
            uint32_t x[16];
            md5_load_block(x, p + (size_t)WB * b);
            step(x);
        }
        return;
    }
    uint32_t buf[kMd5Ring][16];
#pragma unroll
    for (int k = 0; k < kMd5Ahead; k++) {
        md5_load_block(buf[k], p + WB * k);
        __builtin_amdgcn_sched_barrier(0);
    }
    const W *q = p + WB * kMd5Ahead;  // block b + kMd5Ahead of the turn at b
    uint32_t b = 0;
    for (; b + kMd5Ring + kMd5Ahead <= nblk; b += kMd5Ring, q += WB * kMd5Ring) {
#pragma unroll
        for (int k = 0; k < kMd5Ring; k++) {
            md5_load_block(buf[(k + kMd5Ahead) % kMd5Ring], q + WB * k);
            __builtin_amdgcn_sched_barrier(0);  // as in md5_blocks_at
            step(buf[k]);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    for (; b < nblk; b += kMd5Ring) {
#pragma unroll
        for (int k = 0; k < kMd5Ring; k++) {
            if (b + k < nblk) {
                const uint32_t nxt = b + k + kMd5Ahead;
                if (nxt < nblk) md5_load_block(buf[(k + kMd5Ahead) % kMd5Ring], p + (size_t)WB * nxt);
                step(buf[k]);
            }
        }
    }
}

template <typename W, typename Visit = Md5NoVisit>
__device__ __forceinline__ void md5_stream(uint32_t h[4], const W *p, uint32_t nblk,
                                           Visit visit = Visit()) {
    md5_stream_steps(p, nblk, [&](const uint32_t *x) {
        visit(x);
        md5_compress(h, x);
    });
}

template <typename W>
__device__ __forceinline__ void md5_bulk(uint32_t h[4], const uint8_t *p, uint32_t nblk) {
    md5_stream(h, reinterpret_cast<const W *>(p), nblk);
}

// 64 zero bytes: the message block of a slice wholly past its data.
__device__ const uint4 kZeroBlock[4] = {};

// The nmsg message blocks of a slice as one lane reads them: blocks below
// nfull = valid / 64 from the slice, block nfull from its tail block tb,
// zero blocks after that.  The first nmain blocks, nmain = the smallest
// min(nfull, nmsg) over the wave's active lanes (all of them when every
// lane holds a whole slice of one size, the common case), stream from
// contiguous addresses; only the rest select per block.  Lanes may differ
// in nmsg (a batched wave can hold slices of two jobs).
template <typename W, typename Visit = Md5NoVisit>
__device__ __forceinline__ void md5_slice_blocks(uint32_t h[4], const uint8_t *ptr, uint32_t nfull,
                                                 const uint4 *tb, uint32_t nmsg,
                                                 Visit visit = Visit()) {
    const uint32_t mine = min(nfull, nmsg);
    // wave minimum: each turn takes the value of a lane below the current
    // one, so it ends after at most one turn per distinct smaller value
    uint32_t nmain = __builtin_amdgcn_readfirstlane(mine);
    for (;;) {
        const uint64_t m = __ballot(mine < nmain);
        if (m == 0) break;
        nmain = __builtin_amdgcn_readlane((int)mine, __builtin_ctzll(m));
    }
    md5_stream(h, reinterpret_cast<const W *>(ptr), nmain, visit);
    if (nmain == nmsg) return;
    md5_blocks_at<W>(h, [=](uint32_t b) {
        b += nmain;
        const void *a = b < nfull ? (const void *)(ptr + 64 * (size_t)b)
                      : b == nfull ? (const void *)tb : (const void *)kZeroBlock;
        return reinterpret_cast<const W *>(a);
    }, nmsg - nmain, visit);
}

// One MD5 wave per CU: each lane streams its own message, so the waves'
// scattered 64-byte reads would otherwise share (and thrash) a CU's L1 and
// address unit.  The launches request this much dynamic LDS (unused) per
// single-wave workgroup, which keeps the dispatcher from placing two of
// them on one CU.
constexpr uint32_t kMd5LdsReserve = 96 * 1024;

// ======================================================================
// K2a: slice MD5, one lane per slice
// ======================================================================

// MD5 of one zero-padded S-byte slice into out (one lane).  A16: the slice
// start is 16-byte aligned.  tb: the slice's tail block (k_desc_tails).  The
// S / 64 whole message blocks run through md5_slice_blocks whatever `valid`:
// block b comes from the slice while it is whole data, from tb for the block
// holding byte `valid`, and from kZeroBlock after that; only the last
// (padding) blocks take the word-by-word path.
template <bool A16>
__device__ __forceinline__ void slice_md5_lane(const SliceDesc d, const uint4 *tb, uint32_t S,
                                               uint8_t *out) {
    using W = typename std::conditional<A16, uint4, uint32_t>::type;
    const uint32_t nb = md5_blocks(S);
    const uint32_t nmsg = S / 64;         // whole message blocks
    const uint32_t nfull = d.valid / 64;  // of them, whole blocks of real data
    uint32_t h[4];
    md5_init(h);
    md5_slice_blocks<W>(h, d.ptr, nfull, tb, nmsg);
    for (uint32_t b = nmsg; b < nb; b++) {
        const uint32_t base = b * 64u;
        uint32_t x[16];
#pragma unroll
        for (int w = 0; w < 16; w++) {
            const uint32_t o = base + 4u * w;
            x[w] = (o < S) ? load_word(d.ptr, o, d.valid) : md5_pad_word(o, S, w, b, nb);
        }
        md5_compress(h, x);
    }
    store_digest(out, h);
}

template <bool A16>
__global__ __launch_bounds__(64) void k_slice_md5(const SliceDesc *__restrict__ slices,
                                                  const uint4 *__restrict__ tail_blocks, uint32_t n,
                                                  uint32_t S, uint8_t *__restrict__ md5_out) {
    const uint32_t s = blockIdx.x * 64 + threadIdx.x;
    if (s >= n) return;
    slice_md5_lane<A16>(slices[s], tail_blocks + 4 * (size_t)s, S, md5_out + (size_t)s * 16);
}

hipError_t launch_slice_md5(const SliceDesc *d_slices, const uint4 *d_tail_blocks, uint32_t n,
                            uint32_t S, uint8_t *d_md5, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    // Slices are 16-byte aligned when S is a multiple of 16 (file bases are).
    if (S % 16 == 0)
        hipLaunchKernelGGL(k_slice_md5<true>, dim3((n + 63) / 64), dim3(64), kMd5LdsReserve, stream,
                           d_slices, d_tail_blocks, n, S, d_md5);
    else
        hipLaunchKernelGGL(k_slice_md5<false>, dim3((n + 63) / 64), dim3(64), kMd5LdsReserve,
                           stream, d_slices, d_tail_blocks, n, S, d_md5);
    return hipGetLastError();
}

// ======================================================================
// K2b: slice CRC-32, one wave per slice
// ======================================================================
//
// Each lane takes the CRC-32 of a contiguous segment; the 64 segment CRCs are
// merged in a shuffle tree with crc(A || B) = (crc(A) * x^(8 |B|) mod P) ^
// crc(B), which holds for the init/xorout 0xFFFFFFFF form.  Polynomials are
// in reflected form: bit 31 is x^0.

constexpr uint32_t kCrcPoly = 0xEDB88320u;

struct CrcPowers {
    uint32_t x2n[40];  // x^(2^k) mod P, reflected; 8 n < 2^35 for n < 2^32 bytes
};

__device__ __forceinline__ uint32_t crc_mulmod(uint32_t a, uint32_t b) {
    uint32_t p = 0;
#pragma unroll
    for (int i = 0; i < 32; i++) {
        if (a & (0x80000000u >> i)) p ^= b;
        b = (b & 1u) ? ((b >> 1) ^ kCrcPoly) : (b >> 1);
    }
    return p;
}

// x^(8 n) mod P
__device__ __forceinline__ uint32_t crc_shift_op(const CrcPowers &pw, uint32_t n) {
    uint32_t p = 0x80000000u;
    uint32_t k = 3;
    while (n) {
        if (n & 1u) p = crc_mulmod(pw.x2n[k], p);
        n >>= 1;
        k++;
    }
    return p;
}

// Reflected table CRC: t[0] is the byte table and t[k][i] = t[k-1][i] >> 8
// ^ t[0][t[k-1][i] & 0xFF], the effect of byte i followed by k zero bytes;
// four tables step four bytes at a time (slicing by 4), sixteen step sixteen.
// Filled by a whole 256-thread workgroup.
constexpr uint32_t kCrcAhead = 8;  // 16-byte loads in flight per lane ahead of the CRC steps

struct CrcLds {
    uint32_t t[16][256];
    uint32_t part[4];  // per-wave partial CRCs
};

__device__ __forceinline__ void crc_tables_init(CrcLds &L) {
    const uint32_t i = threadIdx.x;
    uint32_t c = i;
    for (int k = 0; k < 8; k++) c = (c & 1u) ? (kCrcPoly ^ (c >> 1)) : (c >> 1);
    L.t[0][i] = c;
    __syncthreads();
    for (int k = 1; k < 16; k++) {
        const uint32_t q = L.t[k - 1][i];
        L.t[k][i] = (q >> 8) ^ L.t[0][q & 0xFFu];
        __syncthreads();
    }
}

__device__ __forceinline__ uint32_t crc_step4(const uint32_t (*t)[256], uint32_t c, uint32_t w) {
    c ^= w;
    return t[3][c & 0xFFu] ^ t[2][(c >> 8) & 0xFFu] ^ t[1][(c >> 16) & 0xFFu] ^ t[0][c >> 24];
}
__device__ __forceinline__ uint32_t crc_step4(const CrcLds &L, uint32_t c, uint32_t w) {
    return crc_step4(L.t, c, w);
}
// Sixteen bytes per step (slicing by 16, tables t[0..15]): byte j of the
// sixteen is followed by 15 - j bytes, so it indexes t[15 - j]; only the
// first word depends on the running CRC.
__device__ __forceinline__ uint32_t crc_step16(const uint32_t (*t)[256], uint32_t c, uint32_t w0,
                                               uint32_t w1, uint32_t w2, uint32_t w3) {
    c ^= w0;
    return (t[15][c & 0xFFu] ^ t[14][(c >> 8) & 0xFFu] ^ t[13][(c >> 16) & 0xFFu] ^ t[12][c >> 24]) ^
           (t[11][w1 & 0xFFu] ^ t[10][(w1 >> 8) & 0xFFu] ^ t[9][(w1 >> 16) & 0xFFu] ^ t[8][w1 >> 24]) ^
           (t[7][w2 & 0xFFu] ^ t[6][(w2 >> 8) & 0xFFu] ^ t[5][(w2 >> 16) & 0xFFu] ^ t[4][w2 >> 24]) ^
           (t[3][w3 & 0xFFu] ^ t[2][(w3 >> 8) & 0xFFu] ^ t[1][(w3 >> 16) & 0xFFu] ^ t[0][w3 >> 24]);
}

// The same four tables filled by one wave alone: entry t[k][i] is byte i
// followed by k zero bytes, i.e. 8 (k + 1) bit steps from i, so each lane
// computes its own entries and only the wave's LDS order has to hold.
template <int NT = 4>
__device__ __forceinline__ void crc_tables_init_wave(uint32_t (*t)[256]) {
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t e = lane; e < 256; e += 64) {
        uint32_t c = e;
#pragma unroll
        for (int k = 0; k < NT; k++) {
#pragma unroll
            for (int b = 0; b < 8; b++) c = (c & 1u) ? (kCrcPoly ^ (c >> 1)) : (c >> 1);
            t[k][e] = c;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// CRC-32 of one zero-padded S-byte slice by one wave: lane l takes bytes
// [l seg, (l + 1) seg), merged in a shuffle tree as in crc_slice_block.  t
// holds the 16 tables of slicing by 16 (crc_tables_init_wave<16>).  A lane's
// 16-byte loads run kCrcAhead chunks ahead of its CRC steps, so the steps
// wait on one load latency per segment rather than one per chunk.  Lane 0
// stores the result.
__device__ __forceinline__ void crc_slice_wave(const uint32_t (*t)[256], const SliceDesc d,
                                               uint32_t S, bool a16, const CrcPowers &pw,
                                               uint32_t *out) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t q = a16 ? 16u : 4u;
    const uint32_t seg = ((S + 63u) / 64u + q - 1) / q * q;
    const uint32_t lo = min(lane * seg, S), hi = min(lo + seg, S);
    uint32_t c = 0xFFFFFFFFu;
    uint32_t o = lo;
    if (a16) {
        // whole 16-byte chunks inside the real bytes, then words
        const uint32_t whole = d.valid > lo ? lo + (min(hi, d.valid) - lo) / 16 * 16 : lo;
        const uint32_t nch = (whole - lo) / 16;
        const uint4 *p = reinterpret_cast<const uint4 *>(d.ptr + lo);
        uint4 ring[kCrcAhead];
#pragma unroll
        for (uint32_t k = 0; k < kCrcAhead; k++)
            ring[k] = k < nch ? gload(p + k) : make_uint4(0, 0, 0, 0);
        for (uint32_t i = 0; i < nch; i += kCrcAhead) {
#pragma unroll
            for (uint32_t k = 0; k < kCrcAhead; k++) {
                if (i + k < nch) {
                    const uint4 v = ring[k];
                    if (i + k + kCrcAhead < nch) ring[k] = gload(p + i + k + kCrcAhead);
                    c = crc_step16(t, c, v.x, v.y, v.z, v.w);
                }
            }
        }
        o = whole;
    }
    for (; o < hi; o += 4) c = crc_step4(t, c, load_word(d.ptr, o, d.valid));
    c ^= 0xFFFFFFFFu;
    for (uint32_t step = 1; step < 64; step <<= 1) {
        const uint32_t other = __shfl_down(c, step, 64);
        if ((lane & (2 * step - 1)) == 0) {
            const uint32_t rlo = min((lane + step) * seg, S), rhi = min((lane + 2 * step) * seg, S);
            c = crc_mulmod(crc_shift_op(pw, rhi - rlo), c) ^ other;
        }
    }
    if (lane == 0) *out = c;
}

// CRC-32 of one zero-padded S-byte slice by a whole 256-thread workgroup:
// thread t takes bytes [t seg, (t + 1) seg) (seg a multiple of 16 when the
// slice is 16-byte aligned), then the 256 segment CRCs are merged with
// crc(A || B) = crc(A) * x^(8 |B|) mod P ^ crc(B), which holds for the
// init / xorout 0xFFFFFFFF form: a shuffle tree inside each wave, then the
// four wave results in order.  Thread 0 stores the result.
__device__ __forceinline__ void crc_slice_block(CrcLds &L, const SliceDesc d, uint32_t S,
                                                bool a16, const CrcPowers &pw, uint32_t *out) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t q = a16 ? 16u : 4u;
    const uint32_t seg = ((S + 255u) / 256u + q - 1) / q * q;
    const uint32_t lo = min(tid * seg, S), hi = min(lo + seg, S);
    uint32_t c = 0xFFFFFFFFu;
    uint32_t o = lo;
    if (a16) {
        // whole 16-byte chunks inside the real bytes (slicing by 16, loads
        // kCrcAhead chunks ahead as in crc_slice_wave), then words
        const uint32_t whole = d.valid > lo ? lo + (min(hi, d.valid) - lo) / 16 * 16 : lo;
        const uint32_t nch = (whole - lo) / 16;
        const uint4 *p = reinterpret_cast<const uint4 *>(d.ptr + lo);
        uint4 ring[kCrcAhead];
#pragma unroll
        for (uint32_t k = 0; k < kCrcAhead; k++)
            ring[k] = k < nch ? gload(p + k) : make_uint4(0, 0, 0, 0);
        for (uint32_t i = 0; i < nch; i += kCrcAhead) {
#pragma unroll
            for (uint32_t k = 0; k < kCrcAhead; k++) {
                if (i + k < nch) {
                    const uint4 v = ring[k];
                    if (i + k + kCrcAhead < nch) ring[k] = gload(p + i + k + kCrcAhead);
                    c = crc_step16(L.t, c, v.x, v.y, v.z, v.w);
                }
            }
        }
        o = whole;
    }
    for (; o < hi; o += 4) c = crc_step4(L, c, load_word(d.ptr, o, d.valid));
    c ^= 0xFFFFFFFFu;
    for (uint32_t step = 1; step < 64; step <<= 1) {
        const uint32_t other = __shfl_down(c, step, 64);
        if ((lane & (2 * step - 1)) == 0) {
            const uint32_t rlo = min((tid + step) * seg, S), rhi = min((tid + 2 * step) * seg, S);
            c = crc_mulmod(crc_shift_op(pw, rhi - rlo), c) ^ other;
        }
    }
    if (lane == 0) L.part[w] = c;
    __syncthreads();
    if (tid == 0) {
        uint32_t r = L.part[0];
        for (uint32_t k = 1; k < 4; k++) {
            const uint32_t blo = min(k * 64 * seg, S), bhi = min((k + 1) * 64 * seg, S);
            r = crc_mulmod(crc_shift_op(pw, bhi - blo), r) ^ L.part[k];
        }
        *out = r;
    }
    __syncthreads();  // L.part free for the next slice
}

__global__ __launch_bounds__(256) void k_slice_crc32(const SliceDesc *__restrict__ slices,
                                                     uint32_t S, uint32_t *__restrict__ crc_out,
                                                     CrcPowers pw) {
    __shared__ CrcLds L;
    crc_tables_init(L);
    crc_slice_block(L, slices[blockIdx.x], S, S % 16 == 0, pw, crc_out + blockIdx.x);
}

static CrcPowers make_crc_powers() {
    auto mul = [](uint32_t a, uint32_t b) {
        uint32_t p = 0;
        for (int i = 0; i < 32; i++) {
            if (a & (0x80000000u >> i)) p ^= b;
            b = (b & 1u) ? ((b >> 1) ^ kCrcPoly) : (b >> 1);
        }
        return p;
    };
    CrcPowers pw;
    pw.x2n[0] = 0x40000000u;  // x^1
    for (int k = 1; k < 40; k++) pw.x2n[k] = mul(pw.x2n[k - 1], pw.x2n[k - 1]);
    return pw;
}

hipError_t launch_slice_crc32(const SliceDesc *d_slices, uint32_t n, uint32_t S, uint32_t *d_crc,
                              hipStream_t stream) {
    if (n == 0) return hipSuccess;
    static const CrcPowers pw = make_crc_powers();
    hipLaunchKernelGGL(k_slice_crc32, dim3(n), dim3(256), 0, stream, d_slices, S, d_crc, pw);
    return hipGetLastError();
}

// ======================================================================
// K3: recovery packet MD5, one lane per packet
// ======================================================================

struct PacketPrefix {
    uint32_t w[8];  // set_id (4 words) || "PAR 2.0\0RecvSlic" (4 words)
};

// The packet hash covers set_id || type || le32(exp) || S data bytes:
// len = 36 + S.  Block 0 holds the 36-byte prefix and data bytes 0..27;
// block k >= 1 holds data bytes 64 k - 36 .. 64 k + 27; blocks k >= nfull =
// len / 64 carry the padding.  md5_packet_words assembles any block k from
// the prefix, the data and the padding.
__device__ __forceinline__ void md5_packet_words(uint32_t x[16], const PacketPrefix &pre,
                                                 uint32_t exp, const uint8_t *data, uint64_t len,
                                                 uint32_t k, uint32_t nb) {
    const uint64_t base = (uint64_t)k * 64u;
#pragma unroll
    for (int w = 0; w < 16; w++) {
        const uint64_t o = base + 4u * w;
        uint32_t v;
        if (o < 32) v = pre.w[o / 4];
        else if (o < 36) v = exp;
        else if (o < len) v = gload(reinterpret_cast<const uint32_t *>(data + (o - 36)));
        else v = md5_pad_word(o, len, w, k, nb);
        x[w] = v;
    }
}

__device__ __forceinline__ void md5_packet_block(uint32_t h[4], const PacketPrefix &pre,
                                                 uint32_t exp, const uint8_t *data, uint64_t len,
                                                 uint32_t k, uint32_t nb) {
    uint32_t x[16];
    md5_packet_words(x, pre, exp, data, len, k, nb);
    md5_compress(h, x);
}

// Packet blocks k0 .. k0 + nblk - 1 (k0 >= 1) when the block data starts
// 16-byte aligned and S % 16 == 0.  Block k is data words 16k - 9 .. 16k + 6:
// the last word of the previous block's aligned 64 bytes, carried, and the
// first 15 words of data[64k - 32, 64k + 32), loaded as four 16-byte loads
// (which end at or before byte S, as S is a multiple of 16).  step(x) takes
// each block's 16 words in order.
template <typename Step>
__device__ __forceinline__ void md5_packet_bulk16_steps(const uint8_t *data, uint32_t k0,
                                                        uint32_t nblk, Step &&step) {
    if (nblk == 0) return;
    uint32_t carry = gload(reinterpret_cast<const uint32_t *>(data + 4 * (16 * (size_t)k0 - 9)));
    md5_stream_steps(reinterpret_cast<const uint4 *>(data + 64 * (size_t)k0 - 32), nblk,
                     [&](const uint32_t *y) {
                         uint32_t x[16];
                         x[0] = carry;
#pragma unroll
                         for (int w = 1; w < 16; w++) x[w] = y[w - 1];
                         carry = y[15];
                         step(x);
                     });
}

__device__ __forceinline__ void md5_packet_bulk16(uint32_t h[4], const uint8_t *data, uint32_t k0,
                                                  uint32_t nblk) {
    md5_packet_bulk16_steps(data, k0, nblk, [&](const uint32_t *x) { md5_compress(h, x); });
}

__global__ __launch_bounds__(64) void k_recovery_md5(const uint8_t *__restrict__ rec,
                                                     uint32_t count, uint32_t S, uint32_t first_exp,
                                                     PacketPrefix pre, uint8_t *__restrict__ md5_out) {
    const uint32_t r = blockIdx.x * 64 + threadIdx.x;
    if (r >= count) return;
    const uint8_t *data = rec + (size_t)r * S;
    const uint64_t len = 36ull + S;
    const uint32_t nb = md5_blocks(len);
    const uint32_t nfull = (uint32_t)(len / 64);
    uint32_t h[4];
    md5_init(h);
    md5_packet_block(h, pre, first_exp + r, data, len, 0, nb);
    if (nfull > 1) {
        if (S % 16 == 0 && reinterpret_cast<uintptr_t>(rec) % 16 == 0)
            md5_packet_bulk16(h, data, 1, nfull - 1);
        else
            md5_bulk<uint32_t>(h, data + 28, nfull - 1);
    }
    for (uint32_t b = nfull > 1 ? nfull : 1; b < nb; b++)
        md5_packet_block(h, pre, first_exp + r, data, len, b, nb);
    store_digest(md5_out + (size_t)r * 16, h);
}

static PacketPrefix make_packet_prefix(const uint8_t set_id[16]) {
    static const uint8_t type_tag[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0,
                                         'R', 'e', 'c', 'v', 'S', 'l', 'i', 'c'};
    PacketPrefix pre;
    for (int w = 0; w < 4; w++) {
        pre.w[w] = (uint32_t)set_id[4 * w] | ((uint32_t)set_id[4 * w + 1] << 8) |
                   ((uint32_t)set_id[4 * w + 2] << 16) | ((uint32_t)set_id[4 * w + 3] << 24);
        pre.w[4 + w] = (uint32_t)type_tag[4 * w] | ((uint32_t)type_tag[4 * w + 1] << 8) |
                       ((uint32_t)type_tag[4 * w + 2] << 16) | ((uint32_t)type_tag[4 * w + 3] << 24);
    }
    return pre;
}

hipError_t launch_recovery_md5(const uint8_t *d_rec, uint32_t count, uint32_t S,
                               uint32_t first_exp, const uint8_t set_id[16], uint8_t *d_md5,
                               hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const PacketPrefix pre = make_packet_prefix(set_id);
    hipLaunchKernelGGL(k_recovery_md5, dim3((count + 63) / 64), dim3(64), kMd5LdsReserve, stream,
                       d_rec, count, S, first_exp, pre, d_md5);
    return hipGetLastError();
}

// ======================================================================
// Fused create pass: encode, slice MD5 and packet MD5 in one launch
// ======================================================================
//
// One workgroup per CU, producers first: workgroups [0, E) take encode
// units, [E, E + h2) hash slices (one lane per slice, on wave 0; waves 1-3
// compute slice CRCs) and then take encode units too, and the last h3 hash
// recovery packets (one lane per packet), so the roles that wait are
// dispatched after the ones they wait on.  A unit is one row block (8 tiles
// = 16 recovery rows) of one 512-byte column group, all K steps.  Units go
// band by band (next_band_width column groups per band, narrow at both ends)
// and row block by row block inside a band, so the workgroups running at one
// time share the A fragments of a few row blocks and the source columns of
// one band; they come from per-XCD work queues (UnitQueue).  Finished units
// are counted in their band's counter (stores released at agent scope
// first); a packet lane hashes the blocks of band b once that count is full
// (acquire), so the packet hashes run under the encode and only the last,
// narrowest band is hashed after it.  The MD5 workgroups keep one CU each,
// one wave per SIMD, for the serial hashes of their lanes.

struct PassArgs {
    const SliceDesc *inputs;
    const uint4 *tail_blocks;  // 4 per slice (k_desc_tails)
    const uint4 *afrag;
    uint8_t *rec;
    uint8_t *slice_md5;
    uint32_t *slice_crc;
    uint8_t *pkt_md5;
    uint32_t *sync;  // stamps, [kPassBands + b] band counts, band starts
    uint32_t *err;   // sticky error word: kept across passes until the host reads it
    PacketPrefix pre;
    uint32_t n, nq, R, S;
    uint32_t row_blocks, col_groups, band_cols, n_bands, n_units;
    uint32_t h3_blocks, h2_blocks;
    uint32_t stall_ticks;  // give up after this long without progress (100 MHz ticks)
    CrcPowers pw;
};

// Device work queues of the encode units.  Workgroups are dealt to the 8
// XCDs round-robin (XCD = blockIdx % 8), and units are taken in rounds of E
// (the encode workgroups): round r holds units [r E, (r + 1) E), of which
// XCD x owns cnt_x slots (cnt_x = its encode workgroups).  Inside an XCD the
// units are not pre-assigned: each of its workgroups takes the XCD's next
// slot from a counter when it is free, so faster workgroups take more units
// and a unit that costs more delays only the workgroup that took it; a
// slice-hash workgroup whose hashes are done takes units too.  Queue
// slots increase, so each workgroup's units come in increasing order (the
// band walks rely on it).  XCD x owns the contiguous slots [off_x, off_x +
// cnt_x) of a round (off_x = encode workgroups of lower XCDs), so units
// adjacent in order run on one XCD and share its L2 copy of their operands:
// one row block's column groups in the batched pass (whose small-slice jobs
// stream large A fragments through few column groups), a tile of row blocks
// and column groups in the fused pass.
struct UnitQueue {
    uint32_t *ctr;
    uint32_t e_blocks, cnt, off;
    __device__ __forceinline__ UnitQueue(uint32_t *sync, uint32_t blk, uint32_t e) {
        const uint32_t x = blk % kPassXcds;
        ctr = sync + kPassQueues + x;
        e_blocks = e;
        // 0 only on an XCD without encode workgroups (grids of fewer than 8),
        // whose hash workgroups then find no units
        cnt = e / kPassXcds + (x < e % kPassXcds ? 1u : 0u);
        off = x * (e / kPassXcds) + min(x, e % kPassXcds);
    }
    // The next unit for the calling workgroup (all threads call it): thread
    // 0 takes a slot and broadcasts it through `slot`; readfirstlane makes the
    // value a scalar, so every branch on it is uniform and every thread
    // reaches the barriers of the unit loop.
    __device__ __forceinline__ uint32_t next(uint32_t *slot) const {
        if (cnt == 0) return 0xFFFFFFFFu;
        if (threadIdx.x == 0) {
            const uint32_t k = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            *slot = (k / cnt) * e_blocks + off + k % cnt;
        }
        __syncthreads();
        const uint32_t u = __builtin_amdgcn_readfirstlane(*(volatile uint32_t *)slot);
        __syncthreads();  // `slot` is free for the next call
        return u;
    }
};

// Waits until *cnt >= target, then acquires at agent scope.  The count is
// read wave-uniformly.  Returns false (and sets kPassErrWait) when it has not
// moved for stall_ticks of the 100 MHz real-time counter, so a wait always
// ends.
__device__ __forceinline__ bool pass_wait(const uint32_t *cnt, uint32_t target, uint32_t stall_ticks,
                          uint32_t *flags) {
    uint32_t seen = __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    uint64_t since = __builtin_amdgcn_s_memrealtime();
    while (seen < target) {
        __builtin_amdgcn_s_sleep(16);
        const uint32_t now = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        const uint64_t t = __builtin_amdgcn_s_memrealtime();
        if (now != seen) {
            seen = now;
            since = t;
        } else if (t - since > stall_ticks) {
            atomicOr(flags, kPassErrWait);
            return false;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

// Packet r (exponent r): the 16 message words of every block in order, each
// band's blocks as the band completes, then the padding, to emit(x).  Blocks
// [1, bulk_end) are whole data blocks (streamed).  False when a band wait
// gave up; stops early once `ok` is false.
template <typename Emit>
__device__ __forceinline__ bool pass_packet_walk(const PassArgs &a, uint32_t r, const bool &ok,
                                                 Emit &&emit) {
    const uint8_t *data = a.rec + (size_t)r * a.S;
    const bool a16 = a.S % 16 == 0 && reinterpret_cast<uintptr_t>(a.rec) % 16 == 0;
    const uint64_t len = 36ull + a.S;
    const uint32_t nb = md5_blocks(len), nfull = (uint32_t)(len / 64);
    const uint32_t bulk_end = nfull > 1 ? nfull : 1;
    uint32_t k = 0;  // next block
    const uint32_t *bstart = a.sync + kPassBands + a.n_bands;
    for (uint32_t b = 0; b < a.n_bands && ok; b++) {
        const uint32_t c1 = bstart[b + 1], w = c1 - bstart[b];
        if (!pass_wait(a.sync + kPassBands + b, w * a.row_blocks, a.stall_ticks, a.err))
            return false;
        // data bytes [0, e) of every block are final; block k needs up to 64 k + 28
        const uint64_t e = min((uint64_t)c1 * (kMfThreads * 2), (uint64_t)a.S);
        const uint32_t avail = b + 1 == a.n_bands ? nb : e >= 28 ? (uint32_t)((e - 28) / 64) + 1 : 0;
        if (k == 0 && avail > 0) {
            uint32_t x[16];
            md5_packet_words(x, a.pre, r, data, len, 0, nb);
            emit(x);
            k = 1;
        }
        const uint32_t stop = min(avail, bulk_end);
        if (stop > k) {
            if (a16) md5_packet_bulk16_steps(data, k, stop - k, emit);
            else md5_stream_steps(reinterpret_cast<const uint32_t *>(data + 64 * (size_t)k - 36),
                                  stop - k, emit);
            k = stop;
        }
    }
    for (; k < nb && ok; k++) {
        uint32_t x[16];
        md5_packet_words(x, a.pre, r, data, len, k, nb);
        emit(x);
    }
    return true;
}

// A packet workgroup hashes 64 packets with two waves.  The feeder (wave 1)
// walks the bands and the data (pass_packet_walk) and writes each block's 64
// round inputs x[md5_word(i)] + K[i] to a ring of kPktRing blocks in LDS; the
// chain (wave 0) runs the rounds from them (md5_compress_w).  The chain is
// the pass's serial path and a lone wave issues one instruction per ~4
// cycles whatever its kind, so moving the loads and the message additions to
// the other wave cuts its instructions per block from ~331 to ~300 (260
// round VALU).  Two counters in LDS pace the waves: prod (blocks written)
// and cons (blocks whose inputs the chain has used), both counted over all
// packets of the workgroup; the feeder writes block j once cons > j -
// kPktRing, the chain reads it once prod > j.  A wave that gives up (a band
// wait that gave up, kPktAbort from the other wave, or no movement for the
// wait limit) stores kPktAbort in its own counter, so the other's waits end
// too.
constexpr uint32_t kPktRing = 4;  // blocks of 64 lanes x 64 round inputs (16 KiB each)
constexpr uint32_t kPktAbort = 0xFFFFFFFFu;
// word offsets in the flag area: the two counters, then one word per lane so
// that a counter is stored by every lane of a wave in one conflict-free
// store (lane 0 to the counter, the others to their own word)
constexpr uint32_t kPktProd = 0, kPktCons = 1, kPktLaneWords = 64;

// The counters are read and written as relaxed workgroup-scope atomics (LDS
// reads and writes; volatile accesses through these pointers would be flat
// memory operations); the ring's data is ordered around them by explicit
// waits.
__device__ __forceinline__ uint32_t pkt_load(uint32_t *f) {
    return __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void pkt_store(uint32_t *f, uint32_t v) {
    __hip_atomic_store(f, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Waits until the counter f >= target; false when the other wave gave up or
// f has not moved for `ticks` of the 100 MHz real-time counter.
__device__ __forceinline__ bool pkt_wait(uint32_t *f, uint32_t target, uint32_t ticks) {
    uint32_t seen = __builtin_amdgcn_readfirstlane(pkt_load(f));
    if (seen >= target) return seen != kPktAbort;
    uint64_t since = __builtin_amdgcn_s_memrealtime();
    for (;;) {
        __builtin_amdgcn_s_sleep(1);
        const uint32_t now = __builtin_amdgcn_readfirstlane(pkt_load(f));
        if (now >= target) return now != kPktAbort;
        const uint64_t t = __builtin_amdgcn_s_memrealtime();
        if (now != seen) {
            seen = now;
            since = t;
        } else if (t - since > ticks) {
            return false;
        }
    }
}

__device__ __forceinline__ uint4 *pkt_slot(uint4 *ring, uint32_t j) {
    return ring + (size_t)(j % kPktRing) * 16 * 64 + (threadIdx.x & 63);
}

// Feeder lane of packet r: blocks [J, J + nb) of the ring.
__device__ __forceinline__ bool pkt_feed(const PassArgs &a, uint32_t r, uint4 *ring,
                                         uint32_t *flags, uint32_t *own, uint32_t J,
                                         uint32_t ticks) {
    uint32_t j = J;
    bool ok = true;
    const bool walked = pass_packet_walk(a, r, ok, [&](const uint32_t *x) {
        if (!ok) return;
        if (j >= kPktRing && !pkt_wait(flags + kPktCons, j + 1 - kPktRing, ticks)) {
            ok = false;
            return;
        }
        uint4 *slot = pkt_slot(ring, j);
#pragma unroll
        for (int q = 0; q < 16; q++)
            slot[q * 64] = make_uint4(x[md5_word(4 * q)] + kMd5K[4 * q],
                                      x[md5_word(4 * q + 1)] + kMd5K[4 * q + 1],
                                      x[md5_word(4 * q + 2)] + kMd5K[4 * q + 2],
                                      x[md5_word(4 * q + 3)] + kMd5K[4 * q + 3]);
        // the block's inputs are in LDS before prod says so
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
        asm volatile("" ::: "memory");
        pkt_store(own, j + 1);
        j++;
    });
    if (walked && ok) return true;
    if (walked) atomicOr(a.err, kPassErrWait);  // a band wait that gave up flagged it already
    pkt_store(own, kPktAbort);
    return false;
}

// Chain lane: the digest of blocks [J, J + nb) of the ring, stored when
// `store`.  Block j first reads the inputs of block j + 1 (into the other of
// two register sets) and then runs its own rounds, so the reads land while
// the rounds run; whether block j + 1 is there is judged from prod as read
// one block earlier (the feeder, up to kPktRing blocks ahead, has had a
// block's time to write it), so that read does not stall either.
__device__ __forceinline__ bool pkt_chain(const PassArgs &a, uint8_t *digest, bool store,
                                          const uint4 *ring, uint32_t *flags, uint32_t *own,
                                          uint32_t J, uint32_t nb, uint32_t ticks) {
    static_assert((kPktRing & (kPktRing - 1)) == 0, "ring slots by mask");
    const uint4 *lane_ring = ring + (threadIdx.x & 63);
    auto read = [&](uint32_t *w, uint32_t j) {
        const uint4 *src = lane_ring + (size_t)(j & (kPktRing - 1)) * 16 * 64;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const uint4 v = src[q * 64];
            w[4 * q] = v.x;
            w[4 * q + 1] = v.y;
            w[4 * q + 2] = v.z;
            w[4 * q + 3] = v.w;
        }
    };
    uint32_t h[4];
    md5_init(h);
    uint32_t wa[64], wb[64];
    const uint32_t end = J + nb;
    uint32_t j = J, p;  // p: prod as read in the previous block
    // Block j from cur when block j + 1 follows.  `seen` is taken first on
    // every such block: its wait also covers the reads of cur, made before
    // p's, so the rounds need no waits of their own.  seen - (j + 2) >=
    // ~(j + 2) (unsigned) holds exactly when seen < j + 2 or seen is
    // kPktAbort.  The store of cons = j + 1 depends on the rounds' result,
    // so it cannot move ahead of them (block j's inputs are then used).
#define PKT_STEP(CUR, NXT)                                                                    \
    do {                                                                                      \
        const uint32_t seen = __builtin_amdgcn_readfirstlane(p);                              \
        if (seen - (j + 2) >= ~(j + 2) && !pkt_wait(flags + kPktProd, j + 2, ticks))          \
            goto failed;                                                                      \
        asm volatile("" ::: "memory");                                                        \
        read(NXT, j + 1);                                                                     \
        p = pkt_load(flags + kPktProd);                                                       \
        md5_compress_w(h, CUR);                                                               \
        uint32_t c = j + 1;                                                                   \
        asm volatile("" : "+v"(c) : "v"(h[1]));                                               \
        pkt_store(own, c);                                                                    \
        j++;                                                                                  \
    } while (0)
    // the last block: no successor to read
#define PKT_LAST(CUR)                                                                         \
    do {                                                                                      \
        md5_compress_w(h, CUR);                                                               \
        uint32_t c = j + 1;                                                                   \
        asm volatile("" : "+v"(c) : "v"(h[1]));                                               \
        pkt_store(own, c);                                                                    \
    } while (0)
    if (!pkt_wait(flags + kPktProd, J + 1, ticks)) goto failed;
    asm volatile("" ::: "memory");
    read(wa, J);
    p = pkt_load(flags + kPktProd);
    for (;;) {
        if (j + 1 == end) {
            PKT_LAST(wa);
            break;
        }
        PKT_STEP(wa, wb);
        if (j + 1 == end) {
            PKT_LAST(wb);
            break;
        }
        PKT_STEP(wb, wa);
    }
#undef PKT_STEP
#undef PKT_LAST
    if (store) store_digest(digest, h);
    return true;
failed:
    // a feeder that gave up flagged the error itself; a stall here flags it
    if ((uint32_t)__builtin_amdgcn_readfirstlane(pkt_load(flags + kPktProd)) != kPktAbort)
        atomicOr(a.err, kPassErrWait);
    pkt_store(own, kPktAbort);
    return false;
}

// Real-time stamps of the pass (one thread per workgroup calls these): the
// first start and the last end of each role's work.
__device__ __forceinline__ unsigned long long *pass_times(const PassArgs &a) {
    return reinterpret_cast<unsigned long long *>(a.sync + kPassTimes);
}
__device__ __forceinline__ void pass_stamp_start(const PassArgs &a, int role) {
    atomicMin(pass_times(a) + 2 * role, (unsigned long long)__builtin_amdgcn_s_memrealtime());
}
__device__ __forceinline__ void pass_stamp_end(const PassArgs &a, int role) {
    atomicMax(pass_times(a) + 2 * role + 1, (unsigned long long)__builtin_amdgcn_s_memrealtime());
}

template <bool FULL>
__global__ __launch_bounds__(kMfThreads, 1) void k_create_pass(PassArgs a) {
    // the encode's descriptors and A buffers (and the slice role's CRC
    // tables); the packet role's ring of round inputs shares the space
    __shared__ union {
        MfLds mf;
        uint4 pkt[kPktRing * 16 * 64];
    } lds_u;
    MfLds &lds = lds_u.mf;
    // the A ring of mf_unit_dma: four separate objects (see there)
    __shared__ uint4 s_a0[kMfRing], s_a1[kMfRing], s_a2[kMfRing], s_a3[kMfRing];
    const uint32_t tid = threadIdx.x;
    const PassRole role = pass_role(blockIdx.x, gridDim.x, a.h2_blocks, a.h3_blocks);
    // MD5 workgroups run their hash on wave 0 (64 lanes: each lane streams its
    // own message).  In the slice-hash workgroups waves 1-3, on the other
    // SIMDs, take the slice CRCs, one slice per wave, with their tables in the
    // LDS the encode would use; the packet-hash workgroups keep their CU's
    // memory path to themselves (they must keep pace with the encode).
    if (role.role == kPassPacket) {
        // waves 0 (chain) and 1 (feeder) of pkt_chain / pkt_feed; the ring
        // takes lds, unused by this role, and the counters s_a0
        static_assert(kMfRing * 4 >= 3 * kPktLaneWords, "packet counters");
        uint4 *ring = lds_u.pkt;
        uint32_t *flags = reinterpret_cast<uint32_t *>(s_a0);
        if (tid < 2) flags[tid] = 0;
        __syncthreads();
        if (tid >= 128) return;
        // the wave index as a scalar: each wave's code is then uniform
        const uint32_t lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
        // each lane stores its wave's counter: lane 0 to the counter, the
        // others to words of their own
        uint32_t *own = lane == 0 ? flags + (wave ? kPktProd : kPktCons)
                                  : flags + (1 + wave) * kPktLaneWords + lane;
        // the chain's waits cover the feeder's, which reset on any band progress
        const uint32_t ticks = min(a.stall_ticks, 0x3FFFFFFFu) * 4u;
        const uint32_t nb = md5_blocks(36ull + a.S);
        if (tid == 0) pass_stamp_start(a, kPassPacket);
        uint32_t J = 0;
        for (uint32_t r0 = role.index * 64; r0 < a.R; r0 += a.h3_blocks * 64, J += nb) {
            const uint32_t r = min(r0 + lane, a.R - 1);
            const bool ok = wave ? pkt_feed(a, r, ring, flags, own, J, ticks)
                                 : pkt_chain(a, a.pkt_md5 + (size_t)r * 16, r0 + lane < a.R, ring,
                                             flags, own, J, nb, ticks);
            if (!ok) break;
        }
        if (tid == 0) pass_stamp_end(a, kPassPacket);
        return;
    }
    if (role.role == kPassSlice) {
        if (tid >= 64) {
            const uint32_t cw = role.index * 3 + (tid >> 6) - 1, n_cw = 3 * a.h2_blocks;
            // 16 tables (16 KiB) per CRC wave, 48 KiB in all: the descriptor
            // table and the register-ring A buffer of lds, unused until the
            // workgroup encodes
            static_assert(sizeof(MfLds) >= 3 * 16 * 256 * 4, "three waves of 16 CRC tables");
            uint32_t(*t)[256] = reinterpret_cast<uint32_t(*)[256]>(&lds) + 16 * ((tid >> 6) - 1);
            crc_tables_init_wave<16>(t);
            if (tid % 64 == 0) pass_stamp_start(a, kPassCrc);
            for (uint32_t sl = cw; sl < a.n; sl += n_cw)
                crc_slice_wave(t, a.inputs[sl], a.S, a.S % 16 == 0, a.pw, a.slice_crc + sl);
            if (tid % 64 == 0) pass_stamp_end(a, kPassCrc);
        } else {
            if (tid == 0) pass_stamp_start(a, kPassSlice);
            const uint32_t hb = role.index;
            for (uint32_t s = hb * 64 + tid; s < a.n; s += a.h2_blocks * 64) {
                // slices start 16-byte aligned when S is a multiple of 16 (file bases are)
                const uint4 *tb = a.tail_blocks + 4 * (size_t)s;
                if (a.S % 16 == 0) slice_md5_lane<true>(a.inputs[s], tb, a.S, a.slice_md5 + (size_t)s * 16);
                else slice_md5_lane<false>(a.inputs[s], tb, a.S, a.slice_md5 + (size_t)s * 16);
            }
            if (tid == 0) pass_stamp_end(a, kPassSlice);
        }
        // Its slice hashes done, the workgroup encodes what is left in its
        // XCD's queue (the CRC tables in lds.desc are not needed any more).
        __syncthreads();
    }
    if (tid == 0) pass_stamp_start(a, kPassEncode);
    mf_load_desc<FULL>(lds.desc, a.inputs, a.n, a.nq);
    __syncthreads();  // descriptor table ready
    // Units come from the per-XCD work queues (contiguous slots), band by
    // band (the queues hand out units in increasing order).
    __shared__ uint32_t s_unit;
    const uint32_t e_blocks = gridDim.x - a.h3_blocks - a.h2_blocks;
    const UnitQueue queue(a.sync, blockIdx.x, e_blocks);
    // Inside a band, units go tile by tile: tiles of kPassTileRows row blocks
    // x tc column groups, about one XCD's run of queue slots per round, so
    // the units an XCD runs at once share their A fragments (per row block)
    // and source words (per column group) in its L2.  Row-block order inside
    // a band with slots strided over the XCDs had each XCD run ~7 row blocks
    // x ~4 column groups at once: 11.5 GB of L2 fills per headline launch
    // against 8.9 GB in 4 x 8 tiles (3 x 10: 8.9 GB, 2 x 15: 9.4 GB), the
    // step the same or 0.2-0.4% shorter.
    const uint32_t tc = max(1u, (e_blocks / kPassXcds + kPassTileRows - 1) / kPassTileRows);
    // Band b holds units [row_blocks * bstart[b], row_blocks * bstart[b + 1]).
    const uint32_t *bstart = a.sync + kPassBands + a.n_bands;
    uint32_t band = 0, b_lo = 0, b_hi = a.row_blocks * bstart[1];
    // Finished units are counted per band and published once per band: when
    // this workgroup's next unit lies in another band, or there is none.
    // Publishing = every thread's stores released to the workgroup (they
    // reach L2), a barrier, then one agent-scope release (L2 write-back) and
    // one add by thread 0.
    uint32_t done = 0;  // units of band `band` finished and not yet counted
    uint32_t u = queue.next(&s_unit);
    while (u < a.n_units) {
        while (u >= b_hi) {
            band++;
            b_lo = b_hi;
            b_hi = a.row_blocks * bstart[band + 1];
        }
        const uint32_t c0 = bstart[band], w = bstart[band + 1] - c0;
        uint32_t y, x;
        band_tile_unit(u - b_lo, w, a.row_blocks, tc, &y, &x);
        x += c0;
        if (FULL)
            mf_unit_dma<false>(lds.desc, s_a0, s_a1, s_a2, s_a3, a.inputs, a.afrag, a.nq, a.nq, a.R, a.S,
                        a.rec, 0, y * kMfTiles, x * (kMfThreads * 2));
        else
            mf_unit<FULL>(lds, a.inputs, a.afrag, a.nq, a.nq, a.R, a.S, a.rec, 0, y * kMfTiles,
                          x * (kMfThreads * 2));
        done++;
        const uint32_t nxt = queue.next(&s_unit);
        if (nxt >= a.n_units || nxt >= b_hi) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __syncthreads();
            if (tid == 0) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __hip_atomic_fetch_add(a.sync + kPassBands + band, done, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                pass_stamp_end(a, kPassEncode);
            }
            done = 0;
        }
        u = nxt;
    }
}

// How long a waiting role keeps waiting without progress before it gives up
// and flags the launch: 2 s of the 100 MHz real-time counter.
// P2G_STALL_TICKS overrides it (read at each launch), so that tests can
// drive the give-up path from the host with a limit that every wait hits
// while the rest of the launch still runs to its normal end.
static uint32_t stall_ticks() {
    if (const char *v = getenv("P2G_STALL_TICKS")) return (uint32_t)strtoul(v, nullptr, 10);
    return 200000000u;
}

// Zeroes flags and counts (start stamps, the even u64 slots, get the maximum:
// they are minima over workgroups), and thread 0 writes the band starts.
__global__ void k_pass_reset(uint32_t *sync, uint32_t n_bands, uint32_t col_groups,
                             uint32_t band_cols) {
    const uint32_t words = kPassBands + n_bands;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x)
        sync[i] = (i >= kPassTimes && i < kPassQueues && ((i - kPassTimes) / 2) % 2 == 0)
                      ? 0xFFFFFFFFu
                      : 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        uint32_t *bstart = sync + kPassBands + n_bands, c = 0;
        for (uint32_t b = 0; b < n_bands; b++) {
            bstart[b] = c;
            c += next_band_width(c, col_groups - c, band_cols);
        }
        bstart[n_bands] = c;  // = col_groups
    }
}

uint32_t pass_sync_words(const PassPlan &p) { return kPassBands + 2 * p.n_bands + 1; }

hipError_t launch_create_pass(const PassPlan &p, const SliceDesc *d_inputs,
                              const uint4 *d_tail_blocks, uint32_t n,
                              const uint4 *d_afrag, uint32_t R, uint32_t S, uint8_t *d_rec,
                              uint8_t *d_slice_md5, uint32_t *d_slice_crc, const uint8_t set_id[16],
                              uint8_t *d_pkt_md5, uint32_t *d_sync, uint32_t *d_err,
                              bool all_full, hipStream_t stream) {
    const uint32_t words = kPassBands + p.n_bands;
    hipLaunchKernelGGL(k_pass_reset, dim3(std::min<uint32_t>((words + 255) / 256, 64)), dim3(256),
                       0, stream, d_sync, p.n_bands, p.col_groups, p.band_cols);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    PassArgs a;
    a.inputs = d_inputs;
    a.tail_blocks = d_tail_blocks;
    a.afrag = d_afrag;
    a.rec = d_rec;
    a.slice_md5 = d_slice_md5;
    a.slice_crc = d_slice_crc;
    a.pkt_md5 = d_pkt_md5;
    a.sync = d_sync;
    a.err = d_err;
    a.pre = make_packet_prefix(set_id);
    a.n = n;
    a.nq = mfma_steps(n);
    a.R = R;
    a.S = S;
    a.row_blocks = p.row_blocks;
    a.col_groups = p.col_groups;
    a.band_cols = p.band_cols;
    a.n_bands = p.n_bands;
    a.n_units = p.n_units;
    a.h3_blocks = p.h3_blocks;
    a.h2_blocks = p.h2_blocks;
    a.stall_ticks = stall_ticks();  // without a finished unit
    static const CrcPowers pw = make_crc_powers();
    a.pw = pw;
    if (all_full)
        hipLaunchKernelGGL(k_create_pass<true>, dim3(p.grid), dim3(kMfThreads), 0, stream, a);
    else
        hipLaunchKernelGGL(k_create_pass<false>, dim3(p.grid), dim3(kMfThreads), 0, stream, a);
    return hipGetLastError();
}

// ======================================================================
// Batched create pass: many jobs in one persistent launch
// ======================================================================
//
// The roles of k_create_pass over a table of pieces (a job's recovery rows,
// or a 16-row-aligned part of them) and a table of slice jobs.  Units of all
// pieces form one list in piece order, band by band inside a piece; encode
// workgroups take them in a fixed stride and reload the descriptor table
// from the piece when it changes, in K chunks of kMfmaChunkInputs inputs
// (the first chunk stores, later ones XOR in).  Packet hashing runs as wave
// tasks of up to 64 packets of one piece, so a task's band waits are wave
// uniform; slice lanes (MD5 and CRC-32 of a slice from one read) walk the
// slice jobs.  Every role
// raises its job's completion stamp (atomicMax on a u64 real-time count),
// which gives per-job latency inside the batch.

// MD5 and CRC-32 of one zero-padded S-byte slice by one lane, from one read
// of its bytes: the CRC takes each message block's words as the MD5 ring
// hands them over, sixteen bytes per dependent step, then the words of the
// final partial block below S.  t: this wave's 16 CRC tables
// (crc_tables_init_wave<16>).
template <bool A16>
__device__ __forceinline__ void slice_md5_crc_lane(const SliceDesc d, const uint4 *tb, uint32_t S,
                                                   const uint32_t (*t)[256], uint8_t *md5_out,
                                                   uint32_t *crc_out) {
    using W = typename std::conditional<A16, uint4, uint32_t>::type;
    const uint32_t nb = md5_blocks(S);
    const uint32_t nmsg = S / 64;
    const uint32_t nfull = d.valid / 64;
    uint32_t h[4];
    md5_init(h);
    uint32_t c = 0xFFFFFFFFu;
    md5_slice_blocks<W>(h, d.ptr, nfull, tb, nmsg, [&](const uint32_t *x) {
#pragma unroll
        for (int w = 0; w < 16; w += 4) c = crc_step16(t, c, x[w], x[w + 1], x[w + 2], x[w + 3]);
    });
    for (uint32_t b = nmsg; b < nb; b++) {
        const uint32_t base = b * 64u;
        uint32_t x[16];
#pragma unroll
        for (int w = 0; w < 16; w++) {
            const uint32_t o = base + 4u * w;
            if (o < S) {
                x[w] = load_word(d.ptr, o, d.valid);
                c = crc_step4(t, c, x[w]);
            } else {
                x[w] = md5_pad_word(o, S, w, b, nb);
            }
        }
        md5_compress(h, x);
    }
    store_digest(md5_out, h);
    *crc_out = c ^ 0xFFFFFFFFu;
}

struct BatchArgs {
    const BatchPiece *pieces;
    const BatchSlices *sjobs;
    const SliceDesc *descs;      // whole batch, indexed by global slice index
    const uint4 *tail_blocks;    // 4 per global slice
    uint32_t *sync;
    uint32_t *err;  // sticky error word of the plan: kept across launches until read
    unsigned long long *job_done;
    uint32_t n_pieces, n_sjobs, slice_lo, slice_hi;
    uint32_t n_units, n_tasks, h3_blocks, h2_blocks;
    uint32_t n_jobs, stall_ticks;
};

__device__ __forceinline__ void batch_stamp(unsigned long long *slot) {
    atomicMax(slot, (unsigned long long)__builtin_amdgcn_s_memrealtime());
}

// First start / last end of each role's work in the group's sync words
// (kPassTimes, as in the fused pass).
__device__ __forceinline__ void batch_role(uint32_t *sync, int role, bool end) {
    unsigned long long *t = reinterpret_cast<unsigned long long *>(sync + kPassTimes) + 2 * role;
    const unsigned long long now = __builtin_amdgcn_s_memrealtime();
    if (end) atomicMax(t + 1, now);
    else atomicMin(t, now);
}

// pass_wait, but the stall clock also restarts whenever any unit of the
// launch is published (progress), so a task queued behind a long batch
// waits as long as the encode keeps moving.
__device__ __forceinline__ bool batch_wait(const uint32_t *cnt, uint32_t target,
                                           const uint32_t *progress, uint32_t stall_ticks,
                                           uint32_t *flags) {
    uint32_t seen = __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    uint32_t prog = __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(progress, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    uint64_t since = __builtin_amdgcn_s_memrealtime();
    while (seen < target) {
        __builtin_amdgcn_s_sleep(16);
        const uint32_t now = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        const uint32_t pnow = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(progress, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        const uint64_t t = __builtin_amdgcn_s_memrealtime();
        if (now != seen || pnow != prog) {
            seen = now;
            prog = pnow;
            since = t;
        } else if (t - since > stall_ticks) {
            atomicOr(flags, kPassErrWait);
            return false;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

// MD5 blocks of a packet body whose data bytes lie in [0, e): block k holds
// data bytes up to 64 k + 28.
__device__ __forceinline__ uint32_t packet_blocks_within(uint64_t e) {
    return e >= 28 ? (uint32_t)((e - 28) / 64) + 1 : 0u;
}

// Packet r of piece P (exponent P.first_exp + r), band by band as in
// pass_packet_walk, on one wave.  A lane with store = false hashes along (its wave waits
// together) and writes nothing.  A piece of column groups [cg0, cg0 +
// col_groups) continues the row's hash from P.md5_state when cg0 > 0 (the
// earlier pieces hashed every block within column group cg0's start) and,
// when it ends before the row does, saves the state there instead of the
// digest.
__device__ __forceinline__ void batch_packet_lane(const BatchArgs &a, const BatchPiece &P,
                                                  uint32_t r, bool store) {
    const uint32_t S = P.S;
    const uint8_t *data = P.rec + (size_t)r * S;
    const bool a16 = S % 16 == 0 && reinterpret_cast<uintptr_t>(P.rec) % 16 == 0;
    const uint64_t len = 36ull + S;
    const uint32_t nb = md5_blocks(len), nfull = (uint32_t)(len / 64);
    const uint32_t bulk_end = nfull > 1 ? nfull : 1;
    const uint32_t exp = P.first_exp + r;
    const bool last = P.cg0 + P.col_groups >= P.cg_total;
    PacketPrefix pre;
#pragma unroll
    for (int w = 0; w < 8; w++) pre.w[w] = P.pre[w];
    uint32_t h[4];
    uint32_t k = 0;
    if (P.cg0 == 0) {
        md5_init(h);
    } else {
        const uint4 st = P.md5_state[r];
        h[0] = st.x; h[1] = st.y; h[2] = st.z; h[3] = st.w;
        k = min(packet_blocks_within((uint64_t)P.cg0 * (kMfThreads * 2)), bulk_end);
    }
    const uint32_t *bstart = a.sync + P.bstart0;
    for (uint32_t b = 0; b < P.n_bands; b++) {
        const uint32_t c1 = bstart[b + 1], w = c1 - bstart[b];
        if (!batch_wait(a.sync + P.band0 + b, w * P.row_blocks, a.sync + kBatchProgress,
                        a.stall_ticks, a.err))
            return;
        const uint64_t e = min((uint64_t)(P.cg0 + c1) * (kMfThreads * 2), (uint64_t)S);
        const uint32_t avail = b + 1 == P.n_bands && last ? nb : packet_blocks_within(e);
        if (k == 0 && avail > 0) {
            md5_packet_block(h, pre, exp, data, len, 0, nb);
            k = 1;
        }
        const uint32_t stop = min(avail, bulk_end);
        if (stop > k) {
            if (a16) md5_packet_bulk16(h, data, k, stop - k);
            else md5_bulk<uint32_t>(h, data + 64 * (size_t)k - 36, stop - k);
            k = stop;
        }
    }
    if (!last) {
        if (store) P.md5_state[r] = make_uint4(h[0], h[1], h[2], h[3]);
        return;
    }
    for (; k < nb; k++) md5_packet_block(h, pre, exp, data, len, k, nb);
    if (store) store_digest(P.pkt_md5 + (size_t)r * 16, h);
}

// LDS of the batched kernel beside the A ring: one K chunk's descriptors
// (the slice role keeps its CRC tables there instead).
struct BatchLds {
    uint4 desc[kMfmaChunkInputs];
};

__global__ __launch_bounds__(kMfThreads, 1) void k_create_batch(BatchArgs a) {
    __shared__ BatchLds lds;
    __shared__ uint4 s_a0[kMfRing], s_a1[kMfRing], s_a2[kMfRing], s_a3[kMfRing];
    const uint32_t blk = blockIdx.x, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const PassRole role = pass_role(blk, gridDim.x, a.h2_blocks, a.h3_blocks);
    if (tid == 0) {
        unsigned long long *t0 = a.job_done + a.n_jobs;
        atomicMin(t0, (unsigned long long)__builtin_amdgcn_s_memrealtime());
    }
    if (role.role == kPassPacket) {
        // packet tasks: wave task t = 64 packets of one piece
        if (lane == 0) batch_role(a.sync, kPassPacket, false);
        uint32_t pi = 0;
        for (uint32_t t = role.index * 4 + wave; t < a.n_tasks; t += a.h3_blocks * 4) {
            while (t >= a.pieces[pi].task_end) pi++;
            const BatchPiece &P = a.pieces[pi];
            const uint32_t r = (t - P.task0) * 64 + lane;
            batch_packet_lane(a, P, min(r, P.R - 1), r < P.R);
            if (lane == 0) batch_stamp(a.job_done + P.job);
        }
        if (lane == 0) batch_role(a.sync, kPassPacket, true);
        return;
    }
    if (role.role == kPassSlice && wave < kBatchSliceWaves) {
        // slice MD5 + CRC-32 lanes on waves 0 .. kBatchSliceWaves - 1 (CRC
        // tables per wave in lds.desc).  A wave takes tasks of 64 consecutive
        // slices (one per lane) from the group's queue; the table runs from
        // the largest slices to the smallest, so the waves end within about
        // one small slice of each other.
        static_assert(sizeof(BatchLds::desc) >= kBatchSliceWaves * 16 * 256 * 4,
                      "16 CRC tables per slice wave");
        uint32_t(*t)[256] = reinterpret_cast<uint32_t(*)[256]>(lds.desc) + 16 * wave;
        crc_tables_init_wave<16>(t);
        const uint32_t n_tasks = (a.slice_hi - a.slice_lo + 63) / 64;
        if (lane == 0) batch_role(a.sync, kPassSlice, false);
        uint32_t sj = 0;
        for (;;) {
            uint32_t task = 0;
            if (lane == 0)
                task = __hip_atomic_fetch_add(a.sync + kBatchSliceQueue, 1u, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
            task = __builtin_amdgcn_readfirstlane(task);
            if (task >= n_tasks) break;
            const uint32_t sl = a.slice_lo + task * 64 + lane;
            if (sl >= a.slice_hi) continue;
            while (sl >= a.sjobs[sj].slice_end) sj++;
            const BatchSlices J = a.sjobs[sj];
            if (sl < J.slice0) continue;  // descriptors of an item without slice checksums
            const uint4 *tb = a.tail_blocks + 4 * (size_t)sl;
            uint8_t *mo = J.md5 + (size_t)(sl - J.slice0) * 16;
            uint32_t *co = J.crc + (sl - J.slice0);
            if (J.S % 16 == 0) slice_md5_crc_lane<true>(a.descs[sl], tb, J.S, t, mo, co);
            else slice_md5_crc_lane<false>(a.descs[sl], tb, J.S, t, mo, co);
            batch_stamp(a.job_done + J.job);
        }
        if (lane == 0) batch_role(a.sync, kPassSlice, true);
    }
    // A slice workgroup whose hashes are done encodes what is left in its
    // XCD's queue, as the encode workgroups do (once all its slice waves are
    // done).
    if (role.role == kPassSlice) __syncthreads();

    // Encode.
    const uint32_t e_blocks = gridDim.x - a.h3_blocks - a.h2_blocks;
    if (tid == 0) batch_role(a.sync, kPassEncode, false);
    // Units come from the per-XCD work queues (UnitQueue, contiguous slots):
    // the units of one round that fall to an XCD (one row block's column
    // groups) run on that XCD and share its L2 copy of the row block's A
    // fragments, and inside
    // an XCD a workgroup takes the next unit when it is free, so units of
    // different jobs' sizes even out.  Every value the loop branches on is a
    // scalar (the unit from the queue, the tables), so the barriers inside
    // are reached by the whole workgroup.
    __shared__ uint32_t s_unit;
    const UnitQueue queue(a.sync, blk, e_blocks);
    uint32_t pi = 0;
    const SliceDesc *loaded = nullptr;  // first descriptor of the chunk in lds.desc
    uint32_t band = 0, b_lo = 0, b_hi = 0, unit_end = 0;
    uint32_t done = 0;
    uint32_t u = queue.next(&s_unit);
    while (u < a.n_units) {
        if (u >= unit_end) {
            while (u >= a.pieces[pi].unit_end) pi++;
            const BatchPiece &P = a.pieces[pi];
            unit_end = P.unit_end;
            band = 0;
            b_lo = P.unit0;
            b_hi = P.unit0 + P.row_blocks * a.sync[P.bstart0 + 1];
        }
        const BatchPiece &P = a.pieces[pi];
        const uint32_t *bstart = a.sync + P.bstart0;
        while (u >= b_hi) {
            band++;
            b_lo = b_hi;
            b_hi = P.unit0 + P.row_blocks * bstart[band + 1];
        }
        const uint32_t c0 = bstart[band], w = bstart[band + 1] - c0;
        const uint32_t v = u - b_lo, y = v / w, x = P.cg0 + c0 + v % w;
        const uint4 *afrag = static_cast<const uint4 *>(P.afrag);
        for (uint32_t k0 = 0; k0 < P.n; k0 += kMfmaChunkInputs) {
            const SliceDesc *cin = P.inputs + k0;
            const uint32_t cn = min(P.n - k0, kMfmaChunkInputs);
            const uint32_t cq = min(P.nq - k0 / 4, kMfmaChunkInputs / 4);
            if (cin != loaded) {
                __syncthreads();  // every unit is done reading the old table
                mf_load_desc<true>(lds.desc, cin, cn, cq);
                __syncthreads();
                loaded = cin;
            }
            mf_unit_dma<true>(lds.desc, s_a0, s_a1, s_a2, s_a3, cin, afrag + (size_t)(k0 / 4) * 64, cq, P.nq, P.R, P.S, P.rec,
                          k0 > 0, y * kMfTiles, x * (kMfThreads * 2));
        }
        done++;
        const uint32_t nxt = queue.next(&s_unit);
        if (nxt >= a.n_units || nxt >= b_hi) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __syncthreads();
            if (tid == 0) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __hip_atomic_fetch_add(a.sync + P.band0 + band, done, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(a.sync + kBatchProgress, done, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                batch_role(a.sync, kPassEncode, true);
            }
            done = 0;
        }
        u = nxt;
    }
}

// Zeroes the group's flags, progress, stamps and band counters, writes each
// piece's band starts (thread per piece) and, for the first group of a
// launch, clears the job completion stamps (the batch start slot gets the
// maximum: it is a minimum over workgroups).
__global__ void k_batch_reset(uint32_t *sync, uint32_t words, const BatchPiece *__restrict__ pieces,
                              uint32_t n_pieces, unsigned long long *job_done, uint32_t n_jobs,
                              int first) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, gsz = gridDim.x * blockDim.x;
    for (uint32_t i = gid; i < words; i += gsz)
        sync[i] = (i >= kPassTimes && i < kPassQueues && ((i - kPassTimes) / 2) % 2 == 0) ? ~0u : 0u;
    if (first)
        for (uint32_t j = gid; j <= n_jobs; j += gsz) job_done[j] = j == n_jobs ? ~0ull : 0ull;
    for (uint32_t p = gid; p < n_pieces; p += gsz) {
        const BatchPiece &P = pieces[p];
        uint32_t *bs = sync + P.bstart0, c = 0;
        for (uint32_t b = 0; b < P.n_bands; b++) {
            bs[b] = c;
            c += batch_band_width(c, P.col_groups - c, P.band_cols);
        }
        bs[P.n_bands] = c;
    }
}

// Fragments of every piece of a group, grid-stride per piece.
__global__ __launch_bounds__(256) void k_batch_afrag(const uint16_t *__restrict__ antilog,
                                                     const uint16_t *__restrict__ logbase,
                                                     const BatchPiece *__restrict__ pieces,
                                                     uint32_t n_pieces) {
    AfragLane L;
    L.load(threadIdx.x & 63);
    for (uint32_t p = 0; p < n_pieces; p++) {
        const BatchPiece &P = pieces[p];
        afrag_fill(L, antilog, logbase, P.n, P.R, P.first_exp, P.nq, mfma_row_tiles(P.R),
                   static_cast<uint4 *>(const_cast<void *>(P.afrag)));
    }
}

// Copies of partial slices: blockIdx.y = part, x strides its words.
__global__ __launch_bounds__(256) void k_stage_parts(const StagePart *__restrict__ parts) {
    const StagePart d = parts[blockIdx.y];
    const uint32_t words = d.S / 4;
    for (uint32_t w = blockIdx.x * 256 + threadIdx.x; w < words; w += gridDim.x * 256)
        reinterpret_cast<uint32_t *>(d.dst)[w] = load_word(d.src, 4 * w, d.valid);
}

hipError_t launch_stage_parts(const StagePart *d_parts, uint32_t nparts, uint32_t max_S,
                              hipStream_t stream) {
    for (uint32_t p0 = 0; p0 < nparts; p0 += 65535) {
        const uint32_t np = std::min<uint32_t>(nparts - p0, 65535);
        const uint32_t bx = std::min<uint32_t>((max_S / 4 + 255) / 256, 64);
        hipLaunchKernelGGL(k_stage_parts, dim3(bx, np), dim3(256), 0, stream, d_parts + p0);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_batch_afrag(const uint16_t *d_antilog, const uint16_t *d_logbase,
                              const BatchGroupLaunch &g, hipStream_t stream) {
    if (g.n_pieces == 0 || g.afrag_entries == 0) return hipSuccess;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((g.afrag_entries + 255) / 256, 16384);
    hipLaunchKernelGGL(k_batch_afrag, dim3(blocks), dim3(256), 0, stream, d_antilog, d_logbase,
                       g.pieces, g.n_pieces);
    return hipGetLastError();
}

int pass_occupancy() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_create_pass<true>, kMfThreads, 0) != hipSuccess)
        return 0;
    int nb2 = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb2, k_create_pass<false>, kMfThreads, 0) != hipSuccess)
        return 0;
    return std::min(nb, nb2);
}

int batch_occupancy() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_create_batch, kMfThreads, 0) != hipSuccess)
        return 0;
    return nb;
}

hipError_t launch_create_batch(const BatchGroupLaunch &g, const SliceDesc *d_descs,
                               const uint4 *d_tail_blocks, hipStream_t stream) {
    hipLaunchKernelGGL(k_batch_reset, dim3(64), dim3(256), 0, stream, g.sync, g.sync_words,
                       g.pieces, g.n_pieces, g.job_done, g.n_jobs, g.first ? 1 : 0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    BatchArgs a;
    a.pieces = g.pieces;
    a.sjobs = g.sjobs;
    a.descs = d_descs;
    a.tail_blocks = d_tail_blocks;
    a.sync = g.sync;
    a.err = g.err;
    a.job_done = g.job_done;
    a.n_pieces = g.n_pieces;
    a.n_sjobs = g.n_sjobs;
    a.slice_lo = g.slice_lo;
    a.slice_hi = g.slice_hi;
    a.n_units = g.n_units;
    a.n_tasks = g.n_tasks;
    a.h3_blocks = g.h3_blocks;
    a.h2_blocks = g.h2_blocks;
    a.n_jobs = g.n_jobs;
    a.stall_ticks = stall_ticks();  // without any unit published
    hipLaunchKernelGGL(k_create_batch, dim3(g.grid), dim3(kMfThreads), 0, stream, a);
    return hipGetLastError();
}

// ======================================================================
// Synthetic input: the xorshift64 stream on the device
// ======================================================================
//
// Word k of the stream is the state after k + 1 steps (x ^= x << 13;
// x ^= x >> 7; x ^= x << 17) from the seed, stored little-endian.  The step
// is linear over GF(2), so lane j starts from step^(j W) (seed), built from
// the jumps step^(W 2^b) for the set bits b of j, and writes words
// [j W, (j + 1) W).  A jump is applied as 8 byte-table lookups (the XOR of
// each state byte's image): one select-and-XOR per state bit had made the
// jumps ~87% of a fill wave's instructions, and a serving run of seeded
// requests spent 7% of its device time in fills.

__device__ __forceinline__ uint64_t xs_apply(const uint64_t *__restrict__ tables, uint64_t v) {
    uint64_t out = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) out ^= tables[256 * q + ((v >> (8 * q)) & 0xFFu)];
    return out;
}

// Lane j of one file: words [j W, (j + 1) W) of the stream from `seed`.
__device__ __forceinline__ void xs_fill_lane(uint8_t *__restrict__ dst, uint64_t nbytes,
                                             uint64_t seed, const uint64_t *__restrict__ jumps,
                                             uint32_t j) {
    uint64_t s = seed;
    for (int b = 0; b < 32; b++)
        if ((j >> b) & 1u) s = xs_apply(jumps + (size_t)kXsJumpTableWords * b, s);
    const uint64_t w0 = (uint64_t)j * kXorshiftLaneWords;
    const uint64_t words = nbytes / 8;
    uint64_t *out = reinterpret_cast<uint64_t *>(dst);
    for (uint32_t k = 0; k < kXorshiftLaneWords; k += 2) {
        uint64_t a, c;
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        a = s;
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        c = s;
        const uint64_t w = w0 + k;
        if (w + 2 <= words) {
            *reinterpret_cast<ulonglong2 *>(out + w) = make_ulonglong2(a, c);
        } else {
            // the end of the buffer: whole words, then the bytes of a partial one
            const uint64_t v[2] = {a, c};
            for (int q = 0; q < 2; q++) {
                const uint64_t o = 8 * (w + q);
                for (int t = 0; t < 8; t++)
                    if (o + t < nbytes) dst[o + t] = (uint8_t)(v[q] >> (8 * t));
            }
            return;
        }
    }
}

__global__ __launch_bounds__(256) void k_xorshift_fill(uint8_t *__restrict__ dst, uint64_t nbytes,
                                                       uint64_t seed,
                                                       const uint64_t *__restrict__ jumps,
                                                       uint32_t lanes) {
    const uint32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j < lanes) xs_fill_lane(dst, nbytes, seed, jumps, j);
}

hipError_t launch_xorshift_fill(uint8_t *d_dst, uint64_t nbytes, uint64_t seed,
                                const uint64_t *d_jumps, hipStream_t stream) {
    const uint64_t words = (nbytes + 7) / 8;
    const uint64_t lanes = (words + kXorshiftLaneWords - 1) / kXorshiftLaneWords;
    if (lanes == 0) return hipSuccess;
    if (lanes > 0xFFFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_xorshift_fill, dim3((uint32_t)((lanes + 255) / 256)), dim3(256), 0, stream,
                       d_dst, nbytes, seed, d_jumps, (uint32_t)lanes);
    return hipGetLastError();
}

// Several files in one launch: the files' lanes are numbered one after the
// other (file f from lane0[f]), so short files fill side by side instead of
// one small launch each, whose serial jump lookups set its time.
__global__ __launch_bounds__(256) void k_xorshift_fill_files(XsFillFiles f,
                                                             const uint64_t *__restrict__ jumps) {
    const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= f.lane0[f.n]) return;
    uint32_t lo = 0, hi = f.n;  // the file with lane0[k] <= g < lane0[k + 1]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (f.lane0[mid] <= g) lo = mid;
        else hi = mid;
    }
    xs_fill_lane(f.dst[lo], f.nbytes[lo], f.seed[lo], jumps, (uint32_t)(g - f.lane0[lo]));
}

hipError_t launch_xorshift_fill_files(const XsFillFiles &f, const uint64_t *d_jumps,
                                      hipStream_t stream) {
    const uint64_t lanes = f.lane0[f.n];
    if (lanes == 0) return hipSuccess;
    hipLaunchKernelGGL(k_xorshift_fill_files, dim3((uint32_t)((lanes + 255) / 256)), dim3(256), 0,
                       stream, f, d_jumps);
    return hipGetLastError();
}

}  // namespace p2g
