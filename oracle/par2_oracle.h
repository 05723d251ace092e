/*
 * par2_oracle.h -- CPU restatement of javi11/par2rust's PAR2 create path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the GPU
 * product (libpar2gpu.so) is compared against.  Only tests/, the smoke()
 * entry point and bench.py's CPU-baseline leg may load it.  Nothing in the
 * product links it, calls it or falls back to it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the par2rust repository root).
 *
 * Parity pins (what fixes each function independently of this file):
 *   or_gf_*            closed form of GF(2^16)/0x1100B, textbook carry-less
 *                      multiply in tests, invariants of src/galois.rs:101-183
 *   or_input_bases     src/reedsolomon.rs:159-200 known answers, phi(65535)
 *   or_gf_mul_xor_*    brute force vs carry-less multiply; 0/1 cases and the
 *                      0x1234 example of src/reedsolomon.rs:222-259
 *   or_md5             RFC 1321 vectors, Python hashlib, vendored md5_ref.c
 *   or_crc32           CRC-32 check value 0xCBF43926, Python zlib.crc32
 *   packets / splits   layout tests of src/format.rs, src/packet/ modules and
 *                      src/creator.rs:1347-1615 restated in tests/
 *   or_encode_simd     equal to or_encode (scalar) on every tested input
 */
#ifndef PAR2_ORACLE_H
#define PAR2_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes: one per Par2Error variant (src/error.rs:7-41). */
enum {
    OR_OK = 0,
    OR_ERR_IO = 1,
    OR_ERR_INVALID_SLICE_SIZE = 2,
    OR_ERR_INVALID_FILE_NAME = 3,
    OR_ERR_NO_INPUT_FILES = 4,
    OR_ERR_EMPTY_FILE = 5,
    OR_ERR_TOO_MANY_FILES = 6,
    OR_ERR_TOO_MANY_RECOVERY_BLOCKS = 7,
    OR_ERR_TOO_MANY_INPUT_BLOCKS = 8,
    OR_ERR_INVALID_VOLUME_SCHEME = 9,
};

/* ---- GF(2^16), src/galois.rs ---------------------------------------- */
void or_gf_tables(uint16_t *log_out, uint16_t *antilog_out);
uint16_t or_gf_mul(uint16_t a, uint16_t b);
uint16_t or_gf_pow(uint16_t base, uint32_t exp);
uint16_t or_gf_inv(uint16_t a);

/* ---- Reed-Solomon, src/reedsolomon.rs -------------------------------- */
int or_input_bases(uint32_t count, uint16_t *out);
uint16_t or_coefficient(const uint16_t *bases, uint32_t recovery_exp, uint32_t input_idx);
void or_gf_mul_xor_scalar(uint16_t coeff, const uint8_t *in, uint8_t *out, size_t len);

/* ---- Hashes: src/md5_impl.rs (RFC 1321), crc32fast (CRC-32/ISO-HDLC) -- */
typedef struct {
    uint32_t h[4];
    uint64_t total;
    uint8_t buf[64];
    uint32_t fill;
} or_md5_ctx;
void or_md5_init(or_md5_ctx *c);
void or_md5_update(or_md5_ctx *c, const uint8_t *data, size_t len);
void or_md5_final(or_md5_ctx *c, uint8_t out[16]);
void or_md5(const uint8_t *data, size_t len, uint8_t out[16]);
uint32_t or_crc32(const uint8_t *data, size_t len);

/* ---- Scan, src/source.rs:145-309 ------------------------------------ */
uint64_t or_slice_count(uint64_t len, uint64_t slice_size);
void or_slice_checksums(const uint8_t *data, uint64_t len, uint64_t slice_size,
                        uint8_t *md5_out, uint32_t *crc_out);
void or_file_hashes(const uint8_t *data, uint64_t len, uint8_t hash_full[16],
                    uint8_t hash16k[16]);
void or_file_id(const uint8_t hash16k[16], uint64_t len, const uint8_t *name,
                size_t name_len, uint8_t out[16]);

/* ---- Encode, src/creator.rs:883-1018 -------------------------------- */
/* files[f] holds lens[f] bytes; slices are taken in file order then slice
 * order, the last slice of each file zero padded.  rec_out receives R
 * blocks of slice_size bytes, exponent-major. */
int or_encode(const uint8_t *const *files, const uint64_t *lens, uint32_t nfiles,
              uint64_t slice_size, uint32_t recovery_count, uint8_t *rec_out);
/* Blocks of exponents first_exp .. first_exp + count - 1 only. */
int or_encode_range(const uint8_t *const *files, const uint64_t *lens, uint32_t nfiles,
                    uint64_t slice_size, uint32_t first_exp, uint32_t count, uint8_t *rec_out);
/* The reference's SIMD CPU path (GFNI-AVX512 > GFNI > AVX2 > table) with the
 * recovery buffers split in ceil(R/T) chunks over T threads and a barrier per
 * input slice.  force_dispatch < 0 picks by CPU features (or_simd_detect);
 * or_simd_dispatch_name names a dispatch.  Returns OR_OK. */
int or_encode_simd(const uint8_t *const *files, const uint64_t *lens, uint32_t nfiles,
                   uint64_t slice_size, uint32_t recovery_count, uint8_t *rec_out,
                   uint32_t threads, int force_dispatch);
int or_simd_detect(void);
/* Threaded forms of the scan's per-slice checksums and of the recovery
 * packet hashes, for the timed CPU baseline (threads = 0: all cores). */
void or_slice_checksums_mt(const uint8_t *data, uint64_t len, uint64_t slice_size,
                           uint8_t *md5_out, uint32_t *crc_out, uint32_t threads);
void or_recovery_md5_mt(const uint8_t *rec, uint32_t count, uint64_t slice_size,
                        uint32_t first_exp, const uint8_t set_id[16], uint8_t *md5_out,
                        uint32_t threads);
const char *or_simd_dispatch_name(int d);

/* ---- Packets, src/format.rs and src/packet/ modules ---------------------- */
size_t or_build_packet(const uint8_t set_id[16], const uint8_t type_tag[16],
                       const uint8_t *body, size_t body_len, uint8_t *out);
size_t or_main_packet(uint64_t slice_size, const uint8_t *file_ids, uint32_t n,
                      uint8_t *out, uint8_t set_id_out[16]);
size_t or_recovery_packet(const uint8_t set_id[16], uint16_t exponent,
                          const uint8_t *slice, size_t slice_len, uint8_t *out);
size_t or_creator_packet(const uint8_t set_id[16], uint8_t *out);

/* ---- Volume schemes, src/creator.rs:110-139, 325-431 ----------------- */
enum { OR_SCHEME_SINGLE = 0, OR_SCHEME_EXPONENTIAL = 1, OR_SCHEME_UNIFORM = 2,
       OR_SCHEME_EXPLICIT = 3 };
typedef struct {
    int kind;                 /* OR_SCHEME_* */
    uint32_t uniform_count;   /* Uniform { count } */
    int limited;              /* wrap in Limited { max_blocks_per_volume, inner } */
    uint32_t limit_cap;
    const uint32_t *explicit_sizes;
    uint32_t n_explicit;
    const uint32_t *outer_caps; /* further Limited wrappers around that one, innermost first */
    uint32_t n_outer_caps;
} or_scheme;
/* Returns the number of volumes written into sizes_out (capacity cap), or a
 * negative OR_ERR_* code. */
int or_resolve_volume_sizes(const or_scheme *s, uint32_t total, uint32_t *sizes_out,
                            uint32_t cap);

/* ---- run_create, src/creator.rs:182-321 + SourceFile::scan ----------- */
/* Scans every input path (display names given separately), then writes the
 * index and the volume files next to `output`.  The names of the files
 * written go, NUL separated, into names_out. */
int or_run_create(const char *output, uint64_t slice_size, uint32_t recovery_count,
                  const or_scheme *scheme, uint32_t nfiles, const char *const *paths,
                  const uint8_t *const *names, const uint32_t *name_lens,
                  char *names_out, size_t names_cap, char *msg, size_t msg_cap);

/* ---- Synthetic data: xorshift64 (13, 7, 17), 8 little-endian bytes per
 * step.  Not part of the reference; the workload generator of the tests. */
uint64_t or_xorshift64_fill(uint8_t *buf, size_t len, uint64_t state);

#ifdef __cplusplus
}
#endif
#endif
