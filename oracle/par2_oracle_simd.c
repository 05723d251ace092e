/*
 * par2_oracle_simd.c -- restatement of the reference's SIMD CPU encode path,
 * used as the timed CPU baseline and checked against the scalar or_encode.
 *
 * TEST INFRASTRUCTURE ONLY (see par2_oracle.h).
 *
 * Follows src/creator.rs:883-1018 (per input slice: one coefficient table
 * per recovery row, recovery buffers split in ceil(R/T) chunks over T
 * workers, a barrier per slice) and src/galois_simd.rs:605-636 for the
 * dispatch order GFNI-AVX512 > GFNI > AVX2 > table-scalar.  The GFNI form
 * uses four 8x8 GF(2) matrices per coefficient exactly as
 * src/galois_simd/gfni.rs:78-118 defines them; the byte shuffling around
 * the affine instructions is this file's own (byte-swap + masked affine).
 */
#define _GNU_SOURCE
#include "par2_oracle.h"

#include <immintrin.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

enum { D_TABLE = 0, D_AVX2 = 1, D_GFNI = 2, D_GFNI_AVX512 = 3 };

const char *or_simd_dispatch_name(int d) {
    switch (d) {
    case D_GFNI_AVX512: return "gfni-avx512";
    case D_GFNI: return "gfni-sse";
    case D_AVX2: return "avx2-nibble";
    default: return "table-scalar";
    }
}

int or_simd_detect(void) {
    __builtin_cpu_init();
    if (__builtin_cpu_supports("gfni") && __builtin_cpu_supports("avx512f") &&
        __builtin_cpu_supports("avx512bw"))
        return D_GFNI_AVX512;
    if (__builtin_cpu_supports("gfni") && __builtin_cpu_supports("ssse3")) return D_GFNI;
    if (__builtin_cpu_supports("avx2")) return D_AVX2;
    return D_TABLE;
}

/* Per-coefficient tables for every dispatch. */
typedef struct {
    uint16_t coeff;
    uint64_t m_ll, m_lh, m_hl, m_hh;  /* GFNI matrices, gfni.rs:78-118 */
    uint8_t nib[8][16];               /* AVX2 nibble tables */
    uint16_t lo[256], hi[256];        /* CoeffTables, galois_simd.rs:45-53 */
} coeff_tables;

static void build_coeff_tables(coeff_tables *t, uint16_t c, int dispatch) {
    t->coeff = c;
    if (c <= 1) return; /* Trivial, galois_simd.rs:684-686 */
    uint16_t cols[16];
    for (int j = 0; j < 16; j++) cols[j] = or_gf_mul(c, (uint16_t)(1u << j));
    if (dispatch >= D_GFNI) {
        /* row i of a block (output bit out_base+i) sits in byte 7-i; bit ip of
         * the row is output bit out_base+i of column in_base+ip. */
        for (int blk = 0; blk < 4; blk++) {
            int out_base = (blk >> 1) * 8, in_base = (blk & 1) * 8;
            uint64_t acc = 0;
            for (int i = 0; i < 8; i++) {
                uint8_t row = 0;
                for (int ip = 0; ip < 8; ip++)
                    row |= (uint8_t)(((cols[in_base + ip] >> (out_base + i)) & 1u) << ip);
                acc |= (uint64_t)row << ((7 - i) * 8);
            }
            if (blk == 0) t->m_ll = acc;
            else if (blk == 1) t->m_lh = acc;
            else if (blk == 2) t->m_hl = acc;
            else t->m_hh = acc;
        }
        return;
    }
    for (int s = 0; s < 256; s++) {
        uint16_t l = 0, h = 0;
        for (int b = 0; b < 8; b++) {
            if (s >> b & 1) { l ^= cols[b]; h ^= cols[8 + b]; }
        }
        t->lo[s] = l;
        t->hi[s] = h;
    }
    if (dispatch == D_AVX2) {
        /* nib[0..3]: product bytes for the low/high nibble of the symbol's
         * low byte; nib[4..7] the same for the symbol's high byte.  Even
         * index = product low byte, odd = product high byte. */
        for (int n = 0; n < 16; n++) {
            uint16_t a = t->lo[n], b = t->lo[n << 4], c2 = t->hi[n], d = t->hi[n << 4];
            t->nib[0][n] = (uint8_t)a; t->nib[1][n] = (uint8_t)(a >> 8);
            t->nib[2][n] = (uint8_t)b; t->nib[3][n] = (uint8_t)(b >> 8);
            t->nib[4][n] = (uint8_t)c2; t->nib[5][n] = (uint8_t)(c2 >> 8);
            t->nib[6][n] = (uint8_t)d; t->nib[7][n] = (uint8_t)(d >> 8);
        }
    }
}

static void mul_xor_table(const coeff_tables *t, const uint8_t *in, uint8_t *out, size_t len) {
    for (size_t k = 0; k + 1 < len; k += 2) {
        uint16_t p = t->lo[in[k]] ^ t->hi[in[k + 1]];
        out[k] ^= (uint8_t)p;
        out[k + 1] ^= (uint8_t)(p >> 8);
    }
}

static void mul_xor_trivial(uint16_t c, const uint8_t *in, uint8_t *out, size_t len) {
    if (c == 0) return;
    size_t k = 0;
    for (; k + 8 <= len; k += 8) {
        uint64_t a, b;
        memcpy(&a, in + k, 8);
        memcpy(&b, out + k, 8);
        b ^= a;
        memcpy(out + k, &b, 8);
    }
    for (; k < len; k++) out[k] ^= in[k];
}

/* Symbol-pair byte swap pattern (per 128-bit lane). */
#define SWAP_PAIRS 1, 0, 3, 2, 5, 4, 7, 6, 9, 8, 11, 10, 13, 12, 15, 14

__attribute__((target("gfni,avx512f,avx512bw")))
static void mul_xor_gfni512(const coeff_tables *t, const uint8_t *in, uint8_t *out, size_t len) {
    /* At even bytes x holds the symbol's low byte and xs its high byte, at odd
     * bytes the other way round.  Even result = M_ll(x) ^ M_lh(xs); odd
     * result = M_hh(x) ^ M_hl(xs).  Merge-masked affines pick per byte. */
    const __m512i swap = _mm512_broadcast_i32x4(_mm_setr_epi8(SWAP_PAIRS));
    const __m512i ll = _mm512_set1_epi64((long long)t->m_ll), lh = _mm512_set1_epi64((long long)t->m_lh);
    const __m512i hl = _mm512_set1_epi64((long long)t->m_hl), hh = _mm512_set1_epi64((long long)t->m_hh);
    const __mmask64 odd = 0xAAAAAAAAAAAAAAAAull;
    size_t k = 0;
    for (; k + 64 <= len; k += 64) {
        __m512i x = _mm512_loadu_si512((const void *)(in + k));
        __m512i xs = _mm512_shuffle_epi8(x, swap);
        __m512i p = _mm512_gf2p8affine_epi64_epi8(x, ll, 0);
        p = _mm512_mask_gf2p8affine_epi64_epi8(p, odd, x, hh, 0);
        __m512i q = _mm512_gf2p8affine_epi64_epi8(xs, lh, 0);
        q = _mm512_mask_gf2p8affine_epi64_epi8(q, odd, xs, hl, 0);
        __m512i o = _mm512_loadu_si512((const void *)(out + k));
        o = _mm512_ternarylogic_epi32(o, p, q, 0x96);
        _mm512_storeu_si512((void *)(out + k), o);
    }
    for (; k + 1 < len; k += 2) {
        uint16_t sym = (uint16_t)(in[k] | (in[k + 1] << 8));
        uint16_t p = or_gf_mul(t->coeff, sym);
        out[k] ^= (uint8_t)p;
        out[k + 1] ^= (uint8_t)(p >> 8);
    }
}

__attribute__((target("gfni,sse4.1,ssse3")))
static void mul_xor_gfni128(const coeff_tables *t, const uint8_t *in, uint8_t *out, size_t len) {
    const __m128i swap = _mm_setr_epi8(SWAP_PAIRS);
    const __m128i ll = _mm_set1_epi64x((long long)t->m_ll), lh = _mm_set1_epi64x((long long)t->m_lh);
    const __m128i hl = _mm_set1_epi64x((long long)t->m_hl), hh = _mm_set1_epi64x((long long)t->m_hh);
    const __m128i oddm = _mm_set1_epi16((short)0xFF00);
    size_t k = 0;
    for (; k + 16 <= len; k += 16) {
        __m128i x = _mm_loadu_si128((const __m128i *)(in + k));
        __m128i xs = _mm_shuffle_epi8(x, swap);
        __m128i pe = _mm_gf2p8affine_epi64_epi8(x, ll, 0), po = _mm_gf2p8affine_epi64_epi8(x, hh, 0);
        __m128i qe = _mm_gf2p8affine_epi64_epi8(xs, lh, 0), qo = _mm_gf2p8affine_epi64_epi8(xs, hl, 0);
        __m128i r = _mm_xor_si128(_mm_blendv_epi8(pe, po, oddm), _mm_blendv_epi8(qe, qo, oddm));
        __m128i o = _mm_loadu_si128((const __m128i *)(out + k));
        _mm_storeu_si128((__m128i *)(out + k), _mm_xor_si128(o, r));
    }
    for (; k + 1 < len; k += 2) {
        uint16_t sym = (uint16_t)(in[k] | (in[k + 1] << 8));
        uint16_t p = or_gf_mul(t->coeff, sym);
        out[k] ^= (uint8_t)p;
        out[k + 1] ^= (uint8_t)(p >> 8);
    }
}

__attribute__((target("avx2")))
static void mul_xor_avx2(const coeff_tables *t, const uint8_t *in, uint8_t *out, size_t len) {
    const __m256i swap = _mm256_setr_epi8(SWAP_PAIRS, SWAP_PAIRS);
    const __m256i m4 = _mm256_set1_epi8(0x0F);
    const __m256i oddm = _mm256_set1_epi16((short)0xFF00);
    __m256i tb[8];
    for (int i = 0; i < 8; i++)
        tb[i] = _mm256_broadcastsi128_si256(_mm_loadu_si128((const __m128i *)t->nib[i]));
    size_t k = 0;
    for (; k + 32 <= len; k += 32) {
        __m256i x = _mm256_loadu_si256((const __m256i *)(in + k));
        __m256i xs = _mm256_shuffle_epi8(x, swap);
        /* own byte (low byte at even, high byte at odd) and partner byte */
        __m256i xl = _mm256_and_si256(x, m4), xh = _mm256_and_si256(_mm256_srli_epi16(x, 4), m4);
        __m256i sl = _mm256_and_si256(xs, m4), sh = _mm256_and_si256(_mm256_srli_epi16(xs, 4), m4);
        /* even bytes: product low byte = T0[xl]^T2[xh]^T4[sl]^T6[sh] */
        __m256i e = _mm256_xor_si256(
            _mm256_xor_si256(_mm256_shuffle_epi8(tb[0], xl), _mm256_shuffle_epi8(tb[2], xh)),
            _mm256_xor_si256(_mm256_shuffle_epi8(tb[4], sl), _mm256_shuffle_epi8(tb[6], sh)));
        /* odd bytes: own byte is the symbol's high byte, partner its low byte;
         * product high byte = T5[xl]^T7[xh]^T1[sl]^T3[sh] */
        __m256i o = _mm256_xor_si256(
            _mm256_xor_si256(_mm256_shuffle_epi8(tb[5], xl), _mm256_shuffle_epi8(tb[7], xh)),
            _mm256_xor_si256(_mm256_shuffle_epi8(tb[1], sl), _mm256_shuffle_epi8(tb[3], sh)));
        __m256i r = _mm256_blendv_epi8(e, o, oddm);
        __m256i d = _mm256_loadu_si256((const __m256i *)(out + k));
        _mm256_storeu_si256((__m256i *)(out + k), _mm256_xor_si256(d, r));
    }
    for (; k + 1 < len; k += 2) {
        uint16_t p = t->lo[in[k]] ^ t->hi[in[k + 1]];
        out[k] ^= (uint8_t)p;
        out[k + 1] ^= (uint8_t)(p >> 8);
    }
}

static void mul_xor(const coeff_tables *t, int dispatch, const uint8_t *in, uint8_t *out, size_t len) {
    if (t->coeff <= 1) { mul_xor_trivial(t->coeff, in, out, len); return; }
    switch (dispatch) {
    case D_GFNI_AVX512: mul_xor_gfni512(t, in, out, len); break;
    case D_GFNI: mul_xor_gfni128(t, in, out, len); break;
    case D_AVX2: mul_xor_avx2(t, in, out, len); break;
    default: mul_xor_table(t, in, out, len); break;
    }
}

/* ---- threaded driver: src/creator.rs:931-995 ------------------------- */

typedef struct {
    const uint8_t *const *files;
    const uint64_t *lens;
    uint32_t nfiles;
    uint64_t S;
    uint32_t R;
    uint8_t *rec;
    uint32_t T, chunk;
    int dispatch;
    const uint16_t *bases;
    uint64_t ninputs;
    const uint8_t *volatile cur_slice;
    uint8_t *pad;
    pthread_barrier_t bar;
} enc_shared;

typedef struct { enc_shared *sh; uint32_t tid; } enc_arg;

static void *enc_worker(void *p) {
    enc_arg *a = (enc_arg *)p;
    enc_shared *sh = a->sh;
    uint32_t r0 = a->tid * sh->chunk;
    uint32_t r1 = r0 + sh->chunk > sh->R ? sh->R : r0 + sh->chunk;
    coeff_tables *tabs = NULL;
    if (r1 > r0) tabs = (coeff_tables *)malloc(sizeof(coeff_tables) * (r1 - r0));
    uint64_t idx = 0;
    for (uint32_t f = 0; f < sh->nfiles; f++) {
        uint64_t n = or_slice_count(sh->lens[f], sh->S);
        for (uint64_t s = 0; s < n; s++, idx++) {
            if (a->tid == 0) { /* SourceReader::read_next_slice, zero padded */
                uint64_t off = s * sh->S;
                uint64_t have = sh->lens[f] - off < sh->S ? sh->lens[f] - off : sh->S;
                if (have == sh->S) sh->cur_slice = sh->files[f] + off;
                else {
                    memcpy(sh->pad, sh->files[f] + off, have);
                    memset(sh->pad + have, 0, sh->S - have);
                    sh->cur_slice = sh->pad;
                }
            }
            pthread_barrier_wait(&sh->bar);
            const uint8_t *slice = sh->cur_slice;
            for (uint32_t r = r0; r < r1; r++)
                build_coeff_tables(&tabs[r - r0], or_gf_pow(sh->bases[idx], r), sh->dispatch);
            for (uint32_t r = r0; r < r1; r++)
                mul_xor(&tabs[r - r0], sh->dispatch, slice, sh->rec + (size_t)r * sh->S, sh->S);
            pthread_barrier_wait(&sh->bar);
        }
    }
    free(tabs);
    return NULL;
}

int or_encode_simd(const uint8_t *const *files, const uint64_t *lens, uint32_t nfiles,
                   uint64_t slice_size, uint32_t recovery_count, uint8_t *rec_out,
                   uint32_t threads, int force_dispatch) {
    enc_shared sh;
    memset(&sh, 0, sizeof sh);
    uint64_t total = 0;
    for (uint32_t f = 0; f < nfiles; f++) total += or_slice_count(lens[f], slice_size);
    memset(rec_out, 0, (size_t)recovery_count * slice_size);
    if (recovery_count == 0) return OR_OK;
    uint16_t *bases = (uint16_t *)malloc(sizeof(uint16_t) * (total ? total : 1));
    if (or_input_bases((uint32_t)total, bases) != 0) { free(bases); return OR_ERR_TOO_MANY_INPUT_BLOCKS; }
    if (threads == 0) {
        long n = sysconf(_SC_NPROCESSORS_ONLN);
        threads = n > 0 ? (uint32_t)n : 1;
    }
    sh.files = files; sh.lens = lens; sh.nfiles = nfiles; sh.S = slice_size;
    sh.R = recovery_count; sh.rec = rec_out; sh.T = threads;
    sh.chunk = (recovery_count + threads - 1) / threads;
    if (sh.chunk == 0) sh.chunk = 1;
    sh.dispatch = force_dispatch >= 0 ? force_dispatch : or_simd_detect();
    sh.bases = bases;
    sh.ninputs = total;
    sh.pad = (uint8_t *)malloc(slice_size);
    pthread_barrier_init(&sh.bar, NULL, threads);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    enc_arg *args = (enc_arg *)malloc(sizeof(enc_arg) * threads);
    for (uint32_t t = 0; t < threads; t++) {
        args[t].sh = &sh;
        args[t].tid = t;
        pthread_create(&th[t], NULL, enc_worker, &args[t]);
    }
    for (uint32_t t = 0; t < threads; t++) pthread_join(th[t], NULL);
    pthread_barrier_destroy(&sh.bar);
    free(th);
    free(args);
    free(sh.pad);
    free(bases);
    return OR_OK;
}

/* ---- threaded scan and packet hashes (the rest of the CPU step) ------ */
/* Per-slice MD5 + CRC-32 spread over T threads, as scan_via_mmap's
 * par_chunks does (src/source.rs:165-220).  CRC-32 uses slicing-by-8
 * tables (crc32fast is SIMD-accelerated; this is the portable form). */

static uint32_t g_crc8[8][256];
static pthread_once_t g_crc8_once = PTHREAD_ONCE_INIT;

static void crc8_build(void) {
    for (uint32_t n = 0; n < 256; n++) {
        uint32_t c = n;
        for (int k = 0; k < 8; k++) c = (c & 1) ? (0xEDB88320u ^ (c >> 1)) : (c >> 1);
        g_crc8[0][n] = c;
    }
    for (uint32_t n = 0; n < 256; n++)
        for (int t = 1; t < 8; t++) g_crc8[t][n] = (g_crc8[t - 1][n] >> 8) ^ g_crc8[0][g_crc8[t - 1][n] & 0xFF];
}

static uint32_t crc32_s8(uint32_t c, const uint8_t *p, size_t n) {
    c = ~c;
    while (n >= 8) {
        uint32_t lo, hi;
        memcpy(&lo, p, 4);
        memcpy(&hi, p + 4, 4);
        lo ^= c;
        c = g_crc8[7][lo & 0xFF] ^ g_crc8[6][(lo >> 8) & 0xFF] ^ g_crc8[5][(lo >> 16) & 0xFF] ^
            g_crc8[4][lo >> 24] ^ g_crc8[3][hi & 0xFF] ^ g_crc8[2][(hi >> 8) & 0xFF] ^
            g_crc8[1][(hi >> 16) & 0xFF] ^ g_crc8[0][hi >> 24];
        p += 8;
        n -= 8;
    }
    while (n--) c = g_crc8[0][(c ^ *p++) & 0xFF] ^ (c >> 8);
    return ~c;
}

typedef struct {
    const uint8_t *data; uint64_t len, S, nslices;
    uint8_t *md5; uint32_t *crc;
    uint32_t T, tid;
    /* recovery packets */
    const uint8_t *rec; uint32_t first_exp; const uint8_t *set_id;
} mt_arg;

static void *scan_worker(void *p) {
    mt_arg *a = (mt_arg *)p;
    uint8_t *pad = NULL;
    for (uint64_t s = a->tid; s < a->nslices; s += a->T) {
        uint64_t off = s * a->S;
        uint64_t have = a->len - off < a->S ? a->len - off : a->S;
        const uint8_t *q = a->data + off;
        if (have < a->S) {
            if (!pad) pad = (uint8_t *)malloc(a->S);
            memcpy(pad, q, have);
            memset(pad + have, 0, a->S - have);
            q = pad;
        }
        or_md5(q, a->S, a->md5 + 16 * s);
        a->crc[s] = crc32_s8(0, q, a->S);
    }
    free(pad);
    return NULL;
}

void or_slice_checksums_mt(const uint8_t *data, uint64_t len, uint64_t S, uint8_t *md5_out,
                           uint32_t *crc_out, uint32_t threads) {
    pthread_once(&g_crc8_once, crc8_build);
    if (threads == 0) { long n = sysconf(_SC_NPROCESSORS_ONLN); threads = n > 0 ? (uint32_t)n : 1; }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    mt_arg *args = (mt_arg *)calloc(threads, sizeof(mt_arg));
    for (uint32_t t = 0; t < threads; t++) {
        args[t] = (mt_arg){data, len, S, or_slice_count(len, S), md5_out, crc_out, threads, t, 0, 0, 0};
        pthread_create(&th[t], NULL, scan_worker, &args[t]);
    }
    for (uint32_t t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
    free(args);
}

static void *pkt_worker(void *p) {
    static const uint8_t type_tag[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0,
                                         'R', 'e', 'c', 'v', 'S', 'l', 'i', 'c'};
    mt_arg *a = (mt_arg *)p;
    for (uint64_t r = a->tid; r < a->nslices; r += a->T) {
        or_md5_ctx c;
        uint8_t le[4];
        uint32_t e = a->first_exp + (uint32_t)r;
        for (int i = 0; i < 4; i++) le[i] = (uint8_t)(e >> (8 * i));
        or_md5_init(&c);
        or_md5_update(&c, a->set_id, 16);
        or_md5_update(&c, type_tag, 16);
        or_md5_update(&c, le, 4);
        or_md5_update(&c, a->rec + r * a->S, a->S);
        or_md5_final(&c, a->md5 + 16 * r);
    }
    return NULL;
}

/* Recovery packet hashes (build_packet via build_recovery_packet) over T
 * threads; the reference parallelises across volumes (src/creator.rs:296). */
void or_recovery_md5_mt(const uint8_t *rec, uint32_t count, uint64_t S, uint32_t first_exp,
                        const uint8_t set_id[16], uint8_t *md5_out, uint32_t threads) {
    if (threads == 0) { long n = sysconf(_SC_NPROCESSORS_ONLN); threads = n > 0 ? (uint32_t)n : 1; }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    mt_arg *args = (mt_arg *)calloc(threads, sizeof(mt_arg));
    for (uint32_t t = 0; t < threads; t++) {
        args[t] = (mt_arg){0, 0, S, count, md5_out, 0, threads, t, rec, first_exp, set_id};
        pthread_create(&th[t], NULL, pkt_worker, &args[t]);
    }
    for (uint32_t t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
    free(args);
}
