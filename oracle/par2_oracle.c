/*
 * par2_oracle.c -- CPU restatement of javi11/par2rust's PAR2 create path.
 *
 * TEST INFRASTRUCTURE ONLY (see par2_oracle.h): the checker for the GPU
 * product, never linked into it.  Scalar code first; the reference's SIMD
 * encode loop is restated in par2_oracle_simd.c for the timed CPU baseline.
 */
#define _GNU_SOURCE
#include "par2_oracle.h"

#include <errno.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

/* ===================================================================== */
/* GF(2^16) -- src/galois.rs:15-99                                        */
/* ===================================================================== */

#define FIELD_LIMIT 65535u
#define FIELD_GENERATOR 0x1100Bu

static uint16_t g_log[65536];
static uint16_t g_antilog[65536];
static pthread_once_t g_tables_once = PTHREAD_ONCE_INIT;

/* build_tables, src/galois.rs:33-53: walk alpha^l, alpha = 2, reducing by
 * the generator when bit 16 is set; sentinels log[0] = 65535 and
 * antilog[65535] = 0. */
static void build_tables(void) {
    uint32_t b = 1;
    for (uint32_t l = 0; l < FIELD_LIMIT; l++) {
        g_log[b] = (uint16_t)l;
        g_antilog[l] = (uint16_t)b;
        b <<= 1;
        if (b & 0x10000u) b ^= FIELD_GENERATOR;
    }
    g_log[0] = (uint16_t)FIELD_LIMIT;
    g_antilog[FIELD_LIMIT] = 0;
}

static void tables(void) { pthread_once(&g_tables_once, build_tables); }

void or_gf_tables(uint16_t *log_out, uint16_t *antilog_out) {
    tables();
    memcpy(log_out, g_log, sizeof g_log);
    memcpy(antilog_out, g_antilog, sizeof g_antilog);
}

/* gf_mul, src/galois.rs:57-69 */
uint16_t or_gf_mul(uint16_t a, uint16_t b) {
    tables();
    if (a == 0 || b == 0) return 0;
    uint32_t s = (uint32_t)g_log[a] + g_log[b];
    if (s >= FIELD_LIMIT) s -= FIELD_LIMIT;
    return g_antilog[s];
}

/* gf_pow, src/galois.rs:74-87: x^0 = 1 (also 0^0), 0^e = 0. */
uint16_t or_gf_pow(uint16_t base, uint32_t exp) {
    tables();
    if (exp == 0) return 1;
    if (base == 0) return 0;
    uint64_t prod = (uint64_t)g_log[base] * exp;
    return g_antilog[prod % FIELD_LIMIT];
}

/* gf_inv, src/galois.rs:91-99 (a != 0). */
uint16_t or_gf_inv(uint16_t a) {
    tables();
    if (a == 0) return 0;
    uint32_t l = g_log[a];
    return g_antilog[(FIELD_LIMIT - l) % FIELD_LIMIT];
}

/* ===================================================================== */
/* Reed-Solomon -- src/reedsolomon.rs                                     */
/* ===================================================================== */

static uint32_t gcd_u32(uint32_t a, uint32_t b) {
    while (b) { uint32_t t = a % b; a = b; b = t; }
    return a;
}

/* generate_input_bases, src/reedsolomon.rs:130-149: base_i = antilog[L_i],
 * L_i the i-th non-negative integer coprime to 65535 (gcd(65535, 0) = 65535,
 * so L = 0 is skipped).  Returns -1 when the 32768 bases run out. */
int or_input_bases(uint32_t count, uint16_t *out) {
    tables();
    uint32_t n = 0, logbase = 0;
    while (n < count) {
        if (logbase >= FIELD_LIMIT) return -1;
        if (gcd_u32(FIELD_LIMIT, logbase) == 1) out[n++] = g_antilog[logbase];
        logbase++;
    }
    return 0;
}

/* RsEncoder::coefficient, src/reedsolomon.rs:56-60 with first_exponent 0
 * (src/creator.rs:915): coef(r, i) = base_i ^ r. */
uint16_t or_coefficient(const uint16_t *bases, uint32_t recovery_exp, uint32_t input_idx) {
    return or_gf_pow(bases[input_idx], recovery_exp);
}

/* gf_mul_xor_scalar, src/reedsolomon.rs:87-126: out ^= coeff * in over
 * little-endian u16 symbols. */
void or_gf_mul_xor_scalar(uint16_t coeff, const uint8_t *in, uint8_t *out, size_t len) {
    tables();
    if (coeff == 0) return;
    if (coeff == 1) {
        for (size_t k = 0; k < len; k++) out[k] ^= in[k];
        return;
    }
    uint32_t lc = g_log[coeff];
    for (size_t k = 0; k + 1 < len; k += 2) {
        uint32_t sym = (uint32_t)in[k] | ((uint32_t)in[k + 1] << 8);
        uint32_t prod = 0;
        if (sym != 0) {
            uint32_t s = lc + g_log[sym];
            if (s >= FIELD_LIMIT) s -= FIELD_LIMIT;
            prod = g_antilog[s];
        }
        out[k] ^= (uint8_t)prod;
        out[k + 1] ^= (uint8_t)(prod >> 8);
    }
}

/* ===================================================================== */
/* MD5 -- RFC 1321 (the md-5 crate behind src/md5_impl.rs:24-48)         */
/* ===================================================================== */

static uint32_t g_md5_t[64];
static pthread_once_t g_md5_once = PTHREAD_ONCE_INIT;

/* RFC 1321 section 3.4: T[i] = floor(4294967296 * |sin(i)|), i = 1..64. */
static void md5_build(void) {
    for (int i = 0; i < 64; i++)
        g_md5_t[i] = (uint32_t)(uint64_t)floor(4294967296.0 * fabs(sin((double)(i + 1))));
}

static const uint8_t md5_shift[64] = {
    7, 12, 17, 22, 7, 12, 17, 22, 7, 12, 17, 22, 7, 12, 17, 22,
    5, 9, 14, 20, 5, 9, 14, 20, 5, 9, 14, 20, 5, 9, 14, 20,
    4, 11, 16, 23, 4, 11, 16, 23, 4, 11, 16, 23, 4, 11, 16, 23,
    6, 10, 15, 21, 6, 10, 15, 21, 6, 10, 15, 21, 6, 10, 15, 21};

static uint32_t rotl32(uint32_t x, unsigned s) { return (x << s) | (x >> (32 - s)); }

static void md5_block(uint32_t h[4], const uint8_t *p) {
    uint32_t x[16];
    for (int i = 0; i < 16; i++)
        x[i] = (uint32_t)p[4 * i] | ((uint32_t)p[4 * i + 1] << 8) |
               ((uint32_t)p[4 * i + 2] << 16) | ((uint32_t)p[4 * i + 3] << 24);
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3];
    for (int i = 0; i < 64; i++) {
        uint32_t f;
        int g;
        if (i < 16) { f = (b & c) | (~b & d); g = i; }
        else if (i < 32) { f = (b & d) | (c & ~d); g = (5 * i + 1) & 15; }
        else if (i < 48) { f = b ^ c ^ d; g = (3 * i + 5) & 15; }
        else { f = c ^ (b | ~d); g = (7 * i) & 15; }
        uint32_t t = d;
        d = c;
        c = b;
        b = b + rotl32(a + f + g_md5_t[i] + x[g], md5_shift[i]);
        a = t;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d;
}

void or_md5_init(or_md5_ctx *c) {
    pthread_once(&g_md5_once, md5_build);
    c->h[0] = 0x67452301u; c->h[1] = 0xefcdab89u;
    c->h[2] = 0x98badcfeu; c->h[3] = 0x10325476u;
    c->total = 0;
    c->fill = 0;
}

void or_md5_update(or_md5_ctx *c, const uint8_t *data, size_t len) {
    c->total += len;
    if (c->fill) {
        size_t take = 64 - c->fill;
        if (take > len) take = len;
        memcpy(c->buf + c->fill, data, take);
        c->fill += (uint32_t)take;
        data += take;
        len -= take;
        if (c->fill == 64) { md5_block(c->h, c->buf); c->fill = 0; }
    }
    while (len >= 64) { md5_block(c->h, data); data += 64; len -= 64; }
    if (len) { memcpy(c->buf, data, len); c->fill = (uint32_t)len; }
}

void or_md5_final(or_md5_ctx *c, uint8_t out[16]) {
    uint64_t bits = c->total * 8;
    uint8_t pad[72] = {0x80};
    size_t padlen = (c->fill < 56) ? (56 - c->fill) : (120 - c->fill);
    or_md5_update(c, pad, padlen);
    uint8_t lenb[8];
    for (int i = 0; i < 8; i++) lenb[i] = (uint8_t)(bits >> (8 * i));
    or_md5_update(c, lenb, 8);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) out[4 * i + j] = (uint8_t)(c->h[i] >> (8 * j));
}

void or_md5(const uint8_t *data, size_t len, uint8_t out[16]) {
    or_md5_ctx c;
    or_md5_init(&c);
    or_md5_update(&c, data, len);
    or_md5_final(&c, out);
}

/* ===================================================================== */
/* CRC-32/ISO-HDLC (crc32fast::hash, src/source.rs:172,181,275):           */
/* reflected polynomial 0xEDB88320, init and xorout 0xFFFFFFFF.            */
/* ===================================================================== */

static uint32_t g_crc_table[256];
static pthread_once_t g_crc_once = PTHREAD_ONCE_INIT;

static void crc_build(void) {
    for (uint32_t n = 0; n < 256; n++) {
        uint32_t c = n;
        for (int k = 0; k < 8; k++) c = (c & 1) ? (0xEDB88320u ^ (c >> 1)) : (c >> 1);
        g_crc_table[n] = c;
    }
}

uint32_t or_crc32(const uint8_t *data, size_t len) {
    pthread_once(&g_crc_once, crc_build);
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < len; i++) c = g_crc_table[(c ^ data[i]) & 0xFF] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

/* ===================================================================== */
/* Scan -- src/source.rs:145-309                                          */
/* ===================================================================== */

uint64_t or_slice_count(uint64_t len, uint64_t slice_size) {
    return (len + slice_size - 1) / slice_size;
}

/* scan_via_mmap per-slice closure, src/source.rs:165-220: MD5 and CRC-32
 * of each slice, the trailing partial slice zero padded to slice_size. */
void or_slice_checksums(const uint8_t *data, uint64_t len, uint64_t slice_size,
                        uint8_t *md5_out, uint32_t *crc_out) {
    uint64_t n = or_slice_count(len, slice_size);
    uint8_t *pad = NULL;
    for (uint64_t s = 0; s < n; s++) {
        uint64_t off = s * slice_size;
        uint64_t have = len - off < slice_size ? len - off : slice_size;
        const uint8_t *p = data + off;
        if (have < slice_size) {
            pad = (uint8_t *)calloc(slice_size, 1);
            memcpy(pad, p, have);
            p = pad;
        }
        or_md5(p, slice_size, md5_out + 16 * s);
        crc_out[s] = or_crc32(p, slice_size);
        if (pad) { free(pad); pad = NULL; }
    }
}

/* file_hashes, src/source.rs:223-228: MD5 of the unpadded file and of its
 * first min(16384, len) bytes. */
void or_file_hashes(const uint8_t *data, uint64_t len, uint8_t hash_full[16],
                    uint8_t hash16k[16]) {
    or_md5(data, len, hash_full);
    or_md5(data, len < 16384 ? len : 16384, hash16k);
}

/* compute_file_id, src/source.rs:303-309: MD5(hash16k || le64(len) || name). */
void or_file_id(const uint8_t hash16k[16], uint64_t len, const uint8_t *name,
                size_t name_len, uint8_t out[16]) {
    or_md5_ctx c;
    or_md5_init(&c);
    or_md5_update(&c, hash16k, 16);
    uint8_t l[8];
    for (int i = 0; i < 8; i++) l[i] = (uint8_t)(len >> (8 * i));
    or_md5_update(&c, l, 8);
    or_md5_update(&c, name, name_len);
    or_md5_final(&c, out);
}

/* ===================================================================== */
/* Encode -- src/creator.rs:883-1018 (scalar kernel)                      */
/* ===================================================================== */

int or_encode(const uint8_t *const *files, const uint64_t *lens, uint32_t nfiles,
              uint64_t slice_size, uint32_t recovery_count, uint8_t *rec_out) {
    return or_encode_range(files, lens, nfiles, slice_size, 0, recovery_count, rec_out);
}

/* Recovery blocks of exponents first_exp .. first_exp + count - 1: block k
 * uses coef = gf_pow(base_i, first_exp + k), as RsEncoder::new(first_exp, ..)
 * (src/reedsolomon.rs:39-60) would; the reference's create always starts at
 * exponent 0 (src/creator.rs:915), so block k here equals block
 * first_exp + k of a create with more recovery blocks. */
int or_encode_range(const uint8_t *const *files, const uint64_t *lens, uint32_t nfiles,
                    uint64_t slice_size, uint32_t first_exp, uint32_t count, uint8_t *rec_out) {
    const uint32_t recovery_count = count;
    uint64_t total = 0;
    for (uint32_t f = 0; f < nfiles; f++) total += or_slice_count(lens[f], slice_size);
    memset(rec_out, 0, (size_t)recovery_count * slice_size);
    if (recovery_count == 0) return OR_OK;
    uint16_t *bases = (uint16_t *)malloc(sizeof(uint16_t) * (total ? total : 1));
    if (or_input_bases((uint32_t)total, bases) != 0) { free(bases); return OR_ERR_TOO_MANY_INPUT_BLOCKS; }
    uint8_t *slice = (uint8_t *)malloc(slice_size);
    uint64_t input_idx = 0;
    /* flat (file, slice) order, src/creator.rs:895-900 */
    for (uint32_t f = 0; f < nfiles; f++) {
        uint64_t n = or_slice_count(lens[f], slice_size);
        for (uint64_t s = 0; s < n; s++, input_idx++) {
            uint64_t off = s * slice_size;
            uint64_t have = lens[f] - off < slice_size ? lens[f] - off : slice_size;
            memcpy(slice, files[f] + off, have);
            memset(slice + have, 0, slice_size - have); /* src/creator.rs:79-85 */
            for (uint32_t r = 0; r < recovery_count; r++)
                or_gf_mul_xor_scalar(or_coefficient(bases, first_exp + r, (uint32_t)input_idx), slice,
                                     rec_out + (size_t)r * slice_size, slice_size);
        }
    }
    free(slice);
    free(bases);
    return OR_OK;
}

/* ===================================================================== */
/* Packets -- src/format.rs:33-54, src/packet/ modules                         */
/* ===================================================================== */

static const uint8_t PACKET_MAGIC[8] = {'P', 'A', 'R', '2', 0, 'P', 'K', 'T'};
static const uint8_t TYPE_MAIN[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0,
                                      'M', 'a', 'i', 'n', 0, 0, 0, 0};
static const uint8_t TYPE_FILE_DESC[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0,
                                           'F', 'i', 'l', 'e', 'D', 'e', 's', 'c'};
static const uint8_t TYPE_FILE_VERIFY[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0,
                                             'I', 'F', 'S', 'C', 0, 0, 0, 0};
static const uint8_t TYPE_RECOVERY[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0,
                                          'R', 'e', 'c', 'v', 'S', 'l', 'i', 'c'};
static const uint8_t TYPE_CREATOR[16] = {'P', 'A', 'R', ' ', '2', '.', '0', 0,
                                         'C', 'r', 'e', 'a', 't', 'o', 'r', 0};
static const char CREATOR_STRING[] = "Created by par2rust v0.1.0"; /* src/packet/creator.rs:3 */

static void put_le64(uint8_t *p, uint64_t v) { for (int i = 0; i < 8; i++) p[i] = (uint8_t)(v >> (8 * i)); }
static void put_le32(uint8_t *p, uint32_t v) { for (int i = 0; i < 4; i++) p[i] = (uint8_t)(v >> (8 * i)); }

/* build_packet, src/format.rs:33-54: magic || le64(64+len) ||
 * MD5(set_id || type || body) || set_id || type || body. */
size_t or_build_packet(const uint8_t set_id[16], const uint8_t type_tag[16],
                       const uint8_t *body, size_t body_len, uint8_t *out) {
    size_t total = 64 + body_len;
    memcpy(out, PACKET_MAGIC, 8);
    put_le64(out + 8, total);
    or_md5_ctx c;
    or_md5_init(&c);
    or_md5_update(&c, set_id, 16);
    or_md5_update(&c, type_tag, 16);
    or_md5_update(&c, body, body_len);
    or_md5_final(&c, out + 16);
    memcpy(out + 32, set_id, 16);
    memcpy(out + 48, type_tag, 16);
    memmove(out + 64, body, body_len);
    return total;
}

static int cmp16(const void *a, const void *b) { return memcmp(a, b, 16); }

/* build_main_packet, src/packet/main_packet.rs:25-42: le64(S) || le32(n) ||
 * file ids sorted ascending; set_id = MD5(body). */
size_t or_main_packet(uint64_t slice_size, const uint8_t *file_ids, uint32_t n,
                      uint8_t *out, uint8_t set_id_out[16]) {
    size_t blen = 12 + 16 * (size_t)n;
    uint8_t *body = (uint8_t *)malloc(blen);
    put_le64(body, slice_size);
    put_le32(body + 8, n);
    memcpy(body + 12, file_ids, 16 * (size_t)n);
    qsort(body + 12, n, 16, cmp16);
    or_md5(body, blen, set_id_out);
    size_t t = or_build_packet(set_id_out, TYPE_MAIN, body, blen, out);
    free(body);
    return t;
}

/* build_recovery_packet, src/packet/recovery.rs:11-19: le32(exp) || data. */
size_t or_recovery_packet(const uint8_t set_id[16], uint16_t exponent,
                          const uint8_t *slice, size_t slice_len, uint8_t *out) {
    uint8_t *body = (uint8_t *)malloc(4 + slice_len);
    put_le32(body, exponent);
    memcpy(body + 4, slice, slice_len);
    size_t t = or_build_packet(set_id, TYPE_RECOVERY, body, 4 + slice_len, out);
    free(body);
    return t;
}

/* build_creator_packet, src/packet/creator.rs:8-14: string zero padded to 4. */
size_t or_creator_packet(const uint8_t set_id[16], uint8_t *out) {
    size_t n = sizeof(CREATOR_STRING) - 1;
    size_t padded = (n + 3) & ~(size_t)3;
    uint8_t body[64] = {0};
    memcpy(body, CREATOR_STRING, n);
    return or_build_packet(set_id, TYPE_CREATOR, body, padded, out);
}

/* ===================================================================== */
/* Volume schemes -- src/creator.rs:325-431                               */
/* ===================================================================== */

typedef struct { uint32_t *v; uint32_t n, cap; } u32vec;
static int vpush(u32vec *a, uint32_t x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? a->cap * 2 : 16;
        a->v = (uint32_t *)realloc(a->v, sizeof(uint32_t) * a->cap);
    }
    a->v[a->n++] = x;
    return 0;
}

/* exponential_split, src/creator.rs:367-385: 1, 1, 2, 4, ... capped by what
 * remains; the doubling starts once two volumes exist. */
static void exponential_split(uint32_t total, u32vec *out) {
    uint32_t remaining = total, next = 1;
    while (remaining > 0) {
        uint32_t take = next < remaining ? next : remaining;
        vpush(out, take);
        remaining -= take;
        if (out->n >= 2) next = next > 0x7FFFFFFFu ? 0xFFFFFFFFu : next * 2;
    }
}

static int resolve(const or_scheme *s, uint32_t total, u32vec *out) {
    switch (s->kind) {
    case OR_SCHEME_SINGLE:
        vpush(out, total);
        break;
    case OR_SCHEME_EXPONENTIAL:
        exponential_split(total, out);
        break;
    case OR_SCHEME_UNIFORM: { /* uniform_split, src/creator.rs:390-409 */
        uint32_t count = s->uniform_count;
        if (count == 0 || count > total) return OR_ERR_INVALID_VOLUME_SCHEME;
        uint32_t base = total / count, rem = total - base * count;
        for (uint32_t i = 0; i + 1 < count; i++) vpush(out, base);
        vpush(out, base + rem);
        break;
    }
    case OR_SCHEME_EXPLICIT: { /* src/creator.rs:346-360 */
        if (s->n_explicit == 0) return OR_ERR_INVALID_VOLUME_SCHEME;
        uint64_t sum = 0;
        for (uint32_t i = 0; i < s->n_explicit; i++) {
            if (s->explicit_sizes[i] == 0) return OR_ERR_INVALID_VOLUME_SCHEME;
            sum += s->explicit_sizes[i];
        }
        if (sum != total) return OR_ERR_INVALID_VOLUME_SCHEME;
        for (uint32_t i = 0; i < s->n_explicit; i++) vpush(out, s->explicit_sizes[i]);
        break;
    }
    default:
        return OR_ERR_INVALID_VOLUME_SCHEME;
    }
    return OR_OK;
}

/* cap_volume_sizes, src/creator.rs:414-431 */
static void cap_volume_sizes(const u32vec *in, uint32_t cap, u32vec *out) {
    for (uint32_t i = 0; i < in->n; i++) {
        uint32_t n = in->v[i];
        if (n <= cap) { vpush(out, n); continue; }
        uint32_t rem = n;
        while (rem > cap) { vpush(out, cap); rem -= cap; }
        if (rem > 0) vpush(out, rem);
    }
}

/* resolve_volume_sizes, src/creator.rs:325-362, for the scheme under `k`
 * Limited wrappers (wrapper k - 1 is the outermost of them; wrapper 0 has
 * limit_cap, wrapper j >= 1 outer_caps[j - 1]).  As in the reference's
 * recursion through `inner: Box<VolumeScheme>` (:330-344), a wrapper checks
 * its cap before its inner scheme is resolved, and caps the inner sizes. */
static int resolve_wrapped(const or_scheme *s, uint32_t k, uint32_t total, u32vec *out) {
    if (k == 0) return resolve(s, total, out);
    uint32_t cap = k == 1 ? s->limit_cap : s->outer_caps[k - 2];
    if (cap == 0) return OR_ERR_INVALID_VOLUME_SCHEME;
    u32vec inner = {0};
    int rc = resolve_wrapped(s, k - 1, total, &inner);
    if (rc == OR_OK) cap_volume_sizes(&inner, cap, out);
    free(inner.v);
    return rc;
}

int or_resolve_volume_sizes(const or_scheme *s, uint32_t total, uint32_t *sizes_out,
                            uint32_t cap) {
    if (s->n_outer_caps && !s->limited) return -OR_ERR_INVALID_VOLUME_SCHEME;
    u32vec out = {0};
    int rc = resolve_wrapped(s, s->limited ? 1 + s->n_outer_caps : 0, total, &out);
    if (rc != OR_OK) { free(out.v); return -rc; }
    uint32_t n = out.n;
    for (uint32_t i = 0; i < n && i < cap; i++) sizes_out[i] = out.v[i];
    free(out.v);
    return (int)n;
}

/* ===================================================================== */
/* run_create -- src/creator.rs:182-321, 1025-1087; scan src/source.rs:59 */
/* ===================================================================== */

typedef struct {
    uint8_t *data;
    uint64_t len;
    const uint8_t *name;
    uint32_t name_len;
    uint8_t hash_full[16], hash16k[16], file_id[16];
    uint64_t nslices;
    uint8_t *md5s;
    uint32_t *crcs;
} or_source;

static void set_msg(char *msg, size_t cap, const char *fmt, const char *a) {
    if (msg && cap) snprintf(msg, cap, fmt, a);
}

/* write_atomic, src/creator.rs:1070-1087: tmp file + fsync + rename. */
static int write_atomic(const char *path, const uint8_t *bytes, size_t len) {
    size_t pl = strlen(path);
    char *tmp = (char *)malloc(pl + 32);
    memcpy(tmp, path, pl);
    memcpy(tmp + pl, ".par2rust.tmp", 14);
    FILE *f = fopen(tmp, "wb");
    if (!f) { free(tmp); return OR_ERR_IO; }
    if (len && fwrite(bytes, 1, len, f) != len) { fclose(f); free(tmp); return OR_ERR_IO; }
    fflush(f);
    fsync(fileno(f));
    fclose(f);
    int rc = rename(tmp, path) == 0 ? OR_OK : OR_ERR_IO;
    free(tmp);
    return rc;
}

/* derive_volume_filename, src/creator.rs:1046-1057, using Rust's
 * Path::file_stem / extension split at the last '.' of the file name (a
 * leading dot alone does not start an extension). */
static void volume_name(const char *index, uint32_t first, uint32_t count, char *out, size_t cap) {
    const char *slash = strrchr(index, '/');
    const char *fname = slash ? slash + 1 : index;
    size_t dirlen = (size_t)(fname - index);
    const char *dot = strrchr(fname, '.');
    int has_ext = dot && dot != fname;
    size_t stemlen = has_ext ? (size_t)(dot - fname) : strlen(fname);
    const char *ext = has_ext ? dot + 1 : "par2";
    snprintf(out, cap, "%.*s%.*s.vol%u+%u.%s", (int)dirlen, index, (int)stemlen, fname, first,
             count, ext);
}

static uint8_t *read_file(const char *path, uint64_t *len_out, int *err) {
    struct stat st;
    if (stat(path, &st) != 0) { *err = OR_ERR_IO; return NULL; }
    FILE *f = fopen(path, "rb");
    if (!f) { *err = OR_ERR_IO; return NULL; }
    uint64_t len = (uint64_t)st.st_size;
    uint8_t *buf = (uint8_t *)malloc(len ? len : 1);
    if (len && fread(buf, 1, len, f) != len) { fclose(f); free(buf); *err = OR_ERR_IO; return NULL; }
    fclose(f);
    *len_out = len;
    *err = OR_OK;
    return buf;
}

int or_run_create(const char *output, uint64_t slice_size, uint32_t recovery_count,
                  const or_scheme *scheme, uint32_t nfiles, const char *const *paths,
                  const uint8_t *const *names, const uint32_t *name_lens,
                  char *names_out, size_t names_cap, char *msg, size_t msg_cap) {
    int rc = OR_OK;
    or_source *src = (or_source *)calloc(nfiles ? nfiles : 1, sizeof(or_source));
    uint8_t *critical = NULL;
    size_t names_used = 0;
    if (names_out && names_cap) names_out[0] = 0;

    /* SourceFile::scan for each input, src/source.rs:59-136 */
    for (uint32_t f = 0; f < nfiles; f++) {
        or_source *s = &src[f];
        if (name_lens[f] == 0 || memchr(names[f], 0, name_lens[f])) {
            set_msg(msg, msg_cap, "invalid file name %s", paths[f]);
            rc = OR_ERR_INVALID_FILE_NAME; goto done;
        }
        if (slice_size == 0 || slice_size % 4) { rc = OR_ERR_INVALID_SLICE_SIZE; goto done; }
        s->data = read_file(paths[f], &s->len, &rc);
        if (rc != OR_OK) { set_msg(msg, msg_cap, "cannot read %s", paths[f]); goto done; }
        if (s->len == 0) { set_msg(msg, msg_cap, "empty file %s", paths[f]); rc = OR_ERR_EMPTY_FILE; goto done; }
        s->name = names[f];
        s->name_len = name_lens[f];
        s->nslices = or_slice_count(s->len, slice_size);
        s->md5s = (uint8_t *)malloc(16 * s->nslices);
        s->crcs = (uint32_t *)malloc(4 * s->nslices);
        or_slice_checksums(s->data, s->len, slice_size, s->md5s, s->crcs);
        or_file_hashes(s->data, s->len, s->hash_full, s->hash16k);
        or_file_id(s->hash16k, s->len, s->name, s->name_len, s->file_id);
    }

    /* validation order, src/creator.rs:193-217 */
    if (nfiles == 0) { rc = OR_ERR_NO_INPUT_FILES; goto done; }
    if (nfiles > 32768) { rc = OR_ERR_TOO_MANY_FILES; goto done; }
    if (recovery_count > 65535) { rc = OR_ERR_TOO_MANY_RECOVERY_BLOCKS; goto done; }
    if (slice_size == 0 || slice_size % 4) { rc = OR_ERR_INVALID_SLICE_SIZE; goto done; }
    uint64_t total_blocks = 0;
    for (uint32_t f = 0; f < nfiles; f++) total_blocks += src[f].nslices;
    if (total_blocks > 32768) { rc = OR_ERR_TOO_MANY_INPUT_BLOCKS; goto done; }

    /* critical packets, src/creator.rs:219-235 */
    uint8_t set_id[16];
    size_t crit_cap = 64 + 12 + 16 * (size_t)nfiles + 128;
    for (uint32_t f = 0; f < nfiles; f++)
        crit_cap += 2 * 64 + 56 + src[f].name_len + 4 + 16 + 20 * src[f].nslices;
    critical = (uint8_t *)malloc(crit_cap);
    size_t cl = 0;
    {
        uint8_t *ids = (uint8_t *)malloc(16 * (size_t)nfiles);
        for (uint32_t f = 0; f < nfiles; f++) memcpy(ids + 16 * f, src[f].file_id, 16);
        cl += or_main_packet(slice_size, ids, nfiles, critical + cl, set_id);
        free(ids);
    }
    for (uint32_t f = 0; f < nfiles; f++) {
        or_source *s = &src[f];
        /* build_file_desc_packet, src/packet/file_desc.rs:17-28 */
        size_t np = ((size_t)s->name_len + 3) & ~(size_t)3;
        uint8_t *body = (uint8_t *)calloc(56 + np, 1);
        memcpy(body, s->file_id, 16);
        memcpy(body + 16, s->hash_full, 16);
        memcpy(body + 32, s->hash16k, 16);
        put_le64(body + 48, s->len);
        memcpy(body + 56, s->name, s->name_len);
        cl += or_build_packet(set_id, TYPE_FILE_DESC, body, 56 + np, critical + cl);
        free(body);
        /* build_file_verify_packet, src/packet/file_verify.rs:12-23 */
        size_t vl = 16 + 20 * s->nslices;
        body = (uint8_t *)malloc(vl);
        memcpy(body, s->file_id, 16);
        for (uint64_t k = 0; k < s->nslices; k++) {
            memcpy(body + 16 + 20 * k, s->md5s + 16 * k, 16);
            put_le32(body + 16 + 20 * k + 16, s->crcs[k]);
        }
        cl += or_build_packet(set_id, TYPE_FILE_VERIFY, body, vl, critical + cl);
        free(body);
    }
    cl += or_creator_packet(set_id, critical + cl);

    rc = write_atomic(output, critical, cl);
    if (rc != OR_OK) goto done;
    if (names_out) {
        size_t l = strlen(output);
        if (names_used + l + 1 <= names_cap) { memcpy(names_out + names_used, output, l + 1); names_used += l + 1; }
    }

    if (recovery_count > 0) {
        uint32_t sizes[70000];
        int nv = or_resolve_volume_sizes(scheme, recovery_count, sizes, 70000);
        if (nv < 0) { rc = -nv; goto done; }
        uint8_t *rec = (uint8_t *)malloc((size_t)recovery_count * slice_size);
        const uint8_t **fp = (const uint8_t **)malloc(sizeof(void *) * nfiles);
        uint64_t *fl = (uint64_t *)malloc(sizeof(uint64_t) * nfiles);
        for (uint32_t f = 0; f < nfiles; f++) { fp[f] = src[f].data; fl[f] = src[f].len; }
        rc = or_encode_simd(fp, fl, nfiles, slice_size, recovery_count, rec, 0, -1);
        free(fp);
        free(fl);
        if (rc != OR_OK) { free(rec); goto done; }
        uint32_t first = 0;
        for (int v = 0; v < nv; v++) {
            /* build_volume_packet, src/creator.rs:1025-1042 */
            size_t vlen = 2 * cl + (size_t)sizes[v] * (68 + slice_size);
            uint8_t *vb = (uint8_t *)malloc(vlen);
            size_t o = 0;
            memcpy(vb, critical, cl); o += cl;
            for (uint32_t k = 0; k < sizes[v]; k++)
                o += or_recovery_packet(set_id, (uint16_t)(first + k),
                                        rec + (size_t)(first + k) * slice_size, slice_size, vb + o);
            memcpy(vb + o, critical, cl); o += cl;
            char vname[4096];
            volume_name(output, first, sizes[v], vname, sizeof vname);
            rc = write_atomic(vname, vb, o);
            free(vb);
            if (rc != OR_OK) { free(rec); goto done; }
            if (names_out) {
                size_t l = strlen(vname);
                if (names_used + l + 1 <= names_cap) { memcpy(names_out + names_used, vname, l + 1); names_used += l + 1; }
            }
            first += sizes[v];
        }
        free(rec);
    }

done:
    for (uint32_t f = 0; f < nfiles; f++) { free(src[f].data); free(src[f].md5s); free(src[f].crcs); }
    free(src);
    free(critical);
    return rc;
}

/* ===================================================================== */
/* xorshift64 workload generator                                          */
/* ===================================================================== */

uint64_t or_xorshift64_fill(uint8_t *buf, size_t len, uint64_t state) {
    size_t i = 0;
    while (i < len) {
        state ^= state << 13;
        state ^= state >> 7;
        state ^= state << 17;
        for (int b = 0; b < 8 && i < len; b++, i++) buf[i] = (uint8_t)(state >> (8 * b));
    }
    return state;
}
