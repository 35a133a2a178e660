/*
 * CPU oracle for the LXMF vanity GPU search: TEST INFRASTRUCTURE ONLY.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it, as the checker; the product library never links or calls it.
 *
 * Restates the reference's CPU path (src/main.rs) in plain C. The function
 * groups, the reference lines they follow and what pins each of them are
 * listed in oracle.h.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#if defined(__x86_64__)
#include <cpuid.h>
#include <immintrin.h>
#endif

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------
 * SHA-256 and SHA-512, FIPS 180-4.
 */

static const uint32_t K256[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

static const uint32_t H256[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                                 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};

static inline uint32_t ror32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
static inline uint64_t ror64(uint64_t x, int n) { return (x >> n) | (x << (64 - n)); }
static inline uint32_t be32(const uint8_t* p) {
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}
static inline uint64_t be64(const uint8_t* p) { return (uint64_t)be32(p) << 32 | be32(p + 4); }
static inline void put_be32(uint8_t* p, uint32_t v) {
    p[0] = v >> 24; p[1] = v >> 16; p[2] = v >> 8; p[3] = v;
}
static inline uint64_t le64(const uint8_t* p) {
    uint64_t v = 0;
    for (int i = 7; i >= 0; i--) v = v << 8 | p[i];
    return v;
}

/* One compression of the 64-byte block `data` into `s` (portable C). */
static void sha256_block_c(uint32_t s[8], const uint8_t* data) {
    uint32_t w[64];
    for (int i = 0; i < 16; i++) w[i] = be32(data + 4 * i);
    for (int i = 16; i < 64; i++) {
        uint32_t s0 = ror32(w[i - 15], 7) ^ ror32(w[i - 15], 18) ^ (w[i - 15] >> 3);
        uint32_t s1 = ror32(w[i - 2], 17) ^ ror32(w[i - 2], 19) ^ (w[i - 2] >> 10);
        w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t a = s[0], b = s[1], c = s[2], d = s[3], e = s[4], f = s[5], g = s[6], h = s[7];
    for (int i = 0; i < 64; i++) {
        uint32_t t1 = h + (ror32(e, 6) ^ ror32(e, 11) ^ ror32(e, 25)) + ((e & f) ^ (~e & g)) + K256[i] + w[i];
        uint32_t t2 = (ror32(a, 2) ^ ror32(a, 13) ^ ror32(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
        h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    s[0] += a; s[1] += b; s[2] += c; s[3] += d; s[4] += e; s[5] += f; s[6] += g; s[7] += h;
}

#if defined(__x86_64__)
/* The same compression on the x86 SHA extensions (as the reference's sha2
 * crate selects at run time when the CPU has them). */
__attribute__((target("sha,sse4.1,ssse3"))) static void sha256_block_ni(uint32_t s[8], const uint8_t* data) {
    const __m128i bswap = _mm_set_epi64x(0x0c0d0e0f08090a0bULL, 0x0405060700010203ULL);
    __m128i tmp = _mm_loadu_si128((const __m128i*)&s[0]);
    __m128i st1 = _mm_loadu_si128((const __m128i*)&s[4]);
    tmp = _mm_shuffle_epi32(tmp, 0xB1);
    st1 = _mm_shuffle_epi32(st1, 0x1B);
    __m128i st0 = _mm_alignr_epi8(tmp, st1, 8);
    st1 = _mm_blend_epi16(st1, tmp, 0xF0);
    const __m128i save0 = st0, save1 = st1;
    __m128i m[4];
    for (int i = 0; i < 16; i++) {
        __m128i wk;
        if (i < 4) {
            m[i] = _mm_shuffle_epi8(_mm_loadu_si128((const __m128i*)(data + 16 * i)), bswap);
        } else {
            __m128i t = _mm_sha256msg1_epu32(m[i & 3], m[(i + 1) & 3]);
            t = _mm_add_epi32(t, _mm_alignr_epi8(m[(i + 3) & 3], m[(i + 2) & 3], 4));
            m[i & 3] = _mm_sha256msg2_epu32(t, m[(i + 3) & 3]);
        }
        wk = _mm_add_epi32(m[i & 3], _mm_loadu_si128((const __m128i*)&K256[4 * i]));
        st1 = _mm_sha256rnds2_epu32(st1, st0, wk);
        wk = _mm_shuffle_epi32(wk, 0x0E);
        st0 = _mm_sha256rnds2_epu32(st0, st1, wk);
    }
    st0 = _mm_add_epi32(st0, save0);
    st1 = _mm_add_epi32(st1, save1);
    tmp = _mm_shuffle_epi32(st0, 0x1B);
    st1 = _mm_shuffle_epi32(st1, 0xB1);
    st0 = _mm_blend_epi16(tmp, st1, 0xF0);
    st1 = _mm_alignr_epi8(st1, tmp, 8);
    _mm_storeu_si128((__m128i*)&s[0], st0);
    _mm_storeu_si128((__m128i*)&s[4], st1);
}

static int cpu_has_shani(void) {
    unsigned a, b, c, d;
    if (!__get_cpuid_count(7, 0, &a, &b, &c, &d)) return 0;
    int sha = (b >> 29) & 1;
    if (!__get_cpuid(1, &a, &b, &c, &d)) return 0;
    int sse41 = (c >> 19) & 1, ssse3 = (c >> 9) & 1;
    return sha && sse41 && ssse3;
}
#endif

typedef void (*sha256_block_fn)(uint32_t s[8], const uint8_t* data);
static sha256_block_fn sha256_block = 0;
static pthread_once_t sha_once = PTHREAD_ONCE_INIT;
static void pick_sha256(void) {
    sha256_block = sha256_block_c;
#if defined(__x86_64__)
    const char* force = getenv("LXV_ORACLE_NO_SHANI");
    if (cpu_has_shani() && !(force && force[0] == '1')) sha256_block = sha256_block_ni;
#endif
}
static inline sha256_block_fn sha256_impl(void) {
    pthread_once(&sha_once, pick_sha256);
    return sha256_block;
}

int orc_sha256_uses_shani(void) {
#if defined(__x86_64__)
    return sha256_impl() == sha256_block_ni;
#else
    return 0;
#endif
}

void orc_sha256(const uint8_t* msg, size_t len, uint8_t out[32]) {
    sha256_block_fn blk = sha256_impl();
    uint32_t s[8];
    memcpy(s, H256, sizeof s);
    size_t full = len / 64;
    for (size_t i = 0; i < full; i++) blk(s, msg + 64 * i);
    uint8_t tail[128];
    size_t rem = len - 64 * full;
    memset(tail, 0, sizeof tail);
    memcpy(tail, msg + 64 * full, rem);
    tail[rem] = 0x80;
    size_t tl = (rem + 9 <= 64) ? 64 : 128;
    uint64_t bits = (uint64_t)len * 8;
    for (int i = 0; i < 8; i++) tail[tl - 1 - i] = (uint8_t)(bits >> (8 * i));
    blk(s, tail);
    if (tl == 128) blk(s, tail + 64);
    for (int i = 0; i < 8; i++) put_be32(out + 4 * i, s[i]);
}

static const uint64_t K512[80] = {
    0x428a2f98d728ae22ULL, 0x7137449123ef65cdULL, 0xb5c0fbcfec4d3b2fULL, 0xe9b5dba58189dbbcULL,
    0x3956c25bf348b538ULL, 0x59f111f1b605d019ULL, 0x923f82a4af194f9bULL, 0xab1c5ed5da6d8118ULL,
    0xd807aa98a3030242ULL, 0x12835b0145706fbeULL, 0x243185be4ee4b28cULL, 0x550c7dc3d5ffb4e2ULL,
    0x72be5d74f27b896fULL, 0x80deb1fe3b1696b1ULL, 0x9bdc06a725c71235ULL, 0xc19bf174cf692694ULL,
    0xe49b69c19ef14ad2ULL, 0xefbe4786384f25e3ULL, 0x0fc19dc68b8cd5b5ULL, 0x240ca1cc77ac9c65ULL,
    0x2de92c6f592b0275ULL, 0x4a7484aa6ea6e483ULL, 0x5cb0a9dcbd41fbd4ULL, 0x76f988da831153b5ULL,
    0x983e5152ee66dfabULL, 0xa831c66d2db43210ULL, 0xb00327c898fb213fULL, 0xbf597fc7beef0ee4ULL,
    0xc6e00bf33da88fc2ULL, 0xd5a79147930aa725ULL, 0x06ca6351e003826fULL, 0x142929670a0e6e70ULL,
    0x27b70a8546d22ffcULL, 0x2e1b21385c26c926ULL, 0x4d2c6dfc5ac42aedULL, 0x53380d139d95b3dfULL,
    0x650a73548baf63deULL, 0x766a0abb3c77b2a8ULL, 0x81c2c92e47edaee6ULL, 0x92722c851482353bULL,
    0xa2bfe8a14cf10364ULL, 0xa81a664bbc423001ULL, 0xc24b8b70d0f89791ULL, 0xc76c51a30654be30ULL,
    0xd192e819d6ef5218ULL, 0xd69906245565a910ULL, 0xf40e35855771202aULL, 0x106aa07032bbd1b8ULL,
    0x19a4c116b8d2d0c8ULL, 0x1e376c085141ab53ULL, 0x2748774cdf8eeb99ULL, 0x34b0bcb5e19b48a8ULL,
    0x391c0cb3c5c95a63ULL, 0x4ed8aa4ae3418acbULL, 0x5b9cca4f7763e373ULL, 0x682e6ff3d6b2b8a3ULL,
    0x748f82ee5defb2fcULL, 0x78a5636f43172f60ULL, 0x84c87814a1f0ab72ULL, 0x8cc702081a6439ecULL,
    0x90befffa23631e28ULL, 0xa4506cebde82bde9ULL, 0xbef9a3f7b2c67915ULL, 0xc67178f2e372532bULL,
    0xca273eceea26619cULL, 0xd186b8c721c0c207ULL, 0xeada7dd6cde0eb1eULL, 0xf57d4f7fee6ed178ULL,
    0x06f067aa72176fbaULL, 0x0a637dc5a2c898a6ULL, 0x113f9804bef90daeULL, 0x1b710b35131c471bULL,
    0x28db77f523047d84ULL, 0x32caab7b40c72493ULL, 0x3c9ebe0a15c9bebcULL, 0x431d67c49c100d4cULL,
    0x4cc5d4becb3e42b6ULL, 0x597f299cfc657e2aULL, 0x5fcb6fab3ad6faecULL, 0x6c44198c4a475817ULL};

static void sha512_block(uint64_t s[8], const uint8_t* data) {
    uint64_t w[80];
    for (int i = 0; i < 16; i++) w[i] = be64(data + 8 * i);
    for (int i = 16; i < 80; i++) {
        uint64_t s0 = ror64(w[i - 15], 1) ^ ror64(w[i - 15], 8) ^ (w[i - 15] >> 7);
        uint64_t s1 = ror64(w[i - 2], 19) ^ ror64(w[i - 2], 61) ^ (w[i - 2] >> 6);
        w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint64_t a = s[0], b = s[1], c = s[2], d = s[3], e = s[4], f = s[5], g = s[6], h = s[7];
    for (int i = 0; i < 80; i++) {
        uint64_t t1 = h + (ror64(e, 14) ^ ror64(e, 18) ^ ror64(e, 41)) + ((e & f) ^ (~e & g)) + K512[i] + w[i];
        uint64_t t2 = (ror64(a, 28) ^ ror64(a, 34) ^ ror64(a, 39)) + ((a & b) ^ (a & c) ^ (b & c));
        h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    s[0] += a; s[1] += b; s[2] += c; s[3] += d; s[4] += e; s[5] += f; s[6] += g; s[7] += h;
}

void orc_sha512(const uint8_t* msg, size_t len, uint8_t out[64]) {
    uint64_t s[8] = {0x6a09e667f3bcc908ULL, 0xbb67ae8584caa73bULL, 0x3c6ef372fe94f82bULL,
                     0xa54ff53a5f1d36f1ULL, 0x510e527fade682d1ULL, 0x9b05688c2b3e6c1fULL,
                     0x1f83d9abfb41bd6bULL, 0x5be0cd19137e2179ULL};
    size_t full = len / 128;
    for (size_t i = 0; i < full; i++) sha512_block(s, msg + 128 * i);
    uint8_t tail[256];
    size_t rem = len - 128 * full;
    memset(tail, 0, sizeof tail);
    memcpy(tail, msg + 128 * full, rem);
    tail[rem] = 0x80;
    size_t tl = (rem + 17 <= 128) ? 128 : 256;
    uint64_t bits = (uint64_t)len * 8;
    for (int i = 0; i < 8; i++) tail[tl - 1 - i] = (uint8_t)(bits >> (8 * i));
    sha512_block(s, tail);
    if (tl == 256) sha512_block(s, tail + 128);
    for (int i = 0; i < 8; i++) {
        put_be32(out + 8 * i, (uint32_t)(s[i] >> 32));
        put_be32(out + 8 * i + 4, (uint32_t)s[i]);
    }
}

/* ------------------------------------------------------------------------
 * GF(2^255 - 19) in 5 limbs of 51 bits with 128-bit products: the layout of
 * fiat-crypto's curve25519_64 (main.rs:31, 43-120). Limbs of stored values
 * stay below 2^52; add/sub results stay below 2^54, which fe_mul accepts.
 */

typedef struct { uint64_t v[5]; } fe;
#define M51 ((1ULL << 51) - 1)

static void fe_frombytes(fe* h, const uint8_t s[32]) {
    uint64_t w0 = le64(s), w1 = le64(s + 8), w2 = le64(s + 16), w3 = le64(s + 24);
    h->v[0] = w0 & M51;
    h->v[1] = ((w0 >> 51) | (w1 << 13)) & M51;
    h->v[2] = ((w1 >> 38) | (w2 << 26)) & M51;
    h->v[3] = ((w2 >> 25) | (w3 << 39)) & M51;
    h->v[4] = (w3 >> 12) & M51; /* bit 255 dropped (RFC 7748 5) */
}

static void fe_tobytes(uint8_t s[32], const fe* f) {
    uint64_t t[5];
    memcpy(t, f->v, sizeof t);
    for (int pass = 0; pass < 2; pass++) {
        for (int i = 0; i < 4; i++) { t[i + 1] += t[i] >> 51; t[i] &= M51; }
        t[0] += 19 * (t[4] >> 51);
        t[4] &= M51;
    }
    /* Now value < 2^255 + 2^13: subtract p once if value >= p. */
    uint64_t q = (t[0] + 19) >> 51;
    q = (t[1] + q) >> 51;
    q = (t[2] + q) >> 51;
    q = (t[3] + q) >> 51;
    q = (t[4] + q) >> 51;
    t[0] += 19 * q;
    for (int i = 0; i < 4; i++) { t[i + 1] += t[i] >> 51; t[i] &= M51; }
    t[4] &= M51;
    uint64_t w[4];
    w[0] = t[0] | (t[1] << 51);
    w[1] = (t[1] >> 13) | (t[2] << 38);
    w[2] = (t[2] >> 26) | (t[3] << 25);
    w[3] = (t[3] >> 39) | (t[4] << 12);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 8; j++) s[8 * i + j] = (uint8_t)(w[i] >> (8 * j));
}

/* One carry pass: limbs back below 2^51 + 2^18. */
static inline void fe_weak(fe* h) {
    for (int i = 0; i < 4; i++) { h->v[i + 1] += h->v[i] >> 51; h->v[i] &= M51; }
    h->v[0] += 19 * (h->v[4] >> 51);
    h->v[4] &= M51;
}

static inline void fe_one(fe* h) { memset(h, 0, sizeof *h); h->v[0] = 1; }
static inline void fe_zero(fe* h) { memset(h, 0, sizeof *h); }

static inline void fe_add(fe* h, const fe* f, const fe* g) {
    for (int i = 0; i < 5; i++) h->v[i] = f->v[i] + g->v[i];
}

/* f - g + 4p, limbwise; g limbs below 2^53. */
static inline void fe_sub(fe* h, const fe* f, const fe* g) {
    h->v[0] = f->v[0] + ((1ULL << 53) - 76) - g->v[0];
    for (int i = 1; i < 5; i++) h->v[i] = f->v[i] + ((1ULL << 53) - 4) - g->v[i];
}

static inline void fe_carry128(fe* h, u128 r0, u128 r1, u128 r2, u128 r3, u128 r4) {
    r1 += (uint64_t)(r0 >> 51);
    uint64_t h0 = (uint64_t)r0 & M51;
    r2 += (uint64_t)(r1 >> 51);
    uint64_t h1 = (uint64_t)r1 & M51;
    r3 += (uint64_t)(r2 >> 51);
    uint64_t h2 = (uint64_t)r2 & M51;
    r4 += (uint64_t)(r3 >> 51);
    uint64_t h3 = (uint64_t)r3 & M51;
    u128 c = (r4 >> 51) * 19 + h0;
    uint64_t h4 = (uint64_t)r4 & M51;
    h0 = (uint64_t)c & M51;
    h1 += (uint64_t)(c >> 51);
    h->v[0] = h0; h->v[1] = h1; h->v[2] = h2; h->v[3] = h3; h->v[4] = h4;
}

static void fe_mul(fe* h, const fe* f, const fe* g) {
    const uint64_t f0 = f->v[0], f1 = f->v[1], f2 = f->v[2], f3 = f->v[3], f4 = f->v[4];
    const uint64_t g0 = g->v[0], g1 = g->v[1], g2 = g->v[2], g3 = g->v[3], g4 = g->v[4];
    const uint64_t g1x = 19 * g1, g2x = 19 * g2, g3x = 19 * g3, g4x = 19 * g4;
    u128 r0 = (u128)f0 * g0 + (u128)f1 * g4x + (u128)f2 * g3x + (u128)f3 * g2x + (u128)f4 * g1x;
    u128 r1 = (u128)f0 * g1 + (u128)f1 * g0 + (u128)f2 * g4x + (u128)f3 * g3x + (u128)f4 * g2x;
    u128 r2 = (u128)f0 * g2 + (u128)f1 * g1 + (u128)f2 * g0 + (u128)f3 * g4x + (u128)f4 * g3x;
    u128 r3 = (u128)f0 * g3 + (u128)f1 * g2 + (u128)f2 * g1 + (u128)f3 * g0 + (u128)f4 * g4x;
    u128 r4 = (u128)f0 * g4 + (u128)f1 * g3 + (u128)f2 * g2 + (u128)f3 * g1 + (u128)f4 * g0;
    fe_carry128(h, r0, r1, r2, r3, r4);
}

static void fe_sq(fe* h, const fe* f) {
    const uint64_t f0 = f->v[0], f1 = f->v[1], f2 = f->v[2], f3 = f->v[3], f4 = f->v[4];
    const uint64_t d0 = 2 * f0, d1 = 2 * f1, d3x = 38 * f3, f4x = 19 * f4, d4x = 38 * f4;
    /* r0 = f0^2 + 38 f1 f4 + 38 f2 f3 */
    u128 r0 = (u128)f0 * f0 + (u128)f1 * d4x + (u128)f2 * d3x;
    /* r1 = 2 f0 f1 + 38 f2 f4 + 19 f3^2 */
    u128 r1 = (u128)d0 * f1 + (u128)f2 * d4x + (u128)f3 * (19 * f3);
    /* r2 = 2 f0 f2 + f1^2 + 38 f3 f4 */
    u128 r2 = (u128)d0 * f2 + (u128)f1 * f1 + (u128)f3 * d4x;
    /* r3 = 2 f0 f3 + 2 f1 f2 + 19 f4^2 */
    u128 r3 = (u128)d0 * f3 + (u128)d1 * f2 + (u128)f4 * f4x;
    /* r4 = 2 f0 f4 + 2 f1 f3 + f2^2 */
    u128 r4 = (u128)d0 * f4 + (u128)d1 * f3 + (u128)f2 * f2;
    fe_carry128(h, r0, r1, r2, r3, r4);
}

static void fe_sq_n(fe* h, const fe* f, int n) {
    fe t = *f;
    for (int i = 0; i < n; i++) fe_sq(&t, &t);
    *h = t;
}

/* z^(p-2) by ref10's addition chain (main.rs:92-107). */
static void fe_invert(fe* out, const fe* z) {
    fe z2, z9, z11, z2_5_0, z2_10_0, z2_20_0, z2_40_0, z2_50_0, z2_100_0, z2_200_0, z2_250_0, t;
    fe_sq(&z2, z);
    fe_sq_n(&t, &z2, 2); fe_mul(&z9, &t, z);
    fe_mul(&z11, &z9, &z2);
    fe_sq(&t, &z11); fe_mul(&z2_5_0, &t, &z9);
    fe_sq_n(&t, &z2_5_0, 5); fe_mul(&z2_10_0, &t, &z2_5_0);
    fe_sq_n(&t, &z2_10_0, 10); fe_mul(&z2_20_0, &t, &z2_10_0);
    fe_sq_n(&t, &z2_20_0, 20); fe_mul(&z2_40_0, &t, &z2_20_0);
    fe_sq_n(&t, &z2_40_0, 10); fe_mul(&z2_50_0, &t, &z2_10_0);
    fe_sq_n(&t, &z2_50_0, 50); fe_mul(&z2_100_0, &t, &z2_50_0);
    fe_sq_n(&t, &z2_100_0, 100); fe_mul(&z2_200_0, &t, &z2_100_0);
    fe_sq_n(&t, &z2_200_0, 50); fe_mul(&z2_250_0, &t, &z2_50_0);
    fe_sq_n(&t, &z2_250_0, 5); fe_mul(out, &t, &z11);
}

/* a^e for a 256-bit little-endian exponent (square and multiply). */
static void fe_pow(fe* out, const fe* a, const uint8_t e[32]) {
    fe r;
    fe_one(&r);
    for (int bit = 255; bit >= 0; bit--) {
        fe_sq(&r, &r);
        if ((e[bit >> 3] >> (bit & 7)) & 1) fe_mul(&r, &r, a);
    }
    *out = r;
}

static int fe_equal(const fe* a, const fe* b) {
    uint8_t x[32], y[32];
    fe_tobytes(x, a);
    fe_tobytes(y, b);
    return memcmp(x, y, 32) == 0;
}

static void fe_from_u64(fe* h, uint64_t v) {
    fe_zero(h);
    h->v[0] = v & M51;
    h->v[1] = v >> 51;
}

void orc_fe_mul(const uint8_t a[32], const uint8_t b[32], uint8_t out[32]) {
    fe x, y, r;
    fe_frombytes(&x, a); fe_frombytes(&y, b); fe_mul(&r, &x, &y); fe_tobytes(out, &r);
}
void orc_fe_sq(const uint8_t a[32], uint8_t out[32]) {
    fe x, r;
    fe_frombytes(&x, a); fe_sq(&r, &x); fe_tobytes(out, &r);
}
void orc_fe_add(const uint8_t a[32], const uint8_t b[32], uint8_t out[32]) {
    fe x, y, r;
    fe_frombytes(&x, a); fe_frombytes(&y, b); fe_add(&r, &x, &y); fe_tobytes(out, &r);
}
void orc_fe_sub(const uint8_t a[32], const uint8_t b[32], uint8_t out[32]) {
    fe x, y, r;
    fe_frombytes(&x, a); fe_frombytes(&y, b); fe_sub(&r, &x, &y); fe_tobytes(out, &r);
}
void orc_fe_invert(const uint8_t a[32], uint8_t out[32]) {
    fe x, r;
    fe_frombytes(&x, a); fe_invert(&r, &x); fe_tobytes(out, &r);
}
void orc_fe_canon(const uint8_t a[32], uint8_t out[32]) {
    fe x;
    fe_frombytes(&x, a); fe_tobytes(out, &x);
}

/* ------------------------------------------------------------------------
 * Montgomery-curve ladder, RFC 7748 section 5 (A = 486662, a24 = 121665).
 */

static void fe_cswap(fe* a, fe* b, uint64_t swap) {
    uint64_t m = 0 - swap;
    for (int i = 0; i < 5; i++) {
        uint64_t t = m & (a->v[i] ^ b->v[i]);
        a->v[i] ^= t;
        b->v[i] ^= t;
    }
}

static void ladder(uint8_t out[32], const uint8_t k[32], const uint8_t u[32], int top_bit) {
    fe x1, x2, z2, x3, z3, a, aa, b, bb, e, c, d, da, cb, t, a24;
    fe_frombytes(&x1, u);
    fe_one(&x2); fe_zero(&z2);
    x3 = x1; fe_one(&z3);
    fe_from_u64(&a24, 121665);
    uint64_t swap = 0;
    for (int i = top_bit; i >= 0; i--) {
        uint64_t bit = (k[i >> 3] >> (i & 7)) & 1;
        swap ^= bit;
        fe_cswap(&x2, &x3, swap);
        fe_cswap(&z2, &z3, swap);
        swap = bit;
        fe_add(&a, &x2, &z2); fe_sq(&aa, &a);
        fe_sub(&b, &x2, &z2); fe_sq(&bb, &b);
        fe_sub(&e, &aa, &bb);
        fe_add(&c, &x3, &z3); fe_sub(&d, &x3, &z3);
        fe_mul(&da, &d, &a); fe_mul(&cb, &c, &b);
        fe_add(&t, &da, &cb); fe_sq(&x3, &t);
        fe_sub(&t, &da, &cb); fe_sq(&t, &t); fe_mul(&z3, &x1, &t);
        fe_mul(&x2, &aa, &bb);
        fe_mul(&t, &a24, &e); fe_add(&t, &aa, &t); fe_mul(&z2, &e, &t);
    }
    fe_cswap(&x2, &x3, swap);
    fe_cswap(&z2, &z3, swap);
    fe_invert(&z2, &z2);
    fe_mul(&x2, &x2, &z2);
    fe_tobytes(out, &x2);
}

void orc_x25519(const uint8_t k[32], const uint8_t u[32], uint8_t out[32]) {
    uint8_t kc[32];
    memcpy(kc, k, 32);
    kc[0] &= 248;
    kc[31] &= 127;
    kc[31] |= 64;
    ladder(out, kc, u, 254);
}

void orc_mont_ladder(const uint8_t k[32], const uint8_t u[32], uint8_t out[32]) { ladder(out, k, u, 255); }

static const uint8_t BASE_U[32] = {9};

/* ------------------------------------------------------------------------
 * Ed25519 public key, RFC 8032 5.1.5: twisted Edwards -x^2 + y^2 = 1 + d x^2 y^2
 * in extended coordinates, complete addition (a = -1).
 */

typedef struct { fe X, Y, Z, T; } ge;

typedef struct { fe d2, bx, by; } ed_consts;

static void fe_neg(fe* h, const fe* f) {
    fe z;
    fe_zero(&z);
    fe_sub(h, &z, f);
}

static void ed_setup(ed_consts* k) {
    /* p - 2, (p - 1) / 4, (p + 3) / 8 as little-endian bytes. */
    uint8_t pm2[32], pm1q[32], pp3e[32];
    memset(pm2, 0xff, 32); pm2[0] = 0xeb; pm2[31] = 0x7f;
    memset(pm1q, 0xff, 32); pm1q[0] = 0xfb; pm1q[31] = 0x1f;
    memset(pp3e, 0xff, 32); pp3e[0] = 0xfe; pp3e[31] = 0x0f;
    fe n121665, n121666, d, t;
    fe_from_u64(&n121665, 121665);
    fe_from_u64(&n121666, 121666);
    fe_pow(&t, &n121666, pm2);
    fe_mul(&d, &n121665, &t);
    fe_neg(&d, &d);
    fe_add(&k->d2, &d, &d);
    /* y = 4/5; x = sqrt((y^2 - 1) / (d y^2 + 1)), the even root. */
    fe four, five, one, y2, num, den, x2, x, chk, sqrtm1, two;
    fe_from_u64(&four, 4); fe_from_u64(&five, 5); fe_one(&one); fe_from_u64(&two, 2);
    fe_pow(&t, &five, pm2);
    fe_mul(&k->by, &four, &t);
    fe_sq(&y2, &k->by);
    fe_sub(&num, &y2, &one);
    fe_mul(&den, &d, &y2); fe_add(&den, &den, &one);
    fe_pow(&t, &den, pm2);
    fe_mul(&x2, &num, &t);
    fe_pow(&x, &x2, pp3e);
    fe_sq(&chk, &x);
    if (!fe_equal(&chk, &x2)) {
        fe_pow(&sqrtm1, &two, pm1q);
        fe_mul(&x, &x, &sqrtm1);
    }
    uint8_t xb[32];
    fe_tobytes(xb, &x);
    if (xb[0] & 1) fe_neg(&x, &x);
    fe_weak(&x);
    fe_weak(&k->d2);
    k->bx = x;
}

static void ge_add(ge* r, const ge* p, const ge* q, const fe* d2) {
    fe a, b, c, d, e, f, g, h, t;
    fe_sub(&a, &p->Y, &p->X); fe_sub(&t, &q->Y, &q->X); fe_mul(&a, &a, &t);
    fe_add(&b, &p->Y, &p->X); fe_add(&t, &q->Y, &q->X); fe_mul(&b, &b, &t);
    fe_mul(&c, &p->T, &q->T); fe_mul(&c, &c, d2);
    fe_mul(&d, &p->Z, &q->Z); fe_add(&d, &d, &d);
    fe_sub(&e, &b, &a); fe_sub(&f, &d, &c); fe_add(&g, &d, &c); fe_add(&h, &b, &a);
    fe_mul(&r->X, &e, &f); fe_mul(&r->Y, &g, &h); fe_mul(&r->T, &e, &h); fe_mul(&r->Z, &f, &g);
}

void orc_ed25519_pub(const uint8_t seed[32], uint8_t out[32]) {
    ed_consts k;
    ed_setup(&k);
    uint8_t h[64];
    orc_sha512(seed, 32, h);
    h[0] &= 248;
    h[31] &= 127;
    h[31] |= 64;
    ge base, acc;
    base.X = k.bx; base.Y = k.by; fe_one(&base.Z); fe_mul(&base.T, &k.bx, &k.by);
    fe_zero(&acc.X); fe_one(&acc.Y); fe_one(&acc.Z); fe_zero(&acc.T);
    for (int i = 255; i >= 0; i--) {
        ge_add(&acc, &acc, &acc, &k.d2);
        if ((h[i >> 3] >> (i & 7)) & 1) ge_add(&acc, &acc, &base, &k.d2);
    }
    fe zi, x, y;
    fe_invert(&zi, &acc.Z);
    fe_mul(&x, &acc.X, &zi);
    fe_mul(&y, &acc.Y, &zi);
    uint8_t xb[32];
    fe_tobytes(out, &y);
    fe_tobytes(xb, &x);
    out[31] |= (uint8_t)((xb[0] & 1) << 7);
}

/* ------------------------------------------------------------------------
 * Scalars, main.rs:124-170.
 */

void orc_step_scalar(const uint8_t k[32], uint64_t n, uint8_t out[32]) {
    memcpy(out, k, 32);
    u128 carry = (u128)n * 8;
    for (int i = 0; i < 32 && carry; i++) {
        u128 sum = out[i] + (carry & 0xff);
        out[i] = (uint8_t)sum;
        carry = (carry >> 8) + (sum >> 8);
    }
}

void orc_previous_scalar(const uint8_t k[32], uint8_t out[32]) {
    memcpy(out, k, 32);
    int borrow = 8;
    for (int i = 0; i < 32 && borrow; i++) {
        int v = (int)out[i] - borrow;
        out[i] = (uint8_t)(v & 0xff);
        borrow = v < 0;
    }
}

int orc_start_scalar_ok(const uint8_t k[32]) {
    if ((k[0] & 7) != 0 || (k[31] & 0xc0) != 0x40) return 0;
    int high_all_set = k[31] == 127;
    for (int i = 9; i < 31; i++) high_all_set &= k[i] == 0xff;
    int low_all_clear = k[31] == 64;
    for (int i = 0; i < 31; i++) low_all_clear &= k[i] == 0;
    return !high_all_set && !low_all_clear;
}

/* ------------------------------------------------------------------------
 * Addresses and prefixes, main.rs:174-225 and gpu/mod.rs:74-84.
 */

static const char ASPECTS[] = "lxmf.delivery";

void orc_name_hash(uint8_t out[10]) {
    uint8_t h[32];
    orc_sha256((const uint8_t*)ASPECTS, strlen(ASPECTS), h);
    memcpy(out, h, 10);
}

void orc_lxmf_address(const uint8_t x_pub[32], const uint8_t ed_pub[32], uint8_t out[16]) {
    uint8_t pk[64], ih[32], material[26], addr[32];
    memcpy(pk, x_pub, 32);
    memcpy(pk + 32, ed_pub, 32);
    orc_sha256(pk, 64, ih);
    orc_name_hash(material);
    memcpy(material + 10, ih, 16);
    orc_sha256(material, 26, addr);
    memcpy(out, addr, 16);
}

void orc_address_of_identity(const uint8_t identity[64], uint8_t out[16]) {
    uint8_t x_pub[32], ed_pub[32];
    orc_x25519(identity, BASE_U, x_pub);
    orc_ed25519_pub(identity + 32, ed_pub);
    orc_lxmf_address(x_pub, ed_pub, out);
}

static int hexval(char c) {
    if (c >= '0' && c <= '9') return c - '0';
    if (c >= 'a' && c <= 'f') return c - 'a' + 10;
    if (c >= 'A' && c <= 'F') return c - 'A' + 10;
    return -1;
}

int orc_prefix_parse(const char* s, uint8_t bytes[16], int* nibbles) {
    size_t n = strlen(s);
    if (n == 0 || n > 32) return -1;
    memset(bytes, 0, 16);
    for (size_t i = 0; i < n; i++) {
        int v = hexval(s[i]);
        if (v < 0) return -1;
        bytes[i / 2] |= (uint8_t)(v << ((i % 2) ? 0 : 4));
    }
    *nibbles = (int)n;
    return 0;
}

int orc_prefix_matches(const uint8_t bytes[16], int nibbles, const uint8_t addr[16]) {
    int whole = nibbles / 2;
    if (memcmp(addr, bytes, (size_t)whole) != 0) return 0;
    return nibbles % 2 == 0 || (addr[whole] >> 4) == (bytes[whole] >> 4);
}

void orc_prefix_words(const uint8_t bytes[16], int nibbles, uint32_t want[4], uint32_t mask[4]) {
    uint8_t w[16] = {0}, m[16] = {0};
    for (int i = 0; i < nibbles; i++) {
        int shift = (i % 2 == 0) ? 4 : 0;
        m[i / 2] |= (uint8_t)(0xF << shift);
        w[i / 2] |= (uint8_t)(((bytes[i / 2] >> shift) & 0xF) << shift);
    }
    for (int i = 0; i < 4; i++) { want[i] = be32(w + 4 * i); mask[i] = be32(m + 4 * i); }
}

/* ------------------------------------------------------------------------
 * The key walk, main.rs:229-290.
 */

struct orc_walk {
    uint8_t start[32];
    uint64_t next_n;
    fe xp, zp, xc, zc;
    fe q_plus, q_minus;
    size_t batch;
    fe* xs;
    fe* zs;
    fe* prefix;
};

orc_walk* orc_walk_new_batch(const uint8_t start[32], size_t batch) {
    if (batch == 0) return NULL;
    orc_walk* w = (orc_walk*)calloc(1, sizeof *w);
    if (!w) return NULL;
    w->xs = (fe*)malloc(batch * sizeof(fe));
    w->zs = (fe*)malloc(batch * sizeof(fe));
    w->prefix = (fe*)malloc(batch * sizeof(fe));
    if (!w->xs || !w->zs || !w->prefix) { orc_walk_free(w); return NULL; }
    w->batch = batch;
    memcpy(w->start, start, 32);
    uint8_t prev[32], u[32], eight[32] = {8}, q[32];
    orc_previous_scalar(start, prev);
    orc_x25519(prev, BASE_U, u);
    fe_frombytes(&w->xp, u);
    fe_one(&w->zp);
    orc_x25519(start, BASE_U, u);
    fe_frombytes(&w->xc, u);
    fe_one(&w->zc);
    orc_mont_ladder(eight, BASE_U, q);
    fe qf, one;
    fe_frombytes(&qf, q);
    fe_one(&one);
    fe_add(&w->q_plus, &qf, &one);
    fe_sub(&w->q_minus, &qf, &one);
    return w;
}

orc_walk* orc_walk_new(const uint8_t start[32]) { return orc_walk_new_batch(start, 256); }

void orc_walk_free(orc_walk* w) {
    if (!w) return;
    free(w->xs);
    free(w->zs);
    free(w->prefix);
    free(w);
}

uint64_t orc_walk_next_batch(orc_walk* w, uint8_t (*out)[32], size_t len) {
    uint64_t first = w->next_n;
    if (len > w->batch) len = w->batch;
    for (size_t i = 0; i < len; i++) {
        w->xs[i] = w->xc;
        w->zs[i] = w->zc;
        /* xADD: P(n+1) = P(n) + Q, with P(n) - Q = P(n-1) and Q affine. */
        fe a, b, t1, t2, s, xn, zn;
        fe_sub(&a, &w->xc, &w->zc);
        fe_mul(&t1, &a, &w->q_plus);
        fe_add(&b, &w->xc, &w->zc);
        fe_mul(&t2, &b, &w->q_minus);
        fe_add(&s, &t1, &t2); fe_sq(&s, &s); fe_mul(&xn, &w->zp, &s);
        fe_sub(&s, &t1, &t2); fe_sq(&s, &s); fe_mul(&zn, &w->xp, &s);
        w->xp = w->xc; w->zp = w->zc;
        w->xc = xn; w->zc = zn;
    }
    /* Montgomery's trick: one inversion for the batch. */
    fe acc;
    fe_one(&acc);
    for (size_t i = 0; i < len; i++) {
        w->prefix[i] = acc;
        fe_mul(&acc, &acc, &w->zs[i]);
    }
    fe inv;
    fe_invert(&inv, &acc);
    for (size_t i = len; i-- > 0;) {
        fe z_inv, u;
        fe_mul(&z_inv, &inv, &w->prefix[i]);
        fe_mul(&inv, &inv, &w->zs[i]);
        fe_mul(&u, &w->xs[i], &z_inv);
        fe_tobytes(out[i], &u);
    }
    w->next_n += len;
    return first;
}

/* ------------------------------------------------------------------------
 * Range scans (the search loop of main.rs:292-323 over explicit ranges).
 */

typedef struct {
    uint64_t g;
    uint8_t addr[16];
} hit_rec;

typedef struct {
    const uint8_t* start;
    const uint8_t* ed_pub;
    const uint64_t* g0;
    const uint64_t* len;
    uint32_t n_chunks;
    const uint8_t* pbytes;
    int pnib;
    pthread_mutex_t lock;
    uint32_t next_chunk;
    /* results */
    uint8_t xor_acc[16];
    hit_rec* hits;
    size_t n_hits, cap_hits;
    int failed;
} scan_job;

static void* scan_worker(void* arg) {
    scan_job* job = (scan_job*)arg;
    sha256_block_fn blk = sha256_impl();
    uint8_t nh[10];
    orc_name_hash(nh);
    /* Block 2 of SHA-256(x_pub || ed_pub) is pure padding for 64 bytes. */
    uint8_t pad64[64] = {0x80};
    pad64[62] = 0x02; /* 512 bits */
    uint8_t (*keys)[32] = (uint8_t(*)[32])malloc(256 * 32);
    uint8_t xor_acc[16] = {0};
    hit_rec* local = NULL;
    size_t n_local = 0, cap_local = 0;
    if (!keys) { job->failed = 1; return NULL; }
    for (;;) {
        pthread_mutex_lock(&job->lock);
        uint32_t c = job->next_chunk++;
        pthread_mutex_unlock(&job->lock);
        if (c >= job->n_chunks) break;
        uint8_t k0[32];
        orc_step_scalar(job->start, job->g0[c], k0);
        orc_walk* w = orc_walk_new(k0);
        if (!w) { job->failed = 1; break; }
        uint64_t left = job->len[c];
        while (left) {
            size_t take = left < 256 ? (size_t)left : 256;
            uint64_t first = orc_walk_next_batch(w, keys, take);
            for (size_t i = 0; i < take; i++) {
                uint8_t blk1[64], m3[64];
                uint32_t s[8], t[8];
                memcpy(blk1, keys[i], 32);
                memcpy(blk1 + 32, job->ed_pub, 32);
                memcpy(s, H256, sizeof s);
                blk(s, blk1);
                blk(s, pad64);
                memset(m3, 0, 64);
                memcpy(m3, nh, 10);
                for (int j = 0; j < 4; j++) put_be32(m3 + 10 + 4 * j, s[j]);
                m3[26] = 0x80;
                m3[62] = 0x00;
                m3[63] = 26 * 8;
                memcpy(t, H256, sizeof t);
                blk(t, m3);
                uint8_t addr[16];
                for (int j = 0; j < 4; j++) put_be32(addr + 4 * j, t[j]);
                for (int j = 0; j < 16; j++) xor_acc[j] ^= addr[j];
                if (orc_prefix_matches(job->pbytes, job->pnib, addr)) {
                    if (n_local == cap_local) {
                        cap_local = cap_local ? 2 * cap_local : 64;
                        hit_rec* nl = (hit_rec*)realloc(local, cap_local * sizeof *local);
                        if (!nl) { job->failed = 1; break; }
                        local = nl;
                    }
                    local[n_local].g = job->g0[c] + first + i;
                    memcpy(local[n_local].addr, addr, 16);
                    n_local++;
                }
            }
            left -= take;
        }
        orc_walk_free(w);
    }
    free(keys);
    pthread_mutex_lock(&job->lock);
    for (int j = 0; j < 16; j++) job->xor_acc[j] ^= xor_acc[j];
    if (n_local) {
        if (job->n_hits + n_local > job->cap_hits) {
            size_t nc = job->cap_hits ? job->cap_hits : 64;
            while (nc < job->n_hits + n_local) nc *= 2;
            hit_rec* nh2 = (hit_rec*)realloc(job->hits, nc * sizeof *nh2);
            if (!nh2) job->failed = 1;
            else { job->hits = nh2; job->cap_hits = nc; }
        }
        if (!job->failed) {
            memcpy(job->hits + job->n_hits, local, n_local * sizeof *local);
            job->n_hits += n_local;
        }
    }
    pthread_mutex_unlock(&job->lock);
    free(local);
    return NULL;
}

static int cmp_hit(const void* a, const void* b) {
    uint64_t x = ((const hit_rec*)a)->g, y = ((const hit_rec*)b)->g;
    return x < y ? -1 : x > y;
}

int orc_scan_ranges(const uint8_t start[32], const uint8_t ed_pub[32],
                    const uint64_t* g0, const uint64_t* len, uint32_t n_ranges,
                    const uint8_t prefix_bytes[16], int prefix_nibbles,
                    uint64_t* hit_g, uint8_t (*hit_addr)[16], uint32_t cap,
                    uint32_t* n_hits, uint8_t xor_out[16], int threads) {
    if (prefix_nibbles < 0 || prefix_nibbles > 32) return -1;
    if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (threads < 1) threads = 1;
    /* Split ranges into chunks so every thread gets work; each chunk starts
     * its own walk (two ladders), so chunks stay at least 2^14 keys. */
    uint64_t total = 0;
    for (uint32_t r = 0; r < n_ranges; r++) total += len[r];
    uint64_t chunk = total / ((uint64_t)threads * 16) + 1;
    if (chunk < (1u << 14)) chunk = 1u << 14;
    uint64_t n_chunks = 0;
    for (uint32_t r = 0; r < n_ranges; r++) n_chunks += (len[r] + chunk - 1) / chunk;
    uint64_t* cg = (uint64_t*)malloc((n_chunks + 1) * sizeof(uint64_t));
    uint64_t* cl = (uint64_t*)malloc((n_chunks + 1) * sizeof(uint64_t));
    if (!cg || !cl) { free(cg); free(cl); return -1; }
    uint64_t c = 0;
    for (uint32_t r = 0; r < n_ranges; r++) {
        for (uint64_t off = 0; off < len[r]; off += chunk) {
            cg[c] = g0[r] + off;
            cl[c] = (len[r] - off < chunk) ? len[r] - off : chunk;
            c++;
        }
    }
    scan_job job;
    memset(&job, 0, sizeof job);
    job.start = start; job.ed_pub = ed_pub; job.g0 = cg; job.len = cl;
    job.n_chunks = (uint32_t)n_chunks; job.pbytes = prefix_bytes; job.pnib = prefix_nibbles;
    pthread_mutex_init(&job.lock, NULL);
    pthread_t* tids = (pthread_t*)malloc((size_t)threads * sizeof(pthread_t));
    int started = 0;
    for (int i = 0; tids && i < threads; i++) {
        if (pthread_create(&tids[i], NULL, scan_worker, &job) != 0) break;
        started++;
    }
    if (started == 0) scan_worker(&job);
    for (int i = 0; i < started; i++) pthread_join(tids[i], NULL);
    free(tids);
    pthread_mutex_destroy(&job.lock);
    free(cg);
    free(cl);
    if (job.failed) { free(job.hits); return -1; }
    qsort(job.hits, job.n_hits, sizeof *job.hits, cmp_hit);
    *n_hits = (uint32_t)job.n_hits;
    for (size_t i = 0; i < job.n_hits && i < cap; i++) {
        if (hit_g) hit_g[i] = job.hits[i].g;
        if (hit_addr) memcpy(hit_addr[i], job.hits[i].addr, 16);
    }
    memcpy(xor_out, job.xor_acc, 16);
    free(job.hits);
    return 0;
}
