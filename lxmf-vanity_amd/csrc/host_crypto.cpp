// Host-side crypto of the search driver (see host_crypto.h).
#include "host_crypto.h"

#include <sys/random.h>

#include <cstring>

namespace lxv {

typedef unsigned __int128 u128;

// ---------------------------------------------------------------------------
// SHA-256 / SHA-512 (FIPS 180-4)

static const uint32_t K256[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

static inline uint32_t rr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
static inline uint64_t rr64(uint64_t x, int n) { return (x >> n) | (x << (64 - n)); }

static void sha256_compress(uint32_t s[8], const uint8_t* p) {
    uint32_t w[64];
    for (int i = 0; i < 16; i++)
        w[i] = (uint32_t)p[4 * i] << 24 | (uint32_t)p[4 * i + 1] << 16 | (uint32_t)p[4 * i + 2] << 8 | p[4 * i + 3];
    for (int i = 16; i < 64; i++)
        w[i] = w[i - 16] + (rr32(w[i - 15], 7) ^ rr32(w[i - 15], 18) ^ (w[i - 15] >> 3)) + w[i - 7] +
               (rr32(w[i - 2], 17) ^ rr32(w[i - 2], 19) ^ (w[i - 2] >> 10));
    uint32_t a = s[0], b = s[1], c = s[2], d = s[3], e = s[4], f = s[5], g = s[6], h = s[7];
    for (int i = 0; i < 64; i++) {
        uint32_t t1 = h + (rr32(e, 6) ^ rr32(e, 11) ^ rr32(e, 25)) + ((e & f) ^ (~e & g)) + K256[i] + w[i];
        uint32_t t2 = (rr32(a, 2) ^ rr32(a, 13) ^ rr32(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
        h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    s[0] += a; s[1] += b; s[2] += c; s[3] += d; s[4] += e; s[5] += f; s[6] += g; s[7] += h;
}

void sha256(const uint8_t* msg, size_t len, uint8_t out[32]) {
    uint32_t s[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    size_t i = 0;
    for (; i + 64 <= len; i += 64) sha256_compress(s, msg + i);
    uint8_t tail[128] = {0};
    size_t rem = len - i;
    std::memcpy(tail, msg + i, rem);
    tail[rem] = 0x80;
    size_t tl = rem + 9 <= 64 ? 64 : 128;
    uint64_t bits = (uint64_t)len * 8;
    for (int k = 0; k < 8; k++) tail[tl - 1 - k] = (uint8_t)(bits >> (8 * k));
    sha256_compress(s, tail);
    if (tl == 128) sha256_compress(s, tail + 64);
    for (int k = 0; k < 8; k++) {
        out[4 * k] = s[k] >> 24; out[4 * k + 1] = s[k] >> 16; out[4 * k + 2] = s[k] >> 8; out[4 * k + 3] = s[k];
    }
}

static const uint64_t K512[80] = {
    0x428a2f98d728ae22ULL, 0x7137449123ef65cdULL, 0xb5c0fbcfec4d3b2fULL, 0xe9b5dba58189dbbcULL, 0x3956c25bf348b538ULL,
    0x59f111f1b605d019ULL, 0x923f82a4af194f9bULL, 0xab1c5ed5da6d8118ULL, 0xd807aa98a3030242ULL, 0x12835b0145706fbeULL,
    0x243185be4ee4b28cULL, 0x550c7dc3d5ffb4e2ULL, 0x72be5d74f27b896fULL, 0x80deb1fe3b1696b1ULL, 0x9bdc06a725c71235ULL,
    0xc19bf174cf692694ULL, 0xe49b69c19ef14ad2ULL, 0xefbe4786384f25e3ULL, 0x0fc19dc68b8cd5b5ULL, 0x240ca1cc77ac9c65ULL,
    0x2de92c6f592b0275ULL, 0x4a7484aa6ea6e483ULL, 0x5cb0a9dcbd41fbd4ULL, 0x76f988da831153b5ULL, 0x983e5152ee66dfabULL,
    0xa831c66d2db43210ULL, 0xb00327c898fb213fULL, 0xbf597fc7beef0ee4ULL, 0xc6e00bf33da88fc2ULL, 0xd5a79147930aa725ULL,
    0x06ca6351e003826fULL, 0x142929670a0e6e70ULL, 0x27b70a8546d22ffcULL, 0x2e1b21385c26c926ULL, 0x4d2c6dfc5ac42aedULL,
    0x53380d139d95b3dfULL, 0x650a73548baf63deULL, 0x766a0abb3c77b2a8ULL, 0x81c2c92e47edaee6ULL, 0x92722c851482353bULL,
    0xa2bfe8a14cf10364ULL, 0xa81a664bbc423001ULL, 0xc24b8b70d0f89791ULL, 0xc76c51a30654be30ULL, 0xd192e819d6ef5218ULL,
    0xd69906245565a910ULL, 0xf40e35855771202aULL, 0x106aa07032bbd1b8ULL, 0x19a4c116b8d2d0c8ULL, 0x1e376c085141ab53ULL,
    0x2748774cdf8eeb99ULL, 0x34b0bcb5e19b48a8ULL, 0x391c0cb3c5c95a63ULL, 0x4ed8aa4ae3418acbULL, 0x5b9cca4f7763e373ULL,
    0x682e6ff3d6b2b8a3ULL, 0x748f82ee5defb2fcULL, 0x78a5636f43172f60ULL, 0x84c87814a1f0ab72ULL, 0x8cc702081a6439ecULL,
    0x90befffa23631e28ULL, 0xa4506cebde82bde9ULL, 0xbef9a3f7b2c67915ULL, 0xc67178f2e372532bULL, 0xca273eceea26619cULL,
    0xd186b8c721c0c207ULL, 0xeada7dd6cde0eb1eULL, 0xf57d4f7fee6ed178ULL, 0x06f067aa72176fbaULL, 0x0a637dc5a2c898a6ULL,
    0x113f9804bef90daeULL, 0x1b710b35131c471bULL, 0x28db77f523047d84ULL, 0x32caab7b40c72493ULL, 0x3c9ebe0a15c9bebcULL,
    0x431d67c49c100d4cULL, 0x4cc5d4becb3e42b6ULL, 0x597f299cfc657e2aULL, 0x5fcb6fab3ad6faecULL, 0x6c44198c4a475817ULL};

static void sha512_compress(uint64_t s[8], const uint8_t* p) {
    uint64_t w[80];
    for (int i = 0; i < 16; i++) {
        uint64_t v = 0;
        for (int k = 0; k < 8; k++) v = v << 8 | p[8 * i + k];
        w[i] = v;
    }
    for (int i = 16; i < 80; i++)
        w[i] = w[i - 16] + (rr64(w[i - 15], 1) ^ rr64(w[i - 15], 8) ^ (w[i - 15] >> 7)) + w[i - 7] +
               (rr64(w[i - 2], 19) ^ rr64(w[i - 2], 61) ^ (w[i - 2] >> 6));
    uint64_t a = s[0], b = s[1], c = s[2], d = s[3], e = s[4], f = s[5], g = s[6], h = s[7];
    for (int i = 0; i < 80; i++) {
        uint64_t t1 = h + (rr64(e, 14) ^ rr64(e, 18) ^ rr64(e, 41)) + ((e & f) ^ (~e & g)) + K512[i] + w[i];
        uint64_t t2 = (rr64(a, 28) ^ rr64(a, 34) ^ rr64(a, 39)) + ((a & b) ^ (a & c) ^ (b & c));
        h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    s[0] += a; s[1] += b; s[2] += c; s[3] += d; s[4] += e; s[5] += f; s[6] += g; s[7] += h;
}

void sha512(const uint8_t* msg, size_t len, uint8_t out[64]) {
    uint64_t s[8] = {0x6a09e667f3bcc908ULL, 0xbb67ae8584caa73bULL, 0x3c6ef372fe94f82bULL, 0xa54ff53a5f1d36f1ULL,
                     0x510e527fade682d1ULL, 0x9b05688c2b3e6c1fULL, 0x1f83d9abfb41bd6bULL, 0x5be0cd19137e2179ULL};
    size_t i = 0;
    for (; i + 128 <= len; i += 128) sha512_compress(s, msg + i);
    uint8_t tail[256] = {0};
    size_t rem = len - i;
    std::memcpy(tail, msg + i, rem);
    tail[rem] = 0x80;
    size_t tl = rem + 17 <= 128 ? 128 : 256;
    uint64_t bits = (uint64_t)len * 8;
    for (int k = 0; k < 8; k++) tail[tl - 1 - k] = (uint8_t)(bits >> (8 * k));
    sha512_compress(s, tail);
    if (tl == 256) sha512_compress(s, tail + 128);
    for (int k = 0; k < 8; k++)
        for (int j = 0; j < 8; j++) out[8 * k + j] = (uint8_t)(s[k] >> (56 - 8 * j));
}

// ---------------------------------------------------------------------------
// GF(2^255 - 19): 4 x 64-bit limbs holding any value below 2^256 that is
// congruent to the element; 2^256 = 38 (mod p) folds the top.

struct F {
    uint64_t v[4];
};

static inline F f_from_u64(uint64_t x) { return F{{x, 0, 0, 0}}; }

static inline F f_frombytes(const uint8_t b[32]) {
    F r;
    for (int i = 0; i < 4; i++) {
        uint64_t x = 0;
        for (int k = 7; k >= 0; k--) x = x << 8 | b[8 * i + k];
        r.v[i] = x;
    }
    r.v[3] &= 0x7fffffffffffffffULL;  // bit 255 dropped (RFC 7748 5)
    return r;
}

// r = t + 38 c, where t < 2^256 and c * 38 < 2^64; result below 2^256.
static inline void f_fold(uint64_t t[4], uint64_t c) {
    u128 acc = (u128)c * 38;
    for (int i = 0; i < 4; i++) {
        acc += t[i];
        t[i] = (uint64_t)acc;
        acc >>= 64;
    }
    t[0] += (uint64_t)acc * 38;  // a wrap leaves t[0] < 38 * 39: no carry
}

static inline F f_add(const F& a, const F& b) {
    F r;
    u128 acc = 0;
    for (int i = 0; i < 4; i++) {
        acc += (u128)a.v[i] + b.v[i];
        r.v[i] = (uint64_t)acc;
        acc >>= 64;
    }
    f_fold(r.v, (uint64_t)acc);
    return r;
}

static inline F f_sub(const F& a, const F& b) {
    F r;
    unsigned borrow = 0;
    for (int i = 0; i < 4; i++) {
        uint64_t d1, d2;
        unsigned b1 = __builtin_sub_overflow(a.v[i], b.v[i], &d1);
        unsigned b2 = __builtin_sub_overflow(d1, (uint64_t)borrow, &d2);
        r.v[i] = d2;
        borrow = b1 | b2;
    }
    // a - b + 2^256 = a - b + 38 (mod p): take 38 away for each wrap.
    for (int pass = 0; pass < 2 && borrow; pass++) {
        uint64_t sub = 38;
        borrow = 0;
        for (int i = 0; i < 4; i++) {
            uint64_t d1, d2;
            unsigned b1 = __builtin_sub_overflow(r.v[i], sub, &d1);
            unsigned b2 = __builtin_sub_overflow(d1, (uint64_t)borrow, &d2);
            r.v[i] = d2;
            borrow = b1 | b2;
            sub = 0;
        }
    }
    return r;
}

static inline F f_mul(const F& a, const F& b) {
    uint64_t t[8] = {0};
    for (int i = 0; i < 4; i++) {
        u128 c = 0;
        for (int j = 0; j < 4; j++) {
            c += (u128)a.v[i] * b.v[j] + t[i + j];
            t[i + j] = (uint64_t)c;
            c >>= 64;
        }
        t[i + 4] = (uint64_t)c;
    }
    F r;
    u128 c = 0;
    for (int i = 0; i < 4; i++) {
        c += (u128)t[i + 4] * 38 + t[i];
        r.v[i] = (uint64_t)c;
        c >>= 64;
    }
    f_fold(r.v, (uint64_t)c);
    return r;
}

static inline F f_sq(const F& a) { return f_mul(a, a); }

static void f_tobytes(uint8_t out[32], const F& a) {
    uint64_t t[4] = {a.v[0], a.v[1], a.v[2], a.v[3]};
    // value = low 255 bits + 19 * bit255 < 2^255 + 19
    uint64_t top = t[3] >> 63;
    t[3] &= 0x7fffffffffffffffULL;
    u128 acc = (u128)top * 19;
    for (int i = 0; i < 4; i++) {
        acc += t[i];
        t[i] = (uint64_t)acc;
        acc >>= 64;
    }
    // subtract p if value >= p, i.e. if value + 19 reaches 2^255
    uint64_t s[4];
    acc = 19;
    for (int i = 0; i < 4; i++) {
        acc += t[i];
        s[i] = (uint64_t)acc;
        acc >>= 64;
    }
    uint64_t ge = s[3] >> 63;
    s[3] &= 0x7fffffffffffffffULL;
    uint64_t m = 0 - ge;
    for (int i = 0; i < 4; i++) t[i] = (s[i] & m) | (t[i] & ~m);
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 8; k++) out[8 * i + k] = (uint8_t)(t[i] >> (8 * k));
}

static bool f_equal(const F& a, const F& b) {
    uint8_t x[32], y[32];
    f_tobytes(x, a);
    f_tobytes(y, b);
    return std::memcmp(x, y, 32) == 0;
}

// a^e, e a 256-bit little-endian exponent (public exponents only).
static F f_pow(const F& a, const uint8_t e[32]) {
    F r = f_from_u64(1);
    for (int bit = 255; bit >= 0; bit--) {
        r = f_sq(r);
        if ((e[bit >> 3] >> (bit & 7)) & 1) r = f_mul(r, a);
    }
    return r;
}

static const uint8_t P_MINUS_2[32] = {0xeb, 0xff, 0xff, 0xff, 0xff, 0xff, 0xff, 0xff, 0xff, 0xff, 0xff,
                                      0xff, 0xff, 0xff, 0xff, 0xff, 0xff, 0xff, 0xff, 0xff, 0xff, 0xff,
                                      0xff, 0xff, 0xff, 0xff, 0xff, 0xff, 0xff, 0xff, 0xff, 0x7f};

static inline F f_inv(const F& a) { return f_pow(a, P_MINUS_2); }

static inline void f_cswap(F& a, F& b, uint64_t swap) {
    uint64_t m = 0 - swap;
    for (int i = 0; i < 4; i++) {
        uint64_t t = m & (a.v[i] ^ b.v[i]);
        a.v[i] ^= t;
        b.v[i] ^= t;
    }
}

// Montgomery ladder (RFC 7748 5) over bits top..0 of k.
static void ladder(uint8_t out[32], const uint8_t k[32], const uint8_t u[32], int top) {
    const F x1 = f_frombytes(u);
    F x2 = f_from_u64(1), z2 = f_from_u64(0), x3 = x1, z3 = f_from_u64(1);
    const F a24 = f_from_u64(121665);
    uint64_t swap = 0;
    for (int i = top; i >= 0; i--) {
        uint64_t bit = (k[i >> 3] >> (i & 7)) & 1;
        swap ^= bit;
        f_cswap(x2, x3, swap);
        f_cswap(z2, z3, swap);
        swap = bit;
        F a = f_add(x2, z2), aa = f_sq(a), b = f_sub(x2, z2), bb = f_sq(b), e = f_sub(aa, bb);
        F c = f_add(x3, z3), d = f_sub(x3, z3), da = f_mul(d, a), cb = f_mul(c, b);
        x3 = f_sq(f_add(da, cb));
        z3 = f_mul(x1, f_sq(f_sub(da, cb)));
        x2 = f_mul(aa, bb);
        z2 = f_mul(e, f_add(aa, f_mul(a24, e)));
    }
    f_cswap(x2, x3, swap);
    f_cswap(z2, z3, swap);
    f_tobytes(out, f_mul(x2, f_inv(z2)));
}

void clamp(uint8_t k[32]) {
    k[0] &= 248;
    k[31] &= 127;
    k[31] |= 64;
}

void x25519(const uint8_t k[32], const uint8_t u[32], uint8_t out[32]) {
    uint8_t kc[32];
    std::memcpy(kc, k, 32);
    clamp(kc);
    ladder(out, kc, u, 254);
}

static const uint8_t BASE_U[32] = {9};

void x25519_base(const uint8_t k[32], uint8_t out[32]) { x25519(k, BASE_U, out); }

void mont_mul(const uint8_t k[32], const uint8_t u[32], uint8_t out[32]) { ladder(out, k, u, 255); }

// ---------------------------------------------------------------------------
// Ed25519 public keys: twisted Edwards -x^2 + y^2 = 1 + d x^2 y^2, extended
// coordinates (X : Y : Z : T), complete addition for a = -1.

namespace {
struct Pt {
    F X, Y, Z, T;
};
struct EdConst {
    F d2;
    Pt base;
};
}  // namespace

static F f_neg(const F& a) { return f_sub(f_from_u64(0), a); }

static EdConst ed_const() {
    EdConst k;
    const F one = f_from_u64(1);
    F d = f_neg(f_mul(f_from_u64(121665), f_inv(f_from_u64(121666))));
    k.d2 = f_add(d, d);
    F y = f_mul(f_from_u64(4), f_inv(f_from_u64(5)));
    F y2 = f_sq(y);
    F x2 = f_mul(f_sub(y2, one), f_inv(f_add(f_mul(d, y2), one)));
    // x = x2^((p+3)/8), times sqrt(-1) if needed; the even root is B's.
    uint8_t e[32];
    std::memset(e, 0xff, 32);
    e[0] = 0xfe;
    e[31] = 0x0f;
    F x = f_pow(x2, e);
    if (!f_equal(f_sq(x), x2)) {
        uint8_t q[32];
        std::memset(q, 0xff, 32);
        q[0] = 0xfb;
        q[31] = 0x1f;
        x = f_mul(x, f_pow(f_from_u64(2), q));
    }
    uint8_t xb[32];
    f_tobytes(xb, x);
    if (xb[0] & 1) x = f_neg(x);
    k.base = Pt{x, y, one, f_mul(x, y)};
    return k;
}

static Pt pt_add(const Pt& p, const Pt& q, const F& d2) {
    F a = f_mul(f_sub(p.Y, p.X), f_sub(q.Y, q.X));
    F b = f_mul(f_add(p.Y, p.X), f_add(q.Y, q.X));
    F c = f_mul(f_mul(p.T, q.T), d2);
    F d = f_mul(p.Z, q.Z);
    d = f_add(d, d);
    F e = f_sub(b, a), f = f_sub(d, c), g = f_add(d, c), h = f_add(b, a);
    return Pt{f_mul(e, f), f_mul(g, h), f_mul(f, g), f_mul(e, h)};
}

static void pt_cmov(Pt& r, const Pt& p, uint64_t flag) {
    uint64_t m = 0 - flag;
    F* rf[4] = {&r.X, &r.Y, &r.Z, &r.T};
    const F* pf[4] = {&p.X, &p.Y, &p.Z, &p.T};
    for (int k = 0; k < 4; k++)
        for (int i = 0; i < 4; i++) rf[k]->v[i] ^= m & (rf[k]->v[i] ^ pf[k]->v[i]);
}

void ed25519_public(const uint8_t seed[32], uint8_t out[32]) {
    static const EdConst K = ed_const();
    uint8_t h[64];
    sha512(seed, 32, h);
    clamp(h);
    // Double and always add, selecting the sum by the scalar bit.
    Pt acc{f_from_u64(0), f_from_u64(1), f_from_u64(1), f_from_u64(0)};
    for (int i = 254; i >= 0; i--) {
        acc = pt_add(acc, acc, K.d2);
        Pt s = pt_add(acc, K.base, K.d2);
        pt_cmov(acc, s, (h[i >> 3] >> (i & 7)) & 1);
    }
    F zi = f_inv(acc.Z);
    uint8_t xb[32];
    f_tobytes(out, f_mul(acc.Y, zi));
    f_tobytes(xb, f_mul(acc.X, zi));
    out[31] |= (uint8_t)((xb[0] & 1) << 7);
    std::memset(h, 0, sizeof h);
}

// ---------------------------------------------------------------------------
// Scalars, addresses, prefixes.

void step_scalar(const uint8_t k[32], uint64_t n, uint8_t out[32]) {
    std::memcpy(out, k, 32);
    u128 carry = (u128)n * 8;
    for (int i = 0; i < 32 && carry; i++) {
        u128 sum = out[i] + (carry & 0xff);
        out[i] = (uint8_t)sum;
        carry = (carry >> 8) + (sum >> 8);
    }
}

void previous_scalar(const uint8_t k[32], uint8_t out[32]) {
    std::memcpy(out, k, 32);
    int borrow = 8;
    for (int i = 0; i < 32 && borrow; i++) {
        int v = (int)out[i] - borrow;
        out[i] = (uint8_t)(v & 0xff);
        borrow = v < 0;
    }
}

bool start_scalar_ok(const uint8_t k[32]) {
    if ((k[0] & 7) != 0 || (k[31] & 0xc0) != 0x40) return false;
    bool high = k[31] == 127, low = k[31] == 64;
    for (int i = 9; i < 31; i++) high = high && k[i] == 0xff;
    for (int i = 0; i < 31; i++) low = low && k[i] == 0;
    return !high && !low;
}

bool random_bytes(uint8_t* out, size_t n) {
    size_t got = 0;
    while (got < n) {
        ssize_t r = getrandom(out + got, n - got, 0);
        if (r <= 0) return false;
        got += (size_t)r;
    }
    return true;
}

void name_hash(uint8_t out[10]) {
    static const char aspects[] = "lxmf.delivery";
    uint8_t h[32];
    sha256((const uint8_t*)aspects, sizeof aspects - 1, h);
    std::memcpy(out, h, 10);
}

void lxmf_address_nh(const uint8_t nh[10], const uint8_t x_pub[32], const uint8_t ed_pub[32], uint8_t out[16]) {
    uint8_t pk[64], ih[32], material[26], a[32];
    std::memcpy(pk, x_pub, 32);
    std::memcpy(pk + 32, ed_pub, 32);
    sha256(pk, 64, ih);
    std::memcpy(material, nh, 10);
    std::memcpy(material + 10, ih, 16);
    sha256(material, 26, a);
    std::memcpy(out, a, 16);
}

void lxmf_address(const uint8_t x_pub[32], const uint8_t ed_pub[32], uint8_t out[16]) {
    uint8_t nh[10];
    name_hash(nh);
    lxmf_address_nh(nh, x_pub, ed_pub, out);
}

void address_of_identity(const uint8_t identity[64], uint8_t out[16]) {
    uint8_t x_pub[32], ed_pub[32];
    x25519_base(identity, x_pub);
    ed25519_public(identity + 32, ed_pub);
    lxmf_address(x_pub, ed_pub, out);
}

static int hexval(char c) {
    if (c >= '0' && c <= '9') return c - '0';
    if (c >= 'a' && c <= 'f') return c - 'a' + 10;
    if (c >= 'A' && c <= 'F') return c - 'A' + 10;
    return -1;
}

bool Prefix::parse(const std::string& s, Prefix* out, std::string* err) {
    Prefix p;
    bool ok = !s.empty() && s.size() <= 32;
    for (size_t i = 0; ok && i < s.size(); i++) {
        int v = hexval(s[i]);
        if (v < 0) ok = false;
        else p.bytes[i / 2] |= (uint8_t)(v << (i % 2 ? 0 : 4));
    }
    if (!ok) {
        if (err) {
            std::string low = s;
            for (auto& c : low) c = (char)std::tolower((unsigned char)c);
            *err = "'" + low + "' isn't 1 to 32 hex digits: use 0-9 and a-f only (zero, not the letter O).";
        }
        return false;
    }
    p.nibbles = (int)s.size();
    *out = p;
    return true;
}

bool Prefix::matches(const uint8_t addr[16]) const {
    int whole = nibbles / 2;
    if (std::memcmp(addr, bytes, (size_t)whole) != 0) return false;
    return nibbles % 2 == 0 || (addr[whole] >> 4) == (bytes[whole] >> 4);
}

void Prefix::words(uint32_t want[4], uint32_t mask[4]) const {
    uint8_t w[16] = {0}, m[16] = {0};
    for (int i = 0; i < nibbles; i++) {
        int shift = i % 2 == 0 ? 4 : 0;
        m[i / 2] |= (uint8_t)(0xF << shift);
        w[i / 2] |= (uint8_t)(((bytes[i / 2] >> shift) & 0xF) << shift);
    }
    for (int i = 0; i < 4; i++) {
        want[i] = (uint32_t)w[4 * i] << 24 | (uint32_t)w[4 * i + 1] << 16 | (uint32_t)w[4 * i + 2] << 8 | w[4 * i + 3];
        mask[i] = (uint32_t)m[4 * i] << 24 | (uint32_t)m[4 * i + 1] << 16 | (uint32_t)m[4 * i + 2] << 8 | m[4 * i + 3];
    }
}

// ---------------------------------------------------------------------------
// The seeded workload generator.

void drbg(uint64_t seed, const char* label, uint32_t counter, uint8_t out[32]) {
    static const char tag[] = "lxmf-vanity-amd/v1";
    uint8_t msg[128];
    size_t n = 0;
    std::memcpy(msg, tag, sizeof tag);  // includes the NUL
    n += sizeof tag;
    for (int i = 0; i < 8; i++) msg[n++] = (uint8_t)(seed >> (8 * i));
    size_t ll = std::strlen(label);
    if (ll > 64) ll = 64;
    std::memcpy(msg + n, label, ll);
    n += ll;
    msg[n++] = 0;
    for (int i = 0; i < 4; i++) msg[n++] = (uint8_t)(counter >> (8 * i));
    sha256(msg, n, out);
}

void seeded_start(uint64_t seed, uint8_t out[32]) {
    for (uint32_t c = 0;; c++) {
        drbg(seed, "start", c, out);
        clamp(out);
        if (start_scalar_ok(out)) return;
    }
}

void seeded_ed_seed(uint64_t seed, uint8_t out[32]) { drbg(seed, "ed25519", 0, out); }

std::string seeded_prefix(uint64_t seed, int nibbles) {
    uint8_t b[32];
    drbg(seed, "prefix", 0, b);
    std::string h = hex(b, 32);
    if (nibbles < 0) nibbles = 0;
    if (nibbles > 64) nibbles = 64;
    return h.substr(0, (size_t)nibbles);
}

// ---------------------------------------------------------------------------
// The CPU key walk.

static inline F f_load(const uint64_t v[4]) { return F{{v[0], v[1], v[2], v[3]}}; }
static inline void f_store(uint64_t v[4], const F& a) {
    for (int i = 0; i < 4; i++) v[i] = a.v[i];
}

KeyWalk::KeyWalk(const uint8_t start[32]) {
    std::memcpy(start_, start, 32);
    uint8_t prev[32], u[32], q[32], eight[32] = {8};
    previous_scalar(start, prev);
    x25519_base(prev, u);
    f_store(prev_x_, f_frombytes(u));
    x25519_base(start, u);
    f_store(cur_x_, f_frombytes(u));
    f_store(prev_z_, f_from_u64(1));
    f_store(cur_z_, f_from_u64(1));
    mont_mul(eight, BASE_U, q);  // u(8B): the walk's fixed step
    const F qf = f_frombytes(q), one = f_from_u64(1);
    f_store(q_plus_, f_add(qf, one));
    f_store(q_minus_, f_sub(qf, one));
}

uint64_t KeyWalk::next_batch(uint8_t (*out)[32], size_t len) {
    if (len > kBatch) len = kBatch;
    const uint64_t first = next_n_;
    F xs[kBatch], zs[kBatch], pre[kBatch];
    F xp = f_load(prev_x_), zp = f_load(prev_z_), x = f_load(cur_x_), z = f_load(cur_z_);
    const F qp = f_load(q_plus_), qm = f_load(q_minus_);
    for (size_t i = 0; i < len; i++) {
        xs[i] = x;
        zs[i] = z;
        // P(n+1) = P(n) + Q with P(n) - Q = P(n-1), Q = 8B affine
        const F t1 = f_mul(f_sub(x, z), qp), t2 = f_mul(f_add(x, z), qm);
        const F xn = f_mul(zp, f_sq(f_add(t1, t2)));
        const F zn = f_mul(xp, f_sq(f_sub(t1, t2)));
        xp = x;
        zp = z;
        x = xn;
        z = zn;
    }
    // one inversion for the batch (Montgomery's trick)
    F acc = f_from_u64(1);
    for (size_t i = 0; i < len; i++) {
        pre[i] = acc;
        acc = f_mul(acc, zs[i]);
    }
    F inv = f_inv(acc);
    for (size_t i = len; i-- > 0;) {
        const F zinv = f_mul(inv, pre[i]);
        inv = f_mul(inv, zs[i]);
        f_tobytes(out[i], f_mul(xs[i], zinv));
    }
    f_store(prev_x_, xp);
    f_store(prev_z_, zp);
    f_store(cur_x_, x);
    f_store(cur_z_, z);
    next_n_ += len;
    return first;
}

std::string hex(const uint8_t* b, size_t n) {
    static const char* d = "0123456789abcdef";
    std::string s;
    s.reserve(2 * n);
    for (size_t i = 0; i < n; i++) {
        s.push_back(d[b[i] >> 4]);
        s.push_back(d[b[i] & 15]);
    }
    return s;
}

}  // namespace lxv
