// GF(2^255 - 19) in eight 32-bit limbs, radix 2^32: a value below 2^256
// that is congruent to the element, reduced with 2^256 = 38 (mod p). The
// layout chosen for gfx950's 32-bit VALU: a multiplication is 64 32 x 32 -> 64
// products on v_mad_u64_u32 with carry-out chains (fe8_asm.inc, generated by
// gen_fe8.py), a squaring 43. Sums and differences fold their carry or
// borrow back in at once, so every value stays below 2^256.
#pragma once

#include "lxv_common.h"

namespace lxv {

struct Fe8 {
    uint32_t v[8];
};

#if defined(__HIP_DEVICE_COMPILE__)
#include "fe8_asm.inc"
#endif

LXV_HD uint32_t addc32(uint32_t a, uint32_t b, uint32_t cin, uint32_t* cout) {
#if defined(__clang__)
    unsigned c;
    const uint32_t r = __builtin_addc(a, b, cin, &c);
    *cout = c;
    return r;
#else
    const uint64_t s = (uint64_t)a + b + cin;
    *cout = (uint32_t)(s >> 32);
    return (uint32_t)s;
#endif
}

LXV_HD uint32_t subb32(uint32_t a, uint32_t b, uint32_t bin, uint32_t* bout) {
#if defined(__clang__)
    unsigned c;
    const uint32_t r = __builtin_subc(a, b, bin, &c);
    *bout = c;
    return r;
#else
    const uint64_t s = (uint64_t)a - b - bin;
    *bout = (uint32_t)(s >> 63);
    return (uint32_t)s;
#endif
}

LXV_HD Fe8 fe8_zero() {
    Fe8 h;
#pragma unroll
    for (int i = 0; i < 8; i++) h.v[i] = 0;
    return h;
}

LXV_HD Fe8 fe8_one() {
    Fe8 h = fe8_zero();
    h.v[0] = 1;
    return h;
}

// r = t + 38 c (mod 2^256, with the wrap folded again): t < 2^256, c < 2^26.
LXV_HD void fe8_fold(uint32_t t[8], uint32_t c) {
    uint32_t carry;
    t[0] = addc32(t[0], c * 38u, 0, &carry);
#pragma unroll
    for (int i = 1; i < 8; i++) t[i] = addc32(t[i], 0, carry, &carry);
    // a wrap leaves t < 38 c < 2^32 - 38: one more 38 cannot carry
    t[0] += carry * 38u;
}

LXV_HD Fe8 fe8_add(const Fe8& a, const Fe8& b) {
    Fe8 r;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = addc32(a.v[i], b.v[i], c, &c);
    fe8_fold(r.v, c);
    return r;
}

// a - b: a borrow means 2^256 was added, i.e. 38 too many (mod p).
LXV_HD Fe8 fe8_sub(const Fe8& a, const Fe8& b) {
    Fe8 r;
    uint32_t bw = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = subb32(a.v[i], b.v[i], bw, &bw);
    uint32_t bw2;
    r.v[0] = subb32(r.v[0], bw * 38u, 0, &bw2);
#pragma unroll
    for (int i = 1; i < 8; i++) r.v[i] = subb32(r.v[i], 0, bw2, &bw2);
    // a second wrap leaves r >= 2^256 - 38: one more 38 cannot borrow
    r.v[0] -= bw2 * 38u;
    return r;
}

// 512-bit product to a value below 2^256 (host build of the multiply).
LXV_HD void fe8_reduce512(uint32_t out[8], const uint32_t w[16]) {
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint64_t t = (uint64_t)w[8 + i] * 38u + w[i] + c;
        out[i] = (uint32_t)t;
        c = t >> 32;
    }
    fe8_fold(out, (uint32_t)c);
}

LXV_HD Fe8 fe8_mul(const Fe8& a, const Fe8& b) {
    Fe8 r;
#if defined(__HIP_DEVICE_COMPILE__)
    fe8_mul_asm(r.v, a.v, b.v);
#else
    uint32_t w[16] = {0};
    for (int i = 0; i < 8; i++) {
        uint64_t c = 0;
        for (int j = 0; j < 8; j++) {
            const uint64_t t = (uint64_t)a.v[i] * b.v[j] + w[i + j] + c;
            w[i + j] = (uint32_t)t;
            c = t >> 32;
        }
        w[i + 8] = (uint32_t)c;
    }
    fe8_reduce512(r.v, w);
#endif
    return r;
}

LXV_HD Fe8 fe8_sq(const Fe8& a) {
#if defined(__HIP_DEVICE_COMPILE__)
    Fe8 r;
    fe8_sq_asm(r.v, a.v);
    return r;
#else
    return fe8_mul(a, a);
#endif
}

LXV_HD Fe8 fe8_sq_n(Fe8 f, int n) {
    for (int i = 0; i < n; i++) f = fe8_sq(f);
    return f;
}

LXV_HD Fe8 fe8_mul_small(const Fe8& a, uint32_t k) {
    Fe8 r;
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint64_t t = (uint64_t)a.v[i] * k + c;
        r.v[i] = (uint32_t)t;
        c = t >> 32;
    }
    fe8_fold(r.v, (uint32_t)c);
    return r;
}

LXV_HD Fe8 fe8_from_words(const uint32_t w[8]) {
    Fe8 h;
#pragma unroll
    for (int i = 0; i < 8; i++) h.v[i] = w[i];
    h.v[7] &= 0x7fffffffu;  // bit 255 dropped (RFC 7748 5)
    return h;
}

// Canonical encoding: value = low 255 bits + 19 * bit 255 < 2^255 + 19; then
// subtract p when value + 19 reaches 2^255.
LXV_HD void fe8_to_words(const Fe8& a, uint32_t out[8]) {
    uint32_t t[8], s[8];
    const uint32_t top = a.v[7] >> 31;
    uint32_t c;
    t[0] = addc32(a.v[0], top * 19u, 0, &c);
#pragma unroll
    for (int i = 1; i < 8; i++) t[i] = addc32(i == 7 ? (a.v[7] & 0x7fffffffu) : a.v[i], 0, c, &c);
    s[0] = addc32(t[0], 19u, 0, &c);
#pragma unroll
    for (int i = 1; i < 8; i++) s[i] = addc32(t[i], 0, c, &c);
    const uint32_t ge = s[7] >> 31;
    s[7] &= 0x7fffffffu;
    const uint32_t m = 0u - ge;
#pragma unroll
    for (int i = 0; i < 8; i++) out[i] = (s[i] & m) | (t[i] & ~m);
}

// z^(p-2) by the ref10 addition chain (254 squarings, 11 multiplications):
// the inversion of the reference's walk (src/gpu/common.h:109-128), kept for
// comparison with fe8_invert below.
LXV_HD Fe8 fe8_invert_chain(const Fe8& z) {
    const Fe8 z2 = fe8_sq(z);
    const Fe8 z9 = fe8_mul(fe8_sq_n(z2, 2), z);
    const Fe8 z11 = fe8_mul(z9, z2);
    const Fe8 z2_5_0 = fe8_mul(fe8_sq(z11), z9);
    const Fe8 z2_10_0 = fe8_mul(fe8_sq_n(z2_5_0, 5), z2_5_0);
    const Fe8 z2_20_0 = fe8_mul(fe8_sq_n(z2_10_0, 10), z2_10_0);
    const Fe8 z2_40_0 = fe8_mul(fe8_sq_n(z2_20_0, 20), z2_20_0);
    const Fe8 z2_50_0 = fe8_mul(fe8_sq_n(z2_40_0, 10), z2_10_0);
    const Fe8 z2_100_0 = fe8_mul(fe8_sq_n(z2_50_0, 50), z2_50_0);
    const Fe8 z2_200_0 = fe8_mul(fe8_sq_n(z2_100_0, 100), z2_100_0);
    const Fe8 z2_250_0 = fe8_mul(fe8_sq_n(z2_200_0, 50), z2_50_0);
    return fe8_mul(fe8_sq_n(z2_250_0, 5), z11);
}

// ---------------------------------------------------------------------------
// Inversion by divsteps (Bernstein and Yang, "Fast constant-time gcd
// computation and modular inversion", 2019), in the half-delta form: 20
// batches of 30 steps, 600 in all, above the 590 that bound every input below
// 2^256 for this form. The steps are the same instructions for every lane
// (no data-dependent branch), and a batch costs about 21 VALU operations per
// step on 32-bit words plus two 2x2 matrix products on signed 30-bit limbs:
// about a third of the 254 squarings of the addition chain.
//
// A batch works on the low 32 bits of f and g only (30 steps consume 30 low
// bits, and each step reads only bit 0 of g), accumulating the transition
// matrix (u v; q r) with 2^30 (f', g') = (u f + v g, q f + r g). The matrix
// is then applied to the full f, g (exact division by 2^30) and to the
// coefficients d, e with f = d z, g = e z (mod p), where the division by 2^30
// is made exact by adding a multiple of p chosen from p^-1 mod 2^30; adding
// u p (v p) when d (e) is negative keeps d, e in (-2p, p). At the end g = 0,
// f = +-1 and 1/z = d f.

struct S30 {
    int32_t v[9];  // v[0..7] in [0, 2^30) after an update, v[8] signed
};

constexpr uint32_t kM30 = (1u << 30) - 1u;
constexpr uint32_t kPInv30 = 0x179435e5u;  // p^-1 mod 2^30 (p = 2^255 - 19)

LXV_HD int32_t s30_p_limb(int i) { return i == 0 ? (int32_t)((1u << 30) - 19u) : (i == 8 ? 32767 : (int32_t)kM30); }

// 30 divsteps on the low words of f and g; returns zeta = -(delta + 1/2)
// after them and the matrix t = (u, v, q, r).
LXV_HD uint32_t fe8_divsteps30(uint32_t zeta, uint32_t f, uint32_t g, int32_t t[4]) {
    uint32_t u = 1, v = 0, q = 0, r = 1;
#pragma unroll
    for (int i = 0; i < 30; i++) {
        const uint32_t c1 = (uint32_t)((int32_t)zeta >> 31);  // delta > 0
        const uint32_t c2 = 0u - (g & 1u);                    // g odd
        const uint32_t n1 = c1 & 1u;
        // with c1: (f, g) -> (g, -f) first; with c2: g += f; then g /= 2
        const uint32_t x = (f ^ c1) + n1, y = (u ^ c1) + n1, z = (v ^ c1) + n1;
        g += x & c2;
        q += y & c2;
        r += z & c2;
        const uint32_t c3 = c1 & c2;
        zeta = (zeta ^ c3) - 1u;
        f += g & c3;
        u = (u + (q & c3)) << 1;
        v = (v + (r & c3)) << 1;
        g >>= 1;
    }
    t[0] = (int32_t)u;
    t[1] = (int32_t)v;
    t[2] = (int32_t)q;
    t[3] = (int32_t)r;
    return zeta;
}

// (f, g) <- (u f + v g, q f + r g) / 2^30.
LXV_HD void s30_update_fg(S30& f, S30& g, const int32_t t[4]) {
    int64_t cf = (int64_t)t[0] * f.v[0] + (int64_t)t[1] * g.v[0];
    int64_t cg = (int64_t)t[2] * f.v[0] + (int64_t)t[3] * g.v[0];
    cf >>= 30;
    cg >>= 30;
#pragma unroll
    for (int i = 1; i < 9; i++) {
        cf += (int64_t)t[0] * f.v[i] + (int64_t)t[1] * g.v[i];
        cg += (int64_t)t[2] * f.v[i] + (int64_t)t[3] * g.v[i];
        f.v[i - 1] = (int32_t)((uint32_t)cf & kM30);
        g.v[i - 1] = (int32_t)((uint32_t)cg & kM30);
        cf >>= 30;
        cg >>= 30;
    }
    f.v[8] = (int32_t)cf;
    g.v[8] = (int32_t)cg;
}

// (d, e) <- (u d + v e, q d + r e) / 2^30 (mod p), kept in (-2p, p).
LXV_HD void s30_update_de(S30& d, S30& e, const int32_t t[4]) {
    const int32_t u = t[0], v = t[1], q = t[2], r = t[3];
    const int32_t sd = d.v[8] >> 31, se = e.v[8] >> 31;
    int32_t md = (u & sd) + (v & se);
    int32_t me = (q & sd) + (r & se);
    int64_t cd = (int64_t)u * d.v[0] + (int64_t)v * e.v[0];
    int64_t ce = (int64_t)q * d.v[0] + (int64_t)r * e.v[0];
    md -= (int32_t)((kPInv30 * (uint32_t)cd + (uint32_t)md) & kM30);
    me -= (int32_t)((kPInv30 * (uint32_t)ce + (uint32_t)me) & kM30);
    cd += (int64_t)s30_p_limb(0) * md;
    ce += (int64_t)s30_p_limb(0) * me;
    cd >>= 30;
    ce >>= 30;
#pragma unroll
    for (int i = 1; i < 9; i++) {
        cd += (int64_t)u * d.v[i] + (int64_t)v * e.v[i] + (int64_t)s30_p_limb(i) * md;
        ce += (int64_t)q * d.v[i] + (int64_t)r * e.v[i] + (int64_t)s30_p_limb(i) * me;
        d.v[i - 1] = (int32_t)((uint32_t)cd & kM30);
        e.v[i - 1] = (int32_t)((uint32_t)ce & kM30);
        cd >>= 30;
        ce >>= 30;
    }
    d.v[8] = (int32_t)cd;
    e.v[8] = (int32_t)ce;
}

// Carries every limb into [0, 2^30) except the signed top one.
LXV_HD void s30_normalize(S30& a) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
        a.v[i + 1] += a.v[i] >> 30;
        a.v[i] = (int32_t)((uint32_t)a.v[i] & kM30);
    }
}

// a + p if a < 0 (a normalized), normalized again.
LXV_HD void s30_add_p_if_negative(S30& a) {
    const int32_t m = a.v[8] >> 31;
#pragma unroll
    for (int i = 0; i < 9; i++) a.v[i] += s30_p_limb(i) & m;
    s30_normalize(a);
}

LXV_HD Fe8 fe8_invert(const Fe8& z) {
    uint32_t w[8];
    fe8_to_words(z, w);  // canonical, so g < p
    S30 f, g, d, e;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const int k = (30 * i) / 32, s = (30 * i) % 32;
        const uint32_t lo = w[k], hi = k + 1 < 8 ? w[k + 1] : 0u;
        g.v[i] = (int32_t)(funnel32(hi, lo, (uint32_t)s) & kM30);
        f.v[i] = s30_p_limb(i);
        d.v[i] = 0;
        e.v[i] = i == 0 ? 1 : 0;
    }
    uint32_t zeta = 0xffffffffu;  // delta = 1/2
    for (int b = 0; b < 20; b++) {
        int32_t t[4];
        zeta = fe8_divsteps30(zeta, (uint32_t)f.v[0] | ((uint32_t)f.v[1] << 30),
                              (uint32_t)g.v[0] | ((uint32_t)g.v[1] << 30), t);
        s30_update_de(d, e, t);
        s30_update_fg(f, g, t);
    }
    // 1/z = d f with f = +-1: negate d when f < 0, then bring it into [0, 2p)
    const int32_t neg = f.v[8] >> 31;
#pragma unroll
    for (int i = 0; i < 9; i++) d.v[i] = (d.v[i] ^ neg) - neg;
    s30_normalize(d);
    s30_add_p_if_negative(d);
    s30_add_p_if_negative(d);
    Fe8 r;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int i = (32 * j) / 30, s = (32 * j) % 30;
        const uint32_t lo = (uint32_t)d.v[i] >> s;
        const uint32_t hi = i + 1 < 9 ? (uint32_t)d.v[i + 1] << (30 - s) : 0u;
        r.v[j] = lo | hi;
    }
    return r;
}

struct Field8 {
    typedef Fe8 fe;
    static constexpr int kLimbs = 8;
    LXV_HD static fe zero() { return fe8_zero(); }
    LXV_HD static fe one() { return fe8_one(); }
    LXV_HD static fe add(const fe& a, const fe& b) { return fe8_add(a, b); }
    LXV_HD static fe sub(const fe& a, const fe& b) { return fe8_sub(a, b); }
    // Every value is already below 2^256, a valid input and encoding.
    LXV_HD static fe tight(const fe& a) { return a; }
    LXV_HD static fe mul(const fe& a, const fe& b) { return fe8_mul(a, b); }
    LXV_HD static fe sq(const fe& a) { return fe8_sq(a); }
    LXV_HD static fe invert(const fe& a) { return fe8_invert(a); }
    LXV_HD static fe from_words(const uint32_t w[8]) { return fe8_from_words(w); }
    LXV_HD static void to_words(const fe& a, uint32_t out[8]) { fe8_to_words(a, out); }
    LXV_HD static fe mul_small(const fe& a, int32_t k) { return fe8_mul_small(a, (uint32_t)k); }
    LXV_HD static void cswap(fe& a, fe& b, uint32_t swap) {
        const uint32_t m = 0u - swap;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint32_t t = m & (a.v[i] ^ b.v[i]);
            a.v[i] ^= t;
            b.v[i] ^= t;
        }
    }
    LXV_HD static void store(const fe& a, uint32_t out[kLimbs]) {
#pragma unroll
        for (int i = 0; i < 8; i++) out[i] = a.v[i];
    }
    LXV_HD static fe load(const uint32_t in[kLimbs]) {
        fe a;
#pragma unroll
        for (int i = 0; i < 8; i++) a.v[i] = in[i];
        return a;
    }
};

}  // namespace lxv
