// GF(2^255 - 19) in eight 32-bit limbs, radix 2^32: a value below 2^256
// that is congruent to the element, reduced with 2^256 = 38 (mod p). The
// layout chosen for gfx950's 32-bit VALU: a multiplication is 64 32 x 32 -> 64
// products on v_mad_u64_u32 with carry-out chains (fe8_asm.inc, generated by
// gen_fe8.py), a squaring 43. Sums and differences fold their carry or
// borrow back in at once, so every value stays below 2^256.
#pragma once

#include "fe_divsteps.h"
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

// Sums and differences: one carry chain plus the fold of the carry (borrow)
// as 38, propagated by a second chain. On the device that is 19 VALU
// instructions, the chains carrying through VCC (written out because the
// compiler materialises the second chain's carries one by one); v_cndmask
// reads VCC two wait states after the VALU write, as gfx950 requires.
LXV_HD Fe8 fe8_add(const Fe8& a, const Fe8& b) {
#if defined(__HIP_DEVICE_COMPILE__)
    Fe8 r;
    uint32_t t;
    asm("v_add_co_u32_e32 %[r0], vcc, %[a0], %[b0]\n"
        "v_addc_co_u32_e32 %[r1], vcc, %[a1], %[b1], vcc\n"
        "v_addc_co_u32_e32 %[r2], vcc, %[a2], %[b2], vcc\n"
        "v_addc_co_u32_e32 %[r3], vcc, %[a3], %[b3], vcc\n"
        "v_addc_co_u32_e32 %[r4], vcc, %[a4], %[b4], vcc\n"
        "v_addc_co_u32_e32 %[r5], vcc, %[a5], %[b5], vcc\n"
        "v_addc_co_u32_e32 %[r6], vcc, %[a6], %[b6], vcc\n"
        "v_addc_co_u32_e32 %[r7], vcc, %[a7], %[b7], vcc\n"
        "s_nop 1\n"
        "v_cndmask_b32_e32 %[t], 0, %[k38], vcc\n"
        "v_add_co_u32_e32 %[r0], vcc, %[r0], %[t]\n"
        "v_addc_co_u32_e32 %[r1], vcc, 0, %[r1], vcc\n"
        "v_addc_co_u32_e32 %[r2], vcc, 0, %[r2], vcc\n"
        "v_addc_co_u32_e32 %[r3], vcc, 0, %[r3], vcc\n"
        "v_addc_co_u32_e32 %[r4], vcc, 0, %[r4], vcc\n"
        "v_addc_co_u32_e32 %[r5], vcc, 0, %[r5], vcc\n"
        "v_addc_co_u32_e32 %[r6], vcc, 0, %[r6], vcc\n"
        "v_addc_co_u32_e32 %[r7], vcc, 0, %[r7], vcc\n"
        "s_nop 1\n"
        "v_cndmask_b32_e32 %[t], 0, %[k38], vcc\n"
        "v_add_u32_e32 %[r0], %[r0], %[t]\n"
        : [r0] "=&v"(r.v[0]), [r1] "=&v"(r.v[1]), [r2] "=&v"(r.v[2]), [r3] "=&v"(r.v[3]), [r4] "=&v"(r.v[4]),
          [r5] "=&v"(r.v[5]), [r6] "=&v"(r.v[6]), [r7] "=&v"(r.v[7]), [t] "=&v"(t)
        : [a0] "v"(a.v[0]), [a1] "v"(a.v[1]), [a2] "v"(a.v[2]), [a3] "v"(a.v[3]), [a4] "v"(a.v[4]),
          [a5] "v"(a.v[5]), [a6] "v"(a.v[6]), [a7] "v"(a.v[7]), [b0] "v"(b.v[0]), [b1] "v"(b.v[1]),
          [b2] "v"(b.v[2]), [b3] "v"(b.v[3]), [b4] "v"(b.v[4]), [b5] "v"(b.v[5]), [b6] "v"(b.v[6]),
          [b7] "v"(b.v[7]), [k38] "v"(38u)
        : "vcc");
    return r;
#else
    Fe8 r;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = addc32(a.v[i], b.v[i], c, &c);
    fe8_fold(r.v, c);
    return r;
#endif
}

// a - b: a borrow means 2^256 was added, i.e. 38 too many (mod p); a second
// borrow leaves the value >= 2^256 - 38, so one more 38 cannot borrow.
LXV_HD Fe8 fe8_sub(const Fe8& a, const Fe8& b) {
#if defined(__HIP_DEVICE_COMPILE__)
    Fe8 r;
    uint32_t t;
    asm("v_sub_co_u32_e32 %[r0], vcc, %[a0], %[b0]\n"
        "v_subb_co_u32_e32 %[r1], vcc, %[a1], %[b1], vcc\n"
        "v_subb_co_u32_e32 %[r2], vcc, %[a2], %[b2], vcc\n"
        "v_subb_co_u32_e32 %[r3], vcc, %[a3], %[b3], vcc\n"
        "v_subb_co_u32_e32 %[r4], vcc, %[a4], %[b4], vcc\n"
        "v_subb_co_u32_e32 %[r5], vcc, %[a5], %[b5], vcc\n"
        "v_subb_co_u32_e32 %[r6], vcc, %[a6], %[b6], vcc\n"
        "v_subb_co_u32_e32 %[r7], vcc, %[a7], %[b7], vcc\n"
        "s_nop 1\n"
        "v_cndmask_b32_e32 %[t], 0, %[k38], vcc\n"
        "v_sub_co_u32_e32 %[r0], vcc, %[r0], %[t]\n"
        "v_subbrev_co_u32_e32 %[r1], vcc, 0, %[r1], vcc\n"
        "v_subbrev_co_u32_e32 %[r2], vcc, 0, %[r2], vcc\n"
        "v_subbrev_co_u32_e32 %[r3], vcc, 0, %[r3], vcc\n"
        "v_subbrev_co_u32_e32 %[r4], vcc, 0, %[r4], vcc\n"
        "v_subbrev_co_u32_e32 %[r5], vcc, 0, %[r5], vcc\n"
        "v_subbrev_co_u32_e32 %[r6], vcc, 0, %[r6], vcc\n"
        "v_subbrev_co_u32_e32 %[r7], vcc, 0, %[r7], vcc\n"
        "s_nop 1\n"
        "v_cndmask_b32_e32 %[t], 0, %[k38], vcc\n"
        "v_sub_u32_e32 %[r0], %[r0], %[t]\n"
        : [r0] "=&v"(r.v[0]), [r1] "=&v"(r.v[1]), [r2] "=&v"(r.v[2]), [r3] "=&v"(r.v[3]), [r4] "=&v"(r.v[4]),
          [r5] "=&v"(r.v[5]), [r6] "=&v"(r.v[6]), [r7] "=&v"(r.v[7]), [t] "=&v"(t)
        : [a0] "v"(a.v[0]), [a1] "v"(a.v[1]), [a2] "v"(a.v[2]), [a3] "v"(a.v[3]), [a4] "v"(a.v[4]),
          [a5] "v"(a.v[5]), [a6] "v"(a.v[6]), [a7] "v"(a.v[7]), [b0] "v"(b.v[0]), [b1] "v"(b.v[1]),
          [b2] "v"(b.v[2]), [b3] "v"(b.v[3]), [b4] "v"(b.v[4]), [b5] "v"(b.v[5]), [b6] "v"(b.v[6]),
          [b7] "v"(b.v[7]), [k38] "v"(38u)
        : "vcc");
    return r;
#else
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
#endif
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

// 1/z by divsteps (fe_divsteps.h).
LXV_HD Fe8 fe8_invert(const Fe8& z) {
    uint32_t w[8];
    fe8_to_words(z, w);
    Fe8 r;
    divsteps_invert(w, r.v);
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
    LXV_HD static fe invert_chain(const fe& a) { return fe8_invert_chain(a); }
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
