// GF(2^255 - 19) in ten signed 32-bit limbs, radix 2^25.5: limb i sits at
// bit ceil(25.5 i) and holds 26 bits (even i) or 25 bits (odd i). This is the
// layout of the reference's device field (src/gpu/common.h:20-28, after
// ref10). Products are 32 x 32 -> 64-bit multiply-adds (v_mad_i64_i32) into
// 64-bit column sums that cannot overflow, so no per-product carry handling
// is needed; one carry pass per multiply brings the columns back to 32 bits.
//
// "Tight" values (what fe10_mul and fe10_sq return) have |limb| <= 2^25 +
// 2^14; the sum or difference of two tight values is "loose" and is still a
// valid input to fe10_mul/fe10_sq (column sums stay below 2^61).
#pragma once

#include "fe_divsteps.h"
#include "lxv_common.h"

namespace lxv {

struct Fe10 {
    int32_t v[10];
};

// Bit offset of limb i: ceil(25.5 i).
LXV_HD constexpr int fe10_bit(int i) { return (51 * i + 1) >> 1; }
LXV_HD constexpr int fe10_width(int i) { return (i & 1) ? 25 : 26; }

LXV_HD Fe10 fe10_zero() {
    Fe10 h;
#pragma unroll
    for (int i = 0; i < 10; i++) h.v[i] = 0;
    return h;
}

LXV_HD Fe10 fe10_one() {
    Fe10 h = fe10_zero();
    h.v[0] = 1;
    return h;
}

LXV_HD Fe10 fe10_add(const Fe10& f, const Fe10& g) {
    Fe10 h;
#pragma unroll
    for (int i = 0; i < 10; i++) h.v[i] = f.v[i] + g.v[i];
    return h;
}

LXV_HD Fe10 fe10_sub(const Fe10& f, const Fe10& g) {
    Fe10 h;
#pragma unroll
    for (int i = 0; i < 10; i++) h.v[i] = f.v[i] - g.v[i];
    return h;
}

// One rounding carry out of limb i into limb i + 1 (limb 9 wraps into limb 0
// times 19, since 2^255 = 19 mod p).
LXV_HD void fe10_carry_step(int64_t h[10], int i) {
    const int w = fe10_width(i);
    const int64_t c = (h[i] + ((int64_t)1 << (w - 1))) >> w;
    h[i] -= c * ((int64_t)1 << w);
    if (i == 9)
        h[0] += 19 * c;
    else
        h[i + 1] += c;
}

// 64-bit column sums (|h_i| < 2^62) to a tight element. Two interleaved
// chains, from limb 0 and from limb 5, then the two limbs that received a
// late carry are settled again.
LXV_HD Fe10 fe10_carry(int64_t h[10]) {
    fe10_carry_step(h, 0);
    fe10_carry_step(h, 5);
    fe10_carry_step(h, 1);
    fe10_carry_step(h, 6);
    fe10_carry_step(h, 2);
    fe10_carry_step(h, 7);
    fe10_carry_step(h, 3);
    fe10_carry_step(h, 8);
    fe10_carry_step(h, 4);
    fe10_carry_step(h, 9);
    fe10_carry_step(h, 5);
    fe10_carry_step(h, 0);
    Fe10 out;
#pragma unroll
    for (int i = 0; i < 10; i++) out.v[i] = (int32_t)h[i];
    return out;
}

// f * g. Limb i times limb j lands on limb (i + j) mod 10, doubled when i and
// j are both odd (the half bits of their offsets add up to one) and times 19
// when i + j >= 10.
LXV_HD Fe10 fe10_mul(const Fe10& f, const Fe10& g) {
    int32_t g19[10], fd[10];
#pragma unroll
    for (int j = 0; j < 10; j++) g19[j] = 19 * g.v[j];
#pragma unroll
    for (int i = 0; i < 10; i++) fd[i] = (i & 1) ? 2 * f.v[i] : f.v[i];
    int64_t h[10];
#pragma unroll
    for (int k = 0; k < 10; k++) {
        int64_t s = 0;
#pragma unroll
        for (int i = 0; i < 10; i++) {
            // j with (i + j) mod 10 == k
            const int j = (k - i + 10) % 10;
            const int32_t a = ((i & 1) && (j & 1)) ? fd[i] : f.v[i];
            const int32_t b = (i + j >= 10) ? g19[j] : g.v[j];
            s += (int64_t)a * b;
        }
        h[k] = s;
    }
    return fe10_carry(h);
}

// f^2 with each cross product computed once and doubled.
LXV_HD Fe10 fe10_sq(const Fe10& f) {
    int32_t f2[10], f19[10];
#pragma unroll
    for (int i = 0; i < 10; i++) {
        f2[i] = 2 * f.v[i];
        f19[i] = 19 * f.v[i];
    }
    int64_t h[10];
#pragma unroll
    for (int k = 0; k < 10; k++) {
        int64_t s = 0;
#pragma unroll
        for (int i = 0; i < 10; i++) {
            const int j = (k - i + 10) % 10;
            if (j < i) continue;
            // scale = (i < j ? 2 : 1) * (both odd ? 2 : 1), wrap = 19
            const int dbl = (i < j ? 1 : 0) + (((i & 1) && (j & 1)) ? 1 : 0);
            const int32_t a = dbl == 0 ? f.v[i] : (dbl == 1 ? f2[i] : 2 * f2[i]);
            const int32_t b = (i + j >= 10) ? f19[j] : f.v[j];
            s += (int64_t)a * b;
        }
        h[k] = s;
    }
    return fe10_carry(h);
}

LXV_HD Fe10 fe10_sq_n(Fe10 f, int n) {
    for (int i = 0; i < n; i++) f = fe10_sq(f);
    return f;
}

// 32 little-endian bytes as 8 little-endian words to limbs; bit 255 dropped.
LXV_HD Fe10 fe10_from_words(const uint32_t w[8]) {
    Fe10 h;
#pragma unroll
    for (int i = 0; i < 10; i++) {
        const int bit = fe10_bit(i), word = bit >> 5, shift = bit & 31;
        const uint32_t hi = word + 1 < 8 ? w[word + 1] : 0u;
        const uint32_t v = funnel32(hi, w[word], shift);
        h.v[i] = (int32_t)(v & ((1u << fe10_width(i)) - 1));
    }
    return h;
}

// A tight element to its canonical encoding (8 little-endian words). With
// q = floor(h / p) in {-1, 0, 1}, computed from the top limb and rippled up
// through every limb, h - q p = h + 19 q - 2^255 q is carried with floor
// semantics and the 2^255 bit dropped.
LXV_HD void fe10_to_words(const Fe10& f, uint32_t out[8]) {
    int32_t h[10];
#pragma unroll
    for (int i = 0; i < 10; i++) h[i] = f.v[i];
    int32_t q = (19 * h[9] + (1 << 24)) >> 25;
#pragma unroll
    for (int i = 0; i < 10; i++) q = (h[i] + q) >> fe10_width(i);
    h[0] += 19 * q;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const int w = fe10_width(i);
        const int32_t c = h[i] >> w;
        h[i + 1] += c;
        h[i] -= c * (1 << w);
    }
    h[9] &= (1 << 25) - 1;
#pragma unroll
    for (int i = 0; i < 8; i++) out[i] = 0;
#pragma unroll
    for (int i = 0; i < 10; i++) {
        const int bit = fe10_bit(i), word = bit >> 5, shift = bit & 31;
        const uint32_t v = (uint32_t)h[i];
        out[word] |= v << shift;
        if (shift && word + 1 < 8) out[word + 1] |= v >> (32 - shift);
    }
}

}  // namespace lxv

namespace lxv {

// z^(p - 2) = z^-1 by the addition chain of ref10 (254 squarings, 11
// multiplications), as the reference's fe_invert (src/gpu/common.h:115-128).
LXV_HD Fe10 fe10_invert_chain(const Fe10& z) {
    const Fe10 z2 = fe10_sq(z);
    const Fe10 z9 = fe10_mul(fe10_sq_n(z2, 2), z);
    const Fe10 z11 = fe10_mul(z9, z2);
    const Fe10 z2_5_0 = fe10_mul(fe10_sq(z11), z9);
    const Fe10 z2_10_0 = fe10_mul(fe10_sq_n(z2_5_0, 5), z2_5_0);
    const Fe10 z2_20_0 = fe10_mul(fe10_sq_n(z2_10_0, 10), z2_10_0);
    const Fe10 z2_40_0 = fe10_mul(fe10_sq_n(z2_20_0, 20), z2_20_0);
    const Fe10 z2_50_0 = fe10_mul(fe10_sq_n(z2_40_0, 10), z2_10_0);
    const Fe10 z2_100_0 = fe10_mul(fe10_sq_n(z2_50_0, 50), z2_50_0);
    const Fe10 z2_200_0 = fe10_mul(fe10_sq_n(z2_100_0, 100), z2_100_0);
    const Fe10 z2_250_0 = fe10_mul(fe10_sq_n(z2_200_0, 50), z2_50_0);
    return fe10_mul(fe10_sq_n(z2_250_0, 5), z11);
}

// The operations the walk kernels are written against, for this layout.
struct Field10 {
    typedef Fe10 fe;
    static constexpr int kLimbs = 10;
    LXV_HD static fe zero() { return fe10_zero(); }
    LXV_HD static fe one() { return fe10_one(); }
    LXV_HD static fe add(const fe& a, const fe& b) { return fe10_add(a, b); }
    LXV_HD static fe sub(const fe& a, const fe& b) { return fe10_sub(a, b); }
    // A sum of up to four tight values made tight again (for encoding or for
    // further additions).
    LXV_HD static fe tight(const fe& a) {
        int64_t h[10];
#pragma unroll
        for (int i = 0; i < 10; i++) h[i] = a.v[i];
        return fe10_carry(h);
    }
    LXV_HD static fe mul(const fe& a, const fe& b) { return fe10_mul(a, b); }
    LXV_HD static fe sq(const fe& a) { return fe10_sq(a); }
    LXV_HD static fe invert(const fe& a) {
        uint32_t w[8], r[8];
        fe10_to_words(a, w);
        divsteps_invert(w, r);
        return fe10_from_words(r);
    }
    LXV_HD static fe invert_chain(const fe& a) { return fe10_invert_chain(a); }
    LXV_HD static fe from_words(const uint32_t w[8]) { return fe10_from_words(w); }
    LXV_HD static void to_words(const fe& a, uint32_t out[8]) { fe10_to_words(a, out); }
    // a24 * a for the ladder (a24 = 121665 < 2^17): a plain limb scale and a carry.
    LXV_HD static fe mul_small(const fe& a, int32_t k) {
        int64_t h[10];
#pragma unroll
        for (int i = 0; i < 10; i++) h[i] = (int64_t)a.v[i] * k;
        return fe10_carry(h);
    }
    LXV_HD static void cswap(fe& a, fe& b, uint32_t swap) {
        const int32_t m = -(int32_t)swap;
#pragma unroll
        for (int i = 0; i < 10; i++) {
            const int32_t t = m & (a.v[i] ^ b.v[i]);
            a.v[i] ^= t;
            b.v[i] ^= t;
        }
    }
    // Raw limbs for storage: kLimbs 32-bit words.
    LXV_HD static void store(const fe& a, uint32_t out[kLimbs]) {
#pragma unroll
        for (int i = 0; i < 10; i++) out[i] = (uint32_t)a.v[i];
    }
    LXV_HD static fe load(const uint32_t in[kLimbs]) {
        fe a;
#pragma unroll
        for (int i = 0; i < 10; i++) a.v[i] = (int32_t)in[i];
        return a;
    }
};

}  // namespace lxv
