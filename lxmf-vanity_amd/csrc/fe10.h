// GF(2^255 - 19) in ten unsigned 32-bit limbs, radix 2^25.5: limb i sits at
// bit ceil(25.5 i) and is 26 bits wide (even i) or 25 (odd i), the limb
// positions of the reference's device field (src/gpu/common.h:20-28). The
// arithmetic on them is built for gfx950's 32-bit VALU:
//
//  - A product is scanned column by column. Column k collects the limb
//    products i + j = k (mod 10), those that wrap past 2^255 times 19 and the
//    odd x odd ones doubled (their half bits add up to one), as ten
//    v_mad_u64_u32 into one 64-bit accumulator that starts at the previous
//    column's carry: the carry costs no addition of its own. The column's
//    limb and carry are split off with an and, an alignbit and a shift, and
//    the carry out of column 9 (bit 255) is folded back into limb 0 times 19.
//  - Limbs are unsigned. A difference adds 2p limb by limb (f + 2p - g), so
//    no limb goes negative; sums and differences are not carried.
//
// Bounds (w_i the limb width; the column-sum bounds are checked in the tests):
//  - "tight": every limb below 2^w_i + 2^18. mul, sq, mul_small, tight and
//    from_words return tight values.
//  - sub(f, g) needs g tight (2p's limbs cover a tight limb).
//  - mul and sq take "loose" values: tight, the sum of two tight values, or
//    f + 2p - g for tight f and g. Their column sums stay below 2^62.2 and
//    every pre-scaled operand (2 f_i, 4 f_i, 19 g_j) below 2^32.
//  - tight and to_words take any limbs below 2^32 - 2^12.
#pragma once

#include "fe_divsteps.h"
#include "lxv_common.h"

namespace lxv {

struct Fe10 {
    uint32_t v[10];
};

// Bit offset of limb i: ceil(25.5 i).
LXV_HD constexpr int fe10_bit(int i) { return (51 * i + 1) >> 1; }
LXV_HD constexpr int fe10_width(int i) { return (i & 1) ? 25 : 26; }
LXV_HD constexpr uint32_t fe10_mask(int i) { return (1u << fe10_width(i)) - 1; }
// Limb i of 2p = 2^256 - 38: 2^27 - 38, then 2 (2^w_i - 1).
LXV_HD constexpr uint32_t fe10_2p(int i) { return i == 0 ? (1u << 27) - 38 : 2 * fe10_mask(i); }

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

// f + 2p - g, g tight.
LXV_HD Fe10 fe10_sub(const Fe10& f, const Fe10& g) {
    Fe10 h;
#pragma unroll
    for (int i = 0; i < 10; i++) h.v[i] = (f.v[i] + fe10_2p(i)) - g.v[i];
    return h;
}

// a * b + c on v_mad_u64_u32. Written as one instruction so that the sum is
// never reassociated (the compiler otherwise sums a column's products first
// and adds the carry with an extra 64-bit addition); its carry-out goes to
// VCC, unread. The result never shares registers with an operand, as the
// compiler's own v_mad_u64_u32 is constrained.
LXV_HD uint64_t fe10_mad(uint32_t a, uint32_t b, uint64_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t r;
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %3" : "=&v"(r) : "v"(a), "v"(b), "v"(c) : "vcc");
    return r;
#else
    return (uint64_t)a * b + c;
#endif
}

// c + sum a[i] b[i] over one column, on the device as one asm block of
// chained v_mad_u64_u32 (between separate asm statements the compiler puts a
// wait state before each). N is 10 (multiplication) or 6 and 5 (squaring).
#define LXV_MAD0(x, y, c) "v_mad_u64_u32 %0, vcc, %" #x ", %" #y ", %" #c "\n"
#define LXV_MAD(x, y) "v_mad_u64_u32 %0, vcc, %" #x ", %" #y ", %0\n"
#define LXV_AB(i) "v"(a[i]), "v"(b[i])

LXV_HD uint64_t fe10_col10(const uint32_t a[10], const uint32_t b[10], uint64_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t r;
    asm(LXV_MAD0(1, 2, 21) LXV_MAD(3, 4) LXV_MAD(5, 6) LXV_MAD(7, 8) LXV_MAD(9, 10) LXV_MAD(11, 12) LXV_MAD(13, 14)
            LXV_MAD(15, 16) LXV_MAD(17, 18) LXV_MAD(19, 20)
        : "=&v"(r)
        : LXV_AB(0), LXV_AB(1), LXV_AB(2), LXV_AB(3), LXV_AB(4), LXV_AB(5), LXV_AB(6), LXV_AB(7), LXV_AB(8), LXV_AB(9),
          "v"(c)
        : "vcc");
    return r;
#else
    for (int i = 0; i < 10; i++) c = fe10_mad(a[i], b[i], c);
    return c;
#endif
}

LXV_HD uint64_t fe10_col6(const uint32_t a[6], const uint32_t b[6], uint64_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t r;
    asm(LXV_MAD0(1, 2, 13) LXV_MAD(3, 4) LXV_MAD(5, 6) LXV_MAD(7, 8) LXV_MAD(9, 10) LXV_MAD(11, 12)
        : "=&v"(r)
        : LXV_AB(0), LXV_AB(1), LXV_AB(2), LXV_AB(3), LXV_AB(4), LXV_AB(5), "v"(c)
        : "vcc");
    return r;
#else
    for (int i = 0; i < 6; i++) c = fe10_mad(a[i], b[i], c);
    return c;
#endif
}

LXV_HD uint64_t fe10_col5(const uint32_t a[5], const uint32_t b[5], uint64_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t r;
    asm(LXV_MAD0(1, 2, 11) LXV_MAD(3, 4) LXV_MAD(5, 6) LXV_MAD(7, 8) LXV_MAD(9, 10)
        : "=&v"(r)
        : LXV_AB(0), LXV_AB(1), LXV_AB(2), LXV_AB(3), LXV_AB(4), "v"(c)
        : "vcc");
    return r;
#else
    for (int i = 0; i < 5; i++) c = fe10_mad(a[i], b[i], c);
    return c;
#endif
}

#undef LXV_MAD0
#undef LXV_MAD
#undef LXV_AB

// x >> n for a VGPR x, kept a 32-bit shift (the compiler otherwise merges it
// with the alignbit below into a 64-bit shift).
LXV_HD uint32_t fe10_shr(uint32_t x, int n) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t r;
    asm("v_lshrrev_b32 %0, %1, %2" : "=v"(r) : "i"(n), "v"(x));
    return r;
#else
    return x >> n;
#endif
}

// The low w bits of a column sum as its limb; the rest (two words) as the
// next column's starting value.
LXV_HD uint32_t fe10_split(uint64_t acc, int w, uint64_t& carry) {
    const uint32_t lo = (uint32_t)acc, hi = (uint32_t)(acc >> 32);
    carry = ((uint64_t)fe10_shr(hi, w) << 32) | funnel32(hi, lo, w);
    return lo & ((1u << w) - 1);
}

// Limbs h with the carry c (< 2^38) out of bit 255: c times 19 into limb 0,
// whose own carry (< 2^16) goes to limb 1.
LXV_HD Fe10 fe10_fold(const uint32_t h[10], uint64_t c) {
    const uint64_t t = (uint64_t)(uint32_t)c * 19u + h[0] + ((uint64_t)(19u * (uint32_t)(c >> 32)) << 32);
    Fe10 r;
    r.v[0] = (uint32_t)t & fe10_mask(0);
    r.v[1] = h[1] + funnel32((uint32_t)(t >> 32), (uint32_t)t, 26);
#pragma unroll
    for (int i = 2; i < 10; i++) r.v[i] = h[i];
    return r;
}

LXV_HD Fe10 fe10_mul(const Fe10& f, const Fe10& g) {
    uint32_t g19[10], f2[10];
#pragma unroll
    for (int j = 1; j < 10; j++) g19[j] = 19u * g.v[j];
#pragma unroll
    for (int i = 1; i < 10; i += 2) f2[i] = 2u * f.v[i];
    uint32_t h[10];
    uint64_t c = 0;
#pragma unroll
    for (int k = 0; k < 10; k++) {
        uint32_t a[10], b[10];
#pragma unroll
        for (int i = 0; i < 10; i++) {
            const int j = (k - i + 10) % 10;
            a[i] = ((i & 1) && (j & 1)) ? f2[i] : f.v[i];
            b[i] = (i + j >= 10) ? g19[j] : g.v[j];
        }
        h[k] = fe10_split(fe10_col10(a, b, c), fe10_width(k), c);
    }
    return fe10_fold(h, c);
}

// f^2: each cross product once, against 2 f_i (4 f_i when both limbs are
// odd); an even column has six products, an odd one five.
LXV_HD Fe10 fe10_sq(const Fe10& f) {
    uint32_t f2[10], f4[10], f19[10];
#pragma unroll
    for (int i = 0; i < 10; i++) f2[i] = 2u * f.v[i];
#pragma unroll
    for (int i = 1; i < 10; i += 2) f4[i] = 4u * f.v[i];
#pragma unroll
    for (int j = 5; j < 10; j++) f19[j] = 19u * f.v[j];
    uint32_t h[10];
    uint64_t c = 0;
#pragma unroll
    for (int k = 0; k < 10; k++) {
        uint32_t a[6], b[6];
        int n = 0;
#pragma unroll
        for (int i = 0; i < 10; i++) {
            const int j = (k - i + 10) % 10;
            if (j < i) continue;
            const bool odd = (i & 1) && (j & 1);
            a[n] = i == j ? (odd ? f2[i] : f.v[i]) : (odd ? f4[i] : f2[i]);
            b[n] = (i + j >= 10) ? f19[j] : f.v[j];
            n++;
        }
        const uint64_t acc = (k & 1) ? fe10_col5(a, b, c) : fe10_col6(a, b, c);
        h[k] = fe10_split(acc, fe10_width(k), c);
    }
    return fe10_fold(h, c);
}

LXV_HD Fe10 fe10_sq_n(Fe10 f, int n) {
    for (int i = 0; i < n; i++) f = fe10_sq(f);
    return f;
}

// Limbs below 2^32 - 2^12 to a tight element: limb 9's carry wraps into limb
// 0 times 19 first, then one floor carry pass runs from limb 0 up. Limbs 0..8
// end below 2^w_i, limb 9 below 2^25 + 2^8.
LXV_HD Fe10 fe10_carry(const Fe10& a) {
    Fe10 h = a;
    const uint32_t c9 = h.v[9] >> 25;
    h.v[9] &= fe10_mask(9);
    h.v[0] += 19u * c9;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        h.v[i + 1] += h.v[i] >> fe10_width(i);
        h.v[i] &= fe10_mask(i);
    }
    return h;
}

// a * k for a loose a and k < 2^17 (the ladder's a24 = 121665).
LXV_HD Fe10 fe10_mul_small(const Fe10& a, uint32_t k) {
    uint32_t h[10];
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 10; i++) h[i] = fe10_split(fe10_mad(a.v[i], k, c), fe10_width(i), c);
    return fe10_fold(h, c);
}

// 32 little-endian bytes as 8 little-endian words to limbs; bit 255 dropped.
LXV_HD Fe10 fe10_from_words(const uint32_t w[8]) {
    Fe10 h;
#pragma unroll
    for (int i = 0; i < 10; i++) {
        const int bit = fe10_bit(i), word = bit >> 5, shift = bit & 31;
        const uint32_t hi = word + 1 < 8 ? w[word + 1] : 0u;
        h.v[i] = funnel32(hi, w[word], shift) & fe10_mask(i);
    }
    return h;
}

// The canonical encoding (8 little-endian words) of any element with limbs
// below 2^32 - 2^12. After the carry pass the value V is below 2^255 + 2^238
// and limbs 0..8 hold disjoint bits. V >= p exactly when V + 19 carries into
// bit 255, which needs limb 9 at 2^25 - 1 or above; only lanes with such a
// limb take the exact path (add 19 q, carry, drop bit 255 for
// q = floor((V + 19) / 2^255)), so the others pay one compare.
LXV_HD void fe10_to_words(const Fe10& f, uint32_t out[8]) {
    Fe10 h = fe10_carry(f);
    if (h.v[9] >= fe10_mask(9)) {
        uint32_t q = (h.v[0] + 19u) >> 26;
#pragma unroll
        for (int i = 1; i < 10; i++) q = (h.v[i] + q) >> fe10_width(i);
        h.v[0] += 19u * q;
#pragma unroll
        for (int i = 0; i < 9; i++) {
            h.v[i + 1] += h.v[i] >> fe10_width(i);
            h.v[i] &= fe10_mask(i);
        }
        h.v[9] &= fe10_mask(9);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) out[i] = 0;
#pragma unroll
    for (int i = 0; i < 10; i++) {
        const int bit = fe10_bit(i), word = bit >> 5, shift = bit & 31;
        out[word] |= h.v[i] << shift;
        if (shift && word + 1 < 8) out[word + 1] |= h.v[i] >> (32 - shift);
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
    // Sums and differences of a few tight values made tight again (for
    // encoding or for further additions).
    LXV_HD static fe tight(const fe& a) { return fe10_carry(a); }
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
    LXV_HD static fe mul_small(const fe& a, int32_t k) { return fe10_mul_small(a, (uint32_t)k); }
    LXV_HD static void cswap(fe& a, fe& b, uint32_t swap) {
        const uint32_t m = 0u - swap;
#pragma unroll
        for (int i = 0; i < 10; i++) {
            const uint32_t t = m & (a.v[i] ^ b.v[i]);
            a.v[i] ^= t;
            b.v[i] ^= t;
        }
    }
    // Raw limbs for storage: kLimbs 32-bit words.
    LXV_HD static void store(const fe& a, uint32_t out[kLimbs]) {
#pragma unroll
        for (int i = 0; i < 10; i++) out[i] = a.v[i];
    }
    LXV_HD static fe load(const uint32_t in[kLimbs]) {
        fe a;
#pragma unroll
        for (int i = 0; i < 10; i++) a.v[i] = in[i];
        return a;
    }
};

}  // namespace lxv
