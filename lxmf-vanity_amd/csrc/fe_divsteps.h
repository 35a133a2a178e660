// Inversion modulo p = 2^255 - 19 by divsteps, shared by both limb layouts
// (fe8.h, fe10.h). It replaces the reference's exponentiation chain
// (src/gpu/common.h:109-128) with the same result for every input.
#pragma once

#include "lxv_common.h"

namespace lxv {

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

// acc + a b for signed 32-bit a, b: one v_mad_i64_i32. Written out because
// the compiler, knowing a limb is non-negative, otherwise widens the product
// to a 64 x 64-bit multiply.
LXV_HD int64_t mad_i64(int32_t a, int32_t b, int64_t acc) {
#if defined(__HIP_DEVICE_COMPILE__)
    int64_t r;
    uint64_t cc;
    asm("v_mad_i64_i32 %0, %1, %2, %3, %4" : "=v"(r), "=s"(cc) : "v"(a), "v"(b), "v"(acc));
    (void)cc;
    return r;
#else
    return acc + (int64_t)a * b;
#endif
}

// 30 divsteps on the low words of f and g; returns zeta = -(delta + 1/2)
// after them and the matrix t = (u, v, q, r).
LXV_HD uint32_t divsteps30(uint32_t zeta, uint32_t f, uint32_t g, int32_t t[4]) {
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
    int64_t cf = mad_i64(t[1], g.v[0], mad_i64(t[0], f.v[0], 0));
    int64_t cg = mad_i64(t[3], g.v[0], mad_i64(t[2], f.v[0], 0));
    cf >>= 30;
    cg >>= 30;
#pragma unroll
    for (int i = 1; i < 9; i++) {
        cf = mad_i64(t[1], g.v[i], mad_i64(t[0], f.v[i], cf));
        cg = mad_i64(t[3], g.v[i], mad_i64(t[2], f.v[i], cg));
        f.v[i - 1] = (int32_t)((uint32_t)cf & kM30);
        g.v[i - 1] = (int32_t)((uint32_t)cg & kM30);
        cf >>= 30;
        cg >>= 30;
    }
    f.v[8] = (int32_t)cf;
    g.v[8] = (int32_t)cg;
}

// (d, e) <- (u d + v e, q d + r e) / 2^30 (mod p), kept in (-2p, p). p is
// sparse in signed 30-bit limbs (p = 2^15 2^240 - 19), so adding md p adds
// -19 md to limb 0 and 2^15 md to limb 8 only.
LXV_HD void s30_update_de(S30& d, S30& e, const int32_t t[4]) {
    const int32_t u = t[0], v = t[1], q = t[2], r = t[3];
    const int32_t sd = d.v[8] >> 31, se = e.v[8] >> 31;
    int32_t md = (u & sd) + (v & se);
    int32_t me = (q & sd) + (r & se);
    int64_t cd = mad_i64(v, e.v[0], mad_i64(u, d.v[0], 0));
    int64_t ce = mad_i64(r, e.v[0], mad_i64(q, d.v[0], 0));
    md -= (int32_t)((kPInv30 * (uint32_t)cd + (uint32_t)md) & kM30);
    me -= (int32_t)((kPInv30 * (uint32_t)ce + (uint32_t)me) & kM30);
    cd = mad_i64(-19, md, cd);
    ce = mad_i64(-19, me, ce);
    cd >>= 30;
    ce >>= 30;
#pragma unroll
    for (int i = 1; i < 9; i++) {
        cd = mad_i64(v, e.v[i], mad_i64(u, d.v[i], cd));
        ce = mad_i64(r, e.v[i], mad_i64(q, d.v[i], ce));
        if (i == 8) {
            cd = mad_i64(md, 32768, cd);
            ce = mad_i64(me, 32768, ce);
        }
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

// out = 1/w mod p, canonical (< p), for w canonical (< p) as 8 little-endian
// words; 0 maps to 0.
LXV_HD void divsteps_invert(const uint32_t w[8], uint32_t out[8]) {
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
        zeta = divsteps30(zeta, (uint32_t)f.v[0] | ((uint32_t)f.v[1] << 30),
                          (uint32_t)g.v[0] | ((uint32_t)g.v[1] << 30), t);
        s30_update_de(d, e, t);
        s30_update_fg(f, g, t);
        // Once g = 0 further steps leave f and d unchanged (mod p): stop when
        // every lane of the wave is there (random inputs need about 515
        // steps, so 18 batches; the 20 cover the proven worst case).
        if (b >= 16) {
            int32_t any = 0;
#pragma unroll
            for (int i = 0; i < 9; i++) any |= g.v[i];
#if defined(__HIP_DEVICE_COMPILE__)
            if (__all(any == 0)) break;
#else
            if (any == 0) break;
#endif
        }
    }
    // 1/w = d f with f = +-1: negate d when f < 0, bring it into [0, 2p),
    // then subtract p once if that leaves it non-negative
    const int32_t neg = f.v[8] >> 31;
#pragma unroll
    for (int i = 0; i < 9; i++) d.v[i] = (d.v[i] ^ neg) - neg;
    s30_normalize(d);
    s30_add_p_if_negative(d);
    s30_add_p_if_negative(d);
    S30 m;
#pragma unroll
    for (int i = 0; i < 9; i++) m.v[i] = d.v[i] - s30_p_limb(i);
    s30_normalize(m);
    const int32_t keep = m.v[8] >> 31;  // d < p: keep d
#pragma unroll
    for (int i = 0; i < 9; i++) d.v[i] = (d.v[i] & keep) | (m.v[i] & ~keep);
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int i = (32 * j) / 30, s = (32 * j) % 30;
        const uint32_t lo = (uint32_t)d.v[i] >> s;
        const uint32_t hi = i + 1 < 9 ? (uint32_t)d.v[i + 1] << (30 - s) : 0u;
        out[j] = lo | hi;
    }
}

}  // namespace lxv
