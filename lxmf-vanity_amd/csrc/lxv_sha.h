// SHA-256 (FIPS 180-4) specialised for the LXMF address of one walked key:
//
//   identity hash = SHA-256(u || ed)[0..16]       2 compressions: block 1 is
//                                                 u || ed, block 2 is padding
//   address       = SHA-256(nh || ih16)[0..16]    1 compression (block 3)
//
// with the word layout of src/gpu/common.h:218-253. Everything that does not
// depend on the key is computed once per search on the host (ShaSearch):
// block 1's words 8..15 (the Ed25519 key) and the parts of its message
// schedule they fix, block 3's first two rounds and its constant schedule
// terms; block 2's schedule is a compile-time table. Rotates are
// v_alignbit_b32, byte swaps v_perm_b32, Ch/Maj/three-way XOR v_bitop3_b32.
#pragma once

#include "lxv_common.h"

namespace lxv {

LXV_HD constexpr uint32_t sha_k(int t) {
    constexpr uint32_t K[64] = {
        0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
        0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
        0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
        0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
        0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
        0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
        0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
        0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
    return K[t];
}

LXV_HD constexpr uint32_t sha_h0(int i) {
    constexpr uint32_t H[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                               0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    return H[i];
}

LXV_HD constexpr uint32_t cx_rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
LXV_HD constexpr uint32_t cx_s0(uint32_t w) { return cx_rotr(w, 7) ^ cx_rotr(w, 18) ^ (w >> 3); }
LXV_HD constexpr uint32_t cx_s1(uint32_t w) { return cx_rotr(w, 17) ^ cx_rotr(w, 19) ^ (w >> 10); }

// K[t] + W[t] of block 2, the padding block of a 64-byte message (W0 =
// 0x80000000, W15 = 512, the rest zero) and its schedule.
LXV_HD constexpr uint32_t sha_kw_pad64(int t) {
    uint32_t w[64] = {0};
    w[0] = 0x80000000u;
    w[15] = 512;
    for (int i = 16; i < 64; i++) w[i] = cx_s1(w[i - 2]) + w[i - 7] + cx_s0(w[i - 15]) + w[i - 16];
    return sha_k(t) + w[t];
}

// Three-input boolean functions in one v_bitop3_b32 each (the immediate is
// the truth table indexed by a * 4 + b * 2 + c): XOR3 0x96, Ch 0xca, Maj 0xe8.
LXV_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}
LXV_HD uint32_t sha_ch(uint32_t e, uint32_t f, uint32_t g) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(e, f, g, 0xca);
#else
    return (e & f) ^ (~e & g);
#endif
}
LXV_HD uint32_t sha_maj(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0xe8);
#else
    return (a & b) | (c & (a | b));
#endif
}
LXV_HD uint32_t sha_S0(uint32_t a) { return xor3(rotr32(a, 2), rotr32(a, 13), rotr32(a, 22)); }
LXV_HD uint32_t sha_S1(uint32_t e) { return xor3(rotr32(e, 6), rotr32(e, 11), rotr32(e, 25)); }
LXV_HD uint32_t sha_s0(uint32_t w) { return xor3(rotr32(w, 7), rotr32(w, 18), w >> 3); }
LXV_HD uint32_t sha_s1(uint32_t w) { return xor3(rotr32(w, 17), rotr32(w, 19), w >> 10); }

// Round t on the state ring S: at round t, a..h are S[(0 - t) & 7] ..
// S[(7 - t) & 7], so a round writes two slots and nothing moves. With t known
// at compile time (unrolled loops) the ring is plain registers.
LXV_HD void sha_round(uint32_t S[8], int t, uint32_t kw) {
    const uint32_t a = S[(8 - (t & 7)) & 7], b = S[(9 - (t & 7)) & 7], c = S[(10 - (t & 7)) & 7];
    const uint32_t e = S[(12 - (t & 7)) & 7], f = S[(13 - (t & 7)) & 7], g = S[(14 - (t & 7)) & 7];
    uint32_t& d = S[(11 - (t & 7)) & 7];
    uint32_t& h = S[(15 - (t & 7)) & 7];
    const uint32_t t1 = (h + kw) + sha_S1(e) + sha_ch(e, f, g);
    d += t1;
    h = t1 + sha_S0(a) + sha_maj(a, b, c);
}

// Per-search constants of the address hash, computed on the host
// (sha_search_init). Words are big-endian as SHA-256 reads them.
struct ShaSearch {
    uint32_t kw1[8];    // block 1, rounds 8..15: K[t] + ed[t - 8]
    uint32_t c16, c17;  // block 1: s1(W14) + W9, s1(W15) + W10
    uint32_t w11[5];    // block 1: W11..W15 (ed[3..7])
    uint32_t c23;       // block 1: s0(W8)
    uint32_t c24[7];    // block 1, t = 24..30: s0(W[t-15]) + W[t-16]
    uint32_t st3[8];    // block 3: state ring after rounds 0 and 1
    uint32_t nh2;       // block 3: W2's top half (name hash bytes 8, 9)
    uint32_t d16;       // block 3: W16
    uint32_t d17;       // block 3: s1(W15) + W1
    uint32_t d18;       // block 3: s1(W16)
    uint32_t d30;       // block 3: s0(W15)
    uint32_t d31;       // block 3: s0(W16) + W15
    uint32_t want[4];   // prefix test: (address & mask) == want
    uint32_t mask[4];
};

// The general schedule word t (in place in the 16-word ring).
LXV_HD uint32_t sha_wnext(uint32_t W[16], int t) {
    W[t & 15] = sha_s1(W[(t - 2) & 15]) + W[(t - 7) & 15] + sha_s0(W[(t - 15) & 15]) + W[t & 15];
    return W[t & 15];
}

// Block 1: compress u (8 little-endian words, byte-swapped to big-endian)
// || ed from the initial state; writes the 8-word state.
LXV_HD void sha_block1(const uint32_t u[8], const ShaSearch& c, uint32_t st[8]) {
    uint32_t W[16], S[8];
#pragma unroll
    for (int i = 0; i < 8; i++) W[i] = bswap32(u[i]);
#pragma unroll
    for (int i = 0; i < 8; i++) S[i] = sha_h0(i);
#pragma unroll
    for (int t = 0; t < 8; t++) sha_round(S, t, sha_k(t) + W[t]);
#pragma unroll
    for (int t = 8; t < 16; t++) sha_round(S, t, c.kw1[t - 8]);
#pragma unroll
    for (int t = 16; t < 64; t++) {
        uint32_t w;
        if (t == 16)
            w = sha_s0(W[1]) + W[0] + c.c16;
        else if (t == 17)
            w = sha_s0(W[2]) + W[1] + c.c17;
        else if (t <= 22)
            w = sha_s1(W[(t - 2) & 15]) + sha_s0(W[t - 15]) + W[t - 16] + c.w11[t - 18];
        else if (t == 23)
            w = sha_s1(W[21 & 15]) + W[16 & 15] + W[7] + c.c23;
        else if (t <= 30)
            w = sha_s1(W[(t - 2) & 15]) + W[(t - 7) & 15] + c.c24[t - 24];
        else if (t == 31)
            w = sha_s1(W[29 & 15]) + W[24 & 15] + sha_s0(W[16 & 15]) + c.w11[4];
        else
            w = sha_s1(W[(t - 2) & 15]) + W[(t - 7) & 15] + sha_s0(W[(t - 15) & 15]) + W[t & 15];
        W[t & 15] = w;
        sha_round(S, t, sha_k(t) + w);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) st[i] = S[i] + sha_h0(i);
}

// Block 2: the padding block of the 64-byte public key.
LXV_HD void sha_block2(uint32_t st[8]) {
    uint32_t S[8];
#pragma unroll
    for (int i = 0; i < 8; i++) S[i] = st[i];
#pragma unroll
    for (int t = 0; t < 64; t++) sha_round(S, t, sha_kw_pad64(t));
#pragma unroll
    for (int i = 0; i < 8; i++) st[i] += S[i];
}

// Block 3: SHA-256(name hash (10 bytes) || identity hash (16 bytes)), one
// block with its padding; ih = the first 4 state words of the identity hash.
// Writes the first 4 words of the digest: the 16-byte address.
LXV_HD void sha_block3(const uint32_t ih[4], const ShaSearch& c, uint32_t out[4]) {
    uint32_t W[16], S[8];
    W[2] = c.nh2 | (ih[0] >> 16);
    W[3] = funnel32(ih[0], ih[1], 16);
    W[4] = funnel32(ih[1], ih[2], 16);
    W[5] = funnel32(ih[2], ih[3], 16);
    W[6] = (ih[3] << 16) | 0x8000u;
#pragma unroll
    for (int i = 0; i < 8; i++) S[i] = c.st3[i];
#pragma unroll
    for (int t = 2; t < 7; t++) sha_round(S, t, sha_k(t) + W[t]);
#pragma unroll
    for (int t = 7; t < 15; t++) sha_round(S, t, sha_k(t));
    sha_round(S, 15, sha_k(15) + 26 * 8);
#pragma unroll
    for (int t = 16; t < 64; t++) {
        uint32_t w;
        if (t == 16)
            w = c.d16;
        else if (t == 17)
            w = sha_s0(W[2]) + c.d17;
        else if (t == 18)
            w = sha_s0(W[3]) + W[2] + c.d18;
        else if (t <= 21)
            w = sha_s1(W[(t - 2) & 15]) + sha_s0(W[t - 15]) + W[t - 16];
        else if (t == 22)
            w = sha_s1(W[20 & 15]) + W[6] + 26 * 8;
        else if (t == 23)
            w = sha_s1(W[21 & 15]) + c.d16;
        else if (t <= 29)
            w = sha_s1(W[(t - 2) & 15]) + W[(t - 7) & 15];
        else if (t == 30)
            w = sha_s1(W[28 & 15]) + W[23 & 15] + c.d30;
        else if (t == 31)
            w = sha_s1(W[29 & 15]) + W[24 & 15] + c.d31;
        else
            w = sha_s1(W[(t - 2) & 15]) + W[(t - 7) & 15] + sha_s0(W[(t - 15) & 15]) + W[t & 15];
        W[t & 15] = w;
        sha_round(S, t, sha_k(t) + w);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = S[i] + sha_h0(i);
}

// The 16-byte LXMF address (4 big-endian words) of X25519 public key u.
LXV_HD void lxmf_address_words(const uint32_t u[8], const ShaSearch& c, uint32_t out[4]) {
    uint32_t st[8];
    sha_block1(u, c, st);
    sha_block2(st);
    sha_block3(st, c, out);
}

// (x & m) ^ w: zero exactly when (x & m) == w (v_bitop3 truth table 0x6c).
LXV_HD uint32_t masked_diff(uint32_t x, uint32_t w, uint32_t m) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(x, w, m, 0x6c);
#else
    return (x & m) ^ w;
#endif
}

// (address & mask) == want, word by word. Words 1..3 are tested only when
// the search's mask or want words there are not all zero (a uniform branch),
// so a prefix of up to 8 hex digits costs one bitop3 and one compare.
LXV_HD bool prefix_match(const uint32_t addr[4], const ShaSearch& c) {
    uint32_t d = masked_diff(addr[0], c.want[0], c.mask[0]);
    if (__builtin_expect((c.mask[1] | c.mask[2] | c.mask[3] | c.want[1] | c.want[2] | c.want[3]) != 0, 0))
        d |= masked_diff(addr[1], c.want[1], c.mask[1]) | masked_diff(addr[2], c.want[2], c.mask[2]) |
             masked_diff(addr[3], c.want[3], c.mask[3]);
    return d == 0;
}

// Host side: the per-search constants from the Ed25519 public key words (big
// endian), the name hash words (big-endian, the last holding 2 bytes in its
// top half) and the prefix want/mask words.
inline void sha_search_init(ShaSearch& c, const uint32_t ed[8], const uint32_t nh[3], const uint32_t want[4],
                            const uint32_t mask[4]) {
    for (int i = 0; i < 8; i++) c.kw1[i] = sha_k(8 + i) + ed[i];
    c.c16 = cx_s1(ed[6]) + ed[1];
    c.c17 = cx_s1(ed[7]) + ed[2];
    for (int i = 0; i < 5; i++) c.w11[i] = ed[3 + i];
    c.c23 = cx_s0(ed[0]);
    for (int t = 24; t <= 30; t++) c.c24[t - 24] = cx_s0(ed[t - 15 - 8]) + ed[t - 16 - 8];
    uint32_t S[8];
    for (int i = 0; i < 8; i++) S[i] = sha_h0(i);
    sha_round(S, 0, sha_k(0) + nh[0]);
    sha_round(S, 1, sha_k(1) + nh[1]);
    for (int i = 0; i < 8; i++) c.st3[i] = S[i];
    c.nh2 = nh[2];
    c.d16 = cx_s0(nh[1]) + nh[0];
    c.d17 = cx_s1(26 * 8) + nh[1];
    c.d18 = cx_s1(c.d16);
    c.d30 = cx_s0(26 * 8);
    c.d31 = cx_s0(c.d16) + 26 * 8;
    for (int i = 0; i < 4; i++) {
        c.want[i] = want[i];
        c.mask[i] = mask[i];
    }
}

}  // namespace lxv
