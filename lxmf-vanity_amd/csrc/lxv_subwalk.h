// The sub-walk form of the key walk: each lane runs K independent walks
// ("sub-walks") in affine x coordinates, one key per sub-walk per step, and
// inverts across its K sub-walks (Montgomery's trick) instead of across K
// consecutive keys of one walk.
//
// Sub-walk t (t = j x lanes + lane, j < K) walks keys k_t + 8 n with
// u_n = u(k_t + 8 n). With Q = 8B (u = q) and the x-only differential
// addition identity u(P+Q) u(P-Q) (u(P) - q)^2 = (u(P) q - 1)^2,
//
//     u_{n+1} = N / D,   N = (u_n q - 1)^2,   D = (u_n - q)^2 u_{n-1},
//
// so a step costs 2M + 2S and the batched inversion 4M per key: 6M + 2S per
// key against 8M + 2S for the projective walk of lxv_walk.h (whose step
// needs 4M + 2S because it carries Z), for the same one inversion per K keys.
// The price is memory traffic: the affine state (u_n, u_{n-1}) of every
// sub-walk stays in HBM (64 B per sub-walk) and is read and rewritten each
// step, plus the a_j / D_j batch arrays, about 224 B per key in all.
//
// State layout: two slots of canonical u words per sub-walk, uint2 pairs
// lane-interleaved, [slot][j][pair][lane]; at step n, slot (n & 1) holds
// u_n and the other slot u_{n-1}, which the step overwrites with u_{n+1}.
// Every sub-walk of a lane is at the same n (ns is per lane).
#pragma once

#include "lxv_walk.h"

namespace lxv {

LXV_D void sw_load_words(const uint2* st, uint32_t lanes, uint32_t K, uint32_t slot, uint32_t j, uint32_t lane,
                         uint32_t w[8]) {
    const uint2* b = st + ((size_t)(slot * K + j) * 4) * lanes + lane;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint2 v = b[(size_t)i * lanes];
        w[2 * i] = v.x;
        w[2 * i + 1] = v.y;
    }
}

LXV_HD void sw_store_words(uint2* st, uint32_t lanes, uint32_t K, uint32_t slot, uint32_t j, uint32_t lane,
                           const uint32_t w[8]) {
    uint2* b = st + ((size_t)(slot * K + j) * 4) * lanes + lane;
#pragma unroll
    for (int i = 0; i < 4; i++) b[(size_t)i * lanes] = make_uint2(w[2 * i], w[2 * i + 1]);
}

// `iters` steps of the K sub-walks of one lane. The forward pass hashes u_n
// (already affine and canonical in the state) and forms N_j, D_j; one
// inversion; the backward pass writes u_{n+1} = a_j * inv over u_{n-1}.
template <class F, int K>
LXV_D void walk_sw_lane(const WalkParams& p, uint32_t lane) {
    typedef typename F::fe fe;
    const uint32_t L = p.lanes;
    uint2* st = reinterpret_cast<uint2*>(p.state);
    const uint64_t n0 = p.ns[lane];
    uint64_t n = n0;
    uint32_t x0 = p.xor_acc[0 * L + lane], x1 = p.xor_acc[1 * L + lane];
    uint32_t x2 = p.xor_acc[2 * L + lane], x3 = p.xor_acc[3 * L + lane];
    const fe q = F::load(p.q);
    constexpr int kPairs = F::kLimbs / 2;
    const size_t slot_stride = (size_t)2 * kPairs * L;  // uint2 units per sub-walk in batch_mem

    for (uint32_t it = 0; it < p.iters; it++) {
        if (stop_requested(p.stop)) {
            *p.stopped = 1u;
            break;
        }
        const uint32_t cur = (uint32_t)(n & 1u), prev = cur ^ 1u;
        fe prod;
        // u_{n-1} is loaded with u_n and arrives during the hashing of u_n;
        // the backward pass loads a_j one sub-walk ahead, and D_j arrives
        // during the product that does not need it.
        for (int j = 0; j < K; j++) {
            uint32_t uw[8], vw[8];
            sw_load_words(st, L, K, cur, j, lane, uw);
            sw_load_words(st, L, K, prev, j, lane, vw);
            const uint32_t t = (uint32_t)j * L + lane;
            if (it == 0 && t < p.dump_threads) {
                uint32_t* d = p.dump + (size_t)t * 8;
#pragma unroll
                for (int k = 0; k < 8; k++) d[k] = uw[k];
            }
            uint32_t addr[4];
            lxmf_address_words(uw, p.sha, addr);
            x0 ^= addr[0];
            x1 ^= addr[1];
            x2 ^= addr[2];
            x3 ^= addr[3];
            append_hit(prefix_match(addr, p.sha), t, n, addr, p.hits, p.hit_count, p.max_hits);
            // the step: N = (u q - 1)^2, D = (u - q)^2 u_{n-1}
            const fe u = F::from_words(uw), v = F::from_words(vw);
            const fe num = F::sq(F::sub(F::mul(u, q), F::one()));
            const fe den = F::mul(F::sq(F::sub(u, q)), v);
            uint2* slot = p.batch_mem + (size_t)j * slot_stride;
            if (j == 0) {
                store_fe<F>(slot, L, lane, num);
                prod = den;
            } else {
                store_fe<F>(slot, L, lane, F::mul(num, prod));
                prod = F::mul(prod, den);
            }
            store_fe<F>(slot + (size_t)kPairs * L, L, lane, den);
        }
        fe inv = F::invert(prod);
        fe a_next = load_fe<F>(p.batch_mem + (size_t)(K - 1) * slot_stride, L, lane);
        for (int j = K - 1; j >= 0; j--) {
            const uint2* slot = p.batch_mem + (size_t)j * slot_stride;
            const fe a = a_next;
            fe d;
            if (j > 0) {
                d = load_fe<F>(slot + (size_t)kPairs * L, L, lane);
                a_next = load_fe<F>(slot - slot_stride, L, lane);
            }
            const fe u = F::mul(a, inv);
            if (j > 0) inv = F::mul(inv, d);
            uint32_t uw[8];
            F::to_words(u, uw);
            sw_store_words(st, L, K, prev, j, lane, uw);
        }
        n += 1;
    }

    p.ns[lane] = n;
    p.xor_acc[0 * L + lane] = x0;
    p.xor_acc[1 * L + lane] = x1;
    p.xor_acc[2 * L + lane] = x2;
    p.xor_acc[3 * L + lane] = x3;
    // keys walked: one per sub-walk per step
    atomicAdd(p.batches, (unsigned long long)((n - n0) * K));
}

// Sub-walk t = j x lanes + lane from u(k - 8) and u(k) (16 little-endian
// words per sub-walk, the reference's init format), made canonical.
template <class F>
LXV_D void init_sw_from_u(uint2* st, uint64_t* ns, uint32_t* xor_acc, const uint32_t* upc, uint32_t lanes,
                          uint32_t K, uint32_t t) {
    const uint32_t j = t / lanes, lane = t % lanes;
    uint32_t w[8], c[8];
#pragma unroll
    for (int i = 0; i < 8; i++) w[i] = upc[(size_t)t * 16 + i];
    F::to_words(F::from_words(w), c);
    sw_store_words(st, lanes, K, 1, j, lane, c);  // u_{-1}
#pragma unroll
    for (int i = 0; i < 8; i++) w[i] = upc[(size_t)t * 16 + 8 + i];
    F::to_words(F::from_words(w), c);
    sw_store_words(st, lanes, K, 0, j, lane, c);  // u_0
    if (j == 0) {
        ns[lane] = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) xor_acc[(size_t)k * lanes + lane] = 0;
    }
}

// Sub-walk t by a Montgomery ladder on the device (as init_lane_ladder):
// k_t = start + 8 t stride, m = k_t / 8 - 1; the ladder over Q gives (X:Z)
// of mQ and (m + 1)Q, made affine with one inversion of Z_m Z_{m+1}.
template <class F>
LXV_D void init_sw_ladder(uint2* st, uint64_t* ns, uint32_t* xor_acc, const uint32_t base[8], uint64_t stride,
                          const uint32_t qw[8], uint32_t lanes, uint32_t K, uint32_t t) {
    typedef typename F::fe fe;
    const uint32_t j = t / lanes, lane = t % lanes;
    fe x2, z2, x3, z3;
    ladder_mq<F>(base, (uint64_t)t * stride, qw, x2, z2, x3, z3);
    const fe inv = F::invert(F::mul(z2, z3));
    uint32_t w[8];
    F::to_words(F::mul(F::mul(x2, z3), inv), w);
    sw_store_words(st, lanes, K, 1, j, lane, w);  // u(k - 8)
    F::to_words(F::mul(F::mul(x3, z2), inv), w);
    sw_store_words(st, lanes, K, 0, j, lane, w);  // u(k)
    if (j == 0) {
        ns[lane] = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) xor_acc[(size_t)k * lanes + lane] = 0;
    }
}

}  // namespace lxv
