// The per-lane key walk and everything a walk kernel needs, written against a
// field policy (Field10 or Field8) so both limb layouts share one walk.
//
// Each lane walks X25519 private keys k_t + 8 n. Public keys u(k + 8n) come
// from x-only Montgomery differential addition (P(n+1) = P(n) + Q with
// difference P(n-1), Q = 8B affine), and the Z coordinates of B consecutive
// keys are inverted together (Montgomery's trick), as the reference's walk
// (src/gpu/common.h:280-365, src/main.rs:255-289). The difference from the
// reference's storage: the forward pass keeps a_i = X_i * (Z_0 ... Z_{i-1})
// and Z_i (two field elements per key, in HBM, lane-interleaved), so the
// backward pass needs u_i = a_i * inv and inv *= Z_i: the same four
// multiplications per key with two thirds of the memory.
#pragma once

#include "lxv_common.h"
#include "lxv_sha.h"
#include "../../include/lxmf_vanity_hip.h"

namespace lxv {

// Everything one launch of a walk kernel reads, passed by value (kernel
// arguments land in SGPRs).
struct WalkParams {
    ShaSearch sha;
    uint32_t qp[10];             // u(8B) + 1 as raw limbs of the layout
    uint32_t qm[10];             // u(8B) - 1 as raw limbs of the layout
    uint32_t q[10];              // u(8B) as raw limbs of the layout (sub-walk form)
    uint32_t* state;             // [4 * kLimbs][lanes]: X(n-1), Z(n-1), X(n), Z(n)
    uint64_t* ns;                // [lanes]: n of the lane's next key
    uint32_t* xor_acc;           // [4][lanes]: XOR of every address walked
    uint2* batch_mem;            // [B][2][kLimbs / 2][lanes]: a_i and Z_i
    lxv_hit* hits;               // [max_hits]
    uint32_t* hit_count;         // matches found in this launch
    uint32_t* dump;              // [dump_threads][B][8] or null
    const uint32_t* stop;        // host-mapped: nonzero asks lanes to stop
    uint32_t* stopped;           // set to 1 if any lane stopped early
    unsigned long long* batches; // batches of B keys completed, all lanes
    uint32_t* queue;             // device queue counter (persistent pair walk)
    uint32_t max_hits;
    uint32_t dump_threads;
    uint32_t lanes;
    uint32_t iters;
    uint32_t resident;           // lanes of the persistent grid (its batch arrays' stride)
};

template <class F>
LXV_HD void store_fe(uint2* base, uint32_t lanes, uint32_t lane, const typename F::fe& a) {
    uint32_t w[F::kLimbs];
    F::store(a, w);
#pragma unroll
    for (int i = 0; i < F::kLimbs / 2; i++) base[(size_t)i * lanes + lane] = make_uint2(w[2 * i], w[2 * i + 1]);
}

template <class F>
LXV_HD typename F::fe load_fe(const uint2* base, uint32_t lanes, uint32_t lane) {
    uint32_t w[F::kLimbs];
#pragma unroll
    for (int i = 0; i < F::kLimbs / 2; i++) {
        const uint2 v = base[(size_t)i * lanes + lane];
        w[2 * i] = v.x;
        w[2 * i + 1] = v.y;
    }
    return F::load(w);
}

template <class F>
LXV_HD void store_state(uint32_t* st, uint32_t lanes, uint32_t lane, int which, const typename F::fe& a) {
    uint32_t w[F::kLimbs];
    F::store(a, w);
#pragma unroll
    for (int i = 0; i < F::kLimbs; i++) st[(size_t)(which * F::kLimbs + i) * lanes + lane] = w[i];
}

template <class F>
LXV_HD typename F::fe load_state(const uint32_t* st, uint32_t lanes, uint32_t lane, int which) {
    uint32_t w[F::kLimbs];
#pragma unroll
    for (int i = 0; i < F::kLimbs; i++) w[i] = st[(size_t)(which * F::kLimbs + i) * lanes + lane];
    return F::load(w);
}

// One xADD step: (xp:zp), (xc:zc) = P(n-1), P(n)  ->  P(n), P(n+1).
template <class F>
LXV_D void xadd_step(typename F::fe& xp, typename F::fe& zp, typename F::fe& xc, typename F::fe& zc,
                     const typename F::fe& qp, const typename F::fe& qm) {
    typedef typename F::fe fe;
    const fe t1 = F::mul(F::sub(xc, zc), qp);
    const fe t2 = F::mul(F::add(xc, zc), qm);
    const fe xn = F::mul(zp, F::sq(F::add(t1, t2)));
    const fe zn = F::mul(xp, F::sq(F::sub(t1, t2)));
    xp = xc;
    zp = zc;
    xc = xn;
    zc = zn;
}

// Hit compaction: one atomic per wavefront with a match, then each matching
// lane writes its record at its rank among the wave's matches.
LXV_D void append_hit(bool match, uint32_t lane, uint64_t n, const uint32_t addr[4], lxv_hit* hits,
                      uint32_t* hit_count, uint32_t max_hits) {
    const uint64_t ball = __ballot(match);
    if (ball == 0) return;
    const uint32_t leader = (uint32_t)__builtin_ctzll(ball);
    const uint32_t me = __lane_id();
    uint32_t base = 0;
    if (me == leader) base = atomicAdd(hit_count, (uint32_t)__builtin_popcountll(ball));
    base = __shfl(base, (int)leader);
    if (match) {
        const uint64_t below = ball & ((1ull << me) - 1ull);
        const uint32_t slot = base + (uint32_t)__builtin_popcountll(below);
        if (slot < max_hits) {
            lxv_hit h;
            h.thread = lane;
            h.reserved = 0;
            h.n = n;
            for (int k = 0; k < 4; k++) h.address[k] = addr[k];
            hits[slot] = h;
        }
    }
}

LXV_D bool stop_requested(const uint32_t* stop) {
    return __hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0;
}

// The walk of one lane: `iters` batches of B keys. The four walk
// coordinates live in registers only during the forward pass; they go back
// to HBM before the backward pass (inversion, encoding, hashing), which then
// has the register file to itself.
template <class F, int B>
LXV_D void walk_lane(const WalkParams& p, uint32_t lane) {
    typedef typename F::fe fe;
    const uint32_t L = p.lanes;
    const uint64_t n0 = p.ns[lane];
    uint64_t n = n0;
    uint32_t x0 = p.xor_acc[0 * L + lane], x1 = p.xor_acc[1 * L + lane];
    uint32_t x2 = p.xor_acc[2 * L + lane], x3 = p.xor_acc[3 * L + lane];
    const fe qp = F::load(p.qp);
    const fe qm = F::load(p.qm);
    constexpr int kPairs = F::kLimbs / 2;
    const size_t slot_stride = (size_t)2 * kPairs * L;  // uint2 units per key

    for (uint32_t it = 0; it < p.iters; it++) {
        if (stop_requested(p.stop)) {
            *p.stopped = 1u;
            break;
        }
        fe prod;
        {
            fe xp = load_state<F>(p.state, L, lane, 0);
            fe zp = load_state<F>(p.state, L, lane, 1);
            fe xc = load_state<F>(p.state, L, lane, 2);
            fe zc = load_state<F>(p.state, L, lane, 3);
            // Forward: store a_i = X_i * (Z_0 .. Z_{i-1}) and Z_i, step the walk.
            prod = zc;
            store_fe<F>(p.batch_mem, L, lane, xc);
            store_fe<F>(p.batch_mem + (size_t)kPairs * L, L, lane, zc);
            xadd_step<F>(xp, zp, xc, zc, qp, qm);
            for (int i = 1; i < B; i++) {
                uint2* slot = p.batch_mem + (size_t)i * slot_stride;
                store_fe<F>(slot, L, lane, F::mul(xc, prod));
                store_fe<F>(slot + (size_t)kPairs * L, L, lane, zc);
                prod = F::mul(prod, zc);
                xadd_step<F>(xp, zp, xc, zc, qp, qm);
            }
            store_state<F>(p.state, L, lane, 0, xp);
            store_state<F>(p.state, L, lane, 1, zp);
            store_state<F>(p.state, L, lane, 2, xc);
            store_state<F>(p.state, L, lane, 3, zc);
        }
        // One inversion for the batch, then backwards: u_i = a_i * inv,
        // inv *= Z_i.
        fe inv = F::invert(prod);
        for (int i = B - 1; i >= 0; i--) {
            const uint2* slot = p.batch_mem + (size_t)i * slot_stride;
            const fe a = load_fe<F>(slot, L, lane);
            const fe u = F::mul(a, inv);
            if (i > 0) inv = F::mul(inv, load_fe<F>(slot + (size_t)kPairs * L, L, lane));
            uint32_t uw[8];
            F::to_words(u, uw);
            if (it == 0 && lane < p.dump_threads) {
                uint32_t* d = p.dump + ((size_t)lane * B + i) * 8;
#pragma unroll
                for (int k = 0; k < 8; k++) d[k] = uw[k];
            }
            uint32_t addr[4];
            lxmf_address_words(uw, p.sha, addr);
            x0 ^= addr[0];
            x1 ^= addr[1];
            x2 ^= addr[2];
            x3 ^= addr[3];
            append_hit(prefix_match(addr, p.sha), lane, n + (uint64_t)i, addr, p.hits, p.hit_count, p.max_hits);
        }
        n += B;
    }

    p.ns[lane] = n;
    p.xor_acc[0 * L + lane] = x0;
    p.xor_acc[1 * L + lane] = x1;
    p.xor_acc[2 * L + lane] = x2;
    p.xor_acc[3 * L + lane] = x3;
    atomicAdd(p.batches, (unsigned long long)((n - n0) / B));
}

// State of lane t from u(k - 8) and u(k) (16 little-endian words): as
// init_thread, src/gpu/common.h:260-278.
template <class F>
LXV_D void init_lane_from_u(uint32_t* state, uint64_t* ns, uint32_t* xor_acc, const uint32_t* upc, uint32_t lanes,
                            uint32_t lane) {
    uint32_t w[8];
#pragma unroll
    for (int i = 0; i < 8; i++) w[i] = upc[(size_t)lane * 16 + i];
    store_state<F>(state, lanes, lane, 0, F::from_words(w));
#pragma unroll
    for (int i = 0; i < 8; i++) w[i] = upc[(size_t)lane * 16 + 8 + i];
    store_state<F>(state, lanes, lane, 2, F::from_words(w));
    store_state<F>(state, lanes, lane, 1, F::one());
    store_state<F>(state, lanes, lane, 3, F::one());
    ns[lane] = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) xor_acc[(size_t)k * lanes + lane] = 0;
}

// One step of the X25519 ladder (RFC 7748 section 5) over the point with
// u-coordinate x1, after the conditional swap: (x2:z2) doubles and (x3:z3)
// becomes their sum.
template <class F>
LXV_HD void ladder_step(const typename F::fe& x1, typename F::fe& x2, typename F::fe& z2, typename F::fe& x3,
                        typename F::fe& z3) {
    typedef typename F::fe fe;
    const fe a = F::add(x2, z2), aa = F::sq(a), b = F::sub(x2, z2), bb = F::sq(b);
    const fe e = F::sub(aa, bb);
    const fe c = F::add(x3, z3), d = F::sub(x3, z3);
    const fe da = F::mul(d, a), cb = F::mul(c, b);
    x3 = F::sq(F::add(da, cb));
    z3 = F::mul(x1, F::sq(F::sub(da, cb)));
    x2 = F::mul(aa, bb);
    z2 = F::mul(e, F::add(aa, F::mul_small(e, 121665)));
}

// The ladder over the point with u-coordinate x1 for the low `bits` bits of
// m: (X:Z) of mP and (m + 1)P.
template <class F>
LXV_HD void ladder_u64(uint64_t m, int bits, const typename F::fe& x1, typename F::fe& x2, typename F::fe& z2,
                       typename F::fe& x3, typename F::fe& z3) {
    x2 = F::one();
    z2 = F::zero();
    x3 = x1;
    z3 = F::one();
    uint32_t swap = 0;
    for (int i = bits - 1; i >= 0; i--) {
        const uint32_t bit = (uint32_t)(m >> i) & 1u;
        swap ^= bit;
        F::cswap(x2, x3, swap);
        F::cswap(z2, z3, swap);
        swap = bit;
        ladder_step<F>(x1, x2, z2, x3, z3);
    }
    F::cswap(x2, x3, swap);
    F::cswap(z2, z3, swap);
}

// The X25519 ladder over Q = u(qw) for the scalar m = base + off (base as 8
// little-endian words, off < 2^63, no wrap): (X:Z) of mQ and (m + 1)Q.
template <class F>
LXV_HD void ladder_mq(const uint32_t base[8], uint64_t off, const uint32_t qw[8], typename F::fe& x2,
                     typename F::fe& z2, typename F::fe& x3, typename F::fe& z3) {
    typedef typename F::fe fe;
    uint32_t m[8];
    uint64_t carry = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint64_t add = i == 0 ? (off & 0xffffffffu) : (i == 1 ? (off >> 32) : 0);
        const uint64_t v = (uint64_t)base[i] + add + carry;
        m[i] = (uint32_t)v;
        carry = v >> 32;
    }
    const fe x1 = F::from_words(qw);
    x2 = F::one();
    z2 = F::zero();
    x3 = x1;
    z3 = F::one();
    uint32_t swap = 0;
    for (int wi = 7; wi >= 0; wi--) {
        const uint32_t word = m[wi];
        for (int bi = 31; bi >= 0; bi--) {
            const uint32_t bit = (word >> bi) & 1u;
            swap ^= bit;
            F::cswap(x2, x3, swap);
            F::cswap(z2, z3, swap);
            swap = bit;
            ladder_step<F>(x1, x2, z2, x3, z3);
        }
    }
    F::cswap(x2, x3, swap);
    F::cswap(z2, z3, swap);
}

// State of lane t by a Montgomery ladder on the device: with k = start + 8 t
// stride and m = k / 8 - 1, the ladder over Q = 8B gives (X:Z) of mQ =
// u(k - 8) and (m + 1)Q = u(k) with difference Q, which is exactly the
// walk's (P(n-1), P(n)) in projective form, so no inversion is needed.
// base = start / 8 - 1 as 8 little-endian words (start is a multiple of 8
// with bit 254 set, so this never wraps); lane * stride < 2^63.
template <class F>
LXV_D void init_lane_ladder(uint32_t* state, uint64_t* ns, uint32_t* xor_acc, const uint32_t base[8],
                            uint64_t stride, const uint32_t qw[8], uint32_t lanes, uint32_t lane) {
    typename F::fe x2, z2, x3, z3;
    ladder_mq<F>(base, (uint64_t)lane * stride, qw, x2, z2, x3, z3);
    store_state<F>(state, lanes, lane, 0, x2);
    store_state<F>(state, lanes, lane, 1, z2);
    store_state<F>(state, lanes, lane, 2, x3);
    store_state<F>(state, lanes, lane, 3, z3);
    ns[lane] = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) xor_acc[(size_t)k * lanes + lane] = 0;
}

}  // namespace lxv
