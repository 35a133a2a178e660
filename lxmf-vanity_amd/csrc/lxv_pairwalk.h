// The pair walk: one walk per lane over consecutive keys k + 8n, as the
// reference's walk (src/gpu/common.h:280-365, src/main.rs:255-289), with the
// public keys of a batch of B keys computed around an affine centre instead
// of by B chained differential additions.
//
// A lane holds the affine point C = (x, y) = (k/8 + c) Q of its batch centre,
// key index c = n + H with H = B/2 - 1 (Q = 8B; T_i = i Q from
// pair_table.inc). The batch's keys are
//
//     c                       x(C)
//     c + i, c - i (i <= H)   x(C + T_i), x(C - T_i)
//     c + H + 1 = n + B - 1   x(C + T_{H+1})
//
// and the next centre is C + T_B. For T = (u, v) the affine sums are
//
//     x(C +- T) = l^2 - A - x - u,   l = (v -+ y) / (u - x)
//
// (only l^2 is needed, so the sign of l for C - T is free), so C + T_i and
// C - T_i share the inverse of u_i - x. The B/2 + 1 denominators of a batch
// are inverted together with Montgomery's trick: per key about 2.5 M + 1 S
// plus the inversion's share, against 8 M + 2 S for the projective walk of
// lxv_walk.h at the same B keys per inversion.
//
// Denominators are numbered e = 0 (the centre step, T_B) and e = 1 .. H + 1
// (T_e). The forward pass stores the prefix products d_0 ... d_{e-1} for
// e >= 1 (B/2 field elements per lane, lane-interleaved in HBM as the
// projective walk's batch arrays); the backward pass walks e downwards,
// two keys per pair.
#pragma once

#include "fe10.h"
#include "lxv_walk.h"
#include "pair_table.inc"

namespace lxv {

#if defined(__HIP_DEVICE_COMPILE__)
static __constant__ uint32_t kPairTab[LXV_PAIR_TABLE_ENTRIES * 16] = LXV_PAIR_TABLE_INIT;
#else
static const uint32_t kPairTab[LXV_PAIR_TABLE_ENTRIES * 16] = LXV_PAIR_TABLE_INIT;
#endif

constexpr uint32_t kMontA = 486662;  // Curve25519: v^2 = u^3 + A u^2 + u

// -v(T_i) = p - v(T_i) as canonical little-endian words (v is never 0: no
// table point has order 2). The second key of a pair, x(C - T_i), takes its
// slope from (-v - y) / (u - x), the negative of (v + y) / (u - x), whose
// square is the same: both keys of a pair are then (v' - y) times the pair's
// inverse, with v' picked from one table or the other by a uniform index.
struct PairTabNV {
    uint32_t w[LXV_PAIR_TABLE_ENTRIES * 8];
};

constexpr PairTabNV make_pair_tab_nv() {
    constexpr uint32_t src[LXV_PAIR_TABLE_ENTRIES * 16] = LXV_PAIR_TABLE_INIT;
    // p = 2^255 - 19
    constexpr uint32_t p[8] = {0xffffffedu, 0xffffffffu, 0xffffffffu, 0xffffffffu,
                               0xffffffffu, 0xffffffffu, 0xffffffffu, 0x7fffffffu};
    PairTabNV t{};
    for (int e = 0; e < LXV_PAIR_TABLE_ENTRIES; e++) {
        int64_t b = 0;
        for (int i = 0; i < 8; i++) {
            const int64_t d = (int64_t)p[i] - src[e * 16 + 8 + i] + b;
            t.w[e * 8 + i] = (uint32_t)d;
            b = d < 0 ? -1 : 0;
        }
    }
    return t;
}

// The table in the limbs of the 10 x 25.5 layout (30 words per entry: u, v,
// -v), converted at compile time, so that the walk loads the limbs as they
// are instead of converting words on every use.
struct PairTab10 {
    uint32_t w[LXV_PAIR_TABLE_ENTRIES * 30];
};

constexpr PairTab10 make_pair_tab10() {
    constexpr uint32_t src[LXV_PAIR_TABLE_ENTRIES * 16] = LXV_PAIR_TABLE_INIT;
    constexpr PairTabNV nv = make_pair_tab_nv();
    PairTab10 t{};
    for (int e = 0; e < LXV_PAIR_TABLE_ENTRIES; e++)
        for (int h = 0; h < 3; h++) {
            const uint32_t* words = h < 2 ? src + e * 16 + h * 8 : nv.w + e * 8;
            for (int i = 0; i < 10; i++) {
                const int bit = fe10_bit(i), word = bit >> 5, shift = bit & 31;
                const uint32_t lo = words[word];
                const uint32_t hi = word + 1 < 8 ? words[word + 1] : 0u;
                const uint32_t v = shift ? (lo >> shift) | (hi << (32 - shift)) : lo;
                t.w[e * 30 + h * 10 + i] = v & fe10_mask(i);
            }
        }
    return t;
}

#if defined(__HIP_DEVICE_COMPILE__)
static __constant__ PairTabNV kPairTabNV = make_pair_tab_nv();
static __constant__ PairTab10 kPairTab10 = make_pair_tab10();
#else
static const PairTabNV kPairTabNV = make_pair_tab_nv();
static const PairTab10 kPairTab10 = make_pair_tab10();
#endif

LXV_HD int pair_tab_entry(int i) { return i <= 128 ? i - 1 : 128; }

// u and v of T_i (1 <= i <= 128, or i = 256) in the layout of F.
template <class F>
LXV_HD typename F::fe pair_tab_u(int i) {
    if constexpr (F::kLimbs == 10)
        return F::load(kPairTab10.w + 30 * pair_tab_entry(i));
    else
        return F::from_words(kPairTab + 16 * pair_tab_entry(i));
}
template <class F>
LXV_HD typename F::fe pair_tab_v(int i) {
    if constexpr (F::kLimbs == 10)
        return F::load(kPairTab10.w + 30 * pair_tab_entry(i) + 10);
    else
        return F::from_words(kPairTab + 16 * pair_tab_entry(i) + 8);
}
// v(T_i), or -v(T_i) when neg (a uniform choice between two tables).
template <class F>
LXV_HD typename F::fe pair_tab_v_signed(int i, bool neg) {
    if constexpr (F::kLimbs == 10)
        return F::load(kPairTab10.w + 30 * pair_tab_entry(i) + (neg ? 20 : 10));
    else
        return F::from_words(neg ? kPairTabNV.w + 8 * pair_tab_entry(i) : kPairTab + 16 * pair_tab_entry(i) + 8);
}

template <class F>
LXV_HD typename F::fe fe_small(uint32_t k) {
    const uint32_t w[8] = {k, 0, 0, 0, 0, 0, 0, 0};
    return F::from_words(w);
}

// One batch of B keys around the centre (xc, yc), which moves on to the next
// batch's centre. emit(i, x) receives the public key of key n + i (i < B) as
// a field element, not necessarily tight (for encoding only); mem is the
// lane's prefix-product array (B/2 elements, lane-interleaved with stride
// lanes).
template <class F, int B, class Emit>
LXV_HD void pair_batch(typename F::fe& xc, typename F::fe& yc, uint2* mem, uint32_t lanes, uint32_t lane,
                       Emit&& emit) {
    typedef typename F::fe fe;
    static_assert(B >= 4 && B <= 256 && (B & (B - 1)) == 0, "B is a power of two in [4, 256]");
    constexpr int H = B / 2 - 1;
    constexpr int E = B / 2 + 1;
    constexpr int kPairs = F::kLimbs / 2;
    const size_t slot = (size_t)kPairs * lanes;

    // forward: d_0 = u(T_B) - x, then d_e = u(T_e) - x
    fe prod = F::sub(pair_tab_u<F>(B), xc);
    for (int e = 1; e < E; e++) {
        store_fe<F>(mem + (size_t)(e - 1) * slot, lanes, lane, prod);
        prod = F::mul(prod, F::sub(pair_tab_u<F>(e), xc));
    }
    fe inv = F::invert(prod);
    const fe sa = F::tight(F::add(xc, fe_small<F>(kMontA)));

    // backward, by denominator: e = H + 1 (its single key C + T_{H+1}), then
    // e = H .. 1 (the pair C + T_e, C - T_e sharing its inverse), then e = 0
    // for the centre's own key (the centre step's inverse is left in inv).
    // Every key is x = (ie (v' - y))^2 - s with v' = v_e or -v_e: for e = 0,
    // ie = 0 and s = -x(C) give x(C) itself. The keys of one e run in an
    // inner loop that is not unrolled, so a single copy of the hash follows
    // and the inverse, s and the running product stay put in registers.
    fe ie, s;
    for (int e = H + 1; e >= 0; e--) {
        if (e > 0) {
            const fe ue = pair_tab_u<F>(e);
            ie = F::mul(inv, load_fe<F>(mem + (size_t)(e - 1) * slot, lanes, lane));
            inv = F::mul(inv, F::sub(ue, xc));
            s = F::tight(F::add(sa, ue));
        } else {
            ie = F::zero();
            s = F::tight(F::sub(F::zero(), xc));
        }
        const int keys = (e == H + 1 || e == 0) ? 1 : 2;
#pragma clang loop unroll(disable)
        for (int h = 0; h < keys; h++) {
            const fe x = F::sub(F::sq(F::mul(F::sub(pair_tab_v_signed<F>(e > 0 ? e : 1, h != 0), yc), ie)), s);
            emit(e == 0 ? H : h == 0 ? H + e : H - e, x);
        }
    }

    // the next centre, C + T_B: inv is now 1 / d_0
    const fe ub = pair_tab_u<F>(B), vb = pair_tab_v<F>(B);
    const fe l = F::mul(F::sub(vb, yc), inv);
    const fe x2 = F::tight(F::sub(F::sq(l), F::tight(F::add(sa, ub))));
    yc = F::tight(F::sub(F::mul(l, F::sub(xc, x2)), yc));
    xc = x2;
}

// y of R from x(R) = xr / zr and x(R + T) = xs / zs for T = (ut, vt) with
// vt != 0 (Okeya-Sakurai recovery: the affine sum formula solved for y(R)),
//     2 vt y = (ut x + 1)(ut + x + 2A) - 2A - x(R + T) (x - ut)^2,
// returned as the affine pair (x, y) with one inversion.
template <class F>
LXV_HD void recover_affine(const typename F::fe& xr, const typename F::fe& zr, const typename F::fe& xs,
                           const typename F::fe& zs, const typename F::fe& ut, const typename F::fe& vt,
                           typename F::fe& x, typename F::fe& y) {
    typedef typename F::fe fe;
    const fe a2 = fe_small<F>(2 * kMontA);
    const fe uz = F::mul(ut, zr);
    const fe t1 = F::add(F::mul(ut, xr), zr);                  // (ut x + 1) zr
    const fe t2 = F::tight(F::add(F::add(uz, xr), F::mul(a2, zr)));  // (ut + x + 2A) zr
    const fe zz = F::mul(F::sq(zr), zs);                       // zr^2 zs
    const fe d = F::sub(xr, uz);                               // (x - ut) zr
    const fe num = F::tight(F::sub(F::sub(F::mul(F::mul(t1, t2), zs), F::mul(a2, zz)), F::mul(xs, F::sq(d))));
    const fe v2 = F::tight(F::add(vt, vt));
    const fe den = F::mul(v2, zz);                             // 2 vt zr^2 zs
    const fe inv = F::invert(den);
    y = F::mul(num, inv);
    x = F::mul(F::mul(F::mul(xr, zr), F::mul(zs, v2)), inv);
}

// (x, y) + (ut, vt) by the affine addition formula (x != ut).
template <class F>
LXV_HD void affine_add(typename F::fe& x, typename F::fe& y, const typename F::fe& ut, const typename F::fe& vt) {
    typedef typename F::fe fe;
    const fe l = F::mul(F::sub(vt, y), F::invert(F::sub(ut, x)));
    const fe x2 = F::tight(F::sub(F::sub(F::sub(F::sq(l), fe_small<F>(kMontA)), x), ut));
    y = F::tight(F::sub(F::mul(l, F::sub(x, x2)), y));
    x = x2;
}

// (x, y) + T_i (x != u(T_i)).
template <class F>
LXV_HD void affine_add_tab(typename F::fe& x, typename F::fe& y, int i) {
    affine_add<F>(x, y, pair_tab_u<F>(i), pair_tab_v<F>(i));
}

// The centre of a lane's first batch from the reference's start data: u(k - 8)
// and u(k) (8 little-endian words each). R = kB = (k/8) Q has R - Q = u(k - 8),
// so y(R) follows from the recovery with T = -Q; the centre is R + T_H.
template <class F, int B>
LXV_HD void pair_centre_from_u(const uint32_t uprev[8], const uint32_t ucur[8], typename F::fe& xc,
                               typename F::fe& yc) {
    typedef typename F::fe fe;
    const fe uq = pair_tab_u<F>(1), vq = pair_tab_v<F>(1);
    recover_affine<F>(F::from_words(ucur), F::one(), F::from_words(uprev), F::one(), uq, F::sub(F::zero(), vq), xc,
                      yc);
    affine_add_tab<F>(xc, yc, B / 2 - 1);
}

// The centre of a lane's first batch by a ladder: m = base + off with
// base = k_0 / 8 + H; the ladder over Q gives mQ and (m + 1)Q projectively,
// and the recovery with T = Q makes mQ affine with its y.
template <class F>
LXV_HD void pair_centre_ladder(const uint32_t base[8], uint64_t off, typename F::fe& xc, typename F::fe& yc) {
    typedef typename F::fe fe;
    uint32_t qw[8];
#pragma unroll
    for (int k = 0; k < 8; k++) qw[k] = kPairTab[k];
    fe x2, z2, x3, z3;
    ladder_mq<F>(base, off, qw, x2, z2, x3, z3);
    recover_affine<F>(x2, z2, x3, z3, pair_tab_u<F>(1), pair_tab_v<F>(1), xc, yc);
}

// An affine point as canonical little-endian words.
struct PairAnchor {
    uint32_t x[8], y[8];
};

// The anchor mQ of a lane set-up, for a scalar m (8 little-endian words) with
// mQ and (m + 1)Q not the identity.
template <class F>
LXV_HD void pair_anchor(const uint32_t m[8], PairAnchor& out) {
    typename F::fe x, y;
    pair_centre_ladder<F>(m, 0, x, y);
    F::to_words(x, out.x);
    F::to_words(y, out.y);
}

// Lane t's first centre (b + t s) Q = P0 + t S from two anchors computed once
// per set-up: P0 = bQ (lane 0's centre) and S = sQ. A ladder over the `bits`
// low bits of t (t < 2^bits) on S gives tS and (t + 1)S projectively, the
// recovery with T = S makes tS affine, and one affine addition adds P0: about
// a sixth of a full 255-step ladder per lane at 2^24 lanes. The addition needs
// x(tS) != x(P0), i.e. t s != +-b (mod l). That holds for every set-up here:
// b = k_0 / 8 + c with k_0 a valid start (2^254 <= k_0 < 2^255 - 2^72, which
// random_start_scalar and start_scalar_ok guarantee) and |c| < 2^7 lies in
// [2^251 - 2^7, l - 2^69), while 1 <= t s < 2^63.
template <class F>
LXV_HD void pair_centre_offset(const PairAnchor& p0, const PairAnchor& s, uint64_t t, int bits, typename F::fe& xc,
                               typename F::fe& yc) {
    typedef typename F::fe fe;
    xc = F::from_words(p0.x);
    yc = F::from_words(p0.y);
    if (t == 0) return;
    const fe us = F::from_words(s.x), vs = F::from_words(s.y);
    fe x2, z2, x3, z3, xt, yt;
    ladder_u64<F>(t, bits, us, x2, z2, x3, z3);
    recover_affine<F>(x2, z2, x3, z3, us, vs, xt, yt);
    affine_add<F>(xc, yc, xt, yt);
}

#if defined(__HIPCC__)
// The launch's parameters, re-read from the kernel argument segment (scalar
// loads) where a key uses them, rather than held in SGPRs across the whole
// walk: with the hash's per-search constants (45 words) live throughout, the
// SGPRs ran out, the compiler spilled them to VGPR lanes, and every key paid
// about 60 v_readlane to fetch them back. The empty asm makes the pointer
// opaque, so the loads stay next to their use. Only for kernels whose one
// argument is the WalkParams (it then sits at the segment's start).
__device__ __forceinline__ const WalkParams& walk_params_reloaded() {
    typedef __attribute__((address_space(4))) const WalkParams* kernarg_t;
    kernarg_t kp = (kernarg_t)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(kp));
    return *(const WalkParams*)kp;
}

// The walk of one lane: `iters` batches of B keys. The centre is the lane's
// whole state (state slots 0 and 1: x and y); the batch arrays are those of
// lane mem_lane of mem_lanes.
template <class F, int B>
__device__ __forceinline__ void walk_pair_lane(const WalkParams& p, uint32_t lane, uint32_t mem_lanes,
                                               uint32_t mem_lane) {
    typedef typename F::fe fe;
    const uint32_t L = p.lanes;
    const uint64_t n0 = p.ns[lane];
    uint64_t n = n0;
    uint32_t x0 = p.xor_acc[0 * L + lane], x1 = p.xor_acc[1 * L + lane];
    uint32_t x2 = p.xor_acc[2 * L + lane], x3 = p.xor_acc[3 * L + lane];
    fe xc = load_state<F>(p.state, L, lane, 0);
    fe yc = load_state<F>(p.state, L, lane, 1);

    for (uint32_t it = 0; it < p.iters; it++) {
        if (stop_requested(p.stop)) {
            *p.stopped = 1u;
            break;
        }
        const bool dumping = it == 0 && lane < p.dump_threads;
        pair_batch<F, B>(xc, yc, p.batch_mem, mem_lanes, mem_lane, [&](int i, const fe& x) {
            uint32_t uw[8];
            F::to_words(x, uw);
            if (dumping) {
                uint32_t* d = p.dump + ((size_t)lane * B + i) * 8;
#pragma unroll
                for (int k = 0; k < 8; k++) d[k] = uw[k];
            }
            const WalkParams& kp = walk_params_reloaded();
            uint32_t addr[4];
            lxmf_address_words(uw, kp.sha, addr);
            x0 ^= addr[0];
            x1 ^= addr[1];
            x2 ^= addr[2];
            x3 ^= addr[3];
            append_hit(prefix_match(addr, kp.sha), lane, n + (uint64_t)i, addr, kp.hits, kp.hit_count, kp.max_hits);
        });
        n += B;
    }

    store_state<F>(p.state, L, lane, 0, xc);
    store_state<F>(p.state, L, lane, 1, yc);
    p.ns[lane] = n;
    p.xor_acc[0 * L + lane] = x0;
    p.xor_acc[1 * L + lane] = x1;
    p.xor_acc[2 * L + lane] = x2;
    p.xor_acc[3 * L + lane] = x3;
    atomicAdd(p.batches, (unsigned long long)((n - n0) / B));
}

// The persistent form of a launch: the grid is only the resident waves
// (p.resident lanes), and each wave takes groups of 64 lanes from a device
// counter until every group has walked its batches. Waves that run faster
// take more groups, so the launch does not end with its slowest wave, and
// the batch arrays belong to the resident lanes only. Every wave leaves once
// the counter passes the last group.
template <class F, int B>
__device__ __forceinline__ void walk_pair_queue(const WalkParams& p, uint32_t plane) {
    const uint32_t groups = p.lanes / 64u;
    for (;;) {
        uint32_t u = 0;
        if (__lane_id() == 0) u = atomicAdd(p.queue, 1u);
        u = __builtin_amdgcn_readfirstlane(__shfl(u, 0));
        if (u >= groups) break;
        walk_pair_lane<F, B>(p, u * 64u + __lane_id(), p.resident, plane);
    }
}
#endif

}  // namespace lxv
