// The batched search kernel: many independent searches in one persistent
// kernel (no counterpart in the reference, which runs one search per process).
//
// Search s owns V_s virtual lanes; virtual lane v walks keys
// start_s + 8 (v * stride + n), n = 0, 1, 2, ..., and all of a search's virtual
// lanes advance together in blocks of R batches of B keys (block b covers
// n in [b R B, (b + 1) R B)). The walk state of every virtual lane stays in
// an HBM pool between launches. A work unit is (search, virtual wave of 64
// lanes, first block, block count); physical waves take units from a device
// queue one at a time, so when a search is finished (or has a hit in an
// earlier block) its waves move on to other searches within the launch.
//
// A search's answer is its first hit in walk order: the lowest block with a
// hit, then the lowest (n, v) in it. A unit stops before a block above the
// search's lowest hit block, never inside one, so every block up to the one
// with the answer is walked completely and the answer does not depend on what
// else runs.
//
// Two walk forms share the unit logic: the reference's projective walk
// (multi_wave: four field elements X(n-1), Z(n-1), X(n), Z(n) of pool state
// per virtual lane) and the pair walk of lxv_pairwalk.h (multi_pair_wave: the
// affine batch centre (x, y), two field elements per virtual lane).
#pragma once

#include "lxv_pairwalk.h"
#include "lxv_walk.h"

namespace lxv {

struct SlotConst {
    ShaSearch sha;
    uint32_t base[8];     // ladder set-up, little-endian words: start / 8 - 1
                          // (projective) or start / 8 + B/2 - 1 (pair walk)
    uint64_t stride;      // key distance between virtual lanes
    uint32_t vlanes;      // V (a multiple of 64)
    uint32_t state_off;   // first pool lane of this search
    uint32_t rounds;      // R: batches per block
    uint32_t slot;        // index of this slot
};

struct SlotDyn {
    uint32_t found_block;  // lowest block with a hit (0xffffffff: none yet)
    uint32_t pad[15];      // one 64-byte line per slot
};

// One admitted search's virtual lanes, set up by the ladder in one batched
// launch per step (block row = entry).
struct MultiInitEntry {
    uint32_t base[8];    // as SlotConst::base
    uint32_t state_off;  // first pool lane
    uint32_t vlanes;
    uint32_t pad[2];
    PairAnchor p0;       // the pair walk: base Q
};

struct Unit {
    uint32_t slot;
    uint32_t vwave;
    uint32_t block0;
    uint32_t nblocks;
};

// A search's best hit of one step, as one 64-bit key ordered as walk order
// within a block: key = n << kHitVlaneBits | v (n < 2^40 keys per virtual
// lane, v < 2^22 virtual lanes), so the lowest key is the lowest (n, v); the
// block is n / (R B). No key: kNoHit.
constexpr int kHitVlaneBits = 22;
constexpr unsigned long long kNoHit = ~0ull;

LXV_HD unsigned long long hit_key(uint64_t n, uint32_t v) { return (n << kHitVlaneBits) | v; }

struct MultiParams {
    uint32_t qp[10], qm[10];   // u(8B) +- 1 as raw limbs of the layout
    const SlotConst* __restrict__ slots;
    SlotDyn* dyn;
    const Unit* __restrict__ units;
    const uint32_t* n_units;   // device word written with the unit list
    uint32_t* counters;        // [0] next unit, [2..3] keys (u64)
    unsigned long long* best;  // [max slots]: this step's lowest hit key per slot (kNoHit: none)
    uint32_t* state;           // pool: [4 * kLimbs][pool_lanes] (pair walk: [2 * kLimbs][pool_lanes])
    uint2* batch_mem;          // per physical lane
    const uint32_t* stop;
    uint32_t pool_lanes;
    uint32_t phys_lanes;
};

LXV_D uint32_t wave_uniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// Hashes key n of virtual lane v (block blk) and records it if it matches.
// All lanes of a wave hash the same n (their v are consecutive), so the
// wave's lowest matching (n, v) is its lowest matching lane: that lane alone
// lowers the search's best key of this step and its lowest-hit block, with
// two atomics per wave with matches and nothing stored per hit, so the work
// of a step does not grow with its matches.
//
// The slot's hash constants are re-read (scalar loads) for every key rather
// than held in SGPRs across the unit's walk, which would spill them to VGPR
// lanes (see walk_params_reloaded); the slot table is not written while a
// step runs, so it is read through the constant address space.
LXV_D void multi_emit(const MultiParams& p, uint32_t slot, uint32_t v, uint32_t blk, uint64_t n,
                      const uint32_t uw[8]) {
    typedef __attribute__((address_space(4))) const SlotConst* slot_cptr;
    slot_cptr sp = (slot_cptr)(p.slots + slot);
    asm volatile("" : "+s"(sp));
    const ShaSearch& sha = ((const SlotConst*)sp)->sha;
    uint32_t addr[4];
    lxmf_address_words(uw, sha, addr);
    const uint64_t ball = __ballot(prefix_match(addr, sha));
    if (ball && __lane_id() == (uint32_t)__builtin_ctzll(ball)) {
        atomicMin(&p.best[slot], hit_key(n, v));
        atomicMin(&p.dyn[slot].found_block, blk);
    }
}

// The next unit of the device queue for this wave (all lanes agree), or
// false when the list is exhausted or a stop was asked for.
LXV_D bool multi_next_unit(const MultiParams& p, Unit& unit) {
    uint32_t u = 0;
    if (__lane_id() == 0) u = atomicAdd(&p.counters[0], 1u);
    u = wave_uniform(__shfl(u, 0));
    if (u >= *p.n_units || stop_requested(p.stop)) return false;
    unit.slot = wave_uniform(p.units[u].slot);
    unit.vwave = wave_uniform(p.units[u].vwave);
    unit.block0 = wave_uniform(p.units[u].block0);
    unit.nblocks = wave_uniform(p.units[u].nblocks);
    return true;
}

// Whether block blk may still be walked: not above the search's lowest block
// with a hit.
LXV_D bool multi_block_open(const MultiParams& p, uint32_t slot, uint32_t blk) {
    const uint32_t fb = __hip_atomic_load(&p.dyn[slot].found_block, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return wave_uniform(fb) >= blk;
}

LXV_D void multi_count_keys(const MultiParams& p, unsigned long long keys) {
    // keys walked by this lane, summed per wave then once into the counter
    unsigned long long tot = keys;
    for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
    if (__lane_id() == 0 && tot) atomicAdd((unsigned long long*)&p.counters[2], tot);
}

template <class F, int B>
LXV_D void multi_wave(const MultiParams& p, uint32_t plane) {
    typedef typename F::fe fe;
    const fe qp = F::load(p.qp), qm = F::load(p.qm);
    constexpr int kPairs = F::kLimbs / 2;
    const uint32_t PL = p.phys_lanes, SL = p.pool_lanes;
    const size_t slot_stride = (size_t)2 * kPairs * PL;
    unsigned long long keys = 0;
    Unit unit;
    while (multi_next_unit(p, unit)) {
        const uint32_t slot = unit.slot;
        const SlotConst& sc = p.slots[slot];
        const uint32_t rounds = wave_uniform(sc.rounds);
        const uint32_t v = unit.vwave * 64u + __lane_id();
        const uint32_t sl = wave_uniform(sc.state_off) + v;
        for (uint32_t blk = unit.block0; blk < unit.block0 + unit.nblocks; blk++) {
            if (!multi_block_open(p, slot, blk)) break;
            for (uint32_t r = 0; r < rounds; r++) {
                const uint64_t n0 = ((uint64_t)blk * rounds + r) * B;
                fe prod;
                {
                    fe xp = load_state<F>(p.state, SL, sl, 0);
                    fe zp = load_state<F>(p.state, SL, sl, 1);
                    fe xc = load_state<F>(p.state, SL, sl, 2);
                    fe zc = load_state<F>(p.state, SL, sl, 3);
                    prod = zc;
                    store_fe<F>(p.batch_mem, PL, plane, xc);
                    store_fe<F>(p.batch_mem + (size_t)kPairs * PL, PL, plane, zc);
                    xadd_step<F>(xp, zp, xc, zc, qp, qm);
                    for (int i = 1; i < B; i++) {
                        uint2* bs = p.batch_mem + (size_t)i * slot_stride;
                        store_fe<F>(bs, PL, plane, F::mul(xc, prod));
                        store_fe<F>(bs + (size_t)kPairs * PL, PL, plane, zc);
                        prod = F::mul(prod, zc);
                        xadd_step<F>(xp, zp, xc, zc, qp, qm);
                    }
                    store_state<F>(p.state, SL, sl, 0, xp);
                    store_state<F>(p.state, SL, sl, 1, zp);
                    store_state<F>(p.state, SL, sl, 2, xc);
                    store_state<F>(p.state, SL, sl, 3, zc);
                }
                fe inv = F::invert(prod);
                for (int i = B - 1; i >= 0; i--) {
                    const uint2* bs = p.batch_mem + (size_t)i * slot_stride;
                    const fe u_ = F::mul(load_fe<F>(bs, PL, plane), inv);
                    if (i > 0) inv = F::mul(inv, load_fe<F>(bs + (size_t)kPairs * PL, PL, plane));
                    uint32_t uw[8];
                    F::to_words(u_, uw);
                    multi_emit(p, slot, v, blk, n0 + (uint64_t)i, uw);
                }
                keys += B;
            }
        }
    }
    multi_count_keys(p, keys);
}

// The same units walked with the pair walk: the pool holds each virtual
// lane's batch centre (x, y).
template <class F, int B>
LXV_D void multi_pair_wave(const MultiParams& p, uint32_t plane) {
    typedef typename F::fe fe;
    const uint32_t PL = p.phys_lanes, SL = p.pool_lanes;
    unsigned long long keys = 0;
    Unit unit;
    while (multi_next_unit(p, unit)) {
        const uint32_t slot = unit.slot;
        const SlotConst& sc = p.slots[slot];
        const uint32_t rounds = wave_uniform(sc.rounds);
        const uint32_t v = unit.vwave * 64u + __lane_id();
        const uint32_t sl = wave_uniform(sc.state_off) + v;
        fe xc = load_state<F>(p.state, SL, sl, 0);
        fe yc = load_state<F>(p.state, SL, sl, 1);
        for (uint32_t blk = unit.block0; blk < unit.block0 + unit.nblocks; blk++) {
            if (!multi_block_open(p, slot, blk)) break;
            for (uint32_t r = 0; r < rounds; r++) {
                const uint64_t n0 = ((uint64_t)blk * rounds + r) * B;
                pair_batch<F, B>(xc, yc, p.batch_mem, PL, plane, [&](int i, const fe& x) {
                    uint32_t uw[8];
                    F::to_words(x, uw);
                    multi_emit(p, slot, v, blk, n0 + (uint64_t)i, uw);
                });
                keys += B;
            }
        }
        store_state<F>(p.state, SL, sl, 0, xc);
        store_state<F>(p.state, SL, sl, 1, yc);
    }
    multi_count_keys(p, keys);
}

// Virtual lanes [first, first + count) of one search set up by the ladder
// (lane v at g = v * stride): the pool analogue of init_lane_ladder.
template <class F>
LXV_D void multi_init_lane(uint32_t* pool, uint32_t pool_lanes, uint32_t state_off, const uint32_t base[8],
                           uint64_t stride, const uint32_t qw[8], uint32_t v) {
    typename F::fe x2, z2, x3, z3;
    ladder_mq<F>(base, (uint64_t)v * stride, qw, x2, z2, x3, z3);
    const uint32_t lane = state_off + v;
    store_state<F>(pool, pool_lanes, lane, 0, x2);
    store_state<F>(pool, pool_lanes, lane, 1, z2);
    store_state<F>(pool, pool_lanes, lane, 2, x3);
    store_state<F>(pool, pool_lanes, lane, 3, z3);
}

// Virtual lane v of one search set up for the pair walk: the centre of its
// first batch, (base + v stride) Q = P0 + v S with P0 = base Q (base = start /
// 8 + B/2 - 1) and S = stride Q.
template <class F>
LXV_D void multi_pair_init_lane(uint32_t* pool, uint32_t pool_lanes, uint32_t state_off, const PairAnchor& p0,
                                const PairAnchor& s, uint32_t bits, uint32_t v) {
    typename F::fe xc, yc;
    pair_centre_offset<F>(p0, s, v, (int)bits, xc, yc);
    store_state<F>(pool, pool_lanes, state_off + v, 0, xc);
    store_state<F>(pool, pool_lanes, state_off + v, 1, yc);
}

}  // namespace lxv
