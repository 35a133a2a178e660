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
#pragma once

#include "lxv_walk.h"

namespace lxv {

struct SlotConst {
    ShaSearch sha;
    uint32_t base[8];     // start / 8 - 1, little-endian words (ladder setup)
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
    uint32_t base[8];    // start / 8 - 1, little-endian words
    uint32_t state_off;  // first pool lane
    uint32_t vlanes;
    uint32_t pad[2];
};

struct Unit {
    uint32_t slot;
    uint32_t vwave;
    uint32_t block0;
    uint32_t nblocks;
};

struct MultiHit {
    uint32_t slot;
    uint32_t vlane;
    uint64_t n;
    uint32_t address[4];
    uint32_t block;
    uint32_t pad[3];
};

struct MultiParams {
    uint32_t qp[10], qm[10];   // u(8B) +- 1 as raw limbs of the layout
    const SlotConst* __restrict__ slots;
    SlotDyn* dyn;
    const Unit* __restrict__ units;
    const uint32_t* n_units;   // device word written with the unit list
    uint32_t* counters;        // [0] next unit, [1] hit count, [2..3] keys (u64)
    uint32_t* state;           // pool: [4 * kLimbs][pool_lanes]
    uint2* batch_mem;          // per physical lane
    MultiHit* hits;
    const uint32_t* stop;
    uint32_t pool_lanes;
    uint32_t phys_lanes;
    uint32_t max_hits;
};

LXV_D uint32_t wave_uniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

template <class F, int B>
LXV_D void multi_wave(const MultiParams& p, uint32_t plane) {
    typedef typename F::fe fe;
    const fe qp = F::load(p.qp), qm = F::load(p.qm);
    constexpr int kPairs = F::kLimbs / 2;
    const uint32_t PL = p.phys_lanes, SL = p.pool_lanes;
    const size_t slot_stride = (size_t)2 * kPairs * PL;
    const uint32_t me = __lane_id();
    unsigned long long keys = 0;
    for (;;) {
        uint32_t u = 0;
        if (me == 0) u = atomicAdd(&p.counters[0], 1u);
        u = wave_uniform(__shfl(u, 0));
        if (u >= *p.n_units || stop_requested(p.stop)) break;
        const uint32_t slot = wave_uniform(p.units[u].slot);
        const uint32_t vwave = wave_uniform(p.units[u].vwave);
        const uint32_t block0 = wave_uniform(p.units[u].block0);
        const uint32_t nblocks = wave_uniform(p.units[u].nblocks);
        const SlotConst& sc = p.slots[slot];
        const uint32_t rounds = wave_uniform(sc.rounds);
        const uint32_t v = vwave * 64u + me;
        const uint32_t sl = wave_uniform(sc.state_off) + v;
        for (uint32_t blk = block0; blk < block0 + nblocks; blk++) {
            const uint32_t fb =
                __hip_atomic_load(&p.dyn[slot].found_block, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (wave_uniform(fb) < blk) break;
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
                    uint32_t uw[8], addr[4];
                    F::to_words(u_, uw);
                    lxmf_address_words(uw, sc.sha, addr);
                    const bool match = prefix_match(addr, sc.sha);
                    const uint64_t ball = __ballot(match);
                    if (ball) {
                        const uint32_t leader = (uint32_t)__builtin_ctzll(ball);
                        uint32_t base = 0;
                        if (me == leader) {
                            base = atomicAdd(&p.counters[1], (uint32_t)__builtin_popcountll(ball));
                            atomicMin(&p.dyn[slot].found_block, blk);
                        }
                        base = __shfl(base, (int)leader);
                        if (match) {
                            const uint32_t k = base + (uint32_t)__builtin_popcountll(ball & ((1ull << me) - 1ull));
                            if (k < p.max_hits) {
                                MultiHit h;
                                h.slot = slot;
                                h.vlane = v;
                                h.n = n0 + (uint64_t)i;
                                for (int w = 0; w < 4; w++) h.address[w] = addr[w];
                                h.block = blk;
                                h.pad[0] = h.pad[1] = h.pad[2] = 0;
                                p.hits[k] = h;
                            }
                        }
                    }
                }
                keys += B;
            }
        }
    }
    // keys walked by this lane, summed per wave then once into the counter
    unsigned long long tot = keys;
    for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
    if (me == 0 && tot) atomicAdd((unsigned long long*)&p.counters[2], tot);
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

}  // namespace lxv
