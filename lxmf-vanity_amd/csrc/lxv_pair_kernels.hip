// HIP kernels of the pair walk (lxv_pairwalk.h) and their launchers.
//
//   walk_pair         one walk per lane, B keys per inversion around an affine
//                     centre; one instantiation per (limb layout, B, waves per
//                     SIMD the registers are budgeted for)
//   init_pair_from_u  the lane's first centre from host-computed u(k-8), u(k)
//                     (init_walk, src/gpu/kernel.cu:21-24)
//   init_pair_ladder  the lane's first centre from two anchors (pair_anchor)
//
// The file is compiled twice, once per limb layout (LXV_PAIR_PART = 8 or 10),
// so the two halves of the kernel table build in parallel; the launchers are
// in the 8 x 32 part and look in both halves.
#ifndef LXV_PAIR_PART
#define LXV_PAIR_PART 8
#endif
#include "fe10.h"
#include "fe8.h"
#include "lxv_kernels.h"
#include "lxv_pairwalk.h"

namespace lxv {

// Both walk kernels take the WalkParams as their only argument: walk_pair_lane
// re-reads the per-search hash constants from the start of the kernel
// argument segment (walk_params_reloaded).
template <class F, int B, int W>
__global__ __launch_bounds__(64, W) void walk_pair_kernel(const WalkParams p) {
    const uint32_t lane = blockIdx.x * 64u + threadIdx.x;
    if (lane < p.lanes) walk_pair_lane<F, B>(p, lane, p.lanes, lane);
}

// One wave per 64 resident lanes; p.lanes is a multiple of 64.
template <class F, int B, int W>
__global__ __launch_bounds__(64, W) void walk_pair_queue_kernel(const WalkParams p) {
    walk_pair_queue<F, B>(p, blockIdx.x * 64u + threadIdx.x);
}

template <class F>
LXV_D void pair_lane_reset(uint32_t* state, uint64_t* ns, uint32_t* xor_acc, uint32_t lanes, uint32_t lane,
                           const typename F::fe& xc, const typename F::fe& yc) {
    store_state<F>(state, lanes, lane, 0, xc);
    store_state<F>(state, lanes, lane, 1, yc);
    ns[lane] = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) xor_acc[(size_t)k * lanes + lane] = 0;
}

template <class F, int B>
__global__ __launch_bounds__(64) void init_pair_from_u_kernel(uint32_t* state, uint64_t* ns, uint32_t* xor_acc,
                                                              const uint32_t* upc, uint32_t lanes) {
    const uint32_t lane = blockIdx.x * 64u + threadIdx.x;
    if (lane >= lanes) return;
    uint32_t prev[8], cur[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        prev[i] = upc[(size_t)lane * 16 + i];
        cur[i] = upc[(size_t)lane * 16 + 8 + i];
    }
    typename F::fe xc, yc;
    pair_centre_from_u<F, B>(prev, cur, xc, yc);
    pair_lane_reset<F>(state, ns, xor_acc, lanes, lane, xc, yc);
}

struct PairLadderArgs {
    const PairAnchor* anchors;  // [0] P0 = (k_0 / 8 + H) Q, [1] S = stride Q
    uint32_t bits;              // lanes - 1 < 2^bits
    uint32_t lanes;
};

template <class F>
__global__ __launch_bounds__(64) void init_pair_ladder_kernel(uint32_t* state, uint64_t* ns, uint32_t* xor_acc,
                                                              const PairLadderArgs a) {
    const uint32_t lane = blockIdx.x * 64u + threadIdx.x;
    if (lane >= a.lanes) return;
    typename F::fe xc, yc;
    pair_centre_offset<F>(a.anchors[0], a.anchors[1], lane, (int)a.bits, xc, yc);
    pair_lane_reset<F>(state, ns, xor_acc, a.lanes, lane, xc, yc);
}

#if LXV_PAIR_PART == 8
// Lane i < n: out[i] = m_i Q for the scalar m_i at scalars + i * words_apart
// (canonical words, so one build serves both limb layouts).
__global__ __launch_bounds__(64) void pair_anchor_kernel(const uint32_t* scalars, uint32_t words_apart, uint32_t n,
                                                         PairAnchor* out) {
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    if (i >= n) return;
    uint32_t m[8];
#pragma unroll
    for (int k = 0; k < 8; k++) m[k] = scalars[(size_t)i * words_apart + k];
    pair_anchor<Field8>(m, out[i]);
}
#endif

typedef void (*pair_fn)(const WalkParams);
typedef void (*pair_init_fn)(uint32_t*, uint64_t*, uint32_t*, const uint32_t*, uint32_t);
typedef void (*pair_ladder_fn)(uint32_t*, uint64_t*, uint32_t*, const PairLadderArgs);

struct PairEntry {
    int batch;
    int waves;
    pair_fn fn;
    pair_fn queue_fn;  // the persistent form, or null
};

// One layout's kernels: the walks (B x waves), the per-B initialisers from
// u words, and the ladder initialiser.
struct PairPart {
    int layout;
    const PairEntry* walks;
    int n_walks;
    pair_init_fn init_from_u[5];  // B = 16, 32, 64, 128, 256
    pair_ladder_fn init_ladder;
};

#if LXV_PAIR_PART == 8
#define LXV_PF Field8
#define LXV_PL LXV_LAYOUT_8X32
#define LXV_PART_NAME kPairPart8
#else
#define LXV_PF Field10
#define LXV_PL LXV_LAYOUT_10X26
#define LXV_PART_NAME kPairPart10
#endif

#define LXV_PQ(B, W) walk_pair_queue_kernel<LXV_PF, B, W>
#define LXV_P3(B) {B, 1, walk_pair_kernel<LXV_PF, B, 1>, LXV_PQ(B, 1)}, {B, 2, walk_pair_kernel<LXV_PF, B, 2>, LXV_PQ(B, 2)}, \
                  {B, 4, walk_pair_kernel<LXV_PF, B, 4>, LXV_PQ(B, 4)}
static const PairEntry kWalksHere[] = {LXV_P3(16), LXV_P3(32), LXV_P3(64), LXV_P3(128), LXV_P3(256)};
#undef LXV_P3

extern const PairPart LXV_PART_NAME;
const PairPart LXV_PART_NAME = {
    LXV_PL,
    kWalksHere,
    (int)(sizeof kWalksHere / sizeof kWalksHere[0]),
    {init_pair_from_u_kernel<LXV_PF, 16>, init_pair_from_u_kernel<LXV_PF, 32>, init_pair_from_u_kernel<LXV_PF, 64>,
     init_pair_from_u_kernel<LXV_PF, 128>, init_pair_from_u_kernel<LXV_PF, 256>},
    init_pair_ladder_kernel<LXV_PF>,
};

#if LXV_PAIR_PART == 8
extern const PairPart kPairPart10;

static const PairPart* pair_part(int layout) {
    if (layout == LXV_LAYOUT_8X32) return &kPairPart8;
    if (layout == LXV_LAYOUT_10X26) return &kPairPart10;
    return nullptr;
}

static const PairEntry* pair_entry(int layout, int batch, int waves) {
    const PairPart* part = pair_part(layout);
    if (!part) return nullptr;
    for (int i = 0; i < part->n_walks; i++)
        if (part->walks[i].batch == batch && part->walks[i].waves == waves) return &part->walks[i];
    return nullptr;
}

static pair_fn pair_walk_fn(int layout, int batch, int waves) {
    const PairEntry* e = pair_entry(layout, batch, waves);
    return e ? e->fn : nullptr;
}

int walk_pair_queue_waves_per_cu(int layout, int batch, int waves) {
    const PairEntry* e = pair_entry(layout, batch, waves);
    if (!e || !e->queue_fn) return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, e->queue_fn, 64, 0) != hipSuccess) return 0;
    return blocks;
}

hipError_t launch_walk_pair_queue(int layout, int batch, int waves, const WalkParams& p, hipStream_t s) {
    const PairEntry* e = pair_entry(layout, batch, waves);
    if (!e || !e->queue_fn || p.resident % 64 || p.lanes % 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(e->queue_fn, dim3(p.resident / 64), dim3(64), 0, s, p);
    return hipGetLastError();
}

bool walk_pair_supported(int layout, int batch, int waves) { return pair_walk_fn(layout, batch, waves) != nullptr; }

hipError_t walk_pair_attributes(int layout, int batch, int waves, hipFuncAttributes* out) {
    const pair_fn fn = pair_walk_fn(layout, batch, waves);
    if (!fn) return hipErrorInvalidValue;
    return hipFuncGetAttributes(out, reinterpret_cast<const void*>(fn));
}

hipError_t launch_walk_pair(int layout, int batch, int waves, const WalkParams& p, hipStream_t s) {
    const pair_fn fn = pair_walk_fn(layout, batch, waves);
    if (!fn) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fn, dim3((p.lanes + 63) / 64), dim3(64), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_init_pair_from_u(int layout, int batch, uint32_t* state, uint64_t* ns, uint32_t* xor_acc,
                                   const uint32_t* upc, uint32_t lanes, hipStream_t s) {
    const PairPart* part = pair_part(layout);
    int slot = -1;
    for (int b = 16, k = 0; b <= 256; b *= 2, k++)
        if (b == batch) slot = k;
    if (!part || slot < 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(part->init_from_u[slot], dim3((lanes + 63) / 64), dim3(64), 0, s, state, ns, xor_acc, upc,
                       lanes);
    return hipGetLastError();
}

uint32_t bit_length(uint64_t v) {
    uint32_t b = 0;
    while (v) {
        b++;
        v >>= 1;
    }
    return b;
}

hipError_t launch_pair_anchors(const uint32_t* scalars, uint32_t words_apart, uint32_t n, PairAnchor* out,
                               hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(pair_anchor_kernel, dim3((n + 63) / 64), dim3(64), 0, s, scalars, words_apart, n, out);
    return hipGetLastError();
}

hipError_t launch_init_pair_ladder(int layout, uint32_t* state, uint64_t* ns, uint32_t* xor_acc,
                                   const uint32_t base[8], uint64_t stride, uint32_t lanes, uint32_t* scratch,
                                   hipStream_t s) {
    const PairPart* part = pair_part(layout);
    if (!part || !lanes) return hipErrorInvalidValue;
    // the two anchors' scalars (base, stride) in scratch[0..15], the anchors
    // after them
    uint32_t m[16] = {0};
    for (int i = 0; i < 8; i++) m[i] = base[i];
    m[8] = (uint32_t)stride;
    m[9] = (uint32_t)(stride >> 32);
    hipError_t e = hipMemcpy(scratch, m, sizeof m, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    PairLadderArgs a;
    a.anchors = reinterpret_cast<const PairAnchor*>(scratch + 16);
    a.bits = bit_length(lanes - 1);
    a.lanes = lanes;
    e = launch_pair_anchors(scratch, 8, 2, reinterpret_cast<PairAnchor*>(scratch + 16), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(part->init_ladder, dim3((lanes + 63) / 64), dim3(64), 0, s, state, ns, xor_acc, a);
    return hipGetLastError();
}
#endif  // LXV_PAIR_PART == 8

}  // namespace lxv
