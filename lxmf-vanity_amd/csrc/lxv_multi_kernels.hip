// HIP kernels of the batched search (lxv_multi.h) and their launchers:
//
//   multi_pair    the pair walk over work units from the device queue, one
//                 instantiation per (limb layout, keys per inversion B, waves
//                 per SIMD the registers are budgeted for)
//   multi_walk    the same units walked by the reference's projective walk
//   multi_init    the admitted searches' virtual lanes set up by the ladder
//                 (projective) or from the search's anchor (pair walk)
//
// The grid of a batched launch is the resident waves only (persistent
// kernel), so the host sizes it from the occupancy of the instantiation it
// launches (multi_waves_per_cu).
#include <cstring>

#include "fe10.h"
#include "fe8.h"
#include "lxv_kernels.h"
#include "lxv_multi.h"

namespace lxv {

template <class F, int B, int W>
__global__ __launch_bounds__(64, W) void multi_walk_kernel(const MultiParams p) {
    const uint32_t plane = blockIdx.x * 64u + threadIdx.x;
    if (plane < p.phys_lanes) multi_wave<F, B>(p, plane);
}

template <class F, int B, int W>
__global__ __launch_bounds__(64, W) void multi_pair_kernel(const MultiParams p) {
    const uint32_t plane = blockIdx.x * 64u + threadIdx.x;
    if (plane < p.phys_lanes) multi_pair_wave<F, B>(p, plane);
}

struct MultiInitArgs {
    uint32_t q[8];
    uint64_t stride;
    uint32_t pool_lanes;
    uint32_t bits;  // pair walk: every entry's vlanes - 1 < 2^bits
    const MultiInitEntry* entries;
    PairAnchor s_anchor;  // pair walk: stride Q
};

// Block (x, y): lanes 64 x .. 64 x + 63 of admitted search y.
template <class F>
__global__ __launch_bounds__(64) void multi_init_kernel(uint32_t* pool, const MultiInitArgs a) {
    const MultiInitEntry& e = a.entries[blockIdx.y];
    const uint32_t v = blockIdx.x * 64u + threadIdx.x;
    if (v < e.vlanes) multi_init_lane<F>(pool, a.pool_lanes, e.state_off, e.base, a.stride, a.q, v);
}

template <class F>
__global__ __launch_bounds__(64) void multi_pair_init_kernel(uint32_t* pool, const MultiInitArgs a) {
    const MultiInitEntry& e = a.entries[blockIdx.y];
    const uint32_t v = blockIdx.x * 64u + threadIdx.x;
    if (v < e.vlanes) multi_pair_init_lane<F>(pool, a.pool_lanes, e.state_off, e.p0, a.s_anchor, a.bits, v);
}

typedef void (*multi_fn)(const MultiParams);
struct MultiEntry {
    int layout;
    int batch;
    int walk;
    int waves;  // per SIMD, as compiled
    multi_fn fn;
};
static const MultiEntry kMultis[] = {
    {LXV_LAYOUT_8X32, 64, LXV_WALK_PROJECTIVE, 4, multi_walk_kernel<Field8, 64, 4>},
    {LXV_LAYOUT_8X32, 256, LXV_WALK_PROJECTIVE, 4, multi_walk_kernel<Field8, 256, 4>},
    {LXV_LAYOUT_10X26, 256, LXV_WALK_PROJECTIVE, 4, multi_walk_kernel<Field10, 256, 4>},
    {LXV_LAYOUT_8X32, 64, LXV_WALK_PAIRS, 2, multi_pair_kernel<Field8, 64, 2>},
    {LXV_LAYOUT_8X32, 64, LXV_WALK_PAIRS, 4, multi_pair_kernel<Field8, 64, 4>},
    {LXV_LAYOUT_8X32, 128, LXV_WALK_PAIRS, 2, multi_pair_kernel<Field8, 128, 2>},
    {LXV_LAYOUT_8X32, 128, LXV_WALK_PAIRS, 4, multi_pair_kernel<Field8, 128, 4>},
    {LXV_LAYOUT_8X32, 256, LXV_WALK_PAIRS, 2, multi_pair_kernel<Field8, 256, 2>},
    {LXV_LAYOUT_8X32, 256, LXV_WALK_PAIRS, 4, multi_pair_kernel<Field8, 256, 4>},
    {LXV_LAYOUT_10X26, 256, LXV_WALK_PAIRS, 2, multi_pair_kernel<Field10, 256, 2>},
    {LXV_LAYOUT_10X26, 256, LXV_WALK_PAIRS, 4, multi_pair_kernel<Field10, 256, 4>},
};

static const MultiEntry* multi_entry(int layout, int batch, int walk, int waves) {
    for (const auto& e : kMultis)
        if (e.layout == layout && e.batch == batch && e.walk == walk && e.waves == waves) return &e;
    return nullptr;
}

bool multi_supported(int layout, int batch, int walk, int waves) {
    return multi_entry(layout, batch, walk, waves) != nullptr;
}

int multi_waves_per_cu(int layout, int batch, int walk, int waves) {
    const MultiEntry* e = multi_entry(layout, batch, walk, waves);
    if (!e) return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, e->fn, 64, 0) != hipSuccess) return 0;
    return blocks;
}

hipError_t launch_multi(int layout, int batch, int walk, int waves, const MultiParams& p, hipStream_t s) {
    const MultiEntry* e = multi_entry(layout, batch, walk, waves);
    if (!e || p.phys_lanes % 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(e->fn, dim3(p.phys_lanes / 64), dim3(64), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_multi_init(int layout, int walk, uint32_t* pool, uint32_t pool_lanes, const uint32_t q[8],
                             uint64_t stride, const MultiInitEntry* entries, uint32_t n, uint32_t max_vlanes,
                             const PairAnchor* s_anchor, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (n > 65535 || max_vlanes == 0) return hipErrorInvalidValue;
    MultiInitArgs a;
    for (int i = 0; i < 8; i++) a.q[i] = q[i];
    a.stride = stride;
    a.pool_lanes = pool_lanes;
    a.bits = bit_length(max_vlanes - 1);
    a.entries = entries;
    std::memset(&a.s_anchor, 0, sizeof a.s_anchor);
    const dim3 grid((max_vlanes + 63) / 64, n), block(64);
    if (walk == LXV_WALK_PAIRS) {
        if (!s_anchor) return hipErrorInvalidValue;
        a.s_anchor = *s_anchor;
        if (layout == LXV_LAYOUT_8X32)
            hipLaunchKernelGGL(multi_pair_init_kernel<Field8>, grid, block, 0, s, pool, a);
        else
            hipLaunchKernelGGL(multi_pair_init_kernel<Field10>, grid, block, 0, s, pool, a);
    } else if (layout == LXV_LAYOUT_8X32) {
        hipLaunchKernelGGL(multi_init_kernel<Field8>, grid, block, 0, s, pool, a);
    } else {
        hipLaunchKernelGGL(multi_init_kernel<Field10>, grid, block, 0, s, pool, a);
    }
    return hipGetLastError();
}

}  // namespace lxv
