// HIP kernels for gfx950 and the host-side launchers the runtime calls.
//
//   walk           the per-lane walk (src/gpu/common.h:280-365 / kernel.cu:26-36),
//                  one instantiation per (limb layout, keys per inversion,
//                  waves per SIMD the registers are budgeted for)
//   init_from_u    lane state from host-computed u(k-8), u(k)
//                  (init_walk, src/gpu/kernel.cu:21-24)
//   init_ladder    lane state from an on-device X25519 ladder
//   clear          zero the per-launch counters (no hipMemsetAsync, so the
//                  launch sequence can be captured in a graph)
//   walk_sw        the sub-walk form: K affine walks per lane, inverted across
//                  (lxv_subwalk.h), with its two initialisers
//   debug kernels  field operations and the address hash on given inputs, for
//                  parity tests against the CPU
//   valu_rate      VALU issue-rate probes (single instruction classes, the
//                  SHA-256 mix, the pair walk's mix) with the in-kernel clock
//   copy           a plain copy of known size, to calibrate HBM byte counters
//
// The batched search's kernels are in lxv_multi_kernels.hip, the pair walk's
// in lxv_pair_kernels.hip.
#include <cstring>

#include "fe10.h"
#include "fe8.h"
#include "lxv_kernels.h"
#include "lxv_walk.h"
#include "lxv_subwalk.h"

namespace lxv {

template <class F, int B, int W>
__global__ __launch_bounds__(64, W) void walk_kernel(const WalkParams p) {
    const uint32_t lane = blockIdx.x * 64u + threadIdx.x;
    if (lane < p.lanes) walk_lane<F, B>(p, lane);
}

template <class F>
__global__ __launch_bounds__(64) void init_from_u_kernel(uint32_t* state, uint64_t* ns, uint32_t* xor_acc,
                                                         const uint32_t* upc, uint32_t lanes) {
    const uint32_t lane = blockIdx.x * 64u + threadIdx.x;
    if (lane < lanes) init_lane_from_u<F>(state, ns, xor_acc, upc, lanes, lane);
}

struct LadderArgs {
    uint32_t base[8];  // start / 8 - 1
    uint32_t q[8];
    uint64_t stride;
    uint32_t lanes;
};

template <class F>
__global__ __launch_bounds__(64) void init_ladder_kernel(uint32_t* state, uint64_t* ns, uint32_t* xor_acc,
                                                         const LadderArgs a) {
    const uint32_t lane = blockIdx.x * 64u + threadIdx.x;
    if (lane < a.lanes) init_lane_ladder<F>(state, ns, xor_acc, a.base, a.stride, a.q, a.lanes, lane);
}

template <class F, int K, int W>
__global__ __launch_bounds__(64, W) void walk_sw_kernel(const WalkParams p) {
    const uint32_t lane = blockIdx.x * 64u + threadIdx.x;
    if (lane < p.lanes) walk_sw_lane<F, K>(p, lane);
}

template <class F>
__global__ __launch_bounds__(64) void init_sw_from_u_kernel(uint2* st, uint64_t* ns, uint32_t* xor_acc,
                                                            const uint32_t* upc, uint32_t lanes, uint32_t K) {
    const uint32_t t = blockIdx.x * 64u + threadIdx.x;
    if (t < lanes * K) init_sw_from_u<F>(st, ns, xor_acc, upc, lanes, K, t);
}

template <class F>
__global__ __launch_bounds__(64) void init_sw_ladder_kernel(uint2* st, uint64_t* ns, uint32_t* xor_acc,
                                                            const LadderArgs a, uint32_t K) {
    const uint32_t t = blockIdx.x * 64u + threadIdx.x;
    if (t < a.lanes * K) init_sw_ladder<F>(st, ns, xor_acc, a.base, a.stride, a.q, a.lanes, K, t);
}

__global__ void clear_kernel(uint32_t* words, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) words[i] = 0;
}

template <class F>
__global__ void debug_fe_kernel(int op, const uint32_t* a, const uint32_t* b, uint32_t* out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int w = debug_fe_stride<F>(op);
    uint32_t wa[F::kLimbs > 8 ? F::kLimbs : 8], wb[F::kLimbs > 8 ? F::kLimbs : 8], wo[F::kLimbs > 8 ? F::kLimbs : 8];
    for (int k = 0; k < w; k++) {
        wa[k] = a[(size_t)i * w + k];
        wb[k] = b[(size_t)i * w + k];
    }
    debug_fe_op<F>(op, wa, wb, wo);
    for (int k = 0; k < w; k++) out[(size_t)i * w + k] = wo[k];
}

__global__ void debug_addr_kernel(const ShaSearch c, const uint32_t* u, uint32_t* out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t w[8], a[4];
    for (int k = 0; k < 8; k++) w[k] = u[i * 8 + k];
    lxmf_address_words(w, c, a);
    for (int k = 0; k < 4; k++) out[i * 4 + k] = a[k];
}

// Issue-rate probes of single VALU instructions and of the SHA-256 mix, at
// a chosen occupancy: 8 independent chains per lane (so a wave never waits on
// its own results), each loop iteration kValuOpsPerIter[op] instructions,
// exactly as written (register operands only). Thread 0 of every block
// records the shader clock (s_memtime) and the constant-rate clock
// (s_memrealtime) at its start and end, so the host can tell the clock the
// SIMDs actually ran at from the kernel itself.
enum ValuProbeOp {
    kVpAlignbit = 0, kVpBitop3 = 1, kVpAdd3 = 2, kVpAdd = 3, kVpXor = 4, kVpPerm = 5, kVpShift = 6,
    kVpMad64 = 7, kVpShaMix = 8, kVpAddCo = 9, kVpCarryChain = 10, kVpWalkMix = 11, kVpMulLo = 12,
    kVpLshlAdd = 13, kVpLshlOr = 14, kVpMad64Chain = 15, kVpOps = 16
};
constexpr int kValuOpsPerIter[kVpOps] = {32, 32, 32, 32, 32, 32, 32, 32, 56, 32, 36, 200, 32, 32, 32, 32};

#define LXV_R8(I) I(0) I(1) I(2) I(3) I(4) I(5) I(6) I(7)
#define LXV_R32(I) LXV_R8(I) LXV_R8(I) LXV_R8(I) LXV_R8(I)

template <int kOp>
__global__ __launch_bounds__(256) void valu_rate_kernel(uint32_t* out, unsigned long long* clocks, uint32_t seed,
                                                        uint32_t iters) {
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long c0 = 0, r0 = 0;
    if (threadIdx.x == 0) {
        c0 = clock64();
        r0 = wall_clock64();
    }
    uint32_t x0 = seed ^ lane, x1 = x0 * 3u, x2 = x0 * 5u, x3 = x0 * 7u, x4 = x0 * 9u, x5 = x0 * 11u,
             x6 = x0 * 13u, x7 = x0 * 15u;
    uint64_t w0 = x0, w1 = x1, w2 = x2, w3 = x3, w4 = x4, w5 = x5, w6 = x6, w7 = x7;
    const uint32_t y = seed + 0x9e3779b9u, z = seed * 0x85ebca6bu;
    for (uint32_t it = 0; it < iters; it++) {
        if constexpr (kOp == kVpAlignbit) {
#define LXV_I(k) "v_alignbit_b32 %" #k ", %" #k ", %8, 7\n"
            asm volatile(LXV_R32(LXV_I) : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4), "+v"(x5), "+v"(x6),
                         "+v"(x7) : "v"(y), "v"(z));
#undef LXV_I
        } else if constexpr (kOp == kVpBitop3) {
#define LXV_I(k) "v_bitop3_b32 %" #k ", %" #k ", %8, %9 bitop3:0x96\n"
            asm volatile(LXV_R32(LXV_I) : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4), "+v"(x5), "+v"(x6),
                         "+v"(x7) : "v"(y), "v"(z));
#undef LXV_I
        } else if constexpr (kOp == kVpAdd3) {
#define LXV_I(k) "v_add3_u32 %" #k ", %" #k ", %8, %9\n"
            asm volatile(LXV_R32(LXV_I) : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4), "+v"(x5), "+v"(x6),
                         "+v"(x7) : "v"(y), "v"(z));
#undef LXV_I
        } else if constexpr (kOp == kVpAdd) {
#define LXV_I(k) "v_add_u32_e32 %" #k ", %8, %" #k "\n"
            asm volatile(LXV_R32(LXV_I) : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4), "+v"(x5), "+v"(x6),
                         "+v"(x7) : "v"(y), "v"(z));
#undef LXV_I
        } else if constexpr (kOp == kVpXor) {
#define LXV_I(k) "v_xor_b32_e32 %" #k ", %8, %" #k "\n"
            asm volatile(LXV_R32(LXV_I) : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4), "+v"(x5), "+v"(x6),
                         "+v"(x7) : "v"(y), "v"(z));
#undef LXV_I
        } else if constexpr (kOp == kVpPerm) {
#define LXV_I(k) "v_perm_b32 %" #k ", %" #k ", %8, %9\n"
            asm volatile(LXV_R32(LXV_I) : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4), "+v"(x5), "+v"(x6),
                         "+v"(x7) : "v"(y), "v"(z));
#undef LXV_I
        } else if constexpr (kOp == kVpShift) {
#define LXV_I(k) "v_lshrrev_b32_e32 %" #k ", %8, %" #k "\n"
            asm volatile(LXV_R32(LXV_I) : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4), "+v"(x5), "+v"(x6),
                         "+v"(x7) : "v"(y), "v"(z));
#undef LXV_I
        } else if constexpr (kOp == kVpMad64) {
            uint64_t c0_, c1_, c2_, c3_, c4_, c5_, c6_, c7_;
#define LXV_I(k) "v_mad_u64_u32 %" #k ", %[c" #k "], %16, %17, %" #k "\n"
            asm volatile(LXV_R32(LXV_I)
                         : "+v"(w0), "+v"(w1), "+v"(w2), "+v"(w3), "+v"(w4), "+v"(w5), "+v"(w6), "+v"(w7),
                           [c0] "=&s"(c0_), [c1] "=&s"(c1_), [c2] "=&s"(c2_), [c3] "=&s"(c3_), [c4] "=&s"(c4_),
                           [c5] "=&s"(c5_), [c6] "=&s"(c6_), [c7] "=&s"(c7_)
                         : "v"(y), "v"(z));
#undef LXV_I
        } else if constexpr (kOp == kVpMulLo) {
#define LXV_I(k) "v_mul_lo_u32 %" #k ", %" #k ", %8\n"
            asm volatile(LXV_R32(LXV_I) : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4), "+v"(x5), "+v"(x6),
                         "+v"(x7) : "v"(y), "v"(z));
#undef LXV_I
        } else if constexpr (kOp == kVpLshlAdd) {
#define LXV_I(k) "v_lshl_add_u32 %" #k ", %" #k ", 4, %8\n"
            asm volatile(LXV_R32(LXV_I) : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4), "+v"(x5), "+v"(x6),
                         "+v"(x7) : "v"(y), "v"(z));
#undef LXV_I
        } else if constexpr (kOp == kVpLshlOr) {
#define LXV_I(k) "v_lshl_or_b32 %" #k ", %" #k ", 6, %8\n"
            asm volatile(LXV_R32(LXV_I) : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4), "+v"(x5), "+v"(x6),
                         "+v"(x7) : "v"(y), "v"(z));
#undef LXV_I
        } else if constexpr (kOp == kVpMad64Chain) {
            // one dependent chain per lane: each v_mad_u64_u32 adds into the
            // previous one's result (a column of a multiplication)
            uint64_t cj_;
#define LXV_I(k) "v_mad_u64_u32 %0, %1, %2, %3, %0\n"
            asm volatile(LXV_R32(LXV_I) : "+v"(w0), "=&s"(cj_) : "v"(y), "v"(z));
#undef LXV_I
        } else if constexpr (kOp == kVpShaMix) {
            // per chain: 3 rotates, 2 three-input boolean functions, 2
            // three-input adds (the proportions of a SHA-256 round's
            // Sigma + Ch/Maj + sums)
#define LXV_I(k) "v_alignbit_b32 %" #k ", %" #k ", %" #k ", 6\n"
#define LXV_J(k) "v_alignbit_b32 %" #k ", %" #k ", %8, 11\n"
#define LXV_K(k) "v_bitop3_b32 %" #k ", %" #k ", %8, %9 bitop3:0x96\n"
#define LXV_L(k) "v_add3_u32 %" #k ", %" #k ", %8, %9\n"
#define LXV_M(k) "v_bitop3_b32 %" #k ", %" #k ", %9, %8 bitop3:0xca\n"
            asm volatile(LXV_R8(LXV_I) LXV_R8(LXV_J) LXV_R8(LXV_K) LXV_R8(LXV_L) LXV_R8(LXV_I) LXV_R8(LXV_M)
                             LXV_R8(LXV_L)
                         : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4), "+v"(x5), "+v"(x6), "+v"(x7)
                         : "v"(y), "v"(z));
#undef LXV_I
#undef LXV_J
#undef LXV_K
#undef LXV_L
#undef LXV_M
        } else if constexpr (kOp == kVpAddCo) {
#define LXV_I(k) "v_add_co_u32_e32 %" #k ", vcc, %8, %" #k "\n"
            asm volatile(LXV_R32(LXV_I) : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4), "+v"(x5), "+v"(x6),
                         "+v"(x7) : "v"(y), "v"(z) : "vcc");
#undef LXV_I
        } else if constexpr (kOp == kVpWalkMix) {
            // the per-key instruction classes of the pair walk's loop (its
            // gfx950 assembly at B = 64, 4 waves per SIMD: 30% v_alignbit, 19%
            // v_bitop3, 12% v_add3, 9% v_addc_co_u32_e64, 8% v_mad_u64_u32, 8%
            // v_add_u32, 6% v_mov, 3% v_lshrrev), 25 per chain per iteration
            uint64_t cout_;
#define LXV_PA(k) "v_alignbit_b32 %" #k ", %" #k ", %" #k ", 6\n"
#define LXV_PB(k) "v_bitop3_b32 %" #k ", %" #k ", %17, %18 bitop3:0x96\n"
#define LXV_PC(k) "v_add3_u32 %" #k ", %" #k ", %17, %18\n"
#define LXV_PD(k) "v_addc_co_u32_e64 %" #k ", %16, %" #k ", %17, %19\n"
#define LXV_PE(k) "v_mad_u64_u32 %[w" #k "], %16, %17, %18, %[w" #k "]\n"
#define LXV_PF(k) "v_add_u32_e32 %" #k ", %17, %" #k "\n"
#define LXV_PG(k) "v_mov_b32_e32 %" #k ", %18\n"
#define LXV_PH(k) "v_lshrrev_b32_e32 %" #k ", %17, %" #k "\n"
            asm volatile(LXV_R8(LXV_PA) LXV_R8(LXV_PA) LXV_R8(LXV_PA) LXV_R8(LXV_PB) LXV_R8(LXV_PB) LXV_R8(LXV_PC)
                         LXV_R8(LXV_PA) LXV_R8(LXV_PA) LXV_R8(LXV_PA) LXV_R8(LXV_PB) LXV_R8(LXV_PC) LXV_R8(LXV_PD)
                         LXV_R8(LXV_PE) LXV_R8(LXV_PF) LXV_R8(LXV_PG) LXV_R8(LXV_PA) LXV_R8(LXV_PA) LXV_R8(LXV_PB)
                         LXV_R8(LXV_PB) LXV_R8(LXV_PC) LXV_R8(LXV_PD) LXV_R8(LXV_PE) LXV_R8(LXV_PF) LXV_R8(LXV_PG)
                         LXV_R8(LXV_PH)
                         : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4), "+v"(x5), "+v"(x6), "+v"(x7),
                           [w0] "+v"(w0), [w1] "+v"(w1), [w2] "+v"(w2), [w3] "+v"(w3), [w4] "+v"(w4), [w5] "+v"(w5),
                           [w6] "+v"(w6), [w7] "+v"(w7), "=&s"(cout_)
                         : "v"(y), "v"(z), "s"((uint64_t)0));
#undef LXV_PA
#undef LXV_PB
#undef LXV_PC
#undef LXV_PD
#undef LXV_PE
#undef LXV_PF
#undef LXV_PG
#undef LXV_PH
        } else {
            // one dependent carry chain through VCC (the shape of a field
            // addition): each instruction reads the carry the previous wrote
#define LXV_I(k) "v_addc_co_u32_e32 %" #k ", vcc, %8, %" #k ", vcc\n"
#define LXV_C9 "v_add_co_u32_e32 %0, vcc, %8, %0\n" LXV_R8(LXV_I)
            asm volatile(LXV_C9 LXV_C9 LXV_C9 LXV_C9
                         : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4), "+v"(x5), "+v"(x6), "+v"(x7)
                         : "v"(y), "v"(z) : "vcc");
#undef LXV_C9
#undef LXV_I
        }
    }
    out[lane] = x0 ^ x1 ^ x2 ^ x3 ^ x4 ^ x5 ^ x6 ^ x7 ^ (uint32_t)(w0 ^ w1 ^ w2 ^ w3 ^ w4 ^ w5 ^ w6 ^ (w7 >> 3));
    if (threadIdx.x == 0) {
        unsigned long long* c = clocks + 4ull * blockIdx.x;
        c[0] = c0;
        c[1] = clock64();
        c[2] = r0;
        c[3] = wall_clock64();
    }
}
#undef LXV_R32
#undef LXV_R8

// A plain copy of known size, for scaling HBM byte counters.
__global__ __launch_bounds__(256) void copy_kernel(const uint4* src, uint4* dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Launchers

typedef void (*walk_fn)(const WalkParams);

#define LXV_WALK_ENTRY(F, LAYOUT, B, W) {LAYOUT, B, W, walk_kernel<F, B, W>}

struct WalkEntry {
    int layout;
    int batch;
    int waves;
    walk_fn fn;
};

static const WalkEntry kWalks[] = {
#define LXV_W3(F, LAYOUT, B) LXV_WALK_ENTRY(F, LAYOUT, B, 1), LXV_WALK_ENTRY(F, LAYOUT, B, 2), LXV_WALK_ENTRY(F, LAYOUT, B, 4)
    LXV_W3(Field10, LXV_LAYOUT_10X26, 16),  LXV_W3(Field10, LXV_LAYOUT_10X26, 32),
    LXV_W3(Field10, LXV_LAYOUT_10X26, 64),  LXV_W3(Field10, LXV_LAYOUT_10X26, 128),
    LXV_W3(Field10, LXV_LAYOUT_10X26, 256), LXV_W3(Field8, LXV_LAYOUT_8X32, 16),
    LXV_W3(Field8, LXV_LAYOUT_8X32, 32),    LXV_W3(Field8, LXV_LAYOUT_8X32, 64),
    LXV_W3(Field8, LXV_LAYOUT_8X32, 128),   LXV_W3(Field8, LXV_LAYOUT_8X32, 256),
#undef LXV_W3
};

bool walk_supported(int layout, int batch, int waves) {
    for (const auto& e : kWalks)
        if (e.layout == layout && e.batch == batch && e.waves == waves) return true;
    return false;
}

int limbs_of(int layout) { return layout == LXV_LAYOUT_8X32 ? Field8::kLimbs : Field10::kLimbs; }

hipError_t walk_attributes(int layout, int batch, int waves, hipFuncAttributes* out) {
    for (const auto& e : kWalks)
        if (e.layout == layout && e.batch == batch && e.waves == waves)
            return hipFuncGetAttributes(out, reinterpret_cast<const void*>(e.fn));
    return hipErrorInvalidValue;
}

hipError_t launch_walk(int layout, int batch, int waves, const WalkParams& p, hipStream_t s) {
    for (const auto& e : kWalks) {
        if (e.layout == layout && e.batch == batch && e.waves == waves) {
            const dim3 grid((p.lanes + 63) / 64), block(64);
            hipLaunchKernelGGL(e.fn, grid, block, 0, s, p);
            return hipGetLastError();
        }
    }
    return hipErrorInvalidValue;
}

hipError_t launch_init_from_u(int layout, uint32_t* state, uint64_t* ns, uint32_t* xor_acc, const uint32_t* upc,
                              uint32_t lanes, hipStream_t s) {
    const dim3 grid((lanes + 63) / 64), block(64);
    if (layout == LXV_LAYOUT_8X32)
        hipLaunchKernelGGL(init_from_u_kernel<Field8>, grid, block, 0, s, state, ns, xor_acc, upc, lanes);
    else
        hipLaunchKernelGGL(init_from_u_kernel<Field10>, grid, block, 0, s, state, ns, xor_acc, upc, lanes);
    return hipGetLastError();
}

hipError_t launch_init_ladder(int layout, uint32_t* state, uint64_t* ns, uint32_t* xor_acc, const uint32_t base[8],
                              const uint32_t q[8], uint64_t stride, uint32_t lanes, hipStream_t s) {
    LadderArgs a;
    for (int i = 0; i < 8; i++) {
        a.base[i] = base[i];
        a.q[i] = q[i];
    }
    a.stride = stride;
    a.lanes = lanes;
    const dim3 grid((lanes + 63) / 64), block(64);
    if (layout == LXV_LAYOUT_8X32)
        hipLaunchKernelGGL(init_ladder_kernel<Field8>, grid, block, 0, s, state, ns, xor_acc, a);
    else
        hipLaunchKernelGGL(init_ladder_kernel<Field10>, grid, block, 0, s, state, ns, xor_acc, a);
    return hipGetLastError();
}

// The sub-walk form: 8 x 32-bit limbs, K sub-walks per lane.
struct SwEntry {
    int layout;
    int subwalks;
    int waves;
    walk_fn fn;
};
static const SwEntry kSwWalks[] = {
    {LXV_LAYOUT_8X32, 64, 2, walk_sw_kernel<Field8, 64, 2>},   {LXV_LAYOUT_8X32, 64, 4, walk_sw_kernel<Field8, 64, 4>},
    {LXV_LAYOUT_8X32, 128, 2, walk_sw_kernel<Field8, 128, 2>}, {LXV_LAYOUT_8X32, 128, 4, walk_sw_kernel<Field8, 128, 4>},
    {LXV_LAYOUT_8X32, 256, 2, walk_sw_kernel<Field8, 256, 2>}, {LXV_LAYOUT_8X32, 256, 4, walk_sw_kernel<Field8, 256, 4>},
};

bool walk_sw_supported(int layout, int subwalks, int waves) {
    for (const auto& e : kSwWalks)
        if (e.layout == layout && e.subwalks == subwalks && e.waves == waves) return true;
    return false;
}

hipError_t launch_walk_sw(int layout, int subwalks, int waves, const WalkParams& p, hipStream_t s) {
    for (const auto& e : kSwWalks) {
        if (e.layout == layout && e.subwalks == subwalks && e.waves == waves) {
            hipLaunchKernelGGL(e.fn, dim3((p.lanes + 63) / 64), dim3(64), 0, s, p);
            return hipGetLastError();
        }
    }
    return hipErrorInvalidValue;
}

hipError_t walk_sw_attributes(int layout, int subwalks, int waves, hipFuncAttributes* out) {
    for (const auto& e : kSwWalks)
        if (e.layout == layout && e.subwalks == subwalks && e.waves == waves)
            return hipFuncGetAttributes(out, reinterpret_cast<const void*>(e.fn));
    return hipErrorInvalidValue;
}

hipError_t launch_init_sw_from_u(int layout, uint32_t* state, uint64_t* ns, uint32_t* xor_acc, const uint32_t* upc,
                                 uint32_t lanes, uint32_t subwalks, hipStream_t s) {
    if (layout != LXV_LAYOUT_8X32) return hipErrorInvalidValue;
    const uint64_t total = (uint64_t)lanes * subwalks;
    hipLaunchKernelGGL(init_sw_from_u_kernel<Field8>, dim3((uint32_t)((total + 63) / 64)), dim3(64), 0, s,
                       reinterpret_cast<uint2*>(state), ns, xor_acc, upc, lanes, subwalks);
    return hipGetLastError();
}

hipError_t launch_init_sw_ladder(int layout, uint32_t* state, uint64_t* ns, uint32_t* xor_acc, const uint32_t base[8],
                                 const uint32_t q[8], uint64_t stride, uint32_t lanes, uint32_t subwalks,
                                 hipStream_t s) {
    if (layout != LXV_LAYOUT_8X32) return hipErrorInvalidValue;
    LadderArgs a;
    for (int i = 0; i < 8; i++) {
        a.base[i] = base[i];
        a.q[i] = q[i];
    }
    a.stride = stride;
    a.lanes = lanes;
    const uint64_t total = (uint64_t)lanes * subwalks;
    hipLaunchKernelGGL(init_sw_ladder_kernel<Field8>, dim3((uint32_t)((total + 63) / 64)), dim3(64), 0, s,
                       reinterpret_cast<uint2*>(state), ns, xor_acc, a, subwalks);
    return hipGetLastError();
}

hipError_t launch_clear(uint32_t* words, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(clear_kernel, dim3((n + 255) / 256), dim3(256), 0, s, words, n);
    return hipGetLastError();
}

hipError_t launch_debug_fe(int layout, int op, const uint32_t* a, const uint32_t* b, uint32_t* out, uint32_t n,
                           hipStream_t s) {
    const dim3 grid((n + 63) / 64), block(64);
    if (layout == LXV_LAYOUT_8X32)
        hipLaunchKernelGGL(debug_fe_kernel<Field8>, grid, block, 0, s, op, a, b, out, n);
    else
        hipLaunchKernelGGL(debug_fe_kernel<Field10>, grid, block, 0, s, op, a, b, out, n);
    return hipGetLastError();
}

int valu_probe_ops_per_iter(int op) { return op >= 0 && op < kVpOps ? kValuOpsPerIter[op] : 0; }

hipError_t launch_valu_rate(int op, uint32_t* out, unsigned long long* clocks, uint32_t blocks, uint32_t seed,
                            uint32_t iters, hipStream_t s) {
    typedef void (*probe_fn)(uint32_t*, unsigned long long*, uint32_t, uint32_t);
    static const probe_fn fns[kVpOps] = {
        valu_rate_kernel<0>, valu_rate_kernel<1>, valu_rate_kernel<2>, valu_rate_kernel<3>,
        valu_rate_kernel<4>, valu_rate_kernel<5>, valu_rate_kernel<6>, valu_rate_kernel<7>,
        valu_rate_kernel<8>, valu_rate_kernel<9>, valu_rate_kernel<10>, valu_rate_kernel<11>,
        valu_rate_kernel<12>, valu_rate_kernel<13>, valu_rate_kernel<14>, valu_rate_kernel<15>};
    if (op < 0 || op >= kVpOps) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fns[op], dim3(blocks), dim3(256), 0, s, out, clocks, seed, iters);
    return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, bytes / 16);
    return hipGetLastError();
}

hipError_t launch_debug_addr(const ShaSearch& c, const uint32_t* u, uint32_t* out, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(debug_addr_kernel, dim3((n + 63) / 64), dim3(64), 0, s, c, u, out, n);
    return hipGetLastError();
}

}  // namespace lxv
