// Launchers of the device kernels (defined in lxv_kernels.hip) and the
// field-operation dispatcher shared by the debug kernels and the host checks.
#pragma once

#include <hip/hip_runtime.h>

#include "lxv_sha.h"

namespace lxv {

struct WalkParams;

// The waves per SIMD a kernel is compiled for, from the resident waves per
// CU asked for (4 SIMDs per CU): 16 and up -> 4, 8 -> 2, fewer -> 1.
// (5 and 6 per SIMD, with 96 and 80 registers and more spilling, measured
// +0.3% and -0.2% keys/s for the pair walk at B = 64 and less at B = 256:
// not built.)
inline int waves_per_simd(uint32_t waves_per_cu) {
    return waves_per_cu >= 16 ? 4 : waves_per_cu >= 8 ? 2 : 1;
}

enum DebugFeOp {
    kFeMul = 0, kFeSq = 1, kFeAdd = 2, kFeSub = 3, kFeInvert = 4, kFeCanon = 5, kFeMulSmall = 6, kFeInvertChain = 7
};
// Or-ed into an op code: the inputs are taken as the layout's raw limbs
// (F::kLimbs words per element) and the result is returned as the limbs the
// operation produced (congruent to the result, not reduced further) instead
// of its canonical encoding; kFeCanon still returns the canonical encoding of
// its raw input. For 8 x 32 the limbs are any value below 2^256, bit 255
// included; for 10 x 25.5 they must keep to each operation's input bounds
// (fe10.h).
constexpr int kFeRawLimbs = 0x100;

// Words per element in the debug buffers: 8 for encodings, F::kLimbs for raw
// limbs.
template <class F>
LXV_HD constexpr int debug_fe_stride(int op) {
    return (op & kFeRawLimbs) ? F::kLimbs : 8;
}

// out = canonical words of op(a, b); a and b are decoded as 8 little-endian
// words (bit 255 dropped). Sums and differences are multiplied by one before
// encoding, as the walk only encodes multiplication results. With
// kFeRawLimbs: raw limbs in and out, see above (a, b and out then hold
// debug_fe_stride words; a canonical result fills the first 8, the rest 0).
template <class F>
LXV_HD void debug_fe_op(int op, const uint32_t* a, const uint32_t* b, uint32_t* out) {
    typedef typename F::fe fe;
    const bool raw = (op & kFeRawLimbs) != 0;
    op &= 0xff;
    const fe x = raw ? F::load(a) : F::from_words(a), y = raw ? F::load(b) : F::from_words(b);
    fe r;
    switch (op) {
        case kFeMul: r = F::mul(x, y); break;
        case kFeSq: r = F::sq(x); break;
        case kFeAdd: r = raw ? F::add(x, y) : F::mul(F::add(x, y), F::one()); break;
        case kFeSub: r = raw ? F::sub(x, y) : F::mul(F::sub(x, y), F::one()); break;
        case kFeInvert: r = F::invert(x); break;
        case kFeInvertChain: r = F::invert_chain(x); break;
        case kFeMulSmall: r = F::mul_small(x, 121665); break;
        default: r = raw ? x : F::mul(x, F::one()); break;
    }
    if (raw && op != kFeCanon) {
        F::store(r, out);
        return;
    }
    F::to_words(r, out);
    for (int i = 8; i < (raw ? F::kLimbs : 8); i++) out[i] = 0;
}

bool walk_supported(int layout, int batch, int waves);
int limbs_of(int layout);
hipError_t walk_attributes(int layout, int batch, int waves, hipFuncAttributes* out);
hipError_t launch_walk(int layout, int batch, int waves, const WalkParams& p, hipStream_t s);
hipError_t launch_init_from_u(int layout, uint32_t* state, uint64_t* ns, uint32_t* xor_acc, const uint32_t* upc,
                              uint32_t lanes, hipStream_t s);
hipError_t launch_init_ladder(int layout, uint32_t* state, uint64_t* ns, uint32_t* xor_acc, const uint32_t base[8],
                              const uint32_t q[8], uint64_t stride, uint32_t lanes, hipStream_t s);
hipError_t launch_clear(uint32_t* words, uint32_t n, hipStream_t s);

// The sub-walk form (lxv_subwalk.h): `lanes` physical lanes of `subwalks`
// affine walks each; the state holds lanes x subwalks x 64 bytes.
bool walk_sw_supported(int layout, int subwalks, int waves);
hipError_t launch_walk_sw(int layout, int subwalks, int waves, const WalkParams& p, hipStream_t s);
hipError_t walk_sw_attributes(int layout, int subwalks, int waves, hipFuncAttributes* out);
hipError_t launch_init_sw_from_u(int layout, uint32_t* state, uint64_t* ns, uint32_t* xor_acc, const uint32_t* upc,
                                 uint32_t lanes, uint32_t subwalks, hipStream_t s);
hipError_t launch_init_sw_ladder(int layout, uint32_t* state, uint64_t* ns, uint32_t* xor_acc, const uint32_t base[8],
                                 const uint32_t q[8], uint64_t stride, uint32_t lanes, uint32_t subwalks,
                                 hipStream_t s);

// The pair walk (lxv_pairwalk.h, lxv_pair_kernels.hip): one walk per lane,
// B keys per inversion around an affine centre; the state holds the centre
// (x, y) in slots 0 and 1 of the projective walk's 4-element lane state.
bool walk_pair_supported(int layout, int batch, int waves);
hipError_t walk_pair_attributes(int layout, int batch, int waves, hipFuncAttributes* out);
hipError_t launch_walk_pair(int layout, int batch, int waves, const WalkParams& p, hipStream_t s);
// The persistent form: p.resident / 64 waves take groups of 64 of the p.lanes
// lanes from the counter p.queue (zero at launch); the batch arrays are sized
// for p.resident lanes.
// Waves of 64 of the persistent kernel one CU holds at once (0: not built).
int walk_pair_queue_waves_per_cu(int layout, int batch, int waves);
hipError_t launch_walk_pair_queue(int layout, int batch, int waves, const WalkParams& p, hipStream_t s);
hipError_t launch_init_pair_from_u(int layout, int batch, uint32_t* state, uint64_t* ns, uint32_t* xor_acc,
                                   const uint32_t* upc, uint32_t lanes, hipStream_t s);
// base = k_0 / 8 + (B/2 - 1) as 8 little-endian words; lane t starts at
// k_0 + 8 t stride. scratch: kPairInitScratchWords device words for the two
// anchors and their scalars.
constexpr uint32_t kPairInitScratchWords = 48;
hipError_t launch_init_pair_ladder(int layout, uint32_t* state, uint64_t* ns, uint32_t* xor_acc,
                                   const uint32_t base[8], uint64_t stride, uint32_t lanes, uint32_t* scratch,
                                   hipStream_t s);
// Lane i < n: out[i] = m_i Q for the scalars m_i (8 little-endian words at
// scalars + i * words_apart, device memory).
struct PairAnchor;
hipError_t launch_pair_anchors(const uint32_t* scalars, uint32_t words_apart, uint32_t n, PairAnchor* out,
                               hipStream_t s);
uint32_t bit_length(uint64_t v);

struct MultiParams;
struct MultiInitEntry;
// The batched kernels (lxv_multi_kernels.hip). walk: LXV_WALK_PROJECTIVE or
// LXV_WALK_PAIRS (pool state of 4 or 2 field elements per virtual lane);
// waves: per SIMD the instantiation's registers are budgeted for (2 or 4;
// the projective form is built for 4 only). multi_waves_per_cu: waves of 64
// one CU holds at once (0: not built); the grid of p.phys_lanes / 64 waves
// should not exceed CUs x that.
bool multi_supported(int layout, int batch, int walk, int waves);
int multi_waves_per_cu(int layout, int batch, int walk, int waves);
hipError_t launch_multi(int layout, int batch, int walk, int waves, const MultiParams& p, hipStream_t s);
// Sets up the virtual lanes of n admitted searches (entries in device memory)
// in one launch; max_vlanes bounds every entry's lane count. The pair walk
// takes each entry's anchor p0 and s_anchor = stride Q (host memory, passed
// by value).
hipError_t launch_multi_init(int layout, int walk, uint32_t* pool, uint32_t pool_lanes, const uint32_t q[8],
                             uint64_t stride, const MultiInitEntry* entries, uint32_t n, uint32_t max_vlanes,
                             const PairAnchor* s_anchor, hipStream_t s);
hipError_t launch_debug_fe(int layout, int op, const uint32_t* a, const uint32_t* b, uint32_t* out, uint32_t n,
                           hipStream_t s);
// VALU issue-rate probe `op` (LXV_VALU_* of lxmf_vanity_hip_debug.h) over
// `blocks` blocks of 256 lanes: out[lane] (blocks x 256 words) and, per block,
// 4 clock readings (shader clock at start and end, constant-rate clock at
// start and end) in clocks[4 block ..]. valu_probe_ops_per_iter: the probe's
// VALU instructions per lane per iteration (0: no such probe).
int valu_probe_ops_per_iter(int op);
hipError_t launch_valu_rate(int op, uint32_t* out, unsigned long long* clocks, uint32_t blocks, uint32_t seed,
                            uint32_t iters, hipStream_t s);
// Copies bytes (a multiple of 16) from src to dst.
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
hipError_t launch_debug_addr(const ShaSearch& c, const uint32_t* u, uint32_t* out, uint32_t n, hipStream_t s);

}  // namespace lxv
