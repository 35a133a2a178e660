// The C ABI of include/lxmf_vanity_hip.h: device context, persistent walk
// state, asynchronous launch and poll, stop, checksums, and the host helpers
// of the search driver. Host code only; the kernels are in lxv_kernels.hip.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lxmf_vanity_hip.h"
#include "../../include/lxmf_vanity_hip_debug.h"
#include "fe10.h"
#include "fe8.h"
#include "host_crypto.h"
#include "lxv_kernels.h"
#include "lxv_pairwalk.h"
#include "lxv_walk.h"
#include "tuned_shapes.h"

using namespace lxv;

namespace {

thread_local std::string g_open_error = "no error";

constexpr uint32_t kLaneRounds = 8;    // default lanes = resident lanes x this (hardware dispatch)
constexpr uint32_t kQueueRounds = 64;  // the same for the persistent pair walk

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// Scratch of lxv_launch_noalloc: a file-scope device array, so that entry
// point allocates nothing. Words [0, 64) hold the launch counters (queue,
// stopped flag, batches walked), the rest the pair walk's batch arrays of
// the context's grid. Each GPU holds its own copy from the first use of this
// code object until the process ends. Host code never uses the array's name
// as an address: g_noalloc.ptr[device] is its device address, from
// hipGetSymbolAddress on that device.
constexpr size_t kNoAllocCounterWords = 64;
constexpr size_t kNoAllocBatchBytes = (size_t)32 << 20;
constexpr size_t kNoAllocWords = kNoAllocCounterWords + kNoAllocBatchBytes / sizeof(uint32_t);
constexpr int kNoAllocMaxDevices = 64;

struct NoAllocScratch {
    std::mutex mu;
    uint32_t* ptr[kNoAllocMaxDevices] = {};
    hipError_t err[kNoAllocMaxDevices] = {};
};
NoAllocScratch g_noalloc;

}  // namespace

static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

// Looks up (once per device, with that device current) the device address of
// `scratch` on `device`.
static hipError_t noalloc_scratch(int device, uint32_t** out) {
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_noalloc.mu);
    if (!g_noalloc.ptr[device]) {
        if (g_noalloc.err[device] != hipSuccess) return g_noalloc.err[device];
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            g_noalloc.err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
            return g_noalloc.err[device];
        }
        g_noalloc.ptr[device] = static_cast<uint32_t*>(p);
    }
    *out = g_noalloc.ptr[device];
    return hipSuccess;
}

struct lxv_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string desc;
    std::string err = "no error";
    uint32_t lanes = 0, batch = 64, waves = 4, layout = LXV_LAYOUT_10X26, limbs = 10;
    uint32_t subwalks = 0;  // 0: one walk per lane; K: K affine sub-walks per lane
    uint32_t walk = LXV_WALK_PAIRS;  // the form of the one-walk-per-lane kernels
    uint32_t phys = 0;      // physical lanes (= lanes unless sub-walks)
    uint32_t resident = 0;  // lanes of the persistent pair-walk grid; 0: one thread per lane
    uint32_t* d_state = nullptr;
    uint64_t* d_ns = nullptr;
    uint32_t* d_xor = nullptr;
    uint2* d_batch = nullptr;
    uint32_t* d_counters = nullptr;            // [0] hit count, [1] stopped, [2..3] batches (u64), [4] queue
    uint32_t* d_anchor = nullptr;              // the pair walk's set-up: two scalars and their anchors
    lxv_hit* d_hits = nullptr;
    uint32_t hits_cap = 0;
    uint32_t* d_dump = nullptr;
    size_t dump_cap = 0;                        // words
    uint32_t* h_stop = nullptr;                 // host-mapped stop flag
    uint32_t* d_stop = nullptr;
    bool initialized = false;
    bool in_flight = false;
    lxv_launch_args last{};
    lxv_launch_stats stats{};
};

static int fail(lxv_ctx* c, int code, const std::string& msg) {
    if (c)
        c->err = msg;
    else
        g_open_error = msg;
    return code;
}

static int hip_fail(lxv_ctx* c, hipError_t e, const char* what) {
    return fail(c, LXV_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(c, expr)                                   \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return hip_fail(c, e_, #expr); \
    } while (0)

static void free_ctx(lxv_ctx* c) {
    if (!c) return;
    if (c->device >= 0) (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    (void)hipFree(c->d_state);
    (void)hipFree(c->d_ns);
    (void)hipFree(c->d_xor);
    (void)hipFree(c->d_batch);
    (void)hipFree(c->d_counters);
    (void)hipFree(c->d_anchor);
    (void)hipFree(c->d_hits);
    (void)hipFree(c->d_dump);
    if (c->h_stop) (void)hipHostFree(c->h_stop);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int lxv_open(const lxv_open_opts* opts, lxv_ctx** out) {
    if (!out) return fail(nullptr, LXV_E_ARG, "lxv_open: out is NULL");
    *out = nullptr;
    lxv_open_opts o{};
    if (opts) {
        // Callers built against an older header pass a shorter struct: the
        // fields it lacks (those after `layout`) keep their zero defaults.
        if (opts->size < offsetof(lxv_open_opts, subwalks))
            return fail(nullptr, LXV_E_ARG, "lxv_open: opts->size too small");
        std::memcpy(&o, opts, opts->size < sizeof o ? opts->size : sizeof o);
    }
    auto* c = new (std::nothrow) lxv_ctx();
    if (!c) return fail(nullptr, LXV_E_NOMEM, "lxv_open: out of host memory");
    c->device = o.device < 0 ? 0 : o.device;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) {
        delete c;
        return fail(nullptr, LXV_E_HIP, "no HIP device found (is this an MI355X host?)");
    }
    if (c->device >= ndev) {
        delete c;
        return fail(nullptr, LXV_E_ARG, "lxv_open: no such device");
    }
    hipDeviceProp_t prop;
    if ((e = hipSetDevice(c->device)) != hipSuccess || (e = hipGetDeviceProperties(&prop, c->device)) != hipSuccess) {
        delete c;
        return fail(nullptr, LXV_E_HIP, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
    }
    {
        // the no-allocation launch's scratch on this device (device current);
        // a failure here is reported by lxv_launch_noalloc, not by open
        uint32_t* unused = nullptr;
        (void)noalloc_scratch(c->device, &unused);
    }
    c->layout = o.layout;  // LXV_LAYOUT_8X32 (0) is the default: the faster layout measured on gfx950
    if (c->layout != LXV_LAYOUT_10X26 && c->layout != LXV_LAYOUT_8X32) {
        delete c;
        return fail(nullptr, LXV_E_ARG, "lxv_open: unknown layout");
    }
    const lxv_tuned_shape& tuned = kTunedShapes[LXV_TUNED_SINGLE];
    uint32_t wpc = o.waves_per_cu ? o.waves_per_cu : tuned.waves_per_cu;
    c->waves = (uint32_t)waves_per_simd(wpc);
    c->subwalks = o.subwalks;
    if (c->subwalks) {
        // K sub-walks per lane: a thread is a sub-walk, one key per iteration
        if (o.batch > 1 || !walk_sw_supported((int)c->layout, (int)c->subwalks, (int)c->waves)) {
            delete c;
            return fail(nullptr, LXV_E_ARG,
                        "lxv_open: sub-walks need the 8x32 layout, subwalks 64, 128 or 256, batch 0 or 1");
        }
        c->batch = 1;
    } else {
        c->batch = o.batch ? o.batch : tuned.batch;
        c->walk = o.walk;
        if (c->walk != LXV_WALK_PAIRS && c->walk != LXV_WALK_PROJECTIVE) {
            delete c;
            return fail(nullptr, LXV_E_ARG, "lxv_open: unknown walk form");
        }
    }
    if (!c->subwalks && !(c->walk == LXV_WALK_PAIRS ? walk_pair_supported((int)c->layout, (int)c->batch, (int)c->waves)
                                                    : walk_supported((int)c->layout, (int)c->batch, (int)c->waves))) {
        delete c;
        return fail(nullptr, LXV_E_ARG,
                    "lxv_open: no walk kernel built for this layout / batch / waves (batch must be 16..256, a power of two)");
    }
    c->limbs = (uint32_t)limbs_of((int)c->layout);
    // Default lanes: many times what is resident at once (one walk per lane
    // everywhere but the sub-walk form). Waves do not all run at the same
    // speed, so a launch sized to exactly one resident round ends with the
    // slowest wave. The pair walk launches a persistent grid of the resident
    // waves that take 64-lane groups from a counter, 64 groups per wave
    // (batch arrays for the resident lanes only); the projective walk queues
    // eight rounds of waves for the hardware dispatcher, whose batch arrays
    // cover every lane. Measured on MI355X at B=64: eight hardware rounds
    // +4.5% keys/s over one; the persistent grid over 64 rounds +1.1% more.
    const bool pairs = !c->subwalks && c->walk == LXV_WALK_PAIRS;
    const int queue_wpc = pairs ? walk_pair_queue_waves_per_cu((int)c->layout, (int)c->batch, (int)c->waves) : 0;
    const uint32_t rounds = c->subwalks ? 1u : (queue_wpc > 0 ? kQueueRounds : kLaneRounds);
    c->phys = o.threads ? o.threads : (uint32_t)prop.multiProcessorCount * wpc * 64u * rounds;
    c->phys = (c->phys + 63) / 64 * 64;
    if (queue_wpc > 0) {
        c->resident = o.resident_lanes ? (o.resident_lanes + 63) / 64 * 64
                                       : (uint32_t)prop.multiProcessorCount * (uint32_t)queue_wpc * 64u;
        if (c->resident >= c->phys) c->resident = 0;  // one thread per lane: nothing to balance
    }
    const uint64_t threads = (uint64_t)c->phys * (c->subwalks ? c->subwalks : 1);
    if (threads > 0xffffffffull) {
        delete c;
        return fail(nullptr, LXV_E_ARG, "lxv_open: lanes x subwalks exceeds 2^32");
    }
    c->lanes = (uint32_t)threads;
    char buf[256];
    snprintf(buf, sizeof buf, "%s%s%s, %d CUs (HIP)", prop.name, prop.name[0] ? " " : "AMD GPU ", prop.gcnArchName, prop.multiProcessorCount);
    c->desc = buf;

#define ALLOC(ptr, bytes)                                               \
    do {                                                                \
        hipError_t ea = hipMalloc((void**)&(ptr), (bytes));              \
        if (ea != hipSuccess) {                                         \
            std::string m = std::string("hipMalloc ") + #ptr + ": " + hipGetErrorString(ea); \
            free_ctx(c);                                                \
            return fail(nullptr, LXV_E_HIP, m);                         \
        }                                                               \
    } while (0)
    const size_t L = c->phys;
    if (c->subwalks) {
        // two slots of 8 canonical words per sub-walk; a_j and D_j per sub-walk
        ALLOC(c->d_state, L * c->subwalks * 2 * 8 * sizeof(uint32_t));
        ALLOC(c->d_batch, L * (size_t)c->subwalks * 2 * c->limbs * sizeof(uint32_t));
    } else if (c->walk == LXV_WALK_PAIRS) {
        // the centre (x, y) in slots 0-1 of the lane state; B/2 prefix products
        // per lane of the grid
        ALLOC(c->d_state, L * 2 * c->limbs * sizeof(uint32_t));
        ALLOC(c->d_batch, (c->resident ? c->resident : L) * (size_t)(c->batch / 2) * c->limbs * sizeof(uint32_t));
    } else {
        ALLOC(c->d_state, L * 4 * c->limbs * sizeof(uint32_t));
        ALLOC(c->d_batch, L * (size_t)c->batch * 2 * c->limbs * sizeof(uint32_t));
    }
    ALLOC(c->d_ns, L * sizeof(uint64_t));
    ALLOC(c->d_xor, L * 4 * sizeof(uint32_t));
    ALLOC(c->d_counters, 16 * sizeof(uint32_t));
    ALLOC(c->d_anchor, kPairInitScratchWords * sizeof(uint32_t));
#undef ALLOC
    if ((e = hipHostMalloc((void**)&c->h_stop, 64, hipHostMallocMapped | hipHostMallocCoherent)) != hipSuccess) {
        std::string m = std::string("hipHostMalloc stop flag: ") + hipGetErrorString(e);
        free_ctx(c);
        return fail(nullptr, LXV_E_HIP, m);
    }
    *(volatile uint32_t*)c->h_stop = 0;
    if ((e = hipHostGetDevicePointer((void**)&c->d_stop, c->h_stop, 0)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess) {
        std::string m = std::string("stream/event setup: ") + hipGetErrorString(e);
        free_ctx(c);
        return fail(nullptr, LXV_E_HIP, m);
    }
    *out = c;
    return LXV_OK;
}

extern "C" const char* lxv_describe(lxv_ctx* c) { return c ? c->desc.c_str() : ""; }
extern "C" uint32_t lxv_threads(lxv_ctx* c) { return c ? c->lanes : 0; }
extern "C" uint64_t lxv_batch(lxv_ctx* c) { return c ? c->batch : 0; }

static int check_idle(lxv_ctx* c, const char* what) {
    if (!c) return LXV_E_ARG;
    if (c->in_flight) return fail(c, LXV_E_BUSY, std::string(what) + ": a launch has not been polled");
    return LXV_OK;
}

extern "C" int lxv_init(lxv_ctx* c, const uint32_t* upc, size_t n_words) {
    int rc = check_idle(c, "lxv_init");
    if (rc) return rc;
    if (!upc || n_words != (size_t)c->lanes * 16) return fail(c, LXV_E_ARG, "lxv_init: need 16 words per lane");
    HIPCHK(c, hipSetDevice(c->device));
    uint32_t* d_upc = nullptr;
    HIPCHK(c, hipMalloc((void**)&d_upc, n_words * sizeof(uint32_t)));
    hipError_t e = hipMemcpyAsync(d_upc, upc, n_words * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        if (c->subwalks)
            e = launch_init_sw_from_u((int)c->layout, c->d_state, c->d_ns, c->d_xor, d_upc, c->phys, c->subwalks,
                                      c->stream);
        else if (c->walk == LXV_WALK_PAIRS)
            e = launch_init_pair_from_u((int)c->layout, (int)c->batch, c->d_state, c->d_ns, c->d_xor, d_upc,
                                        c->lanes, c->stream);
        else
            e = launch_init_from_u((int)c->layout, c->d_state, c->d_ns, c->d_xor, d_upc, c->lanes, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_upc);
    if (e != hipSuccess) return hip_fail(c, e, "lxv_init");
    c->initialized = true;
    return LXV_OK;
}

// k / 8 + d as 8 little-endian words, for a multiple of 8 with bit 254 set
// (start scalars) and |d| < 2^31, so nothing wraps.
static void scalar_div8_plus(const uint8_t k[32], int64_t d, uint32_t out[8]) {
    uint8_t s8[32];
    uint32_t carry = 0;
    for (int i = 31; i >= 0; i--) {
        s8[i] = (uint8_t)((k[i] >> 3) | (carry << 5));
        carry = k[i] & 7;
    }
    uint64_t acc = (uint64_t)(int64_t)d;  // two's complement, sign-extended through every word
    const uint64_t ext = d < 0 ? 0xffffffffull : 0;
    uint64_t c = 0;
    for (int i = 0; i < 8; i++) {
        const uint64_t w = (uint64_t)s8[4 * i] | (uint64_t)s8[4 * i + 1] << 8 | (uint64_t)s8[4 * i + 2] << 16 |
                           (uint64_t)s8[4 * i + 3] << 24;
        const uint64_t add = i == 0 ? (acc & 0xffffffffull) : (i == 1 ? (acc >> 32) : ext);
        const uint64_t v = w + add + c;
        out[i] = (uint32_t)v;
        c = v >> 32;
    }
}

extern "C" int lxv_init_seeded(lxv_ctx* c, const uint8_t start[32], uint64_t stride) {
    int rc = check_idle(c, "lxv_init_seeded");
    if (rc) return rc;
    if (!start || !start_scalar_ok(start)) return fail(c, LXV_E_ARG, "lxv_init_seeded: not a valid start scalar");
    if (stride == 0 || (c->lanes > 1 && stride > (UINT64_MAX >> 1) / c->lanes))
        return fail(c, LXV_E_ARG, "lxv_init_seeded: lanes x stride must stay below 2^63");
    // base = start / 8 - 1 (the ladder gives u(k - 8), u(k)), or for the pair
    // walk start / 8 + B/2 - 1 (the first batch's centre)
    uint32_t base[8];
    scalar_div8_plus(start, c->walk == LXV_WALK_PAIRS && !c->subwalks ? (int64_t)c->batch / 2 - 1 : -1, base);
    lxv_launch_args a{};
    uint8_t zero[32] = {0};
    lxv_launch_args_for(zero, &a);
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->subwalks && c->walk == LXV_WALK_PAIRS)
        HIPCHK(c, launch_init_pair_ladder((int)c->layout, c->d_state, c->d_ns, c->d_xor, base, stride, c->lanes,
                                          c->d_anchor, c->stream));
    else if (c->subwalks)
        HIPCHK(c, launch_init_sw_ladder((int)c->layout, c->d_state, c->d_ns, c->d_xor, base, a.q, stride, c->phys,
                                        c->subwalks, c->stream));
    else
        HIPCHK(c, launch_init_ladder((int)c->layout, c->d_state, c->d_ns, c->d_xor, base, a.q, stride, c->lanes,
                                     c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->initialized = true;
    return LXV_OK;
}

template <class F>
static void limbs_of_q(const uint32_t q[8], uint32_t qp[10], uint32_t qm[10], uint32_t qr[10]) {
    typedef typename F::fe fe;
    const fe qf = F::from_words(q);
    uint32_t a[F::kLimbs], b[F::kLimbs], r[F::kLimbs];
    F::store(F::add(qf, F::one()), a);
    F::store(F::sub(qf, F::one()), b);
    F::store(qf, r);
    for (int i = 0; i < 10; i++) {
        qp[i] = i < F::kLimbs ? a[i] : 0;
        qm[i] = i < F::kLimbs ? b[i] : 0;
        qr[i] = i < F::kLimbs ? r[i] : 0;
    }
}

static void fill_q(uint32_t layout, const uint32_t q[8], uint32_t qp[10], uint32_t qm[10], uint32_t qr[10]) {
    if (layout == LXV_LAYOUT_8X32)
        limbs_of_q<Field8>(q, qp, qm, qr);
    else
        limbs_of_q<Field10>(q, qp, qm, qr);
}

extern "C" int lxv_launch(lxv_ctx* c, const lxv_launch_args* a) {
    int rc = check_idle(c, "lxv_launch");
    if (rc) return rc;
    if (!a) return fail(c, LXV_E_ARG, "lxv_launch: args is NULL");
    if (!c->initialized) return fail(c, LXV_E_STATE, "lxv_launch: lanes not initialised (lxv_init first)");
    if (a->dump_threads > c->lanes) return fail(c, LXV_E_ARG, "lxv_launch: dump_threads exceeds lanes");
    HIPCHK(c, hipSetDevice(c->device));
    if (a->max_hits > c->hits_cap) {
        (void)hipFree(c->d_hits);
        c->d_hits = nullptr;
        c->hits_cap = 0;
        HIPCHK(c, hipMalloc((void**)&c->d_hits, (size_t)a->max_hits * sizeof(lxv_hit)));
        c->hits_cap = a->max_hits;
    }
    const size_t dump_words = (size_t)a->dump_threads * c->batch * 8;
    if (dump_words > c->dump_cap) {
        (void)hipFree(c->d_dump);
        c->d_dump = nullptr;
        c->dump_cap = 0;
        HIPCHK(c, hipMalloc((void**)&c->d_dump, dump_words * sizeof(uint32_t)));
        c->dump_cap = dump_words;
    }
    WalkParams p;
    std::memset(&p, 0, sizeof p);
    sha_search_init(p.sha, a->ed, a->name_hash, a->want, a->mask);
    fill_q(c->layout, a->q, p.qp, p.qm, p.q);
    p.state = c->d_state;
    p.ns = c->d_ns;
    p.xor_acc = c->d_xor;
    p.batch_mem = c->d_batch;
    p.hits = c->d_hits;
    p.hit_count = c->d_counters;
    p.stopped = c->d_counters + 1;
    p.batches = (unsigned long long*)(c->d_counters + 2);
    p.dump = c->d_dump;
    p.stop = c->d_stop;
    p.max_hits = a->max_hits;
    p.dump_threads = a->dump_threads;
    p.lanes = c->phys;
    p.iters = a->iters;
    p.queue = c->d_counters + 4;
    p.resident = c->resident;
    *(volatile uint32_t*)c->h_stop = 0;
    HIPCHK(c, launch_clear(c->d_counters, 16, c->stream));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (c->subwalks)
        HIPCHK(c, launch_walk_sw((int)c->layout, (int)c->subwalks, (int)c->waves, p, c->stream));
    else if (c->walk == LXV_WALK_PAIRS && c->resident)
        HIPCHK(c, launch_walk_pair_queue((int)c->layout, (int)c->batch, (int)c->waves, p, c->stream));
    else if (c->walk == LXV_WALK_PAIRS)
        HIPCHK(c, launch_walk_pair((int)c->layout, (int)c->batch, (int)c->waves, p, c->stream));
    else
        HIPCHK(c, launch_walk((int)c->layout, (int)c->batch, (int)c->waves, p, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->last = *a;
    c->in_flight = true;
    return LXV_OK;
}

extern "C" int lxv_poll(lxv_ctx* c, int wait, lxv_hit* hits, uint32_t cap, uint32_t* n_hits, uint32_t* dump,
                        size_t dump_words) {
    if (!c) return LXV_E_ARG;
    if (!c->in_flight) return fail(c, LXV_E_STATE, "lxv_poll: nothing launched");
    // checked before anything is collected, so a too-small buffer leaves the
    // launch in flight and the call can be repeated with a larger one
    const size_t dw = (size_t)c->last.dump_threads * c->batch * 8;
    if (dw && dump && dump_words < dw) return fail(c, LXV_E_ARG, "lxv_poll: dump buffer too small");
    HIPCHK(c, hipSetDevice(c->device));
    if (!wait) {
        hipError_t q = hipEventQuery(c->ev1);
        if (q == hipErrorNotReady) return LXV_E_BUSY;
        if (q != hipSuccess) {
            c->in_flight = false;
            return hip_fail(c, q, "lxv_poll: the walk kernel failed");
        }
    }
    hipError_t e = hipEventSynchronize(c->ev1);
    c->in_flight = false;
    if (e != hipSuccess) return hip_fail(c, e, "lxv_poll: the walk kernel failed");
    uint32_t counters[4];
    HIPCHK(c, hipMemcpy(counters, c->d_counters, sizeof counters, hipMemcpyDeviceToHost));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    uint64_t batches = (uint64_t)counters[2] | (uint64_t)counters[3] << 32;
    c->stats.kernel_ms = ms;
    c->stats.keys = batches * c->batch;
    c->stats.hits_total = counters[0];
    c->stats.stopped = counters[1];
    const uint32_t stored = counters[0] < c->last.max_hits ? counters[0] : c->last.max_hits;
    const uint32_t give = stored < cap ? stored : cap;
    if (n_hits) *n_hits = stored;
    if (give && hits) HIPCHK(c, hipMemcpy(hits, c->d_hits, (size_t)give * sizeof(lxv_hit), hipMemcpyDeviceToHost));
    if (dw && dump) HIPCHK(c, hipMemcpy(dump, c->d_dump, dw * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return LXV_OK;
}

extern "C" int lxv_launch_noalloc(lxv_ctx* c, const lxv_launch_args* a, lxv_hit* d_hits, uint32_t* d_hit_count,
                                  void* stream) {
    int rc = check_idle(c, "lxv_launch_noalloc");
    if (rc) return rc;
    if (!a || !d_hit_count || (a->max_hits && !d_hits))
        return fail(c, LXV_E_ARG, "lxv_launch_noalloc: args, d_hit_count and (for max_hits > 0) d_hits are needed");
    if (!c->initialized) return fail(c, LXV_E_STATE, "lxv_launch_noalloc: lanes not initialised (lxv_init first)");
    if (c->subwalks || c->walk != LXV_WALK_PAIRS)
        return fail(c, LXV_E_ARG, "lxv_launch_noalloc: only the pair walk (LXV_WALK_PAIRS, no sub-walks) is supported");
    if (a->dump_threads) return fail(c, LXV_E_ARG, "lxv_launch_noalloc: dump_threads must be 0 (no dump buffer)");
    const uint32_t grid = c->resident ? c->resident : c->phys;
    const size_t need = (size_t)grid * (c->batch / 2) * c->limbs * sizeof(uint32_t);
    if (need > kNoAllocBatchBytes) {
        char msg[256];
        snprintf(msg, sizeof msg,
                 "lxv_launch_noalloc: the batch arrays of this context's grid (%u lanes x %u keys per inversion) "
                 "need %zu bytes, more than the %zu of the static scratch; open the context with fewer "
                 "resident_lanes or a smaller batch",
                 grid, c->batch, need, kNoAllocBatchBytes);
        return fail(c, LXV_E_ARG, msg);
    }
    HIPCHK(c, hipSetDevice(c->device));
    uint32_t* scr = nullptr;
    HIPCHK(c, noalloc_scratch(c->device, &scr));
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
    WalkParams p;
    std::memset(&p, 0, sizeof p);
    sha_search_init(p.sha, a->ed, a->name_hash, a->want, a->mask);
    fill_q(c->layout, a->q, p.qp, p.qm, p.q);
    p.state = c->d_state;
    p.ns = c->d_ns;
    p.xor_acc = c->d_xor;
    p.batch_mem = reinterpret_cast<uint2*>(scr + kNoAllocCounterWords);
    p.hits = d_hits;
    p.hit_count = d_hit_count;
    p.stopped = scr + 1;
    p.batches = reinterpret_cast<unsigned long long*>(scr + 2);
    p.queue = scr + 4;
    p.dump = nullptr;
    p.stop = c->d_stop;
    p.max_hits = a->max_hits;
    p.dump_threads = 0;
    p.lanes = c->phys;
    p.iters = a->iters;
    p.resident = c->resident;
    *(volatile uint32_t*)c->h_stop = 0;
    // the scratch keeps the last call's contents: clear its counters, and
    // the caller's hit count, with a kernel (capturable, unlike a memset)
    HIPCHK(c, launch_clear(scr, (uint32_t)kNoAllocCounterWords, s));
    HIPCHK(c, launch_clear(d_hit_count, 1, s));
    if (c->resident)
        HIPCHK(c, launch_walk_pair_queue((int)c->layout, (int)c->batch, (int)c->waves, p, s));
    else
        HIPCHK(c, launch_walk_pair((int)c->layout, (int)c->batch, (int)c->waves, p, s));
    return LXV_OK;
}

extern "C" int lxv_stop(lxv_ctx* c) {
    if (!c) return LXV_E_ARG;
    __atomic_store_n(c->h_stop, 1u, __ATOMIC_SEQ_CST);
    return LXV_OK;
}

extern "C" int lxv_checksum(lxv_ctx* c, uint32_t xw[4]) {
    int rc = check_idle(c, "lxv_checksum");
    if (rc) return rc;
    if (!xw) return fail(c, LXV_E_ARG, "lxv_checksum: NULL output");
    std::vector<uint32_t> h((size_t)4 * c->phys);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpy(h.data(), c->d_xor, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (int k = 0; k < 4; k++) {
        uint32_t x = 0;
        for (size_t i = 0; i < c->phys; i++) x ^= h[(size_t)k * c->phys + i];
        xw[k] = x;
    }
    return LXV_OK;
}

extern "C" int lxv_read_ns(lxv_ctx* c, uint32_t first, uint32_t count, uint64_t* ns) {
    int rc = check_idle(c, "lxv_read_ns");
    if (rc) return rc;
    if (!ns || (uint64_t)first + count > c->lanes) return fail(c, LXV_E_ARG, "lxv_read_ns: range out of bounds");
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->subwalks) {
        HIPCHK(c, hipMemcpy(ns, c->d_ns + first, (size_t)count * sizeof(uint64_t), hipMemcpyDeviceToHost));
        return LXV_OK;
    }
    // every sub-walk of a lane is at the lane's n: thread t is lane t % lanes
    std::vector<uint64_t> h(c->phys);
    HIPCHK(c, hipMemcpy(h.data(), c->d_ns, h.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < count; i++) ns[i] = h[(first + i) % c->phys];
    return LXV_OK;
}

extern "C" int lxv_last_stats(lxv_ctx* c, lxv_launch_stats* out) {
    if (!c || !out) return LXV_E_ARG;
    *out = c->stats;
    return LXV_OK;
}

extern "C" void lxv_close(lxv_ctx* c) { free_ctx(c); }

extern "C" const char* lxv_last_error(lxv_ctx* c) { return c ? c->err.c_str() : g_open_error.c_str(); }

extern "C" int lxv_tuned(int workload, lxv_tuned_shape* out) {
    if (!out || workload < LXV_TUNED_SINGLE || workload > LXV_TUNED_SERVING) return LXV_E_ARG;
    *out = kTunedShapes[workload];
    return LXV_OK;
}

// ---------------------------------------------------------------------------
// Host helpers of the search driver.

extern "C" void lxv_x25519_base(const uint8_t k[32], uint8_t out[32]) { x25519_base(k, out); }
extern "C" void lxv_ed25519_public(const uint8_t seed[32], uint8_t out[32]) { ed25519_public(seed, out); }
extern "C" void lxv_address_of_identity(const uint8_t id[64], uint8_t out[16]) { address_of_identity(id, out); }
extern "C" void lxv_step_scalar(const uint8_t k[32], uint64_t n, uint8_t out[32]) { step_scalar(k, n, out); }
extern "C" int lxv_start_scalar_ok(const uint8_t k[32]) { return start_scalar_ok(k) ? 1 : 0; }

extern "C" int lxv_random_start_scalar(uint8_t out[32]) {
    for (;;) {
        if (!random_bytes(out, 32)) return LXV_E_ARG;
        clamp(out);
        if (start_scalar_ok(out)) return LXV_OK;
    }
}

extern "C" int lxv_prefix_words(const char* s, uint32_t want[4], uint32_t mask[4]) {
    Prefix p;
    if (!s || !Prefix::parse(s, &p, nullptr)) return LXV_E_ARG;
    p.words(want, mask);
    return p.nibbles;
}

static uint32_t be_word(const uint8_t* b) {
    return (uint32_t)b[0] << 24 | (uint32_t)b[1] << 16 | (uint32_t)b[2] << 8 | b[3];
}

extern "C" void lxv_launch_args_for(const uint8_t ed_pub[32], lxv_launch_args* a) {
    for (int i = 0; i < 8; i++) a->ed[i] = be_word(ed_pub + 4 * i);
    uint8_t nh[10];
    name_hash(nh);
    a->name_hash[0] = be_word(nh);
    a->name_hash[1] = be_word(nh + 4);
    a->name_hash[2] = (uint32_t)nh[8] << 24 | (uint32_t)nh[9] << 16;
    uint8_t eight[32] = {8}, base[32] = {9}, q[32];
    mont_mul(eight, base, q);
    for (int i = 0; i < 8; i++)
        a->q[i] = (uint32_t)q[4 * i] | (uint32_t)q[4 * i + 1] << 8 | (uint32_t)q[4 * i + 2] << 16 |
                  (uint32_t)q[4 * i + 3] << 24;
}

extern "C" void lxv_seeded_search(uint64_t seed, int nibbles, uint8_t start[32], uint8_t ed_seed[32], char* prefix) {
    if (start) seeded_start(seed, start);
    if (ed_seed) seeded_ed_seed(seed, ed_seed);
    if (prefix) {
        std::string p = seeded_prefix(seed, nibbles);
        std::memcpy(prefix, p.c_str(), p.size() + 1);
    }
}

// ---------------------------------------------------------------------------
// Debug entry points.

extern "C" int lxv_debug_fe(int layout, int op, const uint32_t* a, const uint32_t* b, uint32_t* out, uint32_t n,
                            int on_gpu) {
    if (!a || !b || !out || (op & 0xff) > 7 || (op & ~(0xff | LXV_FE_RAW_LIMBS)) != 0) return LXV_E_ARG;
    if (layout != LXV_LAYOUT_10X26 && layout != LXV_LAYOUT_8X32) return LXV_E_ARG;
    const size_t w = layout == LXV_LAYOUT_8X32 ? debug_fe_stride<Field8>(op) : debug_fe_stride<Field10>(op);
    if (!on_gpu) {
        for (uint32_t i = 0; i < n; i++) {
            if (layout == LXV_LAYOUT_8X32)
                debug_fe_op<Field8>(op, a + w * i, b + w * i, out + w * i);
            else
                debug_fe_op<Field10>(op, a + w * i, b + w * i, out + w * i);
        }
        return LXV_OK;
    }
    if (n == 0) return LXV_OK;
    uint32_t *da = nullptr, *db = nullptr, *dout = nullptr;
    const size_t bytes = (size_t)n * w * sizeof(uint32_t);
    hipError_t e = hipSetDevice(0);
    if (e == hipSuccess) e = hipMalloc((void**)&da, bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&db, bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&dout, bytes);
    if (e == hipSuccess) e = hipMemcpy(da, a, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, b, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_debug_fe(layout, op, da, db, dout, n, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dout);
    return e == hipSuccess ? LXV_OK : LXV_E_HIP;
}

extern "C" int lxv_debug_cpu_walk(const uint8_t start[32], uint32_t batches, uint8_t* out) {
    if (!start || !out || !start_scalar_ok(start)) return LXV_E_ARG;
    KeyWalk w(start);
    for (uint32_t b = 0; b < batches; b++)
        w.next_batch(reinterpret_cast<uint8_t(*)[32]>(out + (size_t)b * KeyWalk::kBatch * 32), KeyWalk::kBatch);
    return LXV_OK;
}

extern "C" int lxv_debug_address(const uint32_t ed[8], const uint32_t nh[3], const uint32_t* u, uint32_t* out,
                                 uint32_t n, int on_gpu) {
    if (!ed || !nh || !u || !out) return LXV_E_ARG;
    ShaSearch c;
    const uint32_t zero[4] = {0, 0, 0, 0};
    sha_search_init(c, ed, nh, zero, zero);
    if (!on_gpu) {
        for (uint32_t i = 0; i < n; i++) lxmf_address_words(u + 8 * i, c, out + 4 * i);
        return LXV_OK;
    }
    if (n == 0) return LXV_OK;
    uint32_t *du = nullptr, *dout = nullptr;
    hipError_t e = hipSetDevice(0);
    if (e == hipSuccess) e = hipMalloc((void**)&du, (size_t)n * 8 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&dout, (size_t)n * 4 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemcpy(du, u, (size_t)n * 8 * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_debug_addr(c, du, dout, n, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out, dout, (size_t)n * 4 * sizeof(uint32_t), hipMemcpyDeviceToHost);
    (void)hipFree(du);
    (void)hipFree(dout);
    return e == hipSuccess ? LXV_OK : LXV_E_HIP;
}

extern "C" int lxv_debug_valu_rate(int device, int op, uint32_t waves_per_cu, double* lane_ops_per_s,
                                   double* clock_hz, double* nominal_clock_hz, int* cus) {
    const int per_iter = valu_probe_ops_per_iter(op);
    if (!lane_ops_per_s || per_iter == 0 || waves_per_cu == 0 || waves_per_cu % 4 || waves_per_cu > 32)
        return LXV_E_ARG;
    hipDeviceProp_t prop;
    int wall_khz = 0;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
        hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess)
        return LXV_E_HIP;
    // one resident round at the asked occupancy: blocks of 4 waves
    const uint32_t blocks = (uint32_t)prop.multiProcessorCount * (waves_per_cu / 4);
    const uint32_t lanes = blocks * 256u, iters = 4096, runs = 5;
    uint32_t* out = nullptr;
    unsigned long long* clocks = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc((void**)&out, (size_t)lanes * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&clocks, (size_t)blocks * runs * 4 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = launch_valu_rate(op, out, clocks, blocks, 1, iters / 8, nullptr);  // warm-up
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    for (uint32_t r = 0; r < runs && e == hipSuccess; r++)
        e = launch_valu_rate(op, out, clocks + (size_t)r * blocks * 4, blocks, 2 + r, iters, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    std::vector<unsigned long long> h((size_t)blocks * runs * 4);
    if (e == hipSuccess) e = hipMemcpy(h.data(), clocks, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    (void)hipFree(clocks);
    if (e != hipSuccess || ms <= 0) return LXV_E_HIP;
    *lane_ops_per_s = (double)runs * lanes * (double)iters * per_iter / (ms * 1e-3);
    // the shader clock each block saw over its own run, against the
    // constant-rate clock, averaged over blocks
    double cyc = 0, real = 0;
    for (size_t b = 0; b < (size_t)blocks * runs; b++) {
        cyc += (double)(h[4 * b + 1] - h[4 * b]);
        real += (double)(h[4 * b + 3] - h[4 * b + 2]);
    }
    if (clock_hz) *clock_hz = real > 0 ? cyc / real * wall_khz * 1e3 : 0;
    if (nominal_clock_hz) *nominal_clock_hz = prop.clockRate * 1e3;
    if (cus) *cus = prop.multiProcessorCount;
    return LXV_OK;
}

extern "C" int lxv_debug_copy(int device, uint64_t bytes, double* ms) {
    if (bytes == 0 || bytes % 16) return LXV_E_ARG;
    void *src = nullptr, *dst = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipMalloc(&src, bytes);
    if (e == hipSuccess) e = hipMalloc(&dst, bytes);
    if (e == hipSuccess) e = hipMemset(src, 1, bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    if (e == hipSuccess) e = launch_copy(src, dst, bytes, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(src);
    (void)hipFree(dst);
    return e == hipSuccess ? LXV_OK : LXV_E_HIP;
}

extern "C" int lxv_debug_kernel_attrs(int layout, uint32_t batch, uint32_t waves_per_cu, uint32_t walk,
                                      uint32_t* vgprs, uint32_t* sgprs, uint32_t* scratch_bytes,
                                      uint32_t* waves_per_simd) {
    const int w = lxv::waves_per_simd(waves_per_cu);
    const bool pairs = walk == LXV_WALK_PAIRS;
    if (pairs ? !walk_pair_supported(layout, (int)batch, w) : !walk_supported(layout, (int)batch, w))
        return LXV_E_ARG;
    hipFuncAttributes a;
    if ((pairs ? walk_pair_attributes(layout, (int)batch, w, &a) : walk_attributes(layout, (int)batch, w, &a)) !=
        hipSuccess)
        return LXV_E_HIP;
    if (vgprs) *vgprs = (uint32_t)a.numRegs;
    if (sgprs) *sgprs = 0;
    if (scratch_bytes) *scratch_bytes = (uint32_t)a.localSizeBytes;
    if (waves_per_simd) *waves_per_simd = (uint32_t)w;
    return LXV_OK;
}

extern "C" int lxv_debug_walk_kernel(const lxv_ctx* c, char* name, size_t name_cap, uint32_t* grid_lanes) {
    if (!c) return LXV_E_ARG;
    const char* fn = c->subwalks ? "walk_sw_kernel"
                     : c->walk == LXV_WALK_PAIRS ? (c->resident ? "walk_pair_queue_kernel" : "walk_pair_kernel")
                                                  : "walk_kernel";
    if (name && name_cap)
        snprintf(name, name_cap, "%s<lxv::%s, %u, %u>", fn, c->layout == LXV_LAYOUT_10X26 ? "Field10" : "Field8",
                 c->subwalks ? c->subwalks : c->batch, c->waves);
    if (grid_lanes) *grid_lanes = c->resident ? c->resident : c->phys;
    return LXV_OK;
}

namespace lxv {

// The host build of the same walk, for CPU checks (lxv_debug_pair_walk).
template <class F, int B>
static void host_pair_walk(const uint32_t* upc, const uint32_t base[8], uint32_t batches, uint32_t* out) {
    typename F::fe xc, yc;
    if (upc)
        pair_centre_from_u<F, B>(upc, upc + 8, xc, yc);
    else
        pair_centre_ladder<F>(base, 0, xc, yc);
    uint2 mem[(B / 2) * (F::kLimbs / 2)];
    for (uint32_t b = 0; b < batches; b++) {
        uint32_t* o = out + (size_t)b * B * 8;
        pair_batch<F, B>(xc, yc, mem, 1, 0, [&](int i, const typename F::fe& x) { F::to_words(x, o + (size_t)i * 8); });
    }
}

template <class F>
static bool host_pair_walk_b(int batch, const uint32_t* upc, const uint32_t base[8], uint32_t batches,
                             uint32_t* out) {
    switch (batch) {
        case 4: host_pair_walk<F, 4>(upc, base, batches, out); return true;
        case 8: host_pair_walk<F, 8>(upc, base, batches, out); return true;
        case 16: host_pair_walk<F, 16>(upc, base, batches, out); return true;
        case 32: host_pair_walk<F, 32>(upc, base, batches, out); return true;
        case 64: host_pair_walk<F, 64>(upc, base, batches, out); return true;
        case 128: host_pair_walk<F, 128>(upc, base, batches, out); return true;
        case 256: host_pair_walk<F, 256>(upc, base, batches, out); return true;
        default: return false;
    }
}

static bool host_pair_walk_run(int layout, int batch, const uint32_t* upc, const uint32_t base[8], uint32_t batches,
                        uint32_t* out) {
    if (layout == LXV_LAYOUT_8X32) return host_pair_walk_b<Field8>(batch, upc, base, batches, out);
    if (layout == LXV_LAYOUT_10X26) return host_pair_walk_b<Field10>(batch, upc, base, batches, out);
    return false;
}

}  // namespace lxv

template <class F>
static void host_pair_centre(const uint32_t base[8], uint64_t stride, uint64_t t, int bits, uint32_t* out) {
    typename F::fe x, y;
    pair_centre_ladder<F>(base, t * stride, x, y);
    F::to_words(x, out);
    F::to_words(y, out + 8);
    PairAnchor a[2];
    const uint32_t sw[8] = {(uint32_t)stride, (uint32_t)(stride >> 32), 0, 0, 0, 0, 0, 0};
    pair_anchor<Field8>(base, a[0]);
    pair_anchor<Field8>(sw, a[1]);
    pair_centre_offset<F>(a[0], a[1], t, bits, x, y);
    F::to_words(x, out + 16);
    F::to_words(y, out + 24);
}

extern "C" int lxv_debug_pair_centre(int layout, uint32_t batch, const uint8_t start[32], uint64_t stride, uint64_t t,
                                     uint32_t bits, uint32_t* out) {
    if (!out || !start || !start_scalar_ok(start)) return LXV_E_ARG;
    if (batch < 4 || batch > 256 || (batch & (batch - 1)) || stride == 0 || bits > 63 || (t >> bits) != 0 ||
        (t && stride > (UINT64_MAX >> 1) / t))
        return LXV_E_ARG;
    uint32_t base[8];
    scalar_div8_plus(start, (int64_t)batch / 2 - 1, base);
    if (layout == LXV_LAYOUT_8X32)
        host_pair_centre<Field8>(base, stride, t, (int)bits, out);
    else if (layout == LXV_LAYOUT_10X26)
        host_pair_centre<Field10>(base, stride, t, (int)bits, out);
    else
        return LXV_E_ARG;
    return LXV_OK;
}

extern "C" int lxv_debug_pair_walk(int layout, uint32_t batch, const uint8_t start[32], const uint32_t* upc,
                                   uint32_t batches, uint32_t* out) {
    if (!out || (!upc && (!start || !start_scalar_ok(start)))) return LXV_E_ARG;
    if (batch < 4 || batch > 256 || (batch & (batch - 1))) return LXV_E_ARG;
    uint32_t base[8] = {0};
    if (!upc) scalar_div8_plus(start, (int64_t)batch / 2 - 1, base);
    return host_pair_walk_run(layout, (int)batch, upc, base, batches, out) ? LXV_OK : LXV_E_ARG;
}
