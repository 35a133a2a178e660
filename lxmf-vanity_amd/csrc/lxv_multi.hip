// Host side of the batched search (lxv_multi_* in include/lxmf_vanity_hip.h):
// admission into slots and the HBM state pool, per-step work planning, the
// hipGraph-captured relaunch of the persistent kernel, and answer selection.
#include <hip/hip_runtime.h>

#include <cstddef>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "../../include/lxmf_vanity_hip.h"
#include "../../include/lxmf_vanity_hip_debug.h"
#include "fe10.h"
#include "fe8.h"
#include "host_crypto.h"
#include "lxv_kernels.h"
#include "lxv_multi.h"
#include "tuned_shapes.h"

using namespace lxv;

namespace {

constexpr uint64_t kStride = 1ull << 40;   // key distance between virtual lanes
constexpr uint32_t kNone = 0xffffffffu;
constexpr size_t kMaxSuspendPerStep = 64;   // suspensions one admission may make

thread_local std::string g_err = "no error";

struct Search {
    uint64_t id = 0, tag = 0, max_keys = 0;
    uint32_t priority = 0;
    uint8_t start[32];
    uint8_t ed_pub[32];
    SlotConst sc;
    uint32_t vlanes = 0, slot = kNone, state_off = 0;
    // Planned progress: once every launched step has run, blocks
    // [0, next_block) are walked completely, and virtual waves
    // [0, wave_cursor) of block next_block as well. done_block / done_cursor:
    // the same for the steps collected so far; after_block / after_cursor:
    // the planned progress at the end of the step launched with parity p.
    uint32_t next_block = 0, wave_cursor = 0;
    uint32_t done_block = 0, done_cursor = 0;
    uint32_t after_block[2] = {0, 0}, after_cursor[2] = {0, 0};
    // The step being planned: virtual waves [w0, w1) of blocks [next_block, + nblocks).
    uint32_t w0 = 0, w1 = 0, nblocks = 0;
    PairAnchor p0;  // the pair walk: (start / 8 + B/2 - 1) Q, from the host at submission
    // Best hit so far in walk order (block, n, v); final once its block is whole.
    bool hit = false;
    uint32_t hit_block = 0, hit_vlane = 0;
    uint64_t hit_n = 0;
    uint64_t keys = 0;
    double submit_s = 0, admit_s = 0;
    uint32_t suspensions = 0;
    bool active = false;
};

// First-fit allocator of pool lanes (ranges are multiples of 64).
struct Pool {
    std::map<uint32_t, uint32_t> free;  // offset -> length
    void init(uint32_t n) {
        free.clear();
        free[0] = n;
    }
    bool take(uint32_t len, uint32_t* off) {
        for (auto it = free.begin(); it != free.end(); ++it) {
            if (it->second >= len) {
                *off = it->first;
                const uint32_t rest = it->second - len, noff = it->first + len;
                free.erase(it);
                if (rest) free[noff] = rest;
                return true;
            }
        }
        return false;
    }
    void give(uint32_t off, uint32_t len) {
        auto it = free.emplace(off, len).first;
        auto nx = std::next(it);
        if (nx != free.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free.erase(nx);
        }
        if (it != free.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free.erase(it);
            }
        }
    }
};

uint32_t vlanes_for(int nibbles, uint32_t cap, uint32_t lane_keys_log2) {
    // about 16^digits / 2^lane_keys_log2 lanes. A lane is set up by a ladder
    // over its index (at most 18 bits) for the pair walk, and by a full
    // scalar ladder for the projective walk (a few hundred keys' worth of
    // field work: about 1% of a lane's 2^14 expected keys).
    const int log2_lanes = 4 * nibbles - (int)lane_keys_log2;
    uint64_t v = log2_lanes <= 6 ? 64 : (log2_lanes >= 18 ? (1u << 18) : (1ull << log2_lanes));
    if (v > cap) v = cap;
    if (v < 64) v = 64;
    return (uint32_t)v;
}

uint32_t be_word(const uint8_t* b) {
    return (uint32_t)b[0] << 24 | (uint32_t)b[1] << 16 | (uint32_t)b[2] << 8 | b[3];
}

}  // namespace

// What a launched step walked: (search id, slot) of every search with units
// in it, so its hits and progress are credited to the searches that owned
// the slots when it was planned.
struct StepRec {
    std::vector<std::pair<uint64_t, uint32_t>> members;
};

struct lxv_multi {
    int device = 0;
    hipStream_t stream = nullptr;
    // Steps alternate between two parities, each with its own events, hit
    // buffer, counters, unit staging and graph, so that with `pipelined` the
    // next step is planned and enqueued before the last one is collected.
    hipEvent_t ev0[2] = {nullptr, nullptr}, ev1[2] = {nullptr, nullptr};
    bool pipelined = false;
    bool inflight = false;  // a launched step not yet collected, of parity 1 - parity
    int parity = 0;         // of the next step to launch
    StepRec rec[2];
    PairAnchor s_anchor;    // the pair walk: stride Q, shared by every search
    std::string err = "no error";
    int layout = LXV_LAYOUT_8X32, batch = 256, walk = LXV_WALK_PAIRS, waves = 4;
    double oldest_share = 0;  // of each step's work, for the longest-waiting resident search
    bool preempt = false;     // a search that does not fit suspends resident ones of lower priority
    uint64_t suspensions = 0;
    uint32_t rounds = 1, limbs = 8;
    uint32_t phys_lanes = 0, pool_lanes = 0, max_slots = 1024, max_vlanes = 1u << 18, lane_keys_log2 = 14;
    bool use_graph = true;
    // device
    SlotConst* d_slots = nullptr;
    SlotDyn* d_dyn = nullptr;
    Unit* d_units = nullptr;
    uint32_t* d_nunits = nullptr;
    uint32_t* d_counters = nullptr;  // 2 x 16 words, one block per parity
    uint32_t* d_pool = nullptr;
    uint2* d_batch = nullptr;
    unsigned long long* d_best = nullptr;  // 2 x max_slots: each step's lowest hit key per slot, one half per parity
    uint32_t* h_stop = nullptr;
    uint32_t* d_stop = nullptr;
    // pinned staging: slot tables mirrored on the host, admission entries
    SlotConst* h_slots = nullptr;
    SlotDyn* h_dyn = nullptr;
    MultiInitEntry* h_init = nullptr;
    MultiInitEntry* d_init = nullptr;
    hipEvent_t ev_admit = nullptr;
    bool admit_pending = false;
    Unit* h_units[2] = {nullptr, nullptr};
    uint32_t* h_nunits[2] = {nullptr, nullptr};
    uint32_t units_cap = 0;
    // What the host reads of a finished step, written by the step's own
    // publish kernel into host-mapped memory (one block per parity): words
    // [0, 4) the counters, then max_slots best keys. Reading it needs no copy,
    // so the read-back of one step never waits behind the next one's kernel,
    // which holds every CU.
    uint32_t* h_pub[2] = {nullptr, nullptr};
    uint32_t* d_pub[2] = {nullptr, nullptr};
    // graphs, one per parity
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    MultiParams params[2];
    // host bookkeeping
    std::deque<uint64_t> queue;
    std::map<uint64_t, Search> searches;
    std::vector<uint64_t> slot_owner;
    std::vector<uint32_t> free_slots;
    std::deque<lxv_search_result> ready;
    Pool pool;
    uint64_t next_id = 1, keys = 0;
    double busy_s = 0, last_ms = 0, rate = 6e9;
    uint32_t done = 0, steps = 0;
    std::chrono::steady_clock::time_point t0;
    uint32_t q[8];

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

// Words of one parity's host-mapped read-back block: 4 counters, then a
// 64-bit best key per slot.
static size_t pub_words(const lxv_multi* m) { return 4 + 2 * (size_t)m->max_slots; }

static int mfail(lxv_multi* m, int code, const std::string& msg) {
    if (m)
        m->err = msg;
    else
        g_err = msg;
    return code;
}

#define MCHK(m, expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return mfail(m, LXV_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static void multi_free(lxv_multi* m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (int p = 0; p < 2; p++) {
        if (m->gexec[p]) (void)hipGraphExecDestroy(m->gexec[p]);
        if (m->graph[p]) (void)hipGraphDestroy(m->graph[p]);
    }
    (void)hipFree(m->d_slots);
    (void)hipFree(m->d_dyn);
    (void)hipFree(m->d_units);
    (void)hipFree(m->d_nunits);
    (void)hipFree(m->d_counters);
    (void)hipFree(m->d_pool);
    (void)hipFree(m->d_batch);
    (void)hipFree(m->d_best);
    (void)hipFree(m->d_init);
    if (m->h_slots) (void)hipHostFree(m->h_slots);
    if (m->h_dyn) (void)hipHostFree(m->h_dyn);
    if (m->h_init) (void)hipHostFree(m->h_init);
    if (m->ev_admit) (void)hipEventDestroy(m->ev_admit);
    if (m->h_stop) (void)hipHostFree(m->h_stop);
    for (int p = 0; p < 2; p++) {
        if (m->h_units[p]) (void)hipHostFree(m->h_units[p]);
        if (m->h_nunits[p]) (void)hipHostFree(m->h_nunits[p]);
        if (m->ev0[p]) (void)hipEventDestroy(m->ev0[p]);
        if (m->ev1[p]) (void)hipEventDestroy(m->ev1[p]);
    }
    for (int p = 0; p < 2; p++)
        if (m->h_pub[p]) (void)hipHostFree(m->h_pub[p]);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

template <class F>
static void q_limbs(const uint32_t q[8], uint32_t qp[10], uint32_t qm[10]) {
    typedef typename F::fe fe;
    const fe qf = F::from_words(q);
    uint32_t a[F::kLimbs], b[F::kLimbs];
    F::store(F::add(qf, F::one()), a);
    F::store(F::sub(qf, F::one()), b);
    for (int i = 0; i < 10; i++) {
        qp[i] = i < F::kLimbs ? a[i] : 0;
        qm[i] = i < F::kLimbs ? b[i] : 0;
    }
}

extern "C" int lxv_multi_open(const lxv_multi_opts* opts, lxv_multi** out) {
    if (!out) return mfail(nullptr, LXV_E_ARG, "lxv_multi_open: out is NULL");
    *out = nullptr;
    lxv_multi_opts o{};
    if (opts) {
        // A shorter struct (an older header) leaves the fields it lacks
        // (those after `lane_keys_log2`) at their zero defaults.
        if (opts->size < offsetof(lxv_multi_opts, walk))
            return mfail(nullptr, LXV_E_ARG, "lxv_multi_open: opts->size too small");
        std::memcpy(&o, opts, opts->size < sizeof o ? opts->size : sizeof o);
    }
    auto* m = new (std::nothrow) lxv_multi();
    if (!m) return mfail(nullptr, LXV_E_NOMEM, "lxv_multi_open: out of host memory");
    m->t0 = std::chrono::steady_clock::now();
    m->device = o.device < 0 ? 0 : o.device;
    const lxv_tuned_shape& tuned = kTunedShapes[LXV_TUNED_BATCHED];
    m->batch = o.batch ? (int)o.batch : (int)tuned.batch;
    const uint32_t wpc = o.waves_per_cu ? o.waves_per_cu : tuned.waves_per_cu;
    m->waves = waves_per_simd(wpc);
    if (o.oldest_share_pct > 90) {
        delete m;
        return mfail(nullptr, LXV_E_ARG, "lxv_multi_open: oldest_share_pct must be 0..90");
    }
    m->oldest_share = o.oldest_share_pct / 100.0;
    m->preempt = o.preempt != 0;
    m->layout = (int)o.layout;
    m->walk = (int)o.walk;
    m->use_graph = !o.no_graph;
    m->pipelined = o.pipelined != 0;
    if (o.max_searches) m->max_slots = o.max_searches;
    if (o.max_vlanes) m->max_vlanes = std::max<uint32_t>(64, o.max_vlanes / 64 * 64);
    if (tuned.lane_keys_log2 >= 8 && tuned.lane_keys_log2 <= 20) m->lane_keys_log2 = tuned.lane_keys_log2;
    if (o.lane_keys_log2) {
        if (o.lane_keys_log2 < 8 || o.lane_keys_log2 > 20) {
            delete m;
            return mfail(nullptr, LXV_E_ARG, "lxv_multi_open: lane_keys_log2 must be 8..20");
        }
        m->lane_keys_log2 = o.lane_keys_log2;
    }
    if (!multi_supported(m->layout, m->batch, m->walk, m->waves)) {
        delete m;
        return mfail(nullptr, LXV_E_ARG,
                     "lxv_multi_open: no batched kernel for this layout, batch, walk and waves per CU");
    }
    int ndev = 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= m->device || hipSetDevice(m->device) != hipSuccess ||
        hipGetDeviceProperties(&prop, m->device) != hipSuccess) {
        delete m;
        return mfail(nullptr, LXV_E_HIP, "no HIP device found (is this an MI355X host?)");
    }
    m->limbs = (uint32_t)limbs_of(m->layout);
    // the persistent grid: exactly the waves one round of the device holds
    const int per_cu = multi_waves_per_cu(m->layout, m->batch, m->walk, m->waves);
    if (per_cu <= 0) {
        delete m;
        return mfail(nullptr, LXV_E_HIP, "lxv_multi_open: cannot read the batched kernel's occupancy");
    }
    m->phys_lanes = (uint32_t)prop.multiProcessorCount * (uint32_t)per_cu * 64u;
    const uint64_t pool_bytes = o.pool_bytes ? o.pool_bytes : (2ull << 30);
    // pool state per virtual lane: the pair walk's centre (x, y), or the
    // projective walk's X(n-1), Z(n-1), X(n), Z(n)
    const uint64_t lane_bytes = (m->walk == LXV_WALK_PAIRS ? 2ull : 4ull) * m->limbs * sizeof(uint32_t);
    m->pool_lanes = (uint32_t)std::min<uint64_t>(pool_bytes / lane_bytes / 64 * 64, 0xffffffc0ull);
    if (m->pool_lanes < 64) {
        delete m;
        return mfail(nullptr, LXV_E_ARG, "lxv_multi_open: pool_bytes too small for one search");
    }
    if (m->max_vlanes > m->pool_lanes) m->max_vlanes = m->pool_lanes;
    m->pool.init(m->pool_lanes);
    m->slot_owner.assign(m->max_slots, 0);
    for (uint32_t s = m->max_slots; s-- > 0;) m->free_slots.push_back(s);
    uint8_t zero[32] = {0};
    lxv_launch_args a{};
    lxv_launch_args_for(zero, &a);
    std::memcpy(m->q, a.q, sizeof m->q);

#define MALLOC(ptr, bytes)                                                        \
    do {                                                                          \
        hipError_t e_ = hipMalloc((void**)&(ptr), (bytes));                        \
        if (e_ != hipSuccess) {                                                   \
            std::string msg = std::string("hipMalloc ") + #ptr + ": " + hipGetErrorString(e_); \
            multi_free(m);                                                        \
            return mfail(nullptr, LXV_E_HIP, msg);                                \
        }                                                                         \
    } while (0)
    MALLOC(m->d_slots, (size_t)m->max_slots * sizeof(SlotConst));
    MALLOC(m->d_dyn, (size_t)m->max_slots * sizeof(SlotDyn));
    MALLOC(m->d_nunits, 64);
    MALLOC(m->d_counters, 2 * 16 * sizeof(uint32_t));
    MALLOC(m->d_pool, (size_t)m->pool_lanes * lane_bytes);
    MALLOC(m->d_batch, (size_t)m->phys_lanes * (m->walk == LXV_WALK_PAIRS ? m->batch / 2 : m->batch * 2) * m->limbs *
                           sizeof(uint32_t));
    MALLOC(m->d_best, 2 * (size_t)m->max_slots * sizeof(unsigned long long));
    MALLOC(m->d_init, (size_t)m->max_slots * sizeof(MultiInitEntry));
#undef MALLOC
    hipError_t e = hipHostMalloc((void**)&m->h_stop, 64, hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&m->d_stop, m->h_stop, 0);
    for (int p = 0; p < 2 && e == hipSuccess; p++) e = hipHostMalloc((void**)&m->h_nunits[p], 64, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&m->h_slots, (size_t)m->max_slots * sizeof(SlotConst), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&m->h_dyn, (size_t)m->max_slots * sizeof(SlotDyn), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&m->h_init, (size_t)m->max_slots * sizeof(MultiInitEntry), hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&m->ev_admit, hipEventDisableTiming);
    for (int p = 0; p < 2 && e == hipSuccess; p++) {
        e = hipHostMalloc((void**)&m->h_pub[p], pub_words(m) * sizeof(uint32_t),
                          hipHostMallocMapped | hipHostMallocCoherent);
        if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&m->d_pub[p], m->h_pub[p], 0);
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    for (int p = 0; p < 2 && e == hipSuccess; p++) {
        e = hipEventCreate(&m->ev0[p]);
        if (e == hipSuccess) e = hipEventCreate(&m->ev1[p]);
    }
    if (m->walk == LXV_WALK_PAIRS) {
        // S = stride Q, shared by every search's lane set-up (host build of
        // the device code)
        const uint32_t w[8] = {(uint32_t)kStride, (uint32_t)(kStride >> 32), 0, 0, 0, 0, 0, 0};
        pair_anchor<Field8>(w, m->s_anchor);
    }
    if (e != hipSuccess) {
        std::string msg = std::string("lxv_multi_open setup: ") + hipGetErrorString(e);
        multi_free(m);
        return mfail(nullptr, LXV_E_HIP, msg);
    }
    *(volatile uint32_t*)m->h_stop = 0;
    for (int p = 0; p < 2; p++) {
        MultiParams& mp = m->params[p];
        std::memset(&mp, 0, sizeof mp);
        if (m->layout == LXV_LAYOUT_8X32)
            q_limbs<Field8>(m->q, mp.qp, mp.qm);
        else
            q_limbs<Field10>(m->q, mp.qp, mp.qm);
        mp.slots = m->d_slots;
        mp.dyn = m->d_dyn;
        mp.n_units = m->d_nunits;
        mp.counters = m->d_counters + 16 * p;
        mp.state = m->d_pool;
        mp.batch_mem = m->d_batch;
        mp.best = m->d_best + (size_t)m->max_slots * p;
        mp.stop = m->d_stop;
        mp.pool_lanes = m->pool_lanes;
        mp.phys_lanes = m->phys_lanes;
    }
    *out = m;
    return LXV_OK;
}

extern "C" int lxv_multi_submit(lxv_multi* m, const lxv_search_spec* spec, uint64_t* id) {
    if (!m || !spec) return LXV_E_ARG;
    if (!start_scalar_ok(spec->start)) return mfail(m, LXV_E_ARG, "lxv_multi_submit: not a valid start scalar");
    char pfx[33];
    std::memcpy(pfx, spec->prefix, 32);
    pfx[32] = 0;
    Prefix p;
    std::string perr;
    if (!Prefix::parse(pfx, &p, &perr)) return mfail(m, LXV_E_ARG, perr);
    Search s;
    s.id = m->next_id++;
    s.tag = spec->tag;
    s.max_keys = spec->max_keys;
    s.priority = spec->priority;
    std::memcpy(s.start, spec->start, 32);
    std::memcpy(s.ed_pub, spec->ed_pub, 32);
    s.vlanes = vlanes_for(p.nibbles, m->max_vlanes, m->lane_keys_log2);
    uint32_t ed[8], want[4], mask[4], nh[3];
    for (int i = 0; i < 8; i++) ed[i] = be_word(spec->ed_pub + 4 * i);
    p.words(want, mask);
    uint8_t nhb[10];
    name_hash(nhb);
    nh[0] = be_word(nhb);
    nh[1] = be_word(nhb + 4);
    nh[2] = (uint32_t)nhb[8] << 24 | (uint32_t)nhb[9] << 16;
    std::memset(&s.sc, 0, sizeof s.sc);
    sha_search_init(s.sc.sha, ed, nh, want, mask);
    // base = start / 8 - 1 (the projective walk's u(k - 8)), or start / 8 +
    // B/2 - 1 (the pair walk's first centre); no wrap for a start scalar
    uint8_t s8[32];
    uint32_t carry = 0;
    for (int i = 31; i >= 0; i--) {
        s8[i] = (uint8_t)((spec->start[i] >> 3) | (carry << 5));
        carry = spec->start[i] & 7;
    }
    for (int i = 0; i < 8; i++)
        s.sc.base[i] = (uint32_t)s8[4 * i] | (uint32_t)s8[4 * i + 1] << 8 | (uint32_t)s8[4 * i + 2] << 16 |
                       (uint32_t)s8[4 * i + 3] << 24;
    if (m->walk == LXV_WALK_PAIRS) {
        uint64_t c = (uint64_t)(m->batch / 2 - 1);
        for (int i = 0; i < 8 && c; i++) {
            const uint64_t v = (uint64_t)s.sc.base[i] + c;
            s.sc.base[i] = (uint32_t)v;
            c = v >> 32;
        }
    } else {
        for (int i = 0; i < 8; i++)
            if (s.sc.base[i]-- != 0) break;
    }
    s.sc.stride = kStride;
    s.sc.vlanes = s.vlanes;
    s.sc.rounds = m->rounds;
    if (m->walk == LXV_WALK_PAIRS) pair_anchor<Field8>(s.sc.base, s.p0);  // canonical words: either layout
    s.submit_s = m->now();
    m->searches[s.id] = s;
    m->queue.push_back(s.id);
    if (id) *id = s.id;
    return LXV_OK;
}

// The LXMF address of key start + 8 g, re-derived from scratch on the host
// (X25519 and both hashes), as big-endian words.
static void answer_address(const Search& s, uint64_t g, uint32_t out[4]) {
    uint8_t k[32], u[32], a[16];
    step_scalar(s.start, g, k);
    x25519_base(k, u);
    lxmf_address(u, s.ed_pub, a);
    for (int i = 0; i < 4; i++) out[i] = be_word(a + 4 * i);
}

static void finish(lxv_multi* m, Search& s, int status, uint64_t g, bool found) {
    lxv_search_result r{};
    r.id = s.id;
    r.tag = s.tag;
    r.g = g;
    if (found) answer_address(s, g, r.address);
    r.keys = s.keys;
    r.vlanes = s.vlanes;
    r.status = status;
    r.submit_s = s.submit_s;
    r.admit_s = s.admit_s;
    r.done_s = m->now();
    m->ready.push_back(r);
    if (s.active) {
        m->pool.give(s.state_off, s.vlanes);
        m->slot_owner[s.slot] = 0;
        m->free_slots.push_back(s.slot);
    }
    m->done++;
    m->searches.erase(s.id);
}

// Whether search r may be suspended to make room: resident, not admitted in
// this call, not walked by the step still in flight (its progress there is
// not collected yet), and not the longest-waiting resident search when that
// one has a share.
static bool suspendable(lxv_multi* m, const Search& r, const std::vector<uint64_t>& admitted_now,
                        const Search* oldest_active) {
    if (!r.active || &r == oldest_active) return false;
    if (std::find(admitted_now.begin(), admitted_now.end(), r.id) != admitted_now.end()) return false;
    if (m->pipelined && m->inflight)
        for (const auto& mem : m->rec[1 - m->parity].members)
            if (mem.first == r.id) return false;
    return true;
}

// Suspends a resident search: its slot and pool lanes are freed and it goes
// back to the queue with the progress collected so far. A partly walked
// block is rolled back to its start (its keys are walked again after the
// resumption, which sets every lane up at that block from the start
// scalar), so its key count and answer are those of an uninterrupted walk.
static void suspend(lxv_multi* m, Search& r) {
    m->pool.give(r.state_off, r.vlanes);
    m->slot_owner[r.slot] = 0;
    m->free_slots.push_back(r.slot);
    r.active = false;
    r.slot = kNone;
    r.next_block = r.done_block;
    r.wave_cursor = 0;
    r.done_cursor = 0;
    r.suspensions++;
    m->suspensions++;
}

// The set-up scalar of a search resumed at key index n0 of every virtual
// lane: base + n0 as 8 little-endian words (base = start / 8 + B/2 - 1 for
// the pair walk, start / 8 - 1 for the projective walk).
static void base_plus(const uint32_t base[8], uint64_t n0, uint32_t out[8]) {
    uint64_t c = n0;
    for (int i = 0; i < 8; i++) {
        const uint64_t v = (uint64_t)base[i] + (c & 0xffffffffull);
        out[i] = (uint32_t)v;
        c = (c >> 32) + (v >> 32);
    }
}

// Admits queued searches by (priority, submission order), each as soon as a
// slot and a pool range of its size are free; one that does not fit yet does
// not hold back smaller ones behind it. With `preempt`, a search that does
// not fit suspends resident searches of a lower priority (a larger value),
// lowest priority and latest submission first, until it fits. The admitted
// searches' slot tables are copied from pinned host mirrors and their lanes
// set up by one launch (at the block each has reached, for resumed ones),
// all enqueued ahead of the step's walk on the same stream.
static int admit(lxv_multi* m) {
    if (m->queue.empty()) return LXV_OK;
    if (m->admit_pending) {
        // the staging entries of the last admission are read by its copy
        MCHK(m, hipEventSynchronize(m->ev_admit));
        m->admit_pending = false;
    }
    std::sort(m->queue.begin(), m->queue.end(), [m](uint64_t a, uint64_t b) {
        const uint32_t pa = m->searches.at(a).priority, pb = m->searches.at(b).priority;
        return pa != pb ? pa < pb : a < b;
    });
    if (m->oldest_share > 0) {
        // the longest-waiting queued search goes first, whatever its priority
        auto first = std::min_element(m->queue.begin(), m->queue.end());
        std::rotate(m->queue.begin(), first, first + 1);
    }
    const Search* oldest_active = nullptr;
    if (m->oldest_share > 0)
        for (auto& kv : m->searches)
            if (kv.second.active && (!oldest_active || kv.first < oldest_active->id)) oldest_active = &kv.second;
    std::vector<uint64_t> admitted_now, suspended;
    uint32_t n = 0, max_vl = 0;
    for (auto qi = m->queue.begin(); qi != m->queue.end();) {
        Search& s = m->searches.at(*qi);
        uint32_t off = 0;
        bool fits = !m->free_slots.empty() && m->pool.take(s.vlanes, &off);
        while (!fits && m->preempt && suspended.size() < kMaxSuspendPerStep) {
            Search* victim = nullptr;
            for (auto& kv : m->searches) {
                Search& r = kv.second;
                if (r.priority <= s.priority || !suspendable(m, r, admitted_now, oldest_active)) continue;
                if (!victim || r.priority > victim->priority || (r.priority == victim->priority && r.id > victim->id))
                    victim = &r;
            }
            if (!victim) break;
            suspend(m, *victim);
            suspended.push_back(victim->id);
            fits = !m->free_slots.empty() && m->pool.take(s.vlanes, &off);
        }
        if (!fits) {  // admission by device memory
            ++qi;
            continue;
        }
        qi = m->queue.erase(qi);
        admitted_now.push_back(s.id);
        s.slot = m->free_slots.back();
        m->free_slots.pop_back();
        s.state_off = off;
        s.sc.state_off = off;
        s.sc.slot = s.slot;
        s.active = true;
        if (s.suspensions == 0) s.admit_s = m->now();
        m->slot_owner[s.slot] = s.id;
        // a slot's host mirror is rewritten only here, after the step that
        // last read it has been waited for
        m->h_slots[s.slot] = s.sc;
        std::memset(&m->h_dyn[s.slot], 0, sizeof(SlotDyn));
        m->h_dyn[s.slot].found_block = s.hit ? s.hit_block : kNone;
        MCHK(m, hipMemcpyAsync(m->d_slots + s.slot, m->h_slots + s.slot, sizeof(SlotConst), hipMemcpyHostToDevice,
                               m->stream));
        MCHK(m, hipMemcpyAsync(m->d_dyn + s.slot, m->h_dyn + s.slot, sizeof(SlotDyn), hipMemcpyHostToDevice,
                               m->stream));
        MultiInitEntry& e = m->h_init[n++];
        const uint64_t n0 = (uint64_t)s.next_block * m->rounds * (uint64_t)m->batch;
        base_plus(s.sc.base, n0, e.base);
        if (m->walk == LXV_WALK_PAIRS) {
            if (n0 == 0)
                e.p0 = s.p0;
            else
                pair_anchor<Field8>(e.base, e.p0);  // the resumed block's centre of virtual lane 0
        }
        e.state_off = off;
        e.vlanes = s.vlanes;
        e.pad[0] = e.pad[1] = 0;
        max_vl = std::max(max_vl, s.vlanes);
    }
    for (uint64_t id : suspended) m->queue.push_back(id);
    if (n) {
        MCHK(m, hipMemcpyAsync(m->d_init, m->h_init, (size_t)n * sizeof(MultiInitEntry), hipMemcpyHostToDevice,
                               m->stream));
        MCHK(m, launch_multi_init(m->layout, m->walk, m->d_pool, m->pool_lanes, m->q, kStride, m->d_init, n, max_vl,
                                  &m->s_anchor, m->stream));
        MCHK(m, hipEventRecord(m->ev_admit, m->stream));
        m->admit_pending = true;
    }
    return LXV_OK;
}

// Every slot's best hit key of the step with parity p set to kNoHit (all
// ones) by the fill kernel, which a captured graph replays correctly.
__global__ void fill_ones_kernel(uint32_t* words, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) words[i] = 0xffffffffu;
}

// Copies the step's counters (4 words) and its best key per slot into the
// host-mapped block `out` with plain stores, after the walk kernel in stream
// order.
__global__ void publish_kernel(const uint32_t* counters, const unsigned long long* best, uint32_t slots,
                               uint32_t* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 4) out[i] = counters[i];
    if (i < slots) {
        const unsigned long long k = best[i];
        out[4 + 2 * i] = (uint32_t)k;
        out[5 + 2 * i] = (uint32_t)(k >> 32);
    }
}

static hipError_t publish(lxv_multi* m, int p) {
    const uint32_t n = std::max<uint32_t>(m->max_slots, 4);
    hipLaunchKernelGGL(publish_kernel, dim3((n + 255) / 256), dim3(256), 0, m->stream, m->d_counters + 16 * p,
                       m->d_best + (size_t)m->max_slots * p, m->max_slots, m->d_pub[p]);
    return hipGetLastError();
}

static hipError_t clear_best(lxv_multi* m, int p) {
    const uint32_t n = 2 * m->max_slots;
    hipLaunchKernelGGL(fill_ones_kernel, dim3((n + 255) / 256), dim3(256), 0, m->stream,
                       reinterpret_cast<uint32_t*>(m->d_best + (size_t)m->max_slots * p), n);
    return hipGetLastError();
}

static int launch_step(lxv_multi* m, int p) {
    uint32_t* counters = m->d_counters + 16 * p;
    if (m->use_graph) {
        if (!m->gexec[p]) {
            // capture on an idle stream (once per parity, and after the unit
            // list moves)
            MCHK(m, hipStreamSynchronize(m->stream));
            MCHK(m, hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = launch_clear(counters, 16, m->stream);
            if (e == hipSuccess) e = clear_best(m, p);
            if (e == hipSuccess) e = launch_multi(m->layout, m->batch, m->walk, m->waves, m->params[p], m->stream);
            if (e == hipSuccess) e = publish(m, p);
            hipError_t e2 = hipStreamEndCapture(m->stream, &m->graph[p]);
            if (e != hipSuccess) return mfail(m, LXV_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            MCHK(m, e2);
            MCHK(m, hipGraphInstantiate(&m->gexec[p], m->graph[p], nullptr, nullptr, 0));
        }
        MCHK(m, hipEventRecord(m->ev0[p], m->stream));
        MCHK(m, hipGraphLaunch(m->gexec[p], m->stream));
        MCHK(m, hipEventRecord(m->ev1[p], m->stream));
    } else {
        MCHK(m, launch_clear(counters, 16, m->stream));
        MCHK(m, clear_best(m, p));
        MCHK(m, hipEventRecord(m->ev0[p], m->stream));
        MCHK(m, launch_multi(m->layout, m->batch, m->walk, m->waves, m->params[p], m->stream));
        MCHK(m, publish(m, p));
        MCHK(m, hipEventRecord(m->ev1[p], m->stream));
    }
    return LXV_OK;
}

// Plans about budget_ms of work from the planned progress and launches it
// with parity p; *launched says whether anything was. Each active search may
// use at most its lanes' worth of the chip for the budget. Priorities are
// served in ascending order; searches of one priority share what is left
// equally (water-filling), so a search may get nothing this step and keep its
// state for later. A search whose share is at least a block walks whole
// blocks; otherwise it walks that many virtual waves of its current block,
// and the block completes over the next steps.
static int plan_step(lxv_multi* m, uint32_t budget_ms, int p, bool* launched) {
    *launched = false;
    m->rec[p].members.clear();
    std::vector<Search*> act;
    for (auto& kv : m->searches)
        if (kv.second.active) act.push_back(&kv.second);
    if (act.empty()) return LXV_OK;
    const double budget_s = budget_ms / 1e3;
    const double wave_keys = 64.0 * m->rounds * m->batch;          // one virtual wave, one block
    const double lane_keys = budget_s * m->rate / m->phys_lanes;  // one lane, busy for the budget
    std::vector<std::pair<double, Search*>> want;
    for (Search* s : act)
        want.push_back({(double)std::min(s->vlanes, m->phys_lanes) * std::max(lane_keys, (double)m->rounds * m->batch), s});
    std::sort(want.begin(), want.end(), [](const std::pair<double, Search*>& a, const std::pair<double, Search*>& b) {
        if (a.second->priority != b.second->priority) return a.second->priority < b.second->priority;
        return a.first != b.first ? a.first < b.first : a.second->id < b.second->id;
    });
    // the longest-waiting resident search (lowest id: submission order)
    // first gets oldest_share of the step, up to what its lanes can take
    Search* oldest = nullptr;
    if (m->oldest_share > 0)
        for (Search* s : act)
            if (!oldest || s->id < oldest->id) oldest = s;
    double left = m->rate * budget_s;
    double reserved = 0;
    if (oldest) {
        for (const auto& w : want)
            if (w.second == oldest) reserved = std::min(w.first, m->oldest_share * left);
        left -= reserved;
    }
    size_t n_units = 0;
    for (size_t i = 0; i < want.size(); i++) {
        Search* s = want[i].second;
        size_t tier_end = i;
        while (tier_end < want.size() && want[tier_end].second->priority == s->priority) tier_end++;
        const double extra = s == oldest ? reserved : 0.0;
        const double share = std::min(want[i].first - extra, left / (double)(tier_end - i));
        left -= share;
        const double total = share + extra;
        const uint32_t vwaves = s->vlanes / 64;
        const double block_keys = wave_keys * vwaves;
        if (total < 0.5 * wave_keys) {
            s->w0 = s->w1 = s->wave_cursor;  // nothing this step
            s->nblocks = 0;
        } else if (s->wave_cursor == 0 && total >= block_keys && !s->hit) {
            s->w0 = 0;
            s->w1 = vwaves;
            s->nblocks = (uint32_t)std::floor(total / block_keys);
        } else {
            const uint32_t k = (uint32_t)std::max(1.0, std::round(total / wave_keys));
            s->w0 = s->wave_cursor;
            s->w1 = std::min(vwaves, s->wave_cursor + k);
            s->nblocks = 1;
        }
        n_units += s->w1 - s->w0;
    }
    if (n_units == 0) {
        // shares below one unit each: the first search in service order
        // still walks one virtual wave, so every step makes progress
        Search* s = want[0].second;
        s->w0 = s->wave_cursor;
        s->w1 = s->wave_cursor + 1;
        s->nblocks = 1;
        n_units = 1;
    }
    if (n_units > m->units_cap) {
        // the staging and device unit lists may still be read by a step in
        // flight: let it finish before they are replaced
        MCHK(m, hipStreamSynchronize(m->stream));
        for (int q = 0; q < 2; q++) {
            if (m->h_units[q]) (void)hipHostFree(m->h_units[q]);
            m->h_units[q] = nullptr;
        }
        (void)hipFree(m->d_units);
        m->d_units = nullptr;
        m->units_cap = 0;
        const size_t capn = std::max<size_t>(n_units, 4096);
        for (int q = 0; q < 2; q++) MCHK(m, hipHostMalloc((void**)&m->h_units[q], capn * sizeof(Unit), hipHostMallocDefault));
        MCHK(m, hipMalloc((void**)&m->d_units, capn * sizeof(Unit)));
        m->units_cap = (uint32_t)capn;
        for (int q = 0; q < 2; q++) {
            m->params[q].units = m->d_units;
            if (m->gexec[q]) {
                (void)hipGraphExecDestroy(m->gexec[q]);
                (void)hipGraphDestroy(m->graph[q]);
                m->gexec[q] = nullptr;
                m->graph[q] = nullptr;
            }
        }
    }
    // Longest units first, so the tail of the launch is short units.
    std::sort(act.begin(), act.end(), [](const Search* a, const Search* b) {
        return a->nblocks != b->nblocks ? a->nblocks > b->nblocks : a->id < b->id;
    });
    size_t k = 0;
    for (Search* s : act) {
        if (s->w1 > s->w0) m->rec[p].members.push_back({s->id, s->slot});
        for (uint32_t w = s->w0; w < s->w1; w++) m->h_units[p][k++] = Unit{s->slot, w, s->next_block, s->nblocks};
        // the planned progress once this step has run
        const uint32_t vwaves = s->vlanes / 64;
        if (s->w1 > s->w0) {
            if (s->w0 == 0 && s->w1 == vwaves) {
                s->next_block += s->nblocks;  // whole blocks (a hit stops units only past its block)
            } else {
                s->wave_cursor = s->w1;
                if (s->wave_cursor == vwaves) {
                    s->wave_cursor = 0;
                    s->next_block += 1;
                }
            }
        }
        s->after_block[p] = s->next_block;
        s->after_cursor[p] = s->wave_cursor;
    }
    *m->h_nunits[p] = (uint32_t)k;
    MCHK(m, hipMemcpyAsync(m->d_units, m->h_units[p], k * sizeof(Unit), hipMemcpyHostToDevice, m->stream));
    MCHK(m, hipMemcpyAsync(m->d_nunits, m->h_nunits[p], sizeof(uint32_t), hipMemcpyHostToDevice, m->stream));
    const int rc = launch_step(m, p);
    if (rc) return rc;
    *launched = true;
    return LXV_OK;
}

// Waits for the step launched with parity p and credits it: keys, the rate
// estimate, every hit to the search that owned its slot when the step was
// planned (searches finished since are skipped), the completed progress, and
// the searches whose answer is final or whose key budget is spent.
static int collect_step(lxv_multi* m, int p) {
    MCHK(m, hipEventSynchronize(m->ev1[p]));
    float ms = 0;
    MCHK(m, hipEventElapsedTime(&ms, m->ev0[p], m->ev1[p]));
    // the step's publish kernel has written its counters and best keys to
    // host-mapped memory before ev1[p]
    const volatile uint32_t* pub = m->h_pub[p];
    const uint64_t kkeys = (uint64_t)pub[2] | (uint64_t)pub[3] << 32;
    m->steps++;
    m->keys += kkeys;
    m->busy_s += ms / 1e3;
    m->last_ms = ms;
    if (ms > 1.0 && kkeys > 0) m->rate = 0.5 * m->rate + 0.5 * (kkeys / (ms / 1e3));
    // the searches this step walked, if they still hold the slot they had
    // when it was planned (searches finished since are skipped)
    std::vector<Search*> walked;
    for (const auto& mem : m->rec[p].members) {
        auto it = m->searches.find(mem.first);
        if (it == m->searches.end() || !it->second.active || it->second.slot != mem.second) continue;
        Search* s = &it->second;
        walked.push_back(s);
        // keep the lowest hit in walk order: block, then n, then v (the key
        // orders by (n, v), and n determines the block)
        const unsigned long long key = (unsigned long long)pub[4 + 2 * mem.second] |
                                       (unsigned long long)pub[5 + 2 * mem.second] << 32;
        if (key == kNoHit) continue;
        const uint64_t n = key >> kHitVlaneBits;
        const uint32_t v = (uint32_t)(key & ((1u << kHitVlaneBits) - 1));
        const uint32_t blk = (uint32_t)(n / ((uint64_t)m->rounds * m->batch));
        if (!s->hit || hit_key(n, v) < hit_key(s->hit_n, s->hit_vlane)) {
            s->hit = true;
            s->hit_block = blk;
            s->hit_n = n;
            s->hit_vlane = v;
        }
    }
    const double block_of = (double)m->rounds * m->batch;  // keys per lane per block
    for (Search* s : walked) {
        s->done_block = s->after_block[p];
        s->done_cursor = s->after_cursor[p];
        if (s->hit && s->done_block > s->hit_block) {
            // every block up to the hit's is walked: the answer is final
            s->keys = (uint64_t)(s->hit_block + 1) * s->vlanes * (uint64_t)block_of;
            finish(m, *s, 1, (uint64_t)s->hit_vlane * kStride + s->hit_n, true);
            continue;
        }
        s->keys = (uint64_t)s->done_block * s->vlanes * (uint64_t)block_of +
                  (uint64_t)s->done_cursor * 64 * (uint64_t)block_of;
        if (s->max_keys && s->keys >= s->max_keys) finish(m, *s, 0, 0, false);
    }
    return LXV_OK;
}

extern "C" int lxv_multi_step(lxv_multi* m, uint32_t budget_ms, lxv_search_result* out, uint32_t cap,
                              uint32_t* n_out) {
    if (!m) return LXV_E_ARG;
    MCHK(m, hipSetDevice(m->device));
    int rc = admit(m);
    if (rc) return rc;
    const int p = m->parity;
    bool launched = false;
    if (budget_ms) {  // 0: launch nothing (with `pipelined`: only collect the step in flight)
        rc = plan_step(m, budget_ms, p, &launched);
        if (rc) return rc;
    }
    if (m->pipelined) {
        // the step launched last call ran meanwhile; collect it while this
        // one runs. The step just launched is tracked first, so a failed
        // collection leaves the engine consistent.
        const bool had = m->inflight;
        m->inflight = launched;
        if (launched) m->parity = 1 - p;
        if (had) {
            rc = collect_step(m, 1 - p);
            if (rc) return rc;
        }
    } else if (launched) {
        rc = collect_step(m, p);
        if (rc) return rc;
    }
    uint32_t n = 0;
    while (out && n < cap && !m->ready.empty()) {
        out[n++] = m->ready.front();
        m->ready.pop_front();
    }
    if (n_out) *n_out = n;
    return LXV_OK;
}

extern "C" int lxv_multi_cancel(lxv_multi* m, uint64_t id) {
    if (!m) return LXV_E_ARG;
    auto it = m->searches.find(id);
    if (it == m->searches.end()) return mfail(m, LXV_E_ARG, "lxv_multi_cancel: no such search");
    if (!it->second.active) m->queue.erase(std::find(m->queue.begin(), m->queue.end(), id));
    finish(m, it->second, -1, 0, false);
    return LXV_OK;
}

extern "C" int lxv_multi_stats(lxv_multi* m, lxv_multi_stat* out) {
    if (!m || !out) return LXV_E_ARG;
    out->keys = m->keys;
    out->busy_s = m->busy_s;
    out->last_step_ms = m->last_ms;
    uint32_t res = 0;
    for (auto& kv : m->searches) res += kv.second.active ? 1 : 0;
    out->resident = res;
    out->queued = (uint32_t)m->queue.size();
    out->done = m->done;
    out->steps = m->steps;
    return LXV_OK;
}

extern "C" void lxv_multi_close(lxv_multi* m) { multi_free(m); }

extern "C" int lxv_debug_multi_suspensions(lxv_multi* m, uint64_t* n) {
    if (!m || !n) return LXV_E_ARG;
    *n = m->suspensions;
    return LXV_OK;
}

extern "C" const char* lxv_multi_last_error(lxv_multi* m) { return m ? m->err.c_str() : g_err.c_str(); }
