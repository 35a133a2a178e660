// lxmf-vanity: finds a Reticulum identity whose LXMF address starts with a
// hex prefix, on an MI355X through the C ABI of include/lxmf_vanity_hip.h,
// and/or on the CPU cores.
//
// The search command of src/main.rs (USAGE 325-335, open_gpu 339-353,
// run_search 362-429, main 431-498, the CPU worker `search` 292-323) and the GPU driver of src/gpu/mod.rs
// (GpuSearch::open/run/self_test, 96-207), with the same output lines, usage
// and error texts, exit codes and identity-file semantics: fleet scripts that
// parse the reference's output parse this one's.
//
// GPU lanes start at one random start scalar k, lane t at k + 8 (t x 2^40),
// set up on the device (lxv_init_seeded) instead of per-lane random starts
// whose public keys the reference computes on the CPU (src/gpu/mod.rs:210-
// 228). Every GPU hit is re-derived from scratch on the CPU before it counts.
#include <fcntl.h>
#include <sched.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lxmf_vanity_hip.h"
#include "host_crypto.h"

namespace {

using Clock = std::chrono::steady_clock;
using Identity = std::array<uint8_t, 64>;

constexpr uint32_t kMaxHits = 4096;                     // src/gpu/mod.rs:23
constexpr double kTargetLaunchS = 0.5;                  // src/gpu/mod.rs:26
constexpr double kReportIntervalS = 10.0;               // src/main.rs:37
constexpr uint64_t kLaneStride = 1ull << 40;            // key distance between GPU lanes, at most

const char* kUsage =
    "usage: lxmf-vanity [--gpu [--cpu]] <hex prefix>\n"
    "       lxmf-vanity --gpu-self-test\n"
    "\n"
    "Finds a Reticulum identity whose LXMF address starts with the prefix and\n"
    "saves it to <prefix>.identity (readable only by you).\n"
    "\n"
    "  (no flag)        search on every CPU core\n"
    "  --gpu            search on the GPU: HIP on an AMD Instinct MI355X (gfx950)\n"
    "  --gpu --cpu      search on the GPU and the remaining CPU cores together\n"
    "  --gpu-self-test  check the GPU's results against the CPU";

// The CPUs this process may use, as Rust's available_parallelism counts
// them (src/main.rs:462): the affinity mask, capped by a cgroup CPU quota.
size_t available_parallelism() {
    size_t n = 0;
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof set, &set) == 0) n = (size_t)CPU_COUNT(&set);
    if (n == 0) {
        const unsigned hw = std::thread::hardware_concurrency();
        n = hw ? hw : 1;
    }
    // cgroup v2: "max" or "<quota> <period>" in cpu.max
    if (FILE* f = std::fopen("/sys/fs/cgroup/cpu.max", "r")) {
        char quota[32] = {0};
        unsigned long long period = 0;
        if (std::fscanf(f, "%31s %llu", quota, &period) == 2 && std::strcmp(quota, "max") != 0 && period) {
            const unsigned long long q = std::strtoull(quota, nullptr, 10) / period;
            n = std::min<size_t>(n, std::max<unsigned long long>(q, 1));
        }
        std::fclose(f);
    }
    return n;
}

double seconds_since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

// Shared by the workers of one search (run_search's AtomicU64, AtomicBool
// and Mutex<Option<[u8; 64]>>).
struct Shared {
    std::atomic<uint64_t> tries{0};
    std::atomic<bool> found{false};
    std::mutex mu;
    bool have_hit = false;
    Identity hit{};
    std::string gpu_error;

    void offer(const Identity& id) {
        std::lock_guard<std::mutex> g(mu);
        if (!have_hit) {
            have_hit = true;
            hit = id;
        }
        found.store(true);
    }
};

// ---- the GPU driver: GpuSearch (src/gpu/mod.rs:86-207)

class GpuSearch {
  public:
    static std::unique_ptr<GpuSearch> open(std::string* err) {
        std::unique_ptr<GpuSearch> g(new GpuSearch());
        // batch and waves_per_cu 0: the library's tuned single-search shape
        lxv_open_opts o{};
        o.size = sizeof o;
        if (lxv_open(&o, &g->ctx_) != LXV_OK) {
            *err = lxv_last_error(nullptr);
            return nullptr;
        }
        if (lxv_random_start_scalar(g->start_) != LXV_OK || !lxv::random_bytes(g->seed_, 32)) {
            *err = "no system randomness";
            return nullptr;
        }
        // lanes x stride stays below 2^63: with many lanes, a smaller stride
        // (2^24 lanes: 2^38 keys per lane, years of walking at any rate)
        const uint64_t lanes = lxv_threads(g->ctx_);
        while (g->stride_ > 1 && lanes > 1 && g->stride_ > (UINT64_MAX >> 1) / lanes) g->stride_ >>= 1;
        if (lxv_init_seeded(g->ctx_, g->start_, g->stride_) != LXV_OK) {
            *err = lxv_last_error(g->ctx_);
            return nullptr;
        }
        uint8_t ed_pub[32];
        lxv::ed25519_public(g->seed_, ed_pub);
        lxv_launch_args_for(ed_pub, &g->base_);
        return g;
    }

    ~GpuSearch() {
        if (ctx_) lxv_close(ctx_);
    }

    std::string describe() const {
        return std::string(lxv_describe(ctx_)) + ": " + std::to_string(lxv_threads(ctx_)) + " GPU threads";
    }

    // The identity for lane t's key number n (src/gpu/mod.rs:139-144).
    Identity identity(uint32_t t, uint64_t n) const {
        Identity id;
        lxv::step_scalar(start_, (uint64_t)t * stride_ + n, id.data());
        std::memcpy(id.data() + 32, seed_, 32);
        return id;
    }

    // Searches until this or another worker finds a match (src/gpu/mod.rs:146-174).
    bool run(const lxv::Prefix& prefix, Shared& sh, std::string* err) {
        uint32_t iters = 1;
        while (!sh.found.load()) {
            const auto launched = Clock::now();
            std::vector<lxv_hit> hits;
            lxv_launch_stats st{};
            if (!launch(prefix, iters, 0, &hits, nullptr, &sh.found, &st, err)) return false;
            const double took = seconds_since(launched);
            sh.tries.fetch_add(st.keys);
            if (!hits.empty()) {
                const lxv_hit& h = hits[0];
                const Identity id = identity(h.thread, h.n);
                uint8_t addr[16];
                lxv::address_of_identity(id.data(), addr);
                if (!prefix.matches(addr)) {
                    *err = "GPU hit (thread " + std::to_string(h.thread) + ", key " + std::to_string(h.n) +
                           ") failed its CPU recheck: a kernel bug";
                    return false;
                }
                sh.offer(id);
                return true;
            }
            // Grow launches towards the target length.
            if (took < kTargetLaunchS / 2 && iters < (1u << 16)) iters *= 2;
        }
        return true;
    }

    // Checks the kernel against the CPU: its public keys against plain
    // X25519, and every hit for a short prefix against the CPU's address
    // (src/gpu/mod.rs:176-207).
    bool self_test(std::string* report, std::string* err) {
        const uint32_t dump_threads = std::min<uint32_t>(lxv_threads(ctx_), 256);
        const uint64_t batch = lxv_batch(ctx_);
        lxv::Prefix any;
        lxv::Prefix::parse("b0", &any, nullptr);  // about 1 key in 256
        std::vector<lxv_hit> hits;
        std::vector<uint32_t> dumped((size_t)dump_threads * batch * 8);
        if (!launch(any, 1, dump_threads, &hits, &dumped, nullptr, nullptr, err)) return false;
        for (uint32_t t = 0; t < dump_threads; t++) {
            for (uint64_t i = 0; i < batch; i++) {
                uint8_t k[32], expected[32], got[32];
                lxv::step_scalar(start_, (uint64_t)t * stride_ + i, k);
                lxv::x25519_base(k, expected);
                const size_t base = (size_t)(t * batch + i) * 8;
                for (int w = 0; w < 8; w++)
                    for (int b = 0; b < 4; b++) got[4 * w + b] = (uint8_t)(dumped[base + w] >> (8 * b));
                if (std::memcmp(got, expected, 32) != 0) {
                    *err = "public key mismatch: thread " + std::to_string(t) + ", key " + std::to_string(i);
                    return false;
                }
            }
        }
        if (hits.size() < 100) {
            *err = "only " + std::to_string(hits.size()) + " hits for a 2-digit prefix; expected thousands";
            return false;
        }
        for (const lxv_hit& h : hits) {
            uint8_t addr[16];
            lxv::address_of_identity(identity(h.thread, h.n).data(), addr);
            if (!any.matches(addr)) {
                *err = "hit (thread " + std::to_string(h.thread) + ", key " + std::to_string(h.n) +
                       ") failed its CPU recheck";
                return false;
            }
        }
        *report = std::to_string((uint64_t)dump_threads * batch) + " public keys match X25519, " +
                  std::to_string(hits.size()) + " hits for 'b0' all recheck on the CPU.";
        return true;
    }

  private:
    GpuSearch() = default;

    // One launch: enqueue, then poll until done; a match found meanwhile by
    // another worker stops the kernel after its current batch.
    bool launch(const lxv::Prefix& prefix, uint32_t iters, uint32_t dump_threads, std::vector<lxv_hit>* hits,
                std::vector<uint32_t>* dump, const std::atomic<bool>* found, lxv_launch_stats* st,
                std::string* err) {
        lxv_launch_args a = base_;
        prefix.words(a.want, a.mask);
        a.iters = iters;
        a.max_hits = kMaxHits;
        a.dump_threads = dump_threads;
        if (lxv_launch(ctx_, &a) != LXV_OK) {
            *err = lxv_last_error(ctx_);
            return false;
        }
        hits->resize(kMaxHits);
        uint32_t n = 0;
        bool stopped = false;
        for (;;) {
            const int rc = lxv_poll(ctx_, 0, hits->data(), kMaxHits, &n, dump ? dump->data() : nullptr,
                                    dump ? dump->size() : 0);
            if (rc == LXV_OK) break;
            if (rc != LXV_E_BUSY) {
                *err = lxv_last_error(ctx_);
                return false;
            }
            if (found && found->load() && !stopped) {
                lxv_stop(ctx_);
                stopped = true;
            }
            std::this_thread::sleep_for(std::chrono::microseconds(500));
        }
        hits->resize(std::min(n, kMaxHits));
        if (st) lxv_last_stats(ctx_, st);
        return true;
    }

    lxv_ctx* ctx_ = nullptr;
    uint8_t start_[32] = {0};
    uint8_t seed_[32] = {0};
    uint64_t stride_ = kLaneStride;
    lxv_launch_args base_{};
};

// ---- the CPU worker (src/main.rs:292-323)

void cpu_search(const lxv::Prefix& prefix, Shared& sh) {
    uint8_t seed[32], ed_pub[32], start[32], nh[10];
    if (!lxv::random_bytes(seed, 32) || lxv_random_start_scalar(start) != LXV_OK) {
        std::fprintf(stderr, "no system randomness\n");
        std::exit(2);
    }
    lxv::ed25519_public(seed, ed_pub);
    lxv::name_hash(nh);
    lxv::KeyWalk walk(start);
    static thread_local uint8_t keys[lxv::KeyWalk::kBatch][32];
    while (!sh.found.load(std::memory_order_relaxed)) {
        const uint64_t first = walk.next_batch(keys, lxv::KeyWalk::kBatch);
        for (size_t i = 0; i < lxv::KeyWalk::kBatch; i++) {
            uint8_t addr[16];
            lxv::lxmf_address_nh(nh, keys[i], ed_pub, addr);
            if (!prefix.matches(addr)) continue;
            Identity id;
            lxv::step_scalar(walk.start(), first + i, id.data());
            std::memcpy(id.data() + 32, seed, 32);
            // Recompute from scratch: never save a key the fast path got wrong.
            lxv::address_of_identity(id.data(), addr);
            if (!prefix.matches(addr)) {
                std::fprintf(stderr, "internal error: a hit failed its recheck; not saving it\n");
                std::exit(2);
            }
            sh.offer(id);
            return;
        }
        sh.tries.fetch_add(lxv::KeyWalk::kBatch, std::memory_order_relaxed);
    }
}

std::unique_ptr<GpuSearch> open_gpu() {
    const auto started = Clock::now();
    std::string err;
    auto gpu = GpuSearch::open(&err);
    if (!gpu) {
        std::fprintf(stderr, "GPU unavailable: %s\n", err.c_str());
        std::exit(1);
    }
    std::printf("%s, ready in %.1f s.\n", gpu->describe().c_str(), seconds_since(started));
    std::fflush(stdout);
    return gpu;
}

std::string join_on(bool use_gpu, size_t cpu_threads) {
    std::vector<std::string> on;
    if (use_gpu) on.push_back("the GPU");
    if (cpu_threads > 0) on.push_back(std::to_string(cpu_threads) + " CPU threads");
    std::string s;
    for (size_t i = 0; i < on.size(); i++) s += (i ? " and " : "") + on[i];
    return s;
}

// run_search (src/main.rs:362-429): returns the identity and the keys tried.
Identity run_search(const lxv::Prefix& prefix, bool use_gpu, size_t cpu_threads, uint64_t* tried) {
    const double expected = std::pow(16.0, prefix.nibbles);
    Shared sh;
    std::printf("Searching on %s; about %.0f tries on average.\n", join_on(use_gpu, cpu_threads).c_str(), expected);
    std::fflush(stdout);
    const auto start = Clock::now();
    std::vector<std::thread> workers;
    for (size_t i = 0; i < cpu_threads; i++) workers.emplace_back([&] { cpu_search(prefix, sh); });
    // The GPU is opened on its own thread: a context stays on one thread.
    if (use_gpu) {
        workers.emplace_back([&] {
            const auto started = Clock::now();
            std::string err;
            auto gpu = GpuSearch::open(&err);
            bool ok = false;
            if (gpu) {
                std::printf("%s, ready in %.1f s.\n", gpu->describe().c_str(), seconds_since(started));
                std::fflush(stdout);
                ok = gpu->run(prefix, sh, &err);
            }
            if (!ok) {
                std::lock_guard<std::mutex> g(sh.mu);
                sh.gpu_error = "GPU: " + err;
                sh.found.store(true);
            }
        });
    }
    double next_report = kReportIntervalS;
    while (!sh.found.load()) {
        std::this_thread::sleep_for(std::chrono::milliseconds(50));
        const double now = seconds_since(start);
        if (now >= next_report) {
            const double done = (double)sh.tries.load();
            const double rate = done / now;
            const double left = std::max(expected - done, 0.0) / rate / 60.0;
            if (done < expected)
                std::printf("%.0f tries, %.0f/s; the average is about %.0f min away\n", done, rate, left);
            else
                std::printf("%.0f tries, %.0f/s; past the average, still searching\n", done, rate);
            std::fflush(stdout);
            next_report += kReportIntervalS;
        }
    }
    for (auto& w : workers) w.join();
    if (!sh.gpu_error.empty()) {
        std::fprintf(stderr, "%s\n", sh.gpu_error.c_str());
        std::exit(2);
    }
    *tried = sh.tries.load();
    return sh.hit;
}

std::string lower(std::string s) {
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

}  // namespace

int main(int argc, char** argv) {
    std::vector<std::string> flags, words;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        (a.rfind("--", 0) == 0 ? flags : words).push_back(a);
    }
    auto has = [&](const char* f) { return std::find(flags.begin(), flags.end(), f) != flags.end(); };
    const bool use_gpu = has("--gpu"), with_cpu = has("--cpu");
    for (const auto& f : flags) {
        if (f != "--gpu" && f != "--cpu" && f != "--gpu-self-test") {
            std::fprintf(stderr, "%s\n", kUsage);
            return 1;
        }
    }
    if (with_cpu && !use_gpu) {
        std::fprintf(stderr, "%s\n", kUsage);
        return 1;
    }

    if (has("--gpu-self-test")) {
        auto gpu = open_gpu();
        std::string report, err;
        if (gpu->self_test(&report, &err)) {
            std::printf("Self-test passed: %s\n", report.c_str());
            return 0;
        }
        std::fprintf(stderr, "Self-test FAILED: %s\n", err.c_str());
        return 2;
    }

    if (words.size() != 1) {
        std::fprintf(stderr, "%s\n", kUsage);
        return 1;
    }
    lxv::Prefix prefix;
    std::string perr;
    if (!lxv::Prefix::parse(words[0], &prefix, &perr)) {
        std::fprintf(stderr, "%s\n", perr.c_str());
        return 1;
    }
    const std::string path = lower(words[0]) + ".identity";
    struct stat sb;
    if (::stat(path.c_str(), &sb) == 0) {
        std::fprintf(stderr, "%s already exists; not overwriting it.\n", path.c_str());
        return 1;
    }
    std::printf("Looking for an LXMF address starting %s.\n", lower(words[0]).c_str());
    std::fflush(stdout);

    const size_t cores = available_parallelism();
    // With the GPU, one core is left to drive it.
    const size_t cpu_threads = !use_gpu ? cores : (with_cpu ? std::max<size_t>(cores - 1, 1) : 0);
    const auto start = Clock::now();
    uint64_t tried = 0;
    const Identity identity = run_search(prefix, use_gpu, cpu_threads, &tried);

    // create_new, mode 0600: the file holds a private key (src/main.rs:484-491)
    const int fd = ::open(path.c_str(), O_WRONLY | O_CREAT | O_EXCL | O_CLOEXEC, 0600);
    if (fd < 0) {
        const int e = errno;
        std::fprintf(stderr, "could not create %s: %s (os error %d)\n", path.c_str(), std::strerror(e), e);
        return 1;
    }
    size_t off = 0;
    while (off < identity.size()) {
        const ssize_t w = ::write(fd, identity.data() + off, identity.size() - off);
        if (w <= 0) {
            std::fprintf(stderr, "could not write the identity\n");
            ::close(fd);
            return 101;
        }
        off += (size_t)w;
    }
    ::close(fd);

    const double elapsed = seconds_since(start);
    uint8_t addr[16];
    lxv::address_of_identity(identity.data(), addr);
    std::printf("Found %s\n", lxv::hex(addr, 16).c_str());
    std::printf("About %llu tries in %.1f s (%.0f/s). Saved to %s\n", (unsigned long long)tried, elapsed,
                (double)tried / elapsed, path.c_str());
    return 0;
}
