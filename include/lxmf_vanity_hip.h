/*
 * lxmf_vanity_hip.h: C ABI of the MI355X (gfx950) backend for the LXMF vanity
 * address search.
 *
 * This is the drop-in boundary for the reference's GPU search path. The
 * reference drives its kernels through the Rust trait `Backend`
 * (src/gpu/mod.rs:40-52), implemented for CUDA in src/gpu/cuda.rs:71-133,
 * with the kernel ABI of src/gpu/kernel.cu:21-36. The entry points below map
 * onto it one to one (each cites the interface it replaces); the Rust
 * binding for them is lxmf-vanity_amd/rust/hip.rs.
 *
 * Conventions (as the reference's backend, made explicit for a C ABI):
 *  - Every call returns LXV_OK (0) or a negative LXV_E_* status; nothing
 *    unwinds across the boundary. The message of the last failure on a
 *    context is lxv_last_error(ctx).
 *  - The caller owns every host buffer it passes; the library owns all
 *    device memory and frees it in lxv_close.
 *  - A context is bound to the thread that uses it, as the reference's
 *    backend lives on one OS thread (src/main.rs:387). Calls on one context
 *    must not overlap.
 *  - Key numbering: lane t walks X25519 private keys start_t + 8 n for
 *    n = 0, 1, 2, ...; a hit is reported as (t, n), exactly the pair
 *    Backend::launch returns (src/gpu/common.h:349-355).
 *  - All words are 32-bit: u-coordinates and q as little-endian words of the
 *    32-byte encoding (src/gpu/mod.rs:69-71); the Ed25519 key, the name hash,
 *    want/mask and addresses as big-endian words (src/gpu/mod.rs:74-84,
 *    104-117).
 */
#ifndef LXMF_VANITY_HIP_H
#define LXMF_VANITY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LXV_OK 0
#define LXV_E_ARG -1        /* an argument is out of range */
#define LXV_E_HIP -2        /* the HIP runtime failed (no GPU, no memory, ...) */
#define LXV_E_STATE -3      /* the call is out of order (e.g. launch before init) */
#define LXV_E_NOMEM -4      /* a host allocation failed */
#define LXV_E_BUSY -5       /* a launch is still running (poll with wait = 0) */

/* Field-element layouts the walk kernels are built for. */
#define LXV_LAYOUT_8X32 0   /* 8 x 32-bit limbs, reduction modulo 2^256 - 38 */
#define LXV_LAYOUT_10X26 1  /* ref10 10-limb signed radix 2^25.5 */

/* How a lane computes the public keys of its batch of B consecutive keys
 * (the keys and their numbering are the same either way). */
#define LXV_WALK_PAIRS 0       /* default: around an affine centre C, with C + iQ and
                                  C - iQ sharing one inverse (Q = 8B): B/2 + 1
                                  denominators per inversion, ~2.5 M + 1 S per key */
#define LXV_WALK_PROJECTIVE 1  /* the reference's form: B chained x-only differential
                                  additions in projective (X:Z), ~8 M + 2 S per key */

typedef struct lxv_ctx lxv_ctx;

/* Options for lxv_open. Zero-initialise, set `size`, override what you need.
 * A shorter struct from an older header (ending at `layout`) is accepted: the
 * fields it lacks take their zero defaults. */
typedef struct lxv_open_opts {
    uint32_t size;         /* sizeof(lxv_open_opts) */
    int32_t device;        /* HIP device ordinal; negative means 0 */
    uint32_t threads;      /* walk lanes; 0 = CUs x waves_per_cu x 64 x R: R = 64 for the
                              pair walk, whose resident waves take 64-lane groups from a
                              queue; R = 8 for the projective walk (rounds of waves for
                              the dispatcher); R = 1 for the sub-walk form */
    uint32_t batch;        /* keys per batched inversion (B); 0 = the tuned pick (LXV_TUNED_SINGLE) */
    uint32_t waves_per_cu; /* resident waves per CU the grid is sized for: 4, 8 or 16 (1, 2 or 4
                              per SIMD); 0 = the tuned pick (LXV_TUNED_SINGLE) */
    uint32_t layout;       /* LXV_LAYOUT_*; 0 (8x32) is the default */
    uint32_t subwalks;     /* 0: one walk per lane, B keys per inversion (the reference's
                              form). K = 64, 128 or 256 (8x32 only): K walks per lane in
                              affine coordinates, inverted across the K, which saves two
                              of the eight multiplications per key; then a thread is one
                              of those walks (lxv_threads = lanes x K), lxv_batch is 1 (one
                              key per thread per iteration) and `threads` above counts
                              lanes */
    uint32_t walk;         /* LXV_WALK_*, for subwalks = 0 */
    uint32_t resident_lanes; /* pair walk: lanes of the persistent grid, rounded up to 64;
                              0 = what the device holds at once. With as many or more
                              than `threads`, one thread walks each lane instead */
} lxv_open_opts;

/* What one launch needs besides the per-lane state: LaunchArgs,
 * src/gpu/mod.rs:28-38, with the same fields and meanings. */
typedef struct lxv_launch_args {
    uint32_t iters;        /* batches of B keys per lane in this launch */
    uint32_t ed[8];        /* Ed25519 public key, big-endian words */
    uint32_t name_hash[3]; /* name hash: 10 bytes, last word holds 2 in its top half */
    uint32_t want[4];      /* (address & mask) == want selects a hit */
    uint32_t mask[4];
    uint32_t q[8];         /* u(8B), little-endian words */
    uint32_t max_hits;     /* hits recorded per launch (more are counted, not kept) */
    uint32_t dump_threads; /* lanes that write their first batch's u to the dump */
} lxv_launch_args;

/* One hit: Backend::launch's (thread, n) pair plus the address the GPU
 * computed for it, so the caller can compare without recomputing. */
typedef struct lxv_hit {
    uint32_t thread;
    uint32_t reserved;
    uint64_t n;
    uint32_t address[4];   /* big-endian words of the 16-byte LXMF address */
} lxv_hit;

/* Timing of the last finished launch, measured with HIP events on the
 * stream the walk kernel runs on. */
typedef struct lxv_launch_stats {
    double kernel_ms;      /* walk kernel time, launch to completion */
    uint64_t keys;         /* keys walked by that launch */
    uint32_t hits_total;   /* matches found (may exceed max_hits) */
    uint32_t stopped;      /* 1 if lxv_stop ended it before iters batches */
} lxv_launch_stats;

/* Opens GPU `opts->device`, builds nothing at run time (kernels are compiled
 * ahead of time for gfx950) and allocates the per-lane walk state.
 * Replaces Cuda::new (src/gpu/cuda.rs:33-64) behind open_backend
 * (src/gpu/mod.rs:60-67). */
int lxv_open(const lxv_open_opts* opts, lxv_ctx** out);

/* The GPU, for the start-up line. Backend::describe (src/gpu/mod.rs:43). */
const char* lxv_describe(lxv_ctx* ctx);

/* Backend::threads (src/gpu/mod.rs:44). */
uint32_t lxv_threads(lxv_ctx* ctx);

/* Keys per batched inversion. Backend::batch (src/gpu/mod.rs:46). */
uint64_t lxv_batch(lxv_ctx* ctx);

/* Sets up every lane from u(k-8) and u(k): 16 little-endian words per lane,
 * u(k-8) first; n_words must be 16 x lxv_threads. Resets n and the
 * checksum. Backend::init (src/gpu/mod.rs:48) and kernel init_walk
 * (src/gpu/kernel.cu:21-24). */
int lxv_init(lxv_ctx* ctx, const uint32_t* u_prev_cur, size_t n_words);

/* Sets up every lane on the GPU itself: lane t starts at private key
 * start + 8 (t x lane_stride), its u(k-8) and u(k) computed by an X25519
 * ladder on the device. `start` must be a clamped scalar that
 * random_start_scalar could return (src/main.rs:157-170) and
 * lxv_threads x lane_stride must stay below 2^63. Replaces the host loop
 * first_public_keys (src/gpu/mod.rs:210-228) for this layout of starts. */
int lxv_init_seeded(lxv_ctx* ctx, const uint8_t start[32], uint64_t lane_stride);

/* Enqueues one run of the walk kernel and returns without waiting: the
 * first half of Backend::launch (src/gpu/mod.rs:51, src/gpu/cuda.rs:93-123).
 * Fails with LXV_E_BUSY if the previous launch has not been polled. */
int lxv_launch(lxv_ctx* ctx, const lxv_launch_args* args);

/* Collects the result of the last launch: the second half of
 * Backend::launch (src/gpu/cuda.rs:124-131). With wait = 0 it returns
 * LXV_E_BUSY while the kernel runs; with wait != 0 it blocks. Fills up to
 * `cap` hits (order unspecified, at most max_hits exist) and sets *n_hits;
 * `dump` (dump_words >= dump_threads x B x 8, or NULL to skip it) receives
 * the first batch's u words of the dumped lanes, lane-major as the
 * reference's dump (src/gpu/common.h:340-343). A dump buffer that is too
 * small fails with LXV_E_ARG before anything is collected: the launch stays
 * in flight and the call can be repeated. */
int lxv_poll(lxv_ctx* ctx, int wait, lxv_hit* hits, uint32_t cap, uint32_t* n_hits,
             uint32_t* dump, size_t dump_words);

/* No counterpart in the reference. The launch of lxv_launch for callers that
 * cannot allocate device memory: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc. It allocates nothing, copies
 * nothing and does not wait: it enqueues two clearing kernels and the walk
 * kernel on `stream` (a hipStream_t; NULL = the context's own stream), so it
 * may be captured into a graph. The walk is the context's own (same kernel,
 * grid and lane state, updated in place as by lxv_launch); the caller passes
 * the outputs in device memory: d_hits (max_hits records, may be NULL when
 * max_hits is 0) and d_hit_count (one word, set to the matches found, which
 * may exceed max_hits). Its scratch (the launch counters and the batch
 * arrays of the grid) is a 32 MiB static device array per GPU, held from the
 * first lxv_open on that GPU until the process ends; contexts whose grid
 * needs more (resident lanes x batch / 2 x 32 bytes for 8 x 32 limbs, e.g.
 * more than 32768 resident lanes at batch 64) are refused with LXV_E_ARG, as
 * are the projective and sub-walk forms and dump_threads > 0. All calls on
 * one GPU share that scratch: calls that overlap on one GPU (two streams,
 * or a launch of this kind still running) race. No lxv_poll follows: the
 * caller waits on its stream and reads its outputs itself; the context's
 * lxv_last_stats are not updated. lxv_stop ends it as it ends lxv_launch;
 * the stop flag is cleared when this function is called (so not again when
 * a captured graph of it is replayed). */
int lxv_launch_noalloc(lxv_ctx* ctx, const lxv_launch_args* args, lxv_hit* d_hits, uint32_t* d_hit_count,
                       void* stream);

/* Asks a running launch to end early: every lane finishes the batch it is
 * in, stores its state and exits, so the walk can resume with the next
 * launch. No counterpart in the reference's trait; it is the "stop" of the
 * host seam and what lets a long launch give way to a found hit. It only
 * sets a flag in host memory, so unlike every other call it may come from
 * another thread while the context's own thread waits in lxv_poll. */
int lxv_stop(lxv_ctx* ctx);

/* XOR of the LXMF addresses of every key walked since the last init, as 4
 * big-endian words (a parity checksum; the reference has no equivalent). */
int lxv_checksum(lxv_ctx* ctx, uint32_t xor_words[4]);

/* Per-lane key counters n, for lanes [first, first + count). */
int lxv_read_ns(lxv_ctx* ctx, uint32_t first, uint32_t count, uint64_t* ns);

/* Timing of the last polled launch. */
int lxv_last_stats(lxv_ctx* ctx, lxv_launch_stats* out);

/* Releases the context and all its device memory. Drop of Cuda. */
void lxv_close(lxv_ctx* ctx);

/* The message of the last failure on ctx (or of the last failed lxv_open
 * when ctx is NULL); never NULL. The String side of Result<_, String>. */
const char* lxv_last_error(lxv_ctx* ctx);

/* The autotuner's launch shapes, compiled into the library from its table
 * (tuned.json): what lxv_open uses for batch = 0 and waves_per_cu = 0
 * (LXV_TUNED_SINGLE), what lxv_multi_open uses for batch = 0 and
 * waves_per_cu = 0 (LXV_TUNED_BATCHED), and the shape a serving loop should
 * open lxv_multi with (LXV_TUNED_SERVING). No counterpart in the reference,
 * whose launch shape is fixed by hand (src/gpu/cuda.rs:12-16). */
#define LXV_TUNED_SINGLE 0   /* one search, the plain kernel */
#define LXV_TUNED_BATCHED 1  /* many searches submitted at once */
#define LXV_TUNED_SERVING 2  /* a stream of requests arriving over time */

typedef struct lxv_tuned_shape {
    uint32_t batch;          /* keys per inversion */
    uint32_t waves_per_cu;   /* resident waves per CU */
    uint32_t layout;         /* LXV_LAYOUT_* */
    uint32_t walk;           /* LXV_WALK_* */
    uint32_t step_ms;        /* batched and serving: the step budget (lxv_multi_step); 0 for a single search */
    uint32_t lane_keys_log2; /* batched and serving: lxv_multi_opts.lane_keys_log2; 0 for a single search */
    double keys_per_s;       /* what the tuner measured at this shape (0: not measured) */
} lxv_tuned_shape;

/* The compiled-in shape for `workload` (LXV_TUNED_*): LXV_OK or LXV_E_ARG. */
int lxv_tuned(int workload, lxv_tuned_shape* out);

/* ------------------------------------------------------------------------
 * Host helpers of the search driver (the parts of GpuSearch and main.rs the
 * backend's caller needs), so a caller of this C ABI needs no second crypto
 * library. All re-derive from scratch on the CPU, independently of the GPU.
 */

/* X25519(k, 9) (RFC 7748); k is clamped. x25519-dalek as main.rs:194. */
void lxv_x25519_base(const uint8_t k[32], uint8_t out[32]);
/* Ed25519 public key of a 32-byte seed (RFC 8032); main.rs:195. */
void lxv_ed25519_public(const uint8_t seed[32], uint8_t out[32]);
/* The LXMF address of a 64-byte identity file (X25519 key || Ed25519 seed),
 * computed the plain way: address_of_identity, src/main.rs:190-197. */
void lxv_address_of_identity(const uint8_t identity[64], uint8_t out[16]);
/* k + 8n as a little-endian 256-bit integer: step_scalar, src/main.rs:124-137. */
void lxv_step_scalar(const uint8_t k[32], uint64_t n, uint8_t out[32]);
/* 1 if k could come from random_start_scalar (src/main.rs:157-170). */
int lxv_start_scalar_ok(const uint8_t k[32]);
/* A random start scalar from the OS (random_start_scalar). Returns 0 or
 * LXV_E_ARG if no system randomness is available. */
int lxv_random_start_scalar(uint8_t out[32]);
/* Parses 1-32 hex digits into want/mask words: Prefix::parse
 * (src/main.rs:206-217) and prefix_words (src/gpu/mod.rs:74-84). Returns
 * the number of digits, or LXV_E_ARG. */
int lxv_prefix_words(const char* prefix, uint32_t want[4], uint32_t mask[4]);
/* Fills the search-invariant words of a launch: ed (from ed_pub), the name
 * hash of "lxmf.delivery" and q = u(8B), as GpuSearch::open does
 * (src/gpu/mod.rs:102-117). */
void lxv_launch_args_for(const uint8_t ed_pub[32], lxv_launch_args* args);
/* The seeded workload generator used by the benchmarks and tests: a
 * SHA-256 counter DRBG over (seed, label). Writes a start scalar that
 * passes the rejection of src/main.rs:164-166, an Ed25519 seed, and the
 * first `nibbles` hex digits of a prefix (prefix needs nibbles + 1 bytes). */
void lxv_seeded_search(uint64_t seed, int nibbles, uint8_t start[32], uint8_t ed_seed[32], char* prefix);

/* ------------------------------------------------------------------------
 * Batched searches: many independent searches (each its own start scalar,
 * Ed25519 key and prefix) in one persistent kernel. No counterpart in the
 * reference (one search per process); the serving loop is built on these.
 *
 * A search owns V virtual lanes (64 to 262144, by prefix length: about
 * 16^digits / 2^lane_keys_log2, by default 2^14 expected keys per lane)
 * whose walk state stays resident in an HBM pool; virtual
 * lane v walks keys start + 8 (v x 2^40 + n) and all of them advance
 * together in blocks of B keys. Each step relaunches the kernel (a captured
 * hipGraph) over a list of work units (search, 64 virtual lanes, blocks);
 * physical waves take units from a device queue, so lanes of searches that
 * are done move on to unfinished ones. A search's answer is its first hit in
 * walk order: the lowest block with a hit, then the lowest (n, v) in it. Every
 * block up to that one is walked completely before the answer is reported, so
 * it does not depend on what else runs. Queued searches are admitted at
 * the start of a step, by (priority, submission order), as slots and pool
 * memory allow; one that does not fit yet does not hold back smaller ones
 * queued after it.
 */

typedef struct lxv_multi lxv_multi;

typedef struct lxv_multi_opts {
    uint32_t size;          /* sizeof(lxv_multi_opts); a struct ending at lane_keys_log2
                               (an older header) is accepted, the later fields then
                               default */
    int32_t device;         /* HIP device ordinal; negative means 0 */
    uint32_t max_searches;  /* slots resident at once; 0 = 1024 */
    uint32_t batch;         /* keys per inversion: pair walk 64, 128 or 256 (8x32),
                               256 (10x25.5); projective 64 or 256 (8x32), 256
                               (10x25.5); 0 = the tuned pick (LXV_TUNED_BATCHED) */
    uint32_t layout;        /* LXV_LAYOUT_* */
    uint32_t no_graph;      /* 1: launch eagerly instead of replaying a hipGraph */
    uint64_t pool_bytes;    /* device bytes for resident walk state; 0 = 2 GiB */
    uint32_t max_vlanes;    /* cap on a search's virtual lanes; 0 = 262144 */
    uint32_t lane_keys_log2;/* expected keys per virtual lane, log2 (8..20); 0 = the
                               tuned pick's (LXV_TUNED_BATCHED), else 14.
                               Smaller: more lanes per search, so a short search
                               finishes in fewer sequential blocks, at the cost of
                               one X25519 ladder per extra lane */
    uint32_t walk;          /* LXV_WALK_*: PAIRS (default; 8x32, batch 64 or 256) keeps
                               2 field elements of state per virtual lane,
                               PROJECTIVE 4 */
    uint32_t pipelined;     /* 1: lxv_multi_step enqueues the next step before it
                               waits for the one before, so the host's planning and
                               read-back overlap the walk; a step's answers then come
                               back from the following call (at the same time on the
                               clock: the call waits for that step). 0 (default):
                               each call runs one step to completion */
    uint32_t waves_per_cu;  /* resident waves per CU the kernel is built for: 8 or 16
                               (the pair walk; the projective walk only 16); 0 = the
                               tuned pick (LXV_TUNED_BATCHED) */
    uint32_t oldest_share_pct; /* 0..90: share of each step's work given to the
                               longest-waiting resident search before the rest is
                               shared by priority, and that search is admitted first;
                               bounds how long a low-priority search can starve.
                               0 (default): strict priority */
    uint32_t preempt;       /* 1: a queued search that finds no free slot or pool range
                               suspends resident searches of a lower priority (larger
                               value) until it fits; a suspended search keeps the
                               progress collected so far (a partly walked block is
                               walked again) and is set up again at that block when it
                               is readmitted, so its answer is unchanged. Searches
                               walked by a pipelined step still in flight are not
                               suspended. 0 (default): admission only into free room */
} lxv_multi_opts;

typedef struct lxv_search_spec {
    uint8_t start[32];      /* first private key of the search (as random_start_scalar) */
    uint8_t ed_pub[32];     /* Ed25519 public key of the identity */
    char prefix[33];        /* 1-32 hex digits, NUL-terminated */
    uint8_t reserved[3];
    uint32_t priority;      /* admission order: lower first, then submission order */
    uint64_t tag;           /* caller's cookie, returned with the result */
    uint64_t max_keys;      /* give up after about this many keys; 0 = no limit */
} lxv_search_spec;

typedef struct lxv_search_result {
    uint64_t id;            /* from lxv_multi_submit */
    uint64_t tag;
    uint64_t g;             /* step index of the answer: key = start + 8 g */
    uint32_t address[4];    /* its LXMF address, big-endian words */
    uint64_t keys;          /* keys this search walked to reach its answer */
    uint32_t vlanes;        /* its virtual lanes (V) */
    int32_t status;         /* 1 found, 0 gave up (max_keys), -1 cancelled */
    double submit_s;        /* seconds since lxv_multi_open at submission */
    double admit_s;         /* ... when it got its slot and lanes */
    double done_s;          /* ... when the answer was known on the host */
} lxv_search_result;

typedef struct lxv_multi_stat {
    uint64_t keys;          /* keys walked so far, all searches */
    double busy_s;          /* GPU time in the walk kernel (HIP events) */
    double last_step_ms;    /* walk kernel time of the last step */
    uint32_t resident;      /* searches holding a slot */
    uint32_t queued;        /* searches waiting for a slot */
    uint32_t done;          /* searches finished (found, gave up or cancelled) */
    uint32_t steps;         /* relaunches so far */
} lxv_multi_stat;

int lxv_multi_open(const lxv_multi_opts* opts, lxv_multi** out);
/* Queues a search; returns its id. */
int lxv_multi_submit(lxv_multi* m, const lxv_search_spec* spec, uint64_t* id);
/* Admits what fits, relaunches the kernel once over about budget_ms of work,
 * and returns up to cap finished searches (more are kept for the next call).
 * budget_ms = 0 launches nothing; with `pipelined` it collects the step still
 * in flight, so a caller can drain before it stops stepping. */
int lxv_multi_step(lxv_multi* m, uint32_t budget_ms, lxv_search_result* out, uint32_t cap, uint32_t* n_out);
int lxv_multi_cancel(lxv_multi* m, uint64_t id);
int lxv_multi_stats(lxv_multi* m, lxv_multi_stat* out);
void lxv_multi_close(lxv_multi* m);
const char* lxv_multi_last_error(lxv_multi* m);

#ifdef __cplusplus
}
#endif
#endif
