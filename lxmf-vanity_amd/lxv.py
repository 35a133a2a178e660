"""Python binding of the MI355X backend (liblxmf_vanity_hip.so) and the host
side of the GPU search, mirroring the reference's src/gpu/mod.rs.

The C ABI (include/lxmf_vanity_hip.h) maps onto the reference's `Backend`
trait; `GpuSearch` below re-creates `GpuSearch::{open, run, self_test}`
(src/gpu/mod.rs:95-207) over it, with the same meaning of every argument and
the same errors:

* open(): a context on GPU 0, every lane set up from a start scalar, the
  Ed25519 key of the identity, q = u(8B) and the name hash words.
* run(prefix): launches until a hit, doubling `iters` while launches take
  less than half the target (src/gpu/mod.rs:148-174); the first hit becomes an
  identity that is re-derived from scratch before it is returned.
* self_test(): one launch with prefix "b0" and a dump of 256 lanes' first
  batch, checked against plain X25519 (src/gpu/mod.rs:176-207).

Nothing here computes keys: the product path is the HIP library, and it
fails loudly (`GpuUnavailable`) if the library or a GPU is missing.
"""

import ctypes
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblxmf_vanity_hip.so")

LXV_OK = 0
LXV_E_BUSY = -5
LAYOUT_8X32 = 0
LAYOUT_10X26 = 1
FE_RAW_LIMBS = 0x100            # lxv_debug_fe: the layout's raw limbs in and out
WALK_PAIRS = 0                  # batch keys around an affine centre (the default)
WALK_PROJECTIVE = 1             # the reference's chained x-only differential additions
MAX_HITS = 4096                 # hits recorded per launch (src/gpu/mod.rs:23)
TARGET_LAUNCH_S = 0.5           # src/gpu/mod.rs:26


class GpuUnavailable(RuntimeError):
    pass


class OpenOpts(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("device", ctypes.c_int32), ("threads", ctypes.c_uint32),
                ("batch", ctypes.c_uint32), ("waves_per_cu", ctypes.c_uint32), ("layout", ctypes.c_uint32),
                ("subwalks", ctypes.c_uint32), ("walk", ctypes.c_uint32), ("resident_lanes", ctypes.c_uint32)]


class LaunchArgs(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_uint32), ("ed", ctypes.c_uint32 * 8), ("name_hash", ctypes.c_uint32 * 3),
                ("want", ctypes.c_uint32 * 4), ("mask", ctypes.c_uint32 * 4), ("q", ctypes.c_uint32 * 8),
                ("max_hits", ctypes.c_uint32), ("dump_threads", ctypes.c_uint32)]


class Hit(ctypes.Structure):
    _fields_ = [("thread", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("n", ctypes.c_uint64),
                ("address", ctypes.c_uint32 * 4)]


class LaunchStats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("keys", ctypes.c_uint64), ("hits_total", ctypes.c_uint32),
                ("stopped", ctypes.c_uint32)]


class MultiOpts(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("device", ctypes.c_int32), ("max_searches", ctypes.c_uint32),
                ("batch", ctypes.c_uint32), ("layout", ctypes.c_uint32), ("no_graph", ctypes.c_uint32),
                ("pool_bytes", ctypes.c_uint64), ("max_vlanes", ctypes.c_uint32), ("lane_keys_log2", ctypes.c_uint32),
                ("walk", ctypes.c_uint32), ("pipelined", ctypes.c_uint32), ("waves_per_cu", ctypes.c_uint32),
                ("oldest_share_pct", ctypes.c_uint32), ("preempt", ctypes.c_uint32)]


class TunedShape(ctypes.Structure):
    _fields_ = [("batch", ctypes.c_uint32), ("waves_per_cu", ctypes.c_uint32), ("layout", ctypes.c_uint32),
                ("walk", ctypes.c_uint32), ("step_ms", ctypes.c_uint32), ("lane_keys_log2", ctypes.c_uint32),
                ("keys_per_s", ctypes.c_double)]


TUNED_WORKLOADS = {"single": 0, "batched": 1, "serving": 2}


class SearchSpec(ctypes.Structure):
    _fields_ = [("start", ctypes.c_uint8 * 32), ("ed_pub", ctypes.c_uint8 * 32), ("prefix", ctypes.c_char * 33),
                ("reserved", ctypes.c_uint8 * 3), ("priority", ctypes.c_uint32), ("tag", ctypes.c_uint64),
                ("max_keys", ctypes.c_uint64)]


class SearchResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("tag", ctypes.c_uint64), ("g", ctypes.c_uint64),
                ("address", ctypes.c_uint32 * 4), ("keys", ctypes.c_uint64), ("vlanes", ctypes.c_uint32),
                ("status", ctypes.c_int32), ("submit_s", ctypes.c_double), ("admit_s", ctypes.c_double),
                ("done_s", ctypes.c_double)]


class MultiStat(ctypes.Structure):
    _fields_ = [("keys", ctypes.c_uint64), ("busy_s", ctypes.c_double), ("last_step_ms", ctypes.c_double),
                ("resident", ctypes.c_uint32), ("queued", ctypes.c_uint32), ("done", ctypes.c_uint32),
                ("steps", ctypes.c_uint32)]


_lib = None


def lib():
    """Loads the backend library, or raises GpuUnavailable."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GpuUnavailable("%s is not built: run __graft_entry__.build() or `make -C lxmf-vanity_amd`" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, u8p, u32p = ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint32)
    L.lxv_open.argtypes = [ctypes.POINTER(OpenOpts), ctypes.POINTER(vp)]
    L.lxv_describe.argtypes = [vp]
    L.lxv_describe.restype = ctypes.c_char_p
    L.lxv_threads.argtypes = [vp]
    L.lxv_threads.restype = ctypes.c_uint32
    L.lxv_batch.argtypes = [vp]
    L.lxv_batch.restype = ctypes.c_uint64
    L.lxv_init.argtypes = [vp, u32p, ctypes.c_size_t]
    L.lxv_init_seeded.argtypes = [vp, u8p, ctypes.c_uint64]
    L.lxv_launch.argtypes = [vp, ctypes.POINTER(LaunchArgs)]
    L.lxv_poll.argtypes = [vp, ctypes.c_int, ctypes.POINTER(Hit), ctypes.c_uint32, u32p, u32p, ctypes.c_size_t]
    L.lxv_stop.argtypes = [vp]
    L.lxv_checksum.argtypes = [vp, u32p]
    L.lxv_read_ns.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)]
    L.lxv_last_stats.argtypes = [vp, ctypes.POINTER(LaunchStats)]
    L.lxv_close.argtypes = [vp]
    L.lxv_close.restype = None
    L.lxv_last_error.argtypes = [vp]
    L.lxv_last_error.restype = ctypes.c_char_p
    for name in ("lxv_x25519_base", "lxv_ed25519_public", "lxv_address_of_identity"):
        getattr(L, name).argtypes = [u8p, vp]
        getattr(L, name).restype = None
    L.lxv_step_scalar.argtypes = [u8p, ctypes.c_uint64, vp]
    L.lxv_step_scalar.restype = None
    L.lxv_start_scalar_ok.argtypes = [u8p]
    L.lxv_random_start_scalar.argtypes = [vp]
    L.lxv_prefix_words.argtypes = [ctypes.c_char_p, u32p, u32p]
    L.lxv_launch_args_for.argtypes = [u8p, ctypes.POINTER(LaunchArgs)]
    L.lxv_launch_args_for.restype = None
    L.lxv_seeded_search.argtypes = [ctypes.c_uint64, ctypes.c_int, vp, vp, vp]
    L.lxv_seeded_search.restype = None
    L.lxv_multi_open.argtypes = [ctypes.POINTER(MultiOpts), ctypes.POINTER(vp)]
    L.lxv_multi_submit.argtypes = [vp, ctypes.POINTER(SearchSpec), ctypes.POINTER(ctypes.c_uint64)]
    L.lxv_multi_step.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(SearchResult), ctypes.c_uint32, u32p]
    L.lxv_multi_cancel.argtypes = [vp, ctypes.c_uint64]
    L.lxv_multi_stats.argtypes = [vp, ctypes.POINTER(MultiStat)]
    L.lxv_multi_close.argtypes = [vp]
    L.lxv_multi_close.restype = None
    L.lxv_multi_last_error.argtypes = [vp]
    L.lxv_multi_last_error.restype = ctypes.c_char_p
    L.lxv_debug_fe.argtypes = [ctypes.c_int, ctypes.c_int, u32p, u32p, u32p, ctypes.c_uint32, ctypes.c_int]
    L.lxv_debug_address.argtypes = [u32p, u32p, u32p, u32p, ctypes.c_uint32, ctypes.c_int]
    L.lxv_debug_pair_walk.argtypes = [ctypes.c_int, ctypes.c_uint32, u8p, u32p, ctypes.c_uint32, u32p]
    L.lxv_debug_walk_kernel.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t, u32p]
    L.lxv_debug_pair_centre.argtypes = [ctypes.c_int, ctypes.c_uint32, u8p, ctypes.c_uint64, ctypes.c_uint64,
                                        ctypes.c_uint32, u32p]
    _lib = L
    return L


# ---- host helpers (CPU, independent of the GPU)

def _buf(n):
    return ctypes.create_string_buffer(n)


def x25519_base(k):
    o = _buf(32); lib().lxv_x25519_base(k, o); return o.raw


def ed25519_public(seed):
    o = _buf(32); lib().lxv_ed25519_public(seed, o); return o.raw


def address_of_identity(identity):
    o = _buf(16); lib().lxv_address_of_identity(identity, o); return o.raw


def step_scalar(k, n):
    o = _buf(32); lib().lxv_step_scalar(k, n, o); return o.raw


def start_scalar_ok(k):
    return bool(lib().lxv_start_scalar_ok(k))


def random_start_scalar():
    o = _buf(32)
    if lib().lxv_random_start_scalar(o) != LXV_OK:
        raise RuntimeError("no system randomness")
    return o.raw


def prefix_words(prefix):
    want = (ctypes.c_uint32 * 4)(); mask = (ctypes.c_uint32 * 4)()
    n = lib().lxv_prefix_words(prefix.encode(), want, mask)
    if n < 0:
        raise ValueError("'%s' isn't 1 to 32 hex digits: use 0-9 and a-f only (zero, not the letter O)."
                         % prefix.lower())
    return list(want), list(mask)


def prefix_matches(prefix, address):
    return address.hex().startswith(prefix.lower())


def seeded_search(seed, nibbles=0):
    """(start scalar, Ed25519 seed, prefix) of workload `seed`."""
    s, e, p = _buf(32), _buf(32), _buf(max(nibbles, 0) + 1)
    lib().lxv_seeded_search(seed, nibbles, s, e, p)
    return s.raw, e.raw, p.value.decode()


def launch_args_for(ed_pub, prefix=None, iters=1, max_hits=MAX_HITS, dump_threads=0):
    a = LaunchArgs()
    lib().lxv_launch_args_for(ed_pub, ctypes.byref(a))
    if prefix:
        want, mask = prefix_words(prefix)
        a.want[:] = want
        a.mask[:] = mask
    a.iters, a.max_hits, a.dump_threads = iters, max_hits, dump_threads
    return a


def words_to_bytes_be(words):
    return b"".join(int(w).to_bytes(4, "big") for w in words)


def words_to_bytes_le(words):
    return b"".join(int(w).to_bytes(4, "little") for w in words)


# ---- the backend (Backend trait over the C ABI)

class Backend:
    """One GPU context: Backend::{describe, threads, batch, init, launch}."""

    def __init__(self, threads=0, batch=64, waves_per_cu=0, layout=LAYOUT_8X32, device=0, subwalks=0,
                 walk=WALK_PAIRS, resident_lanes=0):
        """subwalks=K: K affine sub-walks per lane (threads then counts lanes
        and self.threads is lanes x K; batch is 1 key per thread per launch
        iteration). walk: WALK_PAIRS or WALK_PROJECTIVE for one walk per lane.
        resident_lanes: the pair walk's persistent grid (0: the device's
        occupancy; the grid is used only when it is smaller than threads)."""
        o = OpenOpts(ctypes.sizeof(OpenOpts), device, threads, 0 if subwalks else batch, waves_per_cu, layout,
                     subwalks, walk, resident_lanes)
        h = ctypes.c_void_p()
        rc = lib().lxv_open(ctypes.byref(o), ctypes.byref(h))
        if rc != LXV_OK:
            raise GpuUnavailable(lib().lxv_last_error(None).decode())
        self.h = h
        self.threads = lib().lxv_threads(h)
        self.batch = lib().lxv_batch(h)

    def _check(self, rc, what):
        if rc != LXV_OK:
            raise RuntimeError("%s: %s" % (what, lib().lxv_last_error(self.h).decode()))

    def describe(self):
        return lib().lxv_describe(self.h).decode()

    def walk_kernel(self):
        """(name of the walk kernel a launch runs, lanes of its grid)."""
        buf = ctypes.create_string_buffer(128)
        grid = ctypes.c_uint32(0)
        self._check(lib().lxv_debug_walk_kernel(self.h, buf, len(buf), ctypes.byref(grid)), "walk_kernel")
        return buf.value.decode(), grid.value

    def init(self, u_prev_cur):
        arr = (ctypes.c_uint32 * len(u_prev_cur))(*u_prev_cur)
        self._check(lib().lxv_init(self.h, arr, len(u_prev_cur)), "init")

    def init_seeded(self, start, stride):
        self._check(lib().lxv_init_seeded(self.h, start, stride), "init_seeded")

    def launch_async(self, args):
        self._check(lib().lxv_launch(self.h, ctypes.byref(args)), "launch")

    def poll(self, wait=True, dump_threads=0, cap=MAX_HITS):
        hits = (Hit * max(cap, 1))()
        n = ctypes.c_uint32(0)
        dw = dump_threads * self.batch * 8
        dump = (ctypes.c_uint32 * max(dw, 1))()
        rc = lib().lxv_poll(self.h, 1 if wait else 0, hits, cap, ctypes.byref(n), dump if dw else None, dw)
        if rc == LXV_E_BUSY:
            return None
        self._check(rc, "poll")
        got = [(hits[i].thread, hits[i].n, words_to_bytes_be(hits[i].address)) for i in range(min(n.value, cap))]
        return got, list(dump[:dw]) if dw else []

    def launch(self, args):
        """Backend::launch: run once, return ((thread, n, address) hits, dump)."""
        self.launch_async(args)
        return self.poll(True, args.dump_threads)

    def stop(self):
        self._check(lib().lxv_stop(self.h), "stop")

    def checksum(self):
        w = (ctypes.c_uint32 * 4)()
        self._check(lib().lxv_checksum(self.h, w), "checksum")
        return words_to_bytes_be(w)

    def read_ns(self, first=0, count=None):
        count = self.threads - first if count is None else count
        arr = (ctypes.c_uint64 * count)()
        self._check(lib().lxv_read_ns(self.h, first, count, arr), "read_ns")
        return list(arr)

    def stats(self):
        s = LaunchStats()
        self._check(lib().lxv_last_stats(self.h, ctypes.byref(s)), "stats")
        return s

    def close(self):
        if getattr(self, "h", None):
            lib().lxv_close(self.h)
            self.h = None

    def __del__(self):
        self.close()


# ---- GpuSearch (src/gpu/mod.rs:86-207)

def lane_stride(lanes, stride=1 << 40):
    """The key distance between lanes: stride halved until lanes x stride
    stays below 2^63 (2^24 lanes: 2^38 keys per lane)."""
    while stride > 1 and lanes * stride >= 1 << 63:
        stride >>= 1
    return stride


class GpuSearch:
    """The GPU search over one backend. Lanes start at start + 8 (t x stride)
    (one shared Ed25519 key, as src/gpu/mod.rs:102-104)."""

    LANE_STRIDE = 1 << 40

    def __init__(self, start=None, ed_seed=None, stride=LANE_STRIDE, **backend_opts):
        self.backend = Backend(**backend_opts)
        self.start = start if start is not None else random_start_scalar()
        if not start_scalar_ok(self.start):
            raise ValueError("not a start scalar random_start_scalar could return")
        self.seed = ed_seed if ed_seed is not None else os.urandom(32)
        stride = lane_stride(self.backend.threads, stride)
        self.stride = stride
        self.backend.init_seeded(self.start, stride)
        self.ed_pub = ed25519_public(self.seed)

    def describe(self):
        return "%s: %d GPU threads" % (self.backend.describe(), self.backend.threads)

    def identity(self, t, n):
        """The identity file for lane t's key number n (src/gpu/mod.rs:139-144)."""
        return step_scalar(self.start, t * self.stride + n) + self.seed

    def launch(self, prefix, iters, dump_threads=0):
        return self.backend.launch(launch_args_for(self.ed_pub, prefix, iters, MAX_HITS, dump_threads))

    def run(self, prefix, found=None, tries=None):
        """Searches until a hit; returns (identity, tries)."""
        per_iter = self.backend.threads * self.backend.batch
        iters, total = 1, 0
        while not (found and found()):
            t0 = time.perf_counter()
            hits, _ = self.launch(prefix, iters)
            took = time.perf_counter() - t0
            total += per_iter * iters
            if tries:
                tries(per_iter * iters)
            if hits:
                t, n, _ = hits[0]
                identity = self.identity(t, n)
                if not prefix_matches(prefix, address_of_identity(identity)):
                    raise RuntimeError("GPU hit (thread %d, key %d) failed its CPU recheck: a kernel bug" % (t, n))
                return identity, total
            if took < TARGET_LAUNCH_S / 2 and iters < (1 << 16):
                iters *= 2
        return None, total

    def self_test(self):
        dump_threads = min(self.backend.threads, 256)
        batch = self.backend.batch
        hits, dumped = self.launch("b0", 1, dump_threads)
        for t in range(dump_threads):
            for i in range(batch):
                k = step_scalar(self.start, t * self.stride + i)
                base = (t * batch + i) * 8
                if words_to_bytes_le(dumped[base:base + 8]) != x25519_base(k):
                    raise RuntimeError("public key mismatch: thread %d, key %d" % (t, i))
        if len(hits) < 100:
            raise RuntimeError("only %d hits for a 2-digit prefix; expected thousands" % len(hits))
        for t, n, _ in hits:
            if not prefix_matches("b0", address_of_identity(self.identity(t, n))):
                raise RuntimeError("hit (thread %d, key %d) failed its CPU recheck" % (t, n))
        return "%d public keys match X25519, %d hits for 'b0' all recheck on the CPU." % (dump_threads * batch,
                                                                                          len(hits))

    def close(self):
        self.backend.close()


# ---- batched searches (many searches in one persistent kernel)

STRIDE_MULTI = 1 << 40   # key distance between a search's virtual lanes


class Multi:
    """lxv_multi_*: submit searches, step the device, collect answers.

    A result is a dict: id, tag, g (key = start + 8 g), address (16 bytes),
    keys, vlanes, status (1 found, 0 gave up, -1 cancelled), submit_s,
    admit_s, done_s (seconds since open). pipelined: each step() enqueues
    the next step before collecting the previous one, whose answers it
    returns. batch = 0 and waves_per_cu = 0 take the library's tuned shape
    for batched searches; oldest_share_pct reserves that share of every
    step for the longest-waiting search."""

    def __init__(self, max_searches=0, batch=0, layout=LAYOUT_8X32, no_graph=False, pool_bytes=0,
                 max_vlanes=0, lane_keys_log2=0, device=0, walk=WALK_PAIRS, pipelined=False, waves_per_cu=0,
                 oldest_share_pct=0, preempt=False):
        o = MultiOpts(ctypes.sizeof(MultiOpts), device, max_searches, batch, layout, 1 if no_graph else 0,
                      pool_bytes, max_vlanes, lane_keys_log2, walk, 1 if pipelined else 0, waves_per_cu,
                      oldest_share_pct, 1 if preempt else 0)
        h = ctypes.c_void_p()
        if lib().lxv_multi_open(ctypes.byref(o), ctypes.byref(h)) != LXV_OK:
            raise GpuUnavailable(lib().lxv_multi_last_error(None).decode())
        self.h = h

    def _check(self, rc, what):
        if rc != LXV_OK:
            raise RuntimeError("%s: %s" % (what, lib().lxv_multi_last_error(self.h).decode()))

    def submit(self, start, ed_pub, prefix, tag=0, max_keys=0, priority=0):
        s = SearchSpec()
        s.start[:] = list(start)
        s.ed_pub[:] = list(ed_pub)
        s.prefix = prefix.encode()
        s.tag, s.max_keys, s.priority = tag, max_keys, priority
        i = ctypes.c_uint64(0)
        self._check(lib().lxv_multi_submit(self.h, ctypes.byref(s), ctypes.byref(i)), "submit")
        return i.value

    def step(self, budget_ms=200, cap=4096):
        out = (SearchResult * cap)()
        n = ctypes.c_uint32(0)
        self._check(lib().lxv_multi_step(self.h, budget_ms, out, cap, ctypes.byref(n)), "step")
        return [{"id": r.id, "tag": r.tag, "g": r.g, "address": words_to_bytes_be(r.address), "keys": r.keys,
                 "vlanes": r.vlanes, "status": r.status, "submit_s": r.submit_s, "admit_s": r.admit_s,
                 "done_s": r.done_s} for r in out[:n.value]]

    def cancel(self, sid):
        self._check(lib().lxv_multi_cancel(self.h, sid), "cancel")

    def stats(self):
        s = MultiStat()
        self._check(lib().lxv_multi_stats(self.h, ctypes.byref(s)), "stats")
        return {k: getattr(s, k) for k, _ in MultiStat._fields_}

    def suspensions(self):
        """Searches suspended so far by preemption."""
        n = ctypes.c_uint64(0)
        lib().lxv_debug_multi_suspensions.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]
        self._check(lib().lxv_debug_multi_suspensions(self.h, ctypes.byref(n)), "suspensions")
        return n.value

    def run_all(self, budget_ms=200, timeout_s=600.0):
        """Steps until every submitted search is finished; returns results."""
        res, t0 = [], time.time()
        while True:
            res += self.step(budget_ms)
            st = self.stats()
            if st["resident"] == 0 and st["queued"] == 0:
                return res + self.step(budget_ms)
            if time.time() - t0 > timeout_s:
                raise TimeoutError("searches still running after %.0f s" % timeout_s)

    def close(self):
        if getattr(self, "h", None):
            lib().lxv_multi_close(self.h)
            self.h = None

    def __del__(self):
        self.close()


def debug_fe(op, a_words, b_words, on_gpu, layout=LAYOUT_8X32):
    """op(a, b) per element by the device field code (on the GPU or its host
    build). Elements are 8 words each, or with FE_RAW_LIMBS the layout's raw
    limbs (8 words for 8x32, 10 for 10x26); so is the result."""
    w = 10 if (op & FE_RAW_LIMBS) and layout == LAYOUT_10X26 else 8
    n = len(a_words) // w
    A = (ctypes.c_uint32 * max(len(a_words), 1))(*a_words)
    B = (ctypes.c_uint32 * max(len(b_words), 1))(*b_words)
    O = (ctypes.c_uint32 * max(w * n, 1))()
    rc = lib().lxv_debug_fe(layout, op, A, B, O, n, 1 if on_gpu else 0)
    if rc != LXV_OK:
        raise RuntimeError("lxv_debug_fe failed (%d)" % rc)
    return list(O[:w * n])


def tuned_shape(workload):
    """The launch shape the library was built with for `workload` ("single",
    "batched" or "serving"), from the autotuner's table: a dict of batch,
    waves_per_cu, layout, walk, step_ms, lane_keys_log2, keys_per_s."""
    L = lib()
    L.lxv_tuned.argtypes = [ctypes.c_int, ctypes.POINTER(TunedShape)]
    t = TunedShape()
    if L.lxv_tuned(TUNED_WORKLOADS[workload], ctypes.byref(t)) != LXV_OK:
        raise ValueError("no tuned shape for %r" % workload)
    return {k: getattr(t, k) for k, _ in TunedShape._fields_}


def debug_valu_rate(op, waves_per_cu=16, device=0):
    """VALU issue-rate probe `op` (LXV_VALU_* of the debug header) at
    waves_per_cu resident waves per CU: (lane-level instructions per second,
    the shader clock the probe measured in Hz, the device's nominal clock in
    Hz, CUs)."""
    L = lib()
    d = ctypes.POINTER(ctypes.c_double)
    L.lxv_debug_valu_rate.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, d, d, d,
                                      ctypes.POINTER(ctypes.c_int)]
    rate, clk, nom, cus = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_double(0), ctypes.c_int(0)
    rc = L.lxv_debug_valu_rate(device, op, waves_per_cu, ctypes.byref(rate), ctypes.byref(clk), ctypes.byref(nom),
                               ctypes.byref(cus))
    if rc != LXV_OK:
        raise RuntimeError("lxv_debug_valu_rate failed (%d)" % rc)
    return rate.value, clk.value, nom.value, cus.value


VALU_PROBES = {0: "v_alignbit_b32", 1: "v_bitop3_b32", 2: "v_add3_u32", 3: "v_add_u32", 4: "v_xor_b32",
               5: "v_perm_b32", 6: "v_lshrrev_b32", 7: "v_mad_u64_u32",
               8: "SHA-256 round mix (3 alignbit : 2 bitop3 : 2 add3)", 9: "v_add_co_u32",
               10: "v_addc_co_u32 dependent carry chain",
               11: "pair-walk mix (30% alignbit, 19% bitop3, 12% add3, 9% addc, 8% mad64, 8% add, 6% mov, 3% shr)",
               12: "v_mul_lo_u32", 13: "v_lshl_add_u32", 14: "v_lshl_or_b32",
               15: "v_mad_u64_u32 dependent chain"}


def debug_address(ed_words, nh_words, u_words, on_gpu):
    n = len(u_words) // 8
    E = (ctypes.c_uint32 * 8)(*ed_words)
    N = (ctypes.c_uint32 * 3)(*nh_words)
    U = (ctypes.c_uint32 * max(len(u_words), 1))(*u_words)
    O = (ctypes.c_uint32 * max(4 * n, 1))()
    rc = lib().lxv_debug_address(E, N, U, O, n, 1 if on_gpu else 0)
    if rc != LXV_OK:
        raise RuntimeError("lxv_debug_address failed (%d)" % rc)
    return list(O[:4 * n])


def debug_pair_centre(batch, start, stride, t, bits, layout=LAYOUT_8X32):
    """Lane t's first centre of the lxv_init_seeded pair-walk set-up on the
    CPU, twice: ((x, y) by a full ladder, (x, y) from the set-up's two
    anchors), each coordinate as an integer."""
    out = (ctypes.c_uint32 * 32)()
    rc = lib().lxv_debug_pair_centre(layout, batch, start, stride, t, bits, out)
    if rc != LXV_OK:
        raise RuntimeError("lxv_debug_pair_centre failed (%d)" % rc)
    v = [sum(out[8 * k + i] << (32 * i) for i in range(8)) for k in range(4)]
    return (v[0], v[1]), (v[2], v[3])


def debug_pair_walk(batch, batches, start=None, u_prev_cur=None, layout=LAYOUT_8X32):
    """The pair walk of one lane on the CPU (the host build of the device
    code): u(k + 8n) for n < batches x batch as 32-byte encodings, from a
    start scalar (the lxv_init_seeded set-up) or from 16 words u(k-8), u(k)."""
    out = (ctypes.c_uint32 * (batches * batch * 8))()
    upc = (ctypes.c_uint32 * 16)(*u_prev_cur) if u_prev_cur is not None else None
    rc = lib().lxv_debug_pair_walk(layout, batch, start, upc, batches, out)
    if rc != LXV_OK:
        raise RuntimeError("lxv_debug_pair_walk failed (%d)" % rc)
    return [words_to_bytes_le(out[8 * i:8 * i + 8]) for i in range(batches * batch)]
