/*
 * lxmf_vanity_hip_debug.h: test entry points of the MI355X backend.
 *
 * They have no counterpart in the reference. They run the device field and
 * hash code on given inputs, on the GPU or (the same source compiled for the
 * host) on the CPU, so parity tests can check each piece of the walk against
 * big-integer arithmetic and the oracle without going through a search.
 */
#ifndef LXMF_VANITY_HIP_DEBUG_H
#define LXMF_VANITY_HIP_DEBUG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field operation codes for lxv_debug_fe. */
#define LXV_FE_MUL 0
#define LXV_FE_SQ 1
#define LXV_FE_ADD 2
#define LXV_FE_SUB 3
#define LXV_FE_INVERT 4
#define LXV_FE_CANON 5
#define LXV_FE_MUL121665 6
#define LXV_FE_INVERT_CHAIN 7  /* z^(p-2) by the reference's addition chain (LXV_FE_INVERT: divsteps) */
/* Or-ed into an op: inputs are the layout's raw limbs (8 words per element
 * for LXV_LAYOUT_8X32, any value below 2^256 with bit 255 kept; 10 words for
 * LXV_LAYOUT_10X26, within each operation's input bounds), and the output is
 * the limbs the operation produced (congruent to the result mod p, not
 * canonical), except for LXV_FE_CANON, which still returns the canonical
 * encoding (in the first 8 words of the element, the rest 0). Sums and
 * differences are returned as computed, not multiplied by one. */
#define LXV_FE_RAW_LIMBS 0x100

/* out[i] = canonical encoding of op(a[i], b[i]) in limb layout `layout`
 * (LXV_LAYOUT_*), for n elements of 8 little-endian words each (bit 255 of
 * the inputs dropped; see LXV_FE_RAW_LIMBS for each layout's own
 * representation). on_gpu selects the device kernel (device 0) or the
 * host build of the same code. Returns LXV_OK or a negative status. */
int lxv_debug_fe(int layout, int op, const uint32_t* a, const uint32_t* b, uint32_t* out, uint32_t n, int on_gpu);

/* out[i] = the LXMF address words (4 big-endian) of X25519 public key u[i]
 * (8 little-endian words) with Ed25519 key ed and name hash nh, computed by
 * the walk's specialised SHA-256 on the GPU or its host build. */
int lxv_debug_address(const uint32_t ed[8], const uint32_t nh[3], const uint32_t* u, uint32_t* out, uint32_t n,
                      int on_gpu);

/* VALU issue-rate probes for lxv_debug_valu_rate: one instruction in 8
 * independent chains per lane, or the SHA-256 round mix. */
#define LXV_VALU_ALIGNBIT 0     /* v_alignbit_b32 (rotates) */
#define LXV_VALU_BITOP3 1       /* v_bitop3_b32 (XOR3, Ch, Maj) */
#define LXV_VALU_ADD3 2         /* v_add3_u32 */
#define LXV_VALU_ADD 3          /* v_add_u32 */
#define LXV_VALU_XOR 4          /* v_xor_b32 */
#define LXV_VALU_PERM 5         /* v_perm_b32 (byte swaps) */
#define LXV_VALU_SHIFT 6        /* v_lshrrev_b32 */
#define LXV_VALU_MAD64 7        /* v_mad_u64_u32 (the peak the walk's multiply count is priced against) */
#define LXV_VALU_SHA_MIX 8      /* 3 v_alignbit : 2 v_bitop3 : 2 v_add3, as a SHA-256 round */
#define LXV_VALU_ADD_CO 9       /* v_add_co_u32 (carry out to VCC) */
#define LXV_VALU_CARRY_CHAIN 10 /* v_addc_co_u32 chains through VCC, each reading the last's carry */
#define LXV_VALU_WALK_MIX 11    /* the instruction classes of the pair walk's per-key loop, in its proportions */
#define LXV_VALU_MUL_LO 12      /* v_mul_lo_u32 */
#define LXV_VALU_LSHL_ADD 13    /* v_lshl_add_u32 (shift and add, as in x * 19 = (x << 4) + 3x) */
#define LXV_VALU_LSHL_OR 14     /* v_lshl_or_b32 (packing limbs into words) */
#define LXV_VALU_MAD64_CHAIN 15 /* v_mad_u64_u32 in one dependent chain per lane (a product column) */

/* Measures the issue rate of probe `op` on `device` with one resident round
 * of waves_per_cu waves per CU (a multiple of 4, at most 32): lane-level
 * instructions per second (wave instructions x 64). clock_hz is the shader
 * clock the probe's waves saw (s_memtime against the constant-rate
 * s_memrealtime over each block's run), nominal_clock_hz the device's
 * reported peak clock. */
int lxv_debug_valu_rate(int device, int op, uint32_t waves_per_cu, double* lane_ops_per_s, double* clock_hz,
                        double* nominal_clock_hz, int* cus);

/* Runs one plain device-to-device copy kernel of `bytes` (a multiple of 16)
 * on `device` and returns its time: the known-size reference that HBM byte
 * counters of the walk are scaled by. */
int lxv_debug_copy(int device, uint64_t bytes, double* ms);

/* Register and scratch use of the walk kernel built for (layout, batch,
 * waves_per_cu, walk form LXV_WALK_*): VGPRs per lane, scratch bytes per
 * lane, and the waves per SIMD its register budget was compiled for. Needs a
 * GPU (reads the loaded code object). */
int lxv_debug_kernel_attrs(int layout, uint32_t batch, uint32_t waves_per_cu, uint32_t walk, uint32_t* vgprs,
                           uint32_t* sgprs, uint32_t* scratch_bytes, uint32_t* waves_per_simd);

/* The walk kernel a context launches, named as the profiler names it (e.g.
 * "walk_pair_queue_kernel<lxv::Field8, 64, 4>", NUL-terminated, cut to
 * name_cap), and the lanes of its grid (= lxv_threads unless the pair walk's
 * persistent grid is in use, whose resident waves take the lanes in groups
 * of 64). */
int lxv_debug_walk_kernel(const lxv_ctx* ctx, char* name, size_t name_cap, uint32_t* grid_lanes);

/* The pair walk (LXV_WALK_PAIRS) of one lane, run by the host build of the
 * device code: the lane's first centre from u_prev_cur (u(k-8), u(k) as 16
 * little-endian words, the lxv_init form) or, when u_prev_cur is NULL, from
 * the device ladder's set-up for the clamped scalar `start` (the
 * lxv_init_seeded form); then `batches` batches of `batch` (4..256, a power
 * of two) keys. Writes u(k + 8n) for n < batches x batch as 8 little-endian
 * words each. CPU only. */
int lxv_debug_pair_walk(int layout, uint32_t batch, const uint8_t start[32], const uint32_t* u_prev_cur,
                        uint32_t batches, uint32_t* out);

/* Lane t's first centre for lxv_init_seeded's pair walk set-up with the
 * clamped scalar `start`, keys per inversion `batch` and lane stride
 * `stride`, computed twice by the host build of the device code: out[0..15]
 * = (x, y) by a full ladder over start / 8 + batch / 2 - 1 + t stride, and
 * out[16..31] = (x, y) as the set-up kernel computes it, P0 + t S from its
 * two anchors with a ladder over the `bits` low bits of t (t < 2^bits <=
 * 2^63). Canonical little-endian words. CPU only. */
int lxv_debug_pair_centre(int layout, uint32_t batch, const uint8_t start[32], uint64_t stride, uint64_t t,
                          uint32_t bits, uint32_t* out);

/* The command-line tool's CPU key walk (KeyWalk, src/main.rs:229-290, on
 * the host field arithmetic): the public keys u(start + 8n) for
 * n < batches x 256, as 32-byte little-endian encodings into `out`
 * (batches x 256 x 32 bytes). CPU only. */
int lxv_debug_cpu_walk(const uint8_t start[32], uint32_t batches, uint8_t* out);

/* Searches suspended so far by lxv_multi's preemption (lxv_multi_opts.preempt). */
typedef struct lxv_multi lxv_multi;
int lxv_debug_multi_suspensions(lxv_multi* m, uint64_t* n);

#ifdef __cplusplus
}
#endif
#endif
