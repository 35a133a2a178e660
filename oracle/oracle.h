/*
 * CPU oracle for the LXMF vanity GPU search: TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU key walk and address
 * derivation (src/main.rs), used by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker. Product code never links it.
 *
 * What each group restates, and what pins it (tests/test_oracle.py):
 *   SHA-256 / SHA-512   FIPS 180-4; pinned by NIST short-message vectors and
 *                       Python hashlib on random messages.
 *   Field GF(2^255-19)  5 x 51-bit limbs with 128-bit products, as the
 *                       reference's fiat-crypto curve25519_64 (main.rs:43-120);
 *                       pinned by Python big-integer arithmetic on random and
 *                       edge inputs (0, 1, p-1, p, 2^255-1, 2^256-1).
 *   fe_invert           ref10 addition chain, main.rs:92-107; pinned by
 *                       z * z^-1 == 1 and by pow(z, p-2, p).
 *   X25519              RFC 7748 ladder (x25519-dalek 3.0.0 as called at
 *                       main.rs:194, 243); pinned by RFC 7748 5.2 vectors
 *                       (incl. 1 and 1000 iterations) and 6.1 Alice/Bob.
 *   Ed25519 public key  RFC 8032 5.1.5 (ed25519-dalek 3.0.0, main.rs:195,
 *                       295); pinned by RFC 8032 7.1 tests 1-3.
 *   step/previous       main.rs:124-152; pinned by big-integer arithmetic and
 *                       the clamping cases of main.rs:541-553.
 *   start rejection     main.rs:157-170 (the deterministic test only).
 *   name hash, address  main.rs:174-197; pinned by the RNS identity known
 *                       answer of main.rs:509-525.
 *   Prefix              main.rs:199-225 and gpu/mod.rs:74-84; pinned by the
 *                       cases of main.rs:555-562.
 *   KeyWalk             main.rs:229-290 (B = 256 by default); pinned by
 *                       u(k + 8n) == X25519(k + 8n, 9) for every key of
 *                       several batches (the invariant of main.rs:527-539),
 *                       and by the Python fixtures under tests/golden/.
 *   scan                the per-thread search loop, main.rs:292-323, run
 *                       over explicit key ranges so that its hits and the XOR
 *                       of all addresses can be compared with the GPU's.
 */
#ifndef LXV_ORACLE_H
#define LXV_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- hashes (FIPS 180-4) */
void orc_sha256(const uint8_t* msg, size_t len, uint8_t out[32]);
void orc_sha512(const uint8_t* msg, size_t len, uint8_t out[64]);
/* 1 when the SHA-256 used by orc_scan runs on the x86 SHA extensions. */
int orc_sha256_uses_shani(void);

/* ---- field elements as 32 little-endian bytes (bit 255 ignored on input,
 * canonical on output). Exposed so tests can check them against big ints. */
void orc_fe_mul(const uint8_t a[32], const uint8_t b[32], uint8_t out[32]);
void orc_fe_sq(const uint8_t a[32], uint8_t out[32]);
void orc_fe_add(const uint8_t a[32], const uint8_t b[32], uint8_t out[32]);
void orc_fe_sub(const uint8_t a[32], const uint8_t b[32], uint8_t out[32]);
void orc_fe_invert(const uint8_t a[32], uint8_t out[32]);
void orc_fe_canon(const uint8_t a[32], uint8_t out[32]);

/* ---- curve operations */
/* X25519(k, u), RFC 7748: k is clamped, bit 255 of u is masked. */
void orc_x25519(const uint8_t k[32], const uint8_t u[32], uint8_t out[32]);
/* Montgomery ladder on all 256 bits of k with no clamping: the u-coordinate
 * of [k]P (as X25519_BASEPOINT * Scalar::from(8u8), main.rs:244). */
void orc_mont_ladder(const uint8_t k[32], const uint8_t u[32], uint8_t out[32]);
/* Ed25519 public key from a 32-byte seed, RFC 8032 5.1.5. */
void orc_ed25519_pub(const uint8_t seed[32], uint8_t out[32]);

/* ---- scalars (main.rs:124-170) */
void orc_step_scalar(const uint8_t k[32], uint64_t n, uint8_t out[32]);
void orc_previous_scalar(const uint8_t k[32], uint8_t out[32]);
/* 1 if k is a start scalar random_start_scalar could return (clamped, and
 * neither rejection rule of main.rs:164-166 hits). */
int orc_start_scalar_ok(const uint8_t k[32]);

/* ---- addresses and prefixes (main.rs:174-225, gpu/mod.rs:74-84) */
void orc_name_hash(uint8_t out[10]);
void orc_lxmf_address(const uint8_t x_pub[32], const uint8_t ed_pub[32], uint8_t out[16]);
void orc_address_of_identity(const uint8_t identity[64], uint8_t out[16]);
/* Returns 0 and fills bytes/nibbles, or -1 for anything but 1-32 hex digits. */
int orc_prefix_parse(const char* s, uint8_t bytes[16], int* nibbles);
int orc_prefix_matches(const uint8_t bytes[16], int nibbles, const uint8_t addr[16]);
/* Big-endian want/mask words of the prefix, as gpu/mod.rs:74-84. */
void orc_prefix_words(const uint8_t bytes[16], int nibbles, uint32_t want[4], uint32_t mask[4]);

/* ---- the reference CPU walk (main.rs:229-290) */
typedef struct orc_walk orc_walk;
orc_walk* orc_walk_new(const uint8_t start[32]);
/* Fills out[0..len) with u(k + 8n) for the next len keys (len <= batch);
 * returns n of the first. The batch size is fixed by orc_walk_new_batch. */
uint64_t orc_walk_next_batch(orc_walk* w, uint8_t (*out)[32], size_t len);
orc_walk* orc_walk_new_batch(const uint8_t start[32], size_t batch);
void orc_walk_free(orc_walk* w);

/* ---- scanning key ranges, as the search loop does (main.rs:292-323)
 *
 * Walks the keys k + 8g for every g in each range [g0[r], g0[r] + len[r]),
 * derives each LXMF address with ed_pub, XORs every address into xor_out and
 * records every address that matches the prefix as (g, address). Hits come
 * back sorted by g; *n_hits is the total, of which at most cap are stored.
 * threads <= 0 means one per online CPU. Returns 0, or -1 on bad input. */
int orc_scan_ranges(const uint8_t start[32], const uint8_t ed_pub[32],
                    const uint64_t* g0, const uint64_t* len, uint32_t n_ranges,
                    const uint8_t prefix_bytes[16], int prefix_nibbles,
                    uint64_t* hit_g, uint8_t (*hit_addr)[16], uint32_t cap,
                    uint32_t* n_hits, uint8_t xor_out[16], int threads);

#ifdef __cplusplus
}
#endif
#endif
