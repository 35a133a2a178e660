// Host-side crypto of the search driver: what the reference takes from
// sha2, x25519-dalek, ed25519-dalek and curve25519-dalek (src/main.rs:29-32,
// src/gpu/mod.rs:11-12). Used to set searches up (start points, the Ed25519
// key, q = u(8B)) and to re-derive every hit from scratch before it is
// accepted (src/main.rs:309-313, src/gpu/mod.rs:157-160).
//
// Written independently of the GPU code: 4 x 64-bit limbs with 128-bit
// products here, 8 x 32-bit limbs with v_mad_u64_u32 chains on the device.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

namespace lxv {

void sha256(const uint8_t* msg, size_t len, uint8_t out[32]);
void sha512(const uint8_t* msg, size_t len, uint8_t out[64]);

// X25519(k, u) with k clamped (RFC 7748).
void x25519(const uint8_t k[32], const uint8_t u[32], uint8_t out[32]);
void x25519_base(const uint8_t k[32], uint8_t out[32]);
// u([k]P) over all 256 bits of k, no clamping (X25519_BASEPOINT * Scalar).
void mont_mul(const uint8_t k[32], const uint8_t u[32], uint8_t out[32]);
// Ed25519 public key of a seed (RFC 8032 5.1.5).
void ed25519_public(const uint8_t seed[32], uint8_t out[32]);

// Scalars: src/main.rs:124-170.
void step_scalar(const uint8_t k[32], uint64_t n, uint8_t out[32]);
void previous_scalar(const uint8_t k[32], uint8_t out[32]);
bool start_scalar_ok(const uint8_t k[32]);
void clamp(uint8_t k[32]);
bool random_bytes(uint8_t* out, size_t n);

// Addresses: src/main.rs:174-197.
void name_hash(uint8_t out[10]);
void lxmf_address(const uint8_t x_pub[32], const uint8_t ed_pub[32], uint8_t out[16]);
// The same with the name hash supplied, for loops over many keys.
void lxmf_address_nh(const uint8_t nh[10], const uint8_t x_pub[32], const uint8_t ed_pub[32], uint8_t out[16]);
void address_of_identity(const uint8_t identity[64], uint8_t out[16]);

// A parsed hex prefix: src/main.rs:199-225.
struct Prefix {
    uint8_t bytes[16] = {0};
    int nibbles = 0;
    static bool parse(const std::string& s, Prefix* out, std::string* err);
    bool matches(const uint8_t addr[16]) const;
    // Big-endian want/mask words, src/gpu/mod.rs:74-84.
    void words(uint32_t want[4], uint32_t mask[4]) const;
};

// The seeded workload generator: SHA-256(tag || le64(seed) || label || 0 ||
// le32(counter)), shared with the tests so workloads can be named by seed.
void drbg(uint64_t seed, const char* label, uint32_t counter, uint8_t out[32]);
void seeded_start(uint64_t seed, uint8_t out[32]);
void seeded_ed_seed(uint64_t seed, uint8_t out[32]);
std::string seeded_prefix(uint64_t seed, int nibbles);

std::string hex(const uint8_t* b, size_t n);

// The CPU key walk of the search command's CPU workers: u(k + 8n) for n = 0,
// 1, 2, ... by x-only differential addition with one inversion per batch
// (KeyWalk, src/main.rs:229-290). Host arithmetic is the 4 x 64-bit field
// above; this is product code, independent of the test oracle.
class KeyWalk {
  public:
    explicit KeyWalk(const uint8_t start[32]);
    // Fills out[0..len) (len <= kBatch) with the next public keys; returns
    // the n of the first.
    uint64_t next_batch(uint8_t (*out)[32], size_t len);
    const uint8_t* start() const { return start_; }
    static constexpr size_t kBatch = 256;

  private:
    uint8_t start_[32];
    uint64_t next_n_ = 0;
    uint64_t prev_x_[4], prev_z_[4], cur_x_[4], cur_z_[4], q_plus_[4], q_minus_[4];
};

}  // namespace lxv
