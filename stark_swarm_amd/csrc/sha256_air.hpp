// sha256_air.hpp -- SHA-256 AIR (crates/deep_ali/src/sha256_air.rs): trace
// builder on the host, composition on the device.
#pragma once
#include <stdint.h>

#include <array>
#include <vector>

#include "ctx.hpp"

namespace sg {

constexpr uint32_t SHA256_W = 756;  // trace columns (sha256_air.rs:248)
constexpr uint32_t SHA256_K = 766;  // transition constraints (sha256_air.rs:288)

uint64_t sha256_n_blocks(uint64_t len);
uint64_t sha256_trace_height(uint64_t len);
// 756 x height column-major trace of msg; returns n_blocks, digest = the
// SHA-256 of msg as read from the trace's H-state
uint64_t build_sha256_trace(const uint8_t *msg, uint64_t len, uint64_t *tr, uint8_t digest[32]);
// deep_ali_merge_sha256 (lib.rs:453-545) for K proofs of one n_trace: lde of
// proof k at lde + 756*k*n0, c_eval at c_eval + k*n0
void merge_sha256_batch(Ctx *c, const uint64_t *lde, uint32_t K, uint64_t n_trace, uint64_t blowup,
                        const std::vector<uint32_t> &n_blocks, int phi_mode, uint64_t *c_eval);

}  // namespace sg
