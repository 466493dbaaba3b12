// fri.hpp -- DEEP-FRI / STIR prover on the device (crates/deep_ali/src/fri.rs).
#pragma once
#include <stdint.h>

#include <array>
#include <vector>

#include "ctx.hpp"

namespace sg {

struct FriParams {
    std::vector<uint64_t> schedule;
    uint64_t r = 54, s0 = 54;
    uint64_t seed_z = 0xDEEFBAADULL;
    bool coeff_commit_final = true;
    uint64_t d_final = 1;
    bool stir = true;
    bool has_pi = false;
    uint8_t pi[32] = {0};
};

// STIR DEEP-FRI prove over f0 (device, n0 base-field values, natural order on
// <omega_n0>). Returns the ark-serialize compressed DeepFriProof<SexticExt>.
// raw_out (optional) receives the raw query index of each proximity query.
std::vector<uint8_t> deep_fri_prove_dev(Ctx *c, const uint64_t *d_f0, uint64_t n0, const FriParams &p,
                                        std::vector<uint64_t> *raw_out);

// K proofs of the same n0 and parameters in lockstep: f0 of proof k is
// d_f0[k*n0 .. (k+1)*n0); pis (optional) gives each proof's public-inputs
// hash (otherwise p.pi / none for all). Each output equals the single-proof
// result for that f0.
std::vector<std::vector<uint8_t>> deep_fri_prove_batch(Ctx *c, const uint64_t *d_f0, uint32_t K, uint64_t n0,
                                                       const FriParams &p,
                                                       const std::vector<std::array<uint8_t, 32>> *pis,
                                                       std::vector<std::vector<uint64_t>> *raw_out);

// swarm-dns schedule: 8s plus a residual 2^(log n0 mod 3) (prover.rs:53-67)
std::vector<uint64_t> stir_schedule(uint64_t n0);

}  // namespace sg
