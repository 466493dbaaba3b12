// verify.hpp -- host verifier of DEEP-FRI / STIR proofs and Merkle openings
// (crates/deep_ali/src/fri.rs:2136-2710, crates/merkle/src/lib.rs:754-812).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "fri.hpp"

namespace sg {

// 1 accept, 0 reject, E_INVALID if the bytes do not deserialize as a
// DeepFriProof<SexticExt>
int deep_fri_verify(const FriParams &p, const uint8_t *proof, size_t len);
// n proofs on the host thread pool; pis (optional) gives proof i its own
// 32-byte public-inputs hash
void deep_fri_verify_batch(const FriParams &p, const uint8_t *const *proofs, const size_t *lens, size_t n,
                           const uint8_t *pis, int *results);
// MerkleTreeChannel::verify_opening for a constant-arity tree of `depth`
// levels: 1 / 0, E_INVALID if the opening bytes do not deserialize
int merkle_verify_opening(uint32_t arity, uint32_t depth, uint64_t label, const uint8_t root[32],
                          const uint8_t *opening, size_t len);
// compute_leaf_hash / compress_leaf_standalone (merkle lib.rs:185-212, 806-812)
void host_leaf_hash(uint32_t arity, uint64_t label, uint64_t pos, const uint64_t *vals, size_t n, uint8_t out[32]);
// root of a constant-arity tree over n leaves of vpl values (merkle lib.rs:377-476)
void host_merkle_root(const uint64_t *vals, uint64_t n, size_t vpl, uint32_t arity, uint32_t depth, uint64_t label,
                      uint8_t root[32]);

}  // namespace sg
