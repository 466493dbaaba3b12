// prover.hpp -- whole-proof flows (swarm-dns HashRollup provers, prover core).
#pragma once
#include <stdint.h>

#include <array>
#include <vector>

#include "ctx.hpp"
#include "ed25519_air.hpp"
#include "fri.hpp"
#include "p256_air.hpp"
#include "rsa_air.hpp"

namespace sg {

struct DnsRec {
    const char *domain;
    size_t domain_len;
    uint16_t record_type;
    uint32_t ttl;
    const uint8_t *rdata;
    size_t rdata_len;
};

struct InnerShard {
    uint8_t pi_hash[32], merkle_root[32], root_f0[32];
    size_t record_count = 0, n_trace = 0;
    std::vector<uint8_t> proof;
};

// DsKskOutput (crates/swarm-dns/src/prover.rs:464-490)
struct DsKsk {
    uint8_t pi_hash[32], asserted_digest[32], root_f0[32];
    size_t n_blocks = 0, n_trace = 0;
    std::vector<uint8_t> proof;
};
// One SHA-256 / SHA-512 AIR proof (the prove_ds_ksk_binding pipeline).
struct ShaProof {
    uint8_t digest[64];  // SHA-256 uses the first 32 bytes
    uint8_t root_f0[32];
    size_t n_blocks = 0, n_trace = 0;
    std::vector<uint8_t> proof;
};

std::vector<uint8_t> dns_canonical(const DnsRec &r);
size_t dns_canonical_to(const DnsRec &r, uint8_t *dst);  // returns the length written
void dns_leaf_hash(const DnsRec &r, const uint8_t salt[16], uint8_t out[32]);
void dns_merkle_root(const std::vector<uint8_t> &leaves, uint8_t root[32]);
// stir = false: LdtMode::Fri, the classic arity-2 DEEP-FRI of make_schedule (prover.rs:53-67)
std::vector<uint8_t> rollup_prove(Ctx *c, const std::vector<uint64_t> &leaves, const uint8_t binding[32],
                                  int phi_mode, uint8_t root_f0[32], bool stir = true);
InnerShard prove_inner_shard(Ctx *c, const uint8_t salt[16], const std::vector<DnsRec> &recs,
                             const uint8_t fs_binding[32], int phi_mode, bool stir = true);
// K shards in one lockstep device batch per trace size; host hashing on a thread pool
std::vector<InnerShard> prove_inner_shard_batch(Ctx *c, const uint8_t salt[16],
                                                const std::vector<std::vector<DnsRec>> &shards,
                                                const std::vector<std::array<uint8_t, 32>> &fs_bindings,
                                                int phi_mode, bool stir = true);
std::vector<std::vector<uint8_t>> rollup_prove_batch(Ctx *c, const std::vector<std::vector<uint64_t>> &leaves,
                                                     const std::vector<std::array<uint8_t, 32>> &bindings,
                                                     int phi_mode, std::vector<std::array<uint8_t, 32>> &root_f0, bool stir = true);
// SHA-256 (algs[k] = 256) or SHA-512 (512) AIR proofs: trace of msgs[k] with
// height heights[k] (0 = minimum) -> LDE x32 -> merge -> STIR (54 queries,
// bound to fs_bindings[k]); lockstep batches per (alg, height) in HBM-sized
// waves; result k equals the proof of input k alone.
std::vector<ShaProof> prove_sha_batch(Ctx *c, const std::vector<int> &algs, const std::vector<ByteSpan> &msgs,
                                      const std::vector<uint64_t> &heights,
                                      const std::vector<std::array<uint8_t, 32>> &fs_bindings, int phi_mode, bool stir = true);
// prove_ds_ksk_binding (prover.rs:512-585) for K keys: lockstep device
// batches per trace height; result k equals the single-key proof
std::vector<DsKsk> prove_ds_ksk_binding_batch(Ctx *c, const std::vector<ByteSpan> &dnskeys,
                                              const std::vector<std::array<uint8_t, 32>> &parent_ds_hashes,
                                              const std::vector<std::array<uint8_t, 32>> &fs_bindings, int phi_mode, bool stir = true);
// prove_nsec3_completeness output (crates/swarm-dns/src/prover.rs:327-338)
struct Nsec3Proof {
    uint8_t pi_hash[32], chain_root[32], root_f0[32];
    size_t record_count = 0, n_trace = 0;
    std::vector<uint8_t> proof;
};
// prove_nsec3_completeness (prover.rs:348-446) for K closed NSEC3 chains
// (chain k = counts[k] records of owner_hash || next_hash, 64 bytes each):
// lockstep device batches per trace height; result k equals the proof of
// chain k alone. Throws E_INVALID for an empty or open chain.
std::vector<Nsec3Proof> prove_nsec3_batch(Ctx *c, const std::vector<ByteSpan> &chains, const uint8_t salt[16],
                                          const std::vector<std::array<uint8_t, 32>> &fs_bindings, int phi_mode, bool stir = true);

// One RSA-2048 exponentiation proof (one record, or R stacked records).
struct RsaProof {
    uint8_t pi_hash[32], root_f0[32];
    size_t n_trace = 0;
    std::vector<uint8_t> proof;
};
// prove_record_rsa (crates/swarm-dns/examples/mixed_algorithm_zone_rollup.rs:
// 349-395) from the trace on, for records.size() / R proofs of R stacked
// records each (R = 1: the exponentiation AIR; R > 1: the stacked AIR of
// rsa2048_stacked_air.rs): trace of n_trace rows -> LDE x32 -> merge
// (lambda_j = j + 1) -> STIR (54 queries) bound to pis[k]; pi_hash =
// SHA3-256(pis[k] || root_f0). Lockstep in HBM-sized waves; result k equals
// the proof of its records alone.
std::vector<RsaProof> prove_rsa_batch(Ctx *c, const std::vector<RsaRecord> &records, uint32_t R,
                                      uint64_t n_trace, const std::vector<std::array<uint8_t, 32>> &pis);
// One double-chain ECDSA-P256 proof.
struct P256Proof {
    uint8_t pi_hash[32], root_f0[32];
    std::array<Be32, 6> rproj;  // the r_proj values in the trace
    size_t n_trace = 0;
    std::vector<uint8_t> proof;
};
// The double-chain proof of run_double_chain_streaming
// (crates/swarm-dns/examples/p256_full_ecdsa_stark_bench.rs:133-240) for each
// chain input: trace of n_trace rows -> LDE x32 -> merge (lambda_j = j + 1)
// -> STIR (54 queries) bound to pis[k]; pi_hash = SHA3-256(pis[k] ||
// root_f0). Lockstep in HBM-sized waves; result k equals the proof of input
// k alone.
std::vector<P256Proof> prove_p256_batch(Ctx *c, const std::vector<P256Chains> &in, uint64_t n_trace,
                                        const std::vector<std::array<uint8_t, 32>> &pis);
// prove_record_ecdsa (crates/swarm-dns/examples/mixed_algorithm_zone_rollup.rs:
// 193-322) for each record: the chains of p256_ecdsa_chains at n_trace 256,
// bound to the record's pk_hash_pre. Throws E_INVALID for a record whose
// signature does not verify.
std::vector<P256Proof> prove_ecdsa_p256_batch(Ctx *c, const std::vector<P256Record> &recs);
std::vector<uint8_t> prove_outer_rollup(Ctx *c, const uint8_t *pi_hashes, size_t n, const uint8_t pk_hash[32],
                                        int phi_mode, uint8_t root_f0[32], bool stir = true);
std::vector<uint8_t> prover_core_dev(Ctx *c, const uint64_t *d_trace, uint64_t w, uint64_t n_trace, uint64_t blowup,
                                     uint64_t shift, const FriParams &p, uint64_t *d_lde, uint8_t row_root[32]);

// ZskKskOutput of prove_zsk_ksk_binding_v2 (crates/swarm-dns/src/prover.rs:856-951)
struct ZskKsk {
    uint8_t pi_hash[32], root_f0[32];
    size_t n_trace = 0;
    std::vector<uint8_t> proof;
};
// prove_zsk_ksk_binding_v2 for each signature: the v16 Ed25519 verify trace
// (k_scalar rows per ladder) -> merge (lambda_j = j + 1) at the given blowup
// (the reference uses 32) -> STIR with `queries` queries bound to
// fs_bindings[i]; pi_hash = zsk_ksk_pi_hash_v2_stark (prover.rs:763-781).
// Signatures are grouped by the SHA-512 row count of R || A || M and proved
// in lockstep HBM-sized waves; result i equals the proof of signature i alone.
// Throws E_INVALID when a signature does not verify (the reference asserts).
std::vector<ZskKsk> prove_zsk_ksk_binding_v2_batch(Ctx *c, const std::vector<EdSig> &sigs,
                                                   const std::vector<std::array<uint8_t, 32>> &fs_bindings,
                                                   const std::vector<std::array<uint8_t, 32>> &merkle_roots,
                                                   uint32_t k_scalar, uint64_t blowup, uint32_t queries, bool stir = true);

}  // namespace sg
