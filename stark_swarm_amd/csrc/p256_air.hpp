// p256_air.hpp -- double-chain ECDSA-P256 AIR (crates/deep_ali/src/
// p256_ecdsa_double_multirow_air.rs): two scalar-multiplication chains, one
// double-and-conditional-add step of each per trace row, over the 26-bit-limb
// field gadgets of p256_field_air.rs and the RCB-2016 group gadgets of
// p256_group_air.rs. Witness words on the host, cells and composition on the
// device.
#pragma once
#include <stdint.h>

#include <array>
#include <vector>

#include "ctx.hpp"

namespace sg {

constexpr uint32_t P256_W = 151334;  // columns (build_ecdsa_double_multirow_layout)
constexpr uint32_t P256_K = 155774;  // constraints (ecdsa_double_multirow_constraints)
constexpr uint32_t P256_ECDSA_ROWS = 256;

using Be32 = std::array<uint8_t, 32>;

// The inputs of fill_ecdsa_double_multirow (p256_ecdsa_double_multirow_air.rs:
// 134-222): for chain A then chain B the initial accumulator and the base
// point (projective X, Y, Z), 32-byte big-endian values below p; the scalar
// bits (rows past a chain's bits take bit 0); and the six r_proj values
// (r_a x, y, z, r_b x, y, z) replicated in every row. With bind_outputs the
// r_proj values are the chain outputs at rows k_a - 1 and k_b - 1 (the
// callers' two-pass fill).
struct P256Chains {
    std::array<Be32, 12> pts;
    std::array<Be32, 6> rproj;
    std::vector<uint8_t> a_bits, b_bits;
    bool bind_outputs = true;
};

// One ECDSA-P256 signature record (prove_record_ecdsa,
// crates/swarm-dns/examples/mixed_algorithm_zone_rollup.rs:193-322):
// public key (qx, qy), signature (r, s), SHA-256 message digest.
struct P256Record {
    Be32 qx, qy, r, s, digest;
};

// Column-major trace (P256_W columns of h rows), built on the host; r_proj
// (bound or given) is written to rproj_out when not null.
void build_p256_trace_host(const P256Chains &in, uint64_t h, uint64_t *tr, std::array<Be32, 6> *rproj_out);
// The traces of K chain inputs written to d_tr (proof-major, P256_W x h
// each): the host computes each row's limbs, quotients and carries, the
// device expands them into cells. rproj_out (K entries) gets the r_proj
// values used. With d_odd (same shape as d_tr, h <= 4096) one fused pass
// also writes the odd half of each column's blowup-2 LDE there (what
// lde_odd_coset computes from d_tr), without reading the cells back.
void build_p256_traces_dev(Ctx *c, const std::vector<P256Chains> &in, uint64_t h, uint64_t *d_tr,
                           std::vector<std::array<Be32, 6>> *rproj_out, uint64_t *d_odd = nullptr);
// deep_ali_merge_ecdsa_double_multirow_streaming (crates/deep_ali/src/
// lib.rs:1421-1540) with lambda_j = j + 1 for K proofs: lde of proof k at
// lde + P256_W * k * n0, c_eval at c_eval + k * n0. Phi is evaluated at every
// point (the streaming merge has no thread-count quirk).
// d_lam: the 155 774 combination coefficients (device memory), or null for
// lambda_j = j + 1
void merge_p256_batch(Ctx *c, const uint64_t *lde, uint32_t K, uint64_t n_trace, uint64_t blowup, uint64_t *c_eval,
                      const uint64_t *d_lam = nullptr);
// The same c_eval from the K traces themselves (d_tr proof-major, P256_W x
// n_trace each), without the full LDE: every gadget constraint has degree
// <= 2 in trace polynomials of degree < n_trace, so their weighted sum is a
// polynomial of degree <= 2 n_trace - 2, fixed by its values at the
// 2 n_trace points of <w_{2 n_trace}> (a blowup-2 LDE of all columns) and
// interpolated from there; only the 180 columns of the row-gated boundary and
// transition constraints are extended to all n points. Exact arithmetic
// makes the result equal to merge_p256_batch on the full LDE. d_odd, when
// given, is the odd half of that blowup-2 LDE (build_p256_traces_dev's).
void merge_p256_from_trace_batch(Ctx *c, const uint64_t *d_tr, uint32_t K, uint64_t n_trace, uint64_t blowup,
                                 uint64_t *c_eval, const uint64_t *d_odd = nullptr);
// The chain inputs of an ECDSA record: u1 = e / s and u2 = r / s mod n
// (e = digest mod n), chain A from G with the MSB-first bits of u1, chain B
// from Q with those of u2, 256 steps each, r_proj bound to the outputs.
// Throws E_INVALID if the key is not a curve point below p or the signature
// does not verify (the example asserts native_ecdsa_verify first).
// verify = false skips the signature check itself (a double scalar
// multiplication); the range and curve checks and u1, u2 always run
P256Chains p256_ecdsa_chains(const P256Record &rec, bool verify = true);
// SHA3-256("ECDSAP256-MIXEDZONE" || qx || qy || r || s || digest), the
// record's public-input hash (mixed_algorithm_zone_rollup.rs:293-302)
void p256_pk_hash_pre(const P256Record &rec, uint8_t out[32]);

}  // namespace sg
