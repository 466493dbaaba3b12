// rsa_air.hpp -- RSA-2048 exponentiation AIR (s^65537 mod n, one 2048-bit
// modular product per trace row; crates/deep_ali/src/rsa2048_exp_air.rs over
// the witness-quotient gadget of rsa2048_field_air.rs) and its stacked form
// (rsa2048_stacked_air.rs): row words on the host, cells and composition on
// the device.
#pragma once
#include <stdint.h>

#include <array>
#include <vector>

#include "ctx.hpp"

namespace sg {

constexpr uint32_t RSA_W = 11041;  // columns per record (rsa2048_exp_air.rs:97-123)
constexpr uint32_t RSA_K = 11280;  // constraints per record (rsa2048_exp_air.rs:88-90)
constexpr uint32_t RSA_MAX_RECORDS = 4096;

// One RSA-2048 record: modulus n, signature s < n, encoded message em, each
// 256 bytes big-endian (the PKCS#1 byte form for k = 256).
struct RsaRecord {
    std::array<uint8_t, 256> n, s, em;
};

// Column-major trace (RSA_W * records.size() columns of h rows) of
// fill_rsa_stacked / fill_rsa_exp_multirow, built on the host. h is a power
// of two >= 32. Throws E_INVALID for s >= n or n = 0.
void build_rsa_trace_host(const std::vector<RsaRecord> &records, uint64_t h, uint64_t *tr);
// The same traces for P proofs of R records each (record r of proof p is
// records[p * R + r]) written to d_tr (proof-major, each proof's trace
// RSA_W * R x h): the host computes each row's limbs, quotient and carries,
// the device expands them into cells.
void build_rsa_traces_dev(Ctx *c, const std::vector<RsaRecord> &records, uint32_t R, uint64_t h, uint64_t *d_tr);
// deep_ali_merge_rsa_exp_multirow_streaming (lib.rs:1827-1944; R = 1) and
// deep_ali_merge_rsa_stacked_streaming (lib.rs:1949-2049) with
// lambda_j = j + 1, for K proofs: lde of proof k at lde + RSA_W*R*k*n0,
// c_eval at c_eval + k*n0. The streaming merges always evaluate phi.
void merge_rsa_batch(Ctx *c, const uint64_t *lde, uint32_t K, uint32_t R, uint64_t n_trace, uint64_t blowup,
                     uint64_t *c_eval,
                     const uint64_t *d_lam = nullptr);  // d_lam: 11280 R coefficients, or null (j + 1)
// The same c_eval from the K traces themselves (d_tr: K x RSA_W R x
// n_trace), without the blowup-x LDE: every constraint has degree <= 2 in the
// cells of the current and next row, so each gate group's weighted sum
// (ungated; the EM check of the output row; the transitions) is a
// polynomial of degree <= 2 n_trace - 2, evaluated on a blowup-2 LDE and
// interpolated; the gates of each point's trace row then select the groups.
// Exact arithmetic makes the result equal to merge_rsa_batch. blowup >= 4.
void merge_rsa_from_trace_batch(Ctx *c, const uint64_t *d_tr, uint32_t K, uint32_t R, uint64_t n_trace,
                                uint64_t blowup, uint64_t *c_eval);
// n > 0, s < n and s^65537 mod n == em: the record the exponentiation AIR
// proves (its output-row constraints hold)
bool rsa_record_verifies(const RsaRecord &rec);
// emsa_pkcs1_v1_5_encode_sha256 (crates/deep_ali/src/rsa2048.rs) for k = 256
void rsa_pkcs1_em_sha256(const uint8_t digest[32], uint8_t em[256]);

}  // namespace sg
