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

constexpr uint32_t SHA512_W = 1510;  // sha512_air.rs:478
constexpr uint32_t SHA512_K = 1526;  // sha512_air.rs:477

uint64_t sha256_n_blocks(uint64_t len);
uint64_t sha256_trace_height(uint64_t len);
// 756 x height column-major trace of msg; returns n_blocks, digest = the
// SHA-256 of msg as read from the trace's H-state. height 0 = the minimum
// (next_pow2(128 n_blocks)); a larger power of two pads by repeating the
// last useful row, the builder's own padding rule.
uint64_t build_sha256_trace(const uint8_t *msg, uint64_t len, uint64_t *tr, uint8_t digest[32], uint64_t height = 0);
// K traces of height h built on the device (d_tr: K x 756 x h): the host
// derives each row's 44 words (the sequential rounds), the device expands
// them into cells and repeats the last useful row; returns n_blocks and the
// 32-byte digests (first half of each 64-byte entry)
void build_sha256_traces_dev(Ctx *c, const std::vector<ByteSpan> &msgs, uint64_t h, uint64_t *d_tr,
                             std::vector<uint32_t> &n_blocks, std::vector<std::array<uint8_t, 64>> &digests);
uint64_t sha512_n_blocks(uint64_t len);
uint64_t sha512_trace_height(uint64_t len);
// 1510 x height trace (two 32-bit limbs per word); digest = SHA-512(msg)
uint64_t build_sha512_trace(const uint8_t *msg, uint64_t len, uint64_t *tr, uint8_t digest[64], uint64_t height = 0);
void build_sha512_traces_dev(Ctx *c, const std::vector<ByteSpan> &msgs, uint64_t h, uint64_t *d_tr,
                             std::vector<uint32_t> &n_blocks, std::vector<std::array<uint8_t, 64>> &digests);
// the same rows written into a wider trace: rows h of trace k at
// d_tr + k * pstride, column c at offset c * cstride
void build_sha512_traces_into(Ctx *c, const std::vector<ByteSpan> &msgs, uint64_t h, uint64_t cstride,
                              uint64_t pstride, uint64_t *d_tr, std::vector<uint32_t> &n_blocks,
                              std::vector<std::array<uint8_t, 64>> &digests);
// The SHA-512 slots (0..1525) of an AIR whose first 1510 columns are SHA-512
// rows (wcols columns per proof): phi of those slots at every point of a
// caller's LDE, zero on rows r with r + 1 >= row_limit; d_lam = the slots'
// coefficients (nullptr: j + 1).
void compose_sha512_rows(Ctx *c, const uint64_t *lde, uint64_t wcols, uint32_t K, uint64_t n, uint64_t blowup,
                         const uint32_t *d_nb, uint64_t row_limit, const uint64_t *d_lam, uint64_t *phi);
// ... and their five gate-group sums on a blowup-2 LDE (traces and odd coset,
// n2 points): groups[(k * 5 + g) * n2 + i] (ungated, compression rows,
// finalisation row, other rows, schedule rows), t1_lam = the weights of the
// two T1 identities, whose round constants the caller subtracts
void compose_sha512_groups(Ctx *c, const uint64_t *d_tr, const uint64_t *d_odd, uint64_t wcols, uint32_t K,
                           uint64_t n2, const uint32_t *d_nb, const uint64_t *d_lam, uint64_t *groups,
                           uint64_t *t1_lam);
const uint64_t *sha512_round_constants();  // FIPS 180-4 K[0..80], host memory
// deep_ali_merge_sha256 (lib.rs:453-545) for K proofs of one n_trace: lde of
// proof k at lde + 756*k*n0, c_eval at c_eval + k*n0; d_lam: the 766
// combination coefficients (device memory), or null for lambda_j = j + 1
void merge_sha256_batch(Ctx *c, const uint64_t *lde, uint32_t K, uint64_t n_trace, uint64_t blowup,
                        const std::vector<uint32_t> &n_blocks, int phi_mode, uint64_t *c_eval,
                        const uint64_t *d_lam = nullptr);
// deep_ali_merge_sha512 (lib.rs:561-653), same batching; d_lam: 1526 coefficients
void merge_sha512_batch(Ctx *c, const uint64_t *lde, uint32_t K, uint64_t n_trace, uint64_t blowup,
                        const std::vector<uint32_t> &n_blocks, int phi_mode, uint64_t *c_eval,
                        const uint64_t *d_lam = nullptr);
// The same c_eval from the K traces themselves (d_tr: K x W x n_trace),
// without the blowup-x LDE. Every constraint has degree <= 2 in the cells of
// the current and next row, and the next row's polynomial is T(w_n x), so
// each gate group's weighted sum (ungated; compression rows; finalisation
// row; other rows; schedule rows) is a polynomial of degree <= 2 n_trace - 2
// in the point: it is evaluated on the 2 n_trace points of a blowup-2 LDE
// with the gates ignored and interpolated to all n points, where the gates of
// each point's trace row select the groups and the round constants K[t]
// (the one row-dependent term, in the T1 identity) are subtracted. Exact
// arithmetic makes the result equal to merge_sha256_batch / _sha512_batch
// on the full LDE. blowup >= 4.
void merge_sha256_from_trace_batch(Ctx *c, const uint64_t *d_tr, uint32_t K, uint64_t n_trace, uint64_t blowup,
                                   const std::vector<uint32_t> &n_blocks, int phi_mode, uint64_t *c_eval,
                                   const uint64_t *d_lam = nullptr);
void merge_sha512_from_trace_batch(Ctx *c, const uint64_t *d_tr, uint32_t K, uint64_t n_trace, uint64_t blowup,
                                   const std::vector<uint32_t> &n_blocks, int phi_mode, uint64_t *c_eval,
                                   const uint64_t *d_lam = nullptr);

}  // namespace sg
