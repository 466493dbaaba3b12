// ed25519_air.hpp -- the Ed25519 cofactored-verify AIR, v16
// (crates/deep_ali/src/ed25519_verify_air.rs): host witness words, device
// trace expansion and composition, and prove_zsk_ksk_binding_v2.
#pragma once
#include <stdint.h>

#include <array>
#include <vector>

#include "ctx.hpp"

namespace sg {

// One signature: public key A, signature R || s, message M (the reference's
// (ksk_pubkey, signature, signed_data)).
struct EdSig {
    const uint8_t *pk;   // 32 bytes
    const uint8_t *sig;  // 64 bytes
    const uint8_t *msg;
    size_t len;
};

// verify_air_layout_v16 sizes for a message of len bytes and k_scalar = k
struct EdShape {
    uint64_t width = 0, height = 0, n_constraints = 0, k_hash = 0;
};
EdShape ed25519_shape(size_t msg_len, uint32_t k);

// RFC 8032 verification as the reference's ed25519_verify.rs:54-118
// (cofactored: [8]([s]B - R - [k]A) = O, s < L)
bool ed25519_verify(const EdSig &s);
// RFC 8032 section 5.1.6 signing from a 32-byte seed (synthetic workloads)
void ed25519_sign(const uint8_t seed[32], const uint8_t *msg, size_t len, uint8_t pk[32], uint8_t sig[64]);

// fill_verify_air_v16 on the host: width x height column-major cells (the
// same words and expansion as the device path). Throws E_INVALID when R or A
// does not decompress.
void build_ed25519_trace_host(const EdSig &s, uint32_t k, uint64_t *tr);
// Per-proof pins and SHA-512 block counts of a batch, from the native values
// the trace builder computes (so the merge need not compute them again).
struct EdConsts {
    std::vector<uint64_t> cst;  // G x n_const pins
    std::vector<uint32_t> nb;   // G SHA-512 block counts
};
// G traces of one shape (every message of the same SHA-512 block count class,
// i.e. the same k_hash): d_tr is G x width x height. consts, if given,
// receives the batch's pins for merge_ed25519_from_trace_batch.
void build_ed25519_traces_dev(Ctx *c, const std::vector<EdSig> &sigs, uint32_t k, uint64_t *d_tr,
                              EdConsts *consts = nullptr);

// deep_ali_merge_ed25519_verify_streaming (lib.rs:854-974) on the callers'
// LDEs (G x width x n, n = height * blowup), coefficients d_coeffs
// (n_constraints values on the device, shared by the batch) or nullptr for
// lambda_j = j + 1; c_eval is G x n.
void merge_ed25519_batch(Ctx *c, const std::vector<EdSig> &sigs, uint32_t k, const uint64_t *lde, uint64_t blowup,
                         const uint64_t *d_coeffs, uint64_t *c_eval);
// The same c_eval from the traces (G x width x height): every constraint has
// degree <= 2 in the current and next rows, so each row-gate group's weighted
// sum is a polynomial of degree <= 2 height - 2, evaluated on a blowup-2 LDE
// (the traces and their odd coset) with the gates ignored, interpolated to
// all n points and selected there by each point's trace row. blowup >= 4.
// consts: the trace builder's EdConsts of the same batch, or nullptr to
// compute the pins here.
void merge_ed25519_from_trace_batch(Ctx *c, const std::vector<EdSig> &sigs, uint32_t k, const uint64_t *d_tr,
                                    uint64_t blowup, const uint64_t *d_coeffs, uint64_t *c_eval,
                                    const EdConsts *consts = nullptr);

}  // namespace sg
