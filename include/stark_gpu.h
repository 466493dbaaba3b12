/*
 * stark_gpu.h -- C ABI of the MI355X (gfx950) proving backend for
 * stark-swarm's data-parallel prover core.
 *
 * Conventions (a Rust / C caller binds these with plain pointers and sizes):
 *  - every field value crosses as a canonical u64 in [0, p), p = 2^64-2^32+1;
 *    an Fp6 value is 6 u64 in the reference's to_fp_components order;
 *  - columns are column-major: column c occupies [c*n, (c+1)*n);
 *  - status codes: 0 ok, STARK_E_* < 0 on error; stark_last_error() gives the
 *    message; the library never aborts the process;
 *  - output buffers are caller-owned, except proof bytes, which the library
 *    allocates and the caller releases with stark_free();
 *  - one context owns one HIP stream; calls on different contexts may run
 *    from different threads; a single context is not re-entrant.
 *
 * Citations are to the reference workspace (saholmes/stark-swarm,
 * stark-stir-swarm/); each entry point names the Rust function it replaces.
 */
#ifndef STARK_GPU_H
#define STARK_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STARK_OK 0
#define STARK_E_INVALID (-1)
#define STARK_E_OOM (-2)
#define STARK_E_DEVICE (-3)
#define STARK_E_UNSUPPORTED (-4)
#define STARK_E_INTERNAL (-5)

/* phi_mode of the composition merge: the reference leaves phi identically
 * zero when its rayon build runs with n < 4096 points or on one thread
 * (crates/deep_ali/src/lib.rs:20-33, 374-397). */
#define STARK_PHI_SEQUENTIAL 0 /* non-parallel build: phi always computed */
#define STARK_PHI_RAYON 1      /* parallel build, >1 thread: phi = 0 iff n < 4096 */
#define STARK_PHI_ONE_THREAD 2 /* parallel build, 1 thread: phi always 0 */

#define STARK_LDT_STIR 0 /* LdtMode::Stir (crates/swarm-dns/src/prover.rs:37-41) */
#define STARK_LDT_FRI 1  /* LdtMode::Fri: classic DEEP-FRI with arity-2 folds (prover.rs:53-67) */

/* AirType ids: the variant's position (1-based) in crates/deep_ali/src/air_workloads.rs:20-140 */
#define STARK_AIR_HASH_ROLLUP 5 /* AirType::HashRollup (air_workloads.rs:83) */
#define STARK_AIR_NSEC3_CHAIN 6 /* AirType::Nsec3Chain (air_workloads.rs:117) */

typedef struct stark_ctx stark_ctx;

/* DeepFriParams (crates/deep_ali/src/fri.rs:1303-1314) */
typedef struct {
    const size_t *schedule;
    size_t n_layers;
    size_t r;
    size_t s0;
    uint64_t seed_z;
    int coeff_commit_final;
    size_t d_final;
    int stir;
    const uint8_t *public_inputs_hash; /* 32 bytes, or NULL for None */
} stark_fri_params;

/* DnsRecord (crates/swarm-dns/src/dns.rs:24-31) */
typedef struct {
    const char *domain;
    size_t domain_len;
    uint16_t record_type;
    uint32_t ttl;
    const uint8_t *rdata;
    size_t rdata_len;
} stark_dns_record;

/* InnerShardOutput (crates/swarm-dns/src/prover.rs:105-136) */
typedef struct {
    uint8_t pi_hash[32];
    uint8_t merkle_root[32];
    uint8_t root_f0[32];
    size_t record_count;
    size_t n_trace;
    uint8_t *proof; /* ark-serialize compressed DeepFriProof<SexticExt>; stark_free() it */
    size_t proof_len;
} stark_inner_shard_out;

/* DsKskOutput (crates/swarm-dns/src/prover.rs:464-490) */
typedef struct {
    uint8_t pi_hash[32];
    uint8_t asserted_digest[32]; /* SHA-256(dnskey) as read from the trace */
    uint8_t root_f0[32];
    size_t n_blocks;
    size_t n_trace;
    uint8_t *proof; /* ark-serialize compressed DeepFriProof<SexticExt>; stark_free() it */
    size_t proof_len;
} stark_ds_ksk_out;

/* One SHA-256 / SHA-512 AIR proof of stark_prove_sha_batch */
typedef struct {
    uint8_t digest[64]; /* SHA-256 fills the first 32 bytes */
    uint8_t root_f0[32];
    size_t n_blocks;
    size_t n_trace;
    uint8_t *proof; /* ark-serialize compressed DeepFriProof<SexticExt>; stark_free() it */
    size_t proof_len;
} stark_sha_out;

/* Nsec3Output (crates/swarm-dns/src/prover.rs:327-338) */
typedef struct {
    uint8_t pi_hash[32];
    uint8_t chain_root[32];
    uint8_t root_f0[32];
    size_t record_count;
    size_t n_trace;
    uint8_t *proof; /* stark_free() it */
    size_t proof_len;
} stark_nsec3_out;

/* One RSA-2048 exponentiation proof of stark_prove_rsa_exp_batch */
typedef struct {
    uint8_t pi_hash[32]; /* SHA3-256(public_inputs_hash || root_f0) */
    uint8_t root_f0[32];
    size_t n_trace;
    uint8_t *proof; /* ark-serialize compressed DeepFriProof<SexticExt>; stark_free() it */
    size_t proof_len;
} stark_rsa_out;

/* Inputs of fill_ecdsa_double_multirow (crates/deep_ali/src/
 * p256_ecdsa_double_multirow_air.rs:134-222) for one double-chain trace:
 * 32-byte big-endian field elements below p. */
typedef struct {
    uint8_t pts[12][32];  /* chain A initial x, y, z, base x, y, z; then chain B the same */
    uint8_t rproj[6][32]; /* r_a x, y, z, r_b x, y, z (used when bind_outputs == 0) */
    const uint8_t *a_bits; /* k_a scalar bits of chain A (0/1), row order; rows past them take 0 */
    size_t k_a;
    const uint8_t *b_bits;
    size_t k_b;
    int bind_outputs; /* 1: r_proj = the chain outputs at rows k_a - 1 and k_b - 1 (the callers' two-pass fill) */
} stark_p256_chains;

/* One double-chain ECDSA-P256 proof */
typedef struct {
    uint8_t pi_hash[32]; /* SHA3-256(public_inputs_hash || root_f0) */
    uint8_t root_f0[32];
    uint8_t rproj[6][32]; /* the r_proj values bound in the trace */
    size_t n_trace;
    uint8_t *proof; /* ark-serialize compressed DeepFriProof<SexticExt>; stark_free() it */
    size_t proof_len;
} stark_p256_out;

/* ---- proving server (one GPU, continuous batching) ---- */
typedef struct stark_server stark_server;

#define STARK_SERVE_DS_KSK 0      /* prove_ds_ksk_binding request */
#define STARK_SERVE_INNER_SHARD 1 /* prove_inner_shard request */
#define STARK_SERVE_RSA 2         /* RSA-2048 record (stark_prove_rsa_exp_batch with one record) */
#define STARK_SERVE_P256 3        /* ECDSA-P256 record (stark_prove_ecdsa_p256_batch) */
#define STARK_SERVE_ED25519 4     /* Ed25519 ZSK->KSK binding (stark_prove_zsk_ksk_binding_v2_batch, k_scalar 256) */
#define STARK_SERVER_DRY_RUN 1    /* flag: schedule only, no device work (admission tests) */
#define STARK_SERVER_NO_VALIDATE 2 /* flag: skip the signature checks at submit, so a bad request reaches a
                                      batch (tests of the per-request fallback) */
#define STARK_E_TIMEOUT (-6)

typedef struct {
    int kind;
    const uint8_t *dnskey; /* DS-KSK */
    size_t dnskey_len;
    const uint8_t *parent_ds_hash; /* DS-KSK, 32 bytes */
    const uint8_t *salt;           /* inner shard, 16 bytes */
    const stark_dns_record *records;
    size_t n_records;
    const uint8_t *fs_binding; /* 32 bytes; RSA: the public-inputs hash */
    const uint8_t *rsa_n, *rsa_s, *rsa_em; /* RSA: 256 bytes each, big-endian, s < n */
    const uint8_t *p256_qx, *p256_qy, *p256_r, *p256_s, *p256_digest; /* ECDSA-P256: 32 bytes each, big-endian */
    const uint8_t *ed_pk, *ed_sig; /* Ed25519: 32-byte public key, 64-byte signature */
    const uint8_t *ed_msg;         /* Ed25519: the signed bytes */
    size_t ed_msg_len;
    const uint8_t *ed_merkle_root; /* Ed25519: 32 bytes */
    double deadline_s; /* > 0: the request completes with STARK_E_TIMEOUT unless admitted within this many seconds */
} stark_serve_request;

typedef struct {
    uint64_t id;
    int status; /* 0 or a STARK_E_* code of the proof */
    int kind;
    uint8_t pi_hash[32];
    uint8_t root_f0[32];
    uint8_t aux[32]; /* inner shard: merkle_root; DS-KSK: asserted_digest; RSA, P-256: zero */
    size_t n_trace;
    size_t batch_size; /* requests proven in the same lockstep batch */
    uint64_t batch_seq; /* admission order of that batch (1, 2, ...) */
    uint8_t *proof;    /* stark_free() it */
    size_t proof_len;
    double t_submit, t_start, t_done; /* seconds since the server started; t_start = admission */
    char error[192]; /* status != 0: what failed (this request's own error), NUL-terminated */
} stark_serve_result;

/* ---- context ---- */
int stark_ctx_create(int device, stark_ctx **out);
void stark_ctx_destroy(stark_ctx *ctx);
const char *stark_last_error(const stark_ctx *ctx);
void stark_free(void *p);

/* ---- host-buffer drop-ins ---- */

/* lde_trace_columns (crates/deep_ali/src/trace_import.rs:251-281):
 * cols is w x n_trace, out is w x (n_trace*blowup), natural order on <w_n0>. */
int stark_lde_columns(stark_ctx *ctx, const uint64_t *cols, size_t w, size_t n_trace, size_t blowup, uint64_t *out);

/* deep_ali_merge_general (crates/deep_ali/src/lib.rs:343-437) for
 * air_id = STARK_AIR_HASH_ROLLUP (w 4, k 3) or STARK_AIR_NSEC3_CHAIN (w 8, k 4); coeffs are the k combination coefficients
 * (any values; the reference's provers pass lambda_j = j+1,
 * prover.rs:100-102). lde is w x n0, c_eval receives n0 values. */
int stark_merge(stark_ctx *ctx, int air_id, const uint64_t *lde, size_t w, size_t n_trace, size_t blowup,
                const uint64_t *coeffs, size_t k, int phi_mode, uint64_t *c_eval);

/* deep_fri_prove::<SexticExt> (crates/deep_ali/src/fri.rs:1900-2055), STIR
 * mode. *proof receives the ark-serialize compressed DeepFriProof bytes. */
int stark_deep_fri_prove(stark_ctx *ctx, const uint64_t *f0, size_t n0, const stark_fri_params *params,
                         uint8_t **proof, size_t *proof_len);

/* MerkleTreeChannel::{new, push_leaves_parallel, finalize}
 * (crates/merkle/src/lib.rs:365-476) with constant arity: leaf i hashes
 * leaf_vals[i*vals_per_leaf .. (i+1)*vals_per_leaf]. */
int stark_merkle_commit(stark_ctx *ctx, const uint64_t *leaf_vals, size_t n_leaves, size_t vals_per_leaf,
                        size_t arity, uint64_t tree_label, uint8_t root[32]);

/* MerkleTreeChannel::{new, push_leaves_parallel, finalize, open}
 * (crates/merkle/src/lib.rs:365-512) with constant arity: builds the tree of
 * stark_merkle_commit on the device and opens leaves indices[0..n_idx).
 * *openings (library-allocated, stark_free() it) holds the serialized
 * MerkleOpening of each index (leaf digest, sibling path, index; ark-serialize
 * compressed) back to back, opening q at [offsets[q], offsets[q + 1]). */
int stark_merkle_open(stark_ctx *ctx, const uint64_t *leaf_vals, size_t n_leaves, size_t vals_per_leaf, size_t arity,
                      uint64_t tree_label, const size_t *indices, size_t n_idx, uint8_t root[32], uint8_t **openings,
                      size_t *openings_len, size_t *offsets /* n_idx + 1 */);

/* MerkleTreeChannel::verify_opening (crates/merkle/src/lib.rs:754-799) for a
 * tree of `depth` levels of constant arity: 1 valid, 0 not, STARK_E_INVALID
 * when the opening bytes do not deserialize. Host code, no context. */
int stark_merkle_verify_opening(size_t arity, size_t depth, uint64_t tree_label, const uint8_t root[32],
                                const uint8_t *opening, size_t len);
/* compute_leaf_hash (crates/merkle/src/lib.rs:806-812) */
void stark_host_merkle_leaf_hash(size_t arity, uint64_t tree_label, size_t index, const uint64_t *vals, size_t n,
                                 uint8_t out[32]);

/* deep_fri_verify::<SexticExt> (crates/deep_ali/src/fri.rs:2136-2710) on the
 * ark-serialize compressed DeepFriProof bytes, STIR and classic FRI: 1
 * accept, 0 reject, STARK_E_INVALID when the bytes do not deserialize (the
 * reference's deserialize_with_mode error, swarm-verify/src/main.rs:243-245).
 * A proof that would make the reference panic is rejected. Host code (the
 * reference's verifier runs on the host too), no context. */
int stark_deep_fri_verify(const stark_fri_params *params, const uint8_t *proof, size_t proof_len);
/* n proofs under the same parameters on the library's host thread pool;
 * public_inputs_hashes (32 bytes per proof, or NULL for
 * params->public_inputs_hash) binds proof i to its own hash. results[i] as
 * stark_deep_fri_verify. No single counterpart in the reference (a loop of
 * deep_fri_verify calls, swarm-verify/src/main.rs:266). */
int stark_deep_fri_verify_batch(const stark_fri_params *params, const uint8_t *const *proofs, const size_t *lens,
                                size_t n, const uint8_t *public_inputs_hashes, int *results);

/* prove_inner_shard (crates/swarm-dns/src/prover.rs:148-227). The local
 * self-verify of the reference is not run (proof bytes are byte-identical). */
int stark_prove_inner_shard(stark_ctx *ctx, const uint8_t salt[16], const stark_dns_record *records, size_t n,
                            const uint8_t fs_binding[32], int ldt, int phi_mode, stark_inner_shard_out *out);

/* Batched prove_inner_shard: n_shards independent shards proven together
 * (one lockstep device batch per trace size, host hashing on a thread pool).
 * Shard s owns records [sum(shard_sizes[0..s]), +shard_sizes[s]) of
 * `records` and fs_bindings[32*s .. 32*s+32); outs[s] receives exactly what
 * stark_prove_inner_shard returns for that shard alone. No counterpart in
 * the reference: it replaces a loop of prove_inner_shard calls
 * (crates/swarm-dns/src/prover.rs:148) over the shards a worker holds. */
int stark_prove_inner_shard_batch(stark_ctx *ctx, const uint8_t salt[16], const stark_dns_record *records,
                                  const size_t *shard_sizes, size_t n_shards, const uint8_t *fs_bindings, int ldt,
                                  int phi_mode, stark_inner_shard_out *outs);

/* Batched deep_fri_prove (crates/deep_ali/src/fri.rs:1900): n_proofs f0 vectors
 * of n0 values each, back to back; public_inputs_hashes (32 bytes per proof,
 * or NULL to use params->public_inputs_hash for all). proofs[k] / proof_lens[k]
 * receive library-allocated bytes equal to the single-proof result. */
int stark_deep_fri_prove_batch(stark_ctx *ctx, const uint64_t *f0s, size_t n_proofs, size_t n0,
                               const stark_fri_params *params, const uint8_t *public_inputs_hashes,
                               uint8_t **proofs, size_t *proof_lens);

/* deep_ali_merge_sha256 (crates/deep_ali/src/lib.rs:453-545): lde is the
 * 756 x n0 SHA-256 AIR LDE, coeffs the 766 combination coefficients (any
 * values; the provers use lambda_j = j+1, prover.rs:100-102). */
int stark_merge_sha256(stark_ctx *ctx, const uint64_t *lde, size_t n_trace, size_t blowup, size_t n_blocks,
                       const uint64_t *coeffs, size_t k, int phi_mode, uint64_t *c_eval);

/* prove_ds_ksk_binding (crates/swarm-dns/src/prover.rs:512-585): SHA-256
 * trace of dnskey -> LDE x32 -> SHA-256 merge -> STIR proof bound to
 * fs_binding -> ds_ksk_pi_hash. The reference's local self-verify
 * (prover.rs:563-566) is left to the caller. */
int stark_prove_ds_ksk_binding(stark_ctx *ctx, const uint8_t *dnskey, size_t len, const uint8_t parent_ds_hash[32],
                               const uint8_t fs_binding[32], int ldt, int phi_mode, stark_ds_ksk_out *out);

/* Batched prove_ds_ksk_binding: key i is dnskeys[offsets[i] .. offsets[i] +
 * lens[i]), with parent_ds_hashes / fs_bindings 32 bytes per key; keys of one
 * trace height are proven as one lockstep device batch. outs[i] receives
 * exactly what stark_prove_ds_ksk_binding returns for key i alone. No
 * counterpart in the reference: it replaces a loop of prove_ds_ksk_binding
 * calls (crates/swarm-dns/src/prover.rs:512). */
int stark_prove_ds_ksk_binding_batch(stark_ctx *ctx, const uint8_t *dnskeys, const size_t *offsets, const size_t *lens,
                                     size_t n_keys, const uint8_t *parent_ds_hashes, const uint8_t *fs_bindings,
                                     int ldt, int phi_mode, stark_ds_ksk_out *outs);

/* deep_ali_merge_sha512 (crates/deep_ali/src/lib.rs:561-653): lde is the
 * 1510 x n0 SHA-512 AIR LDE, coeffs the 1526 coefficients (any values). */
int stark_merge_sha512(stark_ctx *ctx, const uint64_t *lde, size_t n_trace, size_t blowup, size_t n_blocks,
                       const uint64_t *coeffs, size_t k, int phi_mode, uint64_t *c_eval);
/* deep_ali_merge_sha256 / _sha512 (alg 256 / 512) with the LDE (width x n0)
 * and c_eval (n0) in device memory; coeffs in host memory. */
int stark_merge_sha_dev(stark_ctx *ctx, int alg, const uint64_t *d_lde, size_t n_trace, size_t blowup,
                        size_t n_blocks, const uint64_t *coeffs, size_t k, int phi_mode, uint64_t *d_c_eval);
/* The same c_eval as stark_merge_sha_dev on lde_trace_columns(trace, blowup)
 * (trace_import.rs:251-281), computed from the width x n_trace trace in
 * device memory without that LDE: the provers' route (gate-group sums on a
 * blowup-2 LDE, interpolated to n0 points). blowup >= 4. No counterpart in
 * the reference, which always builds the LDE first. */
int stark_merge_sha_from_trace_dev(stark_ctx *ctx, int alg, const uint64_t *d_trace, size_t n_trace, size_t blowup,
                                   size_t n_blocks, const uint64_t *coeffs, size_t k, int phi_mode,
                                   uint64_t *d_c_eval);

/* SHA-256 (algs[i] = 256) or SHA-512 (512) AIR proofs over n messages with
 * the prove_ds_ksk_binding pipeline (trace -> LDE x32 -> merge -> STIR with
 * 54 queries bound to fs_bindings[32 i ..]); message i is msgs[offsets[i] ..
 * + lens[i]) and its trace has heights[i] rows (a power of two >= the
 * message's minimum; padding rows repeat the last useful row, the trace
 * builder's own rule, sha256_air.rs:441-479 / sha512_air.rs:551-573), or the
 * minimum when heights is NULL. Proven in lockstep device batches per
 * (alg, height), in waves sized to free HBM. No single counterpart in the
 * reference: the batched SHA-256/512 workload of the swarm's hash AIRs. */
int stark_prove_sha_batch(stark_ctx *ctx, const int *algs, const uint8_t *msgs, const size_t *offsets, const size_t *lens,
                          const size_t *heights, size_t n, const uint8_t *fs_bindings, int phi_mode,
                          stark_sha_out *outs);

/* deep_ali_merge_rsa_exp_multirow_streaming (crates/deep_ali/src/lib.rs:
 * 1827-1944; n_records = 1) and deep_ali_merge_rsa_stacked_streaming
 * (lib.rs:1949-2049; n_records stacked RSA-2048 chains): lde is the
 * (11041 n_records) x n0 LDE, coeffs the 11280 n_records coefficients
 * (any values; lambda_j = j+1 takes the regrouped fast path, others a
 * term-by-term one). The streaming merges always evaluate phi. */
int stark_merge_rsa_exp(stark_ctx *ctx, const uint64_t *lde, size_t n_trace, size_t blowup, size_t n_records,
                        const uint64_t *coeffs, size_t k, uint64_t *c_eval);

/* prove_record_rsa (crates/swarm-dns/examples/mixed_algorithm_zone_rollup.rs:
 * 349-395) from the trace on, for n_proofs proofs of records_per_proof
 * stacked RSA-2048 records each: record j of proof i is (n, s, em) at byte
 * offset 256 (i records_per_proof + j) of n_be, s_be, em_be (big-endian,
 * s < n; em as the chain's expected s^65537 mod n, the caller's PKCS#1
 * encoding). Trace of n_trace rows (a power of two >= 32; the reference
 * uses 32) -> LDE x32 -> merge -> STIR with 54 queries bound to
 * public_inputs_hashes[32 i ..]. Lockstep device batches in HBM-sized
 * waves; result i equals the proof of its records alone. */
int stark_prove_rsa_exp_batch(stark_ctx *ctx, const uint8_t *n_be, const uint8_t *s_be, const uint8_t *em_be,
                              size_t n_proofs, size_t records_per_proof, size_t n_trace,
                              const uint8_t *public_inputs_hashes, stark_rsa_out *outs);

/* deep_ali_merge_ecdsa_double_multirow_streaming (crates/deep_ali/src/lib.rs:
 * 1421-1540): lde is the 151334 x n0 LDE of a double-chain trace, coeffs the
 * 155774 coefficients (any values; lambda_j = j+1 takes the regrouped fast
 * path, others a term-by-term one). Phi is evaluated at every point. */
int stark_merge_ecdsa_p256(stark_ctx *ctx, const uint64_t *lde, size_t n_trace, size_t blowup, const uint64_t *coeffs,
                           size_t k, uint64_t *c_eval);

/* run_double_chain_streaming (crates/swarm-dns/examples/
 * p256_full_ecdsa_stark_bench.rs:133-240) for n_proofs chain inputs: trace of
 * n_trace rows (a power of two >= max(k_a, k_b)) -> LDE x32 -> merge -> STIR
 * with 54 queries bound to public_inputs_hashes[32 i ..]. Lockstep device
 * batches in HBM-sized waves; result i equals the proof of input i alone. */
int stark_prove_p256_chains_batch(stark_ctx *ctx, const stark_p256_chains *in, size_t n_proofs, size_t n_trace,
                                  const uint8_t *public_inputs_hashes, stark_p256_out *outs);

/* prove_record_ecdsa (crates/swarm-dns/examples/mixed_algorithm_zone_rollup.rs:
 * 193-322) from the signature on, for n_records records: record i is
 * (qx, qy, r, s, digest) at byte offset 32 i of each array (big-endian).
 * u1 = e/s, u2 = r/s mod n; chain A from G over the MSB-first bits of u1,
 * chain B from Q over those of u2, n_trace 256, r_proj bound to the outputs;
 * STIR bound to SHA3-256("ECDSAP256-MIXEDZONE" || qx || qy || r || s ||
 * digest). STARK_E_INVALID if a key is not a curve point or a signature does
 * not verify (where the example asserts). */
int stark_prove_ecdsa_p256_batch(stark_ctx *ctx, const uint8_t *qx, const uint8_t *qy, const uint8_t *r,
                                 const uint8_t *s, const uint8_t *digest, size_t n_records, stark_p256_out *outs);

/* ---- Ed25519 cofactored-verify AIR v16 (crates/deep_ali/src/ed25519_verify_air.rs) ---- */

/* One ZskKskOutput of prove_zsk_ksk_binding_v2 (crates/swarm-dns/src/prover.rs:856-951). */
typedef struct {
    uint8_t pi_hash[32]; /* zsk_ksk_pi_hash_v2_stark (prover.rs:763-781) */
    uint8_t root_f0[32];
    size_t n_trace;
    uint8_t *proof; /* ark-serialize compressed DeepFriProof<SexticExt>; stark_free() it */
    size_t proof_len;
} stark_zsk_out;

/* verify_air_layout_v16 sizes (ed25519_verify_air.rs:4741-4818) for a signed
 * message of msg_len bytes (the SHA-512 input is R || A || M) and k_scalar
 * ladder bits (256 for Ed25519): trace width, height and
 * verify_v16_per_row_constraints (:4737). */
int stark_ed25519_shape(size_t msg_len, size_t k_scalar, size_t *width, size_t *height, size_t *n_constraints);

/* prove_zsk_ksk_binding_v2 (crates/swarm-dns/src/prover.rs:856-951) for n
 * signatures: signature i is pks[32 i..], sigs[64 i..] (R || s) over
 * msgs[msg_off[i] .. msg_off[i] + msg_len[i]], bound to fs_bindings[32 i..]
 * and merkle_roots[32 i..]. k_scalar 256 and blowup 32 with 54 queries are the
 * reference's; a smaller k_scalar is its layout-test stub, another blowup /
 * query count a parameter choice (the fold schedule follows n0 and ldt). Lockstep
 * device batches per trace shape in HBM-sized waves; result i equals the
 * proof of signature i alone. STARK_E_INVALID when a signature does not
 * verify (the reference asserts). */
int stark_prove_zsk_ksk_binding_v2_batch(stark_ctx *ctx, const uint8_t *pks, const uint8_t *sigs, const uint8_t *msgs,
                                         const size_t *msg_off, const size_t *msg_len, const uint8_t *fs_bindings,
                                         const uint8_t *merkle_roots, size_t n, size_t k_scalar, size_t blowup,
                                         size_t queries, int ldt, stark_zsk_out *outs);

/* deep_ali_merge_ed25519_verify_streaming (crates/deep_ali/src/lib.rs:854-974)
 * for one signature's layout: lde is width x (height * blowup), coeffs the
 * n_constraints combination coefficients (any values). Phi is evaluated at
 * every point; c_eval gets height * blowup values. */
int stark_merge_ed25519(stark_ctx *ctx, const uint8_t pk[32], const uint8_t sig[64], const uint8_t *msg,
                        size_t msg_len, size_t k_scalar, const uint64_t *lde, size_t blowup, const uint64_t *coeffs,
                        size_t n_coeffs, uint64_t *c_eval);

/* fill_verify_air_v16 (ed25519_verify_air.rs:4820-4865) into trace (width x
 * height, column-major): on the device, copied back (stark_build_...) or on
 * the host (stark_host_build_...), from the same witness words. */
int stark_build_ed25519_trace(stark_ctx *ctx, const uint8_t pk[32], const uint8_t sig[64], const uint8_t *msg,
                              size_t msg_len, size_t k_scalar, uint64_t *trace);
int stark_host_build_ed25519_trace(const uint8_t pk[32], const uint8_t sig[64], const uint8_t *msg, size_t msg_len,
                                   size_t k_scalar, uint64_t *trace);
/* ed25519_verify::verify (crates/deep_ali/src/ed25519_verify.rs:54-118): 1 valid, 0 not */
int stark_host_ed25519_verify(const uint8_t pk[32], const uint8_t sig[64], const uint8_t *msg, size_t msg_len);
/* RFC 8032 section 5.1.6 signing from a 32-byte seed (synthetic workloads);
 * no counterpart in the reference's prover path */
void stark_host_ed25519_sign(const uint8_t seed[32], const uint8_t *msg, size_t msg_len, uint8_t pk[32],
                             uint8_t sig[64]);

/* prove_nsec3_completeness (crates/swarm-dns/src/prover.rs:348-446): records
 * = n x (owner_hash[32] || next_hash[32]), a closed chain (records[n-1]'s
 * next_hash equals records[0]'s owner_hash, else STARK_E_INVALID, where the
 * reference panics). */
int stark_prove_nsec3_completeness(stark_ctx *ctx, const uint8_t *records, size_t n, const uint8_t salt[16],
                                   const uint8_t fs_binding[32], int ldt, int phi_mode, stark_nsec3_out *out);
/* n_chains chains in lockstep device batches per trace height: chain i is
 * counts[i] records starting at record sum(counts[0..i)) of records; no
 * single counterpart in the reference (a loop of prove_nsec3_completeness). */
int stark_prove_nsec3_completeness_batch(stark_ctx *ctx, const uint8_t *records, const size_t *counts,
                                         size_t n_chains, const uint8_t salt[16], const uint8_t *fs_bindings,
                                         int ldt, int phi_mode, stark_nsec3_out *outs);

/* prove_outer_rollup (crates/swarm-dns/src/prover.rs:253-304). */
int stark_prove_outer_rollup(stark_ctx *ctx, const uint8_t *pi_hashes, size_t n, const uint8_t pk_hash[32], int ldt,
                             int phi_mode, uint8_t **proof, size_t *proof_len, uint8_t root_f0[32]);

/* ---- proving server ----
 * No counterpart in the reference: it is the device side of the swarm
 * worker's request loop (crates/swarm-worker/src/main.rs:220-335). `workers`
 * threads own one context (HIP stream) each; a free worker admits the head
 * request and every queued request of the same kind (and salt) that fits the
 * HBM budget, up to max_batch, and proves them as one lockstep batch while
 * other workers' batches run on their own streams. Each result equals the
 * single-call result for that request (stark_prove_ds_ksk_binding /
 * stark_prove_inner_shard). submit copies the request's inputs.
 * Isolation: submit rejects (STARK_E_INVALID) a signature request whose
 * signature does not verify; a batch that still fails is proved again one
 * request at a time, so only the failing request's result carries a status
 * and its own error text. A request with deadline_s > 0 that no worker
 * admits within deadline_s completes with STARK_E_TIMEOUT and no proof.
 * stark_server_last_error gives the calling thread's last submit or wait
 * error (thread-local, so concurrent callers do not see each other's). */
int stark_server_create(int device, int workers, size_t hbm_budget, size_t max_batch, int flags, stark_server **out);
void stark_server_destroy(stark_server *s); /* finishes queued requests first */
int stark_server_submit(stark_server *s, const stark_serve_request *req, uint64_t *id);
/* STARK_E_TIMEOUT if the result is not ready within timeout_s */
int stark_server_wait(stark_server *s, uint64_t id, double timeout_s, stark_serve_result *out);
int stark_server_stats(stark_server *s, size_t *completed, size_t *peak_bytes_admitted, size_t *budget);
const char *stark_server_last_error(const stark_server *s);

/* ---- device-resident API (inputs already in HBM) ---- */
int stark_dev_alloc(stark_ctx *ctx, size_t bytes, void **ptr);
int stark_dev_free(stark_ctx *ctx, void *ptr);
int stark_memcpy_h2d(stark_ctx *ctx, void *dst, const void *src, size_t bytes);
int stark_memcpy_d2h(stark_ctx *ctx, void *dst, const void *src, size_t bytes);
int stark_synchronize(stark_ctx *ctx);

/* Radix2EvaluationDomain::{fft, ifft} on ncols columns of 2^log_n values
 * (ark-poly 0.4.2, used at trace_import.rs:272-279, lib.rs:400-410). */
int stark_ntt_dev(stark_ctx *ctx, const uint64_t *d_in, uint64_t *d_out, size_t log_n, size_t ncols, int inverse);

/* The allocation-free variant of stark_ntt_dev. It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops that forbid hipMalloc). The two-pass
 * transform's scratch is one fixed 32 MiB array per device (held until the
 * process ends) instead of an allocation, so log_n <= 24 with
 * 2^log_n * ncols <= 2^22 (else STARK_E_INVALID); the call makes no
 * hipMalloc / hipFree and times no launch plan. Calls that overlap on one
 * device share the array and race. */
int stark_ntt_dev_noalloc(stark_ctx *ctx, const uint64_t *d_in, uint64_t *d_out, size_t log_n, size_t ncols,
                          int inverse);

/* lde_trace_columns on device buffers; shift != 1 evaluates on the coset
 * shift * <w_n0> (coefficient i scaled by shift^i). */
int stark_lde_columns_dev(stark_ctx *ctx, const uint64_t *d_cols, size_t w, size_t n_trace, size_t blowup,
                          uint64_t shift, uint64_t *d_out);

/* Prover core on device buffers: coset LDE (d_lde, w x n0) -> row Merkle tree
 * (leaf = one LDE row, DsLabel arity 2, label 0) -> f0 = sum_c (c+1) LDE_c ->
 * deep_fri_prove(f0) bound to the row root -> one row opening per STIR query.
 * Output bytes: DeepFriProof || u64 s0 || s0 x (Vec<F> row, MerkleOpening). */
int stark_prover_core_dev(stark_ctx *ctx, const uint64_t *d_trace, size_t w, size_t n_trace, size_t blowup,
                          uint64_t shift, const stark_fri_params *params, uint64_t *d_lde, uint8_t row_root[32],
                          uint8_t **proof, size_t *proof_len);

/* ---- profiling: HIP events around every kernel launch on the context stream ---- */
int stark_profile_enable(stark_ctx *ctx, int on);
int stark_profile_reset(stark_ctx *ctx);
/* Accumulated time (ms), launches and algorithmic bytes of kernels named `name`. */
int stark_profile_get(stark_ctx *ctx, const char *name, double *ms, uint64_t *launches, double *bytes);
/* Newline-separated list of kernel names seen; returns the length written. */
size_t stark_profile_names(stark_ctx *ctx, char *buf, size_t cap);

/* ---- launch-plan autotuner (no counterpart in the reference, whose NTTs are
 * ark-poly's; the pass split of a GPU transform has no effect on its output) ----
 * With autotune on, the first NTT / iNTT / LDE of each shape (kind, log2
 * points, log2 blowup, log2 columns) times every admissible split of its two
 * passes on the context stream and the process keeps the fastest for all
 * later calls. ctx = NULL sets the default for contexts created afterwards
 * (the proving server's workers included); STARK_AUTOTUNE=1 in the
 * environment makes it on from the start. A context with it off times
 * nothing: it takes the plans already tuned in the process, else the fixed
 * default splits. */
int stark_set_autotune(stark_ctx *ctx, int on);
/* Pin the first pass of every two-pass transform that admits it to
 * 2^log_first_pass points (3..12; 0 = no pin); used to check each split. */
int stark_set_ntt_split(stark_ctx *ctx, unsigned log_first_pass);
/* Pin the LDS tile of every tunable transform that admits it: 13 (2^13
 * elements, 512 lanes), 14 (2^14 elements, 1024 lanes; passes of 2^8 points
 * and more), 0 = no pin. A plan is a (split, tile) pair; every plan computes
 * the same values. */
int stark_set_ntt_tile(stark_ctx *ctx, unsigned tile_log);
/* The tuned plans, one line per shape: "<ntt|intt|lde> L LB log_cols ->
 * L1.tile | L1.tile:ms ...". Returns the full length; writes at most cap - 1
 * bytes + NUL. */
size_t stark_autotune_report(char *buf, size_t cap);

/* ---- host-side helpers (no device work; used by CPU tests of the host logic) ---- */
void stark_host_sha3_256(const uint8_t *msg, size_t len, uint8_t out[32]);
uint64_t stark_host_stdrng_first_u64(uint64_t seed);
/* DnsRecord::leaf_hash (dns.rs:72-89) */
/* build_sha256_trace_multi (crates/deep_ali/src/sha256_air.rs:429-479):
 * trace (756 x stark_host_sha256_trace_height(len), column-major) of msg;
 * returns n_blocks and the digest read from the trace (prover.rs:529-538). */
size_t stark_host_sha256_trace_height(size_t len);
size_t stark_host_build_sha256_trace(const uint8_t *msg, size_t len, uint64_t *trace, uint8_t digest[32]);
/* build_sha512_trace_multi (crates/deep_ali/src/sha512_air.rs:551-573):
 * 1510 x stark_host_sha512_trace_height(len) trace; digest = 64 bytes. */
size_t stark_host_sha512_trace_height(size_t len);
size_t stark_host_build_sha512_trace(const uint8_t *msg, size_t len, uint64_t *trace, uint8_t digest[64]);
void stark_host_dns_leaf_hash(const stark_dns_record *rec, const uint8_t salt[16], uint8_t out[32]);
/* fill_ecdsa_double_multirow (crates/deep_ali/src/
 * p256_ecdsa_double_multirow_air.rs:134-222): column-major 151334 x n_trace
 * trace; in->rproj is overwritten with the bound values when
 * in->bind_outputs. */
int stark_host_build_ecdsa_p256_trace(stark_p256_chains *in, size_t n_trace, uint64_t *trace);

/* fill_rsa_exp_multirow (crates/deep_ali/src/rsa2048_exp_air.rs:148-202) /
 * fill_rsa_stacked (rsa2048_stacked_air.rs:64-78): column-major trace of
 * 11041 n_records columns x n_trace rows (n_trace a power of two >= 32) for
 * the records (n, s, em) at 256-byte strides. 0, or STARK_E_INVALID for
 * s >= n, n = 0 or a bad height. */
int stark_host_build_rsa_exp_trace(const uint8_t *n_be, const uint8_t *s_be, const uint8_t *em_be, size_t n_records,
                                   size_t n_trace, uint64_t *trace);
/* emsa_pkcs1_v1_5_encode_sha256 (crates/deep_ali/src/rsa2048.rs) for k = 256 */
void stark_host_rsa_pkcs1_em_sha256(const uint8_t digest[32], uint8_t em[256]);

#ifdef __cplusplus
}
#endif
#endif
