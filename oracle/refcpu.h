/*
 * refcpu.h -- public interface of the CPU oracle ("refcpu").
 *
 * THE ORACLE IS TEST INFRASTRUCTURE. It is a plain-C restatement of the
 * stark-swarm CPU prover's hot path, used only as the checker of the HIP
 * product path by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg. Nothing in the shipped library links, loads or calls it.
 *
 * All field values cross this interface as canonical u64 in [0, p),
 * p = 2^64 - 2^32 + 1. Fp6 values are 6 consecutive u64 in the reference's
 * to_fp_components order (c0.c[0..3], c1.c[0..3]).
 */
#ifndef REFCPU_H
#define REFCPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- field / primitive checks ---- */
uint64_t ref_gl_mul(uint64_t a, uint64_t b);
uint64_t ref_gl_inv(uint64_t a);
uint64_t ref_gl_root_of_unity(size_t n);
/* threads for the loops the reference runs with rayon (1 = serial; the
 * results are the same bytes for any count) */
void ref_set_threads(int n);
int ref_get_threads(void);
void ref_fp6_mul(const uint64_t *a, const uint64_t *b, uint64_t *out);
void ref_fp6_inv(const uint64_t *a, uint64_t *out);
void ref_ntt(uint64_t *a, size_t n, int inverse);
void ref_sha3_256(const uint8_t *msg, size_t len, uint8_t out[32]);
void ref_keccak_f1600(uint64_t st[25]);
void ref_chacha_block(const uint8_t key[32], uint32_t counter, const uint8_t nonce[12], int rounds, uint32_t out[16]);
/* StdRng(seed = LE64(seed) || 0^24).next_u64()  (fri.rs:1044-1051) */
uint64_t ref_stdrng_first_u64(uint64_t seed);

/* ---- LDE: trace_import.rs:251-281 (shift == 1); shift != 1 scales coefficient i by shift^i ---- */
int ref_lde_columns(const uint64_t *cols, size_t w, size_t n_trace, size_t blowup, uint64_t shift, uint64_t *out);

/* ---- HashRollup AIR: air_workloads.rs:571-599, merge lib.rs:343-437 ---- */
void ref_build_hash_rollup_trace(size_t n_trace, const uint64_t *leaves, size_t n_leaves, uint64_t *trace /* 4 x n_trace */);
/* phi_mode: 0 = always compute phi (non-parallel build), 1 = rayon with >1 thread
 * (phi == 0 when n0 < 4096), 2 = rayon with 1 thread (phi always 0). */
int ref_merge_hash_rollup(const uint64_t *lde, size_t n_trace, size_t blowup, int phi_mode, uint64_t *c_eval);
/* the merges with caller coefficients (NULL: lambda_j = j + 1) */
int ref_merge_hash_rollup_lam(const uint64_t *lde, size_t n_trace, size_t blowup, int phi_mode, const uint64_t *lam,
                              uint64_t *c_eval);
int ref_merge_nsec3_lam(const uint64_t *lde, size_t n_trace, size_t blowup, int phi_mode, const uint64_t *lam,
                        uint64_t *c_eval);
int ref_merge_sha256_lam(const uint64_t *lde, size_t n_trace, size_t blowup, size_t n_blocks, int phi_mode,
                         const uint64_t *lam, uint64_t *c_eval);
int ref_merge_sha512_lam(const uint64_t *lde, size_t n_trace, size_t blowup, size_t n_blocks, int phi_mode,
                         const uint64_t *lam, uint64_t *c_eval);

/* ---- Merkle (crates/merkle/src/lib.rs) ---- */
/* root of a tree with constant arity, leaves of `vals_per_leaf` base values */
void ref_merkle_root(const uint64_t *leaf_vals, size_t n_leaves, size_t vals_per_leaf,
                     size_t arity, size_t depth, uint64_t label, uint8_t root[32]);

/* ---- DEEP-FRI / STIR prover (fri.rs:1398-2055) ---- */
typedef struct {
    const size_t *schedule;
    size_t n_layers;
    size_t r;
    size_t s0;
    uint64_t seed_z;
    int coeff_commit_final;
    size_t d_final;
    int stir;
    const uint8_t *public_inputs_hash; /* 32 bytes or NULL */
} ref_fri_params;

/* Writes the ark-serialize compressed DeepFriProof<SexticExt>; *out is malloc'd. */
int ref_deep_fri_prove(const uint64_t *f0, size_t n0, const ref_fri_params *p, uint8_t **out, size_t *out_len);
/* Fold trace of the prover for stage-by-stage checks: writes the layer values
 * f_1 .. f_L (Fp6, concatenated) and the layer roots; sizes follow the schedule. */
int ref_fri_layers(const uint64_t *f0, size_t n0, const ref_fri_params *p, uint64_t *layers_out, uint8_t *roots_out, uint8_t root_f0[32]);
void ref_free(void *p);
/* deep_fri_verify (fri.rs:2136-2710) over the serialized proof: 1 accept,
 * 0 reject (also where the reference panics on a malformed proof), -1 when
 * the bytes do not deserialize as DeepFriProof<SexticExt> */
int ref_deep_fri_verify(const ref_fri_params *p, const uint8_t *proof, size_t len);
/* MerkleTreeChannel::verify_opening (merkle lib.rs:754-799) for a tree of
 * constant arity and `depth` levels; opening = serialized MerkleOpening.
 * 1 / 0, -1 if the bytes do not deserialize */
int ref_merkle_verify_opening(size_t arity, size_t depth, uint64_t label, const uint8_t root[32],
                              const uint8_t *opening, size_t len);
/* MerkleTreeChannel::open (merkle lib.rs:478-512) of leaf `index` of the
 * tree of ref_merkle_root; *out (malloc'd) = serialized MerkleOpening */
int ref_merkle_open(const uint64_t *leaf_vals, size_t n_leaves, size_t vals_per_leaf, size_t arity, size_t depth,
                    uint64_t label, size_t index, uint8_t **out, size_t *out_len);

/* ---- swarm-dns prover (crates/swarm-dns/src/prover.rs, dns.rs) ---- */
/* Synthetic zone record i (swarm-ctrl main.rs:1331-1357), leaf hash h2 (dns.rs:72-89). */
void ref_zone_leaf_hashes(uint64_t first, size_t count, const uint8_t salt[16], uint8_t *out /* count x 32 */);
/* Canonical bytes of synthetic record i; returns length (buffer >= 256 bytes). */
size_t ref_zone_record_canonical(uint64_t i, uint8_t *out);
void ref_dns_merkle_root(const uint8_t *leaves, size_t n, uint8_t root[32]);
void ref_shard_fs_binding(const uint8_t pk_hash[32], const uint8_t job_id[32], uint32_t shard_id, const uint8_t nonce[32], uint8_t out[32]);
/* prove_inner_shard over precomputed leaf hashes; returns proof bytes (malloc'd). */
int ref_prove_inner_shard(const uint8_t salt[16], const uint8_t *leaf_hashes, size_t count,
                          const uint8_t fs_binding[32], int phi_mode,
                          uint8_t **proof, size_t *proof_len, uint8_t pi_hash[32], uint8_t merkle_root[32], uint8_t root_f0[32]);
int ref_prove_outer_rollup(const uint8_t *pi_hashes, size_t count, const uint8_t pk_hash[32], int phi_mode,
                           uint8_t **proof, size_t *proof_len, uint8_t root_f0[32]);

/* c <- FFT(pad(poly_div_zh(iFFT(c), n_trace))), remainder dropped (lib.rs:190-228, 421-429) */
void ref_quotient_from_phi(uint64_t *phi, size_t n, size_t n_trace);
/* LdtMode of the swarm-dns provers (prove_inner_shard, prove_outer_rollup,
 * prove_nsec3_completeness, prove_ds_ksk_binding, prove_zsk_ksk_binding_v2):
 * 0 Stir (the default), 1 Fri (classic DEEP-FRI with arity-2 folds) */
void ref_set_ldt(int fri);
int ref_ldt_stir(void);
/* make_schedule(n0, ldt): Stir = 8s then the 2^(log n0 mod 3) residual, Fri = log2(n0) 2s; returns the layer count */
size_t ref_make_schedule(size_t n0, size_t *s);
size_t ref_stir_schedule(size_t n0, size_t *s); /* make_schedule(n0, Stir) whatever ref_set_ldt says */

/* ---- SHA-256 AIR and prove_ds_ksk_binding (sha256_air.rs, lib.rs:453-545, prover.rs:446-585) ---- */
size_t ref_sha256_n_blocks(size_t len);
size_t ref_sha256_trace_height(size_t len);
/* 756 x height column-major trace of the padded message; returns n_blocks and
 * the digest read from the trace (H-state at block row 65 of the last block). */
int ref_build_sha256_trace(const uint8_t *msg, size_t len, uint64_t *tr, uint8_t digest[32]);
/* the 766 constraint values at one row (cur, nxt: 756 values each) */
void ref_sha256_constraints(const uint64_t *cur, const uint64_t *nxt, size_t row, size_t n_blocks, uint64_t *out);
int ref_merge_sha256(const uint64_t *lde, size_t n_trace, size_t blowup, size_t n_blocks, int phi_mode,
                     uint64_t *c_eval);
int ref_prove_ds_ksk_binding(const uint8_t *dnskey, size_t len, const uint8_t parent_ds_hash[32],
                             const uint8_t fs_binding[32], int phi_mode, uint8_t **proof, size_t *proof_len,
                             uint8_t pi_hash[32], uint8_t asserted_digest[32], uint8_t root_f0[32],
                             size_t *n_blocks_out, size_t *n_trace_out);

/* ---- SHA-512 AIR (sha512_air.rs, lib.rs:561-653) ---- */
size_t ref_sha512_n_blocks(size_t len);
size_t ref_sha512_trace_height(size_t len);
/* 1510 x height trace; returns n_blocks, digest = the 64-byte digest read from the trace */
int ref_build_sha512_trace(const uint8_t *msg, size_t len, uint64_t *tr, uint8_t digest[64]);
void ref_sha512_constraints(const uint64_t *cur, const uint64_t *nxt, size_t row, size_t n_blocks, uint64_t *out);
int ref_merge_sha512(const uint64_t *lde, size_t n_trace, size_t blowup, size_t n_blocks, int phi_mode,
                     uint64_t *c_eval);

/* ---- NSEC3 chain (AirType::Nsec3Chain) and prove_nsec3_completeness ---- */
/* chain: n records x 8 raw u64 (owner limbs, next limbs) */
void ref_build_nsec3_trace(size_t n_trace, const uint64_t *chain, size_t n, uint64_t *tr);
/* eval_nsec3_chain_constraints (air_workloads.rs:862-870): C_k = nxt[k] - cur[4 + k] */
void ref_nsec3_constraints(const uint64_t cur[8], const uint64_t nxt[8], uint64_t out[4]);
int ref_merge_nsec3(const uint64_t *lde, size_t n_trace, size_t blowup, int phi_mode, uint64_t *c_eval);
/* records: n x (owner_hash[32] || next_hash[32]); -2 if not a closed chain */
int ref_prove_nsec3_completeness(const uint8_t *records, size_t n, const uint8_t salt[16], const uint8_t fs_binding[32],
                                 int phi_mode, uint8_t **proof, size_t *proof_len, uint8_t pi_hash[32],
                                 uint8_t chain_root[32], uint8_t root_f0[32], size_t *n_trace_out);

/* ---- RSA-2048 exponentiation AIR (s^65537 mod n; rsa2048_exp_air.rs) ---- */
size_t ref_rsa_exp_width(void);                  /* 11 041 columns per record */
size_t ref_rsa_exp_constraints_per_record(void); /* 11 280 */
/* record at columns [base, base + 11041) of a column-major trace of ht rows;
 * n, s, em: 256-byte big-endian, s < n. 0 or a negative code. */
int ref_fill_rsa_exp(const uint8_t n_be[256], const uint8_t s_be[256], const uint8_t em_be[256], size_t ht,
                     size_t base, uint64_t *tr);
void ref_rsa_exp_constraints(const uint64_t *cur, const uint64_t *nxt, size_t trace_row, size_t n_trace,
                             uint64_t *out);
/* n_records stacked records; coeffs NULL = lambda_j = j + 1 */
int ref_merge_rsa(const uint64_t *lde, size_t n_trace, size_t blowup, size_t n_records, const uint64_t *coeffs,
                  uint64_t *c_eval);
void ref_rsa_pkcs1_em_sha256(const uint8_t digest[32], uint8_t em[256]);
int ref_prove_rsa_exp(const uint8_t n_be[256], const uint8_t s_be[256], const uint8_t em_be[256],
                      const uint8_t pis[32], uint8_t **proof, size_t *proof_len, uint8_t root_f0[32],
                      uint8_t pi_hash[32]);

/* ---- double-chain ECDSA-P256 AIR (p256_ecdsa_double_multirow_air.rs) ---- */
size_t ref_p256_width(void);       /* 151 334 columns */
size_t ref_p256_n_constraints(void); /* 155 774 */
/* column bases: [0] bit A, [1] bit B, [2..7] select outputs of A then B,
 * [8..13] r_proj, [14..19] acc of A then B, [20..25] base of A then B */
void ref_p256_columns(size_t out[26]);
/* sub-gadgets of chain ch (0 = A, 1 = B) in constraint order: kinds (0 mul,
 * 1 add, 2 sub, 3 freeze, 4 select) and 6 column bases each (c, c bits,
 * q / diff limbs, q / diff bits, carries / pos, neg); returns the count */
int ref_p256_gadgets(int ch, int *kinds, size_t *cols);
/* fill_ecdsa_double_multirow: pts = 12 x 32-byte BE values (a_initial xyz,
 * a_base xyz, b_initial xyz, b_base xyz), rproj = 6 x 32 bytes, bits 0/1 */
int ref_fill_p256(const uint8_t *pts, const uint8_t *rproj, const uint8_t *a_bits, size_t k_a, const uint8_t *b_bits,
                  size_t k_b, size_t ht, uint64_t *tr);
void ref_p256_constraints(const uint64_t *cur, const uint64_t *nxt, size_t trace_row, size_t n_trace, uint64_t *out);
int ref_merge_p256(const uint64_t *lde, size_t n_trace, size_t blowup, const uint64_t *coeffs, uint64_t *c_eval);
/* ECDSA record (qx, qy, r, s, digest: 32-byte BE): MSB-first bits of
 * u1 = e/s and u2 = r/s mod n; -30 bad signature, -31 bad key */
int ref_p256_ecdsa_bits(const uint8_t qx[32], const uint8_t qy[32], const uint8_t r[32], const uint8_t s[32],
                        const uint8_t digest[32], uint8_t u1_bits[256], uint8_t u2_bits[256]);
int ref_p256_ecdsa_trace(const uint8_t qx[32], const uint8_t qy[32], const uint8_t r[32], const uint8_t s[32],
                         const uint8_t digest[32], size_t ht, uint64_t *tr, uint8_t rproj[6 * 32]);
int ref_prove_p256_ecdsa(const uint8_t qx[32], const uint8_t qy[32], const uint8_t r[32], const uint8_t s[32],
                         const uint8_t digest[32], uint8_t **proof, size_t *proof_len, uint8_t root_f0[32],
                         uint8_t pi_hash[32]);

/* ---- Ed25519 cofactored-verify AIR v16 and prove_zsk_ksk_binding_v2 (ed25519_verify_air.rs) ---- */
typedef struct ref_ed_ctx ref_ed_ctx;
/* public inputs of one signature (pk, sig = R || s, message); k_scalar = k (256 in production).
 * NULL if R or A does not decompress. */
ref_ed_ctx *ref_ed25519_ctx_new(const uint8_t pk[32], const uint8_t sig[64], const uint8_t *msg, size_t len,
                                size_t k);
void ref_ed25519_ctx_free(ref_ed_ctx *c);
size_t ref_ed25519_width(const ref_ed_ctx *c);
size_t ref_ed25519_height(const ref_ed_ctx *c);
size_t ref_ed25519_n_constraints(const ref_ed_ctx *c);
size_t ref_ed25519_v16_constraints(size_t k); /* verify_v16_per_row_constraints */
void ref_ed25519_rows(const ref_ed_ctx *c, size_t out[11]);
void ref_ed25519_columns(const ref_ed_ctx *c, size_t out[26]);
void ref_ed25519_public(const ref_ed_ctx *c, uint8_t digest[64], uint8_t k_can[32], uint8_t *s_bits,
                        uint8_t *k_bits);
void ref_ed25519_points(const ref_ed_ctx *c, uint8_t out[6 * 128]);
int ref_ed25519_fill(const ref_ed_ctx *c, uint64_t *tr /* width x height */);
void ref_ed25519_constraints(const ref_ed_ctx *c, const uint64_t *cur, const uint64_t *nxt, size_t row,
                             uint64_t *out);
int ref_merge_ed25519(const ref_ed_ctx *c, const uint64_t *lde, size_t blowup, const uint64_t *coeffs,
                      uint64_t *c_eval);
int ref_ed25519_verify(const uint8_t pk[32], const uint8_t sig[64], const uint8_t *msg, size_t len);
void ref_sha512(const uint8_t *msg, size_t len, uint8_t out[64]);
int ref_prove_zsk_ksk_binding_v2(const uint8_t pk[32], const uint8_t sig[64], const uint8_t *msg, size_t len,
                                 const uint8_t fs_binding[32], const uint8_t merkle_root[32], size_t k,
                                 size_t blowup, uint8_t **proof, size_t *proof_len, uint8_t root_f0[32],
                                 uint8_t pi_hash[32]);

/* ---- prover-core microbench (coset LDE -> row Merkle -> STIR prove) ---- */
/* trace: w x n_trace column-major; returns proof bytes of deep_fri_prove over
 * f0 = sum_c (c+1) * LDE_c with the row-tree root as public_inputs_hash, and
 * the row root. lde_out (w x n0) may be NULL. */
int ref_prover_core(const uint64_t *trace, size_t w, size_t n_trace, size_t blowup, uint64_t shift,
                    const ref_fri_params *p, uint8_t row_root[32], uint64_t *f0_out,
                    uint8_t **proof, size_t *proof_len);

#ifdef __cplusplus
}
#endif
#endif
