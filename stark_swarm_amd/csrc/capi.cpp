// capi.cpp -- the extern "C" boundary declared in include/stark_gpu.h.
// Every entry point catches library errors and turns them into status codes.
#include "../../include/stark_gpu.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "air.hpp"
#include "ctx.hpp"
#include "ed25519_air.hpp"
#include "fri.hpp"
#include "gl.hpp"
#include "keccak.hpp"
#include "ntt.hpp"
#include "p256_air.hpp"
#include "prover.hpp"
#include "rsa_air.hpp"
#include "server.hpp"
#include "sha256_air.hpp"
#include "verify.hpp"

struct stark_ctx {
    sg::Ctx *c = nullptr;
    std::string err;
};

struct stark_server {
    sg::Server *s = nullptr;
};

// last submit / wait error of the calling thread (callers submit from many threads)
static thread_local std::string server_err;

namespace {

template <class F>
int guard(stark_ctx *ctx, F &&f) {
    if (!ctx || !ctx->c) return STARK_E_INVALID;
    try {
        hipError_t e = hipSetDevice(ctx->c->device);
        if (e != hipSuccess) throw sg::Error(sg::E_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e));
        f(ctx->c);
        if (ctx->c->profiling) ctx->c->prof_collect();
        return STARK_OK;
    } catch (const sg::Error &e) {
        ctx->err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        ctx->err = "host allocation failed";
        return STARK_E_OOM;
    } catch (const std::exception &e) {
        ctx->err = e.what();
        return STARK_E_INTERNAL;
    }
}

sg::FriParams to_params(const stark_fri_params *p) {
    SG_REQUIRE(p != nullptr, "null fri params");
    SG_REQUIRE(p->schedule != nullptr || p->n_layers == 0, "null schedule");
    sg::FriParams q;
    q.schedule.assign(p->schedule, p->schedule + p->n_layers);
    q.r = p->r;
    q.s0 = p->s0;
    q.seed_z = p->seed_z;
    q.coeff_commit_final = p->coeff_commit_final != 0;
    q.d_final = p->d_final;
    q.stir = p->stir != 0;
    q.has_pi = p->public_inputs_hash != nullptr;
    if (q.has_pi) memcpy(q.pi, p->public_inputs_hash, 32);
    return q;
}

uint8_t *copy_out(const std::vector<uint8_t> &v, size_t *len) {
    uint8_t *p = (uint8_t *)malloc(v.size() ? v.size() : 1);
    if (!p) throw std::bad_alloc();
    memcpy(p, v.data(), v.size());
    *len = v.size();
    return p;
}

void h2d(sg::Ctx *c, void *d, const void *h, size_t bytes) {
    SG_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream));
}
void d2h(sg::Ctx *c, void *h, const void *d, size_t bytes) {
    c->d2h(h, d, bytes);
    SG_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace

extern "C" {

int stark_ctx_create(int device, stark_ctx **out) {
    if (!out) return STARK_E_INVALID;
    *out = nullptr;
    stark_ctx *x = new (std::nothrow) stark_ctx();
    if (!x) return STARK_E_OOM;
    try {
        x->c = sg::ctx_create(device);
    } catch (const sg::Error &e) {
        delete x;
        return e.code;
    } catch (...) {
        delete x;
        return STARK_E_INTERNAL;
    }
    *out = x;
    return STARK_OK;
}

void stark_ctx_destroy(stark_ctx *ctx) {
    if (!ctx) return;
    sg::ctx_destroy(ctx->c);
    delete ctx;
}

const char *stark_last_error(const stark_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void stark_free(void *p) { free(p); }

int stark_lde_columns(stark_ctx *ctx, const uint64_t *cols, size_t w, size_t n_trace, size_t blowup, uint64_t *out) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(cols && out && w > 0, "lde: null buffer or no columns");
        SG_REQUIRE(sg::is_pow2(n_trace), "lde: trace_len must be a power of 2");
        SG_REQUIRE(blowup >= 2 && sg::is_pow2(blowup), "lde: blowup must be a power-of-2 >= 2");
        const size_t n0 = n_trace * blowup;
        sg::Scoped din(c, 8 * w * n_trace), dout(c, 8 * w * n0);
        h2d(c, din.p, cols, 8 * w * n_trace);
        sg::lde_columns(c, din.as<uint64_t>(), dout.as<uint64_t>(), sg::ilog2(n_trace), sg::ilog2(blowup), w, n_trace,
                        n0, 1);
        d2h(c, out, dout.p, 8 * w * n0);
    });
}

// true when coeffs are the reference's own lambda_j = j + 1 (the kernels'
// built-in weights); otherwise the caller's table goes to the device
static bool reference_lambdas(const uint64_t *coeffs, size_t k) {
    for (size_t j = 0; j < k; j++)
        if (coeffs[j] != j + 1) return false;
    return true;
}

int stark_merge(stark_ctx *ctx, int air_id, const uint64_t *lde, size_t w, size_t n_trace, size_t blowup,
                const uint64_t *coeffs, size_t k, int phi_mode, uint64_t *c_eval) {
    return guard(ctx, [&](sg::Ctx *c) {
        if (air_id != STARK_AIR_HASH_ROLLUP && air_id != STARK_AIR_NSEC3_CHAIN)
            throw sg::Error(sg::E_UNSUPPORTED, "merge: only the HashRollup and Nsec3Chain AIRs");
        const size_t ww = air_id == STARK_AIR_HASH_ROLLUP ? 4 : 8, kk = air_id == STARK_AIR_HASH_ROLLUP ? 3 : 4;
        SG_REQUIRE(w == ww && k == kk && lde && coeffs && c_eval, "merge: wrong trace width or constraint count");
        SG_REQUIRE(phi_mode >= 0 && phi_mode <= 2, "merge: bad phi_mode");
        const uint64_t *lam = reference_lambdas(coeffs, k) ? nullptr : coeffs;
        const size_t n0 = n_trace * blowup;
        sg::Scoped dl(c, 8 * ww * n0), dc(c, 8 * n0);
        h2d(c, dl.p, lde, 8 * ww * n0);
        if (air_id == STARK_AIR_HASH_ROLLUP)
            sg::merge_hash_rollup_dev(c, dl.as<uint64_t>(), n_trace, blowup, phi_mode, dc.as<uint64_t>(), lam);
        else
            sg::merge_nsec3_batch(c, dl.as<uint64_t>(), 1, n_trace, blowup, phi_mode, dc.as<uint64_t>(), lam);
        d2h(c, c_eval, dc.p, 8 * n0);
    });
}

int stark_deep_fri_prove(stark_ctx *ctx, const uint64_t *f0, size_t n0, const stark_fri_params *params,
                         uint8_t **proof, size_t *proof_len) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(f0 && proof && proof_len, "deep_fri_prove: null buffer");
        sg::FriParams p = to_params(params);
        SG_REQUIRE(sg::is_pow2(n0), "deep_fri_prove: n0 must be a power of two");
        sg::Scoped d(c, 8 * n0);
        h2d(c, d.p, f0, 8 * n0);
        std::vector<uint8_t> v = sg::deep_fri_prove_dev(c, d.as<uint64_t>(), n0, p, nullptr);
        *proof = copy_out(v, proof_len);
    });
}

int stark_merkle_commit(stark_ctx *ctx, const uint64_t *leaf_vals, size_t n_leaves, size_t vals_per_leaf,
                        size_t arity, uint64_t tree_label, uint8_t root[32]) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(leaf_vals && root && n_leaves >= 1 && arity >= 2 && arity <= 128, "merkle_commit: bad arguments");
        sg::Scoped d(c, 8 * n_leaves * vals_per_leaf + 8);
        h2d(c, d.p, leaf_vals, 8 * n_leaves * vals_per_leaf);
        sg::DevTree t;
        sg::LeafSrc src{d.as<uint64_t>(), vals_per_leaf, 1, (uint32_t)vals_per_leaf, (uint32_t)vals_per_leaf};
        sg::tree_build(c, t, src, n_leaves, (uint32_t)arity, sg::merkle_depth(n_leaves, (uint32_t)arity), tree_label);
        memcpy(root, t.root, 32);
        sg::tree_free(c, t);
    });
}

int stark_merkle_open(stark_ctx *ctx, const uint64_t *leaf_vals, size_t n_leaves, size_t vals_per_leaf, size_t arity,
                      uint64_t tree_label, const size_t *indices, size_t n_idx, uint8_t root[32], uint8_t **openings,
                      size_t *openings_len, size_t *offsets) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(leaf_vals && root && openings && openings_len && offsets && (indices || n_idx == 0) &&
                       n_leaves >= 1 && arity >= 2 && arity <= 128 && vals_per_leaf >= 1,
                   "merkle_open: bad arguments");
        std::vector<std::vector<uint64_t>> idx(1);
        for (size_t q = 0; q < n_idx; q++) {
            SG_REQUIRE(indices[q] < n_leaves, "merkle_open: leaf index out of range");
            idx[0].push_back(indices[q]);
        }
        sg::Scoped d(c, 8 * n_leaves * vals_per_leaf);
        h2d(c, d.p, leaf_vals, 8 * n_leaves * vals_per_leaf);
        sg::DevTree t;
        sg::LeafSrc src{d.as<uint64_t>(), vals_per_leaf, 1, (uint32_t)vals_per_leaf, (uint32_t)vals_per_leaf};
        sg::tree_build(c, t, src, n_leaves, (uint32_t)arity, sg::merkle_depth(n_leaves, (uint32_t)arity), tree_label);
        memcpy(root, t.root, 32);
        std::vector<std::vector<std::vector<uint8_t>>> ops;
        if (n_idx) sg::tree_open_multi(c, t, idx, ops);
        sg::tree_free(c, t);
        std::vector<uint8_t> all;
        offsets[0] = 0;
        for (size_t q = 0; q < n_idx; q++) {
            all.insert(all.end(), ops[0][q].begin(), ops[0][q].end());
            offsets[q + 1] = all.size();
        }
        *openings = copy_out(all, openings_len);
    });
}

int stark_merkle_verify_opening(size_t arity, size_t depth, uint64_t tree_label, const uint8_t root[32],
                                const uint8_t *opening, size_t len) {
    if (!root || !opening || arity < 2 || arity > 128 || depth > 64) return STARK_E_INVALID;
    try {
        return sg::merkle_verify_opening((uint32_t)arity, (uint32_t)depth, tree_label, root, opening, len);
    } catch (...) {
        return STARK_E_INTERNAL;
    }
}

void stark_host_merkle_leaf_hash(size_t arity, uint64_t tree_label, size_t index, const uint64_t *vals, size_t n,
                                 uint8_t out[32]) {
    sg::host_leaf_hash((uint32_t)arity, tree_label, index, vals, n, out);
}

int stark_deep_fri_verify(const stark_fri_params *params, const uint8_t *proof, size_t proof_len) {
    if (!proof) return STARK_E_INVALID;
    try {
        return sg::deep_fri_verify(to_params(params), proof, proof_len);
    } catch (const sg::Error &e) {
        return e.code;
    } catch (const std::bad_alloc &) {
        return STARK_E_OOM;
    } catch (...) {
        return STARK_E_INTERNAL;
    }
}

int stark_deep_fri_verify_batch(const stark_fri_params *params, const uint8_t *const *proofs, const size_t *lens,
                                size_t n, const uint8_t *public_inputs_hashes, int *results) {
    if ((n && (!proofs || !lens || !results))) return STARK_E_INVALID;
    try {
        const sg::FriParams p = to_params(params);
        for (size_t i = 0; i < n; i++)
            if (!proofs[i]) return STARK_E_INVALID;
        sg::deep_fri_verify_batch(p, proofs, lens, n, public_inputs_hashes, results);
        return STARK_OK;
    } catch (const sg::Error &e) {
        return e.code;
    } catch (...) {
        return STARK_E_INTERNAL;
    }
}

int stark_prove_inner_shard(stark_ctx *ctx, const uint8_t salt[16], const stark_dns_record *records, size_t n,
                            const uint8_t fs_binding[32], int ldt, int phi_mode, stark_inner_shard_out *out) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(salt && records && fs_binding && out, "prove_inner_shard: null argument");
        SG_REQUIRE(ldt == STARK_LDT_STIR || ldt == STARK_LDT_FRI, "unknown LdtMode");
        std::vector<sg::DnsRec> recs(n);
        for (size_t i = 0; i < n; i++)
            recs[i] = {records[i].domain, records[i].domain_len, records[i].record_type, records[i].ttl,
                       records[i].rdata, records[i].rdata_len};
        sg::InnerShard s = sg::prove_inner_shard(c, salt, recs, fs_binding, phi_mode, ldt == STARK_LDT_STIR);
        memcpy(out->pi_hash, s.pi_hash, 32);
        memcpy(out->merkle_root, s.merkle_root, 32);
        memcpy(out->root_f0, s.root_f0, 32);
        out->record_count = s.record_count;
        out->n_trace = s.n_trace;
        out->proof = copy_out(s.proof, &out->proof_len);
    });
}

int stark_prove_inner_shard_batch(stark_ctx *ctx, const uint8_t salt[16], const stark_dns_record *records,
                                  const size_t *shard_sizes, size_t n_shards, const uint8_t *fs_bindings, int ldt,
                                  int phi_mode, stark_inner_shard_out *outs) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(salt && shard_sizes && fs_bindings && outs && (records || n_shards == 0),
                   "prove_inner_shard_batch: null argument");
        SG_REQUIRE(ldt == STARK_LDT_STIR || ldt == STARK_LDT_FRI, "unknown LdtMode");
        std::vector<std::vector<sg::DnsRec>> shards(n_shards);
        std::vector<std::array<uint8_t, 32>> fsb(n_shards);
        size_t off = 0;
        for (size_t s = 0; s < n_shards; s++) {
            for (size_t i = 0; i < shard_sizes[s]; i++) {
                const stark_dns_record &r = records[off + i];
                shards[s].push_back({r.domain, r.domain_len, r.record_type, r.ttl, r.rdata, r.rdata_len});
            }
            off += shard_sizes[s];
            memcpy(fsb[s].data(), fs_bindings + 32 * s, 32);
        }
        std::vector<sg::InnerShard> res = sg::prove_inner_shard_batch(c, salt, shards, fsb, phi_mode, ldt == STARK_LDT_STIR);
        for (size_t s = 0; s < n_shards; s++) outs[s].proof = nullptr;
        for (size_t s = 0; s < n_shards; s++) {
            memcpy(outs[s].pi_hash, res[s].pi_hash, 32);
            memcpy(outs[s].merkle_root, res[s].merkle_root, 32);
            memcpy(outs[s].root_f0, res[s].root_f0, 32);
            outs[s].record_count = res[s].record_count;
            outs[s].n_trace = res[s].n_trace;
            outs[s].proof = copy_out(res[s].proof, &outs[s].proof_len);
        }
    });
}

// one SHA-256 / SHA-512 merge: lde (or, from_trace, the trace) and c_eval in
// device memory; coeffs in host memory
static void sha_merge_dev(sg::Ctx *c, int alg, bool from_trace, const uint64_t *d_in, size_t n_trace, size_t blowup,
                          size_t n_blocks, const uint64_t *coeffs, size_t k, int phi_mode, uint64_t *d_c_eval) {
    const char *what = alg == 256 ? "merge_sha256" : "merge_sha512";
    const size_t K = alg == 256 ? sg::SHA256_K : sg::SHA512_K;
    SG_REQUIRE(d_in && coeffs && d_c_eval, std::string(what) + ": null argument");
    SG_REQUIRE(k == K, std::string(what) + ": need one combination coefficient per constraint");
    SG_REQUIRE(phi_mode >= 0 && phi_mode <= 2, std::string(what) + ": bad phi_mode");
    SG_REQUIRE(n_blocks >= 1 && n_blocks <= 0xFFFFFFFFu, std::string(what) + ": bad n_blocks");
    const bool ref = reference_lambdas(coeffs, k);
    sg::Scoped dk(c, ref ? 8 : 8 * k);
    if (!ref) h2d(c, dk.p, coeffs, 8 * k);
    const uint64_t *lam = ref ? nullptr : dk.as<uint64_t>();
    const std::vector<uint32_t> nb = {(uint32_t)n_blocks};
    if (alg == 256 && from_trace)
        sg::merge_sha256_from_trace_batch(c, d_in, 1, n_trace, blowup, nb, phi_mode, d_c_eval, lam);
    else if (alg == 256)
        sg::merge_sha256_batch(c, d_in, 1, n_trace, blowup, nb, phi_mode, d_c_eval, lam);
    else if (from_trace)
        sg::merge_sha512_from_trace_batch(c, d_in, 1, n_trace, blowup, nb, phi_mode, d_c_eval, lam);
    else
        sg::merge_sha512_batch(c, d_in, 1, n_trace, blowup, nb, phi_mode, d_c_eval, lam);
    SG_HIP(hipStreamSynchronize(c->stream));  // the coefficient table goes out of scope
}

static int sha_merge_host(stark_ctx *ctx, int alg, const uint64_t *lde, size_t n_trace, size_t blowup,
                          size_t n_blocks, const uint64_t *coeffs, size_t k, int phi_mode, uint64_t *c_eval) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(lde && c_eval, "merge_sha: null argument");
        const size_t W = alg == 256 ? sg::SHA256_W : sg::SHA512_W, n0 = n_trace * blowup;
        sg::Scoped dl(c, 8ULL * W * n0), dc(c, 8 * n0);
        h2d(c, dl.p, lde, 8ULL * W * n0);
        sha_merge_dev(c, alg, false, dl.as<uint64_t>(), n_trace, blowup, n_blocks, coeffs, k, phi_mode,
                      dc.as<uint64_t>());
        d2h(c, c_eval, dc.p, 8 * n0);
    });
}

int stark_merge_sha256(stark_ctx *ctx, const uint64_t *lde, size_t n_trace, size_t blowup, size_t n_blocks,
                       const uint64_t *coeffs, size_t k, int phi_mode, uint64_t *c_eval) {
    return sha_merge_host(ctx, 256, lde, n_trace, blowup, n_blocks, coeffs, k, phi_mode, c_eval);
}

int stark_merge_sha512(stark_ctx *ctx, const uint64_t *lde, size_t n_trace, size_t blowup, size_t n_blocks,
                       const uint64_t *coeffs, size_t k, int phi_mode, uint64_t *c_eval) {
    return sha_merge_host(ctx, 512, lde, n_trace, blowup, n_blocks, coeffs, k, phi_mode, c_eval);
}

int stark_merge_sha_dev(stark_ctx *ctx, int alg, const uint64_t *d_lde, size_t n_trace, size_t blowup,
                        size_t n_blocks, const uint64_t *coeffs, size_t k, int phi_mode, uint64_t *d_c_eval) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(alg == 256 || alg == 512, "merge_sha_dev: alg 256 or 512");
        sha_merge_dev(c, alg, false, d_lde, n_trace, blowup, n_blocks, coeffs, k, phi_mode, d_c_eval);
    });
}

int stark_merge_sha_from_trace_dev(stark_ctx *ctx, int alg, const uint64_t *d_trace, size_t n_trace, size_t blowup,
                                   size_t n_blocks, const uint64_t *coeffs, size_t k, int phi_mode,
                                   uint64_t *d_c_eval) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(alg == 256 || alg == 512, "merge_sha_from_trace_dev: alg 256 or 512");
        SG_REQUIRE(blowup >= 4, "merge_sha_from_trace_dev: blowup >= 4");
        sha_merge_dev(c, alg, true, d_trace, n_trace, blowup, n_blocks, coeffs, k, phi_mode, d_c_eval);
    });
}

int stark_prove_sha_batch(stark_ctx *ctx, const int *algs, const uint8_t *msgs, const size_t *offsets, const size_t *lens,
                          const size_t *heights, size_t n, const uint8_t *fs_bindings, int phi_mode,
                          stark_sha_out *outs) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(algs && offsets && lens && fs_bindings && outs, "prove_sha_batch: null argument");
        SG_REQUIRE(phi_mode >= 0 && phi_mode <= 2, "prove_sha_batch: bad phi_mode");
        std::vector<sg::ByteSpan> m(n);
        std::vector<uint64_t> h(n, 0);
        std::vector<std::array<uint8_t, 32>> fsb(n);
        for (size_t i = 0; i < n; i++) {
            SG_REQUIRE(msgs || lens[i] == 0, "prove_sha_batch: null messages");
            m[i] = {msgs ? msgs + offsets[i] : nullptr, lens[i]};
            if (heights) h[i] = heights[i];
            memcpy(fsb[i].data(), fs_bindings + 32 * i, 32);
            outs[i].proof = nullptr;
        }
        std::vector<sg::ShaProof> r = sg::prove_sha_batch(c, std::vector<int>(algs, algs + n), m, h, fsb, phi_mode);
        for (size_t i = 0; i < n; i++) {
            memcpy(outs[i].digest, r[i].digest, 64);
            memcpy(outs[i].root_f0, r[i].root_f0, 32);
            outs[i].n_blocks = r[i].n_blocks;
            outs[i].n_trace = r[i].n_trace;
            outs[i].proof = copy_out(r[i].proof, &outs[i].proof_len);
        }
    });
}

static std::vector<sg::RsaRecord> rsa_records(const uint8_t *n_be, const uint8_t *s_be, const uint8_t *em_be,
                                               size_t count) {
    SG_REQUIRE(n_be && s_be && em_be, "rsa: null record bytes");
    SG_REQUIRE(count >= 1 && count <= 65535, "rsa: 1 to 65535 records");
    std::vector<sg::RsaRecord> r(count);
    for (size_t i = 0; i < count; i++) {
        memcpy(r[i].n.data(), n_be + 256 * i, 256);
        memcpy(r[i].s.data(), s_be + 256 * i, 256);
        memcpy(r[i].em.data(), em_be + 256 * i, 256);
    }
    return r;
}

int stark_merge_rsa_exp(stark_ctx *ctx, const uint64_t *lde, size_t n_trace, size_t blowup, size_t n_records,
                        const uint64_t *coeffs, size_t k, uint64_t *c_eval) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(lde && coeffs && c_eval, "merge_rsa_exp: null argument");
        SG_REQUIRE(n_records >= 1 && n_records <= sg::RSA_MAX_RECORDS, "merge_rsa_exp: 1 to 4096 records");
        SG_REQUIRE(k == sg::RSA_K * n_records, "merge_rsa_exp: need one coefficient per constraint (11280 per record)");
        SG_REQUIRE(n_trace <= (1u << 20) && blowup <= (1u << 20), "merge_rsa_exp: sizes out of range");
        const size_t n0 = n_trace * blowup, w = (size_t)sg::RSA_W * n_records;
        const bool ref = reference_lambdas(coeffs, k);
        sg::Scoped dl(c, 8ULL * w * n0), dc(c, 8 * n0), dk(c, ref ? 8 : 8 * k);
        h2d(c, dl.p, lde, 8ULL * w * n0);
        if (!ref) h2d(c, dk.p, coeffs, 8 * k);
        sg::merge_rsa_batch(c, dl.as<uint64_t>(), 1, (uint32_t)n_records, n_trace, blowup, dc.as<uint64_t>(),
                            ref ? nullptr : dk.as<uint64_t>());
        d2h(c, c_eval, dc.p, 8 * n0);
    });
}

static sg::P256Chains p256_chains(const stark_p256_chains &x) {
    sg::P256Chains c;
    SG_REQUIRE((x.a_bits || !x.k_a) && (x.b_bits || !x.k_b), "p256: null bits");
    SG_REQUIRE(x.k_a <= (1u << 16) && x.k_b <= (1u << 16), "p256: too many bits");
    for (int i = 0; i < 12; i++) memcpy(c.pts[i].data(), x.pts[i], 32);
    for (int i = 0; i < 6; i++) memcpy(c.rproj[i].data(), x.rproj[i], 32);
    c.a_bits.assign(x.a_bits, x.a_bits + x.k_a);
    c.b_bits.assign(x.b_bits, x.b_bits + x.k_b);
    c.bind_outputs = x.bind_outputs != 0;
    return c;
}

static void p256_outs(const std::vector<sg::P256Proof> &r, stark_p256_out *outs) {
    for (size_t i = 0; i < r.size(); i++) {
        memcpy(outs[i].pi_hash, r[i].pi_hash, 32);
        memcpy(outs[i].root_f0, r[i].root_f0, 32);
        for (int j = 0; j < 6; j++) memcpy(outs[i].rproj[j], r[i].rproj[j].data(), 32);
        outs[i].n_trace = r[i].n_trace;
        outs[i].proof = copy_out(r[i].proof, &outs[i].proof_len);
    }
}

int stark_merge_ecdsa_p256(stark_ctx *ctx, const uint64_t *lde, size_t n_trace, size_t blowup, const uint64_t *coeffs,
                           size_t k, uint64_t *c_eval) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(lde && coeffs && c_eval, "merge_ecdsa_p256: null argument");
        SG_REQUIRE(k == sg::P256_K, "merge_ecdsa_p256: need one coefficient per constraint (155774)");
        SG_REQUIRE(n_trace <= (1u << 16) && blowup <= (1u << 10), "merge_ecdsa_p256: sizes out of range");
        const size_t n0 = n_trace * blowup;
        const bool ref = reference_lambdas(coeffs, k);
        sg::Scoped dl(c, 8ULL * sg::P256_W * n0), dc(c, 8 * n0), dk(c, ref ? 8 : 8 * k);
        h2d(c, dl.p, lde, 8ULL * sg::P256_W * n0);
        if (!ref) h2d(c, dk.p, coeffs, 8 * k);
        sg::merge_p256_batch(c, dl.as<uint64_t>(), 1, n_trace, blowup, dc.as<uint64_t>(),
                             ref ? nullptr : dk.as<uint64_t>());
        d2h(c, c_eval, dc.p, 8 * n0);
    });
}

int stark_prove_p256_chains_batch(stark_ctx *ctx, const stark_p256_chains *in, size_t n_proofs, size_t n_trace,
                                  const uint8_t *public_inputs_hashes, stark_p256_out *outs) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(in && public_inputs_hashes && outs, "prove_p256_chains_batch: null argument");
        SG_REQUIRE(n_proofs >= 1 && n_proofs <= 65535, "prove_p256_chains_batch: 1 to 65535 proofs");
        for (size_t i = 0; i < n_proofs; i++) outs[i].proof = nullptr;
        std::vector<sg::P256Chains> ch(n_proofs);
        std::vector<std::array<uint8_t, 32>> pis(n_proofs);
        for (size_t i = 0; i < n_proofs; i++) {
            ch[i] = p256_chains(in[i]);
            memcpy(pis[i].data(), public_inputs_hashes + 32 * i, 32);
        }
        p256_outs(sg::prove_p256_batch(c, ch, n_trace, pis), outs);
    });
}

int stark_prove_ecdsa_p256_batch(stark_ctx *ctx, const uint8_t *qx, const uint8_t *qy, const uint8_t *r,
                                 const uint8_t *s, const uint8_t *digest, size_t n_records, stark_p256_out *outs) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(qx && qy && r && s && digest && outs, "prove_ecdsa_p256_batch: null argument");
        SG_REQUIRE(n_records >= 1 && n_records <= 65535, "prove_ecdsa_p256_batch: 1 to 65535 records");
        for (size_t i = 0; i < n_records; i++) outs[i].proof = nullptr;
        std::vector<sg::P256Record> recs(n_records);
        for (size_t i = 0; i < n_records; i++) {
            memcpy(recs[i].qx.data(), qx + 32 * i, 32);
            memcpy(recs[i].qy.data(), qy + 32 * i, 32);
            memcpy(recs[i].r.data(), r + 32 * i, 32);
            memcpy(recs[i].s.data(), s + 32 * i, 32);
            memcpy(recs[i].digest.data(), digest + 32 * i, 32);
        }
        p256_outs(sg::prove_ecdsa_p256_batch(c, recs), outs);
    });
}

int stark_prove_rsa_exp_batch(stark_ctx *ctx, const uint8_t *n_be, const uint8_t *s_be, const uint8_t *em_be,
                              size_t n_proofs, size_t records_per_proof, size_t n_trace,
                              const uint8_t *public_inputs_hashes, stark_rsa_out *outs) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(public_inputs_hashes && outs, "prove_rsa_exp_batch: null argument");
        SG_REQUIRE(records_per_proof >= 1 && records_per_proof <= sg::RSA_MAX_RECORDS,
                   "prove_rsa_exp_batch: 1 to 4096 records per proof");
        SG_REQUIRE(n_proofs >= 1 && n_proofs * records_per_proof <= 65535, "prove_rsa_exp_batch: 1 to 65535 records");
        for (size_t i = 0; i < n_proofs; i++) outs[i].proof = nullptr;
        std::vector<sg::RsaRecord> recs = rsa_records(n_be, s_be, em_be, n_proofs * records_per_proof);
        std::vector<std::array<uint8_t, 32>> pis(n_proofs);
        for (size_t i = 0; i < n_proofs; i++) memcpy(pis[i].data(), public_inputs_hashes + 32 * i, 32);
        std::vector<sg::RsaProof> r = sg::prove_rsa_batch(c, recs, (uint32_t)records_per_proof, n_trace, pis);
        for (size_t i = 0; i < n_proofs; i++) {
            memcpy(outs[i].pi_hash, r[i].pi_hash, 32);
            memcpy(outs[i].root_f0, r[i].root_f0, 32);
            outs[i].n_trace = r[i].n_trace;
            outs[i].proof = copy_out(r[i].proof, &outs[i].proof_len);
        }
    });
}

static void fill_nsec3(const sg::Nsec3Proof &r, stark_nsec3_out *o) {
    memcpy(o->pi_hash, r.pi_hash, 32);
    memcpy(o->chain_root, r.chain_root, 32);
    memcpy(o->root_f0, r.root_f0, 32);
    o->record_count = r.record_count;
    o->n_trace = r.n_trace;
    o->proof = copy_out(r.proof, &o->proof_len);
}

int stark_prove_nsec3_completeness(stark_ctx *ctx, const uint8_t *records, size_t n, const uint8_t salt[16],
                                   const uint8_t fs_binding[32], int ldt, int phi_mode, stark_nsec3_out *out) {
    size_t counts[1] = {n};
    return stark_prove_nsec3_completeness_batch(ctx, records, counts, 1, salt, fs_binding, ldt, phi_mode, out);
}

int stark_prove_nsec3_completeness_batch(stark_ctx *ctx, const uint8_t *records, const size_t *counts,
                                         size_t n_chains, const uint8_t salt[16], const uint8_t *fs_bindings,
                                         int ldt, int phi_mode, stark_nsec3_out *outs) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(records && counts && salt && fs_bindings && outs, "prove_nsec3_completeness: null argument");
        SG_REQUIRE(ldt == STARK_LDT_STIR || ldt == STARK_LDT_FRI, "unknown LdtMode");
        SG_REQUIRE(phi_mode >= 0 && phi_mode <= 2, "prove_nsec3_completeness: bad phi_mode");
        SG_REQUIRE(n_chains >= 1 && n_chains <= 65535, "prove_nsec3_completeness: 1 to 65535 chains");
        std::vector<sg::ByteSpan> ch(n_chains);
        std::vector<std::array<uint8_t, 32>> fsb(n_chains);
        size_t off = 0;
        for (size_t i = 0; i < n_chains; i++) {
            ch[i] = {records + 64 * off, 64 * counts[i]};
            off += counts[i];
            memcpy(fsb[i].data(), fs_bindings + 32 * i, 32);
            outs[i].proof = nullptr;
        }
        std::vector<sg::Nsec3Proof> r = sg::prove_nsec3_batch(c, ch, salt, fsb, phi_mode, ldt == STARK_LDT_STIR);
        for (size_t i = 0; i < n_chains; i++) fill_nsec3(r[i], &outs[i]);
    });
}

static void fill_ds_ksk(const sg::DsKsk &r, stark_ds_ksk_out *o) {
    memcpy(o->pi_hash, r.pi_hash, 32);
    memcpy(o->asserted_digest, r.asserted_digest, 32);
    memcpy(o->root_f0, r.root_f0, 32);
    o->n_blocks = r.n_blocks;
    o->n_trace = r.n_trace;
    o->proof = copy_out(r.proof, &o->proof_len);
}

int stark_prove_ds_ksk_binding(stark_ctx *ctx, const uint8_t *dnskey, size_t len, const uint8_t parent_ds_hash[32],
                               const uint8_t fs_binding[32], int ldt, int phi_mode, stark_ds_ksk_out *out) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE((dnskey || len == 0) && parent_ds_hash && fs_binding && out, "prove_ds_ksk_binding: null argument");
        SG_REQUIRE(ldt == STARK_LDT_STIR || ldt == STARK_LDT_FRI, "unknown LdtMode");
        out->proof = nullptr;
        std::array<uint8_t, 32> par, fsb;
        memcpy(par.data(), parent_ds_hash, 32);
        memcpy(fsb.data(), fs_binding, 32);
        std::vector<sg::DsKsk> r = sg::prove_ds_ksk_binding_batch(c, {{dnskey, len}}, {par}, {fsb}, phi_mode,
                                                                  ldt == STARK_LDT_STIR);
        fill_ds_ksk(r[0], out);
    });
}

int stark_prove_ds_ksk_binding_batch(stark_ctx *ctx, const uint8_t *dnskeys, const size_t *offsets, const size_t *lens,
                                     size_t n_keys, const uint8_t *parent_ds_hashes, const uint8_t *fs_bindings,
                                     int ldt, int phi_mode, stark_ds_ksk_out *outs) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(offsets && lens && parent_ds_hashes && fs_bindings && outs, "prove_ds_ksk_binding_batch: null argument");
        SG_REQUIRE(ldt == STARK_LDT_STIR || ldt == STARK_LDT_FRI, "unknown LdtMode");
        std::vector<sg::ByteSpan> keys(n_keys);
        std::vector<std::array<uint8_t, 32>> par(n_keys), fsb(n_keys);
        for (size_t i = 0; i < n_keys; i++) {
            SG_REQUIRE(dnskeys || lens[i] == 0, "prove_ds_ksk_binding_batch: null keys");
            keys[i] = {dnskeys ? dnskeys + offsets[i] : nullptr, lens[i]};
            memcpy(par[i].data(), parent_ds_hashes + 32 * i, 32);
            memcpy(fsb[i].data(), fs_bindings + 32 * i, 32);
            outs[i].proof = nullptr;
        }
        std::vector<sg::DsKsk> r = sg::prove_ds_ksk_binding_batch(c, keys, par, fsb, phi_mode, ldt == STARK_LDT_STIR);
        for (size_t i = 0; i < n_keys; i++) fill_ds_ksk(r[i], &outs[i]);
    });
}

int stark_deep_fri_prove_batch(stark_ctx *ctx, const uint64_t *f0s, size_t n_proofs, size_t n0,
                               const stark_fri_params *params, const uint8_t *public_inputs_hashes,
                               uint8_t **proofs, size_t *proof_lens) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(f0s && proofs && proof_lens && n_proofs >= 1, "deep_fri_prove_batch: null argument");
        sg::FriParams p = to_params(params);
        SG_REQUIRE(sg::is_pow2(n0), "deep_fri_prove_batch: n0 must be a power of two");
        std::vector<std::array<uint8_t, 32>> pis;
        if (public_inputs_hashes) {
            pis.resize(n_proofs);
            for (size_t k = 0; k < n_proofs; k++) memcpy(pis[k].data(), public_inputs_hashes + 32 * k, 32);
        }
        sg::Scoped d(c, 8 * n0 * n_proofs);
        h2d(c, d.p, f0s, 8 * n0 * n_proofs);
        std::vector<std::vector<uint8_t>> v =
            sg::deep_fri_prove_batch(c, d.as<uint64_t>(), (uint32_t)n_proofs, n0, p, public_inputs_hashes ? &pis : nullptr,
                                     nullptr);
        for (size_t k = 0; k < n_proofs; k++) proofs[k] = nullptr;
        for (size_t k = 0; k < n_proofs; k++) proofs[k] = copy_out(v[k], &proof_lens[k]);
    });
}

int stark_prove_outer_rollup(stark_ctx *ctx, const uint8_t *pi_hashes, size_t n, const uint8_t pk_hash[32], int ldt,
                             int phi_mode, uint8_t **proof, size_t *proof_len, uint8_t root_f0[32]) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE((pi_hashes || n == 0) && pk_hash && proof && proof_len && root_f0, "prove_outer_rollup: null argument");
        SG_REQUIRE(ldt == STARK_LDT_STIR || ldt == STARK_LDT_FRI, "unknown LdtMode");
        std::vector<uint8_t> v = sg::prove_outer_rollup(c, pi_hashes, n, pk_hash, phi_mode, root_f0, ldt == STARK_LDT_STIR);
        *proof = copy_out(v, proof_len);
    });
}

int stark_dev_alloc(stark_ctx *ctx, size_t bytes, void **ptr) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(ptr, "dev_alloc: null out pointer");
        *ptr = c->alloc(bytes);
    });
}

int stark_dev_free(stark_ctx *ctx, void *ptr) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_HIP(hipStreamSynchronize(c->stream));
        c->release(ptr);
    });
}

int stark_memcpy_h2d(stark_ctx *ctx, void *dst, const void *src, size_t bytes) {
    return guard(ctx, [&](sg::Ctx *c) {
        h2d(c, dst, src, bytes);
        SG_HIP(hipStreamSynchronize(c->stream));
    });
}

int stark_memcpy_d2h(stark_ctx *ctx, void *dst, const void *src, size_t bytes) {
    return guard(ctx, [&](sg::Ctx *c) { d2h(c, dst, src, bytes); });
}

int stark_synchronize(stark_ctx *ctx) {
    return guard(ctx, [&](sg::Ctx *c) { SG_HIP(hipStreamSynchronize(c->stream)); });
}

int stark_ntt_dev(stark_ctx *ctx, const uint64_t *d_in, uint64_t *d_out, size_t log_n, size_t ncols, int inverse) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(d_in && d_out && ncols >= 1 && log_n <= 24, "ntt: bad arguments");
        const uint64_t n = 1ULL << log_n;
        sg::Scoped scr(c, 8 * n * ncols);
        sg::ntt_columns(c, d_in, d_out, scr.as<uint64_t>(), (uint32_t)log_n, ncols, n, n, inverse != 0, nullptr);
        SG_HIP(hipStreamSynchronize(c->stream));
    });
}

int stark_ntt_dev_noalloc(stark_ctx *ctx, const uint64_t *d_in, uint64_t *d_out, size_t log_n, size_t ncols,
                          int inverse) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(d_in && d_out && ncols >= 1 && log_n <= 24, "ntt (no alloc): bad arguments");
        sg::ntt_columns_noalloc(c, d_in, d_out, (uint32_t)log_n, ncols, inverse != 0);
        SG_HIP(hipStreamSynchronize(c->stream));
    });
}

int stark_lde_columns_dev(stark_ctx *ctx, const uint64_t *d_cols, size_t w, size_t n_trace, size_t blowup,
                          uint64_t shift, uint64_t *d_out) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(d_cols && d_out && w > 0, "lde: null buffer or no columns");
        SG_REQUIRE(sg::is_pow2(n_trace), "lde: trace_len must be a power of 2");
        SG_REQUIRE(blowup >= 2 && sg::is_pow2(blowup), "lde: blowup must be a power-of-2 >= 2");
        SG_REQUIRE(shift != 0 && shift < sg::P, "lde: shift must be a nonzero canonical field element");
        sg::lde_columns(c, d_cols, d_out, sg::ilog2(n_trace), sg::ilog2(blowup), w, n_trace, n_trace * blowup, shift);
        SG_HIP(hipStreamSynchronize(c->stream));
    });
}

int stark_prover_core_dev(stark_ctx *ctx, const uint64_t *d_trace, size_t w, size_t n_trace, size_t blowup,
                          uint64_t shift, const stark_fri_params *params, uint64_t *d_lde, uint8_t row_root[32],
                          uint8_t **proof, size_t *proof_len) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(d_trace && d_lde && row_root && proof && proof_len, "prover_core: null argument");
        SG_REQUIRE(shift != 0 && shift < sg::P, "prover_core: shift must be a nonzero canonical field element");
        sg::FriParams p = to_params(params);
        std::vector<uint8_t> v = sg::prover_core_dev(c, d_trace, w, n_trace, blowup, shift, p, d_lde, row_root);
        *proof = copy_out(v, proof_len);
    });
}

int stark_profile_enable(stark_ctx *ctx, int on) {
    if (!ctx || !ctx->c) return STARK_E_INVALID;
    ctx->c->profiling = on != 0;
    return STARK_OK;
}

int stark_profile_reset(stark_ctx *ctx) {
    if (!ctx || !ctx->c) return STARK_E_INVALID;
    ctx->c->stats.clear();
    return STARK_OK;
}

int stark_profile_get(stark_ctx *ctx, const char *name, double *ms, uint64_t *launches, double *bytes) {
    if (!ctx || !ctx->c || !name) return STARK_E_INVALID;
    auto it = ctx->c->stats.find(name);
    sg::KernelStat s;
    if (it != ctx->c->stats.end()) s = it->second;
    if (ms) *ms = s.ms;
    if (launches) *launches = s.launches;
    if (bytes) *bytes = s.bytes;
    return STARK_OK;
}

size_t stark_profile_names(stark_ctx *ctx, char *buf, size_t cap) {
    if (!ctx || !ctx->c) return 0;
    std::string all;
    for (auto &kv : ctx->c->stats) all += kv.first + "\n";
    if (buf && cap) {
        size_t n = all.size() < cap - 1 ? all.size() : cap - 1;
        memcpy(buf, all.data(), n);
        buf[n] = 0;
    }
    return all.size();
}

int stark_set_autotune(stark_ctx *ctx, int on) {
    if (!ctx) {
        sg::set_autotune_default(on != 0);
        return STARK_OK;
    }
    if (!ctx->c) return STARK_E_INVALID;
    ctx->c->autotune = on != 0;
    return STARK_OK;
}

int stark_set_ntt_split(stark_ctx *ctx, unsigned log_first_pass) {
    if (!ctx || !ctx->c || log_first_pass > 12) return STARK_E_INVALID;
    ctx->c->force_split = log_first_pass;
    return STARK_OK;
}

int stark_set_ntt_tile(stark_ctx *ctx, unsigned tile_log) {
    if (!ctx || !ctx->c || (tile_log != 0 && tile_log != 13 && tile_log != 14)) return STARK_E_INVALID;
    ctx->c->force_tile = tile_log;
    return STARK_OK;
}

size_t stark_autotune_report(char *buf, size_t cap) {
    const std::string all = sg::autotune_report();
    if (buf && cap) {
        size_t n = all.size() < cap - 1 ? all.size() : cap - 1;
        memcpy(buf, all.data(), n);
        buf[n] = 0;
    }
    return all.size();
}

void stark_host_sha3_256(const uint8_t *msg, size_t len, uint8_t out[32]) { sg::sha3_256_host(msg, len, out); }

uint64_t stark_host_stdrng_first_u64(uint64_t seed) { return sg::stdrng_first_u64(seed); }

size_t stark_host_sha256_trace_height(size_t len) { return sg::sha256_trace_height(len); }

size_t stark_host_build_sha256_trace(const uint8_t *msg, size_t len, uint64_t *trace, uint8_t digest[32]) {
    return sg::build_sha256_trace(msg, len, trace, digest);
}

int stark_server_create(int device, int workers, size_t hbm_budget, size_t max_batch, int flags, stark_server **out) {
    if (!out) return STARK_E_INVALID;
    *out = nullptr;
    stark_server *x = new (std::nothrow) stark_server();
    if (!x) return STARK_E_OOM;
    try {
        x->s = new sg::Server(device, workers, hbm_budget, (uint32_t)max_batch, (flags & STARK_SERVER_DRY_RUN) != 0,
                              (flags & STARK_SERVER_NO_VALIDATE) == 0);
    } catch (const sg::Error &e) {
        delete x;
        return e.code;
    } catch (...) {
        delete x;
        return STARK_E_INTERNAL;
    }
    *out = x;
    return STARK_OK;
}

void stark_server_destroy(stark_server *s) {
    if (!s) return;
    delete s->s;
    delete s;
}

int stark_server_submit(stark_server *s, const stark_serve_request *q, uint64_t *id) {
    if (!s || !s->s || !q || !id) return STARK_E_INVALID;
    try {
        sg::ServeRequest r;
        r.kind = q->kind;
        SG_REQUIRE(q->deadline_s >= 0, "server submit: negative deadline");
        r.deadline_s = q->deadline_s;
        SG_REQUIRE(q->fs_binding || q->kind == STARK_SERVE_P256, "server submit: null fs_binding");
        if (q->fs_binding) memcpy(r.fs_binding.data(), q->fs_binding, 32);
        if (q->kind == STARK_SERVE_P256) {
            SG_REQUIRE(q->p256_qx && q->p256_qy && q->p256_r && q->p256_s && q->p256_digest,
                       "server submit: null ECDSA-P256 input");
            memcpy(r.p256.qx.data(), q->p256_qx, 32);
            memcpy(r.p256.qy.data(), q->p256_qy, 32);
            memcpy(r.p256.r.data(), q->p256_r, 32);
            memcpy(r.p256.s.data(), q->p256_s, 32);
            memcpy(r.p256.digest.data(), q->p256_digest, 32);
        } else if (q->kind == STARK_SERVE_ED25519) {
            SG_REQUIRE(q->ed_pk && q->ed_sig && (q->ed_msg || q->ed_msg_len == 0) && q->ed_merkle_root,
                       "server submit: null Ed25519 input");
            memcpy(r.ed_pk.data(), q->ed_pk, 32);
            memcpy(r.ed_sig.data(), q->ed_sig, 64);
            if (q->ed_msg_len) r.ed_msg.assign(q->ed_msg, q->ed_msg + q->ed_msg_len);
            memcpy(r.ed_root.data(), q->ed_merkle_root, 32);
        } else if (q->kind == STARK_SERVE_DS_KSK) {
            SG_REQUIRE((q->dnskey || q->dnskey_len == 0) && q->parent_ds_hash, "server submit: null DS-KSK input");
            if (q->dnskey_len) r.key.assign(q->dnskey, q->dnskey + q->dnskey_len);
            memcpy(r.parent.data(), q->parent_ds_hash, 32);
        } else if (q->kind == STARK_SERVE_RSA) {
            SG_REQUIRE(q->rsa_n && q->rsa_s && q->rsa_em, "server submit: null RSA input");
            memcpy(r.rsa.n.data(), q->rsa_n, 256);
            memcpy(r.rsa.s.data(), q->rsa_s, 256);
            memcpy(r.rsa.em.data(), q->rsa_em, 256);
        } else {
            SG_REQUIRE(q->salt && (q->records || q->n_records == 0), "server submit: null shard input");
            memcpy(r.salt.data(), q->salt, 16);
            for (size_t i = 0; i < q->n_records; i++) {
                const stark_dns_record &d = q->records[i];
                r.domains.emplace_back(d.domain, d.domain_len);
                r.rdata.emplace_back(d.rdata, d.rdata + d.rdata_len);
                r.rtype.push_back(d.record_type);
                r.ttl.push_back(d.ttl);
            }
        }
        *id = s->s->submit(std::move(r));
        return STARK_OK;
    } catch (const sg::Error &e) {
        server_err = e.what();
        return e.code;
    } catch (const std::exception &e) {
        server_err = e.what();
        return STARK_E_INTERNAL;
    }
}

int stark_server_wait(stark_server *s, uint64_t id, double timeout_s, stark_serve_result *out) {
    if (!s || !s->s || !out) return STARK_E_INVALID;
    sg::ServeResult r;
    if (!s->s->wait(id, r, timeout_s)) {
        server_err = "server wait: result not ready in time";
        return STARK_E_TIMEOUT;
    }
    out->id = r.id;
    out->status = r.status;
    out->kind = r.kind;
    memcpy(out->pi_hash, r.pi_hash, 32);
    memcpy(out->root_f0, r.root_f0, 32);
    memcpy(out->aux, r.aux, 32);
    out->n_trace = r.n_trace;
    out->batch_size = r.batch_size;
    out->batch_seq = r.batch_seq;
    out->t_submit = r.t_submit;
    out->t_start = r.t_start;
    out->t_done = r.t_done;
    out->proof = copy_out(r.proof, &out->proof_len);
    snprintf(out->error, sizeof out->error, "%s", r.error.c_str());
    if (r.status) server_err = r.error;
    return STARK_OK;
}

int stark_server_stats(stark_server *s, size_t *completed, size_t *peak, size_t *budget) {
    if (!s || !s->s) return STARK_E_INVALID;
    if (completed) *completed = s->s->completed();
    if (peak) *peak = s->s->peak_in_use();
    if (budget) *budget = s->s->budget();
    return STARK_OK;
}

const char *stark_server_last_error(const stark_server *s) { return s ? server_err.c_str() : "null server"; }

size_t stark_host_sha512_trace_height(size_t len) { return sg::sha512_trace_height(len); }

size_t stark_host_build_sha512_trace(const uint8_t *msg, size_t len, uint64_t *trace, uint8_t digest[64]) {
    return sg::build_sha512_trace(msg, len, trace, digest);
}

int stark_host_build_rsa_exp_trace(const uint8_t *n_be, const uint8_t *s_be, const uint8_t *em_be, size_t n_records,
                                   size_t n_trace, uint64_t *trace) {
    try {
        SG_REQUIRE(trace, "rsa trace: null output");
        sg::build_rsa_trace_host(rsa_records(n_be, s_be, em_be, n_records), n_trace, trace);
        return STARK_OK;
    } catch (const sg::Error &e) {
        return e.code;
    } catch (const std::bad_alloc &) {
        return STARK_E_OOM;
    } catch (const std::exception &) {
        return STARK_E_INTERNAL;
    }
}

void stark_host_rsa_pkcs1_em_sha256(const uint8_t digest[32], uint8_t em[256]) { sg::rsa_pkcs1_em_sha256(digest, em); }

int stark_host_build_ecdsa_p256_trace(stark_p256_chains *in, size_t n_trace, uint64_t *trace) {
    try {
        SG_REQUIRE(in && trace, "p256 trace: null argument");
        std::array<sg::Be32, 6> rp;
        sg::build_p256_trace_host(p256_chains(*in), n_trace, trace, &rp);
        for (int j = 0; j < 6; j++) memcpy(in->rproj[j], rp[j].data(), 32);
        return STARK_OK;
    } catch (const sg::Error &e) {
        return e.code;
    } catch (const std::bad_alloc &) {
        return STARK_E_OOM;
    } catch (const std::exception &) {
        return STARK_E_INTERNAL;
    }
}

void stark_host_dns_leaf_hash(const stark_dns_record *r, const uint8_t salt[16], uint8_t out[32]) {
    sg::DnsRec d{r->domain, r->domain_len, r->record_type, r->ttl, r->rdata, r->rdata_len};
    sg::dns_leaf_hash(d, salt, out);
}

int stark_ed25519_shape(size_t msg_len, size_t k_scalar, size_t *width, size_t *height, size_t *n_constraints) {
    try {
        SG_REQUIRE(width && height && n_constraints, "ed25519_shape: null argument");
        SG_REQUIRE(k_scalar >= 2 && k_scalar <= 256, "ed25519_shape: k_scalar in [2, 256]");
        const sg::EdShape s = sg::ed25519_shape(msg_len, (uint32_t)k_scalar);
        *width = s.width;
        *height = s.height;
        *n_constraints = s.n_constraints;
        return STARK_OK;
    } catch (const sg::Error &e) {
        return e.code;
    } catch (const std::exception &) {
        return STARK_E_INTERNAL;
    }
}

int stark_prove_zsk_ksk_binding_v2_batch(stark_ctx *ctx, const uint8_t *pks, const uint8_t *sigs, const uint8_t *msgs,
                                         const size_t *msg_off, const size_t *msg_len, const uint8_t *fs_bindings,
                                         const uint8_t *merkle_roots, size_t n, size_t k_scalar, size_t blowup,
                                         size_t queries, int ldt, stark_zsk_out *outs) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(pks && sigs && msg_off && msg_len && fs_bindings && merkle_roots && outs,
                   "prove_zsk_ksk_binding_v2_batch: null argument");
        SG_REQUIRE(n >= 1 && n <= 65535, "prove_zsk_ksk_binding_v2_batch: 1 to 65535 signatures");
        SG_REQUIRE(k_scalar >= 2 && k_scalar <= 256, "prove_zsk_ksk_binding_v2_batch: k_scalar in [2, 256]");
        SG_REQUIRE(ldt == STARK_LDT_STIR || ldt == STARK_LDT_FRI, "unknown LdtMode");
        for (size_t i = 0; i < n; i++) outs[i].proof = nullptr;
        std::vector<sg::EdSig> v(n);
        std::vector<std::array<uint8_t, 32>> fs(n), mr(n);
        for (size_t i = 0; i < n; i++) {
            SG_REQUIRE(msg_len[i] == 0 || msgs, "prove_zsk_ksk_binding_v2_batch: null message");
            v[i] = sg::EdSig{pks + 32 * i, sigs + 64 * i, msg_len[i] ? msgs + msg_off[i] : nullptr, msg_len[i]};
            memcpy(fs[i].data(), fs_bindings + 32 * i, 32);
            memcpy(mr[i].data(), merkle_roots + 32 * i, 32);
        }
        std::vector<sg::ZskKsk> r =
            sg::prove_zsk_ksk_binding_v2_batch(c, v, fs, mr, (uint32_t)k_scalar, blowup, (uint32_t)queries,
                                               ldt == STARK_LDT_STIR);
        for (size_t i = 0; i < n; i++) {
            memcpy(outs[i].pi_hash, r[i].pi_hash, 32);
            memcpy(outs[i].root_f0, r[i].root_f0, 32);
            outs[i].n_trace = r[i].n_trace;
            outs[i].proof = copy_out(r[i].proof, &outs[i].proof_len);
        }
    });
}

int stark_merge_ed25519(stark_ctx *ctx, const uint8_t pk[32], const uint8_t sig[64], const uint8_t *msg,
                        size_t msg_len, size_t k_scalar, const uint64_t *lde, size_t blowup, const uint64_t *coeffs,
                        size_t n_coeffs, uint64_t *c_eval) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(pk && sig && lde && coeffs && c_eval && (msg || !msg_len), "merge_ed25519: null argument");
        SG_REQUIRE(k_scalar >= 2 && k_scalar <= 256, "merge_ed25519: k_scalar in [2, 256]");
        const sg::EdShape sh = sg::ed25519_shape(msg_len, (uint32_t)k_scalar);
        SG_REQUIRE(n_coeffs == sh.n_constraints, "merge_ed25519: need one coefficient per constraint");
        SG_REQUIRE(blowup >= 2 && blowup <= 64 && (blowup & (blowup - 1)) == 0, "merge_ed25519: blowup in {2, ..., 64}");
        const size_t n0 = sh.height * blowup;
        sg::Scoped dl(c, 8ULL * sh.width * n0), dc(c, 8 * n0), dk(c, 8 * n_coeffs);
        h2d(c, dl.p, lde, 8ULL * sh.width * n0);
        h2d(c, dk.p, coeffs, 8 * n_coeffs);
        std::vector<sg::EdSig> v = {sg::EdSig{pk, sig, msg, msg_len}};
        sg::merge_ed25519_batch(c, v, (uint32_t)k_scalar, dl.as<uint64_t>(), blowup, dk.as<uint64_t>(),
                                dc.as<uint64_t>());
        d2h(c, c_eval, dc.p, 8 * n0);
    });
}

int stark_build_ed25519_trace(stark_ctx *ctx, const uint8_t pk[32], const uint8_t sig[64], const uint8_t *msg,
                              size_t msg_len, size_t k_scalar, uint64_t *trace) {
    return guard(ctx, [&](sg::Ctx *c) {
        SG_REQUIRE(pk && sig && trace && (msg || !msg_len), "build_ed25519_trace: null argument");
        SG_REQUIRE(k_scalar >= 2 && k_scalar <= 256, "build_ed25519_trace: k_scalar in [2, 256]");
        const sg::EdShape sh = sg::ed25519_shape(msg_len, (uint32_t)k_scalar);
        sg::Scoped dt(c, 8ULL * sh.width * sh.height);
        std::vector<sg::EdSig> v = {sg::EdSig{pk, sig, msg, msg_len}};
        sg::build_ed25519_traces_dev(c, v, (uint32_t)k_scalar, dt.as<uint64_t>());
        d2h(c, trace, dt.p, 8ULL * sh.width * sh.height);
    });
}

int stark_host_build_ed25519_trace(const uint8_t pk[32], const uint8_t sig[64], const uint8_t *msg, size_t msg_len,
                                   size_t k_scalar, uint64_t *trace) {
    try {
        SG_REQUIRE(pk && sig && trace && (msg || !msg_len), "ed25519 trace: null argument");
        SG_REQUIRE(k_scalar >= 2 && k_scalar <= 256, "ed25519 trace: k_scalar in [2, 256]");
        sg::build_ed25519_trace_host(sg::EdSig{pk, sig, msg, msg_len}, (uint32_t)k_scalar, trace);
        return STARK_OK;
    } catch (const sg::Error &e) {
        return e.code;
    } catch (const std::bad_alloc &) {
        return STARK_E_OOM;
    } catch (const std::exception &) {
        return STARK_E_INTERNAL;
    }
}

int stark_host_ed25519_verify(const uint8_t pk[32], const uint8_t sig[64], const uint8_t *msg, size_t msg_len) {
    if (!pk || !sig || (!msg && msg_len)) return 0;
    return sg::ed25519_verify(sg::EdSig{pk, sig, msg, msg_len}) ? 1 : 0;
}

void stark_host_ed25519_sign(const uint8_t seed[32], const uint8_t *msg, size_t msg_len, uint8_t pk[32],
                             uint8_t sig[64]) {
    sg::ed25519_sign(seed, msg, msg_len, pk, sig);
}

}  // extern "C"
