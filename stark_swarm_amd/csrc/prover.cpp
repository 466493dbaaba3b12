// prover.cpp -- whole-proof flows on the device: the swarm-dns HashRollup
// provers (inner shard, outer rollup) and the prover-core pipeline.
//
//   prove_inner_shard   crates/swarm-dns/src/prover.rs:148-227
//   prove_outer_rollup  crates/swarm-dns/src/prover.rs:253-304
//   DNS leaf hashing    crates/swarm-dns/src/dns.rs:43-118
// Host work here is the step before the path (per-record SHA3, the DNS
// Merkle tree over <= a few thousand leaves, the sequential HashRollup
// recurrence) and the transcript; LDE, composition, quotient, FRI folds and
// all Merkle commitments over the LDE domain run on the device.
#include "prover.hpp"

#include <string.h>

#include "air.hpp"
#include "fri.hpp"
#include "gl.hpp"
#include "keccak.hpp"
#include "ntt.hpp"

namespace sg {

static void put_le(std::vector<uint8_t> &v, uint64_t x, int n) {
    for (int i = 0; i < n; i++) v.push_back((uint8_t)(x >> (8 * i)));
}

void dns_leaf_hash(const DnsRec &r, const uint8_t salt[16], uint8_t out[32]) {
    std::vector<uint8_t> canon;
    const char *tag = "DNS-RECORD-V1";
    canon.insert(canon.end(), tag, tag + 13);
    put_le(canon, r.domain_len, 4);
    canon.insert(canon.end(), (const uint8_t *)r.domain, (const uint8_t *)r.domain + r.domain_len);
    put_le(canon, r.record_type, 2);
    put_le(canon, r.ttl, 4);
    put_le(canon, r.rdata_len, 4);
    canon.insert(canon.end(), r.rdata, r.rdata + r.rdata_len);
    Sha3 h;
    uint8_t h1[32];
    h.update("DNS-LEAF-V1");
    h.update(salt, 16);
    h.update(canon.data(), canon.size());
    h.final(h1);
    h.reset();
    h.update("DNS-LEAF-DOUBLE-V1");
    h.update(salt, 16);
    h.update(h1, 32);
    h.final(out);
}

void dns_merkle_root(const std::vector<uint8_t> &leaves, uint8_t root[32]) {
    std::vector<uint8_t> cur = leaves;
    size_t n = cur.size() / 32;
    while (n > 1) {
        size_t np = (n + 1) / 2;
        std::vector<uint8_t> nxt(32 * np);
        for (size_t i = 0; i < np; i++) {
            const uint8_t *l = &cur[64 * i];
            const uint8_t *r = (2 * i + 1 < n) ? l + 32 : l;
            Sha3 h;
            h.update("DNS-NODE-V1");
            h.update(l, 32);
            h.update(r, 32);
            h.final(&nxt[32 * i]);
        }
        cur.swap(nxt);
        n = np;
    }
    memcpy(root, cur.data(), 32);
}

static const uint64_t SWARM_BLOWUP = 32, SWARM_QUERIES = 54, SWARM_SEED_Z = 0xDEEFBAADULL;

std::vector<uint8_t> rollup_prove(Ctx *c, const std::vector<uint64_t> &leaves, const uint8_t binding[32],
                                  int phi_mode, uint8_t root_f0[32]) {
    uint64_t n_trace = 8;
    while (n_trace < leaves.size()) n_trace <<= 1;  // next_power_of_two().max(8)
    const uint64_t n0 = n_trace * SWARM_BLOWUP;
    std::vector<uint64_t> trace(4 * n_trace);
    build_hash_rollup_trace(n_trace, leaves.data(), leaves.size(), trace.data());
    Scoped d_tr(c, 8 * 4 * n_trace), d_lde(c, 8 * 4 * n0), d_c(c, 8 * n0);
    SG_HIP(hipMemcpyAsync(d_tr.p, trace.data(), 8 * trace.size(), hipMemcpyHostToDevice, c->stream));
    lde_columns(c, d_tr.as<uint64_t>(), d_lde.as<uint64_t>(), ilog2(n_trace), ilog2(SWARM_BLOWUP), 4, n_trace, n0, 1);
    merge_hash_rollup_dev(c, d_lde.as<uint64_t>(), n_trace, SWARM_BLOWUP, phi_mode, d_c.as<uint64_t>());
    FriParams p;
    p.schedule = stir_schedule(n0);
    p.r = p.s0 = SWARM_QUERIES;
    p.seed_z = SWARM_SEED_Z;
    p.coeff_commit_final = true;
    p.d_final = 1;
    p.stir = true;
    p.has_pi = true;
    memcpy(p.pi, binding, 32);
    std::vector<uint8_t> proof = deep_fri_prove_dev(c, d_c.as<uint64_t>(), n0, p, nullptr);
    memcpy(root_f0, proof.data(), 32);
    return proof;
}

static void pack_hash(const uint8_t *h, std::vector<uint64_t> &out) {
    for (int k = 0; k < 4; k++) {
        uint64_t v = 0;
        for (int b = 0; b < 8; b++) v |= (uint64_t)h[8 * k + b] << (8 * b);
        out.push_back(v);
    }
}

InnerShard prove_inner_shard(Ctx *c, const uint8_t salt[16], const std::vector<DnsRec> &recs,
                             const uint8_t fs_binding[32], int phi_mode) {
    SG_REQUIRE(!recs.empty(), "prove_inner_shard: merkle tree needs >= 1 leaf");
    InnerShard o;
    std::vector<uint8_t> lh(32 * recs.size());
    for (size_t i = 0; i < recs.size(); i++) dns_leaf_hash(recs[i], salt, &lh[32 * i]);
    dns_merkle_root(lh, o.merkle_root);
    std::vector<uint64_t> leaves;
    for (size_t i = 0; i < recs.size(); i++) pack_hash(&lh[32 * i], leaves);
    uint64_t n_trace = 8;
    while (n_trace < leaves.size()) n_trace <<= 1;
    o.n_trace = n_trace;
    o.record_count = recs.size();
    o.proof = rollup_prove(c, leaves, fs_binding, phi_mode, o.root_f0);
    // pi_hash = SHA3("DNS-SHARD-PIHASH-V1" || salt || LE64(count) || root || root_f0 || fs_binding)
    Sha3 h;
    uint8_t cnt[8];
    for (int i = 0; i < 8; i++) cnt[i] = (uint8_t)((uint64_t)recs.size() >> (8 * i));
    h.update("DNS-SHARD-PIHASH-V1");
    h.update(salt, 16);
    h.update(cnt, 8);
    h.update(o.merkle_root, 32);
    h.update(o.root_f0, 32);
    h.update(fs_binding, 32);
    h.final(o.pi_hash);
    return o;
}

std::vector<uint8_t> prove_outer_rollup(Ctx *c, const uint8_t *pi_hashes, size_t n, const uint8_t pk_hash[32],
                                        int phi_mode, uint8_t root_f0[32]) {
    std::vector<uint64_t> leaves;
    for (size_t i = 0; i < n; i++) pack_hash(pi_hashes + 32 * i, leaves);
    return rollup_prove(c, leaves, pk_hash, phi_mode, root_f0);
}

std::vector<uint8_t> prover_core_dev(Ctx *c, const uint64_t *d_trace, uint64_t w, uint64_t n_trace, uint64_t blowup,
                                     uint64_t shift, const FriParams &p0, uint64_t *d_lde, uint8_t row_root[32]) {
    SG_REQUIRE(w >= 1 && w <= (1u << 20), "prover_core: bad width");
    SG_REQUIRE(is_pow2(n_trace) && is_pow2(blowup) && blowup >= 2, "prover_core: sizes must be powers of two");
    const uint64_t n0 = n_trace * blowup;
    SG_REQUIRE(n0 <= (1ULL << 24), "prover_core: domain above 2^24");
    lde_columns(c, d_trace, d_lde, ilog2(n_trace), ilog2(blowup), w, n_trace, n0, shift);
    // row tree with the fused f0 combination
    Scoped d_f0(c, 8 * n0);
    DevTree rows;
    LeafSrc src{d_lde, 1, n0, (uint32_t)w, (uint32_t)w, d_f0.as<uint64_t>()};
    const uint32_t depth = ilog2(n0) ? ilog2(n0) : 1;
    tree_build(c, rows, src, n0, 2, depth, 0);
    memcpy(row_root, rows.root, 32);
    FriParams p = p0;
    p.has_pi = true;
    memcpy(p.pi, row_root, 32);
    std::vector<uint64_t> raw;
    std::vector<uint8_t> out = deep_fri_prove_dev(c, d_f0.as<uint64_t>(), n0, p, &raw);
    // row openings at the raw query indices
    std::vector<uint64_t> gi;
    for (uint64_t q : raw)
        for (uint64_t col = 0; col < w; col++) gi.push_back(col * n0 + q);
    std::vector<uint64_t> vals(gi.size());
    if (!gi.empty()) {
        Scoped di(c, 8 * gi.size()), dv(c, 8 * gi.size());
        SG_HIP(hipMemcpyAsync(di.p, gi.data(), 8 * gi.size(), hipMemcpyHostToDevice, c->stream));
        gather_u64(c, d_lde, di.as<uint64_t>(), gi.size(), dv.as<uint64_t>());
        SG_HIP(hipMemcpyAsync(vals.data(), dv.p, 8 * gi.size(), hipMemcpyDeviceToHost, c->stream));
        SG_HIP(hipStreamSynchronize(c->stream));
    }
    std::vector<std::vector<uint8_t>> ops;
    tree_open_batch(c, rows, raw, ops);
    tree_free(c, rows);
    put_le(out, raw.size(), 8);
    for (size_t q = 0; q < raw.size(); q++) {
        put_le(out, w, 8);
        for (uint64_t col = 0; col < w; col++) put_le(out, vals[q * w + col], 8);
        out.insert(out.end(), ops[q].begin(), ops[q].end());
    }
    return out;
}

}  // namespace sg
