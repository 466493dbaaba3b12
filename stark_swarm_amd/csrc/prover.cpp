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

#include <algorithm>
#include <atomic>
#include <map>
#include <thread>

#include "air.hpp"
#include "fri.hpp"
#include "gl.hpp"
#include "keccak.hpp"
#include "ntt.hpp"
#include "rsa_air.hpp"
#include "sha256_air.hpp"

namespace sg {

static void put_le(std::vector<uint8_t> &v, uint64_t x, int n) {
    for (int i = 0; i < n; i++) v.push_back((uint8_t)(x >> (8 * i)));
}

// DnsRecord::canonical_bytes (dns.rs:60-70) written to dst; returns the length
size_t dns_canonical_to(const DnsRec &r, uint8_t *dst) {
    size_t o = 0;
    memcpy(dst, "DNS-RECORD-V1", 13);
    o += 13;
    for (int i = 0; i < 4; i++) dst[o++] = (uint8_t)(r.domain_len >> (8 * i));
    memcpy(dst + o, r.domain, r.domain_len);
    o += r.domain_len;
    for (int i = 0; i < 2; i++) dst[o++] = (uint8_t)(r.record_type >> (8 * i));
    for (int i = 0; i < 4; i++) dst[o++] = (uint8_t)(r.ttl >> (8 * i));
    for (int i = 0; i < 4; i++) dst[o++] = (uint8_t)(r.rdata_len >> (8 * i));
    memcpy(dst + o, r.rdata, r.rdata_len);
    return o + r.rdata_len;
}

std::vector<uint8_t> dns_canonical(const DnsRec &r) {
    std::vector<uint8_t> canon;
    canon.reserve(27 + r.domain_len + r.rdata_len);
    const char *tag = "DNS-RECORD-V1";
    canon.insert(canon.end(), tag, tag + 13);
    put_le(canon, r.domain_len, 4);
    canon.insert(canon.end(), (const uint8_t *)r.domain, (const uint8_t *)r.domain + r.domain_len);
    put_le(canon, r.record_type, 2);
    put_le(canon, r.ttl, 4);
    put_le(canon, r.rdata_len, 4);
    canon.insert(canon.end(), r.rdata, r.rdata + r.rdata_len);
    return canon;
}

void dns_leaf_hash(const DnsRec &r, const uint8_t salt[16], uint8_t out[32]) {
    const std::vector<uint8_t> canon = dns_canonical(r);
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

// build_params (prover.rs:87-98) over make_schedule (:53-67): arity-8 STIR
// folds with a residual, or arity-2 classic folds (LdtMode::Fri); 54 queries
// unless the caller asks for another count
static FriParams swarm_fri_params(uint64_t n0, uint64_t queries, bool stir) {
    FriParams p;
    p.schedule = stir ? stir_schedule(n0) : std::vector<uint64_t>(ilog2(n0), 2);
    p.r = p.s0 = queries;
    p.seed_z = SWARM_SEED_Z;
    p.coeff_commit_final = true;
    p.d_final = 1;
    p.stir = stir;
    p.has_pi = true;
    return p;
}

std::vector<uint8_t> rollup_prove(Ctx *c, const std::vector<uint64_t> &leaves, const uint8_t binding[32],
                                  int phi_mode, uint8_t root_f0[32], bool stir) {
    uint64_t n_trace = 8;
    while (n_trace < leaves.size()) n_trace <<= 1;  // next_power_of_two().max(8)
    const uint64_t n0 = n_trace * SWARM_BLOWUP;
    std::vector<uint64_t> trace(4 * n_trace);
    build_hash_rollup_trace(n_trace, leaves.data(), leaves.size(), trace.data());
    Scoped d_tr(c, 8 * 4 * n_trace), d_lde(c, 8 * 4 * n0), d_c(c, 8 * n0);
    SG_HIP(hipMemcpyAsync(d_tr.p, trace.data(), 8 * trace.size(), hipMemcpyHostToDevice, c->stream));
    lde_columns(c, d_tr.as<uint64_t>(), d_lde.as<uint64_t>(), ilog2(n_trace), ilog2(SWARM_BLOWUP), 4, n_trace, n0, 1);
    merge_hash_rollup_dev(c, d_lde.as<uint64_t>(), n_trace, SWARM_BLOWUP, phi_mode, d_c.as<uint64_t>());
    FriParams p = swarm_fri_params(n0, SWARM_QUERIES, stir);
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

// One proof at a time: the batch path with a batch of one.
InnerShard prove_inner_shard(Ctx *c, const uint8_t salt[16], const std::vector<DnsRec> &recs,
                             const uint8_t fs_binding[32], int phi_mode, bool stir) {
    std::array<uint8_t, 32> f;
    memcpy(f.data(), fs_binding, 32);
    std::vector<InnerShard> v = prove_inner_shard_batch(c, salt, {recs}, {f}, phi_mode, stir);
    return std::move(v[0]);
}

std::vector<std::vector<uint8_t>> rollup_prove_batch(Ctx *c, const std::vector<std::vector<uint64_t>> &leaves,
                                                     const std::vector<std::array<uint8_t, 32>> &bindings,
                                                     int phi_mode, std::vector<std::array<uint8_t, 32>> &root_f0,
                                                     bool stir) {
    const uint32_t K = (uint32_t)leaves.size();
    SG_REQUIRE(K >= 1 && bindings.size() == K, "rollup batch: one binding per proof");
    uint64_t n_trace = 8;
    while (n_trace < leaves[0].size()) n_trace <<= 1;
    for (auto &l : leaves) {
        uint64_t nt = 8;
        while (nt < l.size()) nt <<= 1;
        SG_REQUIRE(nt == n_trace, "rollup batch: proofs of one batch must share n_trace");
    }
    const uint64_t n0 = n_trace * SWARM_BLOWUP;
    std::vector<uint64_t> trace(4 * n_trace * K);
    {
        HostScope h(c, "rollup_trace");
        parallel_for(K, [&](size_t k) {
            build_hash_rollup_trace(n_trace, leaves[k].data(), leaves[k].size(), &trace[4 * n_trace * k]);
        });
    }
    Scoped d_tr(c, 8 * trace.size()), d_lde(c, 8ULL * 4 * n0 * K), d_c(c, 8ULL * n0 * K);
    SG_HIP(hipMemcpyAsync(d_tr.p, trace.data(), 8 * trace.size(), hipMemcpyHostToDevice, c->stream));
    lde_columns(c, d_tr.as<uint64_t>(), d_lde.as<uint64_t>(), ilog2(n_trace), ilog2(SWARM_BLOWUP), 4ULL * K, n_trace,
                n0, 1);
    merge_hash_rollup_batch(c, d_lde.as<uint64_t>(), K, n_trace, SWARM_BLOWUP, phi_mode, d_c.as<uint64_t>());
    FriParams p = swarm_fri_params(n0, SWARM_QUERIES, stir);
    HostScope h(c, "deep_fri_prove_batch");
    std::vector<std::vector<uint8_t>> proofs = deep_fri_prove_batch(c, d_c.as<uint64_t>(), K, n0, p, &bindings, nullptr);
    root_f0.resize(K);
    for (uint32_t k = 0; k < K; k++) memcpy(root_f0[k].data(), proofs[k].data(), 32);
    return proofs;
}

std::vector<InnerShard> prove_inner_shard_batch(Ctx *c, const uint8_t salt[16],
                                                const std::vector<std::vector<DnsRec>> &shards,
                                                const std::vector<std::array<uint8_t, 32>> &fs_bindings,
                                                int phi_mode, bool stir) {
    const size_t K = shards.size();
    SG_REQUIRE(fs_bindings.size() == K, "prove_inner_shard_batch: one fs_binding per shard");
    std::vector<InnerShard> out(K);
    std::vector<std::vector<uint64_t>> leaves(K);
    for (auto &s : shards) SG_REQUIRE(!s.empty(), "prove_inner_shard: merkle tree needs >= 1 leaf");
    // record commitments (the step before the path): canonical bytes on the
    // host, leaf hashes and DNS trees on the device
    RecordSource rs;
    for (auto &s : shards) rs.counts.push_back(s.size());
    rs.size = [&](size_t k, size_t i) { return 27 + shards[k][i].domain_len + shards[k][i].rdata_len; };
    rs.write = [&](size_t k, size_t i, uint8_t *dst) { return dns_canonical_to(shards[k][i], dst); };
    std::vector<std::vector<uint8_t>> lhs;
    std::vector<std::array<uint8_t, 32>> droots;
    {
        HostScope h(c, "dns_commit");
        dns_commit_batch(c, rs, salt, lhs, droots);
    }
    parallel_for(K, [&](size_t k) {
        memcpy(out[k].merkle_root, droots[k].data(), 32);
        for (size_t i = 0; i < shards[k].size(); i++) pack_hash(&lhs[k][32 * i], leaves[k]);
        uint64_t nt = 8;
        while (nt < leaves[k].size()) nt <<= 1;
        out[k].n_trace = nt;
        out[k].record_count = shards[k].size();
    });
    // device: one lockstep batch per trace size
    std::map<uint64_t, std::vector<size_t>> groups;
    for (size_t k = 0; k < K; k++) groups[out[k].n_trace].push_back(k);
    for (auto &g : groups) {
        std::vector<std::vector<uint64_t>> lv;
        std::vector<std::array<uint8_t, 32>> bd;
        for (size_t k : g.second) {
            lv.push_back(std::move(leaves[k]));
            bd.push_back(fs_bindings[k]);
        }
        std::vector<std::array<uint8_t, 32>> rf;
        HostScope h(c, "rollup_prove_batch");
        std::vector<std::vector<uint8_t>> proofs = rollup_prove_batch(c, lv, bd, phi_mode, rf, stir);
        for (size_t j = 0; j < g.second.size(); j++) {
            InnerShard &o = out[g.second[j]];
            o.proof = std::move(proofs[j]);
            memcpy(o.root_f0, rf[j].data(), 32);
        }
    }
    for (size_t k = 0; k < K; k++) {
        InnerShard &o = out[k];
        Sha3 h;
        uint8_t cnt[8];
        for (int i = 0; i < 8; i++) cnt[i] = (uint8_t)((uint64_t)o.record_count >> (8 * i));
        h.update("DNS-SHARD-PIHASH-V1");
        h.update(salt, 16);
        h.update(cnt, 8);
        h.update(o.merkle_root, 32);
        h.update(o.root_f0, 32);
        h.update(fs_bindings[k].data(), 32);
        h.final(o.pi_hash);
    }
    return out;
}

// ds_ksk_pi_hash (prover.rs:446-462)
static void ds_ksk_pi_hash(const ByteSpan &key, const uint8_t parent[32], const uint8_t digest[32],
                           const uint8_t root_f0[32], const uint8_t fsb[32], uint8_t out[32]) {
    Sha3 h;
    uint8_t le[8];
    for (int i = 0; i < 8; i++) le[i] = (uint8_t)((uint64_t)key.n >> (8 * i));
    h.update("DS-KSK-PIHASH-V1");
    h.update(le, 8);
    h.update(key.p, key.n);
    h.update(parent, 32);
    h.update(digest, 32);
    h.update(root_f0, 32);
    h.update(fsb, 32);
    h.final(out);
}

std::vector<ShaProof> prove_sha_batch(Ctx *c, const std::vector<int> &algs, const std::vector<ByteSpan> &msgs,
                                      const std::vector<uint64_t> &heights,
                                      const std::vector<std::array<uint8_t, 32>> &fsbs, int phi_mode, bool stir) {
    const size_t K = msgs.size();
    SG_REQUIRE(algs.size() == K && heights.size() == K && fsbs.size() == K,
               "prove_sha: one alg, height and binding per message");
    std::vector<ShaProof> out(K);
    std::map<std::pair<int, uint64_t>, std::vector<size_t>> groups;  // (alg, trace height) -> inputs
    for (size_t k = 0; k < K; k++) {
        SG_REQUIRE(algs[k] == 256 || algs[k] == 512, "prove_sha: alg must be 256 or 512");
        SG_REQUIRE(msgs[k].p || msgs[k].n == 0, "prove_sha: null message");
        const uint64_t hmin = algs[k] == 256 ? sha256_trace_height(msgs[k].n) : sha512_trace_height(msgs[k].n);
        const uint64_t h = heights[k] ? heights[k] : hmin;
        SG_REQUIRE(is_pow2(h) && h >= hmin, "prove_sha: trace height below the message's minimum");
        SG_REQUIRE(h * SWARM_BLOWUP <= (1ULL << 24), "prove_sha: domain above 2^24");
        groups[{algs[k], h}].push_back(k);
    }
    size_t free_b = 0, total_b = 0;
    SG_HIP(hipMemGetInfo(&free_b, &total_b));
    const uint64_t budget = (uint64_t)((free_b + c->cached_bytes) * 0.6);
    // largest groups first, so the buffers of later waves fit in blocks the
    // caching allocator already holds
    std::vector<std::pair<std::pair<int, uint64_t>, std::vector<size_t>>> order(groups.begin(), groups.end());
    auto wbytes = [](const std::pair<int, uint64_t> &k) { return (k.first == 256 ? SHA256_W : SHA512_W) * k.second; };
    std::stable_sort(order.begin(), order.end(),
                     [&](const auto &a, const auto &b) { return wbytes(a.first) > wbytes(b.first); });
    for (auto &g : order) {
        const int alg = g.first.first;
        const uint64_t h = g.first.second, n0 = h * SWARM_BLOWUP, W = alg == 256 ? SHA256_W : SHA512_W;
        // per proof: trace + its coefficients + its blowup-2 LDE, the gate-group
        // polynomials at 2 h and n0 points, phi, FRI buffers
        const uint64_t per = 8ULL * W * 4 * h + 8ULL * 5 * 2 * h + 8ULL * 24 * n0;
        SG_REQUIRE(per <= budget, "prove_sha: one proof does not fit in device memory");
        // equal waves: every wave of a group reuses the same buffer sizes
        const size_t cap = (size_t)std::min<uint64_t>(4096, budget / per);
        const size_t nw = (g.second.size() + cap - 1) / cap;
        const size_t chunk = (g.second.size() + nw - 1) / nw;
        for (size_t c0 = 0; c0 < g.second.size(); c0 += chunk) {
            const std::vector<size_t> ids(g.second.begin() + c0,
                                          g.second.begin() + std::min(g.second.size(), c0 + chunk));
            const uint32_t G = (uint32_t)ids.size();
            std::vector<ByteSpan> gm(G);
            for (uint32_t j = 0; j < G; j++) gm[j] = msgs[ids[j]];
            std::vector<uint32_t> nb;
            std::vector<std::array<uint8_t, 64>> dgs;
            Scoped d_c(c, 8ULL * n0 * G);
            {
                // traces expanded on the device from the host's row words
                Scoped d_tr(c, 8ULL * W * h * G);
                {
                    HostScope hs(c, "sha_trace");
                    if (alg == 256)
                        build_sha256_traces_dev(c, gm, h, d_tr.as<uint64_t>(), nb, dgs);
                    else
                        build_sha512_traces_dev(c, gm, h, d_tr.as<uint64_t>(), nb, dgs);
                }
                for (uint32_t j = 0; j < G; j++) {
                    ShaProof &o = out[ids[j]];
                    memcpy(o.digest, dgs[j].data(), 64);
                    o.n_blocks = nb[j];
                    o.n_trace = h;
                }
                // the merge from the traces: gate-group polynomials on a
                // blowup-2 LDE, interpolated (no blowup-32 LDE is built)
                if (alg == 256)
                    merge_sha256_from_trace_batch(c, d_tr.as<uint64_t>(), G, h, SWARM_BLOWUP, nb, phi_mode,
                                                  d_c.as<uint64_t>());
                else
                    merge_sha512_from_trace_batch(c, d_tr.as<uint64_t>(), G, h, SWARM_BLOWUP, nb, phi_mode,
                                                  d_c.as<uint64_t>());
            }
            FriParams p = swarm_fri_params(n0, SWARM_QUERIES, stir);
            std::vector<std::array<uint8_t, 32>> pis(G);
            for (uint32_t j = 0; j < G; j++) pis[j] = fsbs[ids[j]];
            std::vector<std::vector<uint8_t>> proofs = deep_fri_prove_batch(c, d_c.as<uint64_t>(), G, n0, p, &pis, nullptr);
            for (uint32_t j = 0; j < G; j++) {
                ShaProof &o = out[ids[j]];
                o.proof = std::move(proofs[j]);
                memcpy(o.root_f0, o.proof.data(), 32);
            }
        }
    }
    return out;
}

std::vector<DsKsk> prove_ds_ksk_binding_batch(Ctx *c, const std::vector<ByteSpan> &keys,
                                              const std::vector<std::array<uint8_t, 32>> &parents,
                                              const std::vector<std::array<uint8_t, 32>> &fsbs, int phi_mode,
                                              bool stir) {
    const size_t K = keys.size();
    SG_REQUIRE(parents.size() == K && fsbs.size() == K, "prove_ds_ksk_binding: one parent hash and binding per key");
    std::vector<ShaProof> sp =
        prove_sha_batch(c, std::vector<int>(K, 256), keys, std::vector<uint64_t>(K, 0), fsbs, phi_mode, stir);
    std::vector<DsKsk> out(K);
    for (size_t k = 0; k < K; k++) {
        DsKsk &o = out[k];
        memcpy(o.asserted_digest, sp[k].digest, 32);
        memcpy(o.root_f0, sp[k].root_f0, 32);
        o.n_blocks = sp[k].n_blocks;
        o.n_trace = sp[k].n_trace;
        o.proof = std::move(sp[k].proof);
        ds_ksk_pi_hash(keys[k], parents[k].data(), o.asserted_digest, o.root_f0, fsbs[k].data(), o.pi_hash);
    }
    return out;
}

std::vector<Nsec3Proof> prove_nsec3_batch(Ctx *c, const std::vector<ByteSpan> &chains, const uint8_t salt[16],
                                          const std::vector<std::array<uint8_t, 32>> &fs_bindings, int phi_mode,
                                          bool stir) {
    const size_t K = chains.size();
    SG_REQUIRE(K >= 1 && fs_bindings.size() == K, "prove_nsec3: one fs_binding per chain");
    std::vector<Nsec3Proof> out(K);
    std::map<uint64_t, std::vector<size_t>> groups;  // n_trace -> chains
    for (size_t k = 0; k < K; k++) {
        const ByteSpan &ch = chains[k];
        SG_REQUIRE(ch.p && ch.n >= 64 && ch.n % 64 == 0, "prove_nsec3: NSEC3 chain must be non-empty");
        const size_t n = ch.n / 64;
        SG_REQUIRE(memcmp(ch.p + 64 * (n - 1) + 32, ch.p, 32) == 0,
                   "prove_nsec3: NSEC3 chain must be closed: records[n-1].next_hash != records[0].owner_hash");
        uint64_t nt = 8;
        while (nt < n) nt <<= 1;
        SG_REQUIRE(nt * SWARM_BLOWUP <= (1ULL << 24), "prove_nsec3: chain above 2^19 records");
        groups[nt].push_back(k);
        // chain root (prover.rs:372-381)
        Nsec3Proof &o = out[k];
        o.record_count = n;
        o.n_trace = nt;
        uint8_t cnt[8];
        for (int b = 0; b < 8; b++) cnt[b] = (uint8_t)((uint64_t)n >> (8 * b));
        Sha3 h;
        h.update("DNS-NSEC3-CHAIN-ROOT-V1", 23);
        h.update(salt, 16);
        h.update(cnt, 8);
        h.update(ch.p, ch.n);
        h.final(o.chain_root);
    }
    for (auto &g : groups) {
        const uint64_t nt = g.first, n0 = nt * SWARM_BLOWUP;
        const uint32_t G = (uint32_t)g.second.size();
        std::vector<uint64_t> trace(8 * nt * G);
        parallel_for(G, [&](size_t j) {
            const ByteSpan &ch = chains[g.second[j]];
            const size_t n = ch.n / 64;
            std::vector<uint64_t> limbs(8 * n);
            for (size_t q = 0; q < 8 * n; q++) {
                uint64_t v = 0;
                for (int b = 0; b < 8; b++) v |= (uint64_t)ch.p[8 * q + b] << (8 * b);
                limbs[q] = v;
            }
            build_nsec3_trace(nt, limbs.data(), n, &trace[8 * nt * j]);
        });
        Scoped d_tr(c, 8 * trace.size()), d_c(c, 8ULL * n0 * G);
        {
            Scoped d_lde(c, 8ULL * 8 * n0 * G);
            SG_HIP(hipMemcpyAsync(d_tr.p, trace.data(), 8 * trace.size(), hipMemcpyHostToDevice, c->stream));
            lde_columns(c, d_tr.as<uint64_t>(), d_lde.as<uint64_t>(), ilog2(nt), ilog2(SWARM_BLOWUP), 8ULL * G, nt,
                        n0, 1);
            merge_nsec3_batch(c, d_lde.as<uint64_t>(), G, nt, SWARM_BLOWUP, phi_mode, d_c.as<uint64_t>());
            SG_HIP(hipStreamSynchronize(c->stream));  // the host trace goes out of scope
        }
        FriParams p = swarm_fri_params(n0, SWARM_QUERIES, stir);
        std::vector<std::array<uint8_t, 32>> pis(G);
        for (uint32_t j = 0; j < G; j++) pis[j] = fs_bindings[g.second[j]];
        std::vector<std::vector<uint8_t>> proofs = deep_fri_prove_batch(c, d_c.as<uint64_t>(), G, n0, p, &pis, nullptr);
        for (uint32_t j = 0; j < G; j++) {
            Nsec3Proof &o = out[g.second[j]];
            o.proof = std::move(proofs[j]);
            memcpy(o.root_f0, o.proof.data(), 32);
            // pi_hash (prover.rs:431-438)
            uint8_t cnt[8];
            for (int b = 0; b < 8; b++) cnt[b] = (uint8_t)((uint64_t)o.record_count >> (8 * b));
            Sha3 h;
            h.update("DNS-NSEC3-PIHASH-V1", 19);
            h.update(salt, 16);
            h.update(cnt, 8);
            h.update(o.chain_root, 32);
            h.update(o.root_f0, 32);
            h.update(pis[j].data(), 32);
            h.final(o.pi_hash);
        }
    }
    return out;
}

std::vector<RsaProof> prove_rsa_batch(Ctx *c, const std::vector<RsaRecord> &records, uint32_t R, uint64_t h,
                                      const std::vector<std::array<uint8_t, 32>> &pis) {
    SG_REQUIRE(R >= 1 && R <= RSA_MAX_RECORDS, "prove_rsa: 1 to 4096 records per proof");
    SG_REQUIRE(!records.empty() && records.size() % R == 0, "prove_rsa: records must fill whole proofs");
    const size_t K = records.size() / R;
    SG_REQUIRE(pis.size() == K, "prove_rsa: one public-input hash per proof");
    SG_REQUIRE(is_pow2(h) && h >= 32 && h * SWARM_BLOWUP <= (1ULL << 24), "prove_rsa: n_trace a power of two >= 32");
    const uint64_t n0 = h * SWARM_BLOWUP, W = (uint64_t)RSA_W * R;
    size_t free_b = 0, total_b = 0;
    SG_HIP(hipMemGetInfo(&free_b, &total_b));
    const uint64_t budget = (uint64_t)((free_b + c->cached_bytes) * 0.6);
    // trace, its coefficients and blowup-2 LDE, the group polynomials, phi, FRI
    const uint64_t per = 8ULL * W * 4 * h + 8ULL * 3 * 2 * h + 8ULL * 24 * n0;
    SG_REQUIRE(per <= budget, "prove_rsa: one proof does not fit in device memory");
    const size_t cap = (size_t)std::min<uint64_t>(std::min<uint64_t>(4096, 65535 / R), budget / per);
    const size_t nw = (K + cap - 1) / cap, chunk = (K + nw - 1) / nw;
    std::vector<RsaProof> out(K);
    for (size_t k0 = 0; k0 < K; k0 += chunk) {
        const uint32_t G = (uint32_t)std::min(chunk, K - k0);
        Scoped d_c(c, 8ULL * n0 * G);
        {
            Scoped d_tr(c, 8ULL * W * h * G);
            {
                HostScope hs(c, "rsa_trace");
                const std::vector<RsaRecord> part(records.begin() + k0 * R, records.begin() + (k0 + G) * R);
                build_rsa_traces_dev(c, part, R, h, d_tr.as<uint64_t>());
            }
            // the merge from the traces: gate-group polynomials on a blowup-2
            // LDE, interpolated (no blowup-32 LDE is built)
            HostScope hs(c, "rsa_lde_merge");
            merge_rsa_from_trace_batch(c, d_tr.as<uint64_t>(), G, R, h, SWARM_BLOWUP, d_c.as<uint64_t>());
            if (c->profiling) SG_HIP(hipStreamSynchronize(c->stream));
        }
        HostScope hf(c, "rsa_fri");
        FriParams p = swarm_fri_params(n0, SWARM_QUERIES, true);
        std::vector<std::array<uint8_t, 32>> gp(pis.begin() + k0, pis.begin() + k0 + G);
        std::vector<std::vector<uint8_t>> proofs = deep_fri_prove_batch(c, d_c.as<uint64_t>(), G, n0, p, &gp, nullptr);
        for (uint32_t j = 0; j < G; j++) {
            RsaProof &o = out[k0 + j];
            o.proof = std::move(proofs[j]);
            o.n_trace = h;
            memcpy(o.root_f0, o.proof.data(), 32);
            Sha3 s;
            s.update(gp[j].data(), 32);
            s.update(o.root_f0, 32);
            s.final(o.pi_hash);
        }
    }
    return out;
}

std::vector<P256Proof> prove_p256_batch(Ctx *c, const std::vector<P256Chains> &in, uint64_t h,
                                        const std::vector<std::array<uint8_t, 32>> &pis) {
    const size_t K = in.size();
    SG_REQUIRE(K >= 1 && K <= 65535, "prove_p256: 1 to 65535 proofs");
    SG_REQUIRE(pis.size() == K, "prove_p256: one public-input hash per proof");
    SG_REQUIRE(is_pow2(h) && h >= 2 && h * SWARM_BLOWUP <= (1ULL << 20), "prove_p256: n_trace a power of two >= 2");
    const uint64_t n0 = h * SWARM_BLOWUP, W = P256_W;
    size_t free_b = 0, total_b = 0;
    SG_HIP(hipMemGetInfo(&free_b, &total_b));
    const uint64_t budget = (uint64_t)((free_b + c->cached_bytes) * 0.6);
    // trace, its LDE coefficients and blowup-2 LDE, the gadget partial sums,
    // the tail LDE, phi, FRI
    const uint64_t per = h >= 32 ? 8ULL * W * 3 * h + 8ULL * 300 * 2 * h + 8ULL * 200 * n0 + 8ULL * 16 * n0
                                 : 8ULL * W * (3 * h + n0) + 8ULL * 300 * n0 + 8ULL * 16 * n0;
    SG_REQUIRE(per <= budget, "prove_p256: one proof does not fit in device memory");
    const size_t cap = (size_t)std::min<uint64_t>(4096, budget / per);
    const size_t nw = (K + cap - 1) / cap, chunk = (K + nw - 1) / nw;
    std::vector<P256Proof> out(K);
    for (size_t k0 = 0; k0 < K; k0 += chunk) {
        const uint32_t G = (uint32_t)std::min(chunk, K - k0);
        Scoped d_c(c, 8ULL * n0 * G);
        {
            std::vector<std::array<Be32, 6>> rp;
            Scoped d_tr(c, 8ULL * W * h * G);
            // the odd half of the blowup-2 LDE, written by the trace expansion itself
            const bool fused = h >= 32 && h <= 4096;
            Scoped d_odd(c, fused ? 8ULL * W * h * G : 0);
            {
                HostScope hs(c, "p256_trace");
                const std::vector<P256Chains> part(in.begin() + k0, in.begin() + k0 + G);
                build_p256_traces_dev(c, part, h, d_tr.as<uint64_t>(), &rp, fused ? d_odd.as<uint64_t>() : nullptr);
            }
            HostScope hs(c, "p256_lde_merge");
            if (h >= 32) {
                merge_p256_from_trace_batch(c, d_tr.as<uint64_t>(), G, h, SWARM_BLOWUP, d_c.as<uint64_t>(),
                                            fused ? d_odd.as<uint64_t>() : nullptr);
            } else {  // below 64 points of <w_{2 n_trace}>: compose on the full LDE
                Scoped d_lde(c, 8ULL * W * n0 * G);
                lde_columns(c, d_tr.as<uint64_t>(), d_lde.as<uint64_t>(), ilog2(h), ilog2(SWARM_BLOWUP), W * G, h,
                            n0, 1);
                merge_p256_batch(c, d_lde.as<uint64_t>(), G, h, SWARM_BLOWUP, d_c.as<uint64_t>());
            }
            if (c->profiling) SG_HIP(hipStreamSynchronize(c->stream));
            for (uint32_t j = 0; j < G; j++) out[k0 + j].rproj = rp[j];
        }
        HostScope hf(c, "p256_fri");
        FriParams p = swarm_fri_params(n0, SWARM_QUERIES, true);
        std::vector<std::array<uint8_t, 32>> gp(pis.begin() + k0, pis.begin() + k0 + G);
        std::vector<std::vector<uint8_t>> proofs = deep_fri_prove_batch(c, d_c.as<uint64_t>(), G, n0, p, &gp, nullptr);
        for (uint32_t j = 0; j < G; j++) {
            P256Proof &o = out[k0 + j];
            o.proof = std::move(proofs[j]);
            o.n_trace = h;
            memcpy(o.root_f0, o.proof.data(), 32);
            Sha3 s;
            s.update(gp[j].data(), 32);
            s.update(o.root_f0, 32);
            s.final(o.pi_hash);
        }
    }
    return out;
}

std::vector<P256Proof> prove_ecdsa_p256_batch(Ctx *c, const std::vector<P256Record> &recs) {
    const size_t K = recs.size();
    SG_REQUIRE(K >= 1 && K <= 65535, "prove_ecdsa_p256: 1 to 65535 records");
    std::vector<P256Chains> in(K);
    std::vector<std::array<uint8_t, 32>> pis(K);
    std::vector<std::string> err(K);
    {
        HostScope hs(c, "p256_ecdsa_inputs");
        parallel_for(K, [&](size_t k) {
            try {
                in[k] = p256_ecdsa_chains(recs[k], false);
            } catch (const std::exception &e) {
                err[k] = e.what();
            }
            p256_pk_hash_pre(recs[k], pis[k].data());
        });
    }
    for (size_t k = 0; k < K; k++)
        if (!err[k].empty()) throw Error(E_INVALID, "prove_ecdsa_p256: record " + std::to_string(k) + ": " + err[k]);
    // The signature checks (a double scalar multiplication each, about
    // 0.4 ms) run on a thread of their own beside the proof pipeline, which
    // needs only u1 and u2; a record that does not verify fails the call
    // when both are done, and its error comes before any other.
    std::vector<std::string> verr(K);
    std::thread checker([&] {
        for (size_t k = 0; k < K; k++) {
            try {
                (void)p256_ecdsa_chains(recs[k], true);
            } catch (const std::exception &e) {
                verr[k] = e.what();
            }
        }
    });
    auto first_bad = [&]() {
        for (size_t k = 0; k < K; k++)
            if (!verr[k].empty())
                throw Error(E_INVALID, "prove_ecdsa_p256: record " + std::to_string(k) + ": " + verr[k]);
    };
    std::vector<P256Proof> out;
    try {
        out = prove_p256_batch(c, in, P256_ECDSA_ROWS, pis);
    } catch (...) {
        checker.join();
        first_bad();
        throw;
    }
    checker.join();
    first_bad();
    return out;
}

std::vector<uint8_t> prove_outer_rollup(Ctx *c, const uint8_t *pi_hashes, size_t n, const uint8_t pk_hash[32],
                                        int phi_mode, uint8_t root_f0[32], bool stir) {
    std::vector<uint64_t> leaves;
    for (size_t i = 0; i < n; i++) pack_hash(pi_hashes + 32 * i, leaves);
    return rollup_prove(c, leaves, pk_hash, phi_mode, root_f0, stir);
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
    // the row values and the row-tree openings behind one wait
    std::vector<GatherJob> jobs(2);
    jobs[0].src = d_lde;
    for (uint64_t q : raw)
        for (uint64_t col = 0; col < w; col++) jobs[0].idx.push_back(col * n0 + q);
    const std::vector<std::vector<uint64_t>> ridx = {raw};
    jobs[1] = tree_open_job(rows, ridx);
    gather_set(c, jobs);
    const HostBuf &vals = jobs[0].out;
    std::vector<std::vector<std::vector<uint8_t>>> opk;
    tree_open_finish(rows, ridx, jobs[1], opk);
    const std::vector<std::vector<uint8_t>> &ops = opk[0];
    tree_free(c, rows);
    put_le(out, raw.size(), 8);
    for (size_t q = 0; q < raw.size(); q++) {
        put_le(out, w, 8);
        for (uint64_t col = 0; col < w; col++) put_le(out, vals[q * w + col], 8);
        out.insert(out.end(), ops[q].begin(), ops[q].end());
    }
    return out;
}

std::vector<ZskKsk> prove_zsk_ksk_binding_v2_batch(Ctx *c, const std::vector<EdSig> &sigs,
                                                   const std::vector<std::array<uint8_t, 32>> &fs_bindings,
                                                   const std::vector<std::array<uint8_t, 32>> &merkle_roots,
                                                   uint32_t k, uint64_t blowup, uint32_t queries, bool stir) {
    const size_t K = sigs.size();
    SG_REQUIRE(K >= 1 && K <= 65535, "prove_zsk_ksk_binding_v2: 1 to 65535 signatures");
    SG_REQUIRE(fs_bindings.size() == K && merkle_roots.size() == K,
               "prove_zsk_ksk_binding_v2: one fs binding and merkle root per signature");
    SG_REQUIRE(is_pow2(blowup) && blowup >= 4 && blowup <= 64, "prove_zsk_ksk_binding_v2: blowup in {4, ..., 64}");
    SG_REQUIRE(queries >= 1 && queries <= 512, "prove_zsk_ksk_binding_v2: 1 to 512 queries");
    {
        HostScope hs(c, "ed25519_verify");
        std::vector<uint8_t> ok(K, 0);
        parallel_for(K, [&](size_t i) { ok[i] = ed25519_verify(sigs[i]); });
        for (size_t i = 0; i < K; i++)
            SG_REQUIRE(ok[i], "prove_zsk_ksk_binding_v2: signature " + std::to_string(i) +
                                  " does not verify (the prover must not commit to an invalid signature)");
    }
    // groups of one trace shape (the SHA-512 row count of R || A || M)
    std::map<uint64_t, std::vector<size_t>> groups;
    for (size_t i = 0; i < K; i++) groups[ed25519_shape(sigs[i].len, k).k_hash].push_back(i);
    std::vector<ZskKsk> out(K);
    for (auto &gr : groups) {
        const EdShape sh = ed25519_shape(sigs[gr.second[0]].len, k);
        const uint64_t h = sh.height, W = sh.width, n0 = h * blowup;
        size_t free_b = 0, total_b = 0;
        SG_HIP(hipMemGetInfo(&free_b, &total_b));
        const uint64_t budget = (uint64_t)((free_b + c->cached_bytes) * 0.6);
        // trace and its odd coset, gate-group sums, phi and FRI
        const uint64_t per = 16ULL * W * h + 8ULL * 400 * 2 * h + 8ULL * 64 * n0;
        SG_REQUIRE(per <= budget, "prove_zsk_ksk_binding_v2: one proof does not fit in device memory");
        const size_t cap = (size_t)std::min<uint64_t>(1024, budget / per);
        const std::vector<size_t> &idx = gr.second;
        const size_t nw = (idx.size() + cap - 1) / cap, chunk = (idx.size() + nw - 1) / nw;
        for (size_t k0 = 0; k0 < idx.size(); k0 += chunk) {
            const uint32_t G = (uint32_t)std::min(chunk, idx.size() - k0);
            std::vector<EdSig> part(G);
            std::vector<std::array<uint8_t, 32>> pis(G);
            for (uint32_t j = 0; j < G; j++) {
                part[j] = sigs[idx[k0 + j]];
                pis[j] = fs_bindings[idx[k0 + j]];
            }
            Scoped d_c(c, 8ULL * n0 * G);
            {
                Scoped d_tr(c, 8ULL * W * h * G);
                EdConsts pins;
                {
                    HostScope hs(c, "ed25519_trace");
                    build_ed25519_traces_dev(c, part, k, d_tr.as<uint64_t>(), &pins);
                }
                HostScope hs(c, "ed25519_merge");
                merge_ed25519_from_trace_batch(c, part, k, d_tr.as<uint64_t>(), blowup, nullptr, d_c.as<uint64_t>(),
                                               &pins);
                if (c->profiling) SG_HIP(hipStreamSynchronize(c->stream));
            }
            HostScope hf(c, "ed25519_fri");
            FriParams p = swarm_fri_params(n0, queries, stir);
            std::vector<std::vector<uint8_t>> proofs = deep_fri_prove_batch(c, d_c.as<uint64_t>(), G, n0, p, &pis, nullptr);
            for (uint32_t j = 0; j < G; j++) {
                const size_t i = idx[k0 + j];
                const EdSig &s = sigs[i];
                ZskKsk &o = out[i];
                o.proof = std::move(proofs[j]);
                o.n_trace = h;
                memcpy(o.root_f0, o.proof.data(), 32);
                // zsk_ksk_pi_hash_v2_stark (prover.rs:763-781)
                uint8_t le[8];
                for (int b = 0; b < 8; b++) le[b] = (uint8_t)((uint64_t)s.len >> (8 * b));
                Sha3 hsh;
                hsh.update((const uint8_t *)"ZSK-KSK-PIHASH-V2-ST", 20);
                hsh.update(s.pk, 32);
                hsh.update(s.sig, 64);
                hsh.update(le, 8);
                if (s.len) hsh.update(s.msg, s.len);
                hsh.update(fs_bindings[i].data(), 32);
                hsh.update(merkle_roots[i].data(), 32);
                hsh.update(o.root_f0, 32);
                hsh.final(o.pi_hash);
            }
        }
    }
    return out;
}

}  // namespace sg
