// keccak.hpp -- SHA3-256 Merkle trees on the device, host SHA3 / transcript /
// ChaCha12 used by the proof orchestration.
#pragma once
#include <stdint.h>
#include <string.h>

#include <array>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "ctx.hpp"

namespace sg {

// ---------------------------------------------------------------- host SHA3
struct Sha3 {
    uint64_t s[25];
    size_t pos;
    Sha3() { reset(); }
    void reset() {
        memset(s, 0, sizeof s);
        pos = 0;
    }
    void update(const void *d, size_t n);
    void update(const std::string &str) { update(str.data(), str.size()); }
    void final(uint8_t out[32]) const;  // leaves the state untouched
};
void keccak_f1600_host(uint64_t s[25]);
void sha3_256_host(const void *d, size_t n, uint8_t out[32]);

// Fiat-Shamir transcript over SHA3-256 (crates/transcript/src/lib.rs:162-212):
// init "FSv1-TRANSCRIPT-INIT"||label, absorb "FSv1-ABSORB-BYTES"||bytes,
// challenges hash a clone with "FSv1-CHALLENGE"||label and leave the state as is.
struct Transcript {
    Sha3 h;
    explicit Transcript(const std::string &label);
    void absorb(const void *b, size_t n);
    void absorb(const std::string &s) { absorb(s.data(), s.size()); }
    void absorb_field(uint64_t x);
    void challenge_bytes(const std::string &label, uint8_t out[32]) const;
    uint64_t challenge_field(const std::string &label) const;  // fri.rs:73-85 (rchunks(7))
    void challenge_ext(const std::string &tag, uint64_t out[6]) const;  // fri.rs:1372-1386
};
uint64_t hash_fields_tagged(const std::string &tag, const uint64_t *f, size_t n);  // fri.rs:123-130
// rand 0.8 StdRng (ChaCha12, key = LE64(seed) || 0^24): first next_u64()
uint64_t stdrng_first_u64(uint64_t seed);

// ---------------------------------------------------------------- device trees
// Levels of one tree, resident on the device. Digests are 4 u64 (LE lanes).
// `batch` trees of identical shape, built together. Level l of proof k holds
// real[l] digests at buf[4 * (off[l] + k * real[l])].
struct DevTree {
    uint32_t arity = 2;
    uint64_t label = 0;
    uint32_t batch = 1;
    std::vector<uint64_t> len;  // nodes per level (one tree) after padding to a multiple of arity
    std::vector<uint64_t> real; // nodes per level (one tree) before padding
    std::vector<uint64_t> off;  // digest offset of each level (all trees)
    uint64_t *buf = nullptr;    // all digests
    std::vector<std::array<uint8_t, 32>> roots;  // per tree
    uint8_t root[32];                            // tree 0
};

// Leaf i hashes DsLabel(arity, u32::MAX, i, label) || LE64(vals[base + i*leaf_stride + c*cs]) for c < w.
// Values c >= nz are zero (a base-field layer hashed as 6 components).
// combo_out (optional): per leaf, sum_c (c+1) * value_c written alongside the
// digest -- the prover-core f0 fused into the row-leaf pass.
struct LeafSrc {
    const uint64_t *vals;
    uint64_t leaf_stride;
    uint64_t cs;
    uint32_t w;
    uint32_t nz;
    uint64_t *combo_out = nullptr;
    uint64_t proof_stride = 0;  // value offset between the trees of a batch
};

// Build `batch` trees of n_leaves leaves each, constant arity
// (crates/merkle/src/lib.rs:377-476); tree k reads its values at
// src.vals + k * src.proof_stride. With read_roots false the kernels are only
// queued and tree_read_roots fills t.roots later (one wait for several
// results).
void tree_build(Ctx *c, DevTree &t, const LeafSrc &src, uint64_t n_leaves, uint32_t arity, uint32_t depth,
                uint64_t label, uint32_t batch = 1, bool read_roots = true);
void tree_read_roots(Ctx *c, DevTree &t);
// the same, with n more bytes from device src copied to host dst behind the
// same wait
void tree_read_roots_with(Ctx *c, DevTree &t, void *dst, const void *src, size_t n);
void tree_free(Ctx *c, DevTree &t);
// Serialized MerkleOpenings (leaf[32], path Vec<Vec<[32]>>, index u64): idx[k]
// lists the leaf indices to open in tree k; out[k][q] is one opening.
void tree_open_multi(Ctx *c, const DevTree &t, const std::vector<std::vector<uint64_t>> &idx,
                     std::vector<std::vector<std::vector<uint8_t>>> &out);
void tree_open_batch(Ctx *c, const DevTree &t, const std::vector<uint64_t> &idx, std::vector<std::vector<uint8_t>> &out);

// A host array of u64 that is not zero-filled when sized (gather_set writes
// every element).
struct HostBuf {
    std::unique_ptr<uint64_t[]> p;
    size_t n = 0;
    void reset(size_t m) {
        p.reset(m ? new uint64_t[m] : nullptr);
        n = m;
    }
    uint64_t &operator[](size_t i) { return p[i]; }
    const uint64_t &operator[](size_t i) const { return p[i]; }
    const uint64_t *data() const { return p.get(); }
    const uint64_t *begin() const { return p.get(); }
    const uint64_t *end() const { return p.get() + n; }
    size_t size() const { return n; }
};

// Several gathers behind one index upload and one wait: job j reads `width`
// u64 per index from src (device; row r at src + r * width) into out.
struct GatherJob {
    const uint64_t *src = nullptr;
    uint32_t width = 1;
    std::vector<uint64_t> idx;  // host row indices
    HostBuf out;                // filled by gather_set: idx.size() * width
};
void gather_set(Ctx *c, std::vector<GatherJob> &jobs);
// tree_open_multi in two halves around a gather_set: the digest gather of the
// openings of leaves idx[k] of tree k, then the serialized openings from it
GatherJob tree_open_job(const DevTree &t, const std::vector<std::vector<uint64_t>> &idx);
void tree_open_finish(const DevTree &t, const std::vector<std::vector<uint64_t>> &idx, const GatherJob &job,
                      std::vector<std::vector<std::vector<uint8_t>>> &out);

// gather out[r] = src[idx[r]] for u64 elements (device index list)
void gather_u64(Ctx *c, const uint64_t *src, const uint64_t *d_idx, uint64_t n, uint64_t *out);
// gather rows of `width` u64: out[r*width + l] = src[idx[r]*width + l]
void gather_u64_w(Ctx *c, const uint64_t *src, const uint64_t *d_idx, uint64_t n, uint32_t width, uint64_t *out);

// Records of K shards: counts[k] records in shard k; size(k, i) is the length
// of DnsRecord::canonical_bytes of record i and write(k, i, dst) writes them.
struct RecordSource {
    std::vector<size_t> counts;
    std::function<size_t(size_t, size_t)> size;
    std::function<size_t(size_t, size_t, uint8_t *)> write;
};
// DNS record commitments of K shards on the device (crates/swarm-dns/src/dns.rs:60-118):
// the doubly salted leaf hashes (32 bytes each) and each shard's DNS-NODE-V1 root.
void dns_commit_batch(Ctx *c, const RecordSource &src, const uint8_t salt[16],
                      std::vector<std::vector<uint8_t>> &leaf_hashes, std::vector<std::array<uint8_t, 32>> &roots);

// STIR query indices of K proofs on the device (fri.rs:1813-1893): raw[k][q] =
// (StdRng(hash_fields_tagged("FRI/index", [seeds[k], 0, q])).next_u64() & (npow2-1)) % n0
void fri_query_indices(Ctx *c, const std::vector<uint64_t> &seeds, uint64_t s0, uint64_t n0,
                       std::vector<std::vector<uint64_t>> &raw);

uint32_t merkle_depth(uint64_t leaves, uint32_t arity);    // fri.rs:950-959
uint32_t pick_arity(uint64_t n, uint64_t requested);       // fri.rs:1094-1103 (.max(2) applied)

}  // namespace sg
