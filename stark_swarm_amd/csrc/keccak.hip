// keccak.hip -- Keccak-f[1600] / SHA3-256 Merkle commitments on gfx950 and the
// host-side SHA3 transcript used by the proof orchestration.
//
// Tree semantics follow crates/merkle/src/lib.rs:95-238,377-512: a leaf is
// SHA3-256(DsLabel(arity, u32::MAX, position, label) || LE64 values), a node is
// SHA3-256(DsLabel(arity, level, position, label) || children), and a level
// whose length is not a multiple of the arity is padded by repeating its last
// node. Every digest is 4 u64 lanes (the first 32 bytes of the state).
//
// One leaf or node per lane: the 25-lane state lives in VGPRs, message lanes
// are read straight from HBM (consecutive lanes read consecutive leaves, so
// each value column is a coalesced stream).
#include "gl.hpp"
#include "keccak.hpp"

namespace sg {

#define KECCAK_RC_LIST                                                                                         \
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,              \
        0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,          \
        0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,          \
        0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,          \
        0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,          \
        0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL

__device__ constexpr uint64_t KRC_D[24] = {KECCAK_RC_LIST};
static const uint64_t KRC_H[24] = {KECCAK_RC_LIST};

SG_HD uint64_t rotl64(uint64_t x, int n) { return n == 0 ? x : (x << n) | (x >> (64 - n)); }

// One round on 25 named lanes: theta, rho+pi, chi, iota.
SG_HD void keccak_round(uint64_t *a, uint64_t rc) {
    uint64_t c0 = a[0] ^ a[5] ^ a[10] ^ a[15] ^ a[20];
    uint64_t c1 = a[1] ^ a[6] ^ a[11] ^ a[16] ^ a[21];
    uint64_t c2 = a[2] ^ a[7] ^ a[12] ^ a[17] ^ a[22];
    uint64_t c3 = a[3] ^ a[8] ^ a[13] ^ a[18] ^ a[23];
    uint64_t c4 = a[4] ^ a[9] ^ a[14] ^ a[19] ^ a[24];
    uint64_t d0 = c4 ^ rotl64(c1, 1), d1 = c0 ^ rotl64(c2, 1), d2 = c1 ^ rotl64(c3, 1);
    uint64_t d3 = c2 ^ rotl64(c4, 1), d4 = c3 ^ rotl64(c0, 1);
    // rho + pi: b[y + 5*((2x+3y)%5)] = rotl(a[x + 5y] ^ d[x], r[x][y])
    uint64_t b[25];
    b[0] = a[0] ^ d0;
    b[10] = rotl64(a[1] ^ d1, 1);
    b[20] = rotl64(a[2] ^ d2, 62);
    b[5] = rotl64(a[3] ^ d3, 28);
    b[15] = rotl64(a[4] ^ d4, 27);
    b[16] = rotl64(a[5] ^ d0, 36);
    b[1] = rotl64(a[6] ^ d1, 44);
    b[11] = rotl64(a[7] ^ d2, 6);
    b[21] = rotl64(a[8] ^ d3, 55);
    b[6] = rotl64(a[9] ^ d4, 20);
    b[7] = rotl64(a[10] ^ d0, 3);
    b[17] = rotl64(a[11] ^ d1, 10);
    b[2] = rotl64(a[12] ^ d2, 43);
    b[12] = rotl64(a[13] ^ d3, 25);
    b[22] = rotl64(a[14] ^ d4, 39);
    b[23] = rotl64(a[15] ^ d0, 41);
    b[8] = rotl64(a[16] ^ d1, 45);
    b[18] = rotl64(a[17] ^ d2, 15);
    b[3] = rotl64(a[18] ^ d3, 21);
    b[13] = rotl64(a[19] ^ d4, 8);
    b[14] = rotl64(a[20] ^ d0, 18);
    b[24] = rotl64(a[21] ^ d1, 2);
    b[9] = rotl64(a[22] ^ d2, 61);
    b[19] = rotl64(a[23] ^ d3, 56);
    b[4] = rotl64(a[24] ^ d4, 14);
#pragma unroll
    for (int y = 0; y < 25; y += 5) {
        uint64_t e0 = b[y], e1 = b[y + 1], e2 = b[y + 2], e3 = b[y + 3], e4 = b[y + 4];
        a[y] = e0 ^ (~e1 & e2);
        a[y + 1] = e1 ^ (~e2 & e3);
        a[y + 2] = e2 ^ (~e3 & e4);
        a[y + 3] = e3 ^ (~e4 & e0);
        a[y + 4] = e4 ^ (~e0 & e1);
    }
    a[0] ^= rc;
}

// Device permutation on 32-bit halves: a 64-bit rotate is two v_alignbit_b32,
// the column parities use three-input XORs, and the rho/pi lane permutation
// is a renaming of registers.
__device__ constexpr int KROT_D[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43,
                                       25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};

__device__ __forceinline__ void rotl2(uint32_t lo, uint32_t hi, int n, uint32_t &olo, uint32_t &ohi) {
    if (n == 0) {
        olo = lo;
        ohi = hi;
    } else if (n < 32) {
        ohi = __builtin_amdgcn_alignbit(hi, lo, 32 - n);
        olo = __builtin_amdgcn_alignbit(lo, hi, 32 - n);
    } else if (n == 32) {
        olo = hi;
        ohi = lo;
    } else {
        const int m = n - 32;
        ohi = __builtin_amdgcn_alignbit(lo, hi, 32 - m);
        olo = __builtin_amdgcn_alignbit(hi, lo, 32 - m);
    }
}

__device__ __forceinline__ void keccak_round32(uint32_t *lo, uint32_t *hi, uint64_t rc) {
    uint32_t clo[5], chi[5];
#pragma unroll
    for (int x = 0; x < 5; x++) {
        // 0x96: three-input XOR truth table
        clo[x] = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_bitop3_b32(lo[x], lo[x + 5], lo[x + 10], 0x96),
                                             lo[x + 15], lo[x + 20], 0x96);
        chi[x] = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_bitop3_b32(hi[x], hi[x + 5], hi[x + 10], 0x96),
                                             hi[x + 15], hi[x + 20], 0x96);
    }
    // D[x] = C[x-1] ^ rot(C[x+1], 1) is folded into one three-input XOR per
    // state half: A[x, y] ^ C[x-1] ^ rot(C[x+1], 1)
    uint32_t rlo[5], rhi[5];
#pragma unroll
    for (int x = 0; x < 5; x++) rotl2(clo[(x + 1) % 5], chi[(x + 1) % 5], 1, rlo[x], rhi[x]);
    uint32_t blo[25], bhi[25];
#pragma unroll
    for (int x = 0; x < 5; x++)
#pragma unroll
        for (int y = 0; y < 5; y++) {
            const int src = x + 5 * y, dst = y + 5 * ((2 * x + 3 * y) % 5);
            rotl2(__builtin_amdgcn_bitop3_b32(lo[src], clo[(x + 4) % 5], rlo[x], 0x96),
                  __builtin_amdgcn_bitop3_b32(hi[src], chi[(x + 4) % 5], rhi[x], 0x96), KROT_D[src], blo[dst],
                  bhi[dst]);
        }
#pragma unroll
    for (int y = 0; y < 25; y += 5)
#pragma unroll
        for (int x = 0; x < 5; x++) {
            lo[y + x] = blo[y + x] ^ (~blo[y + (x + 1) % 5] & blo[y + (x + 2) % 5]);
            hi[y + x] = bhi[y + x] ^ (~bhi[y + (x + 1) % 5] & bhi[y + (x + 2) % 5]);
        }
    lo[0] ^= (uint32_t)rc;
    hi[0] ^= (uint32_t)(rc >> 32);
}

__device__ __forceinline__ void keccak_f_dev(uint64_t *a) {
    uint32_t lo[25], hi[25];
#pragma unroll
    for (int i = 0; i < 25; i++) {
        lo[i] = (uint32_t)a[i];
        hi[i] = (uint32_t)(a[i] >> 32);
    }
#pragma unroll 2
    for (int r = 0; r < 24; r++) keccak_round32(lo, hi, KRC_D[r]);
#pragma unroll
    for (int i = 0; i < 25; i++) a[i] = ((uint64_t)hi[i] << 32) | lo[i];
}

// ---- Keccak-f on two lanes: lane parity h holds the low (h = 0) or high
// (h = 1) 32-bit half of every state word, so a permutation's instruction
// chain is about two thirds as long as on one lane (small tree levels, whose
// latency is the permutation chain). A 64-bit rotation needs the partner's
// half, fetched by a DPP quad permutation [1, 0, 3, 2] (lane i reads lane
// i ^ 1); with m the lane's half and o the partner's, rotl(x, n) is
// alignbit(m, o, 32 - n) for n < 32 and alignbit(o, m, 64 - n) for n > 32 on
// both lanes.
__device__ __forceinline__ uint32_t kx_partner(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t kx_rot(uint32_t m, int n) {
    if (n == 0) return m;
    const uint32_t o = kx_partner(m);
    if (n == 32) return o;
    return n < 32 ? __builtin_amdgcn_alignbit(m, o, 32 - n) : __builtin_amdgcn_alignbit(o, m, 64 - n);
}
__device__ __forceinline__ void keccak_round_half(uint32_t *a, uint32_t rc) {
    uint32_t c[5];
#pragma unroll
    for (int x = 0; x < 5; x++)
        c[x] = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_bitop3_b32(a[x], a[x + 5], a[x + 10], 0x96), a[x + 15],
                                           a[x + 20], 0x96);
    uint32_t r[5];
#pragma unroll
    for (int x = 0; x < 5; x++) r[x] = kx_rot(c[(x + 1) % 5], 1);
    uint32_t b[25];
#pragma unroll
    for (int x = 0; x < 5; x++)
#pragma unroll
        for (int y = 0; y < 5; y++) {
            const int src = x + 5 * y, dst = y + 5 * ((2 * x + 3 * y) % 5);
            b[dst] = kx_rot(__builtin_amdgcn_bitop3_b32(a[src], c[(x + 4) % 5], r[x], 0x96), KROT_D[src]);
        }
#pragma unroll
    for (int y = 0; y < 25; y += 5)
#pragma unroll
        for (int x = 0; x < 5; x++) a[y + x] = b[y + x] ^ (~b[y + (x + 1) % 5] & b[y + (x + 2) % 5]);
    a[0] ^= rc;
}
// every lane of the wave must run this (the DPP reads the partner lane)
__device__ __forceinline__ void keccak_f_half(uint32_t *a, uint32_t h) {
#pragma unroll 2
    for (int r = 0; r < 24; r++) keccak_round_half(a, (uint32_t)(KRC_D[r] >> (32 * h)));
}

void keccak_f1600_host(uint64_t s[25]) {
    for (int r = 0; r < 24; r++) keccak_round(s, KRC_H[r]);
}

// ---------------------------------------------------------------- host SHA3
void Sha3::update(const void *d, size_t n) {
    const uint8_t *p = (const uint8_t *)d;
    for (size_t i = 0; i < n; i++) {
        s[pos >> 3] ^= (uint64_t)p[i] << (8 * (pos & 7));
        if (++pos == 136) {
            keccak_f1600_host(s);
            pos = 0;
        }
    }
}

void Sha3::final(uint8_t out[32]) const {
    uint64_t t[25];
    memcpy(t, s, sizeof t);
    t[pos >> 3] ^= 0x06ULL << (8 * (pos & 7));
    t[16] ^= 0x8000000000000000ULL;
    keccak_f1600_host(t);
    for (int i = 0; i < 32; i++) out[i] = (uint8_t)(t[i >> 3] >> (8 * (i & 7)));
}

void sha3_256_host(const void *d, size_t n, uint8_t out[32]) {
    Sha3 h;
    h.update(d, n);
    h.final(out);
}

Transcript::Transcript(const std::string &label) {
    h.update("FSv1-TRANSCRIPT-INIT");
    h.update(label);
}
void Transcript::absorb(const void *b, size_t n) {
    h.update("FSv1-ABSORB-BYTES");
    h.update(b, n);
}
void Transcript::absorb_field(uint64_t x) {
    uint8_t le[8];
    for (int i = 0; i < 8; i++) le[i] = (uint8_t)(x >> (8 * i));
    absorb(le, 8);
}
void Transcript::challenge_bytes(const std::string &label, uint8_t out[32]) const {
    Sha3 c = h;
    c.update("FSv1-CHALLENGE");
    c.update(label);
    c.final(out);
}
uint64_t Transcript::challenge_field(const std::string &label) const {
    uint8_t d[32];
    challenge_bytes(label, d);
    // chunks of 7 bytes taken from the end (the first 4 bytes form the last chunk);
    // acc = acc * 2^(8 len) + LE(chunk), in the field
    uint64_t acc = 0;
    for (int end = 32; end > 0;) {
        int start = end >= 7 ? end - 7 : 0, len = end - start;
        uint64_t v = 0;
        for (int i = 0; i < len; i++) v |= (uint64_t)d[start + i] << (8 * i);
        acc = add(mul(acc, 1ULL << (8 * len)), v);
        end = start;
    }
    return acc;
}
void Transcript::challenge_ext(const std::string &tag, uint64_t out[6]) const {
    for (int i = 0; i < 6; i++) out[i] = challenge_field(tag + "/c" + std::to_string(i));
}

uint64_t hash_fields_tagged(const std::string &tag, const uint64_t *f, size_t n) {
    Transcript t("FRI/FS");
    t.absorb(tag);
    for (size_t i = 0; i < n; i++) t.absorb_field(f[i]);
    return t.challenge_field("out");
}

uint64_t stdrng_first_u64(uint64_t seed) {
    // ChaCha12 block 0: constants, key = LE64(seed) || 0^24, counter 0, nonce 0
    uint32_t in[16] = {0x61707865, 0x3320646e, 0x79622d32, 0x6b206574, (uint32_t)seed, (uint32_t)(seed >> 32),
                       0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t x[16];
    memcpy(x, in, sizeof x);
    auto qr = [&](int a, int b, int c, int d) {
        auto rl = [](uint32_t v, int n) { return (v << n) | (v >> (32 - n)); };
        x[a] += x[b]; x[d] = rl(x[d] ^ x[a], 16);
        x[c] += x[d]; x[b] = rl(x[b] ^ x[c], 12);
        x[a] += x[b]; x[d] = rl(x[d] ^ x[a], 8);
        x[c] += x[d]; x[b] = rl(x[b] ^ x[c], 7);
    };
    for (int r = 0; r < 6; r++) {
        qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
        qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
    }
    uint32_t w0 = x[0] + in[0], w1 = x[1] + in[1];
    return (uint64_t)w0 | ((uint64_t)w1 << 32);
}

// ---------------------------------------------------------------- kernels
namespace {

constexpr int KT = 256;
// node levels of at most this many nodes (all trees of a batch) use two lanes per node
constexpr uint64_t NODES2_MAX = 1u << 17;

// absorb `total` message lanes, SHA3-256 padding, digest out: lanes t < 4
// (the domain-separation header) come from head(t), the rest from lane(t),
// in order of t. The first block's lane positions are compile-time, so
// neither functor branches on t there.
template <class HeadFn, class LaneFn>
__device__ __forceinline__ void sha3_lanes(uint32_t total, HeadFn head, LaneFn lane, uint64_t *dig) {
    uint64_t st[25];
#pragma unroll
    for (int j = 0; j < 25; j++) st[j] = 0;
    uint32_t t0 = 0;
    while (total - t0 >= 17) {
        if (t0 == 0) {
#pragma unroll
            for (int j = 0; j < 17; j++) st[j] ^= j < 4 ? head(j) : lane(j);
        } else {
#pragma unroll
            for (int j = 0; j < 17; j++) st[j] ^= lane(t0 + j);
        }
        keccak_f_dev(st);
        t0 += 17;
    }
    const uint32_t rem = total - t0;
#pragma unroll
    for (int j = 0; j < 17; j++) {
        if ((uint32_t)j < rem) st[j] ^= (t0 == 0 && j < 4) ? head(j) : lane(t0 + j);
        if ((uint32_t)j == rem) st[j] ^= 0x06ULL;
    }
    st[16] ^= 0x8000000000000000ULL;
    keccak_f_dev(st);
    dig[0] = st[0];
    dig[1] = st[1];
    dig[2] = st[2];
    dig[3] = st[3];
}

// grid.y = tree of the batch; leaf i of tree k is written at out[k * n + i].
// At most 96 VGPRs (5 waves per SIMD): about 3 % faster on 2^23 leaves of 128
// values than the 128 the compiler picks unbounded, despite a few spills.
__global__ __launch_bounds__(KT, 5) void merkle_leaves_k(LeafSrc src, uint64_t n, uint32_t arity, uint64_t label,
                                                      uint64_t *out) {
    const uint64_t i = (uint64_t)blockIdx.x * KT + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = blockIdx.y;
    const uint64_t *base = src.vals + k * src.proof_stride + i * src.leaf_stride;
    const uint64_t cs = src.cs;
    // sum_c (c+1) v_c = w * T_w - sum_{k<w} T_k with T_k = sum_{c<k} v_c,
    // both sums kept as 128-bit integers (w < 2^32 values below 2^64) and
    // reduced once
    Acc T{0, 0}, U{0, 0};
    const bool combo = src.combo_out != nullptr;
    auto head = [&](uint32_t t) -> uint64_t {
        return t == 0 ? (uint64_t)arity : t == 1 ? 0xFFFFFFFFULL : t == 2 ? i : label;
    };
    auto lane = [&](uint32_t t) -> uint64_t {
        const uint64_t v = (t - 4 < src.nz) ? base[(uint64_t)(t - 4) * cs] : 0;
        if (combo) {
            U.lo += T.lo;
            U.hi += T.hi + (U.lo < T.lo);
            acc_add(T, v);
        }
        return v;
    };
    uint64_t d[4];
    sha3_lanes(4 + src.w, head, lane, d);
    if (combo) src.combo_out[k * n + i] = sub(mul(reduce128(T.lo, T.hi), src.w), reduce128(U.lo, U.hi));
    uint64_t *o = out + 4 * (k * n + i);
    o[0] = d[0];
    o[1] = d[1];
    o[2] = d[2];
    o[3] = d[3];
}

// grid.y = tree of the batch; child/out point at the level arrays of all trees
__global__ __launch_bounds__(KT) void merkle_nodes_k(const uint64_t *child, uint64_t n_child, uint64_t n_parent,
                                                     uint32_t arity, uint32_t level, uint64_t label, uint64_t *out) {
    const uint64_t i = (uint64_t)blockIdx.x * KT + threadIdx.x;
    if (i >= n_parent) return;
    const uint64_t k = blockIdx.y;
    child += 4 * k * n_child;
    out += 4 * k * n_parent;
    auto head = [&](uint32_t t) -> uint64_t {
        return t == 0 ? (uint64_t)arity : t == 1 ? (uint64_t)level : t == 2 ? i : label;
    };
    auto lane = [&](uint32_t t) -> uint64_t {
        uint32_t j = (t - 4) >> 2, q = (t - 4) & 3;
        uint64_t ci = i * arity + j;
        if (ci >= n_child) ci = n_child - 1;
        return child[4 * ci + q];
    };
    uint64_t d[4];
    sha3_lanes(4 + 4 * arity, head, lane, d);
    uint64_t *o = out + 4 * i;
    o[0] = d[0];
    o[1] = d[1];
    o[2] = d[2];
    o[3] = d[3];
}

// merkle_nodes_k on two lanes per node (keccak_f_half): lanes 2i and 2i + 1
// of the grid hash node i. A pair past n_parent still runs the permutation
// (its partner reads come from it) but loads and stores nothing.
// One lane's half of a node digest: the node's four header lanes and its
// arity children's digests (ch32: the child level of this tree as u32
// halves, global or LDS) absorbed by the two-lane permutation.
__device__ __forceinline__ void node2_hash(const uint32_t *ch32, uint64_t n_child, uint64_t i, bool live, uint32_t h,
                                           uint32_t arity, uint32_t level, uint64_t label, uint32_t dig[4]) {
    // 64-bit message lane t, this lane's half
    auto word = [&](uint32_t t) -> uint32_t {
        if (t < 4) {
            const uint64_t v = t == 0 ? (uint64_t)arity : t == 1 ? (uint64_t)level : t == 2 ? i : label;
            return (uint32_t)(v >> (32 * h));
        }
        // a dead pair reads node 0's children (in bounds) and hashes garbage
        const uint32_t j = (t - 4) >> 2, q = (t - 4) & 3;
        uint64_t ci = (live ? i : 0) * arity + j;
        if (ci >= n_child) ci = n_child - 1;
        return ch32[2 * (4 * ci + q) + h];
    };
    uint32_t st[25];
#pragma unroll
    for (int j = 0; j < 25; j++) st[j] = 0;
    const uint32_t total = 4 + 4 * arity;
    uint32_t t0 = 0;
    while (total - t0 >= 17) {
#pragma unroll
        for (int j = 0; j < 17; j++) st[j] ^= word(t0 + j);
        keccak_f_half(st, h);
        t0 += 17;
    }
    const uint32_t rem = total - t0;
#pragma unroll
    for (int j = 0; j < 17; j++) {
        if ((uint32_t)j < rem) st[j] ^= word(t0 + j);
        if ((uint32_t)j == rem) st[j] ^= h ? 0u : 0x06u;
    }
    st[16] ^= h ? 0x80000000u : 0u;
    keccak_f_half(st, h);
#pragma unroll
    for (int j = 0; j < 4; j++) dig[j] = st[j];
}

__global__ __launch_bounds__(KT) void merkle_nodes2_k(const uint64_t *child, uint64_t n_child, uint64_t n_parent,
                                                      uint32_t arity, uint32_t level, uint64_t label, uint64_t *out) {
    const uint64_t i = ((uint64_t)blockIdx.x * KT + threadIdx.x) >> 1;
    const uint32_t h = threadIdx.x & 1;
    const bool live = i < n_parent;
    const uint64_t k = blockIdx.y;
    uint32_t d[4];
    node2_hash((const uint32_t *)(child + 4 * k * n_child), n_child, i, live, h, arity, level, label, d);
    if (!live) return;
    uint32_t *o = (uint32_t *)(out + 4 * (k * n_parent + i));
    o[h] = d[0];
    o[2 + h] = d[1];
    o[4 + h] = d[2];
    o[6 + h] = d[3];
}

// The top levels of a tree, each of at most TOP_NODES nodes, in one launch:
// one workgroup per tree of the batch, two lanes per node, the levels one
// after another behind a barrier. Each level is written to the tree buffer
// and kept in LDS, where the next level reads it. Saves the launch gap
// between dependent levels (several microseconds each) on small trees.
constexpr uint32_t TOP_NODES = 128, TOP_MAX_LEVELS = 16;
struct TopPlan {
    uint32_t nlev;                      // levels computed: level0 .. level0 + nlev - 1
    uint32_t level0;
    uint64_t n[TOP_MAX_LEVELS + 1];     // n[s]: nodes of level level0 - 1 + s (one tree)
    uint64_t off[TOP_MAX_LEVELS + 1];   // its digest offset in the tree buffer (all trees)
};
__global__ __launch_bounds__(2 * TOP_NODES) void merkle_top2_k(uint64_t *buf, TopPlan tp, uint32_t arity,
                                                              uint64_t label) {
    __shared__ uint32_t lev[2][8 * TOP_NODES];
    const uint64_t k = blockIdx.y;
    const uint32_t h = threadIdx.x & 1;
    const uint64_t i = threadIdx.x >> 1;
    for (uint32_t s = 0; s < tp.nlev; s++) {
        const uint64_t n_child = tp.n[s], n_parent = tp.n[s + 1];
        const bool live = i < n_parent;
        const uint32_t *ch32 = s == 0 ? (const uint32_t *)(buf + 4 * (tp.off[0] + k * n_child)) : lev[(s - 1) & 1];
        uint32_t d[4];
        node2_hash(ch32, n_child, i, live, h, arity, tp.level0 + s, label, d);
        if (live) {
            uint32_t *o = (uint32_t *)(buf + 4 * (tp.off[s + 1] + k * n_parent + i));
            uint32_t *l = lev[s & 1] + 8 * i;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                o[2 * j + h] = d[j];
                l[2 * j + h] = d[j];
            }
        }
        __syncthreads();
    }
}

__global__ void gather_k(const uint64_t *src, const uint64_t *idx, uint64_t n, uint32_t width, uint64_t *out) {
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n * width) return;
    uint64_t q = r / width, l = r - q * width;
    out[r] = src[idx[q] * width + l];
}

// ---- STIR query indices (crates/fri/src/fri.rs:1813-1893) ----
// Bytes [start, start + len) of a 32-byte digest held as 4 LE lanes, LE value.
__device__ __forceinline__ uint64_t digest_chunk(const uint64_t *d, int start, int len) {
    const int i = start >> 3, sh = (start & 7) * 8;
    uint64_t v = d[i] >> sh;
    if (sh && 64 - sh < 8 * len) v |= d[i + 1] << (64 - sh);
    return len == 8 ? v : v & ((1ULL << (8 * len)) - 1);
}

// One thread per (proof k, query q): seed = hash_fields_tagged("FRI/index",
// [query_seed_k, 0, q]) as a two-block SHA3 over a fixed message with the two
// variable fields XORed in at byte offsets seed_off / q_off, the digest
// folded to a field element by 7-byte chunks from the end, then the first
// u64 of StdRng(seed) (ChaCha12 block 0) gives raw = (r & (npow2-1)) % n0.
constexpr int QI_BLOCKS = 2;
__global__ __launch_bounds__(KT) void query_indices_k(const uint64_t *tmpl, const uint64_t *seeds, uint32_t K,
                                                      uint32_t s0, uint32_t seed_off, uint32_t q_off, uint64_t npow2,
                                                      uint64_t n0, uint64_t *raw) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K * s0) return;
    const uint32_t k = i / s0, q = i - k * s0;
    uint64_t m[17 * QI_BLOCKS];
#pragma unroll
    for (int j = 0; j < 17 * QI_BLOCKS; j++) m[j] = tmpl[j];
    auto put = [&](uint32_t off, uint64_t x) {
        const uint32_t l = off >> 3, sh = (off & 7) * 8;
#pragma unroll
        for (int j = 0; j < 17 * QI_BLOCKS; j++) {
            if ((uint32_t)j == l) m[j] ^= x << sh;
            if (sh && (uint32_t)j == l + 1) m[j] ^= x >> (64 - sh);
        }
    };
    put(seed_off, seeds[k]);
    put(q_off, (uint64_t)q);
    uint64_t st[25];
#pragma unroll
    for (int j = 0; j < 25; j++) st[j] = 0;
#pragma unroll
    for (int b = 0; b < QI_BLOCKS; b++) {
#pragma unroll
        for (int j = 0; j < 17; j++) st[j] ^= m[17 * b + j];
        keccak_f_dev(st);
    }
    uint64_t acc = 0;
    for (int end = 32; end > 0;) {
        const int start = end >= 7 ? end - 7 : 0, len = end - start;
        acc = add(mul(acc, 1ULL << (8 * len)), digest_chunk(st, start, len));
        end = start;
    }
    // ChaCha12 block 0 with key LE64(acc) || 0^24, counter 0, nonce 0
    const uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, (uint32_t)acc, (uint32_t)(acc >> 32),
                             0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t x[16];
#pragma unroll
    for (int j = 0; j < 16; j++) x[j] = in[j];
    auto qr = [&](int a, int b, int c, int d) {
        x[a] += x[b]; x[d] = __builtin_rotateleft32(x[d] ^ x[a], 16);
        x[c] += x[d]; x[b] = __builtin_rotateleft32(x[b] ^ x[c], 12);
        x[a] += x[b]; x[d] = __builtin_rotateleft32(x[d] ^ x[a], 8);
        x[c] += x[d]; x[b] = __builtin_rotateleft32(x[b] ^ x[c], 7);
    };
#pragma unroll
    for (int r = 0; r < 6; r++) {
        qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
        qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
    }
    const uint64_t r64 = (uint64_t)(x[0] + in[0]) | ((uint64_t)(x[1] + in[1]) << 32);
    raw[i] = (r64 & (npow2 - 1)) % n0;
}

// ---- DNS record commitments (crates/swarm-dns/src/dns.rs:43-118) ----
// h1 = SHA3("DNS-LEAF-V1" || salt || canonical(record)) from host-padded
// message lanes (nblk[r] blocks of 17 lanes at msg + 17 * blk_off[r]);
// h2 = SHA3("DNS-LEAF-DOUBLE-V1" || salt || h1): 66 bytes, a fixed 34-byte
// prefix (4 full lanes + 2 bytes) then h1 at byte offset 34.
__global__ __launch_bounds__(KT) void dns_leaf_k(const uint64_t *msg, const uint64_t *blk_off, const uint32_t *nblk,
                                                 uint64_t n, const uint64_t *prefix2, uint64_t *out) {
    const uint64_t r = (uint64_t)blockIdx.x * KT + threadIdx.x;
    if (r >= n) return;
    uint64_t st[25];
#pragma unroll
    for (int j = 0; j < 25; j++) st[j] = 0;
    const uint64_t *m = msg + 17 * blk_off[r];
    for (uint32_t b = 0; b < nblk[r]; b++) {
#pragma unroll
        for (int j = 0; j < 17; j++) st[j] ^= m[17 * b + j];
        keccak_f_dev(st);
    }
    const uint64_t h0 = st[0], h1 = st[1], h2 = st[2], h3 = st[3];
#pragma unroll
    for (int j = 0; j < 25; j++) st[j] = 0;
    st[0] = prefix2[0];
    st[1] = prefix2[1];
    st[2] = prefix2[2];
    st[3] = prefix2[3];
    st[4] = prefix2[4] | (h0 << 16);  // prefix2[4]: the last 2 prefix bytes
    st[5] = (h0 >> 48) | (h1 << 16);
    st[6] = (h1 >> 48) | (h2 << 16);
    st[7] = (h2 >> 48) | (h3 << 16);
    st[8] = (h3 >> 48) | (0x06ULL << 16);
    st[16] = 0x8000000000000000ULL;
    keccak_f_dev(st);
    uint64_t *o = out + 4 * r;
    o[0] = st[0];
    o[1] = st[1];
    o[2] = st[2];
    o[3] = st[3];
}

// parent i = SHA3("DNS-NODE-V1" || left || right), right = left when the level is odd.
// 75-byte message: the 64 child bytes start at byte 11. grid.y = tree of a batch.
__global__ __launch_bounds__(KT) void dns_node_k(const uint64_t *child, uint64_t n_child, uint64_t n_parent,
                                                 uint64_t child_stride, uint64_t parent_stride, uint64_t *out) {
    const uint64_t i = (uint64_t)blockIdx.x * KT + threadIdx.x;
    if (i >= n_parent) return;
    child += blockIdx.y * child_stride;
    out += blockIdx.y * parent_stride;
    const uint64_t li = 2 * i, ri = (2 * i + 1 < n_child) ? 2 * i + 1 : 2 * i;
    uint64_t w[8];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        w[q] = child[4 * li + q];
        w[4 + q] = child[4 * ri + q];
    }
    // "DNS-NODE-V1" = 44 4E 53 2D 4E 4F 44 45 | 2D 56 31
    const uint64_t tag0 = 0x45444F4E2D534E44ULL, tag1 = 0x31562DULL;
    uint64_t st[25];
#pragma unroll
    for (int j = 0; j < 25; j++) st[j] = 0;
    st[0] = tag0;
    st[1] = tag1 | (w[0] << 24);
#pragma unroll
    for (int j = 2; j <= 8; j++) st[j] = (w[j - 2] >> 40) | (w[j - 1] << 24);
    st[9] = (w[7] >> 40) | (0x06ULL << 24);
    st[16] = 0x8000000000000000ULL;
    keccak_f_dev(st);
    uint64_t *o = out + 4 * i;
    o[0] = st[0];
    o[1] = st[1];
    o[2] = st[2];
    o[3] = st[3];
}

}  // namespace

// Leaf hashes (h2) and Merkle roots of K DNS shards on the device. The host
// writes each record's first SHA3 message, padded, straight into the lanes.
void dns_commit_batch(Ctx *c, const RecordSource &src, const uint8_t salt[16],
                      std::vector<std::vector<uint8_t>> &leaf_hashes, std::vector<std::array<uint8_t, 32>> &roots) {
    const size_t K = src.counts.size();
    std::vector<size_t> first(K + 1, 0);
    for (size_t k = 0; k < K; k++) {
        // dns.rs merkle_root unwraps the top level: an empty shard has no root
        SG_REQUIRE(src.counts[k] >= 1, "dns commit: a shard needs at least one record");
        first[k + 1] = first[k] + src.counts[k];
    }
    const size_t nrec = first[K];
    std::vector<uint64_t> blk_off(nrec);
    std::vector<uint32_t> nblk(nrec);
    parallel_for(K, [&](size_t k) {
        for (size_t i = 0; i < src.counts[k]; i++) {
            const size_t len = 11 + 16 + src.size(k, i);
            nblk[first[k] + i] = (uint32_t)(len / 136 + 1);
        }
    });
    uint64_t total_blocks = 0;
    for (size_t r = 0; r < nrec; r++) {
        blk_off[r] = total_blocks;
        total_blocks += nblk[r];
    }
    // the padded sponge blocks of every record, written (zeros included) on
    // the host's threads straight into pinned memory for one upload
    const size_t msg_words = 17 * total_blocks;
    uint64_t *msg = (uint64_t *)c->upload_staging(8 * msg_words);
    {
        HostScope hm(c, "dns_messages");
        parallel_for(K, [&](size_t k) {
            for (size_t i = 0; i < src.counts[k]; i++) {
                const size_t r = first[k] + i;
                uint8_t *p = (uint8_t *)&msg[17 * blk_off[r]];
                memset(p, 0, 136 * (size_t)nblk[r]);
                memcpy(p, "DNS-LEAF-V1", 11);
                memcpy(p + 11, salt, 16);
                const size_t cl = src.write(k, i, p + 27);
                p[27 + cl] ^= 0x06;
                p[136 * nblk[r] - 1] ^= 0x80;
            }
        });
    }
    uint64_t prefix2[5] = {0, 0, 0, 0, 0};
    {
        uint8_t pb[40] = {0};
        memcpy(pb, "DNS-LEAF-DOUBLE-V1", 18);
        memcpy(pb + 18, salt, 16);
        memcpy(prefix2, pb, 40);
        prefix2[4] &= 0xFFFF;  // bytes 32..33
    }
    // level sizes of one shard's tree; all trees of a batch share nothing but the launch
    Scoped d_msg(c, 8 * msg_words), d_off(c, 8 * nrec), d_nb(c, 4 * nrec), d_pre(c, 40), d_leaf(c, 32 * nrec);
    SG_HIP(hipMemcpyAsync(d_msg.p, msg, 8 * msg_words, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(d_off.p, blk_off.data(), 8 * nrec, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(d_nb.p, nblk.data(), 4 * nrec, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(d_pre.p, prefix2, 40, hipMemcpyHostToDevice, c->stream));
    {
        ProfScope ps(c, "dns_leaf", 8.0 * msg_words + 32.0 * nrec);
        hipLaunchKernelGGL(dns_leaf_k, dim3((uint32_t)((nrec + KT - 1) / KT)), dim3(KT), 0, c->stream,
                           d_msg.as<uint64_t>(), d_off.as<uint64_t>(), d_nb.as<uint32_t>(), (uint64_t)nrec,
                           d_pre.as<uint64_t>(), d_leaf.as<uint64_t>());
        SG_HIP(hipGetLastError());
    }
    leaf_hashes.assign(K, {});
    roots.assign(K, {});
    // trees: shards of equal size go up level by level in one launch per
    // level; every shard's root is gathered into d_roots, and the leaf hashes
    // and the roots come back behind one wait at the end
    Scoped d_roots(c, 32 * K), d_ridx(c, 8 * K);
    std::vector<uint64_t> ridx;  // per group: the root rows j * n of its level-0 layout
    std::vector<size_t> rorder;  // shard of each gathered root
    std::map<size_t, std::vector<size_t>> by_size;
    for (size_t k = 0; k < K; k++) by_size[src.counts[k]].push_back(k);
    for (auto &g : by_size)
        for (size_t j = 0; j < g.second.size(); j++) {
            ridx.push_back(j * g.first);
            rorder.push_back(g.second[j]);
        }
    SG_HIP(hipMemcpyAsync(d_ridx.p, ridx.data(), 8 * K, hipMemcpyHostToDevice, c->stream));
    size_t r0 = 0;
    for (auto &g : by_size) {
        const size_t n = g.first, G = g.second.size();
        std::vector<uint64_t> idx;
        for (size_t k : g.second)
            for (size_t i = 0; i < n; i++) idx.push_back(first[k] + i);
        // released at the end of the iteration: later work on the stream runs after this group's
        Scoped d_idx(c, 8 * idx.size()), a(c, 32 * idx.size()), b(c, 32 * idx.size());
        SG_HIP(hipMemcpyAsync(d_idx.p, idx.data(), 8 * idx.size(), hipMemcpyHostToDevice, c->stream));
        gather_u64_w(c, d_leaf.as<uint64_t>(), d_idx.as<uint64_t>(), idx.size(), 4, a.as<uint64_t>());
        uint64_t *cur = a.as<uint64_t>(), *nxt = b.as<uint64_t>();
        size_t len = n;
        while (len > 1) {
            const size_t np = (len + 1) / 2;
            ProfScope ps(c, "dns_node", 32.0 * (len + np) * G);
            hipLaunchKernelGGL(dns_node_k, dim3((uint32_t)((np + KT - 1) / KT), (uint32_t)G), dim3(KT), 0, c->stream,
                               cur, (uint64_t)len, (uint64_t)np, (uint64_t)4 * n, (uint64_t)4 * n, nxt);
            SG_HIP(hipGetLastError());
            std::swap(cur, nxt);
            len = np;
        }
        gather_u64_w(c, cur, d_ridx.as<uint64_t>() + r0, G, 4, d_roots.as<uint64_t>() + 4 * r0);
        r0 += G;
    }
    std::vector<uint8_t> all(32 * nrec), rt(32 * K);
    c->d2h2(all.data(), d_leaf.p, 32 * nrec, rt.data(), d_roots.p, 32 * K);
    for (size_t j = 0; j < K; j++) memcpy(roots[rorder[j]].data(), &rt[32 * j], 32);
    for (size_t k = 0; k < K; k++)
        leaf_hashes[k].assign(all.begin() + 32 * first[k], all.begin() + 32 * first[k + 1]);
}

uint32_t merkle_depth(uint64_t leaves, uint32_t arity) {
    uint32_t d = 1;
    uint64_t cur = leaves;
    while (cur > arity) {
        cur = (cur + arity - 1) / arity;
        d++;
    }
    return d;
}

uint32_t pick_arity(uint64_t n, uint64_t m) {
    static const uint64_t as[6] = {128, 64, 32, 16, 8, 4};
    for (uint64_t a : as)
        if (m >= a && n % a == 0) return (uint32_t)a;
    return 2;  // n % 2 == 0 gives 2; the reference's 1 is raised to 2 by .max(2)
}

void tree_build(Ctx *c, DevTree &t, const LeafSrc &src, uint64_t n_leaves, uint32_t arity, uint32_t depth,
                uint64_t label, uint32_t batch, bool read_roots) {
    SG_REQUIRE(n_leaves >= 1 && arity >= 2, "merkle: empty tree or arity < 2");
    SG_REQUIRE(batch >= 1 && batch <= 65535, "merkle: batch out of range");
    SG_REQUIRE(n_leaves <= (1ULL << 32) * KT, "merkle: too many leaves");
    t.arity = arity;
    t.label = label;
    t.batch = batch;
    t.len.clear();
    t.real.clear();
    t.off.clear();
    // level sizes (one tree); level arrays of the whole batch are contiguous
    uint64_t cur = n_leaves, total = 0;
    for (;;) {
        const uint64_t padded = cur > 1 ? (cur + arity - 1) / arity * arity : cur;
        t.real.push_back(cur);
        t.len.push_back(padded);
        t.off.push_back(total);
        total += cur * batch;
        if (cur <= 1) break;
        SG_REQUIRE(t.real.size() <= depth, "merkle: layer arities exhausted");
        cur = padded / arity;
    }
    t.buf = c->alloc_n<uint64_t>(4 * total);
    {
        ProfScope ps(c, "merkle_leaves", (8.0 * src.w * n_leaves + 32.0 * n_leaves) * batch);
        hipLaunchKernelGGL(merkle_leaves_k, dim3((uint32_t)((n_leaves + KT - 1) / KT), batch), dim3(KT), 0, c->stream,
                           src, n_leaves, arity, label, t.buf);
        SG_HIP(hipGetLastError());
    }
    // the first level from which every level has at most TOP_NODES nodes
    size_t top = t.real.size();
    while (top > 1 && t.real[top - 1] <= TOP_NODES) top--;
    for (size_t l = 1; l < top; l++) {
        const uint64_t np = t.real[l];
        ProfScope ps(c, "merkle_nodes", (32.0 * t.real[l - 1] + 32.0 * np) * batch);
        // levels too small to fill the chip one lane per node take two lanes
        // per node: a shorter permutation chain for the same total work
        if (np * batch <= NODES2_MAX)
            hipLaunchKernelGGL(merkle_nodes2_k, dim3((uint32_t)((2 * np + KT - 1) / KT), batch), dim3(KT), 0,
                               c->stream, t.buf + 4 * t.off[l - 1], t.real[l - 1], np, arity, (uint32_t)l, label,
                               t.buf + 4 * t.off[l]);
        else
            hipLaunchKernelGGL(merkle_nodes_k, dim3((uint32_t)((np + KT - 1) / KT), batch), dim3(KT), 0, c->stream,
                               t.buf + 4 * t.off[l - 1], t.real[l - 1], np, arity, (uint32_t)l, label,
                               t.buf + 4 * t.off[l]);
        SG_HIP(hipGetLastError());
    }
    if (top < t.real.size()) {
        TopPlan tp{};
        tp.nlev = (uint32_t)(t.real.size() - top);
        SG_REQUIRE(tp.nlev <= TOP_MAX_LEVELS, "merkle: too many top levels");
        tp.level0 = (uint32_t)top;
        for (uint32_t s2 = 0; s2 <= tp.nlev; s2++) {
            tp.n[s2] = t.real[top - 1 + s2];
            tp.off[s2] = t.off[top - 1 + s2];
        }
        double bytes = 0;
        for (size_t l = top; l < t.real.size(); l++) bytes += 32.0 * (t.real[l - 1] + t.real[l]) * batch;
        ProfScope ps(c, "merkle_nodes", bytes);
        hipLaunchKernelGGL(merkle_top2_k, dim3(1, batch), dim3(2 * TOP_NODES), 0, c->stream, t.buf, tp, arity, label);
        SG_HIP(hipGetLastError());
    }
    if (read_roots) tree_read_roots(c, t);
}

void tree_read_roots(Ctx *c, DevTree &t) { tree_read_roots_with(c, t, nullptr, nullptr, 0); }

void tree_read_roots_with(Ctx *c, DevTree &t, void *dst, const void *src, size_t n) {
    // the top level holds one root per tree
    std::vector<uint64_t> r(4 * (size_t)t.batch);
    c->d2h2(dst, src, n, r.data(), t.buf + 4 * t.off.back(), 32 * (size_t)t.batch);
    t.roots.resize(t.batch);
    for (uint32_t k = 0; k < t.batch; k++) memcpy(t.roots[k].data(), &r[4 * k], 32);
    memcpy(t.root, t.roots[0].data(), 32);
}

void tree_free(Ctx *c, DevTree &t) {
    if (t.buf) c->release(t.buf);
    t.buf = nullptr;
}

void gather_u64(Ctx *c, const uint64_t *src, const uint64_t *d_idx, uint64_t n, uint64_t *out) {
    gather_u64_w(c, src, d_idx, n, 1, out);
}

void gather_u64_w(Ctx *c, const uint64_t *src, const uint64_t *d_idx, uint64_t n, uint32_t width, uint64_t *out) {
    if (!n) return;
    hipLaunchKernelGGL(gather_k, dim3((uint32_t)((n * width + 255) / 256)), dim3(256), 0, c->stream, src, d_idx, n,
                       width, out);
    SG_HIP(hipGetLastError());
}

void gather_set(Ctx *c, std::vector<GatherJob> &jobs) {
    const size_t J = jobs.size();
    std::vector<size_t> ioff(J + 1, 0), ooff(J + 1, 0);
    for (size_t j = 0; j < J; j++) {
        ioff[j + 1] = ioff[j] + jobs[j].idx.size();
        ooff[j + 1] = ooff[j] + jobs[j].idx.size() * jobs[j].width;
    }
    const size_t ni = ioff[J], no = ooff[J];
    if (!ni) {
        for (GatherJob &j : jobs) j.out.reset(0);
        return;
    }
    // pieces of at most 2^16 u64 (512 KiB), copied on the host's threads
    struct Piece {
        size_t job, off, n;
    };
    auto pieces = [&](const std::vector<size_t> &off) {
        std::vector<Piece> v;
        for (size_t j = 0; j < J; j++)
            for (size_t a = off[j]; a < off[j + 1]; a += (size_t)1 << 16)
                v.push_back({j, a - off[j], std::min(off[j + 1] - a, (size_t)1 << 16)});
        return v;
    };
    // every index list straight into pinned memory, one upload
    uint64_t *up = (uint64_t *)c->upload_staging(8 * ni);
    const std::vector<Piece> pin_in = pieces(ioff);
    parallel_for(pin_in.size(), [&](size_t i) {
        const Piece &q = pin_in[i];
        memcpy(up + ioff[q.job] + q.off, jobs[q.job].idx.data() + q.off, 8 * q.n);
    });
    Scoped d_idx(c, 8 * ni), d_out(c, 8 * no);
    SG_HIP(hipMemcpyAsync(d_idx.p, up, 8 * ni, hipMemcpyHostToDevice, c->stream));
    for (size_t j = 0; j < J; j++)
        gather_u64_w(c, jobs[j].src, d_idx.as<uint64_t>() + ioff[j], jobs[j].idx.size(), jobs[j].width,
                     d_out.as<uint64_t>() + ooff[j]);
    // one download into pinned memory, copied out to the jobs in parallel
    const uint64_t *h = (const uint64_t *)c->d2h_pinned(d_out.p, 8 * no);
    for (GatherJob &j : jobs) j.out.reset(j.idx.size() * j.width);
    const std::vector<Piece> pin_out = pieces(ooff);
    parallel_for(pin_out.size(), [&](size_t i) {
        const Piece &q = pin_out[i];
        memcpy(jobs[q.job].out.p.get() + q.off, h + ooff[q.job] + q.off, 8 * q.n);
    });
}

// digests an opening needs, per query: the leaf, then every level's full
// sibling group (a digest is a row of 4 u64 of t.buf)
GatherJob tree_open_job(const DevTree &t, const std::vector<std::vector<uint64_t>> &idx) {
    const uint32_t a = t.arity;
    const size_t nlev = t.real.size(), K = idx.size();
    SG_REQUIRE(K <= t.batch, "merkle open: more index lists than trees");
    const size_t per = 1 + (nlev - 1) * a;
    std::vector<size_t> qoff(K + 1, 0);
    for (size_t k = 0; k < K; k++) {
        for (uint64_t q : idx[k]) SG_REQUIRE(q < t.real[0], "merkle open: leaf index out of range");
        qoff[k + 1] = qoff[k] + idx[k].size();
    }
    GatherJob job;
    job.src = t.buf;
    job.width = 4;
    job.idx.resize(qoff[K] * per);
    parallel_for(K, [&](size_t k) {
        uint64_t *w = &job.idx[qoff[k] * per];
        for (uint64_t q : idx[k]) {
            *w++ = t.off[0] + k * t.real[0] + q;
            uint64_t x = q;
            for (size_t l = 0; l + 1 < nlev; l++) {
                const uint64_t gs = x / a * a;
                for (uint32_t i = 0; i < a; i++) {
                    uint64_t pos = gs + i;
                    if (pos >= t.real[l]) pos = t.real[l] - 1;
                    *w++ = t.off[l] + k * t.real[l] + pos;
                }
                x /= a;
            }
        }
    });
    return job;
}

void tree_open_finish(const DevTree &t, const std::vector<std::vector<uint64_t>> &idx, const GatherJob &job,
                      std::vector<std::vector<std::vector<uint8_t>>> &out) {
    const uint32_t a = t.arity;
    const size_t nlev = t.real.size(), K = idx.size();
    const size_t per = 1 + (nlev - 1) * a;
    std::vector<size_t> qoff(K + 1, 0);
    for (size_t k = 0; k < K; k++) qoff[k + 1] = qoff[k] + idx[k].size();
    out.assign(K, {});
    // MerkleOpening: leaf[32] || LE64(levels) || per level (LE64(a-1) || a-1 siblings) || LE64(index)
    const size_t osz = 32 + 8 + (nlev - 1) * (8 + 32 * (size_t)(a - 1)) + 8;
    auto put64 = [](uint8_t *p, uint64_t x) {
        for (int i = 0; i < 8; i++) p[i] = (uint8_t)(x >> (8 * i));
    };
    parallel_for(K, [&](size_t k) {
        out[k].resize(idx[k].size());
        const uint64_t *hw = &job.out[4 * qoff[k] * per];
        for (size_t j = 0; j < idx[k].size(); j++) {
            const uint64_t q = idx[k][j];
            std::vector<uint8_t> &o = out[k][j];
            o.resize(osz);
            uint8_t *p = o.data();
            memcpy(p, hw, 32);
            hw += 4;
            put64(p + 32, nlev - 1);
            p += 40;
            uint64_t x = q;
            for (size_t l = 0; l + 1 < nlev; l++) {
                put64(p, a - 1);
                p += 8;
                const uint64_t gs = x / a * a;
                for (uint32_t i = 0; i < a; i++, hw += 4) {
                    if (gs + i == x) continue;
                    memcpy(p, hw, 32);
                    p += 32;
                }
                x /= a;
            }
            put64(p, q);
        }
    });
}

void tree_open_multi(Ctx *c, const DevTree &t, const std::vector<std::vector<uint64_t>> &idx,
                     std::vector<std::vector<std::vector<uint8_t>>> &out) {
    std::vector<GatherJob> jobs(1);
    jobs[0] = tree_open_job(t, idx);
    gather_set(c, jobs);
    tree_open_finish(t, idx, jobs[0], out);
}

void fri_query_indices(Ctx *c, const std::vector<uint64_t> &seeds, uint64_t s0, uint64_t n0,
                       std::vector<std::vector<uint64_t>> &raw) {
    const size_t K = seeds.size();
    raw.assign(K, std::vector<uint64_t>(s0));
    if (!K || !s0) return;
    SG_REQUIRE(K * s0 < (1ULL << 31), "query indices: too many queries");
    uint64_t npow2 = 1;
    while (npow2 < n0) npow2 <<= 1;
    if (K * s0 <= 1024) {
        // few queries: the host's SHA3 and ChaCha12 beat a launch and a wait
        parallel_for(K, [&](size_t k) {
            for (uint64_t q = 0; q < s0; q++) {
                const uint64_t f[3] = {seeds[k], 0, q};
                raw[k][q] = (stdrng_first_u64(hash_fields_tagged("FRI/index", f, 3)) & (npow2 - 1)) % n0;
            }
        });
        return;
    }
    // the byte stream of hash_fields_tagged("FRI/index", [seed, 0, q]) with seed = q = 0
    std::vector<uint8_t> b;
    auto app = [&](const char *str) { b.insert(b.end(), str, str + strlen(str)); };
    auto le8 = [&]() { b.insert(b.end(), 8, 0); };
    app("FSv1-TRANSCRIPT-INIT");
    app("FRI/FS");
    app("FSv1-ABSORB-BYTES");
    app("FRI/index");
    app("FSv1-ABSORB-BYTES");
    const uint32_t seed_off = (uint32_t)b.size();
    le8();
    app("FSv1-ABSORB-BYTES");
    le8();
    app("FSv1-ABSORB-BYTES");
    const uint32_t q_off = (uint32_t)b.size();
    le8();
    app("FSv1-CHALLENGE");
    app("out");
    const size_t nb = b.size() / 136 + 1;
    SG_REQUIRE(nb == (size_t)QI_BLOCKS, "query indices: message block count");
    std::vector<uint8_t> pad(136 * nb, 0);
    memcpy(pad.data(), b.data(), b.size());
    pad[b.size()] ^= 0x06;
    pad[pad.size() - 1] ^= 0x80;
    Scoped d_t(c, pad.size()), d_s(c, 8 * K), d_r(c, 8 * K * s0);
    SG_HIP(hipMemcpyAsync(d_t.p, pad.data(), pad.size(), hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(d_s.p, seeds.data(), 8 * K, hipMemcpyHostToDevice, c->stream));
    {
        ProfScope ps(c, "query_indices", 8.0 * K * s0);
        hipLaunchKernelGGL(query_indices_k, dim3((uint32_t)((K * s0 + KT - 1) / KT)), dim3(KT), 0, c->stream,
                           d_t.as<uint64_t>(), d_s.as<uint64_t>(), (uint32_t)K, (uint32_t)s0, seed_off, q_off, npow2,
                           n0, d_r.as<uint64_t>());
        SG_HIP(hipGetLastError());
    }
    std::vector<uint64_t> all(K * s0);
    c->d2h(all.data(), d_r.p, 8 * K * s0);
    SG_HIP(hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < K; k++) std::copy(all.begin() + k * s0, all.begin() + (k + 1) * s0, raw[k].begin());
}

void tree_open_batch(Ctx *c, const DevTree &t, const std::vector<uint64_t> &idx,
                     std::vector<std::vector<uint8_t>> &out) {
    std::vector<std::vector<std::vector<uint8_t>>> m;
    tree_open_multi(c, t, {idx}, m);
    out = std::move(m[0]);
}

}  // namespace sg
