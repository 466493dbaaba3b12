// sha512_air.hip -- the SHA-512 AIR (1510 columns, 1526 transition
// constraints, every 64-bit word as two 32-bit limbs): the host trace builder
// and the composition kernel on gfx950.
//
//   column layout          crates/deep_ali/src/sha512_air.rs:313-481
//   trace builder          sha512_air.rs:218-236, 551-573, 585-898
//   constraint evaluator   sha512_air.rs:909-1282
//   merge                  crates/deep_ali/src/lib.rs:561-653
//
// Same structure as the SHA-256 path: one LDE point per lane, the
// constraints folded into phi = sum_j (j+1) C_j as they are produced.
#include <string.h>

#include "air.hpp"
#include "gl.hpp"
#include "ntt.hpp"
#include "sha256_air.hpp"

namespace sg {

namespace sha5 {

enum : uint32_t {
    A0 = 0,
    BIT_A = 16, BIT_B = 80, BIT_C = 144, BIT_E = 208, BIT_F = 272, BIT_G = 336,
    S0_T = 400, S0_BIT = 464, S0_LO = 528,
    S1_T = 530, S1_BIT = 594, S1_LO = 658,
    CH_BIT = 660, CH_LO = 724,
    MAJ_AB = 726, MAJ_AC = 790, MAJ_BC = 854, MAJ_U = 918, MAJ_BIT = 982, MAJ_LO = 1046,
    W_LO = 1048,
    T1L_LO = 1050, T1L_C0 = 1051, T1L_CW = 1054,
    T1H_LO = 1055, T1H_C0 = 1056, T1H_CW = 1059,
    T2L_LO = 1060, T2L_C0 = 1061, T2L_CW = 1062, T2H_LO = 1063, T2H_C0 = 1064, T2H_CW = 1065,
    H0 = 1066, WW0 = 1082, BIT_WW14 = 1112, BIT_WW1 = 1176,
    SS0_T = 1240, SS0_BIT = 1304, SS0_LO = 1368,
    SS1_T = 1370, SS1_BIT = 1434, SS1_LO = 1498,
    WSL_C0 = 1500, WSL_CW = 1502, WSH_C0 = 1503, WSH_CW = 1505,
    NA_LO = 1506, NA_HI = 1507, NE_LO = 1508, NE_HI = 1509,
};

#define SG_SHA512_K 0x428a2f98d728ae22ull, 0x7137449123ef65cdull, 0xb5c0fbcfec4d3b2full, 0xe9b5dba58189dbbcull, \
    0x3956c25bf348b538ull, 0x59f111f1b605d019ull, 0x923f82a4af194f9bull, 0xab1c5ed5da6d8118ull, \
    0xd807aa98a3030242ull, 0x12835b0145706fbeull, 0x243185be4ee4b28cull, 0x550c7dc3d5ffb4e2ull, \
    0x72be5d74f27b896full, 0x80deb1fe3b1696b1ull, 0x9bdc06a725c71235ull, 0xc19bf174cf692694ull, \
    0xe49b69c19ef14ad2ull, 0xefbe4786384f25e3ull, 0x0fc19dc68b8cd5b5ull, 0x240ca1cc77ac9c65ull, \
    0x2de92c6f592b0275ull, 0x4a7484aa6ea6e483ull, 0x5cb0a9dcbd41fbd4ull, 0x76f988da831153b5ull, \
    0x983e5152ee66dfabull, 0xa831c66d2db43210ull, 0xb00327c898fb213full, 0xbf597fc7beef0ee4ull, \
    0xc6e00bf33da88fc2ull, 0xd5a79147930aa725ull, 0x06ca6351e003826full, 0x142929670a0e6e70ull, \
    0x27b70a8546d22ffcull, 0x2e1b21385c26c926ull, 0x4d2c6dfc5ac42aedull, 0x53380d139d95b3dfull, \
    0x650a73548baf63deull, 0x766a0abb3c77b2a8ull, 0x81c2c92e47edaee6ull, 0x92722c851482353bull, \
    0xa2bfe8a14cf10364ull, 0xa81a664bbc423001ull, 0xc24b8b70d0f89791ull, 0xc76c51a30654be30ull, \
    0xd192e819d6ef5218ull, 0xd69906245565a910ull, 0xf40e35855771202aull, 0x106aa07032bbd1b8ull, \
    0x19a4c116b8d2d0c8ull, 0x1e376c085141ab53ull, 0x2748774cdf8eeb99ull, 0x34b0bcb5e19b48a8ull, \
    0x391c0cb3c5c95a63ull, 0x4ed8aa4ae3418acbull, 0x5b9cca4f7763e373ull, 0x682e6ff3d6b2b8a3ull, \
    0x748f82ee5defb2fcull, 0x78a5636f43172f60ull, 0x84c87814a1f0ab72ull, 0x8cc702081a6439ecull, \
    0x90befffa23631e28ull, 0xa4506cebde82bde9ull, 0xbef9a3f7b2c67915ull, 0xc67178f2e372532bull, \
    0xca273eceea26619cull, 0xd186b8c721c0c207ull, 0xeada7dd6cde0eb1eull, 0xf57d4f7fee6ed178ull, \
    0x06f067aa72176fbaull, 0x0a637dc5a2c898a6ull, 0x113f9804bef90daeull, 0x1b710b35131c471bull, \
    0x28db77f523047d84ull, 0x32caab7b40c72493ull, 0x3c9ebe0a15c9bebcull, 0x431d67c49c100d4cull, \
    0x4cc5d4becb3e42b6ull, 0x597f299cfc657e2aull, 0x5fcb6fab3ad6faecull, 0x6c44198c4a475817ull

static const uint64_t KH[80] = {SG_SHA512_K};  // FIPS 180-4 section 4.2.3
__constant__ uint64_t KD[80] = {SG_SHA512_K};
static const uint64_t IV[8] = {0x6a09e667f3bcc908ull, 0xbb67ae8584caa73bull, 0x3c6ef372fe94f82bull, 0xa54ff53a5f1d36f1ull, 0x510e527fade682d1ull, 0x9b05688c2b3e6c1full, 0x1f83d9abfb41bd6bull, 0x5be0cd19137e2179ull};  // section 5.3.5

SG_HD uint64_t rotr(uint64_t x, int n) { return (x >> n) | (x << (64 - n)); }
inline uint64_t big0(uint64_t x) { return rotr(x, 28) ^ rotr(x, 34) ^ rotr(x, 39); }
inline uint64_t big1(uint64_t x) { return rotr(x, 14) ^ rotr(x, 18) ^ rotr(x, 41); }
inline uint64_t small0(uint64_t x) { return rotr(x, 1) ^ rotr(x, 8) ^ (x >> 7); }
inline uint64_t small1(uint64_t x) { return rotr(x, 19) ^ rotr(x, 61) ^ (x >> 6); }
inline uint64_t chf(uint64_t e, uint64_t f, uint64_t g) { return (e & f) ^ (~e & g); }
inline uint64_t majf(uint64_t a, uint64_t b, uint64_t c) { return (a & b) ^ (a & c) ^ (b & c); }
SG_HD uint64_t lo(uint64_t x) { return x & 0xFFFFFFFFu; }
SG_HD uint64_t hi(uint64_t x) { return x >> 32; }

// Every cell of a trace row is a 32-bit limb of one of these words, one of
// their bits, or a carry (46 u64 per row against 1510 cells).
struct RowWords {
    uint64_t st[8];                   // a..h
    uint64_t s0t, s1t, chw, mjw, wt;  // rotr-xor halves of Sigma0/Sigma1, Ch, Maj, W[t]
    uint64_t t1l, t1h, t2l, t2h;      // per-limb T1 / T2 sums (lo carry folded into hi)
    uint64_t hs[8], ww[15];           // H-state, W window
    uint64_t ss0t, ss1t;              // rotr-xor halves of sigma0(W[t-14]), sigma1(W[t-1])
    uint64_t sl, sh;                  // per-limb schedule sums
    uint32_t na_l, na_h, ne_l, ne_h;  // new_a / new_e limb carries
};

// The 1510 cells of one row in column order (host builder and device expansion).
template <class Put>
SG_HD void emit_row(const RowWords &r, Put put) {
    uint32_t c = 0;
    auto cell = [&](uint64_t v) { put(c++, v); };
    auto limbs = [&](uint64_t v) { cell(lo(v)), cell(hi(v)); };
    auto bits = [&](uint64_t v) {
        for (int i = 0; i < 64; i++) put(c++, (v >> i) & 1);
    };
    auto carry = [&](uint64_t sum, int nbits) {  // bits of sum >> 32, then the packed carry
        for (int k = 0; k < nbits; k++) cell((sum >> (32 + k)) & 1);
        cell(sum >> 32);
    };
    for (int k = 0; k < 8; k++) limbs(r.st[k]);                            // A
    bits(r.st[0]), bits(r.st[1]), bits(r.st[2]);                           // B
    bits(r.st[4]), bits(r.st[5]), bits(r.st[6]);
    const uint64_t s0 = r.s0t ^ rotr(r.st[0], 39), s1 = r.s1t ^ rotr(r.st[4], 41);
    bits(r.s0t), bits(s0), limbs(s0);                                      // C0
    bits(r.s1t), bits(s1), limbs(s1);                                      // C1
    bits(r.chw), limbs(r.chw);                                             // D
    const uint64_t ab = r.st[0] & r.st[1], ac = r.st[0] & r.st[2], bc = r.st[1] & r.st[2];
    bits(ab), bits(ac), bits(bc), bits(ab ^ ac), bits(r.mjw), limbs(r.mjw);  // E
    limbs(r.wt);                                                           // F
    cell(lo(r.t1l)), carry(r.t1l, 3), cell(lo(r.t1h)), carry(r.t1h, 3);    // G
    cell(lo(r.t2l)), carry(r.t2l, 1), cell(lo(r.t2h)), carry(r.t2h, 1);
    for (int k = 0; k < 8; k++) limbs(r.hs[k]);                           // H
    for (int k = 0; k < 15; k++) limbs(r.ww[k]);                          // I
    bits(r.ww[14]), bits(r.ww[1]);                                         // J, K
    const uint64_t ss0 = r.ss0t ^ (r.ww[1] >> 7), ss1 = r.ss1t ^ (r.ww[14] >> 6);
    bits(r.ss0t), bits(ss0), limbs(ss0);                                   // L
    bits(r.ss1t), bits(ss1), limbs(ss1);                                   // M
    carry(r.sl, 2), carry(r.sh, 2);                                        // N
    cell(r.na_l), cell(r.na_h), cell(r.ne_l), cell(r.ne_h);                // O
}

// Row words of one 128-row block (sha512_air.rs:585-898): rows 0..79 rounds,
// row 80 the post-compression state, rows 81..127 idle; H-state = incoming
// IV through row 80, the block digest from row 81. Returns the
// post-compression state.
std::array<uint64_t, 8> block_words(const uint8_t *blk, const uint64_t iv[8], RowWords *R) {
    uint64_t w[80];
    for (int i = 0; i < 16; i++) {
        uint64_t v = 0;
        for (int b = 0; b < 8; b++) v = v << 8 | blk[8 * i + b];
        w[i] = v;
    }
    for (int t = 16; t < 80; t++) w[t] = small1(w[t - 2]) + w[t - 7] + small0(w[t - 15]) + w[t - 16];
    auto wat = [&](int t) { return (t >= 0 && t < 80) ? w[t] : 0ull; };
    uint64_t st[81][8];
    for (int k = 0; k < 8; k++) st[0][k] = iv[k];
    for (int t = 0; t < 80; t++) {
        const uint64_t *q = st[t];
        const uint64_t t1 = q[7] + big1(q[4]) + chf(q[4], q[5], q[6]) + KH[t] + w[t];
        const uint64_t t2 = big0(q[0]) + majf(q[0], q[1], q[2]);
        const uint64_t nx[8] = {t1 + t2, q[0], q[1], q[2], q[3] + t1, q[4], q[5], q[6]};
        memcpy(st[t + 1], nx, 64);
    }
    uint64_t dg[8];
    for (int k = 0; k < 8; k++) dg[k] = iv[k] + st[80][k];
    for (int br = 0; br < 128; br++) {
        RowWords &r = R[br];
        memcpy(r.st, st[br <= 80 ? br : 80], 64);
        const uint64_t a = r.st[0], b = r.st[1], c = r.st[2], d = r.st[3], e = r.st[4], f = r.st[5], g = r.st[6];
        r.s0t = rotr(a, 28) ^ rotr(a, 34);
        r.s1t = rotr(e, 14) ^ rotr(e, 18);
        r.chw = chf(e, f, g);
        r.mjw = majf(a, b, c);
        r.wt = br < 80 ? w[br] : 0;
        const uint64_t kt = br < 80 ? KH[br] : 0, s0 = big0(a), s1 = big1(e);
        r.t1l = lo(r.st[7]) + lo(s1) + lo(r.chw) + lo(kt) + lo(r.wt);
        r.t1h = hi(r.st[7]) + hi(s1) + hi(r.chw) + hi(kt) + hi(r.wt) + (r.t1l >> 32);
        r.t2l = lo(s0) + lo(r.mjw);
        r.t2h = hi(s0) + hi(r.mjw) + (r.t2l >> 32);
        memcpy(r.hs, br <= 80 ? iv : dg, 64);
        for (int k = 0; k < 15; k++) r.ww[k] = wat(br - 15 + k);
        const uint64_t ww14 = r.ww[14], ww1 = r.ww[1], ss0 = small0(ww1), ss1 = small1(ww14), w9 = wat(br - 6);
        r.ss0t = rotr(ww1, 1) ^ rotr(ww1, 8);
        r.ss1t = rotr(ww14, 19) ^ rotr(ww14, 61);
        r.sl = lo(r.ww[0]) + lo(ss0) + lo(w9) + lo(ss1);
        r.sh = hi(r.ww[0]) + hi(ss0) + hi(w9) + hi(ss1) + (r.sl >> 32);
        if (br < 80) {
            const uint64_t nal = lo(r.t1l) + lo(r.t2l), nah = lo(r.t1h) + lo(r.t2h) + (nal >> 32);
            const uint64_t nel = lo(d) + lo(r.t1l), neh = hi(d) + lo(r.t1h) + (nel >> 32);
            r.na_l = (uint32_t)(nal >> 32);
            r.na_h = (uint32_t)(nah >> 32);
            r.ne_l = (uint32_t)(nel >> 32);
            r.ne_h = (uint32_t)(neh >> 32);
        } else {
            r.na_l = r.na_h = r.ne_l = r.ne_h = 0;
        }
    }
    std::array<uint64_t, 8> post;
    for (int k = 0; k < 8; k++) post[k] = st[80][k];
    return post;
}

// Device expansion (as for SHA-256): lane = row, rows past the last block
// repeat the last useful row.
// rows h of trace k at tr + k * pstride, column c at offset c * cstride
__global__ __launch_bounds__(256) void expand_sha512_k(const RowWords *words, const uint64_t *off,
                                                       const uint32_t *useful, uint64_t h, uint64_t cstride,
                                                       uint64_t pstride, uint64_t *tr) {
    const uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= h) return;
    const uint32_t k = blockIdx.y;
    const uint64_t src = off[k] + (r < useful[k] ? r : useful[k] - 1);
    const RowWords rw = words[src];
    uint64_t *o = tr + (uint64_t)k * pstride + r;
    emit_row(rw, [&](uint32_t c, uint64_t v) { o[(uint64_t)c * cstride] = v; });
}

// ---- constraints (sha512_air.rs:909-1282), one LDE point per lane ----
// phi accumulator: slot w - 1 weighted by lam[w - 1], or by w without a table
// The weighted terms are summed as a 192-bit integer (lo, mid, top; each
// product is below 2^128, so top counts at most one carry per term) and
// reduced once by value(): a 64x64 product and a three-word add per term
// instead of a reduced product and a modular add.
struct Phi {
    uint64_t lo = 0, mid = 0, w = 1;
    uint32_t top = 0;
    const uint64_t *lam = nullptr;
    __device__ __forceinline__ uint64_t weight() const { return lam ? lam[w - 1] : w; }
    __device__ __forceinline__ void put(uint64_t v) {
        uint64_t pl, ph;
        mul128(weight(), v, pl, ph);  // ph <= 2^64 - 2: ph + carry does not wrap
        lo += pl;
        const uint64_t h = ph + (lo < pl);
        mid += h;
        top += mid < h;
        w++;
    }
    // lo + mid 2^64 + top 2^128, with 2^128 = -2^32 (mod p)
    __device__ __forceinline__ uint64_t value() const { return sub(reduce128(lo, mid), (uint64_t)top << 32); }
    __device__ __forceinline__ void skip(int k) { w += k; }
};
__device__ __forceinline__ uint64_t xorp(uint64_t x, uint64_t y) {
    const uint64_t p = mul(x, y);
    return sub(add(x, y), add(p, p));
}
__device__ __forceinline__ uint64_t boolc(uint64_t b) { return mul(b, sub(1, b)); }

// SP waves per group of 64 points, one point per lane: wave `part`
// evaluates the constraint sections s with s % SP == part and skips the
// slots of the others (the slot weights stay in step); the partial sums meet
// in LDS. One point's ~1 550 constraints split over SP waves keep many CUs
// busy even for a small proof's few hundred points.
constexpr int SP = 8;
constexpr int AT = 64 * SP;

// gate groups, as for SHA-256: ungated, compression rows, the finalisation
// row, every other row, message-schedule rows
enum { G_ALL, G_COMP, G_FIN, G_NOTFIN, G_SCHED, NG };

// GROUPS = false: phi at every point. GROUPS = true: the gate groups'
// weighted sums with the gates ignored and the round constant left out,
// phi[(k * NG + g) * n + i]; t1_lam[0..1] get the weights of the two T1
// limb identities.
// GROUPS reads the blowup-2 LDE as its even points (lde: the traces) and
// its odd points (odd), n / 2 rows each, as compose_sha256_k does.
// wcols: columns per proof of the input (the SHA-512 cells are its first
// 1510); row_limit: phi is zero on rows r with r + 1 >= row_limit (the SHA
// rows of a larger AIR); lam: the slots' coefficients, or j + 1 when null.
template <bool GROUPS>
__global__ __launch_bounds__(AT) void compose_sha512_k(const uint64_t *lde, const uint64_t *odd, uint64_t n,
                                                       uint64_t blowup, const uint32_t *n_blocks, uint64_t *phi,
                                                       uint64_t *t1_lam, uint64_t wcols, uint64_t row_limit,
                                                       const uint64_t *lam) {
    __shared__ uint64_t red[5][SP][64];
    const uint32_t part = threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint64_t i = (uint64_t)blockIdx.x * 64 + lane;  // n is a multiple of 64
    // rows past row_limit get phi = 0: every section skipped
    const bool dead = !GROUPS && i / blowup + 1 >= row_limit;
    auto mine = [&](int sec) { return !dead && sec % SP == (int)part; };
    const uint64_t rows = GROUPS ? n / 2 : n, at = GROUPS ? i >> 1 : i;
    const uint64_t *L = (GROUPS && (i & 1) ? odd : lde) + wcols * rows * blockIdx.y;
    const uint64_t ni = GROUPS ? (at + 1) & (rows - 1) : (i + blowup) & (n - 1);
    auto cur = [&](uint32_t c) { return L[(uint64_t)c * rows + at]; };
    auto nxt = [&](uint32_t c) { return L[(uint64_t)c * rows + ni]; };
    const uint64_t row = i / blowup, blk = row >> 7, br = row & 127;
    const bool past = GROUPS || blk >= n_blocks[blockIdx.y];
    const bool comp = GROUPS || (!past && br < 80), fin = !past && br == 80;
    const bool sched = GROUPS || (!past && br >= 15 && br <= 78);
    const uint64_t TWO32 = 1ULL << 32;
    Phi f, fc, ff, fn, fs;  // ungated; compression, finalisation, not finalisation, schedule (GROUPS mode)
    f.lam = fc.lam = ff.lam = fn.lam = fs.lam = lam;
    Phi &gc = GROUPS ? fc : f, &gs = GROUPS ? fs : f;
    const uint32_t boff[8] = {BIT_A, BIT_B, BIT_C, BIT_E, BIT_F, BIT_G, BIT_WW14, BIT_WW1};
#pragma unroll 1
    for (int o = 0; o < 8; o++) {  // sections 0-7
        if (!mine(o)) {
            f.skip(64);
            continue;
        }
#pragma unroll 4
        for (int b = 0; b < 64; b++) f.put(boolc(cur(boff[o] + b)));
    }
    auto pack = [&](uint32_t off) {  // sum_{b<32} 2^b cell(off + b)
        uint64_t s = 0;
#pragma unroll 8
        for (int b = 0; b < 32; b++) s = add(s, mul(1ULL << b, cur(off + b)));
        return s;
    };
    const uint32_t wlo[8] = {A0 + 0, A0 + 2, A0 + 4, A0 + 8, A0 + 10, A0 + 12, WW0 + 28, WW0 + 2};
#pragma unroll 1
    for (int o = 0; o < 8; o++) {  // sections 8-15
        if (mine(8 + o)) {
            f.put(sub(cur(wlo[o]), pack(boff[o])));
            f.put(sub(cur(wlo[o] + 1), pack(boff[o] + 32)));
        } else {
            f.skip(2);
        }
    }
    struct Sg {
        uint32_t src, t, b, lo;
        int r1, r2, r3;
        bool shr;
    };
    const Sg sg[4] = {{BIT_A, S0_T, S0_BIT, S0_LO, 28, 34, 39, false},
                      {BIT_E, S1_T, S1_BIT, S1_LO, 14, 18, 41, false},
                      {BIT_WW1, SS0_T, SS0_BIT, SS0_LO, 1, 8, 7, true},
                      {BIT_WW14, SS1_T, SS1_BIT, SS1_LO, 19, 61, 6, true}};
#pragma unroll 1
    for (int s = 0; s < 4; s++) {  // sections 16-23
        const Sg &g = sg[s];
        if (mine(16 + 2 * s)) {
#pragma unroll 2
            for (int b = 0; b < 64; b++)
                f.put(sub(cur(g.t + b), xorp(cur(g.src + (b + g.r1) % 64), cur(g.src + (b + g.r2) % 64))));
        } else {
            f.skip(64);
        }
        if (mine(17 + 2 * s)) {
#pragma unroll 2
            for (int b = 0; b < 64; b++) {
                const uint64_t third =
                    g.shr ? (b + g.r3 < 64 ? cur(g.src + b + g.r3) : 0) : cur(g.src + (b + g.r3) % 64);
                f.put(sub(cur(g.b + b), xorp(cur(g.t + b), third)));
            }
            f.put(sub(cur(g.lo), pack(g.b)));
            f.put(sub(cur(g.lo + 1), pack(g.b + 32)));
        } else {
            f.skip(66);
        }
    }
    if (mine(24)) {  // Ch (section 24)
#pragma unroll 2
        for (int b = 0; b < 64; b++) {
            const uint64_t e = cur(BIT_E + b), ff = cur(BIT_F + b), g = cur(BIT_G + b);
            f.put(sub(cur(CH_BIT + b), sub(add(mul(e, ff), g), mul(e, g))));
        }
        f.put(sub(cur(CH_LO), pack(CH_BIT)));
        f.put(sub(cur(CH_LO + 1), pack(CH_BIT + 32)));
    } else {
        f.skip(66);
    }
    if (mine(25)) {  // Maj auxiliaries (sections 25-27)
#pragma unroll 2
        for (int b = 0; b < 64; b++) {
            const uint64_t a = cur(BIT_A + b), bb = cur(BIT_B + b), c = cur(BIT_C + b);
            f.put(sub(cur(MAJ_AB + b), mul(a, bb)));
            f.put(sub(cur(MAJ_AC + b), mul(a, c)));
            f.put(sub(cur(MAJ_BC + b), mul(bb, c)));
        }
    } else {
        f.skip(192);
    }
    if (mine(26)) {
#pragma unroll 2
        for (int b = 0; b < 64; b++) f.put(sub(cur(MAJ_U + b), xorp(cur(MAJ_AB + b), cur(MAJ_AC + b))));
    } else {
        f.skip(64);
    }
    if (mine(27)) {
#pragma unroll 2
        for (int b = 0; b < 64; b++) f.put(sub(cur(MAJ_BIT + b), xorp(cur(MAJ_U + b), cur(MAJ_BC + b))));
        f.put(sub(cur(MAJ_LO), pack(MAJ_BIT)));
        f.put(sub(cur(MAJ_LO + 1), pack(MAJ_BIT + 32)));
    } else {
        f.skip(66);
    }
    // T1 per limb (carry-in from the lo limb), T2 per limb
    const uint64_t kt = GROUPS ? 0 : comp ? KD[br] : 0;
    auto carry3 = [&](uint32_t c0, uint32_t cw) {
        f.put(boolc(cur(c0)));
        f.put(boolc(cur(c0 + 1)));
        f.put(boolc(cur(c0 + 2)));
        f.put(sub(sub(sub(cur(cw), cur(c0)), mul(2, cur(c0 + 1))), mul(4, cur(c0 + 2))));
    };
    if (mine(28)) {  // T1 / T2 limbs (section 28)
        if (GROUPS && i == 0 && blockIdx.y == 0) t1_lam[0] = f.weight();
        f.put(sub(add(cur(T1L_LO), mul(TWO32, cur(T1L_CW))),
                  add(add(add(add(cur(A0 + 14), cur(S1_LO)), cur(CH_LO)), kt & 0xFFFFFFFFu), cur(W_LO))));
        carry3(T1L_C0, T1L_CW);
        if (GROUPS && i == 0 && blockIdx.y == 0) t1_lam[1] = f.weight();
        f.put(sub(add(cur(T1H_LO), mul(TWO32, cur(T1H_CW))),
                  add(add(add(add(add(cur(A0 + 15), cur(S1_LO + 1)), cur(CH_LO + 1)), kt >> 32), cur(W_LO + 1)),
                      cur(T1L_CW))));
        carry3(T1H_C0, T1H_CW);
        f.put(sub(add(cur(T2L_LO), mul(TWO32, cur(T2L_CW))), add(cur(S0_LO), cur(MAJ_LO))));
        f.put(boolc(cur(T2L_C0)));
        f.put(sub(cur(T2L_CW), cur(T2L_C0)));
        f.put(sub(add(cur(T2H_LO), mul(TWO32, cur(T2H_CW))),
                  add(add(cur(S0_LO + 1), cur(MAJ_LO + 1)), cur(T2L_CW))));
        f.put(boolc(cur(T2H_C0)));
        f.put(sub(cur(T2H_CW), cur(T2H_C0)));
    } else {
        f.skip(16);
    }
    if (comp && mine(29)) {  // register shifts and new a / e (section 29)
        gc.w = f.w;
        // shifts b<-a, c<-b, d<-c, f<-e, g<-f, h<-g, both limbs
        const uint32_t dst[6] = {1, 2, 3, 5, 6, 7};
#pragma unroll
        for (int k = 0; k < 6; k++) {
            gc.put(sub(nxt(A0 + 2 * dst[k]), cur(A0 + 2 * (dst[k] - 1))));
            gc.put(sub(nxt(A0 + 2 * dst[k] + 1), cur(A0 + 2 * (dst[k] - 1) + 1)));
        }
        gc.put(sub(sub(add(nxt(A0), mul(TWO32, cur(NA_LO))), cur(T1L_LO)), cur(T2L_LO)));
        gc.put(boolc(cur(NA_LO)));
        gc.put(sub(sub(sub(add(nxt(A0 + 1), mul(TWO32, cur(NA_HI))), cur(T1H_LO)), cur(T2H_LO)), cur(NA_LO)));
        gc.put(boolc(cur(NA_HI)));
        gc.put(sub(sub(add(nxt(A0 + 8), mul(TWO32, cur(NE_LO))), cur(A0 + 6)), cur(T1L_LO)));
        gc.put(boolc(cur(NE_LO)));
        gc.put(sub(sub(sub(add(nxt(A0 + 9), mul(TWO32, cur(NE_HI))), cur(A0 + 7)), cur(T1H_LO)), cur(NE_LO)));
        gc.put(boolc(cur(NE_HI)));
        f.w = gc.w;
    } else {
        f.skip(20);
    }
    // H-state: finalisation with the lo carry delta_lo * 2^-32 folded into hi
    constexpr uint64_t INV32 = 0xFFFFFFFE00000002ull;  // 2^-32 mod p
    static_assert((unsigned __int128)INV32 * (1ull << 32) % P == 1, "2^-32");
#pragma unroll 1
    for (int k = 0; k < 8; k++) {  // section 30
        if (!mine(30)) {
            f.skip(2);
            continue;
        }
        if (GROUPS || fin) {
            const uint64_t dl = sub(add(cur(H0 + 2 * k), cur(A0 + 2 * k)), nxt(H0 + 2 * k));
            const uint64_t dh = sub(add(add(cur(H0 + 2 * k + 1), cur(A0 + 2 * k + 1)), mul(dl, INV32)),
                                    nxt(H0 + 2 * k + 1));
            Phi &g = GROUPS ? ff : f;
            g.w = f.w;
            g.put(mul(dl, sub(dl, TWO32)));
            g.put(mul(dh, sub(dh, TWO32)));
        }
        if (GROUPS || !fin) {
            Phi &g = GROUPS ? fn : f;
            if (GROUPS) g.w = f.w;
            g.put(sub(nxt(H0 + 2 * k), cur(H0 + 2 * k)));
            g.put(sub(nxt(H0 + 2 * k + 1), cur(H0 + 2 * k + 1)));
        }
        if (GROUPS) f.skip(2);
    }
    if (mine(31)) {  // W window shift (section 31)
#pragma unroll 2
        for (int k = 0; k < 14; k++) {
            f.put(sub(nxt(WW0 + 2 * k), cur(WW0 + 2 * k + 2)));
            f.put(sub(nxt(WW0 + 2 * k + 1), cur(WW0 + 2 * k + 3)));
        }
        f.put(sub(nxt(WW0 + 28), cur(W_LO)));
        f.put(sub(nxt(WW0 + 29), cur(W_LO + 1)));
    } else {
        f.skip(30);
    }
    if (sched && mine(32)) {  // message schedule (section 32, the last)
        gs.w = f.w;
        const uint64_t sl = add(add(add(cur(WW0), cur(SS0_LO)), cur(WW0 + 18)), cur(SS1_LO));
        gs.put(sub(add(nxt(W_LO), mul(TWO32, cur(WSL_CW))), sl));
        gs.put(boolc(cur(WSL_C0)));
        gs.put(boolc(cur(WSL_C0 + 1)));
        gs.put(sub(sub(cur(WSL_CW), cur(WSL_C0)), mul(2, cur(WSL_C0 + 1))));
        const uint64_t sh = add(add(add(add(cur(WW0 + 1), cur(SS0_LO + 1)), cur(WW0 + 19)), cur(SS1_LO + 1)), cur(WSL_CW));
        gs.put(sub(add(nxt(W_LO + 1), mul(TWO32, cur(WSH_CW))), sh));
        gs.put(boolc(cur(WSH_C0)));
        gs.put(boolc(cur(WSH_C0 + 1)));
        gs.put(sub(sub(cur(WSH_CW), cur(WSH_C0)), mul(2, cur(WSH_C0 + 1))));
        f.w = gs.w;
    }
    // the SP partial sums of the point
    uint64_t sv[NG] = {f.value(), fc.value(), ff.value(), fn.value(), fs.value()};  // G_ALL .. G_SCHED
#pragma unroll
    for (int g = 0; g < NG; g++) red[g][part][lane] = sv[g];
    __syncthreads();
    if (part != 0) return;
#pragma unroll
    for (int q = 1; q < SP; q++)
#pragma unroll
        for (int g = 0; g < NG; g++) sv[g] = add(sv[g], red[g][q][lane]);
    if (GROUPS) {
        uint64_t *o = phi + (uint64_t)blockIdx.y * NG * n + i;
#pragma unroll
        for (int g = 0; g < NG; g++) o[g * n] = sv[g];
    } else {
        phi[n * blockIdx.y + i] = sv[0];
    }
}

// phi at every point from the interpolated gate-group polynomials, less the
// round constant's two 32-bit halves in the T1 limb identities on
// compression rows
__global__ __launch_bounds__(256) void combine_sha512_k(const uint64_t *pg, uint64_t n, uint64_t blowup,
                                                        const uint32_t *n_blocks, const uint64_t *t1_lam,
                                                        uint64_t *phi) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint64_t row = i / blowup, blk = row >> 7, br = row & 127;
    const bool past = blk >= n_blocks[blockIdx.y];
    const bool comp = !past && br < 80, fin = !past && br == 80, sched = !past && br >= 15 && br <= 78;
    const uint64_t *q = pg + (uint64_t)blockIdx.y * NG * n + i;
    uint64_t v = add(q[G_ALL * n], fin ? q[G_FIN * n] : q[G_NOTFIN * n]);
    if (comp) {
        const uint64_t kt = KD[br];
        v = sub(sub(add(v, q[G_COMP * n]), mul(t1_lam[0], kt & 0xFFFFFFFFu)), mul(t1_lam[1], kt >> 32));
    }
    if (sched) v = add(v, q[G_SCHED * n]);
    phi[n * blockIdx.y + i] = v;
}

}  // namespace sha5

uint64_t sha512_n_blocks(uint64_t len) { return (len + 17 + 127) / 128; }

uint64_t sha512_trace_height(uint64_t len) {
    const uint64_t useful = 128 * sha512_n_blocks(len);
    uint64_t h = 128;
    while (h < useful) h <<= 1;
    return h;
}

// row words of the padded message (build_sha512_trace_multi,
// sha512_air.rs:551-573); digest from the H-state at block row 81 of the
// last block (read_digest_from_trace, :1285-1297)
static uint64_t sha512_words(const uint8_t *msg, uint64_t len, std::vector<sha5::RowWords> &rows, uint8_t digest[64]) {
    const uint64_t nb = sha512_n_blocks(len);
    std::vector<uint8_t> m(128 * nb, 0);
    if (len) memcpy(m.data(), msg, len);
    m[len] = 0x80;
    const uint64_t bits = len * 8;  // the upper half of the 128-bit length is zero
    for (int i = 0; i < 8; i++) m[128 * nb - 1 - i] = (uint8_t)(bits >> (8 * i));
    rows.resize(128 * nb);
    uint64_t iv[8];
    memcpy(iv, sha5::IV, 64);
    for (uint64_t b = 0; b < nb; b++) {
        const std::array<uint64_t, 8> post = sha5::block_words(&m[128 * b], iv, &rows[128 * b]);
        for (int k = 0; k < 8; k++) iv[k] += post[k];
    }
    const sha5::RowWords &d = rows[128 * (nb - 1) + 81];
    for (int k = 0; k < 8; k++)
        for (int i = 0; i < 8; i++) digest[8 * k + i] = (uint8_t)(d.hs[k] >> (56 - 8 * i));
    return nb;
}

uint64_t build_sha512_trace(const uint8_t *msg, uint64_t len, uint64_t *tr, uint8_t digest[64], uint64_t height) {
    std::vector<sha5::RowWords> rows;
    const uint64_t nb = sha512_words(msg, len, rows, digest), h = height ? height : sha512_trace_height(len);
    SG_REQUIRE(is_pow2(h) && h >= 128 * nb, "sha512 trace: height must be a power of two >= 128 * n_blocks");
    for (uint64_t r = 0; r < h; r++) {
        const sha5::RowWords &w = rows[r < rows.size() ? r : rows.size() - 1];
        sha5::emit_row(w, [&](uint32_t c, uint64_t v) { tr[(uint64_t)c * h + r] = v; });
    }
    return nb;
}

void build_sha512_traces_dev(Ctx *c, const std::vector<ByteSpan> &msgs, uint64_t h, uint64_t *d_tr,
                             std::vector<uint32_t> &n_blocks, std::vector<std::array<uint8_t, 64>> &digests) {
    build_sha512_traces_into(c, msgs, h, h, (uint64_t)SHA512_W * h, d_tr, n_blocks, digests);
}

void build_sha512_traces_into(Ctx *c, const std::vector<ByteSpan> &msgs, uint64_t h, uint64_t cstride,
                              uint64_t pstride, uint64_t *d_tr, std::vector<uint32_t> &n_blocks,
                              std::vector<std::array<uint8_t, 64>> &digests) {
    const size_t K = msgs.size();
    std::vector<std::vector<sha5::RowWords>> rows(K);
    n_blocks.assign(K, 0);
    digests.assign(K, {});
    parallel_for(K, [&](size_t k) {
        n_blocks[k] = (uint32_t)sha512_words(msgs[k].p, msgs[k].n, rows[k], digests[k].data());
    });
    std::vector<uint64_t> off(K);
    std::vector<uint32_t> useful(K);
    uint64_t total = 0;
    for (size_t k = 0; k < K; k++) {
        SG_REQUIRE(128ULL * n_blocks[k] <= h, "sha512 traces: height below a message's minimum");
        off[k] = total;
        useful[k] = (uint32_t)rows[k].size();
        total += rows[k].size();
    }
    // the row words of every proof copied on the host's threads straight into
    // pinned memory (one DMA upload, no zero-filled staging vector)
    sha5::RowWords *all = (sha5::RowWords *)c->upload_staging(sizeof(sha5::RowWords) * total);
    parallel_for(K, [&](size_t k) { std::copy(rows[k].begin(), rows[k].end(), all + off[k]); });
    Scoped d_w(c, sizeof(sha5::RowWords) * total), d_off(c, 8 * K), d_u(c, 4 * K);
    SG_HIP(hipMemcpyAsync(d_w.p, all, sizeof(sha5::RowWords) * total, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(d_off.p, off.data(), 8 * K, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(d_u.p, useful.data(), 4 * K, hipMemcpyHostToDevice, c->stream));
    ProfScope ps(c, "expand_sha512", 8.0 * SHA512_W * h * K);
    hipLaunchKernelGGL(sha5::expand_sha512_k, dim3((uint32_t)((h + 255) / 256), (uint32_t)K), dim3(256), 0, c->stream,
                       d_w.as<sha5::RowWords>(), d_off.as<uint64_t>(), d_u.as<uint32_t>(), h, cstride, pstride,
                       d_tr);
    SG_HIP(hipGetLastError());
    SG_HIP(hipStreamSynchronize(c->stream));
}

void merge_sha512_batch(Ctx *c, const uint64_t *lde, uint32_t K, uint64_t n_trace, uint64_t blowup,
                        const std::vector<uint32_t> &n_blocks, int phi_mode, uint64_t *c_eval, const uint64_t *d_lam) {
    SG_REQUIRE(is_pow2(n_trace) && n_trace >= 128 && is_pow2(blowup) && blowup >= 2,
               "merge_sha512: n_trace must be a power of two >= 128");
    SG_REQUIRE(K >= 1 && K <= 65535 && n_blocks.size() == K, "merge_sha512: one n_blocks per proof");
    const uint64_t n = n_trace * blowup;
    SG_REQUIRE(n <= (1ULL << 24), "merge_sha512: domain above 2^24");
    for (uint32_t nb : n_blocks) SG_REQUIRE(nb >= 1 && 128 * (uint64_t)nb <= n_trace, "merge_sha512: bad n_blocks");
    if (!phi_computed(phi_mode, n)) {
        fill_u64(c, c_eval, n * K, 0);
        return;
    }
    Scoped phi(c, 8 * n * K), d_nb(c, 4 * K);
    SG_HIP(hipMemcpyAsync(d_nb.p, n_blocks.data(), 4 * K, hipMemcpyHostToDevice, c->stream));
    {
        ProfScope ps(c, "compose_sha512", (8.0 * SHA512_W * n + 8.0 * n) * K);
        hipLaunchKernelGGL(sha5::compose_sha512_k<false>, dim3((uint32_t)(n / 64), K),
                           dim3(sha5::AT), 0, c->stream, lde, (const uint64_t *)nullptr, n, blowup,
                           d_nb.as<uint32_t>(), phi.as<uint64_t>(), (uint64_t *)nullptr, (uint64_t)SHA512_W,
                           ~0ULL, d_lam);
        SG_HIP(hipGetLastError());
    }
    quotient_batch(c, phi.as<uint64_t>(), K, n_trace, n, c_eval);
}

void merge_sha512_from_trace_batch(Ctx *c, const uint64_t *d_tr, uint32_t K, uint64_t n_trace, uint64_t blowup,
                                   const std::vector<uint32_t> &n_blocks, int phi_mode, uint64_t *c_eval,
                                   const uint64_t *d_lam) {
    SG_REQUIRE(is_pow2(n_trace) && n_trace >= 128 && is_pow2(blowup) && blowup >= 4,
               "merge_sha512: n_trace must be a power of two >= 128, blowup >= 4");
    SG_REQUIRE(K >= 1 && K <= 65535 && n_blocks.size() == K, "merge_sha512: one n_blocks per proof");
    const uint64_t n = n_trace * blowup, n2 = 2 * n_trace;
    SG_REQUIRE(n <= (1ULL << 24), "merge_sha512: domain above 2^24");
    for (uint32_t nb : n_blocks) SG_REQUIRE(nb >= 1 && 128 * (uint64_t)nb <= n_trace, "merge_sha512: bad n_blocks");
    if (!phi_computed(phi_mode, n)) {
        fill_u64(c, c_eval, n * K, 0);
        return;
    }
    Scoped phi(c, 8 * n * K), d_nb(c, 4 * K), t1(c, 16);
    SG_HIP(hipMemcpyAsync(d_nb.p, n_blocks.data(), 4 * K, hipMemcpyHostToDevice, c->stream));
    {
        Scoped groups(c, 8ULL * sha5::NG * n2 * K);
        {
            Scoped odd(c, 8ULL * SHA512_W * n_trace * K);
            lde_odd_coset(c, d_tr, odd.as<uint64_t>(), ilog2(n_trace), (uint64_t)SHA512_W * K);
            ProfScope ps(c, "compose_sha512", (8.0 * SHA512_W * n2 + 8.0 * sha5::NG * n2) * K);
            hipLaunchKernelGGL(sha5::compose_sha512_k<true>, dim3((uint32_t)(n2 / 64), K),
                               dim3(sha5::AT), 0, c->stream, d_tr, odd.as<uint64_t>(), n2, (uint64_t)2,
                               d_nb.as<uint32_t>(), groups.as<uint64_t>(), t1.as<uint64_t>(), (uint64_t)SHA512_W,
                               ~0ULL, d_lam);
            SG_HIP(hipGetLastError());
        }
        Scoped full(c, 8ULL * sha5::NG * n * K);
        lde_columns(c, groups.as<uint64_t>(), full.as<uint64_t>(), ilog2(n2), ilog2(blowup / 2),
                    (uint64_t)sha5::NG * K, n2, n, 1);
        ProfScope ps(c, "combine_sha512", 8.0 * (sha5::NG + 1) * n * K);
        hipLaunchKernelGGL(sha5::combine_sha512_k, dim3((uint32_t)((n + 255) / 256), K), dim3(256), 0, c->stream,
                           full.as<uint64_t>(), n, blowup, d_nb.as<uint32_t>(), t1.as<uint64_t>(), phi.as<uint64_t>());
        SG_HIP(hipGetLastError());
    }
    quotient_batch(c, phi.as<uint64_t>(), K, n_trace, n, c_eval);
}

void compose_sha512_rows(Ctx *c, const uint64_t *lde, uint64_t wcols, uint32_t K, uint64_t n, uint64_t blowup,
                         const uint32_t *d_nb, uint64_t row_limit, const uint64_t *d_lam, uint64_t *phi) {
    SG_REQUIRE(n % 64 == 0, "compose_sha512: point count not a multiple of 64");
    ProfScope ps(c, "compose_sha512", (8.0 * SHA512_W * n + 8.0 * n) * K);
    hipLaunchKernelGGL(sha5::compose_sha512_k<false>, dim3((uint32_t)(n / 64), K),
                       dim3(sha5::AT), 0, c->stream, lde, (const uint64_t *)nullptr, n, blowup, d_nb, phi,
                       (uint64_t *)nullptr, wcols, row_limit, d_lam);
    SG_HIP(hipGetLastError());
}

void compose_sha512_groups(Ctx *c, const uint64_t *d_tr, const uint64_t *d_odd, uint64_t wcols, uint32_t K,
                           uint64_t n2, const uint32_t *d_nb, const uint64_t *d_lam, uint64_t *groups,
                           uint64_t *t1_lam) {
    SG_REQUIRE(n2 % 64 == 0, "compose_sha512: point count not a multiple of 64");
    ProfScope ps(c, "compose_sha512", (8.0 * SHA512_W * n2 + 8.0 * sha5::NG * n2) * K);
    hipLaunchKernelGGL(sha5::compose_sha512_k<true>, dim3((uint32_t)(n2 / 64), K),
                       dim3(sha5::AT), 0, c->stream, d_tr, d_odd, n2, (uint64_t)2, d_nb, groups, t1_lam, wcols,
                       ~0ULL, d_lam);
    SG_HIP(hipGetLastError());
}

const uint64_t *sha512_round_constants() { return sha5::KH; }

}  // namespace sg
