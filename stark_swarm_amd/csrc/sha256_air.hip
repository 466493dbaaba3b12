// sha256_air.hip -- the SHA-256 AIR (756 columns, 766 transition
// constraints) for prove_ds_ksk_binding: the host trace builder and the
// composition kernel on gfx950.
//
//   column layout          crates/deep_ali/src/sha256_air.rs:140-253
//   trace builder          sha256_air.rs:392-479, 648-898
//   constraint evaluator   sha256_air.rs:956-1294
//   merge                  crates/deep_ali/src/lib.rs:453-545
//
// The composition runs one LDE point per lane: the lane reads its row and
// the row `blowup` points ahead straight from the column-major LDE (adjacent
// lanes read adjacent addresses of every column) and folds the 766
// constraints into phi = sum_j (j+1) C_j as it goes; constraints a row's
// phase switches off are skipped, not evaluated.
#include <string.h>

#include "air.hpp"
#include "gl.hpp"
#include "ntt.hpp"
#include "sha256_air.hpp"

namespace sg {

namespace sha {

enum : uint32_t {
    A_A = 0, A_D = 3, A_E = 4, A_H = 7,
    BIT_A = 8, BIT_B = 40, BIT_C = 72, BIT_E = 104, BIT_F = 136, BIT_G = 168,
    S0_T = 200, S0_BIT = 232, S0_W = 264,
    S1_T = 265, S1_BIT = 297, S1_W = 329,
    CH_BIT = 330, CH_W = 362,
    MAJ_AB = 363, MAJ_AC = 395, MAJ_BC = 427, MAJ_U = 459, MAJ_BIT = 491, MAJ_W = 523,
    COL_W = 524,
    T1_LO = 525, T1_C0 = 526, T1_C1 = 527, T1_C2 = 528, T1_CW = 529,
    T2_LO = 530, T2_C0 = 531, T2_C1 = 532, T2_CW = 533,
    H0 = 534, WW0 = 542, BIT_WW14 = 557, BIT_WW1 = 589,
    SS0_T = 621, SS0_BIT = 653, SS0_W = 685,
    SS1_T = 686, SS1_BIT = 718, SS1_W = 750,
    WS_C0 = 751, WS_C1 = 752, WS_CW = 753, NA_C = 754, NE_C = 755,
};

#define SG_SHA_K                                                                                                 \
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u,      \
        0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u,  \
        0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,  \
        0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u,  \
        0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u,  \
        0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,  \
        0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,  \
        0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u

static const uint32_t KH[64] = {SG_SHA_K};     // FIPS 180-4 section 4.2.2
__constant__ uint32_t KD[64] = {SG_SHA_K};
static const uint32_t IV[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                               0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};

SG_HD uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
SG_HD uint32_t big0(uint32_t x) { return rotr(x, 2) ^ rotr(x, 13) ^ rotr(x, 22); }
SG_HD uint32_t big1(uint32_t x) { return rotr(x, 6) ^ rotr(x, 11) ^ rotr(x, 25); }
SG_HD uint32_t small0(uint32_t x) { return rotr(x, 7) ^ rotr(x, 18) ^ (x >> 3); }
SG_HD uint32_t small1(uint32_t x) { return rotr(x, 17) ^ rotr(x, 19) ^ (x >> 10); }
inline uint32_t chf(uint32_t e, uint32_t f, uint32_t g) { return (e & f) ^ (~e & g); }
inline uint32_t majf(uint32_t a, uint32_t b, uint32_t c) { return (a & b) ^ (a & c) ^ (b & c); }

// Every cell of a trace row is one of these words or one of their bits
// (44 u32 per row against 756 cells).
struct RowWords {
    uint32_t st[8];           // a..h
    uint32_t s0t, s1t;        // rotr-xor halves of Sigma0(a), Sigma1(e)
    uint32_t chw, mjw, wt;    // Ch, Maj, W[t]
    uint32_t t1lo, t1c, t2lo, t2c;  // T1 / T2 mod 2^32 and their carries
    uint32_t hs[8];           // H-state
    uint32_t ww[15];          // W window W[t-15..t-1]
    uint32_t ss0t, ss1t;      // rotr-xor halves of sigma0(W[t-14]), sigma1(W[t-1])
    uint32_t sc, na, ne;      // schedule carry, new_a / new_e carries
};

// The 756 cells of one row in column order: put(column, value). Shared by the
// host builder and the device expansion, so both lay out the same trace.
template <class Put>
SG_HD void emit_row(const RowWords &r, Put put) {
    uint32_t c = 0;
    auto word = [&](uint32_t v) { put(c++, (uint64_t)v); };
    auto bits = [&](uint32_t v) {
        for (int i = 0; i < 32; i++) put(c++, (uint64_t)((v >> i) & 1));
    };
    for (int k = 0; k < 8; k++) word(r.st[k]);                          // A
    bits(r.st[0]), bits(r.st[1]), bits(r.st[2]);                         // B
    bits(r.st[4]), bits(r.st[5]), bits(r.st[6]);
    const uint32_t s0 = r.s0t ^ rotr(r.st[0], 22), s1 = r.s1t ^ rotr(r.st[4], 25);
    bits(r.s0t), bits(s0), word(s0);                                     // C0
    bits(r.s1t), bits(s1), word(s1);                                     // C1
    bits(r.chw), word(r.chw);                                            // D
    const uint32_t ab = r.st[0] & r.st[1], ac = r.st[0] & r.st[2], bc = r.st[1] & r.st[2];
    bits(ab), bits(ac), bits(bc), bits(ab ^ ac), bits(r.mjw), word(r.mjw);  // E
    word(r.wt);                                                          // F
    word(r.t1lo), word(r.t1c & 1), word((r.t1c >> 1) & 1), word((r.t1c >> 2) & 1), word(r.t1c);  // G
    word(r.t2lo), word(r.t2c & 1), word((r.t2c >> 1) & 1), word(r.t2c);
    for (int k = 0; k < 8; k++) word(r.hs[k]);                          // H
    for (int k = 0; k < 15; k++) word(r.ww[k]);                         // I
    bits(r.ww[14]), bits(r.ww[1]);                                       // J, K
    const uint32_t ss0 = r.ss0t ^ (r.ww[1] >> 3), ss1 = r.ss1t ^ (r.ww[14] >> 10);
    bits(r.ss0t), bits(ss0), word(ss0);                                  // L
    bits(r.ss1t), bits(ss1), word(ss1);                                  // M
    word(r.sc & 1), word((r.sc >> 1) & 1), word(r.sc);                   // N
    word(r.na), word(r.ne);                                              // O
}

// Row words of one 128-row block (sha256_air.rs:648-898): rows 0..63 are the
// rounds (state before round r), row 64 the post-compression state, rows
// 65..127 idle with the state held; H-state is the incoming IV through row
// 64 and the block digest from row 65. Returns the post-compression state.
std::array<uint32_t, 8> block_words(const uint8_t *blk, const uint32_t iv[8], RowWords *R) {
    uint32_t w[64];
    for (int i = 0; i < 16; i++)
        w[i] = (uint32_t)blk[4 * i] << 24 | (uint32_t)blk[4 * i + 1] << 16 | (uint32_t)blk[4 * i + 2] << 8 |
               blk[4 * i + 3];
    for (int t = 16; t < 64; t++) w[t] = small1(w[t - 2]) + w[t - 7] + small0(w[t - 15]) + w[t - 16];
    auto wat = [&](int t) { return (t >= 0 && t < 64) ? w[t] : 0u; };
    uint32_t st[65][8];
    for (int k = 0; k < 8; k++) st[0][k] = iv[k];
    for (int t = 0; t < 64; t++) {
        const uint32_t *q = st[t];
        const uint32_t t1 = q[7] + big1(q[4]) + chf(q[4], q[5], q[6]) + KH[t] + w[t];
        const uint32_t t2 = big0(q[0]) + majf(q[0], q[1], q[2]);
        const uint32_t nx[8] = {t1 + t2, q[0], q[1], q[2], q[3] + t1, q[4], q[5], q[6]};
        memcpy(st[t + 1], nx, 32);
    }
    uint32_t dg[8];
    for (int k = 0; k < 8; k++) dg[k] = iv[k] + st[64][k];
    for (int br = 0; br < 128; br++) {
        RowWords &r = R[br];
        memcpy(r.st, st[br <= 64 ? br : 64], 32);
        const uint32_t a = r.st[0], b = r.st[1], c = r.st[2], d = r.st[3], e = r.st[4], f = r.st[5], g = r.st[6];
        r.s0t = rotr(a, 2) ^ rotr(a, 13);
        r.s1t = rotr(e, 6) ^ rotr(e, 11);
        r.chw = chf(e, f, g);
        r.mjw = majf(a, b, c);
        r.wt = br < 64 ? w[br] : 0;
        const uint64_t t1 = (uint64_t)r.st[7] + big1(e) + r.chw + (br < 64 ? KH[br] : 0) + r.wt;
        const uint64_t t2 = (uint64_t)big0(a) + r.mjw;
        r.t1lo = (uint32_t)t1;
        r.t1c = (uint32_t)(t1 >> 32);
        r.t2lo = (uint32_t)t2;
        r.t2c = (uint32_t)(t2 >> 32);
        memcpy(r.hs, br <= 64 ? iv : dg, 32);
        for (int k = 0; k < 15; k++) r.ww[k] = wat(br - 15 + k);
        const uint32_t ww14 = r.ww[14], ww1 = r.ww[1];
        r.ss0t = rotr(ww1, 7) ^ rotr(ww1, 18);
        r.ss1t = rotr(ww14, 17) ^ rotr(ww14, 19);
        r.sc = (uint32_t)(((uint64_t)r.ww[0] + small0(ww1) + wat(br - 6) + small1(ww14)) >> 32);
        r.na = br < 64 ? (uint32_t)(((uint64_t)r.t1lo + r.t2lo) >> 32) : 0;
        r.ne = br < 64 ? (uint32_t)(((uint64_t)d + r.t1lo) >> 32) : 0;
    }
    std::array<uint32_t, 8> post;
    for (int k = 0; k < 8; k++) post[k] = st[64][k];
    return post;
}

// Device trace expansion: row r of proof k (grid.y) is built from row words
// words[off[k] + min(r, useful[k] - 1)] (rows past the last block repeat the
// last useful row); lanes hold consecutive rows, so every column store is
// coalesced.
__global__ __launch_bounds__(256) void expand_sha256_k(const RowWords *words, const uint64_t *off,
                                                       const uint32_t *useful, uint64_t h, uint64_t *tr) {
    const uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= h) return;
    const uint32_t k = blockIdx.y;
    const uint64_t src = off[k] + (r < useful[k] ? r : useful[k] - 1);
    const RowWords rw = words[src];
    uint64_t *o = tr + (uint64_t)k * SHA256_W * h + r;
    emit_row(rw, [&](uint32_t c, uint64_t v) { o[(uint64_t)c * h] = v; });
}

// ---- constraints (sha256_air.rs:956-1294), one LDE point per lane ----
// phi accumulator: slot w - 1 weighted by lam[w - 1], or by w = j + 1 without a table
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

// SP waves per group of 64 points: wave `part` evaluates the constraint
// sections s with s % SP == part for its 64 points (one per lane, so the
// wave does not diverge) and skips the slots of the others, so every wave's
// slot weights stay in step; the SP partial sums meet in LDS. One point's
// ~770 constraints are a long serial chain for one lane; split, a small
// proof's few hundred points still keep many CUs busy.
constexpr int SP = 8;
constexpr int AT = 64 * SP;

// Gate groups of the constraints: ungated, compression rows, the
// finalisation row, every other row, message-schedule rows.
enum { G_ALL, G_COMP, G_FIN, G_NOTFIN, G_SCHED, NG };

// grid.y = proof: the LDE of proof k is 756 columns of n at lde + 756 k n;
// n_blocks[k] its SHA-256 block count.
// GROUPS = false: phi at every point, gates from the point's trace row.
// GROUPS = true: the five gate groups' weighted sums as polynomials of the
// point, gates ignored and the row constant K[t] left out, written to
// phi[(k * NG + g) * n + i], on the n points of a blowup-2 LDE held as its
// even points (lde: the traces, n / 2 rows) and its odd points (odd, the
// same layout); t1_lam gets the weight of the T1 identity. lam: the 766
// slots' combination coefficients, or j + 1 when null.
template <bool GROUPS>
__global__ __launch_bounds__(AT) void compose_sha256_k(const uint64_t *lde, const uint64_t *odd, uint64_t n,
                                                       uint64_t blowup, const uint32_t *n_blocks, uint64_t *phi,
                                                       uint64_t *t1_lam, const uint64_t *lam) {
    __shared__ uint64_t red[5][SP][64];
    const uint32_t part = threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint64_t i = (uint64_t)blockIdx.x * 64 + lane;  // n is a multiple of 64
    auto mine = [&](int sec) { return sec % SP == (int)part; };
    // GROUPS: point i is row i / 2 of the traces (even i) or of the odd coset
    const uint64_t rows = GROUPS ? n / 2 : n, at = GROUPS ? i >> 1 : i;
    const uint64_t *L = (GROUPS && (i & 1) ? odd : lde) + (uint64_t)SHA256_W * rows * blockIdx.y;
    const uint64_t ni = GROUPS ? (at + 1) & (rows - 1) : (i + blowup) & (n - 1);
    auto cur = [&](uint32_t col) { return L[(uint64_t)col * rows + at]; };
    auto nxt = [&](uint32_t col) { return L[(uint64_t)col * rows + ni]; };
    const uint64_t row = i / blowup, blk = row >> 7, br = row & 127;
    const bool past = GROUPS || blk >= n_blocks[blockIdx.y];
    const bool comp = GROUPS || (!past && br < 64), fin = !past && br == 64;
    const bool sched = GROUPS || (!past && br >= 15 && br <= 62);
    const uint64_t TWO32 = 1ULL << 32;
    Phi f, fc, ff, fn, fs;  // ungated; compression, finalisation, not finalisation, schedule (GROUPS mode)
    f.lam = fc.lam = ff.lam = fn.lam = fs.lam = lam;
    Phi &gc = GROUPS ? fc : f, &gs = GROUPS ? fs : f;
    // booleanity of the 8 bit blocks
    const uint32_t boff[8] = {BIT_A, BIT_B, BIT_C, BIT_E, BIT_F, BIT_G, BIT_WW14, BIT_WW1};
#pragma unroll 1
    for (int o = 0; o < 8; o++) {  // sections 0-7
        if (!mine(o)) {
            f.skip(32);
            continue;
        }
#pragma unroll 4
        for (int b = 0; b < 32; b++) f.put(boolc(cur(boff[o] + b)));
    }
    auto pack = [&](uint32_t off) {
        uint64_t s = 0;
#pragma unroll 8
        for (int b = 0; b < 32; b++) s = add(s, mul(1ULL << b, cur(off + b)));
        return s;
    };
    const uint32_t wcol[8] = {A_A, A_A + 1, A_A + 2, A_E, A_E + 1, A_E + 2, WW0 + 14, WW0 + 1};
#pragma unroll 1
    for (int o = 0; o < 8; o++) {  // sections 8-15
        if (mine(8 + o))
            f.put(sub(cur(wcol[o]), pack(boff[o])));
        else
            f.skip(1);
    }
    // Sigma0(a), Sigma1(e), sigma0(W_win[1]), sigma1(W_win[14])
    struct Sg {
        uint32_t src, t, b, w;
        int r1, r2, r3;
        bool shr;
    };
    const Sg sg[4] = {{BIT_A, S0_T, S0_BIT, S0_W, 2, 13, 22, false},
                      {BIT_E, S1_T, S1_BIT, S1_W, 6, 11, 25, false},
                      {BIT_WW1, SS0_T, SS0_BIT, SS0_W, 7, 18, 3, true},
                      {BIT_WW14, SS1_T, SS1_BIT, SS1_W, 17, 19, 10, true}};
#pragma unroll 1
    for (int s = 0; s < 4; s++) {  // sections 16-23
        const Sg &g = sg[s];
        if (mine(16 + 2 * s)) {
#pragma unroll 2
            for (int b = 0; b < 32; b++)
                f.put(sub(cur(g.t + b), xorp(cur(g.src + (b + g.r1) % 32), cur(g.src + (b + g.r2) % 32))));
        } else {
            f.skip(32);
        }
        if (mine(17 + 2 * s)) {
#pragma unroll 2
            for (int b = 0; b < 32; b++) {
                const uint64_t third =
                    g.shr ? (b + g.r3 < 32 ? cur(g.src + b + g.r3) : 0) : cur(g.src + (b + g.r3) % 32);
                f.put(sub(cur(g.b + b), xorp(cur(g.t + b), third)));
            }
            f.put(sub(cur(g.w), pack(g.b)));
        } else {
            f.skip(33);
        }
    }
    // Ch(e, f, g)_i = e f + g - e g (section 24)
    if (mine(24)) {
#pragma unroll 2
        for (int b = 0; b < 32; b++) {
            const uint64_t e = cur(BIT_E + b), ff = cur(BIT_F + b), g = cur(BIT_G + b);
            f.put(sub(cur(CH_BIT + b), sub(add(mul(e, ff), g), mul(e, g))));
        }
        f.put(sub(cur(CH_W), pack(CH_BIT)));
    } else {
        f.skip(33);
    }
    // Maj(a, b, c) with the ab/ac/bc/u auxiliaries (sections 25-27)
    if (mine(25)) {
#pragma unroll 2
        for (int b = 0; b < 32; b++) {
            const uint64_t a = cur(BIT_A + b), bb = cur(BIT_B + b), c = cur(BIT_C + b);
            f.put(sub(cur(MAJ_AB + b), mul(a, bb)));
            f.put(sub(cur(MAJ_AC + b), mul(a, c)));
            f.put(sub(cur(MAJ_BC + b), mul(bb, c)));
        }
    } else {
        f.skip(96);
    }
    if (mine(26)) {
#pragma unroll 2
        for (int b = 0; b < 32; b++) f.put(sub(cur(MAJ_U + b), xorp(cur(MAJ_AB + b), cur(MAJ_AC + b))));
    } else {
        f.skip(32);
    }
    if (mine(27)) {
#pragma unroll 2
        for (int b = 0; b < 32; b++) f.put(sub(cur(MAJ_BIT + b), xorp(cur(MAJ_U + b), cur(MAJ_BC + b))));
        f.put(sub(cur(MAJ_W), pack(MAJ_BIT)));
    } else {
        f.skip(33);
    }
    // T1 = h + Sigma1 + Ch + K[t] + W[t] = T1_lo + 2^32 T1_cw; T2 = Sigma0 + Maj
    // (section 28)
    if (mine(28)) {
        const uint64_t krow = GROUPS ? 0 : comp ? KD[br] : 0;
        const uint64_t t1s = add(add(add(add(cur(A_H), cur(S1_W)), cur(CH_W)), krow), cur(COL_W));
        if (GROUPS && i == 0 && blockIdx.y == 0) *t1_lam = f.weight();
        f.put(sub(add(cur(T1_LO), mul(TWO32, cur(T1_CW))), t1s));
        f.put(boolc(cur(T1_C0)));
        f.put(boolc(cur(T1_C1)));
        f.put(boolc(cur(T1_C2)));
        f.put(sub(sub(sub(cur(T1_CW), cur(T1_C0)), mul(2, cur(T1_C1))), mul(4, cur(T1_C2))));
        f.put(sub(add(cur(T2_LO), mul(TWO32, cur(T2_CW))), add(cur(S0_W), cur(MAJ_W))));
        f.put(boolc(cur(T2_C0)));
        f.put(boolc(cur(T2_C1)));
        f.put(sub(sub(cur(T2_CW), cur(T2_C0)), mul(2, cur(T2_C1))));
    } else {
        f.skip(9);
    }
    // register shifts and new a / e (section 29)
    if (comp && mine(29)) {
        gc.w = f.w;
        // register shifts b<-a, c<-b, d<-c, f<-e, g<-f, h<-g
        gc.put(sub(nxt(A_A + 1), cur(A_A)));
        gc.put(sub(nxt(A_A + 2), cur(A_A + 1)));
        gc.put(sub(nxt(A_D), cur(A_A + 2)));
        gc.put(sub(nxt(A_E + 1), cur(A_E)));
        gc.put(sub(nxt(A_E + 2), cur(A_E + 1)));
        gc.put(sub(nxt(A_H), cur(A_E + 2)));
        // new_a, new_e with one carry bit each (two zero slots follow)
        gc.put(sub(sub(add(nxt(A_A), mul(TWO32, cur(NA_C))), cur(T1_LO)), cur(T2_LO)));
        gc.put(boolc(cur(NA_C)));
        gc.put(sub(sub(add(nxt(A_E), mul(TWO32, cur(NE_C))), cur(A_D)), cur(T1_LO)));
        gc.put(boolc(cur(NE_C)));
        gc.skip(2);
        f.w = gc.w;
    } else {
        f.skip(12);
    }
    // H-state: add-with-carry at finalisation, hold otherwise (GROUPS: both
    // forms at the same slot, to the finalisation and "not finalisation" sums)
    // (section 30)
#pragma unroll 1
    for (int k = 0; k < 8; k++) {
        if (!mine(30)) {
            f.skip(1);
        } else if (GROUPS) {
            const uint64_t dl = sub(add(cur(H0 + k), cur(A_A + k)), nxt(H0 + k));
            ff.w = fn.w = f.w;
            ff.put(mul(dl, sub(dl, TWO32)));
            fn.put(sub(nxt(H0 + k), cur(H0 + k)));
            f.skip(1);
        } else if (fin) {
            const uint64_t dl = sub(add(cur(H0 + k), cur(A_A + k)), nxt(H0 + k));
            f.put(mul(dl, sub(dl, TWO32)));
        } else {
            f.put(sub(nxt(H0 + k), cur(H0 + k)));
        }
    }
    // W window shift, then W_win[14]' = W (section 31)
    if (mine(31)) {
#pragma unroll 2
        for (int k = 0; k < 14; k++) f.put(sub(nxt(WW0 + k), cur(WW0 + k + 1)));
        f.put(sub(nxt(WW0 + 14), cur(COL_W)));
    } else {
        f.skip(15);
    }
    // message schedule (section 32, the last: no skip needed)
    if (sched && mine(32)) {
        gs.w = f.w;
        const uint64_t sum = add(add(add(cur(WW0), cur(SS0_W)), cur(WW0 + 9)), cur(SS1_W));
        gs.put(sub(add(nxt(COL_W), mul(TWO32, cur(WS_CW))), sum));
        gs.put(boolc(cur(WS_C0)));
        gs.put(boolc(cur(WS_C1)));
        gs.put(sub(sub(cur(WS_CW), cur(WS_C0)), mul(2, cur(WS_C1))));
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

// phi at every point from the interpolated gate-group polynomials
// pg[(k * NG + g) * n + i]: the groups whose gate holds at the point's trace
// row, less lambda_T1 K[t] on compression rows (the T1 identity's constant)
__global__ __launch_bounds__(256) void combine_sha256_k(const uint64_t *pg, uint64_t n, uint64_t blowup,
                                                        const uint32_t *n_blocks, const uint64_t *t1_lam,
                                                        uint64_t *phi) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint64_t row = i / blowup, blk = row >> 7, br = row & 127;
    const bool past = blk >= n_blocks[blockIdx.y];
    const bool comp = !past && br < 64, fin = !past && br == 64, sched = !past && br >= 15 && br <= 62;
    const uint64_t *q = pg + (uint64_t)blockIdx.y * NG * n + i;
    uint64_t v = add(q[G_ALL * n], fin ? q[G_FIN * n] : q[G_NOTFIN * n]);
    if (comp) v = sub(add(v, q[G_COMP * n]), mul(*t1_lam, KD[br]));
    if (sched) v = add(v, q[G_SCHED * n]);
    phi[n * blockIdx.y + i] = v;
}

}  // namespace sha

uint64_t sha256_n_blocks(uint64_t len) { return (len + 9 + 63) / 64; }

uint64_t sha256_trace_height(uint64_t len) {
    const uint64_t useful = 128 * sha256_n_blocks(len);
    uint64_t h = 128;
    while (h < useful) h <<= 1;
    return h;
}

// build_sha256_trace_multi (sha256_air.rs:392-479): FIPS 180-4 padding,
// one 128-row segment per block with the IV chained, rows past the last
// block copy the last useful row. Also returns the digest as
// prove_ds_ksk_binding reads it (prover.rs:529-538): H-state at block row 65
// of the last block, big-endian words.
// row words of the padded message, block by block with the IV chained
// (build_sha256_trace_multi, sha256_air.rs:429-479); returns n_blocks and the
// digest as prove_ds_ksk_binding reads it (prover.rs:529-538): H-state at
// block row 65 of the last block, big-endian words.
static uint64_t sha256_words(const uint8_t *msg, uint64_t len, std::vector<sha::RowWords> &rows, uint8_t digest[32]) {
    const uint64_t nb = sha256_n_blocks(len);
    std::vector<uint8_t> m(64 * nb, 0);
    if (len) memcpy(m.data(), msg, len);
    m[len] = 0x80;
    const uint64_t bits = len * 8;
    for (int i = 0; i < 8; i++) m[64 * nb - 1 - i] = (uint8_t)(bits >> (8 * i));
    rows.resize(128 * nb);
    uint32_t iv[8];
    memcpy(iv, sha::IV, 32);
    for (uint64_t b = 0; b < nb; b++) {
        const std::array<uint32_t, 8> post = sha::block_words(&m[64 * b], iv, &rows[128 * b]);
        for (int k = 0; k < 8; k++) iv[k] += post[k];
    }
    const sha::RowWords &d = rows[128 * (nb - 1) + 65];
    for (int k = 0; k < 8; k++)
        for (int i = 0; i < 4; i++) digest[4 * k + i] = (uint8_t)(d.hs[k] >> (24 - 8 * i));
    return nb;
}

uint64_t build_sha256_trace(const uint8_t *msg, uint64_t len, uint64_t *tr, uint8_t digest[32], uint64_t height) {
    std::vector<sha::RowWords> rows;
    const uint64_t nb = sha256_words(msg, len, rows, digest), h = height ? height : sha256_trace_height(len);
    SG_REQUIRE(is_pow2(h) && h >= 128 * nb, "sha256 trace: height must be a power of two >= 128 * n_blocks");
    for (uint64_t r = 0; r < h; r++) {
        const sha::RowWords &w = rows[r < rows.size() ? r : rows.size() - 1];
        sha::emit_row(w, [&](uint32_t c, uint64_t v) { tr[(uint64_t)c * h + r] = v; });
    }
    return nb;
}

void build_sha256_traces_dev(Ctx *c, const std::vector<ByteSpan> &msgs, uint64_t h, uint64_t *d_tr,
                             std::vector<uint32_t> &n_blocks, std::vector<std::array<uint8_t, 64>> &digests) {
    const size_t K = msgs.size();
    std::vector<std::vector<sha::RowWords>> rows(K);
    n_blocks.assign(K, 0);
    digests.assign(K, {});
    parallel_for(K, [&](size_t k) {
        n_blocks[k] = (uint32_t)sha256_words(msgs[k].p, msgs[k].n, rows[k], digests[k].data());
    });
    std::vector<uint64_t> off(K);
    std::vector<uint32_t> useful(K);
    uint64_t total = 0;
    for (size_t k = 0; k < K; k++) {
        SG_REQUIRE(128ULL * n_blocks[k] <= h, "sha256 traces: height below a message's minimum");
        off[k] = total;
        useful[k] = (uint32_t)rows[k].size();
        total += rows[k].size();
    }
    // the row words of every proof copied on the host's threads straight into
    // pinned memory (one DMA upload, no zero-filled staging vector)
    sha::RowWords *all = (sha::RowWords *)c->upload_staging(sizeof(sha::RowWords) * total);
    parallel_for(K, [&](size_t k) { std::copy(rows[k].begin(), rows[k].end(), all + off[k]); });
    Scoped d_w(c, sizeof(sha::RowWords) * total), d_off(c, 8 * K), d_u(c, 4 * K);
    SG_HIP(hipMemcpyAsync(d_w.p, all, sizeof(sha::RowWords) * total, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(d_off.p, off.data(), 8 * K, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(d_u.p, useful.data(), 4 * K, hipMemcpyHostToDevice, c->stream));
    ProfScope ps(c, "expand_sha256", 8.0 * SHA256_W * h * K);
    hipLaunchKernelGGL(sha::expand_sha256_k, dim3((uint32_t)((h + 255) / 256), (uint32_t)K), dim3(256), 0, c->stream,
                       d_w.as<sha::RowWords>(), d_off.as<uint64_t>(), d_u.as<uint32_t>(), h, d_tr);
    SG_HIP(hipGetLastError());
    SG_HIP(hipStreamSynchronize(c->stream));  // the host word buffers go out of scope
}

void merge_sha256_batch(Ctx *c, const uint64_t *lde, uint32_t K, uint64_t n_trace, uint64_t blowup,
                        const std::vector<uint32_t> &n_blocks, int phi_mode, uint64_t *c_eval, const uint64_t *d_lam) {
    SG_REQUIRE(is_pow2(n_trace) && n_trace >= 128 && is_pow2(blowup) && blowup >= 2,
               "merge_sha256: n_trace must be a power of two >= 128");
    SG_REQUIRE(K >= 1 && K <= 65535 && n_blocks.size() == K, "merge_sha256: one n_blocks per proof");
    const uint64_t n = n_trace * blowup;
    SG_REQUIRE(n <= (1ULL << 24), "merge_sha256: domain above 2^24");
    for (uint32_t nb : n_blocks) SG_REQUIRE(nb >= 1 && 128 * (uint64_t)nb <= n_trace, "merge_sha256: bad n_blocks");
    if (!phi_computed(phi_mode, n)) {
        fill_u64(c, c_eval, n * K, 0);
        return;
    }
    Scoped phi(c, 8 * n * K), d_nb(c, 4 * K);
    SG_HIP(hipMemcpyAsync(d_nb.p, n_blocks.data(), 4 * K, hipMemcpyHostToDevice, c->stream));
    {
        ProfScope ps(c, "compose_sha256", (8.0 * SHA256_W * n + 8.0 * n) * K);
        hipLaunchKernelGGL(sha::compose_sha256_k<false>, dim3((uint32_t)(n / 64), K), dim3(sha::AT), 0,
                           c->stream, lde, (const uint64_t *)nullptr, n, blowup, d_nb.as<uint32_t>(),
                           phi.as<uint64_t>(), (uint64_t *)nullptr, d_lam);
        SG_HIP(hipGetLastError());
    }
    quotient_batch(c, phi.as<uint64_t>(), K, n_trace, n, c_eval);
}

void merge_sha256_from_trace_batch(Ctx *c, const uint64_t *d_tr, uint32_t K, uint64_t n_trace, uint64_t blowup,
                                   const std::vector<uint32_t> &n_blocks, int phi_mode, uint64_t *c_eval,
                                   const uint64_t *d_lam) {
    SG_REQUIRE(is_pow2(n_trace) && n_trace >= 128 && is_pow2(blowup) && blowup >= 4,
               "merge_sha256: n_trace must be a power of two >= 128, blowup >= 4");
    SG_REQUIRE(K >= 1 && K <= 65535 && n_blocks.size() == K, "merge_sha256: one n_blocks per proof");
    const uint64_t n = n_trace * blowup, n2 = 2 * n_trace;
    SG_REQUIRE(n <= (1ULL << 24), "merge_sha256: domain above 2^24");
    for (uint32_t nb : n_blocks) SG_REQUIRE(nb >= 1 && 128 * (uint64_t)nb <= n_trace, "merge_sha256: bad n_blocks");
    if (!phi_computed(phi_mode, n)) {
        fill_u64(c, c_eval, n * K, 0);
        return;
    }
    Scoped phi(c, 8 * n * K), d_nb(c, 4 * K), t1(c, 8);
    SG_HIP(hipMemcpyAsync(d_nb.p, n_blocks.data(), 4 * K, hipMemcpyHostToDevice, c->stream));
    {
        Scoped groups(c, 8ULL * sha::NG * n2 * K);
        {
            // the blowup-2 LDE as the traces (its even points) and the odd coset
            Scoped odd(c, 8ULL * SHA256_W * n_trace * K);
            lde_odd_coset(c, d_tr, odd.as<uint64_t>(), ilog2(n_trace), (uint64_t)SHA256_W * K);
            ProfScope ps(c, "compose_sha256", (8.0 * SHA256_W * n2 + 8.0 * sha::NG * n2) * K);
            hipLaunchKernelGGL(sha::compose_sha256_k<true>, dim3((uint32_t)(n2 / 64), K),
                               dim3(sha::AT), 0, c->stream, d_tr, odd.as<uint64_t>(), n2, (uint64_t)2,
                               d_nb.as<uint32_t>(), groups.as<uint64_t>(), t1.as<uint64_t>(), d_lam);
            SG_HIP(hipGetLastError());
        }
        Scoped full(c, 8ULL * sha::NG * n * K);
        lde_columns(c, groups.as<uint64_t>(), full.as<uint64_t>(), ilog2(n2), ilog2(blowup / 2),
                    (uint64_t)sha::NG * K, n2, n, 1);
        ProfScope ps(c, "combine_sha256", 8.0 * (sha::NG + 1) * n * K);
        hipLaunchKernelGGL(sha::combine_sha256_k, dim3((uint32_t)((n + 255) / 256), K), dim3(256), 0, c->stream,
                           full.as<uint64_t>(), n, blowup, d_nb.as<uint32_t>(), t1.as<uint64_t>(), phi.as<uint64_t>());
        SG_HIP(hipGetLastError());
    }
    quotient_batch(c, phi.as<uint64_t>(), K, n_trace, n, c_eval);
}

}  // namespace sg
