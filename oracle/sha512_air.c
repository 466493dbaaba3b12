/* sha512_air.c -- CPU restatement of the SHA-512 AIR (test infrastructure:
 * only tests/ load it).
 *
 *   column layout          crates/deep_ali/src/sha512_air.rs:313-481
 *   trace builder          sha512_air.rs:218-236 (padding), :551-573 (multi-block), :585-898 (fill_block)
 *   constraint evaluator   sha512_air.rs:909-1282 (1526 constraints, this order)
 *   merge                  crates/deep_ali/src/lib.rs:561-653 (deep_ali_merge_sha512)
 *
 * Every 64-bit word is two 32-bit limb cells (lo, hi); bits are 0/1. */
#include <stdlib.h>
#include <string.h>

#include "refcpu.h"
#include "refcpu_int.h"

enum {
    A0 = 0, /* word x at (A0 + 2x, A0 + 2x + 1), x = a..h */
    BIT_A = 16, BIT_B = 80, BIT_C = 144, BIT_E = 208, BIT_F = 272, BIT_G = 336,
    S0_T = 400, S0_BIT = 464, S0_LO = 528, S0_HI = 529,
    S1_T = 530, S1_BIT = 594, S1_LO = 658, S1_HI = 659,
    CH_BIT = 660, CH_LO = 724, CH_HI = 725,
    MAJ_AB = 726, MAJ_AC = 790, MAJ_BC = 854, MAJ_U = 918, MAJ_BIT = 982, MAJ_LO = 1046, MAJ_HI = 1047,
    W_LO = 1048, W_HI = 1049,
    T1L_LO = 1050, T1L_C0, T1L_C1, T1L_C2, T1L_CW,
    T1H_LO = 1055, T1H_C0, T1H_C1, T1H_C2, T1H_CW,
    T2L_LO = 1060, T2L_C0, T2L_CW, T2H_LO, T2H_C0, T2H_CW,
    H0 = 1066, WW0 = 1082,
    BIT_WW14 = 1112, BIT_WW1 = 1176,
    SS0_T = 1240, SS0_BIT = 1304, SS0_LO = 1368, SS0_HI = 1369,
    SS1_T = 1370, SS1_BIT = 1434, SS1_LO = 1498, SS1_HI = 1499,
    WSL_C0 = 1500, WSL_C1, WSL_CW, WSH_C0, WSH_C1, WSH_CW,
    NA_LO = 1506, NA_HI, NE_LO, NE_HI,
    W512 = 1510, K512 = 1526, RPB = 128, ROUNDS = 80
};

static const uint64_t KC[80] = { /* FIPS 180-4 section 4.2.3 */
0x428a2f98d728ae22ULL, 0x7137449123ef65cdULL, 0xb5c0fbcfec4d3b2fULL, 0xe9b5dba58189dbbcULL,
0x3956c25bf348b538ULL, 0x59f111f1b605d019ULL, 0x923f82a4af194f9bULL, 0xab1c5ed5da6d8118ULL,
0xd807aa98a3030242ULL, 0x12835b0145706fbeULL, 0x243185be4ee4b28cULL, 0x550c7dc3d5ffb4e2ULL,
0x72be5d74f27b896fULL, 0x80deb1fe3b1696b1ULL, 0x9bdc06a725c71235ULL, 0xc19bf174cf692694ULL,
0xe49b69c19ef14ad2ULL, 0xefbe4786384f25e3ULL, 0x0fc19dc68b8cd5b5ULL, 0x240ca1cc77ac9c65ULL,
0x2de92c6f592b0275ULL, 0x4a7484aa6ea6e483ULL, 0x5cb0a9dcbd41fbd4ULL, 0x76f988da831153b5ULL,
0x983e5152ee66dfabULL, 0xa831c66d2db43210ULL, 0xb00327c898fb213fULL, 0xbf597fc7beef0ee4ULL,
0xc6e00bf33da88fc2ULL, 0xd5a79147930aa725ULL, 0x06ca6351e003826fULL, 0x142929670a0e6e70ULL,
0x27b70a8546d22ffcULL, 0x2e1b21385c26c926ULL, 0x4d2c6dfc5ac42aedULL, 0x53380d139d95b3dfULL,
0x650a73548baf63deULL, 0x766a0abb3c77b2a8ULL, 0x81c2c92e47edaee6ULL, 0x92722c851482353bULL,
0xa2bfe8a14cf10364ULL, 0xa81a664bbc423001ULL, 0xc24b8b70d0f89791ULL, 0xc76c51a30654be30ULL,
0xd192e819d6ef5218ULL, 0xd69906245565a910ULL, 0xf40e35855771202aULL, 0x106aa07032bbd1b8ULL,
0x19a4c116b8d2d0c8ULL, 0x1e376c085141ab53ULL, 0x2748774cdf8eeb99ULL, 0x34b0bcb5e19b48a8ULL,
0x391c0cb3c5c95a63ULL, 0x4ed8aa4ae3418acbULL, 0x5b9cca4f7763e373ULL, 0x682e6ff3d6b2b8a3ULL,
0x748f82ee5defb2fcULL, 0x78a5636f43172f60ULL, 0x84c87814a1f0ab72ULL, 0x8cc702081a6439ecULL,
0x90befffa23631e28ULL, 0xa4506cebde82bde9ULL, 0xbef9a3f7b2c67915ULL, 0xc67178f2e372532bULL,
0xca273eceea26619cULL, 0xd186b8c721c0c207ULL, 0xeada7dd6cde0eb1eULL, 0xf57d4f7fee6ed178ULL,
0x06f067aa72176fbaULL, 0x0a637dc5a2c898a6ULL, 0x113f9804bef90daeULL, 0x1b710b35131c471bULL,
0x28db77f523047d84ULL, 0x32caab7b40c72493ULL, 0x3c9ebe0a15c9bebcULL, 0x431d67c49c100d4cULL,
0x4cc5d4becb3e42b6ULL, 0x597f299cfc657e2aULL, 0x5fcb6fab3ad6faecULL, 0x6c44198c4a475817ULL};
static const uint64_t IV0[8] = {0x6a09e667f3bcc908ULL, 0xbb67ae8584caa73bULL, 0x3c6ef372fe94f82bULL, 0xa54ff53a5f1d36f1ULL, 0x510e527fade682d1ULL, 0x9b05688c2b3e6c1fULL, 0x1f83d9abfb41bd6bULL, 0x5be0cd19137e2179ULL}; /* section 5.3.5 */

static uint64_t rotr(uint64_t x, int n) { return (x >> n) | (x << (64 - n)); }
static uint64_t bsig0(uint64_t x) { return rotr(x, 28) ^ rotr(x, 34) ^ rotr(x, 39); }
static uint64_t bsig1(uint64_t x) { return rotr(x, 14) ^ rotr(x, 18) ^ rotr(x, 41); }
static uint64_t ssig0(uint64_t x) { return rotr(x, 1) ^ rotr(x, 8) ^ (x >> 7); }
static uint64_t ssig1(uint64_t x) { return rotr(x, 19) ^ rotr(x, 61) ^ (x >> 6); }
static uint64_t chf(uint64_t e, uint64_t f, uint64_t g) { return (e & f) ^ (~e & g); }
static uint64_t majf(uint64_t a, uint64_t b, uint64_t c) { return (a & b) ^ (a & c) ^ (b & c); }
static uint64_t lo32(uint64_t x) { return x & 0xFFFFFFFFu; }
static uint64_t hi32(uint64_t x) { return x >> 32; }

size_t ref_sha512_n_blocks(size_t len) { return (len + 17 + 127) / 128; }
size_t ref_sha512_trace_height(size_t len) {
    size_t useful = RPB * ref_sha512_n_blocks(len), h = RPB;
    while (h < useful) h <<= 1;
    return h;
}

static void fill_block(uint64_t *tr, size_t h, size_t base, const uint8_t *blk, const uint64_t iv[8],
                       uint64_t post_out[8]) {
#define T(col, row) tr[(size_t)(col) * h + (row)]
#define PUTW(col, row, x) (T((col), (row)) = lo32(x), T((col) + 1, (row)) = hi32(x))
    uint64_t w[80], st[81][8];
    for (int i = 0; i < 16; i++) {
        uint64_t v = 0;
        for (int b = 0; b < 8; b++) v = v << 8 | blk[8 * i + b];
        w[i] = v;
    }
    for (int t = 16; t < 80; t++) w[t] = ssig1(w[t - 2]) + w[t - 7] + ssig0(w[t - 15]) + w[t - 16];
    memcpy(st[0], iv, 64);
    for (int t = 0; t < 80; t++) {
        const uint64_t *s = st[t];
        uint64_t t1 = s[7] + bsig1(s[4]) + chf(s[4], s[5], s[6]) + KC[t] + w[t];
        uint64_t t2 = bsig0(s[0]) + majf(s[0], s[1], s[2]);
        uint64_t n[8] = {t1 + t2, s[0], s[1], s[2], s[3] + t1, s[4], s[5], s[6]};
        memcpy(st[t + 1], n, 64);
    }
    uint64_t dg[8];
    for (int k = 0; k < 8; k++) dg[k] = iv[k] + st[80][k];
    if (post_out) memcpy(post_out, st[80], 64);
#define WAT(i) (((i) >= 0 && (i) < 80) ? w[(i)] : 0ull)
    for (int br = 0; br < RPB; br++) {
        const size_t row = base + br;
        const uint64_t *s = st[br <= 80 ? br : 80];
        const uint64_t a = s[0], b = s[1], c = s[2], d = s[3], e = s[4], f = s[5], g = s[6], hh = s[7];
        for (int k = 0; k < 8; k++) PUTW(A0 + 2 * k, row, s[k]);
        for (int i = 0; i < 64; i++) {
            T(BIT_A + i, row) = (a >> i) & 1;
            T(BIT_B + i, row) = (b >> i) & 1;
            T(BIT_C + i, row) = (c >> i) & 1;
            T(BIT_E + i, row) = (e >> i) & 1;
            T(BIT_F + i, row) = (f >> i) & 1;
            T(BIT_G + i, row) = (g >> i) & 1;
            uint64_t t0 = ((a >> ((i + 28) % 64)) ^ (a >> ((i + 34) % 64))) & 1;
            T(S0_T + i, row) = t0;
            T(S0_BIT + i, row) = t0 ^ ((a >> ((i + 39) % 64)) & 1);
            uint64_t t1 = ((e >> ((i + 14) % 64)) ^ (e >> ((i + 18) % 64))) & 1;
            T(S1_T + i, row) = t1;
            T(S1_BIT + i, row) = t1 ^ ((e >> ((i + 41) % 64)) & 1);
        }
        const uint64_t s0 = bsig0(a), s1 = bsig1(e), chw = chf(e, f, g), mjw = majf(a, b, c);
        PUTW(S0_LO, row, s0);
        PUTW(S1_LO, row, s1);
        for (int i = 0; i < 64; i++) {
            T(CH_BIT + i, row) = (chw >> i) & 1;
            uint64_t ab = (a >> i) & (b >> i) & 1, ac = (a >> i) & (c >> i) & 1, bc = (b >> i) & (c >> i) & 1;
            T(MAJ_AB + i, row) = ab;
            T(MAJ_AC + i, row) = ac;
            T(MAJ_BC + i, row) = bc;
            T(MAJ_U + i, row) = ab ^ ac;
            T(MAJ_BIT + i, row) = ab ^ ac ^ bc;
        }
        PUTW(CH_LO, row, chw);
        PUTW(MAJ_LO, row, mjw);
        const uint64_t wt = br < 80 ? w[br] : 0, kt = br < 80 ? KC[br] : 0;
        PUTW(W_LO, row, wt);
        /* T1, T2 per limb with the lo carry into hi */
        uint64_t t1l = lo32(hh) + lo32(s1) + lo32(chw) + lo32(kt) + lo32(wt), t1lc = t1l >> 32;
        uint64_t t1h = hi32(hh) + hi32(s1) + hi32(chw) + hi32(kt) + hi32(wt) + t1lc, t1hc = t1h >> 32;
        T(T1L_LO, row) = lo32(t1l);
        T(T1L_C0, row) = t1lc & 1;
        T(T1L_C1, row) = (t1lc >> 1) & 1;
        T(T1L_C2, row) = (t1lc >> 2) & 1;
        T(T1L_CW, row) = t1lc;
        T(T1H_LO, row) = lo32(t1h);
        T(T1H_C0, row) = t1hc & 1;
        T(T1H_C1, row) = (t1hc >> 1) & 1;
        T(T1H_C2, row) = (t1hc >> 2) & 1;
        T(T1H_CW, row) = t1hc;
        uint64_t t2l = lo32(s0) + lo32(mjw), t2lc = t2l >> 32;
        uint64_t t2h = hi32(s0) + hi32(mjw) + t2lc, t2hc = t2h >> 32;
        T(T2L_LO, row) = lo32(t2l);
        T(T2L_C0, row) = t2lc & 1;
        T(T2L_CW, row) = t2lc;
        T(T2H_LO, row) = lo32(t2h);
        T(T2H_C0, row) = t2hc & 1;
        T(T2H_CW, row) = t2hc;
        for (int k = 0; k < 8; k++) PUTW(H0 + 2 * k, row, br <= 80 ? iv[k] : dg[k]);
        for (int k = 0; k < 15; k++) PUTW(WW0 + 2 * k, row, WAT(br - 15 + k));
        const uint64_t ww14 = WAT(br - 1), ww1 = WAT(br - 14);
        for (int i = 0; i < 64; i++) {
            T(BIT_WW14 + i, row) = (ww14 >> i) & 1;
            T(BIT_WW1 + i, row) = (ww1 >> i) & 1;
            uint64_t p = ((ww1 >> ((i + 1) % 64)) ^ (ww1 >> ((i + 8) % 64))) & 1;
            T(SS0_T + i, row) = p;
            T(SS0_BIT + i, row) = p ^ (i + 7 < 64 ? (ww1 >> (i + 7)) & 1 : 0);
            uint64_t q = ((ww14 >> ((i + 19) % 64)) ^ (ww14 >> ((i + 61) % 64))) & 1;
            T(SS1_T + i, row) = q;
            T(SS1_BIT + i, row) = q ^ (i + 6 < 64 ? (ww14 >> (i + 6)) & 1 : 0);
        }
        const uint64_t ss0 = ssig0(ww1), ss1 = ssig1(ww14);
        PUTW(SS0_LO, row, ss0);
        PUTW(SS1_LO, row, ss1);
        const uint64_t w0 = WAT(br - 15), w9 = WAT(br - 6);
        uint64_t sl = lo32(w0) + lo32(ss0) + lo32(w9) + lo32(ss1), slc = sl >> 32;
        uint64_t sh = hi32(w0) + hi32(ss0) + hi32(w9) + hi32(ss1) + slc, shc = sh >> 32;
        T(WSL_C0, row) = slc & 1;
        T(WSL_C1, row) = (slc >> 1) & 1;
        T(WSL_CW, row) = slc;
        T(WSH_C0, row) = shc & 1;
        T(WSH_C1, row) = (shc >> 1) & 1;
        T(WSH_CW, row) = shc;
        if (br < 80) {
            uint64_t nal = lo32(t1l) + lo32(t2l), nalc = nal >> 32;
            uint64_t nah = lo32(t1h) + lo32(t2h) + nalc, nahc = nah >> 32;
            uint64_t nel = lo32(d) + lo32(t1l), nelc = nel >> 32;
            uint64_t neh = hi32(d) + lo32(t1h) + nelc, nehc = neh >> 32;
            T(NA_LO, row) = nalc;
            T(NA_HI, row) = nahc;
            T(NE_LO, row) = nelc;
            T(NE_HI, row) = nehc;
        } else {
            T(NA_LO, row) = T(NA_HI, row) = T(NE_LO, row) = T(NE_HI, row) = 0;
        }
    }
#undef WAT
#undef PUTW
#undef T
}

int ref_build_sha512_trace(const uint8_t *msg, size_t len, uint64_t *tr, uint8_t digest[64]) {
    const size_t nb = ref_sha512_n_blocks(len), h = ref_sha512_trace_height(len);
    uint8_t *m = (uint8_t *)calloc(128 * nb, 1);
    memcpy(m, msg, len);
    m[len] = 0x80;
    uint64_t bits = (uint64_t)len * 8; /* the upper 64 bits of the 128-bit length are zero */
    for (int i = 0; i < 8; i++) m[128 * nb - 1 - i] = (uint8_t)(bits >> (8 * i));
    memset(tr, 0, 8 * W512 * h);
    uint64_t iv[8], post[8];
    memcpy(iv, IV0, 64);
    for (size_t b = 0; b < nb; b++) {
        fill_block(tr, h, b * RPB, m + 128 * b, iv, post);
        for (int k = 0; k < 8; k++) iv[k] += post[k];
    }
    const size_t useful = RPB * nb;
    for (size_t r = useful; r < h; r++)
        for (size_t c = 0; c < W512; c++) tr[c * h + r] = tr[c * h + useful - 1];
    /* digest from the H-state at block row 81 of the last block (read_digest_from_trace, :1285-1297) */
    const size_t drow = RPB * (nb - 1) + 81;
    for (int k = 0; k < 8; k++) {
        uint64_t v = tr[(H0 + 2 * k) * h + drow] | tr[(H0 + 2 * k + 1) * h + drow] << 32;
        for (int i = 0; i < 8; i++) digest[8 * k + i] = (uint8_t)(v >> (56 - 8 * i));
    }
    free(m);
    return (int)nb;
}

static uint64_t xorp(uint64_t x, uint64_t y) { return gl_sub(gl_add(x, y), gl_mul(2, gl_mul(x, y))); }
static uint64_t boolc(uint64_t b) { return gl_mul(b, gl_sub(1, b)); }

void ref_sha512_constraints(const uint64_t *cur, const uint64_t *nxt, size_t row, size_t n_blocks, uint64_t *out) {
    size_t j = 0;
#define PUT(v) (out[j++] = (v))
    const size_t blk = row / RPB, br = row % RPB;
    const int past = blk >= n_blocks;
    const int comp = !past && br < 80, fin = !past && br == 80, sched = !past && br >= 15 && br <= 78;
    const uint64_t two32 = (uint64_t)1 << 32;
    static const int boff[8] = {BIT_A, BIT_B, BIT_C, BIT_E, BIT_F, BIT_G, BIT_WW14, BIT_WW1};
    for (int o = 0; o < 8; o++)
        for (int i = 0; i < 64; i++) PUT(boolc(cur[boff[o] + i]));
#define PACK(off, half) ({ uint64_t s_ = 0; for (int i_ = 0; i_ < 32; i_++) s_ = gl_add(s_, gl_mul((uint64_t)1 << i_, cur[(off) + 32 * (half) + i_])); s_; })
    /* a, b, c, e, f, g limbs, then W_win[14] and W_win[1] */
    static const int wlo[8] = {A0 + 0, A0 + 2, A0 + 4, A0 + 8, A0 + 10, A0 + 12, WW0 + 28, WW0 + 2};
    for (int o = 0; o < 8; o++) {
        PUT(gl_sub(cur[wlo[o]], PACK(boff[o], 0)));
        PUT(gl_sub(cur[wlo[o] + 1], PACK(boff[o], 1)));
    }
    struct { int src, t, b, lo, r1, r2, r3, shr; } sg[4] = {
        {BIT_A, S0_T, S0_BIT, S0_LO, 28, 34, 39, 0},
        {BIT_E, S1_T, S1_BIT, S1_LO, 14, 18, 41, 0},
        {BIT_WW1, SS0_T, SS0_BIT, SS0_LO, 1, 8, 7, 1},
        {BIT_WW14, SS1_T, SS1_BIT, SS1_LO, 19, 61, 6, 1}};
    for (int s = 0; s < 4; s++) {
        for (int i = 0; i < 64; i++)
            PUT(gl_sub(cur[sg[s].t + i], xorp(cur[sg[s].src + (i + sg[s].r1) % 64], cur[sg[s].src + (i + sg[s].r2) % 64])));
        for (int i = 0; i < 64; i++) {
            uint64_t r = sg[s].shr ? (i + sg[s].r3 < 64 ? cur[sg[s].src + i + sg[s].r3] : 0)
                                   : cur[sg[s].src + (i + sg[s].r3) % 64];
            PUT(gl_sub(cur[sg[s].b + i], xorp(cur[sg[s].t + i], r)));
        }
        PUT(gl_sub(cur[sg[s].lo], PACK(sg[s].b, 0)));
        PUT(gl_sub(cur[sg[s].lo + 1], PACK(sg[s].b, 1)));
    }
    for (int i = 0; i < 64; i++) {
        uint64_t e = cur[BIT_E + i], f = cur[BIT_F + i], g = cur[BIT_G + i];
        PUT(gl_sub(cur[CH_BIT + i], gl_sub(gl_add(gl_mul(e, f), g), gl_mul(e, g))));
    }
    PUT(gl_sub(cur[CH_LO], PACK(CH_BIT, 0)));
    PUT(gl_sub(cur[CH_HI], PACK(CH_BIT, 1)));
    for (int i = 0; i < 64; i++) {
        uint64_t a = cur[BIT_A + i], b = cur[BIT_B + i], c = cur[BIT_C + i];
        PUT(gl_sub(cur[MAJ_AB + i], gl_mul(a, b)));
        PUT(gl_sub(cur[MAJ_AC + i], gl_mul(a, c)));
        PUT(gl_sub(cur[MAJ_BC + i], gl_mul(b, c)));
    }
    for (int i = 0; i < 64; i++) PUT(gl_sub(cur[MAJ_U + i], xorp(cur[MAJ_AB + i], cur[MAJ_AC + i])));
    for (int i = 0; i < 64; i++) PUT(gl_sub(cur[MAJ_BIT + i], xorp(cur[MAJ_U + i], cur[MAJ_BC + i])));
    PUT(gl_sub(cur[MAJ_LO], PACK(MAJ_BIT, 0)));
    PUT(gl_sub(cur[MAJ_HI], PACK(MAJ_BIT, 1)));
#undef PACK
    const uint64_t kt = comp ? KC[br] : 0;
    /* T1 lo */
    uint64_t s = gl_add(gl_add(gl_add(gl_add(cur[A0 + 14], cur[S1_LO]), cur[CH_LO]), lo32(kt)), cur[W_LO]);
    PUT(gl_sub(gl_add(cur[T1L_LO], gl_mul(two32, cur[T1L_CW])), s));
    PUT(boolc(cur[T1L_C0]));
    PUT(boolc(cur[T1L_C1]));
    PUT(boolc(cur[T1L_C2]));
    PUT(gl_sub(gl_sub(gl_sub(cur[T1L_CW], cur[T1L_C0]), gl_mul(2, cur[T1L_C1])), gl_mul(4, cur[T1L_C2])));
    /* T1 hi */
    s = gl_add(gl_add(gl_add(gl_add(gl_add(cur[A0 + 15], cur[S1_HI]), cur[CH_HI]), hi32(kt)), cur[W_HI]), cur[T1L_CW]);
    PUT(gl_sub(gl_add(cur[T1H_LO], gl_mul(two32, cur[T1H_CW])), s));
    PUT(boolc(cur[T1H_C0]));
    PUT(boolc(cur[T1H_C1]));
    PUT(boolc(cur[T1H_C2]));
    PUT(gl_sub(gl_sub(gl_sub(cur[T1H_CW], cur[T1H_C0]), gl_mul(2, cur[T1H_C1])), gl_mul(4, cur[T1H_C2])));
    /* T2 lo, hi */
    PUT(gl_sub(gl_add(cur[T2L_LO], gl_mul(two32, cur[T2L_CW])), gl_add(cur[S0_LO], cur[MAJ_LO])));
    PUT(boolc(cur[T2L_C0]));
    PUT(gl_sub(cur[T2L_CW], cur[T2L_C0]));
    PUT(gl_sub(gl_add(cur[T2H_LO], gl_mul(two32, cur[T2H_CW])), gl_add(gl_add(cur[S0_HI], cur[MAJ_HI]), cur[T2L_CW])));
    PUT(boolc(cur[T2H_C0]));
    PUT(gl_sub(cur[T2H_CW], cur[T2H_C0]));
    /* register shifts, both limbs: b<-a, c<-b, d<-c, f<-e, g<-f, h<-g */
    static const int dst[6] = {1, 2, 3, 5, 6, 7}, src[6] = {0, 1, 2, 4, 5, 6};
    for (int k = 0; k < 6; k++) {
        PUT(comp ? gl_sub(nxt[A0 + 2 * dst[k]], cur[A0 + 2 * src[k]]) : 0);
        PUT(comp ? gl_sub(nxt[A0 + 2 * dst[k] + 1], cur[A0 + 2 * src[k] + 1]) : 0);
    }
    if (comp) {
        PUT(gl_sub(gl_sub(gl_add(nxt[A0 + 0], gl_mul(two32, cur[NA_LO])), cur[T1L_LO]), cur[T2L_LO]));
        PUT(boolc(cur[NA_LO]));
        PUT(gl_sub(gl_sub(gl_sub(gl_add(nxt[A0 + 1], gl_mul(two32, cur[NA_HI])), cur[T1H_LO]), cur[T2H_LO]), cur[NA_LO]));
        PUT(boolc(cur[NA_HI]));
        PUT(gl_sub(gl_sub(gl_add(nxt[A0 + 8], gl_mul(two32, cur[NE_LO])), cur[A0 + 6]), cur[T1L_LO]));
        PUT(boolc(cur[NE_LO]));
        PUT(gl_sub(gl_sub(gl_sub(gl_add(nxt[A0 + 9], gl_mul(two32, cur[NE_HI])), cur[A0 + 7]), cur[T1H_LO]), cur[NE_LO]));
        PUT(boolc(cur[NE_HI]));
    } else {
        for (int k = 0; k < 8; k++) PUT(0);
    }
    /* H-state: finalisation with the implicit lo carry delta_lo / 2^32 */
    const uint64_t inv32 = gl_inv(two32);
    for (int k = 0; k < 8; k++) {
        if (fin) {
            uint64_t dl = gl_sub(gl_add(cur[H0 + 2 * k], cur[A0 + 2 * k]), nxt[H0 + 2 * k]);
            uint64_t cl = gl_mul(dl, inv32);
            uint64_t dh = gl_sub(gl_add(gl_add(cur[H0 + 2 * k + 1], cur[A0 + 2 * k + 1]), cl), nxt[H0 + 2 * k + 1]);
            PUT(gl_mul(dl, gl_sub(dl, two32)));
            PUT(gl_mul(dh, gl_sub(dh, two32)));
        } else {
            PUT(gl_sub(nxt[H0 + 2 * k], cur[H0 + 2 * k]));
            PUT(gl_sub(nxt[H0 + 2 * k + 1], cur[H0 + 2 * k + 1]));
        }
    }
    for (int k = 0; k < 14; k++) {
        PUT(gl_sub(nxt[WW0 + 2 * k], cur[WW0 + 2 * (k + 1)]));
        PUT(gl_sub(nxt[WW0 + 2 * k + 1], cur[WW0 + 2 * (k + 1) + 1]));
    }
    PUT(gl_sub(nxt[WW0 + 28], cur[W_LO]));
    PUT(gl_sub(nxt[WW0 + 29], cur[W_HI]));
    if (sched) {
        uint64_t sl = gl_add(gl_add(gl_add(cur[WW0], cur[SS0_LO]), cur[WW0 + 18]), cur[SS1_LO]);
        PUT(gl_sub(gl_add(nxt[W_LO], gl_mul(two32, cur[WSL_CW])), sl));
        PUT(boolc(cur[WSL_C0]));
        PUT(boolc(cur[WSL_C1]));
        PUT(gl_sub(gl_sub(cur[WSL_CW], cur[WSL_C0]), gl_mul(2, cur[WSL_C1])));
        uint64_t sh = gl_add(gl_add(gl_add(gl_add(cur[WW0 + 1], cur[SS0_HI]), cur[WW0 + 19]), cur[SS1_HI]), cur[WSL_CW]);
        PUT(gl_sub(gl_add(nxt[W_HI], gl_mul(two32, cur[WSH_CW])), sh));
        PUT(boolc(cur[WSH_C0]));
        PUT(boolc(cur[WSH_C1]));
        PUT(gl_sub(gl_sub(cur[WSH_CW], cur[WSH_C0]), gl_mul(2, cur[WSH_C1])));
    } else {
        for (int k = 0; k < 8; k++) PUT(0);
    }
#undef PUT
}

typedef struct {
    const uint64_t *lde;
    size_t n, blowup, n_blocks;
    uint64_t *c_eval;
    const uint64_t *lam; /* NULL: lambda_k = k + 1 */
} sha512_merge_job;
static void sha512_merge_range(void *ctx, size_t lo, size_t hi) {
    const sha512_merge_job *J = (const sha512_merge_job *)ctx;
    const size_t n = J->n;
    uint64_t *cur = (uint64_t *)malloc(8 * W512), *nxt = (uint64_t *)malloc(8 * W512);
    uint64_t cv[K512];
    for (size_t i = lo; i < hi; i++) {
        const size_t ni = (i + J->blowup) % n;
        for (size_t c = 0; c < W512; c++) {
            cur[c] = J->lde[c * n + i];
            nxt[c] = J->lde[c * n + ni];
        }
        ref_sha512_constraints(cur, nxt, i / J->blowup, J->n_blocks, cv);
        uint64_t acc = 0;
        for (size_t k = 0; k < K512; k++) acc = gl_add(acc, gl_mul(J->lam ? J->lam[k] : k + 1, cv[k]));
        J->c_eval[i] = acc;
    }
    free(cur);
    free(nxt);
}

int ref_merge_sha512(const uint64_t *lde, size_t n_trace, size_t blowup, size_t n_blocks, int phi_mode,
                     uint64_t *c_eval) {
    return ref_merge_sha512_lam(lde, n_trace, blowup, n_blocks, phi_mode, NULL, c_eval);
}

/* the same merge with combination coefficients lam[0 .. K) (NULL: k + 1) */
int ref_merge_sha512_lam(const uint64_t *lde, size_t n_trace, size_t blowup, size_t n_blocks, int phi_mode,
                         const uint64_t *lam, uint64_t *c_eval) {
    const size_t n = n_trace * blowup;
    const int compute = phi_mode == 0 || (phi_mode == 1 && n >= 4096);
    if (compute) {
        sha512_merge_job J = {lde, n, blowup, n_blocks, c_eval, lam};
        ref_par_for(n, PAR_FRI, sha512_merge_range, &J); /* par_iter_mut at >= 2^12 points (lib.rs:601-605) */
    } else {
        memset(c_eval, 0, 8 * n);
    }
    ref_quotient_from_phi(c_eval, n, n_trace);
    return 0;
}
