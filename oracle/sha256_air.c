/* sha256_air.c -- CPU restatement of the SHA-256 AIR and the DS->KSK
 * binding prover (test infrastructure: only tests/ and the benchmark's
 * CPU baseline load it).
 *
 *   column layout          crates/deep_ali/src/sha256_air.rs:140-253
 *   trace builder          sha256_air.rs:392-479 (pad, multi-block), :648-898 (fill_block)
 *   constraint evaluator   sha256_air.rs:956-1294 (766 constraints, this order)
 *   merge                  crates/deep_ali/src/lib.rs:453-545 (deep_ali_merge_sha256)
 *   prover                 crates/swarm-dns/src/prover.rs:446-585
 *
 * All cells hold canonical field elements; bits are 0/1, words are u32. */
#include <stdlib.h>
#include <string.h>

#include "refcpu.h"
#include "refcpu_int.h"

/* ---- column offsets (sha256_air.rs:140-253) ---- */
enum {
    A_A = 0, A_B, A_C, A_D, A_E, A_F, A_G, A_H,
    BIT_A = 8, BIT_B = 40, BIT_C = 72, BIT_E = 104, BIT_F = 136, BIT_G = 168,
    S0_T = 200, S0_BIT = 232, S0_W = 264,
    S1_T = 265, S1_BIT = 297, S1_W = 329,
    CH_BIT = 330, CH_W = 362,
    MAJ_AB = 363, MAJ_AC = 395, MAJ_BC = 427, MAJ_U = 459, MAJ_BIT = 491, MAJ_W = 523,
    COL_W = 524,
    T1_LO = 525, T1_C0, T1_C1, T1_C2, T1_CW,
    T2_LO = 530, T2_C0, T2_C1, T2_CW,
    H0 = 534,
    WW0 = 542,
    BIT_WW14 = 557, BIT_WW1 = 589,
    SS0_T = 621, SS0_BIT = 653, SS0_W = 685,
    SS1_T = 686, SS1_BIT = 718, SS1_W = 750,
    WS_C0 = 751, WS_C1, WS_CW,
    NA_C = 754, NE_C = 755,
    SHA_W = 756, SHA_K = 766, ROWS_PER_BLOCK = 128
};

static const uint32_t SHA_KC[64] = {  /* FIPS 180-4 section 4.2.2 */
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
static const uint32_t SHA_IV[8] = {  /* section 5.3.3 */
    0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};

static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
static uint32_t bsig0(uint32_t x) { return rotr(x, 2) ^ rotr(x, 13) ^ rotr(x, 22); }
static uint32_t bsig1(uint32_t x) { return rotr(x, 6) ^ rotr(x, 11) ^ rotr(x, 25); }
static uint32_t ssig0(uint32_t x) { return rotr(x, 7) ^ rotr(x, 18) ^ (x >> 3); }
static uint32_t ssig1(uint32_t x) { return rotr(x, 17) ^ rotr(x, 19) ^ (x >> 10); }
static uint32_t ch32(uint32_t e, uint32_t f, uint32_t g) { return (e & f) ^ (~e & g); }
static uint32_t maj32(uint32_t a, uint32_t b, uint32_t c) { return (a & b) ^ (a & c) ^ (b & c); }
static uint32_t bit(uint32_t x, int i) { return (x >> i) & 1; }

/* message schedule of one 64-byte block */
static void schedule(const uint8_t *blk, uint32_t w[64]) {
    for (int i = 0; i < 16; i++)
        w[i] = (uint32_t)blk[4 * i] << 24 | (uint32_t)blk[4 * i + 1] << 16 | (uint32_t)blk[4 * i + 2] << 8 | blk[4 * i + 3];
    for (int t = 16; t < 64; t++) w[t] = ssig1(w[t - 2]) + w[t - 7] + ssig0(w[t - 15]) + w[t - 16];
}

/* FIPS 180-4 padding: number of 64-byte blocks, bytes to `out` (caller
 * sizes it with ref_sha256_n_blocks) */
size_t ref_sha256_n_blocks(size_t len) { return (len + 9 + 63) / 64; }
static void pad_message(const uint8_t *msg, size_t len, uint8_t *out) {
    size_t nb = ref_sha256_n_blocks(len);
    memset(out, 0, 64 * nb);
    memcpy(out, msg, len);
    out[len] = 0x80;
    uint64_t bits = (uint64_t)len * 8;
    for (int i = 0; i < 8; i++) out[64 * nb - 1 - i] = (uint8_t)(bits >> (8 * i));
}

/* fill_block (sha256_air.rs:648-898): rows base .. base+127 of a column-major
 * trace of height `h` */
static void fill_block(uint64_t *tr, size_t h, size_t base, const uint8_t *blk, const uint32_t iv[8],
                       uint32_t post_out[8]) {
#define T(col, row) tr[(size_t)(col) * h + (row)]
    uint32_t w[64], st[65][8];
    schedule(blk, w);
    memcpy(st[0], iv, 32);
    for (int t = 0; t < 64; t++) {
        const uint32_t *s = st[t];
        uint32_t t1 = s[7] + bsig1(s[4]) + ch32(s[4], s[5], s[6]) + SHA_KC[t] + w[t];
        uint32_t t2 = bsig0(s[0]) + maj32(s[0], s[1], s[2]);
        uint32_t n[8] = {t1 + t2, s[0], s[1], s[2], s[3] + t1, s[4], s[5], s[6]};
        memcpy(st[t + 1], n, 32);
    }
    uint32_t digest[8];
    for (int k = 0; k < 8; k++) digest[k] = iv[k] + st[64][k];
    if (post_out) memcpy(post_out, st[64], 32);
    for (int br = 0; br < ROWS_PER_BLOCK; br++) {
        const size_t row = base + br;
        const uint32_t *s = st[br <= 64 ? br : 64];
        const uint32_t a = s[0], b = s[1], c = s[2], d = s[3], e = s[4], f = s[5], g = s[6], hh = s[7];
        for (int k = 0; k < 8; k++) T(A_A + k, row) = s[k];
        for (int i = 0; i < 32; i++) {
            T(BIT_A + i, row) = bit(a, i);
            T(BIT_B + i, row) = bit(b, i);
            T(BIT_C + i, row) = bit(c, i);
            T(BIT_E + i, row) = bit(e, i);
            T(BIT_F + i, row) = bit(f, i);
            T(BIT_G + i, row) = bit(g, i);
        }
        for (int i = 0; i < 32; i++) {
            uint32_t t0 = bit(a, (i + 2) % 32) ^ bit(a, (i + 13) % 32);
            T(S0_T + i, row) = t0;
            T(S0_BIT + i, row) = t0 ^ bit(a, (i + 22) % 32);
            uint32_t t1 = bit(e, (i + 6) % 32) ^ bit(e, (i + 11) % 32);
            T(S1_T + i, row) = t1;
            T(S1_BIT + i, row) = t1 ^ bit(e, (i + 25) % 32);
        }
        T(S0_W, row) = bsig0(a);
        T(S1_W, row) = bsig1(e);
        const uint32_t chw = ch32(e, f, g), mjw = maj32(a, b, c);
        for (int i = 0; i < 32; i++) {
            T(CH_BIT + i, row) = bit(chw, i);
            uint32_t ab = bit(a, i) & bit(b, i), ac = bit(a, i) & bit(c, i), bc = bit(b, i) & bit(c, i);
            T(MAJ_AB + i, row) = ab;
            T(MAJ_AC + i, row) = ac;
            T(MAJ_BC + i, row) = bc;
            T(MAJ_U + i, row) = ab ^ ac;
            T(MAJ_BIT + i, row) = ab ^ ac ^ bc;
        }
        T(CH_W, row) = chw;
        T(MAJ_W, row) = mjw;
        const uint32_t wt = br < 64 ? w[br] : 0, kt = br < 64 ? SHA_KC[br] : 0;
        T(COL_W, row) = wt;
        uint64_t t1f = (uint64_t)hh + bsig1(e) + chw + kt + wt;
        uint32_t t1c = (uint32_t)(t1f >> 32);
        T(T1_LO, row) = (uint32_t)t1f;
        T(T1_C0, row) = t1c & 1;
        T(T1_C1, row) = (t1c >> 1) & 1;
        T(T1_C2, row) = (t1c >> 2) & 1;
        T(T1_CW, row) = t1c;
        uint64_t t2f = (uint64_t)bsig0(a) + mjw;
        uint32_t t2c = (uint32_t)(t2f >> 32);
        T(T2_LO, row) = (uint32_t)t2f;
        T(T2_C0, row) = t2c & 1;
        T(T2_C1, row) = (t2c >> 1) & 1;
        T(T2_CW, row) = t2c;
        for (int k = 0; k < 8; k++) T(H0 + k, row) = br <= 64 ? iv[k] : digest[k];
        /* W window: W_win[k] = W[br - 15 + k], zero outside [0, 64) */
#define WAT(i) (((i) >= 0 && (i) < 64) ? w[(i)] : 0u)
        for (int k = 0; k < 15; k++) T(WW0 + k, row) = WAT(br - 15 + k);
        const uint32_t ww14 = WAT(br - 1), ww1 = WAT(br - 14);
        for (int i = 0; i < 32; i++) {
            T(BIT_WW14 + i, row) = bit(ww14, i);
            T(BIT_WW1 + i, row) = bit(ww1, i);
            uint32_t p = bit(ww1, (i + 7) % 32) ^ bit(ww1, (i + 18) % 32);
            T(SS0_T + i, row) = p;
            T(SS0_BIT + i, row) = p ^ (i + 3 < 32 ? bit(ww1, i + 3) : 0);
            uint32_t q = bit(ww14, (i + 17) % 32) ^ bit(ww14, (i + 19) % 32);
            T(SS1_T + i, row) = q;
            T(SS1_BIT + i, row) = q ^ (i + 10 < 32 ? bit(ww14, i + 10) : 0);
        }
        T(SS0_W, row) = ssig0(ww1);
        T(SS1_W, row) = ssig1(ww14);
        uint64_t sum = (uint64_t)WAT(br - 15) + ssig0(ww1) + WAT(br - 6) + ssig1(ww14);
        uint32_t sc = (uint32_t)(sum >> 32);
        T(WS_C0, row) = sc & 1;
        T(WS_C1, row) = (sc >> 1) & 1;
        T(WS_CW, row) = sc;
#undef WAT
        if (br < 64) {
            uint32_t t1lo = (uint32_t)t1f, t2lo = (uint32_t)t2f;
            T(NA_C, row) = (uint32_t)(((uint64_t)t1lo + t2lo) >> 32);
            T(NE_C, row) = (uint32_t)(((uint64_t)d + t1lo) >> 32);
        } else {
            T(NA_C, row) = 0;
            T(NE_C, row) = 0;
        }
    }
#undef T
}

/* build_sha256_trace_multi (sha256_air.rs:429-479): height =
 * next_pow2(128 n_blocks), rows past the last block copy the last useful row */
size_t ref_sha256_trace_height(size_t len) {
    size_t useful = ROWS_PER_BLOCK * ref_sha256_n_blocks(len), h = ROWS_PER_BLOCK;
    while (h < useful) h <<= 1;
    return h;
}
int ref_build_sha256_trace(const uint8_t *msg, size_t len, uint64_t *tr, uint8_t digest[32]) {
    const size_t nb = ref_sha256_n_blocks(len), h = ref_sha256_trace_height(len);
    uint8_t *blocks = (uint8_t *)malloc(64 * nb);
    pad_message(msg, len, blocks);
    memset(tr, 0, 8 * SHA_W * h);
    uint32_t iv[8], post[8];
    memcpy(iv, SHA_IV, 32);
    for (size_t b = 0; b < nb; b++) {
        fill_block(tr, h, b * ROWS_PER_BLOCK, blocks + 64 * b, iv, post);
        for (int k = 0; k < 8; k++) iv[k] += post[k];
    }
    const size_t useful = ROWS_PER_BLOCK * nb;
    for (size_t r = useful; r < h; r++)
        for (size_t c = 0; c < SHA_W; c++) tr[c * h + r] = tr[c * h + useful - 1];
    /* the digest as prove_ds_ksk_binding reads it: H-state at block row 65
     * of the last block, big-endian words (prover.rs:529-538) */
    const size_t drow = ROWS_PER_BLOCK * (nb - 1) + 65;
    for (int k = 0; k < 8; k++) {
        uint32_t wv = (uint32_t)tr[(H0 + k) * h + drow];
        for (int i = 0; i < 4; i++) digest[4 * k + i] = (uint8_t)(wv >> (24 - 8 * i));
    }
    free(blocks);
    return (int)nb;
}

/* ---- eval_sha256_constraints (sha256_air.rs:956-1294) ---- */
static uint64_t xorp(uint64_t x, uint64_t y) { return gl_sub(gl_add(x, y), gl_mul(2, gl_mul(x, y))); }
static uint64_t boolc(uint64_t b) { return gl_mul(b, gl_sub(1, b)); }

/* cur[c] / nxt[c] read column c; out receives the 766 values in the
 * reference's order */
void ref_sha256_constraints(const uint64_t *cur, const uint64_t *nxt, size_t row, size_t n_blocks, uint64_t *out) {
    size_t j = 0;
#define PUT(v) (out[j++] = (v))
    const size_t block_idx = row / ROWS_PER_BLOCK, br = row % ROWS_PER_BLOCK;
    const int past = block_idx >= n_blocks;
    const int comp = !past && br < 64, fin = !past && br == 64, sched = !past && br >= 15 && br <= 62;
    const uint64_t two32 = (uint64_t)1 << 32;
    static const int boff[8] = {BIT_A, BIT_B, BIT_C, BIT_E, BIT_F, BIT_G, BIT_WW14, BIT_WW1};
    for (int o = 0; o < 8; o++)
        for (int i = 0; i < 32; i++) PUT(boolc(cur[boff[o] + i]));
    /* word = sum 2^i bit_i */
#define PACK(off) ({ uint64_t s_ = 0; for (int i_ = 0; i_ < 32; i_++) s_ = gl_add(s_, gl_mul((uint64_t)1 << i_, cur[(off) + i_])); s_; })
    static const int wcol[8] = {A_A, A_B, A_C, A_E, A_F, A_G, WW0 + 14, WW0 + 1};
    for (int o = 0; o < 8; o++) PUT(gl_sub(cur[wcol[o]], PACK(boff[o])));
    /* Sigma0(a), Sigma1(e), sigma0(W_win[1]), sigma1(W_win[14]) */
    struct { int src, t, b, w, r1, r2, r3, shr; } sg[4] = {
        {BIT_A, S0_T, S0_BIT, S0_W, 2, 13, 22, 0},
        {BIT_E, S1_T, S1_BIT, S1_W, 6, 11, 25, 0},
        {BIT_WW1, SS0_T, SS0_BIT, SS0_W, 7, 18, 3, 1},
        {BIT_WW14, SS1_T, SS1_BIT, SS1_W, 17, 19, 10, 1}};
    for (int s = 0; s < 4; s++) {
        for (int i = 0; i < 32; i++)
            PUT(gl_sub(cur[sg[s].t + i], xorp(cur[sg[s].src + (i + sg[s].r1) % 32], cur[sg[s].src + (i + sg[s].r2) % 32])));
        for (int i = 0; i < 32; i++) {
            uint64_t r = sg[s].shr ? (i + sg[s].r3 < 32 ? cur[sg[s].src + i + sg[s].r3] : 0)
                                   : cur[sg[s].src + (i + sg[s].r3) % 32];
            PUT(gl_sub(cur[sg[s].b + i], xorp(cur[sg[s].t + i], r)));
        }
        PUT(gl_sub(cur[sg[s].w], PACK(sg[s].b)));
    }
    /* Ch */
    for (int i = 0; i < 32; i++) {
        uint64_t e = cur[BIT_E + i], f = cur[BIT_F + i], g = cur[BIT_G + i];
        PUT(gl_sub(cur[CH_BIT + i], gl_sub(gl_add(gl_mul(e, f), g), gl_mul(e, g))));
    }
    PUT(gl_sub(cur[CH_W], PACK(CH_BIT)));
    /* Maj */
    for (int i = 0; i < 32; i++) {
        uint64_t a = cur[BIT_A + i], b = cur[BIT_B + i], c = cur[BIT_C + i];
        PUT(gl_sub(cur[MAJ_AB + i], gl_mul(a, b)));
        PUT(gl_sub(cur[MAJ_AC + i], gl_mul(a, c)));
        PUT(gl_sub(cur[MAJ_BC + i], gl_mul(b, c)));
    }
    for (int i = 0; i < 32; i++) PUT(gl_sub(cur[MAJ_U + i], xorp(cur[MAJ_AB + i], cur[MAJ_AC + i])));
    for (int i = 0; i < 32; i++) PUT(gl_sub(cur[MAJ_BIT + i], xorp(cur[MAJ_U + i], cur[MAJ_BC + i])));
    PUT(gl_sub(cur[MAJ_W], PACK(MAJ_BIT)));
#undef PACK
    /* T1, T2 */
    const uint64_t krow = comp ? SHA_KC[br] : 0;
    uint64_t t1s = gl_add(gl_add(gl_add(gl_add(cur[A_H], cur[S1_W]), cur[CH_W]), krow), cur[COL_W]);
    PUT(gl_sub(gl_add(cur[T1_LO], gl_mul(two32, cur[T1_CW])), t1s));
    PUT(boolc(cur[T1_C0]));
    PUT(boolc(cur[T1_C1]));
    PUT(boolc(cur[T1_C2]));
    PUT(gl_sub(gl_sub(gl_sub(cur[T1_CW], cur[T1_C0]), gl_mul(2, cur[T1_C1])), gl_mul(4, cur[T1_C2])));
    PUT(gl_sub(gl_add(cur[T2_LO], gl_mul(two32, cur[T2_CW])), gl_add(cur[S0_W], cur[MAJ_W])));
    PUT(boolc(cur[T2_C0]));
    PUT(boolc(cur[T2_C1]));
    PUT(gl_sub(gl_sub(cur[T2_CW], cur[T2_C0]), gl_mul(2, cur[T2_C1])));
    /* register shifts */
    static const int sh_dst[6] = {A_B, A_C, A_D, A_F, A_G, A_H}, sh_src[6] = {A_A, A_B, A_C, A_E, A_F, A_G};
    for (int k = 0; k < 6; k++) PUT(comp ? gl_sub(nxt[sh_dst[k]], cur[sh_src[k]]) : 0);
    /* new_a, new_e */
    if (comp) {
        PUT(gl_sub(gl_sub(gl_add(nxt[A_A], gl_mul(two32, cur[NA_C])), cur[T1_LO]), cur[T2_LO]));
        PUT(boolc(cur[NA_C]));
        PUT(gl_sub(gl_sub(gl_add(nxt[A_E], gl_mul(two32, cur[NE_C])), cur[A_D]), cur[T1_LO]));
        PUT(boolc(cur[NE_C]));
        PUT(0);
        PUT(0);
    } else {
        for (int k = 0; k < 6; k++) PUT(0);
    }
    /* H-state */
    for (int k = 0; k < 8; k++) {
        if (fin) {
            uint64_t dl = gl_sub(gl_add(cur[H0 + k], cur[A_A + k]), nxt[H0 + k]);
            PUT(gl_mul(dl, gl_sub(dl, two32)));
        } else {
            PUT(gl_sub(nxt[H0 + k], cur[H0 + k]));
        }
    }
    /* W window */
    for (int k = 0; k < 14; k++) PUT(gl_sub(nxt[WW0 + k], cur[WW0 + k + 1]));
    PUT(gl_sub(nxt[WW0 + 14], cur[COL_W]));
    /* schedule recurrence */
    if (sched) {
        uint64_t sum = gl_add(gl_add(gl_add(cur[WW0], cur[SS0_W]), cur[WW0 + 9]), cur[SS1_W]);
        PUT(gl_sub(gl_add(nxt[COL_W], gl_mul(two32, cur[WS_CW])), sum));
        PUT(boolc(cur[WS_C0]));
        PUT(boolc(cur[WS_C1]));
        PUT(gl_sub(gl_sub(cur[WS_CW], cur[WS_C0]), gl_mul(2, cur[WS_C1])));
    } else {
        for (int k = 0; k < 4; k++) PUT(0);
    }
#undef PUT
}

/* deep_ali_merge_sha256 (lib.rs:453-545): phi(i) = sum_j (j+1) C_j(i) over
 * the LDE (cur = point i, nxt = point i + blowup, row = i / blowup), then
 * the Z_H quotient (remainder dropped) evaluated back on the domain. */
typedef struct {
    const uint64_t *lde;
    size_t n, blowup, n_blocks;
    uint64_t *c_eval;
    const uint64_t *lam; /* NULL: lambda_k = k + 1 */
} sha256_merge_job;
static void sha256_merge_range(void *ctx, size_t lo, size_t hi) {
    const sha256_merge_job *J = (const sha256_merge_job *)ctx;
    const size_t n = J->n;
    uint64_t *cur = (uint64_t *)malloc(8 * SHA_W), *nxt = (uint64_t *)malloc(8 * SHA_W);
    uint64_t cv[SHA_K];
    for (size_t i = lo; i < hi; i++) {
        const size_t ni = (i + J->blowup) % n;
        for (size_t c = 0; c < SHA_W; c++) {
            cur[c] = J->lde[c * n + i];
            nxt[c] = J->lde[c * n + ni];
        }
        ref_sha256_constraints(cur, nxt, i / J->blowup, J->n_blocks, cv);
        uint64_t acc = 0;
        for (size_t k = 0; k < SHA_K; k++) acc = gl_add(acc, gl_mul(J->lam ? J->lam[k] : k + 1, cv[k]));
        J->c_eval[i] = acc;
    }
    free(cur);
    free(nxt);
}

int ref_merge_sha256(const uint64_t *lde, size_t n_trace, size_t blowup, size_t n_blocks, int phi_mode,
                     uint64_t *c_eval) {
    return ref_merge_sha256_lam(lde, n_trace, blowup, n_blocks, phi_mode, NULL, c_eval);
}

/* the same merge with combination coefficients lam[0 .. K) (NULL: k + 1) */
int ref_merge_sha256_lam(const uint64_t *lde, size_t n_trace, size_t blowup, size_t n_blocks, int phi_mode,
                         const uint64_t *lam, uint64_t *c_eval) {
    const size_t n = n_trace * blowup;
    const int compute = phi_mode == 0 || (phi_mode == 1 && n >= 4096);
    if (compute) {
        sha256_merge_job J = {lde, n, blowup, n_blocks, c_eval, lam};
        ref_par_for(n, PAR_FRI, sha256_merge_range, &J); /* par_iter_mut at >= 2^12 points (lib.rs:493-497) */
    } else {
        memset(c_eval, 0, 8 * n);
    }
    ref_quotient_from_phi(c_eval, n, n_trace);
    return 0;
}

/* prove_ds_ksk_binding (prover.rs:512-585), STIR, blowup 32, 54 queries */
int ref_prove_ds_ksk_binding(const uint8_t *dnskey, size_t len, const uint8_t parent_ds_hash[32],
                             const uint8_t fs_binding[32], int phi_mode, uint8_t **proof, size_t *proof_len,
                             uint8_t pi_hash[32], uint8_t asserted_digest[32], uint8_t root_f0[32],
                             size_t *n_blocks_out, size_t *n_trace_out) {
    const size_t h = ref_sha256_trace_height(len), n0 = h * 32;
    uint64_t *tr = (uint64_t *)malloc(8 * SHA_W * h);
    const size_t nb = (size_t)ref_build_sha256_trace(dnskey, len, tr, asserted_digest);
    uint64_t *lde = (uint64_t *)malloc(8 * SHA_W * n0), *c = (uint64_t *)malloc(8 * n0);
    ref_lde_columns(tr, SHA_W, h, 32, 1, lde);
    ref_merge_sha256(lde, h, 32, nb, phi_mode, c);
    free(tr);
    free(lde);
    size_t sched[24];
    ref_fri_params p;
    p.schedule = sched;
    p.n_layers = ref_make_schedule(n0, sched);
    p.r = 54;
    p.s0 = 54;
    p.seed_z = 0xDEEFBAADULL;
    p.coeff_commit_final = 1;
    p.d_final = 1;
    p.stir = ref_ldt_stir();
    p.public_inputs_hash = fs_binding;
    int rc = ref_deep_fri_prove(c, n0, &p, proof, proof_len);
    free(c);
    if (rc) return rc;
    memcpy(root_f0, *proof, 32);
    /* ds_ksk_pi_hash (prover.rs:446-462) */
    sha3_ctx s;
    uint8_t le[8];
    for (int i = 0; i < 8; i++) le[i] = (uint8_t)((uint64_t)len >> (8 * i));
    sha3_init(&s);
    sha3_update(&s, "DS-KSK-PIHASH-V1", 16);
    sha3_update(&s, le, 8);
    sha3_update(&s, dnskey, len);
    sha3_update(&s, parent_ds_hash, 32);
    sha3_update(&s, asserted_digest, 32);
    sha3_update(&s, root_f0, 32);
    sha3_update(&s, fs_binding, 32);
    sha3_final(&s, pi_hash);
    *n_blocks_out = nb;
    *n_trace_out = h;
    return 0;
}
