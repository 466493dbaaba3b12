/* ed25519_air.c -- CPU restatement of the Ed25519 cofactored-verify AIR, v16
 * (test infrastructure: only tests/ load it).
 *
 * THIS IS THE ORACLE: the checker of the HIP product path, never the product.
 *
 *   field F_{2^255-19}, 10 limbs 26/25 bits   crates/deep_ali/src/ed25519_field.rs:103-653
 *   Edwards group (extended coordinates)      ed25519_group.rs:55-311
 *   scalar reduction mod L                    ed25519_scalar.rs:33-216
 *   gadgets: add / sub / mul / select         ed25519_field_air.rs:201-1505
 *   point add / double / cond add / decompress ed25519_group_air.rs:84-1100
 *   scalar reduce gadget                      ed25519_scalar_air.rs:61-584
 *   scalar-mult ladder row                    ed25519_scalar_mult_air.rs:60-257
 *   v0..v16 layout, fill, per-row evaluator   ed25519_verify_air.rs:142-4958
 *   streaming merge                           crates/deep_ali/src/lib.rs:854-974
 *   prove_zsk_ksk_binding_v2, pi hash         crates/swarm-dns/src/prover.rs:763-781, 826-951
 *   native verify                             ed25519_verify.rs:54-118
 *
 * Field elements are kept as canonical 256-bit integers (4 x u64); every
 * native operation of the reference (ed25519_field.rs) is exact arithmetic
 * mod p followed, where the trace stores a value, by `freeze`, so the cell
 * values are the canonical limbs. The one loose constant is 2d, whose limbs
 * are twice those of d (ed25519_field.rs:609, `D.add(D)`), as the point-add
 * gadget places it. The layout follows the reference's version chain: the
 * v16 evaluator is v7 + the v10..v16 blocks (v8 / v9 are not on that chain),
 * and the v16 trace is the v7 trace plus the v14 / v15 rows.
 *
 * Parity against the reference is unpinned (no Ed25519 proof or c_eval is
 * stored in the reference). The tests pin this file with RFC 8032 vectors
 * read through the trace (digest, k mod L, decompressed points, ladder
 * results, the identity verdict), by all constraints vanishing on valid
 * traces, and by planted mistakes whose slot and value follow from the
 * reference formulas. */
#include <stdlib.h>
#include <string.h>

#include "refcpu.h"
#include "refcpu_int.h"

/* ------------------------------------------------------------------ field */

enum { NL = 10, EB = 255, EC = 265, MCB = 36 };
static const int LW[NL] = {26, 25, 26, 25, 26, 25, 26, 25, 26, 25};
static const int LOFF[NL] = {0, 26, 51, 77, 102, 128, 153, 179, 204, 230};
static const uint64_t P25[4] = {0xFFFFFFFFFFFFFFEDULL, 0xFFFFFFFFFFFFFFFFULL, 0xFFFFFFFFFFFFFFFFULL,
                                0x7FFFFFFFFFFFFFFFULL};

typedef struct { uint64_t v[4]; } fe; /* canonical, < p */

static int u256_ge(const uint64_t *a, const uint64_t *b) {
    for (int i = 3; i >= 0; i--)
        if (a[i] != b[i]) return a[i] > b[i];
    return 1;
}
static void u256_sub(uint64_t *a, const uint64_t *b) {
    u128 br = 0;
    for (int i = 0; i < 4; i++) {
        u128 d = (u128)a[i] - b[i] - br;
        a[i] = (uint64_t)d;
        br = (d >> 127) & 1;
    }
}
/* reduce an 8-word product with 2^256 = 38 (mod p) (ed25519_field.rs:303-337) */
static fe fe_red512(const uint64_t w[8]) {
    uint64_t lo[4] = {w[0], w[1], w[2], w[3]};
    u128 c = 0;
    for (int i = 0; i < 4; i++) {
        c += (u128)lo[i] + (u128)w[4 + i] * 38;
        lo[i] = (uint64_t)c;
        c >>= 64;
    }
    while (c) {
        u128 d = (u128)c * 38;
        c = 0;
        for (int i = 0; i < 4; i++) {
            d += lo[i];
            lo[i] = (uint64_t)d;
            d >>= 64;
        }
        c = d;
    }
    fe r;
    memcpy(r.v, lo, 32);
    while (u256_ge(r.v, P25)) u256_sub(r.v, P25);
    return r;
}
static fe fe_mul(fe a, fe b) {
    uint64_t w[8] = {0};
    for (int i = 0; i < 4; i++) {
        u128 c = 0;
        for (int j = 0; j < 4; j++) {
            c += (u128)a.v[i] * b.v[j] + w[i + j];
            w[i + j] = (uint64_t)c;
            c >>= 64;
        }
        w[i + 4] = (uint64_t)c;
    }
    return fe_red512(w);
}
static fe fe_add(fe a, fe b) {
    fe r;
    u128 c = 0;
    for (int i = 0; i < 4; i++) {
        c += (u128)a.v[i] + b.v[i];
        r.v[i] = (uint64_t)c;
        c >>= 64;
    }
    if (u256_ge(r.v, P25)) u256_sub(r.v, P25);
    return r;
}
static fe fe_sub(fe a, fe b) {
    fe r = a;
    if (!u256_ge(a.v, b.v)) {
        u128 c = 0;
        for (int i = 0; i < 4; i++) {
            c += (u128)r.v[i] + P25[i];
            r.v[i] = (uint64_t)c;
            c >>= 64;
        }
    }
    u256_sub(r.v, b.v);
    return r;
}
static fe fe_zero(void) { fe r = {{0, 0, 0, 0}}; return r; }
static fe fe_one(void) { fe r = {{1, 0, 0, 0}}; return r; }
static fe fe_neg(fe a) { return fe_sub(fe_zero(), a); }
static int fe_eq(fe a, fe b) { return memcmp(a.v, b.v, 32) == 0; }
static int fe_is_zero(fe a) { return (a.v[0] | a.v[1] | a.v[2] | a.v[3]) == 0; }
static fe fe_pow(fe a, const uint64_t e[4]) {
    fe r = fe_one();
    for (int i = 255; i >= 0; i--) {
        r = fe_mul(r, r);
        if ((e[i / 64] >> (i % 64)) & 1) r = fe_mul(r, a);
    }
    return r;
}
static fe fe_inv(fe a) { /* a^(p-2) */
    const uint64_t e[4] = {0xFFFFFFFFFFFFFFEBULL, ~0ULL, ~0ULL, 0x7FFFFFFFFFFFFFFFULL};
    return fe_pow(a, e);
}
/* from_bytes: 32 LE bytes, bit 255 masked; the value is reduced mod p here
 * (the reference keeps tight limbs of the raw 255-bit value; the two differ
 * only for the 19 non-canonical encodings) */
static fe fe_from_bytes(const uint8_t b[32]) {
    fe r;
    for (int k = 0; k < 4; k++) {
        r.v[k] = 0;
        for (int i = 0; i < 8; i++) r.v[k] |= (uint64_t)b[8 * k + i] << (8 * i);
    }
    r.v[3] &= 0x7FFFFFFFFFFFFFFFULL;
    if (u256_ge(r.v, P25)) u256_sub(r.v, P25);
    return r;
}
static void fe_to_bytes(fe a, uint8_t b[32]) {
    for (int k = 0; k < 4; k++)
        for (int i = 0; i < 8; i++) b[8 * k + i] = (uint8_t)(a.v[k] >> (8 * i));
}
static int fe_is_negative(fe a) { return (int)(a.v[0] & 1); }
static fe fe_small(int64_t k) {
    fe r = fe_zero();
    if (k >= 0) {
        r.v[0] = (uint64_t)k;
        return r;
    }
    r.v[0] = (uint64_t)(-k);
    return fe_neg(r);
}
/* tight limbs of a canonical value (the `freeze`d FieldElement) */
static void fe_limbs(fe a, int64_t l[NL]) {
    for (int i = 0; i < NL; i++) {
        const int off = LOFF[i], w = LW[i], lo = off / 64, sh = off % 64;
        uint64_t x = a.v[lo] >> sh;
        if (sh + w > 64 && lo + 1 < 4) x |= a.v[lo + 1] << (64 - sh);
        l[i] = (int64_t)(x & ((1ULL << w) - 1));
    }
}
/* value of possibly loose limbs: sum_i l_i 2^off_i mod p */
static fe fe_from_limbs(const int64_t l[NL]) {
    fe r = fe_zero();
    for (int i = 0; i < NL; i++) {
        fe t = fe_small(l[i]), s = fe_zero();
        const int off = LOFF[i];
        s.v[off / 64] = 1ULL << (off % 64);
        r = fe_add(r, fe_mul(t, s));
    }
    return r;
}

static fe ED_D, ED_D2, ED_SQRTM1;
static int64_t D_LIMBS[NL], D2_LIMBS[NL];

/* sqrt_ratio_i + decompress (ed25519_field.rs:538-568, ed25519_group.rs:157-179) */
typedef struct { fe X, Y, Z, T; } ept;
static int ept_decompress(const uint8_t in[32], ept *out) {
    const int sign = (in[31] >> 7) & 1;
    fe y = fe_from_bytes(in), one = fe_one();
    fe y2 = fe_mul(y, y), u = fe_sub(y2, one), v = fe_add(fe_mul(ED_D, y2), one);
    fe v3 = fe_mul(v, fe_mul(v, v)), v7 = fe_mul(v3, fe_mul(v3, v));
    const uint64_t e58[4] = {0xFFFFFFFFFFFFFFFDULL, ~0ULL, ~0ULL, 0x0FFFFFFFFFFFFFFFULL}; /* (p-5)/8 */
    fe r = fe_mul(fe_mul(u, v3), fe_pow(fe_mul(u, v7), e58));
    fe check = fe_mul(v, fe_mul(r, r));
    const int pos = fe_eq(check, u), neg = fe_eq(check, fe_neg(u));
    if (!pos && !neg) return 0;
    if (neg) r = fe_mul(r, ED_SQRTM1);
    if (fe_is_negative(r)) r = fe_neg(r);
    if (fe_is_negative(r) != sign) r = fe_neg(r);
    out->X = r;
    out->Y = y;
    out->Z = one;
    out->T = fe_mul(r, y);
    return 1;
}
static ept ept_identity(void) { ept p = {fe_zero(), fe_one(), fe_one(), fe_zero()}; return p; }
/* extended-coordinate add / double exactly as ed25519_group.rs:92-140 */
static ept ept_add(ept p, ept q) {
    fe a = fe_mul(fe_sub(p.Y, p.X), fe_sub(q.Y, q.X));
    fe b = fe_mul(fe_add(p.Y, p.X), fe_add(q.Y, q.X));
    fe c = fe_mul(fe_mul(p.T, q.T), ED_D2);
    fe zz = fe_mul(p.Z, q.Z), d = fe_add(zz, zz);
    fe e = fe_sub(b, a), f = fe_sub(d, c), g = fe_add(d, c), h = fe_add(b, a);
    ept r = {fe_mul(e, f), fe_mul(g, h), fe_mul(f, g), fe_mul(e, h)};
    return r;
}
static ept ept_double(ept p) {
    fe a = fe_mul(p.X, p.X), b = fe_mul(p.Y, p.Y), zz = fe_mul(p.Z, p.Z), c = fe_add(zz, zz);
    fe h = fe_add(a, b), xy = fe_add(p.X, p.Y), e = fe_sub(h, fe_mul(xy, xy)), g = fe_sub(a, b), f = fe_add(c, g);
    ept r = {fe_mul(e, f), fe_mul(g, h), fe_mul(f, g), fe_mul(e, h)};
    return r;
}
static ept ept_neg(ept p) { ept r = {fe_neg(p.X), p.Y, p.Z, fe_neg(p.T)}; return r; }
static int ept_is_identity(ept p) { return fe_is_zero(p.X) && fe_eq(p.Y, p.Z); }
/* MSB-first double-and-add from the identity (ed25519_verify_air.rs:4834-4838) */
static ept ept_ladder(ept base, const uint8_t *bits, size_t k) {
    ept acc = ept_identity();
    for (size_t i = 0; i < k; i++) {
        acc = ept_double(acc);
        if (bits[i]) acc = ept_add(acc, base);
    }
    return acc;
}
static ept ED_B;

static int ed_inited = 0;
static void ed_init(void) {
    if (ed_inited) return;
    /* d = -121665 / 121666 (ed25519_field.rs:601-606) */
    ED_D = fe_mul(fe_small(-121665), fe_inv(fe_small(121666)));
    ED_D2 = fe_add(ED_D, ED_D);
    /* sqrt(-1) = 2^((p-1)/4) (ed25519_field.rs:616-649) */
    const uint64_t e14[4] = {0xFFFFFFFFFFFFFFFBULL, ~0ULL, ~0ULL, 0x1FFFFFFFFFFFFFFFULL};
    ED_SQRTM1 = fe_pow(fe_small(2), e14);
    fe_limbs(ED_D, D_LIMBS);
    for (int i = 0; i < NL; i++) D2_LIMBS[i] = 2 * D_LIMBS[i]; /* D.add(D): loose limbs */
    uint8_t bc[32];
    bc[0] = 0x58;
    for (int i = 1; i < 32; i++) bc[i] = 0x66;
    ept_decompress(bc, &ED_B);
    ed_inited = 1;
}

/* ------------------------------------------------------------------ scalars mod L */

static const uint64_t LL[4] = {0x5812631a5cf5d3edULL, 0x14def9dea2f79cd6ULL, 0, 0x1000000000000000ULL};
/* 512-bit long division by L (ed25519_scalar.rs:156-211); q has <= 260 bits */
static void divmod_l(const uint8_t in[64], uint64_t q[5], uint64_t r[4]) {
    uint64_t x[8], ls[8] = {LL[0], LL[1], LL[2], LL[3], 0, 0, 0, 0};
    for (int k = 0; k < 8; k++) {
        x[k] = 0;
        for (int i = 0; i < 8; i++) x[k] |= (uint64_t)in[8 * k + i] << (8 * i);
    }
    for (int s = 0; s < 259; s++) {
        for (int i = 7; i > 0; i--) ls[i] = ls[i] << 1 | ls[i - 1] >> 63;
        ls[0] <<= 1;
    }
    memset(q, 0, 40);
    for (int step = 259; step >= 0; step--) {
        int ge = 1;
        for (int i = 7; i >= 0; i--)
            if (x[i] != ls[i]) { ge = x[i] > ls[i]; break; }
        if (ge) {
            u128 br = 0;
            for (int i = 0; i < 8; i++) {
                u128 d = (u128)x[i] - ls[i] - br;
                x[i] = (uint64_t)d;
                br = (d >> 127) & 1;
            }
            q[step / 64] |= 1ULL << (step % 64);
        }
        for (int i = 0; i < 7; i++) ls[i] = ls[i] >> 1 | ls[i + 1] << 63;
        ls[7] >>= 1;
    }
    memcpy(r, x, 32);
}
static void reduce_mod_l_wide(const uint8_t in[64], uint8_t out[32]) {
    uint64_t q[5], r[4];
    divmod_l(in, q, r);
    for (int k = 0; k < 4; k++)
        for (int i = 0; i < 8; i++) out[8 * k + i] = (uint8_t)(r[k] >> (8 * i));
}

/* ------------------------------------------------------------------ SHA-512 (FIPS 180-4) */

static const uint64_t K512[80] = {
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
#define ROTR64(x, n) (((x) >> (n)) | ((x) << (64 - (n))))
static void sha512(const uint8_t *msg, size_t len, uint8_t out[64]) {
    uint64_t h[8] = {0x6a09e667f3bcc908ULL, 0xbb67ae8584caa73bULL, 0x3c6ef372fe94f82bULL, 0xa54ff53a5f1d36f1ULL,
                     0x510e527fade682d1ULL, 0x9b05688c2b3e6c1fULL, 0x1f83d9abfb41bd6bULL, 0x5be0cd19137e2179ULL};
    const size_t nb = (len + 17 + 127) / 128;
    uint8_t *m = (uint8_t *)calloc(128 * nb, 1);
    memcpy(m, msg, len);
    m[len] = 0x80;
    for (int i = 0; i < 8; i++) m[128 * nb - 1 - i] = (uint8_t)(((uint64_t)len * 8) >> (8 * i));
    for (size_t b = 0; b < nb; b++) {
        uint64_t w[80], s[8];
        for (int t = 0; t < 16; t++) {
            w[t] = 0;
            for (int i = 0; i < 8; i++) w[t] = w[t] << 8 | m[128 * b + 8 * t + i];
        }
        for (int t = 16; t < 80; t++) {
            uint64_t s0 = ROTR64(w[t - 15], 1) ^ ROTR64(w[t - 15], 8) ^ (w[t - 15] >> 7);
            uint64_t s1 = ROTR64(w[t - 2], 19) ^ ROTR64(w[t - 2], 61) ^ (w[t - 2] >> 6);
            w[t] = w[t - 16] + s0 + w[t - 7] + s1;
        }
        memcpy(s, h, 64);
        for (int t = 0; t < 80; t++) {
            uint64_t S1 = ROTR64(s[4], 14) ^ ROTR64(s[4], 18) ^ ROTR64(s[4], 41);
            uint64_t ch = (s[4] & s[5]) ^ (~s[4] & s[6]);
            uint64_t t1 = s[7] + S1 + ch + K512[t] + w[t];
            uint64_t S0 = ROTR64(s[0], 28) ^ ROTR64(s[0], 34) ^ ROTR64(s[0], 39);
            uint64_t mj = (s[0] & s[1]) ^ (s[0] & s[2]) ^ (s[1] & s[2]);
            memmove(s + 1, s, 56);
            s[4] += t1;
            s[0] = t1 + S0 + mj;
        }
        for (int k = 0; k < 8; k++) h[k] += s[k];
    }
    for (int k = 0; k < 8; k++)
        for (int i = 0; i < 8; i++) out[8 * k + i] = (uint8_t)(h[k] >> (56 - 8 * i));
    free(m);
}

/* ------------------------------------------------------------------ gadget layouts */

typedef struct { size_t a, b, c, cb, pos, neg; } sub_l;
typedef struct { size_t a, b, c, cb, cy; } add_l;
typedef struct { size_t a, b, c, cb, cyb; } mul_l;
typedef struct { size_t a, b, bit, c; } sel_l;
typedef struct {
    size_t p1[4], p2[4], d2;
    sub_l ym1, ym2;
    add_l yp1, yp2;
    mul_l A, B, tt, C, zz;
    add_l D;
    sub_l E, F;
    add_l G, H;
    mul_l X3, Y3, T3, Z3;
    size_t end;
} padd_l;
typedef struct {
    size_t p1[3];
    mul_l sA, sB, szz;
    add_l C, H, xy;
    mul_l sxy;
    sub_l E, G;
    add_l F;
    mul_l X3, Y3, T3, Z3;
    size_t end;
} pdbl_l;
typedef struct { size_t bit, out[4]; padd_l add; sel_l sel[4]; size_t end; } cadd_l;
typedef struct { size_t y, sign, d, x, xb; mul_l sqx, sqy, dx2, dxy; sub_l y2m1, lhs, eq; size_t end; } decomp_l;
typedef struct { size_t in, q, qb, r, rb, sl, slb, cyb; } reduce_l;
typedef struct { size_t base[4], acc[4]; pdbl_l dbl; cadd_l ca; size_t width; } smult_l;

/* CellAllocator (ed25519_group_air.rs:170-200) */
static size_t al(size_t *n, size_t k) { size_t r = *n; *n += k; return r; }
static sub_l al_sub(size_t *n, size_t a, size_t b) {
    sub_l s = {a, b, 0, 0, 0, 0};
    s.c = al(n, NL); s.cb = al(n, EB); s.pos = al(n, NL); s.neg = al(n, NL);
    return s;
}
static add_l al_add(size_t *n, size_t a, size_t b) {
    add_l s = {a, b, 0, 0, 0};
    s.c = al(n, NL); s.cb = al(n, EB); s.cy = al(n, NL);
    return s;
}
static mul_l al_mul(size_t *n, size_t a, size_t b) {
    mul_l s = {a, b, 0, 0, 0};
    s.c = al(n, NL); s.cb = al(n, EB); s.cyb = al(n, MCB * NL);
    return s;
}
/* point_add_layout_at (ed25519_group_air.rs:202-249) */
static padd_l padd_at(size_t base, const size_t p1[4], const size_t p2[4]) {
    padd_l L;
    size_t n = base;
    memcpy(L.p1, p1, sizeof L.p1);
    memcpy(L.p2, p2, sizeof L.p2);
    L.d2 = al(&n, NL);
    L.ym1 = al_sub(&n, p1[1], p1[0]);
    L.ym2 = al_sub(&n, p2[1], p2[0]);
    L.yp1 = al_add(&n, p1[1], p1[0]);
    L.yp2 = al_add(&n, p2[1], p2[0]);
    L.A = al_mul(&n, L.ym1.c, L.ym2.c);
    L.B = al_mul(&n, L.yp1.c, L.yp2.c);
    L.tt = al_mul(&n, p1[3], p2[3]);
    L.C = al_mul(&n, L.tt.c, L.d2);
    L.zz = al_mul(&n, p1[2], p2[2]);
    L.D = al_add(&n, L.zz.c, L.zz.c);
    L.E = al_sub(&n, L.B.c, L.A.c);
    L.F = al_sub(&n, L.D.c, L.C.c);
    L.G = al_add(&n, L.D.c, L.C.c);
    L.H = al_add(&n, L.B.c, L.A.c);
    L.X3 = al_mul(&n, L.E.c, L.F.c);
    L.Y3 = al_mul(&n, L.G.c, L.H.c);
    L.T3 = al_mul(&n, L.E.c, L.H.c);
    L.Z3 = al_mul(&n, L.F.c, L.G.c);
    L.end = n;
    return L;
}
/* point_double_layout_at (:476-505) */
static pdbl_l pdbl_at(size_t base, size_t x, size_t y, size_t z) {
    pdbl_l L;
    size_t n = base;
    L.p1[0] = x; L.p1[1] = y; L.p1[2] = z;
    L.sA = al_mul(&n, x, x);
    L.sB = al_mul(&n, y, y);
    L.szz = al_mul(&n, z, z);
    L.C = al_add(&n, L.szz.c, L.szz.c);
    L.H = al_add(&n, L.sA.c, L.sB.c);
    L.xy = al_add(&n, x, y);
    L.sxy = al_mul(&n, L.xy.c, L.xy.c);
    L.E = al_sub(&n, L.H.c, L.sxy.c);
    L.G = al_sub(&n, L.sA.c, L.sB.c);
    L.F = al_add(&n, L.C.c, L.G.c);
    L.X3 = al_mul(&n, L.E.c, L.F.c);
    L.Y3 = al_mul(&n, L.G.c, L.H.c);
    L.T3 = al_mul(&n, L.E.c, L.H.c);
    L.Z3 = al_mul(&n, L.F.c, L.G.c);
    L.end = n;
    return L;
}
/* cond_add_layout_at (:662-719) */
static cadd_l cadd_at(size_t base, const size_t acc[4], const size_t bp[4]) {
    cadd_l L;
    size_t n = base;
    L.bit = al(&n, 1);
    for (int i = 0; i < 4; i++) L.out[i] = al(&n, NL);
    L.add = padd_at(n, acc, bp);
    n = L.add.end;
    const size_t o3[4] = {L.add.X3.c, L.add.Y3.c, L.add.Z3.c, L.add.T3.c};
    for (int i = 0; i < 4; i++) {
        L.sel[i].a = acc[i];
        L.sel[i].b = o3[i];
        L.sel[i].bit = L.bit;
        L.sel[i].c = L.out[i];
    }
    L.end = n;
    return L;
}
/* point_decompress_layout_at (:887-955) */
static decomp_l decomp_at(size_t base, size_t y, size_t sign) {
    decomp_l L;
    size_t n = base;
    L.y = y;
    L.sign = sign;
    L.d = al(&n, NL);
    L.x = al(&n, NL);
    L.xb = al(&n, EB);
    L.sqx = al_mul(&n, L.x, L.x);
    L.sqy = al_mul(&n, y, y);
    L.dx2 = al_mul(&n, L.d, L.sqx.c);
    L.dxy = al_mul(&n, L.dx2.c, L.sqy.c);
    const size_t one = al(&n, NL);
    L.y2m1 = al_sub(&n, L.sqy.c, one);
    L.lhs = al_sub(&n, L.y2m1.c, L.sqx.c);
    L.eq = al_sub(&n, L.lhs.c, L.dxy.c);
    al(&n, 1); /* sign_match helper cell */
    L.end = n;
    return L;
}
/* scalar_mult_row_layout (ed25519_scalar_mult_air.rs:93-131) */
static smult_l smult_layout(void) {
    smult_l L;
    for (int i = 0; i < 4; i++) {
        L.base[i] = (size_t)i * NL;
        L.acc[i] = (size_t)(4 + i) * NL;
    }
    L.dbl = pdbl_at(8 * NL, L.acc[0], L.acc[1], L.acc[2]);
    const size_t o[4] = {L.dbl.X3.c, L.dbl.Y3.c, L.dbl.Z3.c, L.dbl.T3.c};
    L.ca = cadd_at(L.dbl.end, o, L.base);
    L.width = L.ca.end;
    return L;
}

/* ------------------------------------------------------------------ v16 layout */

enum { SHA_W = 1510, SHA_K = 1526, OFF_H0 = 1066, IN_L = 32, R_L = 16, Q_L = 17, LBITS = 16, P_L = 33, CBITS = 24 };

typedef struct ref_ed_ctx {
    /* public inputs */
    size_t k, msg_len;
    uint8_t *sha_in; /* R || A || M */
    uint8_t digest[64], k_can[32], r_comp[32], a_comp[32];
    uint8_t *s_bits, *k_bits; /* MSB-first, k each */
    /* rows */
    size_t n_blocks, k_hash, reduce_row, dec_r_row, dec_a_row, sb_first, ka_first, sb_last, ka_last;
    size_t residual_row, dbl_row[3], result_row, width, height, h7;
    /* columns */
    reduce_l red;
    size_t digest_bit_base;
    decomp_l dec;
    smult_l mult;
    int64_t bp_limbs[4][NL], a_limbs[4][NL], ry_limbs[NL], ay_limbs[NL];
    int r_sign, a_sign;
    size_t sblk_sb, sblk_ka, r_thread;
    size_t res_x, res_y, res_z, res_t, zero_c;
    sub_l sub_x, sub_t, sub_yz;
    size_t dbl_in[3];
    pdbl_l dbl;
    size_t th_sb[4], th_ka[4], th_r[4], one_r;
    mul_l mul_rt;
    size_t zero_r1;
    sub_l neg_rx, neg_rt;
    padd_l pa_r1;
    sub_l neg_kax, neg_kat;
    padd_l pa_r2;
    size_t n_cons;
    /* native values */
    ept R, A, sB, kA, res1, res2;
    int valid_r, valid_a;
} ref_ed_ctx;

enum {
    POINT_ADD_K = NL + 4 * (EC + 40) + 5 * (EC + 20) + 9 * (EC + NL + MCB * NL),
    MUL_K = EC + NL + MCB * NL, SUB_K = EC + 40, ADD_K = EC + 20,
    DBL_K = 8 * MUL_K + 4 * ADD_K + 2 * SUB_K,
    CADD_K = 4 * (1 + NL) + POINT_ADD_K,
    SMULT_K = DBL_K + CADD_K,
    DECOMP_K = NL + EB + NL + 4 * MUL_K + 3 * SUB_K + NL + 2,
    REDUCE_K = Q_L * LBITS + Q_L + 2 * (R_L * LBITS + R_L) + R_L + P_L + P_L * CBITS
};

size_t ref_ed25519_v16_constraints(size_t k) {
    size_t v2 = SHA_K + REDUCE_K + 560 + 2 * DECOMP_K;
    size_t v5 = v2 + 2 * SMULT_K + 2 * 80 + 2 * 40 + 80 + 2 * (NL + 1);
    size_t v7 = v5 + 4 * k + 3 * k;
    size_t v10 = v7 + NL + 3 * SUB_K + 3 * NL + DBL_K + 3 * NL + 4 * NL;
    size_t v13 = v10 + 80 + 80 + 4 * NL + NL + MUL_K + 4 * NL;
    return v13 + NL + 2 * SUB_K + POINT_ADD_K + 2 * SUB_K + POINT_ADD_K + 3 * NL;
}

static size_t next_pow2(size_t x) {
    size_t h = 1;
    while (h < x) h <<= 1;
    return h;
}

static void pt_limbs(ept p, int64_t out[4][NL]) {
    fe_limbs(p.X, out[0]);
    fe_limbs(p.Y, out[1]);
    fe_limbs(p.Z, out[2]);
    fe_limbs(p.T, out[3]);
}

/* verify_air_layout_v16 and everything fill_verify_air_v16 needs
 * (ed25519_verify_air.rs:162-4865; prover.rs:856-908 for the inputs).
 * k_scalar = k; s_bits = the low k bits of s MSB-first (prover.rs:826-836),
 * k_bits = r_thread_bits_for_kA(k mod L) (ed25519_verify_air.rs:1881-1893). */
ref_ed_ctx *ref_ed25519_ctx_new(const uint8_t pk[32], const uint8_t sig[64], const uint8_t *msg, size_t len,
                                size_t k) {
    ed_init();
    if (k == 0 || k > 256) return NULL;
    ref_ed_ctx *C = (ref_ed_ctx *)calloc(1, sizeof(ref_ed_ctx));
    C->k = k;
    C->msg_len = 64 + len;
    C->sha_in = (uint8_t *)malloc(64 + len + 1);
    memcpy(C->sha_in, sig, 32);
    memcpy(C->sha_in + 32, pk, 32);
    if (len) memcpy(C->sha_in + 64, msg, len);
    memcpy(C->r_comp, sig, 32);
    memcpy(C->a_comp, pk, 32);
    sha512(C->sha_in, C->msg_len, C->digest);
    reduce_mod_l_wide(C->digest, C->k_can);
    C->s_bits = (uint8_t *)malloc(k);
    C->k_bits = (uint8_t *)malloc(k);
    for (size_t i = 0; i < k; i++) {
        const size_t b = k - 1 - i;
        C->s_bits[i] = (sig[32 + b / 8] >> (b % 8)) & 1;
        C->k_bits[i] = (C->k_can[b / 8] >> (b % 8)) & 1;
    }
    C->valid_r = ept_decompress(C->r_comp, &C->R);
    C->valid_a = ept_decompress(C->a_comp, &C->A);
    if (!C->valid_r || !C->valid_a) {
        ref_ed25519_ctx_free(C);
        return NULL;
    }
    /* v0..v3 rows and columns */
    C->n_blocks = (C->msg_len + 1 + 16 + 127) / 128;
    C->k_hash = next_pow2(C->n_blocks * 128);
    if (C->k_hash < 128) C->k_hash = 128;
    C->reduce_row = C->k_hash;
    C->dec_r_row = C->k_hash + 1;
    C->dec_a_row = C->k_hash + 2;
    C->sb_first = C->k_hash + 3;
    C->ka_first = C->sb_first + k;
    C->sb_last = C->sb_first + k - 1;
    C->ka_last = C->ka_first + k - 1;
    size_t n = IN_L;
    C->red.in = 0;
    C->red.q = al(&n, Q_L);
    C->red.qb = al(&n, Q_L * LBITS);
    C->red.r = al(&n, R_L);
    C->red.rb = al(&n, R_L * LBITS);
    C->red.sl = al(&n, R_L);
    C->red.slb = al(&n, R_L * LBITS);
    C->red.cyb = al(&n, P_L * CBITS);
    const size_t reduce_w = n;
    C->digest_bit_base = SHA_W;
    C->dec = decomp_at(NL + 1, 0, NL);
    C->mult = smult_layout();
    size_t w = SHA_W + 512;
    if (reduce_w > w) w = reduce_w;
    if (C->dec.end > w) w = C->dec.end;
    if (C->mult.width > w) w = C->mult.width;
    size_t h3 = next_pow2(C->k_hash + 3 + 2 * k);
    C->h7 = h3; /* v3..v7 height: the v7 r thread is filled on these rows only */
    /* v4 / v5 constants */
    pt_limbs(ED_B, C->bp_limbs);
    pt_limbs(C->A, C->a_limbs);
    fe_limbs(C->R.Y, C->ry_limbs);
    fe_limbs(C->A.Y, C->ay_limbs);
    C->r_sign = (C->r_comp[31] >> 7) & 1;
    C->a_sign = (C->a_comp[31] >> 7) & 1;
    /* v6 / v7 */
    C->sblk_sb = w;
    C->sblk_ka = w + k;
    C->r_thread = w + 2 * k;
    n = w + 3 * k;
    /* v10 (on v7): result block, cofactor doubling */
    size_t h10_needed = C->ka_last + 4 + 1;
    size_t h = h3 >= h10_needed + 1 ? h3 : next_pow2(h10_needed);
    C->res_x = al(&n, NL);
    C->res_y = al(&n, NL);
    C->res_z = al(&n, NL);
    C->res_t = al(&n, NL);
    C->zero_c = al(&n, NL);
    C->sub_x = al_sub(&n, C->res_x, C->zero_c);
    C->sub_t = al_sub(&n, C->res_t, C->zero_c);
    C->sub_yz = al_sub(&n, C->res_y, C->res_z);
    for (int i = 0; i < 3; i++) C->dbl_in[i] = al(&n, NL);
    C->dbl = pdbl_at(n, C->dbl_in[0], C->dbl_in[1], C->dbl_in[2]);
    n = C->dbl.end;
    /* v11 .. v13: threads */
    for (int i = 0; i < 4; i++) C->th_sb[i] = al(&n, NL);
    for (int i = 0; i < 4; i++) C->th_ka[i] = al(&n, NL);
    for (int i = 0; i < 4; i++) C->th_r[i] = al(&n, NL);
    C->one_r = al(&n, NL);
    C->mul_rt = al_mul(&n, C->th_r[0], C->th_r[1]);
    /* v14: rows reordered, residual 1 */
    C->residual_row = C->ka_last + 1;
    for (int i = 0; i < 3; i++) C->dbl_row[i] = C->ka_last + 2 + i;
    C->result_row = C->ka_last + 5;
    C->zero_r1 = al(&n, NL);
    C->neg_rx = al_sub(&n, C->zero_r1, C->th_r[0]);
    C->neg_rt = al_sub(&n, C->zero_r1, C->th_r[3]);
    const size_t p2a[4] = {C->neg_rx.c, C->th_r[1], C->th_r[2], C->neg_rt.c};
    C->pa_r1 = padd_at(n, C->th_sb, p2a);
    n = C->pa_r1.end;
    if (!(h >= C->result_row + 2)) h = next_pow2(C->result_row + 1);
    /* v15: residual 2 */
    C->neg_kax = al_sub(&n, C->zero_r1, C->th_ka[0]);
    C->neg_kat = al_sub(&n, C->zero_r1, C->th_ka[3]);
    const size_t p1b[4] = {C->pa_r1.X3.c, C->pa_r1.Y3.c, C->pa_r1.Z3.c, C->pa_r1.T3.c};
    const size_t p2b[4] = {C->neg_kax.c, C->th_ka[1], C->th_ka[2], C->neg_kat.c};
    C->pa_r2 = padd_at(n, p1b, p2b);
    n = C->pa_r2.end;
    C->width = n;
    C->height = h;
    C->n_cons = ref_ed25519_v16_constraints(k);
    /* native values: sB, kA, residual_1, residual_2 (:4827-4862) */
    C->sB = ept_ladder(ED_B, C->s_bits, k);
    C->kA = ept_ladder(C->A, C->k_bits, k);
    ept negR = ept_neg(C->R);
    C->res1 = ept_add(C->sB, negR);
    C->res2 = ept_add(C->res1, ept_neg(C->kA));
    return C;
}

void ref_ed25519_ctx_free(ref_ed_ctx *C) {
    if (!C) return;
    free(C->sha_in);
    free(C->s_bits);
    free(C->k_bits);
    free(C);
}

size_t ref_ed25519_width(const ref_ed_ctx *C) { return C->width; }
size_t ref_ed25519_height(const ref_ed_ctx *C) { return C->height; }
size_t ref_ed25519_n_constraints(const ref_ed_ctx *C) { return C->n_cons; }
/* rows: k_hash, reduce, decompress R, decompress A, sB first, kA first, residual, dbl 1..3, result */
void ref_ed25519_rows(const ref_ed_ctx *C, size_t out[11]) {
    const size_t r[11] = {C->k_hash, C->reduce_row, C->dec_r_row, C->dec_a_row, C->sb_first, C->ka_first,
                          C->residual_row, C->dbl_row[0], C->dbl_row[1], C->dbl_row[2], C->result_row};
    memcpy(out, r, sizeof r);
}
/* named columns for tests: [0] digest bits, [1] reduce r bits, [2] decomp x limbs, [3] ladder acc x,
 * [4] cond-add out x, [5] scalar block sB, [6] scalar block kA, [7] r thread, [8] result X, [9] result Y,
 * [10] result Z, [11] result T, [12] thread sB X, [13] thread kA X, [14] thread R X, [15] thread R T,
 * [16] sub_X c, [17] sub_YZ c, [18] dbl input X, [19] point_add_R2 X3 c, [20] decomp sign, [21] mul carry bits of
 * the ladder's dbl sq_A, [22] ladder cond-add bit, [23] decomp x bits, [24] reduce q limbs, [25] reduce carry bits */
void ref_ed25519_columns(const ref_ed_ctx *C, size_t out[26]) {
    const size_t c[26] = {C->digest_bit_base, C->red.rb, C->dec.x, C->mult.acc[0], C->mult.ca.out[0], C->sblk_sb,
                          C->sblk_ka, C->r_thread, C->res_x, C->res_y, C->res_z, C->res_t, C->th_sb[0],
                          C->th_ka[0], C->th_r[0], C->th_r[3], C->sub_x.c, C->sub_yz.c, C->dbl_in[0],
                          C->pa_r2.X3.c, C->dec.sign, C->mult.dbl.sA.cyb, C->mult.ca.bit, C->dec.xb, C->red.q,
                          C->red.cyb};
    memcpy(out, c, sizeof c);
}
void ref_ed25519_public(const ref_ed_ctx *C, uint8_t digest[64], uint8_t k_can[32], uint8_t *s_bits,
                        uint8_t *k_bits) {
    memcpy(digest, C->digest, 64);
    memcpy(k_can, C->k_can, 32);
    memcpy(s_bits, C->s_bits, C->k);
    memcpy(k_bits, C->k_bits, C->k);
}
/* canonical bytes of the native points: R, A, sB, kA, residual_2, 8 residual_2 (x, y, z, t each) */
void ref_ed25519_points(const ref_ed_ctx *C, uint8_t out[6 * 128]) {
    ept p8 = ept_double(ept_double(ept_double(C->res2)));
    const ept *ps[6] = {&C->R, &C->A, &C->sB, &C->kA, &C->res2, &p8};
    for (int i = 0; i < 6; i++) {
        fe_to_bytes(ps[i]->X, out + 128 * i);
        fe_to_bytes(ps[i]->Y, out + 128 * i + 32);
        fe_to_bytes(ps[i]->Z, out + 128 * i + 64);
        fe_to_bytes(ps[i]->T, out + 128 * i + 96);
    }
}

/* ------------------------------------------------------------------ fills */

typedef struct { uint64_t *tr; size_t h; } tv;
#define CELL(t, c, r) ((t).tr[(size_t)(c) * (t).h + (r)])

static void put_limbs(tv t, size_t row, size_t base, const int64_t l[NL]) {
    for (int i = 0; i < NL; i++) CELL(t, base + i, row) = gl_from_u64((uint64_t)l[i]);
}
static void put_fe(tv t, size_t row, size_t base, fe x) {
    int64_t l[NL];
    fe_limbs(x, l);
    put_limbs(t, row, base, l);
}
/* c limbs and their bits at cb (the gadgets' element block) */
static void put_elem(tv t, size_t row, size_t c, size_t cb, const int64_t l[NL]) {
    size_t off = 0;
    for (int i = 0; i < NL; i++) {
        CELL(t, c + i, row) = (uint64_t)l[i];
        for (int b = 0; b < LW[i]; b++) CELL(t, cb + off + b, row) = (uint64_t)((l[i] >> b) & 1);
        off += LW[i];
    }
}

/* fill_add_gadget (ed25519_field_air.rs:330-429): c = freeze(a + b), carry chain by trial of the top carry */
static void fill_add(tv t, size_t row, const add_l *L, const int64_t a[NL], const int64_t b[NL]) {
    int64_t c[NL], cy[NL] = {0};
    fe_limbs(fe_add(fe_from_limbs(a), fe_from_limbs(b)), c);
    for (int cand = 0; cand <= 1; cand++) {
        int64_t prev = 19 * cand;
        int ok = 1;
        for (int k = 0; k < NL; k++) {
            const int64_t d = a[k] + b[k] + prev - c[k], rad = (int64_t)1 << LW[k];
            if (d == 0) { cy[k] = 0; prev = 0; }
            else if (d == rad) { cy[k] = 1; prev = 1; }
            else { ok = 0; break; }
        }
        if (ok && cy[NL - 1] == cand) break;
        memset(cy, 0, sizeof cy);
    }
    put_elem(t, row, L->c, L->cb, c);
    for (int k = 0; k < NL; k++) CELL(t, L->cy + k, row) = (uint64_t)cy[k];
}
/* fill_sub_gadget (:584-670): net carries in {-1, 0, 1} as (pos, neg) */
static void fill_sub(tv t, size_t row, const sub_l *L, const int64_t a[NL], const int64_t b[NL]) {
    int64_t c[NL], net[NL] = {0};
    fe_limbs(fe_sub(fe_from_limbs(a), fe_from_limbs(b)), c);
    const int64_t tops[3] = {0, 1, -1};
    for (int ti = 0; ti < 3; ti++) {
        int64_t prev = 19 * tops[ti];
        int ok = 1;
        for (int k = 0; k < NL; k++) {
            const int64_t rad = (int64_t)1 << LW[k], d = a[k] - b[k] - prev - c[k];
            if (d == 0) { net[k] = 0; prev = 0; }
            else if (d == -rad) { net[k] = 1; prev = 1; }
            else if (d == rad) { net[k] = -1; prev = -1; }
            else { ok = 0; break; }
        }
        if (ok && net[NL - 1] == tops[ti]) break;
        memset(net, 0, sizeof net);
    }
    put_elem(t, row, L->c, L->cb, c);
    for (int k = 0; k < NL; k++) {
        CELL(t, L->pos + k, row) = net[k] == 1;
        CELL(t, L->neg + k, row) = net[k] == -1;
    }
}
/* partial products of limb k: (i, j, factor) with i + j = k (factor 2 if both odd)
 * or i + j = k + 10 (19, or 38 if both odd) (ed25519_field_air.rs:847-865) */
static __int128 mul_tk(const int64_t a[NL], const int64_t b[NL], int k) {
    __int128 s = 0;
    for (int i = 0; i < NL; i++)
        for (int j = 0; j < NL; j++) {
            const int odd = (i & 1) && (j & 1);
            if (i + j == k) s += (__int128)(odd ? 2 : 1) * a[i] * b[j];
            else if (i + j == k + NL) s += (__int128)(odd ? 38 : 19) * a[i] * b[j];
        }
    return s;
}
static __int128 rem_euclid(__int128 x, __int128 m) { __int128 r = x % m; return r < 0 ? r + m : r; }
static __int128 div_euclid(__int128 x, __int128 m) { return (x - rem_euclid(x, m)) / m; }
/* fill_mul_gadget (:874-998): the top carry searched as base + k 2^26, k in [-1024, 1024) */
static void fill_mul(tv t, size_t row, const mul_l *L, const int64_t a[NL], const int64_t b[NL]) {
    int64_t c[NL];
    fe_limbs(fe_mul(fe_from_limbs(a), fe_from_limbs(b)), c);
    __int128 tk[NL], cy[NL] = {0};
    for (int k = 0; k < NL; k++) tk[k] = mul_tk(a, b, k);
    const __int128 r0 = (__int128)1 << 26;
    __int128 base = rem_euclid(rem_euclid((__int128)c[0] - tk[0], r0) * 7064091, r0);
    if (base >= ((__int128)1 << 25)) base -= r0;
    for (__int128 ko = -1024; ko < 1024; ko++) {
        const __int128 guess = base + ko * r0;
        __int128 prev = 19 * guess, tent[NL];
        int ok = 1;
        for (int k = 0; k < NL; k++) {
            const __int128 rad = (__int128)1 << LW[k], d = tk[k] + prev - c[k];
            if (rem_euclid(d, rad) != 0) { ok = 0; break; }
            tent[k] = div_euclid(d, rad);
            prev = tent[k];
        }
        if (ok && tent[NL - 1] == guess) {
            memcpy(cy, tent, sizeof cy);
            break;
        }
    }
    put_elem(t, row, L->c, L->cb, c);
    for (int k = 0; k < NL; k++) {
        const uint64_t biased = (uint64_t)(cy[k] + ((__int128)1 << (MCB - 1)));
        for (int bb = 0; bb < MCB; bb++) CELL(t, L->cyb + (size_t)k * MCB + bb, row) = (biased >> bb) & 1;
    }
}
static void fill_sel(tv t, size_t row, const sel_l *L, const int64_t a[NL], const int64_t b[NL], int bit) {
    CELL(t, L->bit, row) = (uint64_t)bit;
    put_limbs(t, row, L->c, bit ? b : a);
}
typedef struct { int64_t x[NL], y[NL], z[NL], t[NL]; } plimbs;
static plimbs plimbs_of(ept p) {
    plimbs r;
    fe_limbs(p.X, r.x);
    fe_limbs(p.Y, r.y);
    fe_limbs(p.Z, r.z);
    fe_limbs(p.T, r.t);
    return r;
}
static void lim(fe x, int64_t l[NL]) { fe_limbs(x, l); }
/* fill_point_add_gadget (ed25519_group_air.rs:251-321) */
static void fill_padd(tv t, size_t row, const padd_l *L, ept p1, ept p2) {
    put_limbs(t, row, L->d2, D2_LIMBS);
    fe ym1 = fe_sub(p1.Y, p1.X), ym2 = fe_sub(p2.Y, p2.X), yp1 = fe_add(p1.Y, p1.X), yp2 = fe_add(p2.Y, p2.X);
    fe A = fe_mul(ym1, ym2), B = fe_mul(yp1, yp2), tt = fe_mul(p1.T, p2.T), Cc = fe_mul(tt, ED_D2);
    fe zz = fe_mul(p1.Z, p2.Z), Dd = fe_add(zz, zz), E = fe_sub(B, A), F = fe_sub(Dd, Cc), G = fe_add(Dd, Cc),
       H = fe_add(B, A);
    plimbs q1 = plimbs_of(p1), q2 = plimbs_of(p2);
    int64_t l_ym1[NL], l_ym2[NL], l_yp1[NL], l_yp2[NL], lA[NL], lB[NL], ltt[NL], lC[NL], lzz[NL], lD[NL], lE[NL],
        lF[NL], lG[NL], lH[NL];
    lim(ym1, l_ym1); lim(ym2, l_ym2); lim(yp1, l_yp1); lim(yp2, l_yp2); lim(A, lA); lim(B, lB); lim(tt, ltt);
    lim(Cc, lC); lim(zz, lzz); lim(Dd, lD); lim(E, lE); lim(F, lF); lim(G, lG); lim(H, lH);
    fill_sub(t, row, &L->ym1, q1.y, q1.x);
    fill_sub(t, row, &L->ym2, q2.y, q2.x);
    fill_add(t, row, &L->yp1, q1.y, q1.x);
    fill_add(t, row, &L->yp2, q2.y, q2.x);
    fill_mul(t, row, &L->A, l_ym1, l_ym2);
    fill_mul(t, row, &L->B, l_yp1, l_yp2);
    fill_mul(t, row, &L->tt, q1.t, q2.t);
    fill_mul(t, row, &L->C, ltt, D2_LIMBS);
    fill_mul(t, row, &L->zz, q1.z, q2.z);
    fill_add(t, row, &L->D, lzz, lzz);
    fill_sub(t, row, &L->E, lB, lA);
    fill_sub(t, row, &L->F, lD, lC);
    fill_add(t, row, &L->G, lD, lC);
    fill_add(t, row, &L->H, lB, lA);
    fill_mul(t, row, &L->X3, lE, lF);
    fill_mul(t, row, &L->Y3, lG, lH);
    fill_mul(t, row, &L->T3, lE, lH);
    fill_mul(t, row, &L->Z3, lF, lG);
}
/* fill_point_double_gadget (:507-546) */
static void fill_pdbl(tv t, size_t row, const pdbl_l *L, ept p) {
    fe A = fe_mul(p.X, p.X), B = fe_mul(p.Y, p.Y), zz = fe_mul(p.Z, p.Z), Cc = fe_add(zz, zz), H = fe_add(A, B);
    fe xy = fe_add(p.X, p.Y), xys = fe_mul(xy, xy), E = fe_sub(H, xys), G = fe_sub(A, B), F = fe_add(Cc, G);
    plimbs q = plimbs_of(p);
    int64_t lA[NL], lB[NL], lzz[NL], lC[NL], lH[NL], lxy[NL], lxys[NL], lE[NL], lG[NL], lF[NL];
    lim(A, lA); lim(B, lB); lim(zz, lzz); lim(Cc, lC); lim(H, lH); lim(xy, lxy); lim(xys, lxys); lim(E, lE);
    lim(G, lG); lim(F, lF);
    fill_mul(t, row, &L->sA, q.x, q.x);
    fill_mul(t, row, &L->sB, q.y, q.y);
    fill_mul(t, row, &L->szz, q.z, q.z);
    fill_add(t, row, &L->C, lzz, lzz);
    fill_add(t, row, &L->H, lA, lB);
    fill_add(t, row, &L->xy, q.x, q.y);
    fill_mul(t, row, &L->sxy, lxy, lxy);
    fill_sub(t, row, &L->E, lH, lxys);
    fill_sub(t, row, &L->G, lA, lB);
    fill_add(t, row, &L->F, lC, lG);
    fill_mul(t, row, &L->X3, lE, lF);
    fill_mul(t, row, &L->Y3, lG, lH);
    fill_mul(t, row, &L->T3, lE, lH);
    fill_mul(t, row, &L->Z3, lF, lG);
}
/* fill_cond_add_gadget (:721-753) */
static void fill_cadd(tv t, size_t row, const cadd_l *L, ept acc, ept base, int bit) {
    fill_padd(t, row, &L->add, acc, base);
    plimbs s = plimbs_of(ept_add(acc, base)), a = plimbs_of(acc);
    fill_sel(t, row, &L->sel[0], a.x, s.x, bit);
    fill_sel(t, row, &L->sel[1], a.y, s.y, bit);
    fill_sel(t, row, &L->sel[2], a.z, s.z, bit);
    fill_sel(t, row, &L->sel[3], a.t, s.t, bit);
}
/* fill_point_decompress_gadget (:957-1029), y and sign from split_compressed */
static void fill_decomp(tv t, size_t row, const decomp_l *L, fe y, int sign, fe x) {
    put_limbs(t, row, L->d, D_LIMBS);
    const size_t one_base = L->y2m1.b;
    int64_t one[NL] = {1, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    put_limbs(t, row, one_base, one);
    int64_t lx[NL], ly[NL];
    fe_limbs(x, lx);
    fe_limbs(y, ly);
    put_elem(t, row, L->x, L->xb, lx);
    CELL(t, L->sign, row) = (uint64_t)sign;
    fe x2 = fe_mul(x, x), y2 = fe_mul(y, y), dx2 = fe_mul(ED_D, x2), dxy = fe_mul(dx2, y2);
    fe y2m1 = fe_sub(y2, fe_one()), lhs = fe_sub(y2m1, x2);
    int64_t lx2[NL], ly2[NL], ldx2[NL], ldxy[NL], ly2m1[NL], llhs[NL];
    lim(x2, lx2); lim(y2, ly2); lim(dx2, ldx2); lim(dxy, ldxy); lim(y2m1, ly2m1); lim(lhs, llhs);
    fill_mul(t, row, &L->sqx, lx, lx);
    fill_mul(t, row, &L->sqy, ly, ly);
    fill_mul(t, row, &L->dx2, D_LIMBS, lx2);
    fill_mul(t, row, &L->dxy, ldx2, ly2);
    fill_sub(t, row, &L->y2m1, ly2, one);
    fill_sub(t, row, &L->lhs, ly2m1, lx2);
    fill_sub(t, row, &L->eq, llhs, ldxy);
}
/* fill_scalar_reduce_gadget (ed25519_scalar_air.rs:168-212) */
static void fill_reduce(tv t, size_t row, const reduce_l *L, const uint8_t in[64]) {
    uint64_t q[5], r[4];
    divmod_l(in, q, r);
    int64_t ql[Q_L], rl[R_L], sl[R_L], lfull[R_L], inl[P_L] = {0};
    for (int i = 0; i < Q_L; i++) ql[i] = (int64_t)((q[i / 4] >> (16 * (i % 4))) & 0xFFFF);
    for (int i = 0; i < R_L; i++) {
        rl[i] = (int64_t)((r[i / 4] >> (16 * (i % 4))) & 0xFFFF);
        lfull[i] = (int64_t)((LL[i / 4] >> (16 * (i % 4))) & 0xFFFF);
    }
    int64_t br = 0;
    for (int i = 0; i < R_L; i++) { /* slack = (L - 1) - r */
        int64_t v = (i == 0 ? lfull[0] - 1 : lfull[i]) - rl[i] - br;
        if (v < 0) { sl[i] = v + 65536; br = 1; }
        else { sl[i] = v; br = 0; }
    }
    for (int i = 0; i < IN_L; i++) inl[i] = (int64_t)in[2 * i] | (int64_t)in[2 * i + 1] << 8;
    int64_t prev = 0;
    for (int k = 0; k < P_L; k++) {
        int64_t prod = 0;
        for (int i = 0; i < Q_L; i++)
            if (k >= i && k - i < R_L) prod += ql[i] * lfull[k - i];
        const int64_t tot = inl[k] - prod - (k < R_L ? rl[k] : 0) + prev;
        const int64_t cout = (tot - (((tot % 65536) + 65536) % 65536)) / 65536;
        const uint64_t biased = (uint64_t)(cout + (1 << (CBITS - 1)));
        for (int b = 0; b < CBITS; b++) CELL(t, L->cyb + (size_t)k * CBITS + b, row) = (biased >> b) & 1;
        prev = cout;
    }
    for (int i = 0; i < Q_L; i++) {
        CELL(t, L->q + i, row) = (uint64_t)ql[i];
        for (int b = 0; b < LBITS; b++) CELL(t, L->qb + (size_t)i * LBITS + b, row) = (uint64_t)(ql[i] >> b) & 1;
    }
    for (int i = 0; i < R_L; i++) {
        CELL(t, L->r + i, row) = (uint64_t)rl[i];
        CELL(t, L->sl + i, row) = (uint64_t)sl[i];
        for (int b = 0; b < LBITS; b++) {
            CELL(t, L->rb + (size_t)i * LBITS + b, row) = (uint64_t)(rl[i] >> b) & 1;
            CELL(t, L->slb + (size_t)i * LBITS + b, row) = (uint64_t)(sl[i] >> b) & 1;
        }
    }
}
/* fill_scalar_mult_trace (ed25519_scalar_mult_air.rs:133-172) at rows first .. first + k */
static void fill_ladder(tv t, size_t first, const smult_l *L, ept base, const uint8_t *bits, size_t k) {
    ept acc = ept_identity();
    plimbs b = plimbs_of(base);
    for (size_t j = 0; j < k; j++) {
        const size_t row = first + j;
        plimbs a = plimbs_of(acc);
        put_limbs(t, row, L->base[0], b.x);
        put_limbs(t, row, L->base[1], b.y);
        put_limbs(t, row, L->base[2], b.z);
        put_limbs(t, row, L->base[3], b.t);
        put_limbs(t, row, L->acc[0], a.x);
        put_limbs(t, row, L->acc[1], a.y);
        put_limbs(t, row, L->acc[2], a.z);
        put_limbs(t, row, L->acc[3], a.t);
        fill_pdbl(t, row, &L->dbl, acc);
        ept d = ept_double(acc);
        fill_cadd(t, row, &L->ca, d, base, bits[j]);
        acc = bits[j] ? ept_add(d, base) : d;
    }
}

/* fill_verify_air_v16 (ed25519_verify_air.rs:4820-4865) = the v7 trace (v0..v7 fills) plus the
 * v14 threads / residual / doubling / result rows and the v15 residual-2 gadgets. tr: width x height. */
int ref_ed25519_fill(const ref_ed_ctx *C, uint64_t *tr) {
    tv t = {tr, C->height};
    memset(tr, 0, 8 * C->width * C->height);
    /* v0: SHA-512 trace rows 0 .. k_hash (its height is k_hash) */
    const size_t hs = ref_sha512_trace_height(C->msg_len);
    uint64_t *sha = (uint64_t *)malloc(8 * SHA_W * hs);
    uint8_t dg[64];
    ref_build_sha512_trace(C->sha_in, C->msg_len, sha, dg);
    for (size_t c = 0; c < SHA_W; c++)
        for (size_t r = 0; r < hs && r < C->k_hash; r++) CELL(t, c, r) = sha[c * hs + r];
    free(sha);
    /* v1: digest bits at row k_hash - 1 (big-endian words, bit i of word k) */
    for (int k = 0; k < 8; k++) {
        uint64_t hw = 0;
        for (int i = 0; i < 8; i++) hw = hw << 8 | C->digest[8 * k + i];
        for (int i = 0; i < 64; i++) CELL(t, C->digest_bit_base + 64 * k + i, C->k_hash - 1) = (hw >> i) & 1;
    }
    /* v0: scalar reduce at reduce_row, input = 32 LE 16-bit limbs of the digest */
    for (int i = 0; i < IN_L; i++)
        CELL(t, C->red.in + i, C->reduce_row) = (uint64_t)C->digest[2 * i] | (uint64_t)C->digest[2 * i + 1] << 8;
    fill_reduce(t, C->reduce_row, &C->red, C->digest);
    /* v2: decompress R and A (y limbs, sign, gadget) */
    const struct { size_t row; const uint8_t *comp; const ept *p; } dr[2] = {
        {C->dec_r_row, C->r_comp, &C->R}, {C->dec_a_row, C->a_comp, &C->A}};
    for (int i = 0; i < 2; i++) {
        const int sign = (dr[i].comp[31] >> 7) & 1;
        put_fe(t, dr[i].row, C->dec.y, dr[i].p->Y);
        CELL(t, C->dec.sign, dr[i].row) = (uint64_t)sign;
        fill_decomp(t, dr[i].row, &C->dec, dr[i].p->Y, sign, dr[i].p->X);
    }
    /* v3: the two ladders */
    fill_ladder(t, C->sb_first, &C->mult, ED_B, C->s_bits, C->k);
    fill_ladder(t, C->ka_first, &C->mult, C->A, C->k_bits, C->k);
    /* v6: shifting scalar blocks */
    for (size_t j = 0; j < C->k; j++)
        for (size_t i = 0; i < C->k; i++) {
            CELL(t, C->sblk_sb + i, C->sb_first + j) = j + i < C->k ? C->s_bits[j + i] : 0;
            CELL(t, C->sblk_ka + i, C->ka_first + j) = j + i < C->k ? C->k_bits[j + i] : 0;
        }
    /* v7: r thread (the k bits, MSB-first) on every row */
    for (size_t r = 0; r < C->h7; r++)
        for (size_t i = 0; i < C->k; i++) CELL(t, C->r_thread + i, r) = C->k_bits[i];
    /* v14: threads on every row */
    const ept rthr = {C->R.X, C->R.Y, fe_one(), fe_mul(C->R.X, C->R.Y)};
    const ept *th[3] = {&C->sB, &C->kA, &rthr};
    const size_t *tb[3] = {C->th_sb, C->th_ka, C->th_r};
    for (int q = 0; q < 3; q++) {
        plimbs pl = plimbs_of(*th[q]);
        for (size_t r = 0; r < C->height; r++) {
            put_limbs(t, r, tb[q][0], pl.x);
            put_limbs(t, r, tb[q][1], pl.y);
            put_limbs(t, r, tb[q][2], pl.z);
            put_limbs(t, r, tb[q][3], pl.t);
        }
    }
    int64_t one[NL] = {1, 0, 0, 0, 0, 0, 0, 0, 0, 0}, zero[NL] = {0};
    put_limbs(t, C->dec_r_row, C->one_r, one);
    plimbs rl = plimbs_of(rthr), kl = plimbs_of(C->kA);
    fill_mul(t, C->dec_r_row, &C->mul_rt, rl.x, rl.y);
    /* v14: residual 1 = sB + (-R) */
    const size_t rr = C->residual_row;
    put_limbs(t, rr, C->zero_r1, zero);
    fill_sub(t, rr, &C->neg_rx, zero, rl.x);
    fill_sub(t, rr, &C->neg_rt, zero, rl.t);
    fill_padd(t, rr, &C->pa_r1, C->sB, ept_neg(C->R));
    /* v14: cofactor doublings of residual_2 and the result row */
    ept p = C->res2;
    for (int i = 0; i < 3; i++) {
        const size_t row = C->dbl_row[i];
        put_fe(t, row, C->dbl_in[0], p.X);
        put_fe(t, row, C->dbl_in[1], p.Y);
        put_fe(t, row, C->dbl_in[2], p.Z);
        fill_pdbl(t, row, &C->dbl, p);
        p = ept_double(p);
    }
    plimbs p8 = plimbs_of(p);
    const size_t rs = C->result_row;
    put_limbs(t, rs, C->res_x, p8.x);
    put_limbs(t, rs, C->res_y, p8.y);
    put_limbs(t, rs, C->res_z, p8.z);
    put_limbs(t, rs, C->res_t, p8.t);
    put_limbs(t, rs, C->zero_c, zero);
    fill_sub(t, rs, &C->sub_x, p8.x, zero);
    fill_sub(t, rs, &C->sub_t, p8.t, zero);
    fill_sub(t, rs, &C->sub_yz, p8.y, p8.z);
    /* v15: residual 2 = residual 1 + (-kA) */
    fill_sub(t, rr, &C->neg_kax, zero, kl.x);
    fill_sub(t, rr, &C->neg_kat, zero, kl.t);
    fill_padd(t, rr, &C->pa_r2, C->res1, ept_neg(C->kA));
    return 0;
}

/* ------------------------------------------------------------------ evaluators */

typedef struct { uint64_t *o; size_t j; } ob;
static void push(ob *b, uint64_t v) { b->o[b->j++] = v; }
static void zeros(ob *b, size_t n) { memset(b->o + b->j, 0, 8 * n); b->j += n; }
static uint64_t boolc(uint64_t x) { return gl_mul(x, gl_sub(1, x)); }
static uint64_t fsm(int64_t x) { return x >= 0 ? gl_from_u64((uint64_t)x) : gl_neg(gl_from_u64((uint64_t)(-x))); }

/* element block range check at (c, cb): 255 booleanity, then 10 packings (e.g. :431-453) */
static void ev_elem(ob *b, const uint64_t *cur, size_t c, size_t cb) {
    for (size_t i = 0; i < EB; i++) push(b, boolc(cur[cb + i]));
    size_t off = 0;
    for (int i = 0; i < NL; i++) {
        uint64_t s = 0;
        for (int k = 0; k < LW[i]; k++) s = gl_add(s, gl_mul(1ULL << k, cur[cb + off + k]));
        push(b, gl_sub(cur[c + i], s));
        off += LW[i];
    }
}
/* eval_add_gadget (ed25519_field_air.rs:431-484) */
static void ev_add(ob *b, const uint64_t *cur, const add_l *L) {
    ev_elem(b, cur, L->c, L->cb);
    for (int k = 0; k < NL; k++) {
        const uint64_t cin = k == 0 ? gl_mul(19, cur[L->cy + NL - 1]) : cur[L->cy + k - 1];
        push(b, gl_sub(gl_sub(gl_add(gl_add(cur[L->a + k], cur[L->b + k]), cin), cur[L->c + k]),
                       gl_mul(1ULL << LW[k], cur[L->cy + k])));
    }
    for (int k = 0; k < NL; k++) push(b, boolc(cur[L->cy + k]));
}
/* eval_sub_gadget (:672-748) */
static void ev_sub(ob *b, const uint64_t *cur, const sub_l *L) {
    ev_elem(b, cur, L->c, L->cb);
#define NET(k) gl_sub(cur[L->pos + (k)], cur[L->neg + (k)])
    for (int k = 0; k < NL; k++) {
        const uint64_t nin = k == 0 ? gl_mul(19, NET(NL - 1)) : NET(k - 1);
        push(b, gl_add(gl_sub(gl_sub(gl_sub(cur[L->a + k], cur[L->b + k]), nin), cur[L->c + k]),
                       gl_mul(1ULL << LW[k], NET(k))));
    }
#undef NET
    for (int k = 0; k < NL; k++) push(b, boolc(cur[L->pos + k]));
    for (int k = 0; k < NL; k++) push(b, boolc(cur[L->neg + k]));
    for (int k = 0; k < NL; k++) push(b, gl_mul(cur[L->pos + k], cur[L->neg + k]));
}
/* eval_mul_gadget (:1000-1076) */
static uint64_t mul_carry(const uint64_t *cur, const mul_l *L, int k) {
    uint64_t s = 0;
    for (int bb = 0; bb < MCB; bb++) s = gl_add(s, gl_mul(1ULL << bb, cur[L->cyb + (size_t)k * MCB + bb]));
    return gl_sub(s, 1ULL << (MCB - 1));
}
static void ev_mul(ob *b, const uint64_t *cur, const mul_l *L) {
    ev_elem(b, cur, L->c, L->cb);
    for (int k = 0; k < NL; k++) {
        uint64_t tk = 0;
        for (int i = 0; i < NL; i++)
            for (int j = 0; j < NL; j++) {
                const int odd = (i & 1) && (j & 1);
                uint64_t f = 0;
                if (i + j == k) f = odd ? 2 : 1;
                else if (i + j == k + NL) f = odd ? 38 : 19;
                if (f) tk = gl_add(tk, gl_mul(f, gl_mul(cur[L->a + i], cur[L->b + j])));
            }
        const uint64_t cin = k == 0 ? gl_mul(19, mul_carry(cur, L, NL - 1)) : mul_carry(cur, L, k - 1);
        push(b, gl_sub(gl_sub(gl_add(tk, cin), cur[L->c + k]), gl_mul(1ULL << LW[k], mul_carry(cur, L, k))));
    }
    for (size_t i = 0; i < (size_t)MCB * NL; i++) push(b, boolc(cur[L->cyb + i]));
}
/* eval_select_gadget (:1457-1477) */
static void ev_sel(ob *b, const uint64_t *cur, const sel_l *L) {
    const uint64_t bit = cur[L->bit];
    push(b, boolc(bit));
    for (int k = 0; k < NL; k++) {
        const uint64_t a = cur[L->a + k], bb = cur[L->b + k];
        push(b, gl_sub(gl_sub(cur[L->c + k], a), gl_mul(bit, gl_sub(bb, a))));
    }
}
/* eval_point_add_gadget (ed25519_group_air.rs:334-381) */
static void ev_padd(ob *b, const uint64_t *cur, const padd_l *L) {
    for (int k = 0; k < NL; k++) push(b, gl_sub(cur[L->d2 + k], gl_from_u64((uint64_t)D2_LIMBS[k])));
    ev_sub(b, cur, &L->ym1);
    ev_sub(b, cur, &L->ym2);
    ev_add(b, cur, &L->yp1);
    ev_add(b, cur, &L->yp2);
    ev_mul(b, cur, &L->A);
    ev_mul(b, cur, &L->B);
    ev_mul(b, cur, &L->tt);
    ev_mul(b, cur, &L->C);
    ev_mul(b, cur, &L->zz);
    ev_add(b, cur, &L->D);
    ev_sub(b, cur, &L->E);
    ev_sub(b, cur, &L->F);
    ev_add(b, cur, &L->G);
    ev_add(b, cur, &L->H);
    ev_mul(b, cur, &L->X3);
    ev_mul(b, cur, &L->Y3);
    ev_mul(b, cur, &L->T3);
    ev_mul(b, cur, &L->Z3);
}
/* eval_point_double_gadget (:548-574) */
static void ev_pdbl(ob *b, const uint64_t *cur, const pdbl_l *L) {
    ev_mul(b, cur, &L->sA);
    ev_mul(b, cur, &L->sB);
    ev_mul(b, cur, &L->szz);
    ev_add(b, cur, &L->C);
    ev_add(b, cur, &L->H);
    ev_add(b, cur, &L->xy);
    ev_mul(b, cur, &L->sxy);
    ev_sub(b, cur, &L->E);
    ev_sub(b, cur, &L->G);
    ev_add(b, cur, &L->F);
    ev_mul(b, cur, &L->X3);
    ev_mul(b, cur, &L->Y3);
    ev_mul(b, cur, &L->T3);
    ev_mul(b, cur, &L->Z3);
}
/* eval_point_decompress_gadget (:1031-1090) */
static void ev_decomp(ob *b, const uint64_t *cur, const decomp_l *L) {
    for (int k = 0; k < NL; k++) push(b, gl_sub(cur[L->d + k], gl_from_u64((uint64_t)D_LIMBS[k])));
    ev_elem(b, cur, L->x, L->xb);
    ev_mul(b, cur, &L->sqx);
    ev_mul(b, cur, &L->sqy);
    ev_mul(b, cur, &L->dx2);
    ev_mul(b, cur, &L->dxy);
    ev_sub(b, cur, &L->y2m1);
    ev_sub(b, cur, &L->lhs);
    ev_sub(b, cur, &L->eq);
    for (int k = 0; k < NL; k++) push(b, cur[L->eq.c + k]);
    const uint64_t sb = cur[L->sign];
    push(b, boolc(sb));
    push(b, gl_sub(cur[L->xb], sb));
}
/* eval_scalar_reduce_gadget (ed25519_scalar_air.rs:412-584) */
static void ev_range16(ob *b, const uint64_t *cur, size_t limbs, size_t bits, int n) {
    for (int i = 0; i < n * LBITS; i++) push(b, boolc(cur[bits + (size_t)i]));
    for (int i = 0; i < n; i++) {
        uint64_t s = 0;
        for (int k = 0; k < LBITS; k++) s = gl_add(s, gl_mul(1ULL << k, cur[bits + (size_t)i * LBITS + k]));
        push(b, gl_sub(cur[limbs + i], s));
    }
}
static uint64_t red_carry(const uint64_t *cur, const reduce_l *L, int k) {
    uint64_t s = 0;
    for (int bb = 0; bb < CBITS; bb++) s = gl_add(s, gl_mul(1ULL << bb, cur[L->cyb + (size_t)k * CBITS + bb]));
    return gl_sub(s, 1ULL << (CBITS - 1));
}
static void ev_reduce(ob *b, const uint64_t *cur, const reduce_l *L) {
    ev_range16(b, cur, L->q, L->qb, Q_L);
    ev_range16(b, cur, L->r, L->rb, R_L);
    ev_range16(b, cur, L->sl, L->slb, R_L);
    zeros(b, R_L);
    for (int k = 0; k < P_L; k++) {
        const uint64_t in_k = k < IN_L ? cur[L->in + k] : 0;
        uint64_t s = 0;
        for (int i = 0; i < Q_L; i++)
            if (k >= i && k - i < R_L) {
                const uint64_t lj = (LL[(k - i) / 4] >> (16 * ((k - i) % 4))) & 0xFFFF;
                s = gl_add(s, gl_mul(cur[L->q + i], lj));
            }
        const uint64_t r_k = k < R_L ? cur[L->r + k] : 0;
        const uint64_t cin = k == 0 ? 0 : red_carry(cur, L, k - 1);
        push(b, gl_sub(gl_add(gl_sub(gl_sub(in_k, s), r_k), cin), gl_mul(1ULL << LBITS, red_carry(cur, L, k))));
    }
    for (size_t i = 0; i < (size_t)P_L * CBITS; i++) push(b, boolc(cur[L->cyb + i]));
}
static void ev_pins(ob *b, const uint64_t *cur, size_t base, const int64_t l[NL]) {
    for (int k = 0; k < NL; k++) push(b, gl_sub(cur[base + k], fsm(l[k])));
}
static void ev_bind(ob *b, const uint64_t *x, size_t xb, const uint64_t *y, size_t yb, size_t n) {
    for (size_t k = 0; k < n; k++) push(b, gl_sub(x[xb + k], y[yb + k]));
}

/* eval_verify_air_v16_per_row: v7 (v0..v7) then the v10..v16 blocks (ed25519_verify_air.rs:273-4958) */
void ref_ed25519_constraints(const ref_ed_ctx *C, const uint64_t *cur, const uint64_t *nxt, size_t row,
                             uint64_t *out) {
    ob b = {out, 0};
    const size_t k = C->k;
    const smult_l *M = &C->mult;
    /* v0 */
    if (row + 1 < C->k_hash) {
        ref_sha512_constraints(cur, nxt, row, C->n_blocks, out + b.j);
        b.j += SHA_K;
    } else zeros(&b, SHA_K);
    if (row == C->reduce_row) ev_reduce(&b, cur, &C->red);
    else zeros(&b, REDUCE_K);
    /* v1: digest format at row k_hash - 1 */
    if (row + 1 == C->k_hash) {
        const size_t db = C->digest_bit_base;
        for (int i = 0; i < 512; i++) push(&b, boolc(cur[db + i]));
        for (int w = 0; w < 8; w++) {
            uint64_t lo = 0, hi = 0;
            for (int i = 0; i < 32; i++) {
                lo = gl_add(lo, gl_mul(1ULL << i, cur[db + 64 * w + i]));
                hi = gl_add(hi, gl_mul(1ULL << i, cur[db + 64 * w + 32 + i]));
            }
            push(&b, gl_sub(cur[OFF_H0 + 2 * w], lo));
            push(&b, gl_sub(cur[OFF_H0 + 2 * w + 1], hi));
        }
        for (int w = 0; w < 8; w++)
            for (int m = 0; m < 4; m++) {
                const int lob = 56 - 16 * m, hib = lob - 8;
                uint64_t s = 0;
                for (int j = 0; j < 8; j++) s = gl_add(s, gl_mul(1ULL << j, cur[db + 64 * w + lob + j]));
                for (int j = 0; j < 8; j++) s = gl_add(s, gl_mul(1ULL << (8 + j), cur[db + 64 * w + hib + j]));
                push(&b, gl_sub(nxt[C->red.in + 4 * w + m], s));
            }
    } else zeros(&b, 560);
    /* v2 */
    if (row == C->dec_r_row) ev_decomp(&b, cur, &C->dec);
    else zeros(&b, DECOMP_K);
    if (row == C->dec_a_row) ev_decomp(&b, cur, &C->dec);
    else zeros(&b, DECOMP_K);
    /* v3: the two ladders */
    const size_t firsts[2] = {C->sb_first, C->ka_first};
    for (int q = 0; q < 2; q++) {
        const size_t f = firsts[q], l = f + k - 1;
        if (row >= f && row <= l) {
            ev_pdbl(&b, cur, &M->dbl);
            ev_padd(&b, cur, &M->ca.add);
            for (int s = 0; s < 4; s++) ev_sel(&b, cur, &M->ca.sel[s]);
        } else zeros(&b, SMULT_K);
        if (row >= f && row < l) {
            for (int kk = 0; kk < NL; kk++)
                for (int s = 0; s < 4; s++) push(&b, gl_sub(nxt[M->base[s] + kk], cur[M->base[s] + kk]));
            for (int kk = 0; kk < NL; kk++)
                for (int s = 0; s < 4; s++) push(&b, gl_sub(nxt[M->acc[s] + kk], cur[M->ca.out[s] + kk]));
        } else zeros(&b, 8 * NL);
        if (row == f) {
            const int64_t idl[4][NL] = {{0}, {1}, {1}, {0}};
            for (int kk = 0; kk < NL; kk++)
                for (int s = 0; s < 4; s++) push(&b, gl_sub(cur[M->acc[s] + kk], fsm(idl[s][kk])));
        } else zeros(&b, 4 * NL);
    }
    /* v4: base pins at the ladder starts */
    if (row == C->sb_first)
        for (int s = 0; s < 4; s++) ev_pins(&b, cur, M->base[s], C->bp_limbs[s]);
    else zeros(&b, 4 * NL);
    if (row == C->ka_first)
        for (int s = 0; s < 4; s++) ev_pins(&b, cur, M->base[s], C->a_limbs[s]);
    else zeros(&b, 4 * NL);
    /* v5: y and sign pins */
    if (row == C->dec_r_row) {
        ev_pins(&b, cur, C->dec.y, C->ry_limbs);
        push(&b, gl_sub(cur[C->dec.sign], (uint64_t)C->r_sign));
    } else zeros(&b, NL + 1);
    if (row == C->dec_a_row) {
        ev_pins(&b, cur, C->dec.y, C->ay_limbs);
        push(&b, gl_sub(cur[C->dec.sign], (uint64_t)C->a_sign));
    } else zeros(&b, NL + 1);
    /* v6: scalar blocks */
    const size_t blk[2] = {C->sblk_sb, C->sblk_ka};
    const uint8_t *bits[2] = {C->s_bits, C->k_bits};
    for (int q = 0; q < 2; q++) {
        const size_t f = firsts[q], l = f + k - 1;
        if (row == f)
            for (size_t i = 0; i < k; i++) push(&b, gl_sub(cur[blk[q] + i], bits[q][i]));
        else zeros(&b, k);
        if (row >= f && row <= l) push(&b, gl_sub(cur[M->ca.bit], cur[blk[q]]));
        else zeros(&b, 1);
        if (row >= f && row < l)
            for (size_t i = 0; i + 1 < k; i++) push(&b, gl_sub(nxt[blk[q] + i], cur[blk[q] + i + 1]));
        else zeros(&b, k - 1);
    }
    /* v7: r thread */
    if (row == C->reduce_row)
        for (size_t i = 0; i < k; i++) push(&b, gl_sub(cur[C->r_thread + i], cur[C->red.rb + (k - 1 - i)]));
    else zeros(&b, k);
    for (size_t i = 0; i < k; i++) push(&b, gl_sub(nxt[C->r_thread + i], cur[C->r_thread + i]));
    if (row == C->ka_first)
        for (size_t i = 0; i < k; i++) push(&b, gl_sub(cur[C->sblk_ka + i], cur[C->r_thread + i]));
    else zeros(&b, k);
    /* v10: result row, cofactor doublings */
    if (row == C->result_row) {
        for (int i = 0; i < NL; i++) push(&b, cur[C->zero_c + i]);
        ev_sub(&b, cur, &C->sub_x);
        ev_sub(&b, cur, &C->sub_t);
        ev_sub(&b, cur, &C->sub_yz);
        for (int i = 0; i < NL; i++) push(&b, cur[C->sub_x.c + i]);
        for (int i = 0; i < NL; i++) push(&b, cur[C->sub_t.c + i]);
        for (int i = 0; i < NL; i++) push(&b, cur[C->sub_yz.c + i]);
    } else zeros(&b, NL + 3 * SUB_K + 3 * NL);
    const int in_dbl = row == C->dbl_row[0] || row == C->dbl_row[1] || row == C->dbl_row[2];
    if (in_dbl) ev_pdbl(&b, cur, &C->dbl);
    else zeros(&b, DBL_K);
    if (row == C->dbl_row[0] || row == C->dbl_row[1]) {
        ev_bind(&b, nxt, C->dbl_in[0], cur, C->dbl.X3.c, NL);
        ev_bind(&b, nxt, C->dbl_in[1], cur, C->dbl.Y3.c, NL);
        ev_bind(&b, nxt, C->dbl_in[2], cur, C->dbl.Z3.c, NL);
    } else zeros(&b, 3 * NL);
    if (row == C->dbl_row[2]) {
        ev_bind(&b, nxt, C->res_x, cur, C->dbl.X3.c, NL);
        ev_bind(&b, nxt, C->res_y, cur, C->dbl.Y3.c, NL);
        ev_bind(&b, nxt, C->res_z, cur, C->dbl.Z3.c, NL);
        ev_bind(&b, nxt, C->res_t, cur, C->dbl.T3.c, NL);
    } else zeros(&b, 4 * NL);
    /* v11 / v12: ladder result threads */
    const size_t lasts[2] = {C->sb_last, C->ka_last};
    const size_t *ths[2] = {C->th_sb, C->th_ka};
    for (int q = 0; q < 2; q++) {
        if (row == lasts[q])
            for (int s = 0; s < 4; s++) ev_bind(&b, cur, ths[q][s], cur, M->ca.out[s], NL);
        else zeros(&b, 4 * NL);
        for (int s = 0; s < 4; s++) ev_bind(&b, nxt, ths[q][s], cur, ths[q][s], NL);
    }
    /* v13: R thread */
    if (row == C->dec_r_row) {
        ev_bind(&b, cur, C->th_r[0], cur, C->dec.x, NL);
        ev_bind(&b, cur, C->th_r[1], cur, C->dec.y, NL);
        ev_bind(&b, cur, C->th_r[2], cur, C->one_r, NL);
        ev_bind(&b, cur, C->th_r[3], cur, C->mul_rt.c, NL);
    } else zeros(&b, 4 * NL);
    if (row == C->dec_r_row) {
        const int64_t one[NL] = {1};
        ev_pins(&b, cur, C->one_r, one);
    } else zeros(&b, NL);
    if (row == C->dec_r_row) ev_mul(&b, cur, &C->mul_rt);
    else zeros(&b, MUL_K);
    for (int s = 0; s < 4; s++) ev_bind(&b, nxt, C->th_r[s], cur, C->th_r[s], NL);
    /* v14 / v15 / v16: the residual row */
    if (row == C->residual_row) {
        for (int i = 0; i < NL; i++) push(&b, cur[C->zero_r1 + i]);
        ev_sub(&b, cur, &C->neg_rx);
        ev_sub(&b, cur, &C->neg_rt);
        ev_padd(&b, cur, &C->pa_r1);
        ev_sub(&b, cur, &C->neg_kax);
        ev_sub(&b, cur, &C->neg_kat);
        ev_padd(&b, cur, &C->pa_r2);
        ev_bind(&b, nxt, C->dbl_in[0], cur, C->pa_r2.X3.c, NL);
        ev_bind(&b, nxt, C->dbl_in[1], cur, C->pa_r2.Y3.c, NL);
        ev_bind(&b, nxt, C->dbl_in[2], cur, C->pa_r2.Z3.c, NL);
    } else zeros(&b, NL + 2 * SUB_K + POINT_ADD_K + 2 * SUB_K + POINT_ADD_K + 3 * NL);
}

/* deep_ali_merge_ed25519_verify_streaming (lib.rs:854-974): every LDE point, nxt = (i + blowup) mod n,
 * trace_row = i / blowup, no phi quirk; coeffs NULL = lambda_j = j + 1. lde: width x (n_trace * blowup). */
typedef struct {
    const ref_ed_ctx *C;
    const uint64_t *lde, *coeffs;
    size_t blowup;
    uint64_t *c_eval;
} ed_merge_job;
static void ed_merge_range(void *ctx, size_t lo, size_t hi) {
    const ed_merge_job *J = (const ed_merge_job *)ctx;
    const ref_ed_ctx *C = J->C;
    const size_t n = C->height * J->blowup, w = C->width, nk = C->n_cons;
    uint64_t *cur = (uint64_t *)malloc(8 * w), *nxt = (uint64_t *)malloc(8 * w), *cv = (uint64_t *)malloc(8 * nk);
    for (size_t i = lo; i < hi; i++) {
        const size_t ni = (i + J->blowup) % n;
        for (size_t c = 0; c < w; c++) {
            cur[c] = J->lde[c * n + i];
            nxt[c] = J->lde[c * n + ni];
        }
        ref_ed25519_constraints(C, cur, nxt, i / J->blowup, cv);
        uint64_t acc = 0;
        for (size_t j = 0; j < nk; j++)
            if (cv[j]) acc = gl_add(acc, gl_mul(J->coeffs ? J->coeffs[j] : (uint64_t)(j + 1), cv[j]));
        J->c_eval[i] = acc;
    }
    free(cur);
    free(nxt);
    free(cv);
}

int ref_merge_ed25519(const ref_ed_ctx *C, const uint64_t *lde, size_t blowup, const uint64_t *coeffs,
                      uint64_t *c_eval) {
    const size_t h = C->height, n = h * blowup;
    ed_init(); /* the shared constants are set before any thread reads them */
    ed_merge_job J = {C, lde, coeffs, blowup, c_eval};
    ref_par_for(n, PAR_FRI, ed_merge_range, &J); /* into_par_iter at >= 2^12 points (lib.rs:884-920) */
    ref_quotient_from_phi(c_eval, n, h);
    return 0;
}

/* verify (ed25519_verify.rs:54-118): decompress R and A, s < L, k = SHA-512(R || A || M) mod L,
 * [8]([s]B - R - [k]A) == identity */
int ref_ed25519_verify(const uint8_t pk[32], const uint8_t sig[64], const uint8_t *msg, size_t len) {
    ed_init();
    ept R, A;
    if (!ept_decompress(sig, &R) || !ept_decompress(pk, &A)) return 0;
    uint64_t s[4];
    for (int k = 0; k < 4; k++) {
        s[k] = 0;
        for (int i = 0; i < 8; i++) s[k] |= (uint64_t)sig[32 + 8 * k + i] << (8 * i);
    }
    if (u256_ge(s, LL)) return 0;
    uint8_t *buf = (uint8_t *)malloc(64 + len + 1), dg[64], kc[32];
    memcpy(buf, sig, 32);
    memcpy(buf + 32, pk, 32);
    if (len) memcpy(buf + 64, msg, len);
    sha512(buf, 64 + len, dg);
    free(buf);
    reduce_mod_l_wide(dg, kc);
    uint8_t sb[256], kb[256];
    for (int i = 0; i < 256; i++) {
        sb[i] = (sig[32 + (255 - i) / 8] >> ((255 - i) % 8)) & 1;
        kb[i] = (kc[(255 - i) / 8] >> ((255 - i) % 8)) & 1;
    }
    ept r = ept_add(ept_add(ept_ladder(ED_B, sb, 256), ept_neg(R)), ept_neg(ept_ladder(A, kb, 256)));
    r = ept_double(ept_double(ept_double(r)));
    return ept_is_identity(r);
}

void ref_sha512(const uint8_t *msg, size_t len, uint8_t out[64]) { sha512(msg, len, out); }

/* prove_zsk_ksk_binding_v2 (prover.rs:856-951) with k_scalar = k, blowup 32 and the STIR defaults;
 * pi_hash = zsk_ksk_pi_hash_v2_stark (:763-781). -30 when the native verify fails. */
int ref_prove_zsk_ksk_binding_v2(const uint8_t pk[32], const uint8_t sig[64], const uint8_t *msg, size_t len,
                                 const uint8_t fs_binding[32], const uint8_t merkle_root[32], size_t k,
                                 size_t blowup, uint8_t **proof, size_t *proof_len, uint8_t root_f0[32],
                                 uint8_t pi_hash[32]) {
    if (!ref_ed25519_verify(pk, sig, msg, len)) return -30;
    ref_ed_ctx *C = ref_ed25519_ctx_new(pk, sig, msg, len, k);
    if (!C) return -31;
    const size_t h = C->height, w = C->width, n0 = h * blowup;
    uint64_t *tr = (uint64_t *)malloc(8 * w * h), *lde = (uint64_t *)malloc(8 * w * n0), *c = (uint64_t *)malloc(8 * n0);
    ref_ed25519_fill(C, tr);
    ref_lde_columns(tr, w, h, blowup, 1, lde);
    free(tr);
    ref_merge_ed25519(C, lde, blowup, NULL, c);
    free(lde);
    ref_ed25519_ctx_free(C);
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
    sha3_ctx s;
    uint8_t le[8];
    for (int i = 0; i < 8; i++) le[i] = (uint8_t)((uint64_t)len >> (8 * i));
    sha3_init(&s);
    sha3_update(&s, "ZSK-KSK-PIHASH-V2-ST", 20);
    sha3_update(&s, pk, 32);
    sha3_update(&s, sig, 64);
    sha3_update(&s, le, 8);
    if (len) sha3_update(&s, msg, len);
    sha3_update(&s, fs_binding, 32);
    sha3_update(&s, merkle_root, 32);
    sha3_update(&s, root_f0, 32);
    sha3_final(&s, pi_hash);
    return 0;
}
