/*
 * refcpu_int.h -- internal helpers of the CPU oracle (test infrastructure).
 *
 * THIS IS THE ORACLE. It exists only to check the HIP product path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library never links or calls anything under oracle/.
 *
 * Canonical-u64 Goldilocks p = 2^64 - 2^32 + 1 and its Fp3 / Fp6 towers.
 *   Reference: crates/ark-goldilocks/build.rs:8-28 (modulus, generator 7,
 *   two-adicity 32), crates/deep_ali/src/cubic_ext.rs:165-244 (Fp3 = Fp[a]/(a^3-7)),
 *   crates/deep_ali/src/sextic_ext.rs:60-100,199-271 (Fp6 = Fp3[u]/(u^2-a)).
 * The reference keeps elements as 2-limb Montgomery numbers; every observable
 * value (hash input, serialized byte) is the canonical residue, which is what
 * this file stores.
 */
#ifndef REFCPU_INT_H
#define REFCPU_INT_H

#include <stdint.h>
#include <stddef.h>
#include <string.h>

#define GL_P 0xFFFFFFFF00000001ULL
#define GL_EPS 0xFFFFFFFFULL /* 2^64 mod p */

typedef unsigned __int128 u128;

static inline uint64_t gl_red128(u128 x) {
    /* x = lo + 2^64*hi, 2^64 = 2^32 - 1, 2^96 = -1 (mod p) */
    uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    uint64_t hh = hi >> 32, hl = hi & GL_EPS;
    uint64_t t0 = lo - hh;
    if (lo < hh) t0 -= GL_EPS; /* borrow: add p == subtract 2^32-1 mod 2^64 */
    uint64_t t1 = hl * GL_EPS;
    uint64_t r = t0 + t1;
    if (r < t1) r += GL_EPS; /* carry: 2^64 = 2^32 - 1 */
    if (r >= GL_P) r -= GL_P;
    return r;
}
static inline uint64_t gl_add(uint64_t a, uint64_t b) {
    uint64_t r = a + b;
    if (r < a || r >= GL_P) r -= GL_P;
    return r;
}
static inline uint64_t gl_sub(uint64_t a, uint64_t b) {
    return a >= b ? a - b : a + (GL_P - b);
}
static inline uint64_t gl_neg(uint64_t a) { return a ? GL_P - a : 0; }
static inline uint64_t gl_mul(uint64_t a, uint64_t b) { return gl_red128((u128)a * b); }
static inline uint64_t gl_from_u64(uint64_t x) { return x >= GL_P ? x - GL_P : x; }
/* fn(ctx, lo, hi) over [0, n) split across the threads of ref_set_threads
 * when n >= threshold, else in one call on the caller */
void ref_par_for(size_t n, size_t threshold, void (*fn)(void *, size_t, size_t), void *ctx);
#define PAR_MERKLE 256  /* merkle lib.rs:345 */
#define PAR_FRI 4096    /* fri.rs:89 */
uint64_t gl_pow(uint64_t a, uint64_t e);
uint64_t gl_inv(uint64_t a);
/* omega_n = 7^((p-1)/n), n a power of two <= 2^32 (ark-poly group_gen). */
uint64_t gl_root_of_unity(size_t n);

/* ---- Fp3 (c0 + c1 a + c2 a^2, a^3 = 7) ---- */
typedef struct { uint64_t c[3]; } fp3;
/* ---- Fp6 (b0 + b1 u, u^2 = a), components in to_fp_components order ---- */
typedef struct { uint64_t c[6]; } fp6;

static inline fp3 fp3_add(fp3 a, fp3 b) { fp3 r; for (int i = 0; i < 3; i++) r.c[i] = gl_add(a.c[i], b.c[i]); return r; }
static inline fp3 fp3_sub(fp3 a, fp3 b) { fp3 r; for (int i = 0; i < 3; i++) r.c[i] = gl_sub(a.c[i], b.c[i]); return r; }
static inline fp3 fp3_mul(fp3 a, fp3 b) {
    fp3 r;
    uint64_t t = gl_add(gl_mul(a.c[1], b.c[2]), gl_mul(a.c[2], b.c[1]));
    r.c[0] = gl_add(gl_mul(a.c[0], b.c[0]), gl_mul(7, t));
    r.c[1] = gl_add(gl_add(gl_mul(a.c[0], b.c[1]), gl_mul(a.c[1], b.c[0])), gl_mul(7, gl_mul(a.c[2], b.c[2])));
    r.c[2] = gl_add(gl_add(gl_mul(a.c[0], b.c[2]), gl_mul(a.c[1], b.c[1])), gl_mul(a.c[2], b.c[0]));
    return r;
}
static inline fp3 fp3_mul_by_a(fp3 x) { fp3 r = {{gl_mul(7, x.c[2]), x.c[0], x.c[1]}}; return r; }
fp3 fp3_inv(fp3 a);

static inline fp6 fp6_zero(void) { fp6 r; memset(&r, 0, sizeof r); return r; }
static inline fp6 fp6_one(void) { fp6 r = fp6_zero(); r.c[0] = 1; return r; }
static inline fp6 fp6_from_base(uint64_t x) { fp6 r = fp6_zero(); r.c[0] = x; return r; }
static inline fp6 fp6_add(fp6 a, fp6 b) { fp6 r; for (int i = 0; i < 6; i++) r.c[i] = gl_add(a.c[i], b.c[i]); return r; }
static inline fp6 fp6_sub(fp6 a, fp6 b) { fp6 r; for (int i = 0; i < 6; i++) r.c[i] = gl_sub(a.c[i], b.c[i]); return r; }
static inline fp6 fp6_scale(fp6 a, uint64_t s) { fp6 r; for (int i = 0; i < 6; i++) r.c[i] = gl_mul(a.c[i], s); return r; }
static inline int fp6_eq(fp6 a, fp6 b) { return memcmp(&a, &b, sizeof a) == 0; }
static inline int fp6_is_zero(fp6 a) { for (int i = 0; i < 6; i++) if (a.c[i]) return 0; return 1; }
static inline fp6 fp6_mul(fp6 a, fp6 b) {
    fp3 a0 = {{a.c[0], a.c[1], a.c[2]}}, a1 = {{a.c[3], a.c[4], a.c[5]}};
    fp3 b0 = {{b.c[0], b.c[1], b.c[2]}}, b1 = {{b.c[3], b.c[4], b.c[5]}};
    fp3 v0 = fp3_mul(a0, b0), v1 = fp3_mul(a1, b1);
    fp3 r0 = fp3_add(v0, fp3_mul_by_a(v1));
    fp3 r1 = fp3_add(fp3_mul(a0, b1), fp3_mul(a1, b0));
    fp6 r = {{r0.c[0], r0.c[1], r0.c[2], r1.c[0], r1.c[1], r1.c[2]}};
    return r;
}
fp6 fp6_inv(fp6 a);
fp6 fp6_pow(fp6 a, uint64_t e);

/* ---- radix-2 DFT on <omega_n>, natural order (ark-poly Radix2EvaluationDomain) ---- */
void ref_ntt_inplace(uint64_t *a, size_t n, int inverse);

/* ---- SHA3-256 (FIPS 202) ---- */
typedef struct { uint64_t s[25]; size_t pos; } sha3_ctx;
void keccak_f1600(uint64_t s[25]);
void sha3_init(sha3_ctx *c);
void sha3_update(sha3_ctx *c, const void *data, size_t len);
void sha3_final(const sha3_ctx *c, uint8_t out[32]); /* does not modify c */
void sha3_256(const void *data, size_t len, uint8_t out[32]);

/* ---- ChaCha block function (RFC 8439 layout: 32-bit counter word 12, nonce words 13-15) ---- */
void chacha_block(const uint8_t key[32], uint32_t counter, const uint8_t nonce[12], int rounds, uint32_t out[16]);

/* ---- growable byte buffer ---- */
typedef struct { uint8_t *p; size_t len, cap; } bbuf;
void bb_put(bbuf *b, const void *d, size_t n);
static inline void bb_u8(bbuf *b, uint8_t v) { bb_put(b, &v, 1); }
static inline void bb_u64(bbuf *b, uint64_t v) { uint8_t t[8]; for (int i = 0; i < 8; i++) t[i] = (uint8_t)(v >> (8 * i)); bb_put(b, t, 8); }
static inline void bb_fp6(bbuf *b, fp6 v) { for (int i = 0; i < 6; i++) bb_u64(b, v.c[i]); }

#endif
