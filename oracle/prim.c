/*
 * prim.c -- oracle primitives: Goldilocks powers/inverses, Fp3/Fp6 inverse,
 * radix-2 DFT, Keccak-f[1600]/SHA3-256, ChaCha block, byte buffer.
 *
 * Test infrastructure only (the oracle); the product path never calls it.
 *
 * Field: crates/ark-goldilocks/build.rs:8-28. DFT: ark-poly 0.4.2
 * Radix2EvaluationDomain (third-party, used at crates/deep_ali/src/
 * trace_import.rs:271-279) -- evaluation at omega^i, natural order, ifft
 * scaled by 1/n. SHA3-256: sha3 0.10.8 (third-party), selected at
 * crates/hash/src/selected.rs:26-38. ChaCha: rand_chacha 0.3.1 (third-party)
 * behind rand 0.8.5 StdRng, used at crates/deep_ali/src/fri.rs:1044-1051.
 */
#include "refcpu_int.h"
#include <pthread.h>
#include <stdlib.h>

/* Threads for the loops the reference runs with rayon (merkle lib.rs:345
 * PARALLEL_THRESHOLD = 256 leaves / parents; fri.rs:89 PARALLEL_MIN_ELEMS =
 * 2^12 fold outputs). 1 (the default) runs everything on the caller. */
static int g_threads = 1;
void ref_set_threads(int n) { g_threads = n < 1 ? 1 : n > 256 ? 256 : n; }
int ref_get_threads(void) { return g_threads; }

typedef struct {
    void (*fn)(void *, size_t, size_t);
    void *ctx;
    size_t lo, hi;
} par_job;
static void *par_run(void *a) {
    par_job *j = (par_job *)a;
    j->fn(j->ctx, j->lo, j->hi);
    return NULL;
}
void ref_par_for(size_t n, size_t threshold, void (*fn)(void *, size_t, size_t), void *ctx) {
    int T = g_threads;
    if (T <= 1 || n < threshold || n < 2) {
        fn(ctx, 0, n);
        return;
    }
    if ((size_t)T > n) T = (int)n;
    pthread_t th[256];
    par_job jobs[256];
    for (int t = 0; t < T; t++) jobs[t] = (par_job){fn, ctx, n * (size_t)t / (size_t)T, n * (size_t)(t + 1) / (size_t)T};
    int started[256] = {0};
    for (int t = 1; t < T; t++) started[t] = pthread_create(&th[t], NULL, par_run, &jobs[t]) == 0;
    par_run(&jobs[0]);
    for (int t = 1; t < T; t++) {
        if (started[t]) pthread_join(th[t], NULL);
        else par_run(&jobs[t]); /* no thread: run its share here */
    }
}

uint64_t gl_pow(uint64_t a, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = gl_mul(r, a);
        a = gl_mul(a, a);
        e >>= 1;
    }
    return r;
}

uint64_t gl_inv(uint64_t a) { return gl_pow(a, GL_P - 2); }

uint64_t gl_root_of_unity(size_t n) {
    /* two-adic root 7^((p-1)/2^32) squared (32 - log n) times == 7^((p-1)/n) */
    return gl_pow(7, (GL_P - 1) / (uint64_t)n);
}

fp3 fp3_inv(fp3 a) {
    /* adjugate / norm (cubic_ext.rs:222-232) */
    uint64_t a0 = a.c[0], a1 = a.c[1], a2 = a.c[2];
    uint64_t d0 = gl_sub(gl_mul(a0, a0), gl_mul(7, gl_mul(a1, a2)));
    uint64_t d1 = gl_sub(gl_mul(7, gl_mul(a2, a2)), gl_mul(a0, a1));
    uint64_t d2 = gl_sub(gl_mul(a1, a1), gl_mul(a0, a2));
    /* N = a0*d0 + 7*(a1*d2 + a2*d1) */
    uint64_t n = gl_add(gl_mul(a0, d0), gl_mul(7, gl_add(gl_mul(a1, d2), gl_mul(a2, d1))));
    uint64_t ni = gl_inv(n);
    fp3 r = {{gl_mul(d0, ni), gl_mul(d1, ni), gl_mul(d2, ni)}};
    return r;
}

fp6 fp6_inv(fp6 a) {
    /* (b0 + b1 u)^-1 = (b0 - b1 u) / (b0^2 - a b1^2) (sextic_ext.rs:233-241) */
    fp3 b0 = {{a.c[0], a.c[1], a.c[2]}}, b1 = {{a.c[3], a.c[4], a.c[5]}};
    fp3 n = fp3_sub(fp3_mul(b0, b0), fp3_mul_by_a(fp3_mul(b1, b1)));
    fp3 ni = fp3_inv(n);
    fp3 r0 = fp3_mul(b0, ni), r1 = fp3_mul(b1, ni);
    fp6 r = {{r0.c[0], r0.c[1], r0.c[2], gl_neg(r1.c[0]), gl_neg(r1.c[1]), gl_neg(r1.c[2])}};
    return r;
}

fp6 fp6_pow(fp6 a, uint64_t e) {
    fp6 r = fp6_one();
    while (e) {
        if (e & 1) r = fp6_mul(r, a);
        a = fp6_mul(a, a);
        e >>= 1;
    }
    return r;
}

static size_t bitrev(size_t x, int bits) {
    size_t r = 0;
    for (int i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

void ref_ntt_inplace(uint64_t *a, size_t n, int inverse) {
    if (n <= 1) return;
    int lg = 0;
    while (((size_t)1 << lg) < n) lg++;
    for (size_t i = 0; i < n; i++) {
        size_t j = bitrev(i, lg);
        if (i < j) { uint64_t t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    uint64_t w_n = gl_root_of_unity(n);
    if (inverse) w_n = gl_inv(w_n);
    uint64_t *tw = (uint64_t *)malloc(sizeof(uint64_t) * (n / 2));
    for (size_t half = 1; half < n; half <<= 1) {
        uint64_t ws = gl_pow(w_n, n / (2 * half));
        uint64_t w = 1;
        for (size_t j = 0; j < half; j++) { tw[j] = w; w = gl_mul(w, ws); }
        for (size_t base = 0; base < n; base += 2 * half) {
            for (size_t j = 0; j < half; j++) {
                uint64_t u = a[base + j], t = gl_mul(a[base + j + half], tw[j]);
                a[base + j] = gl_add(u, t);
                a[base + j + half] = gl_sub(u, t);
            }
        }
    }
    free(tw);
    if (inverse) {
        uint64_t ninv = gl_inv((uint64_t)n);
        for (size_t i = 0; i < n; i++) a[i] = gl_mul(a[i], ninv);
    }
}

/* ------------------------------- Keccak ------------------------------- */
static const uint64_t KRC[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
    0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
    0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
    0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
    0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
    0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};
/* rho offsets r[x][y] indexed by lane x + 5y */
static const int KROT[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43, 25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};

static inline uint64_t rotl(uint64_t v, int r) { return r ? (v << r) | (v >> (64 - r)) : v; }

void keccak_f1600(uint64_t s[25]) {
    for (int rnd = 0; rnd < 24; rnd++) {
        uint64_t C[5], D[5], B[25];
        for (int x = 0; x < 5; x++) C[x] = s[x] ^ s[x + 5] ^ s[x + 10] ^ s[x + 15] ^ s[x + 20];
        for (int x = 0; x < 5; x++) D[x] = C[(x + 4) % 5] ^ rotl(C[(x + 1) % 5], 1);
        for (int i = 0; i < 25; i++) s[i] ^= D[i % 5];
        /* rho + pi: B[y, 2x+3y] = rot(A[x,y], r[x,y]) */
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++)
                B[y + 5 * ((2 * x + 3 * y) % 5)] = rotl(s[x + 5 * y], KROT[x + 5 * y]);
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++)
                s[x + 5 * y] = B[x + 5 * y] ^ (~B[(x + 1) % 5 + 5 * y] & B[(x + 2) % 5 + 5 * y]);
        s[0] ^= KRC[rnd];
    }
}

#define SHA3_RATE 136
void sha3_init(sha3_ctx *c) { memset(c, 0, sizeof *c); }

void sha3_update(sha3_ctx *c, const void *data, size_t len) {
    const uint8_t *p = (const uint8_t *)data;
    for (size_t i = 0; i < len; i++) {
        c->s[c->pos >> 3] ^= (uint64_t)p[i] << (8 * (c->pos & 7));
        if (++c->pos == SHA3_RATE) { keccak_f1600(c->s); c->pos = 0; }
    }
}

void sha3_final(const sha3_ctx *c0, uint8_t out[32]) {
    sha3_ctx c = *c0;
    c.s[c.pos >> 3] ^= (uint64_t)0x06 << (8 * (c.pos & 7));
    c.s[(SHA3_RATE - 1) >> 3] ^= (uint64_t)0x80 << (8 * ((SHA3_RATE - 1) & 7));
    keccak_f1600(c.s);
    for (int i = 0; i < 32; i++) out[i] = (uint8_t)(c.s[i >> 3] >> (8 * (i & 7)));
}

void sha3_256(const void *data, size_t len, uint8_t out[32]) {
    sha3_ctx c;
    sha3_init(&c);
    sha3_update(&c, data, len);
    sha3_final(&c, out);
}

/* ------------------------------- ChaCha ------------------------------- */
#define QR(a, b, c, d) \
    a += b; d ^= a; d = (d << 16) | (d >> 16); \
    c += d; b ^= c; b = (b << 12) | (b >> 20); \
    a += b; d ^= a; d = (d << 8) | (d >> 24);  \
    c += d; b ^= c; b = (b << 7) | (b >> 25);

static uint32_t le32(const uint8_t *p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }

void chacha_block(const uint8_t key[32], uint32_t counter, const uint8_t nonce[12], int rounds, uint32_t out[16]) {
    uint32_t in[16] = {0x61707865, 0x3320646e, 0x79622d32, 0x6b206574};
    for (int i = 0; i < 8; i++) in[4 + i] = le32(key + 4 * i);
    in[12] = counter;
    for (int i = 0; i < 3; i++) in[13 + i] = le32(nonce + 4 * i);
    uint32_t x[16];
    memcpy(x, in, sizeof x);
    for (int r = 0; r < rounds; r += 2) {
        QR(x[0], x[4], x[8], x[12]) QR(x[1], x[5], x[9], x[13]) QR(x[2], x[6], x[10], x[14]) QR(x[3], x[7], x[11], x[15])
        QR(x[0], x[5], x[10], x[15]) QR(x[1], x[6], x[11], x[12]) QR(x[2], x[7], x[8], x[13]) QR(x[3], x[4], x[9], x[14])
    }
    for (int i = 0; i < 16; i++) out[i] = x[i] + in[i];
}

void bb_put(bbuf *b, const void *d, size_t n) {
    if (b->len + n > b->cap) {
        size_t nc = b->cap ? b->cap * 2 : 4096;
        while (nc < b->len + n) nc *= 2;
        b->p = (uint8_t *)realloc(b->p, nc);
        b->cap = nc;
    }
    memcpy(b->p + b->len, d, n);
    b->len += n;
}
