/* rsa_air.c -- CPU restatement of the RSA-2048 exponentiation AIR
 * (s^65537 mod n, one 2048-bit modular product per trace row), its
 * stacked form and the per-record prover of the mixed-algorithm zone
 * example (test infrastructure: only tests/ load it).
 *
 *   limbs, mul gadget      crates/deep_ali/src/rsa2048_field_air.rs:52-98 (sizes),
 *                          :129-162 (layout), :197-219 (limbs),
 *                          :244-344 (fill), :356-432 (constraints, this order)
 *   exponentiation chain   crates/deep_ali/src/rsa2048_exp_air.rs:97-123 (layout),
 *                          :148-202 (fill), :210-285 (constraints, this order)
 *   stacked records        crates/deep_ali/src/rsa2048_stacked_air.rs:44-99
 *   merges                 crates/deep_ali/src/lib.rs:1827-1944, :1949-2049
 *   PKCS#1 v1.5 encoding   crates/deep_ali/src/rsa2048.rs (emsa_pkcs1_v1_5_encode_sha256)
 *   per-record prover      crates/swarm-dns/examples/mixed_algorithm_zone_rollup.rs:324-395
 *
 * Big integers here are little-endian arrays of 32-bit words; the quotient
 * and remainder of the 4096-bit product come from plain shift-and-subtract
 * long division (simple and slow, which suits a checker). */
#include <stdlib.h>
#include <string.h>

#include "refcpu.h"
#include "refcpu_int.h"

enum {
    NL = 80, LB = 26, CARRY_BITS = 40, CARRY_POS = 158, CROSS = 159,
    /* columns of one record (rsa2048_exp_air.rs:97-123, rsa2048_field_air.rs:129-162) */
    C_ACC = 0, C_S = 80, C_N = 160, C_PHASE = 240, C_OP2 = 241, C_EM = 321,
    C_C = 401, C_Q = 481, C_QBITS = 561, C_CBITS = 2641, C_CARRY = 4721, RSA_W = 11041,
    /* constraint slots of one record */
    K_PHASE = 0, K_SEL = 1, K_CBOOL = 81, K_CPACK = 2161, K_QBOOL = 2241, K_QPACK = 4321,
    K_POS = 4401, K_CARRY = 4560, K_EM = 10880, K_TRANS = 10960, RSA_K = 11280,
    OUT_ROW = 16, NW = 64 /* 32-bit words of a 2048-bit value */
};
static const int64_t CARRY_OFF = (int64_t)1 << 39;

/* ---- big integers (LE 32-bit words) ---- */
static void bn_from_be(const uint8_t *be, size_t len, uint32_t *w, size_t nw) {
    memset(w, 0, 4 * nw);
    for (size_t i = 0; i < len; i++) {
        const size_t bit = 8 * (len - 1 - i);
        if (bit / 32 < nw) w[bit / 32] |= (uint32_t)be[i] << (bit % 32);
    }
}

static int bn_cmp(const uint32_t *a, const uint32_t *b, size_t nw) {
    for (size_t i = nw; i-- > 0;)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}

static int bn_is_zero(const uint32_t *a, size_t nw) {
    for (size_t i = 0; i < nw; i++)
        if (a[i]) return 0;
    return 1;
}

/* prod (2 nw words) = a * b */
static void bn_mul(const uint32_t *a, const uint32_t *b, size_t nw, uint32_t *prod) {
    memset(prod, 0, 8 * nw);
    for (size_t i = 0; i < nw; i++) {
        uint64_t carry = 0;
        for (size_t j = 0; j < nw; j++) {
            const uint64_t t = (uint64_t)a[i] * b[j] + prod[i + j] + carry;
            prod[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        prod[i + nw] = (uint32_t)carry;
    }
}

/* x (2 nw words) = q * n + r, r < n: bit-serial long division; q has 2 nw
 * words, r has nw + 1 words (the running remainder before a subtraction can
 * reach 2 n) */
static void bn_divmod(const uint32_t *x, const uint32_t *n, size_t nw, uint32_t *q, uint32_t *r) {
    const size_t xw = 2 * nw, rw = nw + 1;
    uint32_t nn[NW + 1];
    memcpy(nn, n, 4 * nw);
    nn[nw] = 0;
    memset(q, 0, 4 * xw);
    memset(r, 0, 4 * rw);
    for (size_t bit = 32 * xw; bit-- > 0;) {
        for (size_t i = rw; i-- > 1;) r[i] = (r[i] << 1) | (r[i - 1] >> 31);
        r[0] = (r[0] << 1) | ((x[bit / 32] >> (bit % 32)) & 1);
        if (bn_cmp(r, nn, rw) >= 0) {
            uint64_t borrow = 0;
            for (size_t i = 0; i < rw; i++) {
                const uint64_t t = (uint64_t)r[i] - nn[i] - borrow;
                r[i] = (uint32_t)t;
                borrow = (t >> 63) & 1;
            }
            q[bit / 32] |= 1u << (bit % 32);
        }
    }
}

/* 80 limbs of 26 bits (biguint_to_limbs80, rsa2048_field_air.rs:197-210);
 * bits above 2080 are dropped as there */
static void bn_to_limbs(const uint32_t *w, size_t nw, int64_t *limbs) {
    for (int i = 0; i < NL; i++) {
        uint64_t v = 0;
        for (int b = 0; b < LB; b++) {
            const size_t bit = (size_t)(LB * i + b);
            if (bit / 32 < nw) v |= (uint64_t)((w[bit / 32] >> (bit % 32)) & 1) << b;
        }
        limbs[i] = (int64_t)v;
    }
}

static void limbs_to_bn(const int64_t *limbs, uint32_t *w, size_t nw) {
    memset(w, 0, 4 * nw);
    for (int i = 0; i < NL; i++)
        for (int b = 0; b < LB; b++) {
            const size_t bit = (size_t)(LB * i + b);
            if (((limbs[i] >> b) & 1) && bit / 32 < nw) w[bit / 32] |= 1u << (bit % 32);
        }
}

/* ---- trace builder ---- */

/* fill_rsa_mul_gadget (rsa2048_field_air.rs:244-344) on row r; a, b < n */
static int fill_mul(uint64_t *tr, size_t ht, size_t base, size_t r, const uint32_t *a, const uint32_t *b,
                    const uint32_t *n, uint32_t *c_out) {
    uint32_t prod[2 * NW], q[2 * NW], rem[NW + 1];
    bn_mul(a, b, NW, prod);
    bn_divmod(prod, n, NW, q, rem);
    if (rem[NW] || !bn_is_zero(q + NW, NW)) return -1; /* a, b < n keeps q, c below n */
    memcpy(c_out, rem, 4 * NW);
    int64_t al[NL], bl[NL], nl[NL], cl[NL], ql[NL], carries[CARRY_POS];
    bn_to_limbs(a, NW, al);
    bn_to_limbs(b, NW, bl);
    bn_to_limbs(n, NW, nl);
    bn_to_limbs(rem, NW, cl);
    bn_to_limbs(q, NW, ql);
    int64_t prev = 0;
    for (int k = 0; k < CROSS; k++) {
        int64_t pab = 0, pqn = 0;
        const int lo = k > NL - 1 ? k - (NL - 1) : 0, hi = k < NL - 1 ? k : NL - 1;
        for (int i = lo; i <= hi; i++) {
            pab += al[i] * bl[k - i];
            pqn += ql[i] * nl[k - i];
        }
        const int64_t lhs = pab - pqn - (k < NL ? cl[k] : 0) + prev;
        if (k < CROSS - 1) {
            if (lhs % ((int64_t)1 << LB)) return -2;
            carries[k] = lhs / ((int64_t)1 << LB); /* exact, so truncation is the floor */
            if (carries[k] >= CARRY_OFF || carries[k] < -CARRY_OFF) return -3;
            prev = carries[k];
        } else if (lhs != 0) {
            return -4;
        }
    }
    for (int i = 0; i < NL; i++) {
        tr[(base + C_C + i) * ht + r] = (uint64_t)cl[i];
        tr[(base + C_Q + i) * ht + r] = (uint64_t)ql[i];
        for (int b = 0; b < LB; b++) {
            tr[(base + C_CBITS + LB * i + b) * ht + r] = (uint64_t)((cl[i] >> b) & 1);
            tr[(base + C_QBITS + LB * i + b) * ht + r] = (uint64_t)((ql[i] >> b) & 1);
        }
    }
    for (int k = 0; k < CARRY_POS; k++) {
        const uint64_t biased = (uint64_t)(carries[k] + CARRY_OFF);
        for (int b = 0; b < CARRY_BITS; b++)
            tr[(base + C_CARRY + CARRY_BITS * k + b) * ht + r] = (biased >> b) & 1;
    }
    return 0;
}

/* fill_rsa_exp_multirow (rsa2048_exp_air.rs:148-202) for the record at
 * column base `base` of a column-major trace with ht rows; n, s, em are
 * 256-byte big-endian values, s < n */
int ref_fill_rsa_exp(const uint8_t n_be[256], const uint8_t s_be[256], const uint8_t em_be[256], size_t ht,
                     size_t base, uint64_t *tr) {
    if (ht < 17 || (ht & (ht - 1))) return -10;
    uint32_t n[NW], s[NW], em[NW], acc[NW], c[NW];
    bn_from_be(n_be, 256, n, NW);
    bn_from_be(s_be, 256, s, NW);
    bn_from_be(em_be, 256, em, NW);
    if (bn_is_zero(n, NW) || bn_cmp(s, n, NW) >= 0) return -11;
    int64_t sl[NL], nl[NL], el[NL], al[NL], ol[NL];
    bn_to_limbs(s, NW, sl);
    bn_to_limbs(n, NW, nl);
    bn_to_limbs(em, NW, el);
    memcpy(acc, s, sizeof acc);
    for (size_t r = 0; r < ht; r++) {
        const int phase = r == OUT_ROW;
        const uint32_t *op2 = phase ? s : acc;
        bn_to_limbs(acc, NW, al);
        bn_to_limbs(op2, NW, ol);
        for (int i = 0; i < NL; i++) {
            tr[(base + C_S + i) * ht + r] = (uint64_t)sl[i];
            tr[(base + C_N + i) * ht + r] = (uint64_t)nl[i];
            tr[(base + C_EM + i) * ht + r] = (uint64_t)el[i];
            tr[(base + C_ACC + i) * ht + r] = (uint64_t)al[i];
            tr[(base + C_OP2 + i) * ht + r] = (uint64_t)ol[i];
        }
        tr[(base + C_PHASE) * ht + r] = (uint64_t)phase;
        const int rc = fill_mul(tr, ht, base, r, acc, op2, n, c);
        if (rc) return rc;
        /* acc <- the gadget's c, read back through its limbs */
        int64_t cl[NL];
        for (int i = 0; i < NL; i++) cl[i] = (int64_t)tr[(base + C_C + i) * ht + r];
        limbs_to_bn(cl, acc, NW);
    }
    return 0;
}

size_t ref_rsa_exp_width(void) { return RSA_W; }
size_t ref_rsa_exp_constraints_per_record(void) { return RSA_K; }

/* ---- constraints ---- */

/* eval_rsa_mul_gadget (rsa2048_field_air.rs:356-432): 10 799 values */
static void eval_mul(const uint64_t *cur, uint64_t *out) {
    size_t o = 0;
    const size_t lb[2] = {C_C, C_Q}, bb[2] = {C_CBITS, C_QBITS};
    for (int e = 0; e < 2; e++) {
        for (int i = 0; i < NL; i++)
            for (int b = 0; b < LB; b++) {
                const uint64_t x = cur[bb[e] + LB * i + b];
                out[o++] = gl_mul(x, gl_sub(1, x));
            }
        for (int i = 0; i < NL; i++) {
            uint64_t s = 0;
            for (int b = 0; b < LB; b++) s = gl_add(s, gl_mul((uint64_t)1 << b, cur[bb[e] + LB * i + b]));
            out[o++] = gl_sub(cur[lb[e] + i], s);
        }
    }
    uint64_t sc[CARRY_POS];
    for (int k = 0; k < CARRY_POS; k++) {
        uint64_t biased = 0;
        for (int b = 0; b < CARRY_BITS; b++)
            biased = gl_add(biased, gl_mul((uint64_t)1 << b, cur[C_CARRY + CARRY_BITS * k + b]));
        sc[k] = gl_sub(biased, (uint64_t)CARRY_OFF);
    }
    const uint64_t radix = (uint64_t)1 << LB;
    for (int k = 0; k < CROSS; k++) {
        uint64_t pab = 0, pqn = 0;
        const int lo = k > NL - 1 ? k - (NL - 1) : 0, hi = k < NL - 1 ? k : NL - 1;
        for (int i = lo; i <= hi; i++) {
            pab = gl_add(pab, gl_mul(cur[C_ACC + i], cur[C_OP2 + k - i]));
            pqn = gl_add(pqn, gl_mul(cur[C_Q + i], cur[C_N + k - i]));
        }
        const uint64_t ck = k < NL ? cur[C_C + k] : 0;
        const uint64_t cin = k == 0 ? 0 : sc[k - 1];
        const uint64_t cout = k < CARRY_POS ? sc[k] : 0;
        out[o++] = gl_sub(gl_add(gl_sub(gl_sub(pab, pqn), ck), cin), gl_mul(radix, cout));
    }
    for (int k = 0; k < CARRY_POS; k++)
        for (int b = 0; b < CARRY_BITS; b++) {
            const uint64_t x = cur[C_CARRY + CARRY_BITS * k + b];
            out[o++] = gl_mul(x, gl_sub(1, x));
        }
}

/* eval_rsa_exp_multirow_per_row (rsa2048_exp_air.rs:210-285) for the record
 * whose columns start at cur/nxt: 11 280 values */
void ref_rsa_exp_constraints(const uint64_t *cur, const uint64_t *nxt, size_t trace_row, size_t n_trace,
                             uint64_t *out) {
    const uint64_t ph = cur[C_PHASE];
    out[K_PHASE] = gl_mul(ph, gl_sub(1, ph));
    for (int i = 0; i < NL; i++) {
        const uint64_t target = gl_add(gl_mul(ph, gl_sub(cur[C_S + i], cur[C_ACC + i])), cur[C_ACC + i]);
        out[K_SEL + i] = gl_sub(cur[C_OP2 + i], target);
    }
    eval_mul(cur, out + K_CBOOL);
    for (int i = 0; i < NL; i++)
        out[K_EM + i] = trace_row == OUT_ROW ? gl_sub(cur[C_C + i], cur[C_EM + i]) : 0;
    const int tr = trace_row + 1 < n_trace;
    for (int i = 0; i < NL; i++) {
        out[K_TRANS + i] = tr ? gl_sub(nxt[C_ACC + i], cur[C_C + i]) : 0;
        out[K_TRANS + NL + i] = tr ? gl_sub(nxt[C_S + i], cur[C_S + i]) : 0;
        out[K_TRANS + 2 * NL + i] = tr ? gl_sub(nxt[C_N + i], cur[C_N + i]) : 0;
        out[K_TRANS + 3 * NL + i] = tr ? gl_sub(nxt[C_EM + i], cur[C_EM + i]) : 0;
    }
}

/* deep_ali_merge_rsa_exp_multirow_streaming (lib.rs:1827-1944) and
 * deep_ali_merge_rsa_stacked_streaming (lib.rs:1949-2049): n_records
 * records side by side (record r at columns r * 11041, its constraints at
 * slots r * 11280); coeffs = NULL means lambda_j = j + 1. The streaming
 * merges evaluate phi at every point (no thread-count quirk). */
int ref_merge_rsa(const uint64_t *lde, size_t n_trace, size_t blowup, size_t n_records, const uint64_t *coeffs,
                  uint64_t *c_eval) {
    const size_t n = n_trace * blowup, w = RSA_W * n_records;
    uint64_t *cur = (uint64_t *)malloc(8 * w), *nxt = (uint64_t *)malloc(8 * w);
    uint64_t *cv = (uint64_t *)malloc(8 * RSA_K);
    for (size_t i = 0; i < n; i++) {
        const size_t ni = (i + blowup) % n;
        for (size_t c = 0; c < w; c++) {
            cur[c] = lde[c * n + i];
            nxt[c] = lde[c * n + ni];
        }
        uint64_t acc = 0;
        for (size_t r = 0; r < n_records; r++) {
            ref_rsa_exp_constraints(cur + RSA_W * r, nxt + RSA_W * r, i / blowup, n_trace, cv);
            for (size_t k = 0; k < RSA_K; k++) {
                const size_t j = RSA_K * r + k;
                acc = gl_add(acc, gl_mul(coeffs ? coeffs[j] : (uint64_t)(j + 1), cv[k]));
            }
        }
        c_eval[i] = acc;
    }
    free(cur);
    free(nxt);
    free(cv);
    ref_quotient_from_phi(c_eval, n, n_trace);
    return 0;
}

/* emsa_pkcs1_v1_5_encode_sha256 (rsa2048.rs) for k = 256 (RFC 8017 9.2):
 * 00 01 FF..FF 00 || DigestInfo(SHA-256) || digest */
void ref_rsa_pkcs1_em_sha256(const uint8_t digest[32], uint8_t em[256]) {
    static const uint8_t prefix[19] = {0x30, 0x31, 0x30, 0x0D, 0x06, 0x09, 0x60, 0x86, 0x48, 0x01,
                                       0x65, 0x03, 0x04, 0x02, 0x01, 0x05, 0x00, 0x04, 0x20};
    const size_t ps = 256 - 51 - 3;
    em[0] = 0x00;
    em[1] = 0x01;
    memset(em + 2, 0xFF, ps);
    em[2 + ps] = 0x00;
    memcpy(em + 3 + ps, prefix, 19);
    memcpy(em + 256 - 32, digest, 32);
}

/* prove_record_rsa (mixed_algorithm_zone_rollup.rs:349-395) from the trace
 * on: n_trace 32, blowup 32, lambda_j = j + 1, STIR with 54 queries bound to
 * pis; pi_hash = SHA3-256(pis || root_f0) */
int ref_prove_rsa_exp(const uint8_t n_be[256], const uint8_t s_be[256], const uint8_t em_be[256],
                      const uint8_t pis[32], uint8_t **proof, size_t *proof_len, uint8_t root_f0[32],
                      uint8_t pi_hash[32]) {
    const size_t h = 32, n0 = h * 32;
    uint64_t *tr = (uint64_t *)calloc(RSA_W * h, 8);
    int rc = ref_fill_rsa_exp(n_be, s_be, em_be, h, 0, tr);
    if (rc) {
        free(tr);
        return rc;
    }
    uint64_t *lde = (uint64_t *)malloc(8 * RSA_W * n0), *c = (uint64_t *)malloc(8 * n0);
    ref_lde_columns(tr, RSA_W, h, 32, 1, lde);
    free(tr);
    ref_merge_rsa(lde, h, 32, 1, NULL, c);
    free(lde);
    size_t sched[24];
    ref_fri_params p;
    p.schedule = sched;
    p.n_layers = ref_stir_schedule(n0, sched) /* the examples run STIR only */;
    p.r = 54;
    p.s0 = 54;
    p.seed_z = 0xDEEFBAADULL;
    p.coeff_commit_final = 1;
    p.d_final = 1;
    p.stir = 1;
    p.public_inputs_hash = pis;
    rc = ref_deep_fri_prove(c, n0, &p, proof, proof_len);
    free(c);
    if (rc) return rc;
    memcpy(root_f0, *proof, 32);
    sha3_ctx s;
    sha3_init(&s);
    sha3_update(&s, pis, 32);
    sha3_update(&s, root_f0, 32);
    sha3_final(&s, pi_hash);
    return 0;
}
