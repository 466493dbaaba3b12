/*
 * prover.c -- oracle restatement of the LDE, the HashRollup AIR and merge,
 * the DNS record hashing and the swarm-dns whole-proof entry points.
 *
 * Test infrastructure only (the oracle); the product path never calls it.
 *
 *   lde_trace_columns       crates/deep_ali/src/trace_import.rs:251-281
 *   poly_div_zh             crates/deep_ali/src/lib.rs:190-228
 *   merge + phi quirk       crates/deep_ali/src/lib.rs:20-33,270-437
 *   HashRollup AIR          crates/deep_ali/src/air_workloads.rs:571-619
 *   DNS records / tree      crates/swarm-dns/src/dns.rs:43-118
 *   synthetic zone          crates/swarm-ctrl/src/main.rs:1331-1357
 *   shard_fs_binding        crates/swarm-proto/src/messages.rs:162-176
 *   prove_inner_shard       crates/swarm-dns/src/prover.rs:53-67,87-102,148-227
 *   prove_outer_rollup      crates/swarm-dns/src/prover.rs:253-304
 *   NSEC3 chain AIR         crates/deep_ali/src/air_workloads.rs:789-870
 *   prove_nsec3_completeness crates/swarm-dns/src/prover.rs:348-446
 */
#include "refcpu.h"
#include "refcpu_int.h"
#include <stdio.h>
#include <stdlib.h>

uint64_t ref_gl_mul(uint64_t a, uint64_t b) { return gl_mul(a, b); }
uint64_t ref_gl_inv(uint64_t a) { return gl_inv(a); }
uint64_t ref_gl_root_of_unity(size_t n) { return gl_root_of_unity(n); }
void ref_fp6_mul(const uint64_t *a, const uint64_t *b, uint64_t *out) {
    fp6 x, y, r;
    memcpy(&x, a, 48); memcpy(&y, b, 48);
    r = fp6_mul(x, y);
    memcpy(out, &r, 48);
}
void ref_fp6_inv(const uint64_t *a, uint64_t *out) {
    fp6 x, r;
    memcpy(&x, a, 48);
    r = fp6_inv(x);
    memcpy(out, &r, 48);
}
void ref_ntt(uint64_t *a, size_t n, int inverse) { ref_ntt_inplace(a, n, inverse); }
void ref_sha3_256(const uint8_t *msg, size_t len, uint8_t out[32]) { sha3_256(msg, len, out); }
void ref_keccak_f1600(uint64_t st[25]) { keccak_f1600(st); }
void ref_chacha_block(const uint8_t key[32], uint32_t counter, const uint8_t nonce[12], int rounds, uint32_t out[16]) {
    chacha_block(key, counter, nonce, rounds, out);
}

static int is_pow2(size_t x) { return x && !(x & (x - 1)); }

int ref_lde_columns(const uint64_t *cols, size_t w, size_t n_trace, size_t blowup, uint64_t shift, uint64_t *out) {
    if (w == 0) return -1;
    if (!is_pow2(n_trace)) return -1;
    if (blowup < 2 || !is_pow2(blowup)) return -1;
    size_t n0 = n_trace * blowup;
    for (size_t c = 0; c < w; c++) {
        uint64_t *o = out + c * n0;
        memcpy(o, cols + c * n_trace, 8 * n_trace);
        ref_ntt_inplace(o, n_trace, 1);
        if (shift != 1) {
            uint64_t s = 1;
            for (size_t i = 0; i < n_trace; i++) { o[i] = gl_mul(o[i], s); s = gl_mul(s, shift); }
        }
        memset(o + n_trace, 0, 8 * (n0 - n_trace));
        ref_ntt_inplace(o, n0, 0);
    }
    return 0;
}

void ref_build_hash_rollup_trace(size_t n_trace, const uint64_t *leaves, size_t n_leaves, uint64_t *tr) {
    uint64_t s = 0;
    for (size_t i = 0; i < n_trace; i++) {
        uint64_t leaf = i < n_leaves ? gl_from_u64(leaves[i]) : 0;
        tr[0 * n_trace + i] = gl_from_u64(i);
        tr[1 * n_trace + i] = leaf;
        tr[2 * n_trace + i] = s;
        tr[3 * n_trace + i] = gl_mul(s, s);
        s = gl_add(tr[3 * n_trace + i], leaf);
    }
}

/* c = FFT(pad(poly_div_zh(iFFT(phi), n_trace)))  -- remainder dropped */
void ref_quotient_from_phi(uint64_t *phi, size_t n, size_t n_trace) {
    ref_ntt_inplace(phi, n, 1);
    size_t m = n_trace;
    if (n <= m) {
        memset(phi, 0, 8 * n);
    } else {
        size_t ql = n - m;
        uint64_t *q = (uint64_t *)calloc(ql, 8);
        for (size_t k = n; k-- > m;) {
            uint64_t qk = k < ql ? q[k] : 0;
            q[k - m] = gl_add(phi[k], qk);
        }
        memcpy(phi, q, 8 * ql);
        memset(phi + ql, 0, 8 * (n - ql));
        free(q);
    }
    ref_ntt_inplace(phi, n, 0);
}

int ref_merge_hash_rollup(const uint64_t *lde, size_t n_trace, size_t blowup, int phi_mode, uint64_t *c_eval) {
    return ref_merge_hash_rollup_lam(lde, n_trace, blowup, phi_mode, NULL, c_eval);
}

/* the same merge with combination coefficients lam[0 .. 3) (NULL: k + 1) */
int ref_merge_hash_rollup_lam(const uint64_t *lde, size_t n_trace, size_t blowup, int phi_mode, const uint64_t *lam,
                              uint64_t *c_eval) {
    size_t n = n_trace * blowup;
    const uint64_t l0 = lam ? lam[0] : 1, l1 = lam ? lam[1] : 2, l2 = lam ? lam[2] : 3;
    int compute = phi_mode == 0 || (phi_mode == 1 && n >= 4096);
    if (compute) {
        const uint64_t *idx = lde, *leaf = lde + n, *st = lde + 2 * n, *sq = lde + 3 * n;
        for (size_t i = 0; i < n; i++) {
            size_t j = (i + blowup) % n;
            uint64_t c0 = gl_sub(gl_sub(idx[j], idx[i]), 1);
            uint64_t c1 = gl_sub(sq[i], gl_mul(st[i], st[i]));
            uint64_t c2 = gl_sub(gl_sub(st[j], sq[i]), leaf[i]);
            c_eval[i] = gl_add(gl_add(gl_mul(l0, c0), gl_mul(l1, c1)), gl_mul(l2, c2));
        }
    } else {
        memset(c_eval, 0, 8 * n);
    }
    ref_quotient_from_phi(c_eval, n, n_trace);
    return 0;
}

/* ------------------------------- DNS ------------------------------- */
static void put_le(uint8_t *p, uint64_t v, int n) { for (int i = 0; i < n; i++) p[i] = (uint8_t)(v >> (8 * i)); }

size_t ref_zone_record_canonical(uint64_t i, uint8_t *out) {
    char dom[64];
    uint8_t rd[64];
    size_t rdl = 0;
    uint16_t ty;
    uint32_t ttl = 300;
    switch (i % 4) {
    case 0:
        snprintf(dom, sizeof dom, "host-%08llx.example.com", (unsigned long long)i);
        ty = 1;
        rd[0] = 10; rd[1] = (uint8_t)(i >> 16); rd[2] = (uint8_t)(i >> 8); rd[3] = (uint8_t)i;
        rdl = 4;
        break;
    case 1: {
        snprintf(dom, sizeof dom, "v6-%08llx.example.com", (unsigned long long)i);
        ty = 28;
        uint64_t x = i ^ 0xCAFEBABEDEADBEEFULL;
        for (int k = 0; k < 8; k++) { rd[k] = (uint8_t)(i >> (56 - 8 * k)); rd[8 + k] = (uint8_t)(x >> (56 - 8 * k)); }
        rdl = 16;
        break;
    }
    case 2:
        snprintf(dom, sizeof dom, "txt-%08llx.example.com", (unsigned long long)i);
        ty = 16; ttl = 60;
        rdl = (size_t)snprintf((char *)rd, sizeof rd, "v=auth1;rec=%llu", (unsigned long long)i);
        break;
    default: {
        snprintf(dom, sizeof dom, "mail-%08llx.example.com", (unsigned long long)i);
        ty = 15;
        uint16_t prio = (uint16_t)(i % 1000);
        rd[0] = (uint8_t)(prio >> 8); rd[1] = (uint8_t)prio;
        rdl = 2 + (size_t)snprintf((char *)rd + 2, sizeof rd - 2, "mx%llu.example.com", (unsigned long long)(i & 0xff));
        break;
    }
    }
    size_t dl = strlen(dom), o = 0;
    memcpy(out + o, "DNS-RECORD-V1", 13); o += 13;
    put_le(out + o, dl, 4); o += 4;
    memcpy(out + o, dom, dl); o += dl;
    put_le(out + o, ty, 2); o += 2;
    put_le(out + o, ttl, 4); o += 4;
    put_le(out + o, rdl, 4); o += 4;
    memcpy(out + o, rd, rdl); o += rdl;
    return o;
}

void ref_zone_leaf_hashes(uint64_t first, size_t count, const uint8_t salt[16], uint8_t *out) {
    uint8_t canon[256], h1[32];
    for (size_t k = 0; k < count; k++) {
        size_t cl = ref_zone_record_canonical(first + k, canon);
        sha3_ctx c;
        sha3_init(&c);
        sha3_update(&c, "DNS-LEAF-V1", 11);
        sha3_update(&c, salt, 16);
        sha3_update(&c, canon, cl);
        sha3_final(&c, h1);
        sha3_init(&c);
        sha3_update(&c, "DNS-LEAF-DOUBLE-V1", 18);
        sha3_update(&c, salt, 16);
        sha3_update(&c, h1, 32);
        sha3_final(&c, out + 32 * k);
    }
}

void ref_dns_merkle_root(const uint8_t *leaves, size_t n, uint8_t root[32]) {
    uint8_t *cur = (uint8_t *)malloc(32 * n);
    memcpy(cur, leaves, 32 * n);
    while (n > 1) {
        size_t np = (n + 1) / 2;
        for (size_t i = 0; i < np; i++) {
            const uint8_t *l = cur + 64 * i, *r = (2 * i + 1 < n) ? cur + 64 * i + 32 : l;
            sha3_ctx c;
            uint8_t h[32];
            sha3_init(&c);
            sha3_update(&c, "DNS-NODE-V1", 11);
            sha3_update(&c, l, 32);
            sha3_update(&c, r, 32);
            sha3_final(&c, h);
            memcpy(cur + 32 * i, h, 32);
        }
        n = np;
    }
    memcpy(root, cur, 32);
    free(cur);
}

void ref_shard_fs_binding(const uint8_t pk_hash[32], const uint8_t job_id[32], uint32_t shard_id, const uint8_t nonce[32], uint8_t out[32]) {
    sha3_ctx c;
    uint8_t sid[4];
    put_le(sid, shard_id, 4);
    sha3_init(&c);
    sha3_update(&c, "DNS-SWARM-SHARD-FS-V1", 21);
    sha3_update(&c, pk_hash, 32);
    sha3_update(&c, job_id, 32);
    sha3_update(&c, sid, 4);
    sha3_update(&c, nonce, 32);
    sha3_final(&c, out);
}

#define SWARM_BLOWUP 32
#define SWARM_QUERIES 54
#define SWARM_SEED_Z 0xDEEFBAADULL

/* LdtMode of the swarm-dns provers below (prover.rs:37-67): 0 Stir, 1 Fri */
static int g_ldt_fri = 0;
void ref_set_ldt(int fri) { g_ldt_fri = fri != 0; }
int ref_ldt_stir(void) { return !g_ldt_fri; }

size_t ref_stir_schedule(size_t n0, size_t *s) {
    int lg = 0;
    while (((size_t)1 << lg) < n0) lg++;
    size_t k = 0;
    for (int i = 0; i < lg / 3; i++) s[k++] = 8;
    if (lg % 3) s[k++] = (size_t)1 << (lg % 3);
    return k;
}

size_t ref_make_schedule(size_t n0, size_t *s) {
    int lg = 0;
    while (((size_t)1 << lg) < n0) lg++;
    size_t k = 0;
    if (g_ldt_fri) {
        for (int i = 0; i < lg; i++) s[k++] = 2; /* arity-2 binary folds */
        return k;
    }
    for (int i = 0; i < lg / 3; i++) s[k++] = 8;
    if (lg % 3) s[k++] = (size_t)1 << (lg % 3);
    return k;
}

/* HashRollup trace over packed leaves -> LDE -> merge -> deep_fri_prove */
static int rollup_prove(const uint64_t *leaves, size_t n_active, const uint8_t binding[32], int phi_mode,
                        uint8_t **proof, size_t *proof_len, uint8_t root_f0[32]) {
    size_t n_trace = 8;
    while (n_trace < n_active) n_trace <<= 1;
    size_t n0 = n_trace * SWARM_BLOWUP;
    uint64_t *trace = (uint64_t *)malloc(8 * 4 * n_trace);
    uint64_t *lde = (uint64_t *)malloc(8 * 4 * n0);
    uint64_t *c = (uint64_t *)malloc(8 * n0);
    ref_build_hash_rollup_trace(n_trace, leaves, n_active, trace);
    ref_lde_columns(trace, 4, n_trace, SWARM_BLOWUP, 1, lde);
    ref_merge_hash_rollup(lde, n_trace, SWARM_BLOWUP, phi_mode, c);
    size_t sched[24];
    ref_fri_params p;
    p.schedule = sched;
    p.n_layers = ref_make_schedule(n0, sched);
    p.r = SWARM_QUERIES; p.s0 = SWARM_QUERIES; p.seed_z = SWARM_SEED_Z;
    p.coeff_commit_final = 1; p.d_final = 1; p.stir = ref_ldt_stir();
    p.public_inputs_hash = binding;
    int rc = ref_deep_fri_prove(c, n0, &p, proof, proof_len);
    if (rc == 0) memcpy(root_f0, *proof, 32);
    free(trace); free(lde); free(c);
    return rc;
}

static void pack_leaves(const uint8_t *hashes, size_t count, uint64_t *out) {
    for (size_t k = 0; k < 4 * count; k++) {
        uint64_t v = 0;
        for (int b = 0; b < 8; b++) v |= (uint64_t)hashes[8 * k + b] << (8 * b);
        out[k] = v;
    }
}

int ref_prove_inner_shard(const uint8_t salt[16], const uint8_t *leaf_hashes, size_t count,
                          const uint8_t fs_binding[32], int phi_mode,
                          uint8_t **proof, size_t *proof_len, uint8_t pi_hash[32], uint8_t merkle_root[32], uint8_t root_f0[32]) {
    if (count == 0) return -1;
    ref_dns_merkle_root(leaf_hashes, count, merkle_root);
    uint64_t *leaves = (uint64_t *)malloc(8 * 4 * count);
    pack_leaves(leaf_hashes, count, leaves);
    int rc = rollup_prove(leaves, 4 * count, fs_binding, phi_mode, proof, proof_len, root_f0);
    free(leaves);
    if (rc) return rc;
    sha3_ctx c;
    uint8_t cnt[8];
    put_le(cnt, count, 8);
    sha3_init(&c);
    sha3_update(&c, "DNS-SHARD-PIHASH-V1", 19);
    sha3_update(&c, salt, 16);
    sha3_update(&c, cnt, 8);
    sha3_update(&c, merkle_root, 32);
    sha3_update(&c, root_f0, 32);
    sha3_update(&c, fs_binding, 32);
    sha3_final(&c, pi_hash);
    return 0;
}

int ref_prove_outer_rollup(const uint8_t *pi_hashes, size_t count, const uint8_t pk_hash[32], int phi_mode,
                           uint8_t **proof, size_t *proof_len, uint8_t root_f0[32]) {
    uint64_t *leaves = (uint64_t *)malloc(8 * 4 * (count ? count : 1));
    pack_leaves(pi_hashes, count, leaves);
    int rc = rollup_prove(leaves, 4 * count, pk_hash, phi_mode, proof, proof_len, root_f0);
    free(leaves);
    return rc;
}

/* ---------------------------- NSEC3 chain ---------------------------- */
/* build_nsec3_chain_trace (crates/deep_ali/src/air_workloads.rs:811-860):
 * columns 0..3 owner limbs, 4..7 next limbs (raw LE u64, reduced as
 * F::from); rows past the chain repeat the last next limbs as owner and
 * next, and the last row's next is forced to row 0's owner */
void ref_build_nsec3_trace(size_t n_trace, const uint64_t *chain, size_t n, uint64_t *tr) {
    memset(tr, 0, 8 * 8 * n_trace);
    const size_t m = n < n_trace ? n : n_trace;
    for (size_t i = 0; i < m; i++)
        for (int k = 0; k < 8; k++) tr[k * n_trace + i] = gl_from_u64(chain[8 * i + k]);
    if (n < n_trace) {
        uint64_t prev[4] = {0, 0, 0, 0};
        if (n)
            for (int k = 0; k < 4; k++) prev[k] = chain[8 * (n - 1) + 4 + k];
        for (size_t i = n; i < n_trace; i++)
            for (int k = 0; k < 4; k++) tr[k * n_trace + i] = tr[(4 + k) * n_trace + i] = gl_from_u64(prev[k]);
        for (int k = 0; k < 4; k++) tr[(4 + k) * n_trace + n_trace - 1] = tr[k * n_trace];
    }
}

/* deep_ali_merge_general (lib.rs:343-437) for AirType::Nsec3Chain:
 * C_k = nxt[k] - cur[4 + k] (eval_nsec3_chain_constraints,
 * air_workloads.rs:862-870), lambda_k = k + 1, same phi quirk */
void ref_nsec3_constraints(const uint64_t cur[8], const uint64_t nxt[8], uint64_t out[4]) {
    for (int k = 0; k < 4; k++) out[k] = gl_sub(nxt[k], cur[4 + k]);
}

int ref_merge_nsec3(const uint64_t *lde, size_t n_trace, size_t blowup, int phi_mode, uint64_t *c_eval) {
    return ref_merge_nsec3_lam(lde, n_trace, blowup, phi_mode, NULL, c_eval);
}

/* the same merge with combination coefficients lam[0 .. 4) (NULL: k + 1) */
int ref_merge_nsec3_lam(const uint64_t *lde, size_t n_trace, size_t blowup, int phi_mode, const uint64_t *lam,
                        uint64_t *c_eval) {
    const size_t n = n_trace * blowup;
    const int compute = phi_mode == 0 || (phi_mode == 1 && n >= 4096);
    for (size_t i = 0; i < n; i++) {
        uint64_t acc = 0;
        if (compute) {
            const size_t j = (i + blowup) % n;
            uint64_t cur[8], nxt[8], cv[4];
            for (int c = 0; c < 8; c++) {
                cur[c] = lde[c * n + i];
                nxt[c] = lde[c * n + j];
            }
            ref_nsec3_constraints(cur, nxt, cv);
            for (int k = 0; k < 4; k++) acc = gl_add(acc, gl_mul(lam ? lam[k] : (uint64_t)(k + 1), cv[k]));
        }
        c_eval[i] = acc;
    }
    ref_quotient_from_phi(c_eval, n, n_trace);
    return 0;
}

/* prove_nsec3_completeness (crates/swarm-dns/src/prover.rs:348-446):
 * records = n x (owner_hash[32] || next_hash[32]), a closed chain */
int ref_prove_nsec3_completeness(const uint8_t *records, size_t n, const uint8_t salt[16], const uint8_t fs_binding[32],
                                 int phi_mode, uint8_t **proof, size_t *proof_len, uint8_t pi_hash[32],
                                 uint8_t chain_root[32], uint8_t root_f0[32], size_t *n_trace_out) {
    if (n == 0) return -1;
    if (memcmp(records + 64 * (n - 1) + 32, records, 32) != 0) return -2; /* not closed */
    uint8_t cnt[8];
    put_le(cnt, n, 8);
    sha3_ctx c;
    sha3_init(&c);
    sha3_update(&c, "DNS-NSEC3-CHAIN-ROOT-V1", 23);
    sha3_update(&c, salt, 16);
    sha3_update(&c, cnt, 8);
    sha3_update(&c, records, 64 * n);
    sha3_final(&c, chain_root);
    size_t n_trace = 8;
    while (n_trace < n) n_trace <<= 1;
    const size_t n0 = n_trace * SWARM_BLOWUP;
    uint64_t *chain = (uint64_t *)malloc(8 * 8 * n);
    pack_leaves(records, 2 * n, chain);
    uint64_t *tr = (uint64_t *)malloc(8 * 8 * n_trace), *lde = (uint64_t *)malloc(8 * 8 * n0);
    uint64_t *ce = (uint64_t *)malloc(8 * n0);
    ref_build_nsec3_trace(n_trace, chain, n, tr);
    ref_lde_columns(tr, 8, n_trace, SWARM_BLOWUP, 1, lde);
    ref_merge_nsec3(lde, n_trace, SWARM_BLOWUP, phi_mode, ce);
    size_t sched[24];
    ref_fri_params p;
    p.schedule = sched;
    p.n_layers = ref_make_schedule(n0, sched);
    p.r = SWARM_QUERIES; p.s0 = SWARM_QUERIES; p.seed_z = SWARM_SEED_Z;
    p.coeff_commit_final = 1; p.d_final = 1; p.stir = ref_ldt_stir();
    p.public_inputs_hash = fs_binding;
    int rc = ref_deep_fri_prove(ce, n0, &p, proof, proof_len);
    free(chain); free(tr); free(lde); free(ce);
    if (rc) return rc;
    memcpy(root_f0, *proof, 32);
    sha3_init(&c);
    sha3_update(&c, "DNS-NSEC3-PIHASH-V1", 19);
    sha3_update(&c, salt, 16);
    sha3_update(&c, cnt, 8);
    sha3_update(&c, chain_root, 32);
    sha3_update(&c, root_f0, 32);
    sha3_update(&c, fs_binding, 32);
    sha3_final(&c, pi_hash);
    *n_trace_out = n_trace;
    return 0;
}
