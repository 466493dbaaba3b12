/*
 * fri.c -- oracle restatement of the SHA3 transcript, the Merkle channel and
 * the DEEP-FRI / STIR prover with its ark-serialize wire format.
 *
 * Test infrastructure only (the oracle); the product path never calls it.
 *
 *   transcript : crates/transcript/src/lib.rs:13-17,162-212,338-344,427-446
 *   merkle     : crates/merkle/src/lib.rs:95-113,185-238,377-512
 *   fri        : crates/deep_ali/src/fri.rs:73-85 (safe_field_challenge),
 *                :123-130 (tr_hash_fields_tagged), :338-391 (OOD coset),
 *                :670-771 (coset coefficients, fold), :874-905 (ext coeffs),
 *                :950-959 (merkle_depth), :1044-1103 (indices, tree configs),
 *                :1105-1129 (statement binding), :1372-1393 (ext challenges),
 *                :1398-1718 (transcript builder), :1813-1893 (STIR queries),
 *                :1900-2055 (deep_fri_prove)
 *   wire       : fri.rs:1235-1300 + ark-serialize 0.4.2 compressed
 */
#include "refcpu.h"
#include "refcpu_int.h"
#include <stdio.h>
#include <stdlib.h>

/* ------------------------------ transcript ------------------------------ */
typedef struct { sha3_ctx h; } tscript;

static void tr_init(tscript *t, const char *label) {
    sha3_init(&t->h);
    sha3_update(&t->h, "FSv1-TRANSCRIPT-INIT", 20);
    sha3_update(&t->h, label, strlen(label));
}
static void tr_absorb(tscript *t, const void *b, size_t n) {
    sha3_update(&t->h, "FSv1-ABSORB-BYTES", 17);
    sha3_update(&t->h, b, n);
}
static void tr_absorb_field(tscript *t, uint64_t x) {
    uint8_t le[8];
    for (int i = 0; i < 8; i++) le[i] = (uint8_t)(x >> (8 * i));
    tr_absorb(t, le, 8);
}
static void tr_absorb_fp6(tscript *t, fp6 v) { for (int i = 0; i < 6; i++) tr_absorb_field(t, v.c[i]); }
/* challenge hashes a clone: the state is not modified */
static void tr_challenge_bytes(const tscript *t, const void *label, size_t ln, uint8_t out[32]) {
    sha3_ctx c = t->h;
    sha3_update(&c, "FSv1-CHALLENGE", 14);
    sha3_update(&c, label, ln);
    sha3_final(&c, out);
}
static uint64_t safe_field_challenge(const tscript *t, const void *label, size_t ln) {
    uint8_t d[32];
    tr_challenge_bytes(t, label, ln, d);
    /* rchunks(7): chunks taken from the end, the last one may be short */
    uint64_t acc = 0;
    int end = 32;
    while (end > 0) {
        int start = end - 7 < 0 ? 0 : end - 7;
        int len = end - start;
        uint64_t val = 0;
        for (int i = 0; i < len; i++) val |= (uint64_t)d[start + i] << (8 * i);
        acc = gl_add(gl_mul(acc, (uint64_t)1 << (8 * len)), val);
        end = start;
    }
    return acc;
}
static fp6 challenge_ext(const tscript *t, const char *tag) {
    fp6 r;
    char sub[64];
    for (int i = 0; i < 6; i++) {
        int n = snprintf(sub, sizeof sub, "%s/c%d", tag, i);
        r.c[i] = safe_field_challenge(t, sub, (size_t)n);
    }
    return r;
}
static uint64_t tr_hash_fields_tagged(const char *tag, const uint64_t *f, size_t n) {
    tscript t;
    tr_init(&t, "FRI/FS");
    tr_absorb(&t, tag, strlen(tag));
    for (size_t i = 0; i < n; i++) tr_absorb_field(&t, f[i]);
    return safe_field_challenge(&t, "out", 3);
}

uint64_t ref_stdrng_first_u64(uint64_t seed) {
    uint8_t key[32] = {0}, nonce[12] = {0};
    uint32_t blk[16];
    for (int i = 0; i < 8; i++) key[i] = (uint8_t)(seed >> (8 * i));
    chacha_block(key, 0, nonce, 12, blk);
    return (uint64_t)blk[0] | ((uint64_t)blk[1] << 32);
}

/* -------------------------------- merkle -------------------------------- */
typedef struct {
    size_t arity, depth;
    uint64_t label;
    size_t nlev;
    uint8_t **lv;   /* lv[k]: digests of level k (after padding) */
    size_t *len;
} mtree;

static void dslabel(uint8_t out[32], uint64_t arity, uint64_t level, uint64_t pos, uint64_t label) {
    uint64_t v[4] = {arity, level, pos, label};
    for (int k = 0; k < 4; k++)
        for (int i = 0; i < 8; i++) out[8 * k + i] = (uint8_t)(v[k] >> (8 * i));
}

static void leaf_hash(size_t arity, uint64_t label, uint64_t pos, const uint64_t *vals, size_t nv, uint8_t out[32]) {
    sha3_ctx c;
    uint8_t ds[32];
    dslabel(ds, arity, 0xFFFFFFFFULL, pos, label);
    sha3_init(&c);
    sha3_update(&c, ds, 32);
    for (size_t i = 0; i < nv; i++) {
        uint8_t le[8];
        for (int b = 0; b < 8; b++) le[b] = (uint8_t)(vals[i] >> (8 * b));
        sha3_update(&c, le, 8);
    }
    sha3_final(&c, out);
}

static void mt_init(mtree *t, size_t arity, size_t depth, uint64_t label, size_t n_leaves) {
    t->arity = arity; t->depth = depth; t->label = label;
    t->nlev = 1;
    t->lv = (uint8_t **)calloc(depth + 2, sizeof(uint8_t *));
    t->len = (size_t *)calloc(depth + 2, sizeof(size_t));
    t->lv[0] = (uint8_t *)malloc(32 * (n_leaves + arity));
    t->len[0] = n_leaves;
}
static void mt_free(mtree *t) {
    for (size_t k = 0; k < t->nlev; k++) free(t->lv[k]);
    free(t->lv); free(t->len);
}
typedef struct {
    mtree *t;
    size_t level, a;
} node_job;
static void node_range(void *ctx, size_t lo, size_t hi) {
    node_job *j = (node_job *)ctx;
    for (size_t i = lo; i < hi; i++) {
        sha3_ctx c;
        uint8_t ds[32];
        dslabel(ds, j->a, j->level + 1, i, j->t->label);
        sha3_init(&c);
        sha3_update(&c, ds, 32);
        sha3_update(&c, j->t->lv[j->level] + 32 * j->a * i, 32 * j->a);
        sha3_final(&c, j->t->lv[j->level + 1] + 32 * i);
    }
}

/* leaf i of a tree = leaf_hash of vals + i * stride (nv values); leaves
 * hashed in parallel at >= 256 (merkle lib.rs:397-409) */
typedef struct {
    mtree *t;
    const uint64_t *vals;
    size_t stride, nv;
} leaf_job;
static void leaf_range(void *ctx, size_t lo, size_t hi) {
    leaf_job *j = (leaf_job *)ctx;
    for (size_t i = lo; i < hi; i++) leaf_hash(j->t->arity, j->t->label, i, j->vals + i * j->stride, j->nv, j->t->lv[0] + 32 * i);
}
static void hash_leaves(mtree *t, size_t n, const uint64_t *vals, size_t stride, size_t nv) {
    leaf_job j = {t, vals, stride, nv};
    ref_par_for(n, PAR_MERKLE, leaf_range, &j);
}
/* f0 coset-packed leaves: leaf b = (f0[b + j nn])_{j < m} */
typedef struct {
    mtree *t;
    const uint64_t *f0;
    size_t nn, m;
} packed_job;
static void packed_range(void *ctx, size_t lo, size_t hi) {
    packed_job *j = (packed_job *)ctx;
    uint64_t *fib = (uint64_t *)malloc(8 * j->m);
    for (size_t b = lo; b < hi; b++) {
        for (size_t k = 0; k < j->m; k++) fib[k] = j->f0[b + k * j->nn];
        leaf_hash(j->t->arity, j->t->label, b, fib, j->m, j->t->lv[0] + 32 * b);
    }
    free(fib);
}
static void hash_packed_leaves(mtree *t, const uint64_t *f0, size_t nn, size_t m) {
    packed_job j = {t, f0, nn, m};
    ref_par_for(nn, PAR_MERKLE, packed_range, &j);
}

/* finalize: levels padded by repeating the last node (merkle lib.rs:423-476) */
static void mt_finalize(mtree *t, uint8_t root[32]) {
    size_t level = 0;
    while (t->len[level] > 1) {
        if (level >= t->depth) { fprintf(stderr, "merkle: layer_arities exhausted\n"); abort(); }
        size_t a = t->arity, cur = t->len[level];
        if (cur % a) {
            size_t pad = a - cur % a;
            for (size_t i = 0; i < pad; i++) memcpy(t->lv[level] + 32 * (cur + i), t->lv[level] + 32 * (cur - 1), 32);
            cur += pad;
            t->len[level] = cur;
        }
        size_t np = cur / a;
        t->lv[level + 1] = (uint8_t *)malloc(32 * (np + a));
        node_job nj = {t, level, a};
        ref_par_for(np, PAR_MERKLE, node_range, &nj); /* par_chunks at >= 256 parents (lib.rs:443-449) */
        t->len[level + 1] = np;
        t->nlev = level + 2;
        level++;
    }
    memcpy(root, t->lv[t->nlev - 1], 32);
}
/* MerkleOpening serialized: leaf[32], path Vec<Vec<[32]>>, index u64 (merkle lib.rs:478-512) */
static void mt_open_ser(const mtree *t, size_t index, bbuf *b) {
    bb_put(b, t->lv[0] + 32 * index, 32);
    bb_u64(b, t->nlev - 1);
    size_t idx = index;
    for (size_t level = 0; level + 1 < t->nlev; level++) {
        size_t a = t->arity, gs = (idx / a) * a;
        bb_u64(b, a - 1);
        for (size_t i = 0; i < a; i++) {
            size_t pos = gs + i;
            if (pos == idx) continue;
            size_t p2 = pos < t->len[level] ? pos : t->len[level] - 1;
            bb_put(b, t->lv[level] + 32 * p2, 32);
        }
        idx /= a;
    }
    bb_u64(b, index);
}

static size_t pick_arity(size_t n, size_t m) {
    static const size_t as[6] = {128, 64, 32, 16, 8, 4};
    for (int i = 0; i < 6; i++)
        if (m >= as[i] && n % as[i] == 0) return as[i];
    return n % 2 == 0 ? 2 : 1;
}
static size_t merkle_depth(size_t leaves, size_t arity) {
    size_t d = 1, cur = leaves;
    while (cur > arity) { cur = (cur + arity - 1) / arity; d++; }
    return d;
}
static size_t pick_arity2(size_t n, size_t m) { size_t a = pick_arity(n, m); return a < 2 ? 2 : a; }

void ref_merkle_root(const uint64_t *leaf_vals, size_t n_leaves, size_t vpl, size_t arity, size_t depth, uint64_t label, uint8_t root[32]) {
    mtree t;
    mt_init(&t, arity, depth, label, n_leaves);
    hash_leaves(&t, n_leaves, leaf_vals, vpl, vpl);
    mt_finalize(&t, root);
    mt_free(&t);
}

int ref_merkle_open(const uint64_t *leaf_vals, size_t n_leaves, size_t vpl, size_t arity, size_t depth, uint64_t label,
                    size_t index, uint8_t **out, size_t *out_len) {
    if (index >= n_leaves) return -1;
    mtree t;
    uint8_t root[32];
    bbuf b = {0};
    mt_init(&t, arity, depth, label, n_leaves);
    hash_leaves(&t, n_leaves, leaf_vals, vpl, vpl);
    mt_finalize(&t, root);
    mt_open_ser(&t, index, &b);
    mt_free(&t);
    *out = b.p;
    *out_len = b.len;
    return 0;
}

/* ------------------------------ FRI pieces ------------------------------ */
/* f(x) interpolated on <omega_n> (natural order), evaluated at m points
 * zeta^j * z, zeta = omega^(n/m)  (fri.rs:338-391) */
static void ood_coset_evals(const fp6 *f, size_t n, fp6 z, size_t m, fp6 *out) {
    uint64_t *comp = (uint64_t *)malloc(sizeof(uint64_t) * n);
    fp6 *coef = (fp6 *)malloc(sizeof(fp6) * n);
    for (int c = 0; c < 6; c++) {
        for (size_t i = 0; i < n; i++) comp[i] = f[i].c[c];
        ref_ntt_inplace(comp, n, 1);
        for (size_t i = 0; i < n; i++) coef[i].c[c] = comp[i];
    }
    uint64_t zeta = gl_pow(gl_root_of_unity(n), n / m);
    fp6 zp = fp6_one();
    for (size_t j = 0; j < m; j++) {
        fp6 pt = fp6_mul(zp, z), acc = fp6_zero();
        for (size_t k = n; k-- > 0;) acc = fp6_add(fp6_mul(acc, pt), coef[k]);
        out[j] = acc;
        zp = fp6_scale(zp, zeta);
    }
    free(comp); free(coef);
}

/* per coset b: monomial coefficients of the degree-(m-1) interpolant of
 * f[b + j n/m] on the points omega^b zeta^j  (fri.rs:670-752) */
typedef struct {
    const fp6 *f;
    size_t nn, m;
    uint64_t minv, omega_inv;
    const uint64_t *zi;
    fp6 *tuples;
} cc_job;
static void cc_range(void *ctx, size_t lo, size_t hi) {
    cc_job *J = (cc_job *)ctx;
    const size_t m = J->m, nn = J->nn;
    uint64_t obinv = gl_pow(J->omega_inv, lo); /* omega^{-b} */
    for (size_t b = lo; b < hi; b++) {
        uint64_t sh = J->minv; /* (1/m) * omega^{-bk} */
        for (size_t k = 0; k < m; k++) {
            fp6 acc = fp6_zero();
            for (size_t j = 0; j < m; j++) acc = fp6_add(acc, fp6_scale(J->f[b + j * nn], J->zi[(j * k) % m]));
            J->tuples[b * m + k] = fp6_scale(acc, sh);
            sh = gl_mul(sh, obinv);
        }
        obinv = gl_mul(obinv, J->omega_inv);
    }
}
static void coset_coeffs(const fp6 *f, size_t n, size_t m, fp6 *tuples /* (n/m) x m */) {
    size_t nn = n / m;
    uint64_t omega = gl_root_of_unity(n);
    uint64_t zeta_inv = gl_inv(gl_pow(omega, nn));
    uint64_t *zi = (uint64_t *)malloc(sizeof(uint64_t) * m);
    zi[0] = 1;
    for (size_t k = 1; k < m; k++) zi[k] = gl_mul(zi[k - 1], zeta_inv);
    cc_job J = {f, nn, m, gl_inv(m), gl_inv(omega), zi, tuples};
    ref_par_for(nn, PAR_FRI, cc_range, &J); /* into_par_iter over cosets at >= 2^12 (fri.rs:700-703) */
    free(zi);
}

typedef struct {
    const fp6 *tuples;
    size_t m;
    fp6 alpha, *out;
} fold_job;
static void fold_range(void *ctx, size_t lo, size_t hi) {
    fold_job *J = (fold_job *)ctx;
    for (size_t b = lo; b < hi; b++) {
        fp6 acc = fp6_zero(), ap = fp6_one();
        for (size_t k = 0; k < J->m; k++) {
            acc = fp6_add(acc, fp6_mul(J->tuples[b * J->m + k], ap));
            ap = fp6_mul(ap, J->alpha);
        }
        J->out[b] = acc;
    }
}
static void fold_tuples(const fp6 *tuples, size_t nn, size_t m, fp6 alpha, fp6 *out) {
    fold_job J = {tuples, m, alpha, out};
    ref_par_for(nn, PAR_FRI, fold_range, &J); /* par_iter_mut at >= 2^12 outputs (fri.rs:609-620) */
}

static void bind_statement(tscript *t, const ref_fri_params *p, size_t n0) {
    if (p->public_inputs_hash) {
        tr_absorb(t, "CAIRO-PUBLIC-INPUTS-V1", 22);
        tr_absorb(t, p->public_inputs_hash, 32);
    }
    tr_absorb(t, "DEEP-FRI-STATEMENT-V2", 21);
    tr_absorb_field(t, n0);
    tr_absorb_field(t, p->n_layers);
    for (size_t i = 0; i < p->n_layers; i++) tr_absorb_field(t, p->schedule[i]);
    tr_absorb_field(t, gl_from_u64(p->seed_z));
    tr_absorb_field(t, 6);
    tr_absorb_field(t, p->coeff_commit_final ? 1 : 0);
    tr_absorb_field(t, p->stir ? 1 : 0);
}

typedef struct {
    size_t L, n0;
    fp6 **f;            /* f[0..L]: layer values (f[0] = f0 lifted) */
    size_t *sz;
    uint8_t root_f0[32];
    uint8_t (*roots)[32];
    fp6 *fz;            /* L */
    fp6 *coset_evals;   /* L x m_l, ragged: offsets */
    size_t *ce_off;
    fp6 *coeff_tuples;  /* final layer tuples */
    size_t n_final_tuples, m_final;
    uint8_t coeff_root[32];
    int has_coeff;
    fp6 final_coeffs[64];
    size_t n_final_coeffs;
    uint64_t query_seed;
    fp6 **q;            /* classic FRI: DEEP quotient layers q_l, l < L */
    fp6 alpha_final;
} fri_state;

static void fs_free(fri_state *s) {
    if (s->q)
        for (size_t i = 0; i < s->L; i++) free(s->q[i]);
    free(s->q);
    for (size_t i = 0; i <= s->L; i++) free(s->f[i]);
    free(s->f); free(s->sz); free(s->roots); free(s->fz); free(s->coset_evals); free(s->ce_off); free(s->coeff_tuples);
}

/* the layer's polynomial (6 component iFFTs of f on <omega_n>) at z, by Horner
 * from the top coefficient (fri.rs:255-299) */
static fp6 poly_eval_ext(const fp6 *f, size_t n, fp6 z) {
    uint64_t *comp = (uint64_t *)malloc(sizeof(uint64_t) * n);
    fp6 *coef = (fp6 *)malloc(sizeof(fp6) * n);
    for (int c = 0; c < 6; c++) {
        for (size_t i = 0; i < n; i++) comp[i] = f[i].c[c];
        ref_ntt_inplace(comp, n, 1);
        for (size_t i = 0; i < n; i++) coef[i].c[c] = comp[i];
    }
    fp6 acc = fp6_zero();
    for (size_t k = n; k-- > 0;) acc = fp6_add(fp6_mul(acc, z), coef[k]);
    free(comp); free(coef);
    return acc;
}

/* DEEP quotient q_i = (f_i - f(z)) / (omega^i - z)  (fri.rs:301-333) */
typedef struct {
    const fp6 *f;
    fp6 z, fz, *q;
    uint64_t omega;
} q_job;
static void q_range(void *ctx, size_t lo, size_t hi) {
    q_job *J = (q_job *)ctx;
    uint64_t x = gl_pow(J->omega, lo);
    for (size_t i = lo; i < hi; i++) {
        fp6 den = fp6_sub(fp6_from_base(x), J->z);
        J->q[i] = fp6_mul(fp6_sub(J->f[i], J->fz), fp6_inv(den));
        x = gl_mul(x, J->omega);
    }
}

/* classic-FRI leaf of layer ell at i: f_i, s_i, q_i (ext_leaf_fields,
 * fri.rs:1356-1362); s_i = f_{ell+1}[i mod n/m], or 0 on the final
 * coefficient-committed layer (fri.rs:773-789, :1600) */
static void classic_leaves(const fri_state *s, size_t ell, size_t cur, size_t m, int final_cc, uint64_t *out) {
    size_t nn = cur / m;
    for (size_t i = 0; i < cur; i++) {
        uint64_t *o = out + 18 * i;
        memcpy(o, s->f[ell][i].c, 48);
        if (final_cc) memset(o + 6, 0, 48);
        else memcpy(o + 6, s->f[ell + 1][i % nn].c, 48);
        memcpy(o + 12, s->q[ell][i].c, 48);
    }
}

static int fri_build(const uint64_t *f0, size_t n0, const ref_fri_params *p, fri_state *s) {
    size_t L = p->n_layers;
    if (L == 0) return -2;
    memset(s, 0, sizeof *s);
    s->L = L; s->n0 = n0;
    int use_cc = p->coeff_commit_final && L > 0;
    const int stir = p->stir != 0;
    s->f = (fp6 **)calloc(L + 1, sizeof(fp6 *));
    s->sz = (size_t *)calloc(L + 1, sizeof(size_t));
    s->roots = (uint8_t(*)[32])calloc(L, 32);
    s->fz = (fp6 *)calloc(L, sizeof(fp6));
    s->ce_off = (size_t *)calloc(L + 1, sizeof(size_t));
    if (!stir) s->q = (fp6 **)calloc(L, sizeof(fp6 *));
    size_t tot = 0, n = n0;
    for (size_t l = 0; l < L; l++) {
        if (p->schedule[l] < 2 || n % p->schedule[l]) return -1;
        s->ce_off[l] = tot; tot += p->schedule[l]; n /= p->schedule[l];
    }
    s->ce_off[L] = tot;
    s->coset_evals = (fp6 *)calloc(tot, sizeof(fp6));

    tscript tr;
    tr_init(&tr, "FRI/FS");
    bind_statement(&tr, p, n0);

    size_t m0 = p->schedule[0], nn0 = n0 / m0;
    if (stir) {
        /* coset-packed f0 tree (fri.rs:1420-1431) */
        size_t a = pick_arity2(nn0, 16), d = merkle_depth(nn0, a);
        mtree t;
        mt_init(&t, a, d, 0xFF, nn0);
        hash_packed_leaves(&t, f0, nn0, m0);
        mt_finalize(&t, s->root_f0);
        mt_free(&t);
    } else {
        /* one value per leaf (fri.rs:1432-1440, f0_tree_config :1068-1072) */
        size_t a = pick_arity2(n0, 16), d = merkle_depth(n0, a);
        mtree t;
        mt_init(&t, a, d, 0xFF, n0);
        hash_leaves(&t, n0, f0, 1, 1);
        mt_finalize(&t, s->root_f0);
        mt_free(&t);
    }
    tr_absorb(&tr, s->root_f0, 32);
    fp6 z = challenge_ext(&tr, "z_fp3");
    const fp6 z_ext = z;
    uint8_t seed_bytes[32];
    tr_challenge_bytes(&tr, "FRI/seed", 8, seed_bytes); /* trace_hash: unused by MerkleTreeChannel */

    s->f[0] = (fp6 *)malloc(sizeof(fp6) * n0);
    for (size_t i = 0; i < n0; i++) s->f[0][i] = fp6_from_base(f0[i]);
    s->sz[0] = n0;
    size_t cur = n0;
    for (size_t ell = 0; ell < L; ell++) {
        size_t m = p->schedule[ell], nn = cur / m;
        int final_cc = use_cc && ell == L - 1;
        fp6 alpha = fp6_zero();
        if (!final_cc) alpha = challenge_ext(&tr, "alpha");
        fp6 *ce = s->coset_evals + s->ce_off[ell];
        if (stir) {
            ood_coset_evals(s->f[ell], cur, z, m, ce);
            s->fz[ell] = ce[0];
            z = fp6_pow(z, m);
        } else {
            /* compute_q_layer_ext_on_ext at the fixed z (fri.rs:1481-1484) */
            s->fz[ell] = poly_eval_ext(s->f[ell], cur, z_ext);
            s->q[ell] = (fp6 *)malloc(sizeof(fp6) * cur);
            q_job qj = {s->f[ell], z_ext, s->fz[ell], s->q[ell], gl_root_of_unity(cur)};
            ref_par_for(cur, PAR_FRI, q_range, &qj);
        }
        fp6 *tuples = (fp6 *)malloc(sizeof(fp6) * cur);
        coset_coeffs(s->f[ell], cur, m, tuples);
        s->f[ell + 1] = (fp6 *)malloc(sizeof(fp6) * nn);
        /* s_l (classic) is the fold repeated over the fibres, so the fold with
         * this layer's alpha comes before the layer tree (fri.rs:1519-1524) */
        if (!final_cc) fold_tuples(tuples, nn, m, alpha, s->f[ell + 1]);
        {
            size_t a = pick_arity2(cur, m), d = merkle_depth(cur, a);
            mtree t;
            mt_init(&t, a, d, ell, cur);
            if (stir) {
                /* STIR leaf = 6 components of f (fri.rs:1364-1366,1506-1516) */
                hash_leaves(&t, cur, s->f[ell][0].c, 6, 6);
            } else {
                uint64_t *lv = (uint64_t *)malloc(8 * 18 * cur);
                classic_leaves(s, ell, cur, m, final_cc, lv);
                hash_leaves(&t, cur, lv, 18, 18);
                free(lv);
            }
            mt_finalize(&t, s->roots[ell]);
            mt_free(&t);
        }
        if (stir)
            for (size_t j = 0; j < m; j++) tr_absorb_fp6(&tr, ce[j]);
        else
            tr_absorb_fp6(&tr, s->fz[ell]);
        tr_absorb(&tr, s->roots[ell], 32);
        if (final_cc) {
            /* coefficient-tuple tree, label 0xFE (fri.rs:860-868,1649-1664) */
            size_t a = pick_arity2(nn, 16), d = merkle_depth(nn, a);
            mtree t;
            mt_init(&t, a, d, 0xFE, nn);
            hash_leaves(&t, nn, (const uint64_t *)tuples, 6 * m, 6 * m);
            mt_finalize(&t, s->coeff_root);
            mt_free(&t);
            s->has_coeff = 1;
            tr_absorb(&tr, s->coeff_root, 32);
            alpha = challenge_ext(&tr, "alpha");
            (void)challenge_ext(&tr, "beta_deg");
            fold_tuples(tuples, nn, m, alpha, s->f[ell + 1]);
            s->alpha_final = alpha;
        }
        s->sz[ell + 1] = nn;
        if (final_cc) { s->coeff_tuples = tuples; s->n_final_tuples = nn; s->m_final = m; }
        else free(tuples);
        cur = nn;
    }
    /* final polynomial: first d_final coefficients of the last layer (fri.rs:1919-1937) */
    size_t nf = s->sz[L];
    {
        fp6 *co = (fp6 *)malloc(sizeof(fp6) * nf);
        uint64_t *comp = (uint64_t *)malloc(8 * nf);
        for (int c = 0; c < 6; c++) {
            for (size_t i = 0; i < nf; i++) comp[i] = s->f[L][i].c[c];
            if (nf > 1) ref_ntt_inplace(comp, nf, 1);
            for (size_t i = 0; i < nf; i++) co[i].c[c] = comp[i];
        }
        size_t d = p->d_final < nf ? p->d_final : nf;
        if (d > 64) d = 64;
        for (size_t i = 0; i < d; i++) s->final_coeffs[i] = co[i];
        s->n_final_coeffs = d;
        free(co); free(comp);
    }
    /* the replayed transcript equals this one: challenges never mutate it (fri.rs:1939-1993) */
    for (size_t i = 0; i < s->n_final_coeffs; i++) tr_absorb_fp6(&tr, s->final_coeffs[i]);
    s->query_seed = safe_field_challenge(&tr, "query_seed", 10);
    return 0;
}

int ref_fri_layers(const uint64_t *f0, size_t n0, const ref_fri_params *p, uint64_t *layers_out, uint8_t *roots_out, uint8_t root_f0[32]) {
    fri_state s;
    int rc = fri_build(f0, n0, p, &s);
    if (rc) return rc;
    size_t off = 0;
    for (size_t l = 1; l <= s.L; l++) {
        memcpy(layers_out + off, s.f[l], sizeof(fp6) * s.sz[l]);
        off += 6 * s.sz[l];
    }
    memcpy(roots_out, s.roots, 32 * s.L);
    memcpy(root_f0, s.root_f0, 32);
    fs_free(&s);
    return 0;
}

/* raw query index q (fri.rs:1044-1058): StdRng seeded by index_seed(seed, 0, q) */
static uint64_t query_raw(uint64_t query_seed, size_t q, uint64_t npow2, size_t n0) {
    uint64_t fl[3] = {query_seed, 0, gl_from_u64(q)};
    uint64_t seed = tr_hash_fields_tagged("FRI/index", fl, 3);
    return (ref_stdrng_first_u64(seed) & (npow2 - 1)) % n0;
}

/* Classic-FRI openings (fri_prove_queries, fri.rs:1724-1807, and the payloads
 * of deep_fri_prove :2008-2032): the f0 tree and every layer tree are rebuilt
 * as committed; r queries, each followed down the layers by i <- i mod n/m.
 * Writes layer_proofs, f0_openings and queries of the DeepFriProof. */
static void classic_queries(const fri_state *s, const uint64_t *f0, const ref_fri_params *p, uint64_t npow2,
                            bbuf *b, uint64_t *raw_out) {
    const size_t L = s->L, n0 = s->n0, r = p->r;
    const int use_cc = p->coeff_commit_final && L > 0;
    size_t *idx = (size_t *)malloc(sizeof(size_t) * (r ? r : 1) * (L + 1));
    for (size_t q = 0; q < r; q++) {
        size_t i = query_raw(s->query_seed, q, npow2, n0);
        if (raw_out) raw_out[q] = i;
        for (size_t ell = 0; ell < L; ell++) {
            idx[q * (L + 1) + ell] = i;
            i %= s->sz[ell] / p->schedule[ell];
        }
        idx[q * (L + 1) + L] = i;
    }
    mtree tf;
    {
        size_t a = pick_arity2(n0, 16), d = merkle_depth(n0, a);
        uint8_t rt[32];
        mt_init(&tf, a, d, 0xFF, n0);
        hash_leaves(&tf, n0, f0, 1, 1);
        mt_finalize(&tf, rt);
    }
    /* layer_proofs.layers: per layer, the r openings */
    bb_u64(b, L);
    for (size_t ell = 0; ell < L; ell++) {
        size_t cur = s->sz[ell], m = p->schedule[ell];
        int final_cc = use_cc && ell == L - 1;
        size_t a = pick_arity2(cur, m), d = merkle_depth(cur, a);
        uint64_t *lv = (uint64_t *)malloc(8 * 18 * cur);
        uint8_t rt[32];
        mtree t;
        classic_leaves(s, ell, cur, m, final_cc, lv);
        mt_init(&t, a, d, ell, cur);
        hash_leaves(&t, cur, lv, 18, 18);
        mt_finalize(&t, rt);
        free(lv);
        bb_u64(b, r);
        for (size_t q = 0; q < r; q++) mt_open_ser(&t, idx[q * (L + 1) + ell], b);
        mt_free(&t);
    }
    /* f0_openings */
    bb_u64(b, r);
    for (size_t q = 0; q < r; q++) mt_open_ser(&tf, idx[q * (L + 1)], b);
    /* queries: FriQueryPayload { per_layer_refs, per_layer_payloads, f0_opening, final_index } */
    bb_u64(b, r);
    for (size_t q = 0; q < r; q++) {
        bb_u64(b, L);
        for (size_t ell = 0; ell < L; ell++) {
            size_t i = idx[q * (L + 1) + ell], m = p->schedule[ell], nn = s->sz[ell] / m;
            bb_u64(b, i);
            bb_u64(b, i % m);
            bb_u64(b, i % nn);
            bb_u64(b, 0);
        }
        bb_u64(b, L);
        for (size_t ell = 0; ell < L; ell++) {
            size_t i = idx[q * (L + 1) + ell], nn = s->sz[ell] / p->schedule[ell];
            int final_cc = use_cc && ell == L - 1;
            bb_fp6(b, s->f[ell][i]);
            bb_fp6(b, final_cc ? fp6_zero() : s->f[ell + 1][i % nn]);
            bb_fp6(b, s->q[ell][i]);
        }
        mt_open_ser(&tf, idx[q * (L + 1)], b);
        bb_u64(b, idx[q * (L + 1) + L]);
    }
    mt_free(&tf);
    free(idx);
}

int ref_deep_fri_prove_raw(const uint64_t *f0, size_t n0, const ref_fri_params *p, uint8_t **out, size_t *out_len,
                           uint64_t *raw_out) {
    fri_state s;
    int rc = fri_build(f0, n0, p, &s);
    if (rc) return rc;
    size_t L = s.L;
    bbuf b = {0};
    bb_put(&b, s.root_f0, 32);
    bb_u64(&b, L);
    for (size_t l = 0; l < L; l++) bb_put(&b, s.roots[l], 32);
    uint64_t npow2 = 1;
    while (npow2 < n0) npow2 <<= 1;
    if (p->stir) {
        bb_u64(&b, 0); /* layer_proofs.layers */
        bb_u64(&b, 0); /* f0_openings */
        bb_u64(&b, 0); /* queries */
    } else {
        classic_queries(&s, f0, p, npow2, &b, NULL); /* the raw indices are the STIR queries' */
    }
    bb_u64(&b, L);
    for (size_t l = 0; l < L; l++) bb_fp6(&b, s.fz[l]);
    bb_u64(&b, s.n_final_coeffs);
    for (size_t i = 0; i < s.n_final_coeffs; i++) bb_fp6(&b, s.final_coeffs[i]);
    bb_u64(&b, n0);
    bb_u64(&b, gl_root_of_unity(n0));
    if (s.has_coeff) {
        bb_u8(&b, 1);
        bb_u64(&b, s.n_final_tuples);
        for (size_t t = 0; t < s.n_final_tuples; t++) {
            bb_u64(&b, s.m_final);
            for (size_t k = 0; k < s.m_final; k++) bb_fp6(&b, s.coeff_tuples[t * s.m_final + k]);
        }
        bb_u8(&b, 1);
        bb_put(&b, s.coeff_root, 32);
    } else {
        bb_u8(&b, 0);
        bb_u8(&b, 0);
    }
    if (!p->stir) {
        bb_u8(&b, 0); /* stir_coset_evals: None */
        bb_u8(&b, 0); /* stir_proximity_queries: None */
        fs_free(&s);
        *out = b.p;
        *out_len = b.len;
        return 0;
    }
    bb_u8(&b, 1);
    bb_u64(&b, L);
    for (size_t l = 0; l < L; l++) {
        size_t m = p->schedule[l];
        bb_u64(&b, m);
        for (size_t j = 0; j < m; j++) bb_fp6(&b, s.coset_evals[s.ce_off[l] + j]);
    }
    /* STIR proximity queries (fri.rs:1813-1893): the f0 packed tree and the
     * layer-1 tree are rebuilt exactly as committed. */
    size_t m0 = p->schedule[0], nn0 = n0 / m0;
    mtree t0, t1;
    {
        size_t a = pick_arity2(nn0, 16), d = merkle_depth(nn0, a);
        uint8_t r[32];
        mt_init(&t0, a, d, 0xFF, nn0);
        hash_packed_leaves(&t0, f0, nn0, m0);
        mt_finalize(&t0, r);
    }
    if (L >= 2) {
        size_t n1 = s.sz[1], a = pick_arity2(n1, p->schedule[1]), d = merkle_depth(n1, a);
        uint8_t r[32];
        mt_init(&t1, a, d, 1, n1);
        hash_leaves(&t1, n1, s.f[1][0].c, 6, 6);
        mt_finalize(&t1, r);
    }
    bb_u8(&b, 1);
    bb_u64(&b, p->s0);
    for (size_t q = 0; q < p->s0; q++) {
        uint64_t fl[3] = {s.query_seed, 0, gl_from_u64(q)};
        uint64_t seed = tr_hash_fields_tagged("FRI/index", fl, 3);
        uint64_t raw = (ref_stdrng_first_u64(seed) & (npow2 - 1)) % n0;
        uint64_t base = raw % nn0;
        if (raw_out) raw_out[q] = raw;
        bb_u64(&b, base);
        bb_u64(&b, raw);
        bb_u64(&b, m0);
        for (size_t j = 0; j < m0; j++) bb_u64(&b, base + j * nn0);
        bb_u64(&b, m0);
        for (size_t j = 0; j < m0; j++) bb_u64(&b, f0[base + j * nn0]);
        mt_open_ser(&t0, base, &b);
        if (L >= 2) {
            bb_fp6(&b, s.f[1][base]);
            bb_u8(&b, 1);
            mt_open_ser(&t1, base, &b);
        } else {
            bb_fp6(&b, fp6_zero());
            bb_u8(&b, 0);
        }
    }
    mt_free(&t0);
    if (L >= 2) mt_free(&t1);
    fs_free(&s);
    *out = b.p;
    *out_len = b.len;
    return 0;
}

int ref_deep_fri_prove(const uint64_t *f0, size_t n0, const ref_fri_params *p, uint8_t **out, size_t *out_len) {
    return ref_deep_fri_prove_raw(f0, n0, p, out, out_len, NULL);
}

void ref_free(void *p) { free(p); }

typedef struct {
    mtree *t;
    const uint64_t *lde;
    uint64_t *f0;
    size_t w, n0;
} row_job;
static void row_range(void *ctx, size_t lo, size_t hi) {
    row_job *J = (row_job *)ctx;
    uint64_t *row = (uint64_t *)malloc(8 * J->w);
    for (size_t i = lo; i < hi; i++) {
        uint64_t acc = 0;
        for (size_t c = 0; c < J->w; c++) {
            row[c] = J->lde[c * J->n0 + i];
            acc = gl_add(acc, gl_mul(row[c], c + 1));
        }
        J->f0[i] = acc;
        leaf_hash(2, 0, i, row, J->w, J->t->lv[0] + 32 * i);
    }
    free(row);
}

/* Prover-core microbench: coset LDE -> row tree (DsLabel arity 2, label 0,
 * leaf = the w values of one LDE row) -> f0 = sum_c (c+1) LDE_c ->
 * deep_fri_prove(f0) bound to the row root -> one row opening per STIR query
 * at its raw query index. Output: proof || u64 s0 || s0 x (Vec<F> row, MerkleOpening). */
int ref_prover_core(const uint64_t *trace, size_t w, size_t n_trace, size_t blowup, uint64_t shift,
                    const ref_fri_params *p, uint8_t row_root[32], uint64_t *f0_out,
                    uint8_t **proof, size_t *proof_len) {
    size_t n0 = n_trace * blowup;
    uint64_t *lde = (uint64_t *)malloc(8 * w * n0);
    if (!lde) return -3;
    int rc = ref_lde_columns(trace, w, n_trace, blowup, shift, lde);
    if (rc) { free(lde); return rc; }
    size_t depth = 0;
    while (((size_t)1 << depth) < n0) depth++;
    if (depth == 0) depth = 1;
    mtree t;
    mt_init(&t, 2, depth, 0, n0);
    uint64_t *f0 = f0_out ? f0_out : (uint64_t *)malloc(8 * n0);
    row_job rj = {&t, lde, f0, w, n0};
    ref_par_for(n0, PAR_MERKLE, row_range, &rj); /* row leaves: push_leaves_parallel (lib.rs:390-409) */
    mt_finalize(&t, row_root);
    ref_fri_params q = *p;
    q.public_inputs_hash = row_root;
    uint64_t *raw = (uint64_t *)malloc(8 * (p->s0 ? p->s0 : 1));
    uint8_t *pb = NULL;
    size_t pl = 0;
    rc = ref_deep_fri_prove_raw(f0, n0, &q, &pb, &pl, raw);
    if (rc == 0) {
        bbuf b = {0};
        bb_put(&b, pb, pl);
        bb_u64(&b, p->s0);
        for (size_t k = 0; k < p->s0; k++) {
            bb_u64(&b, w);
            for (size_t c = 0; c < w; c++) bb_u64(&b, lde[c * n0 + raw[k]]);
            mt_open_ser(&t, raw[k], &b);
        }
        *proof = b.p;
        *proof_len = b.len;
    }
    free(pb); free(raw); free(lde);
    if (!f0_out) free(f0);
    mt_free(&t);
    return rc;
}
