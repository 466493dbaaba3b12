/*
 * verify.c -- oracle restatement of deep_fri_verify and the Merkle opening
 * check, over the ark-serialize compressed DeepFriProof<SexticExt> bytes.
 *
 * Test infrastructure only (the oracle); the product path never calls it.
 *
 *   proof types : crates/deep_ali/src/fri.rs:1235-1300 (DeepFriProof and its
 *                 payloads), crates/merkle/src/lib.rs:136-141 (MerkleOpening);
 *                 ark-serialize 0.4.2 CanonicalDeserialize, Compress::Yes,
 *                 Validate::Yes (swarm-verify/src/main.rs:243-245): usize as
 *                 u64 LE, Vec as u64 length + items, Option as a 0/1 byte,
 *                 [u8; 32] as 32 bytes, a field element as its canonical 8
 *                 LE bytes (a value >= p fails)
 *   verify      : fri.rs:2136-2710 (deep_fri_verify), merkle lib.rs:754-799
 *                 (verify_opening), :806-812 (compute_leaf_hash)
 *   helpers     : fri.rs:468-502 (interpolate_stir_coset), :714-752
 *                 (interpolate_coset_ext), :811-858 (batched_degree_check_ext),
 *                 :908-915 (eval_final_poly_ext), :1023-1033 (layer sizes)
 *
 * Where the reference would panic on a malformed proof (an index past a
 * vector's end, an unwrap of None, a zero inverse) this restatement returns
 * "reject"; a byte string that does not deserialize returns -1.
 */
#include "refcpu.h"
#include "refcpu_int.h"
#include <stdio.h>
#include <stdlib.h>

/* ------------------------------------------------------------ arena + reader */
typedef struct ablk { struct ablk *next; } ablk;
typedef struct { ablk *head; } arena;
static void *ar_alloc(arena *a, size_t n) {
    ablk *b = (ablk *)malloc(sizeof(ablk) + (n ? n : 1));
    if (!b) return NULL;
    b->next = a->head;
    a->head = b;
    return b + 1;
}
static void ar_free(arena *a) {
    while (a->head) { ablk *n = a->head->next; free(a->head); a->head = n; }
}

typedef struct { const uint8_t *p; size_t len, pos; int bad; arena *ar; } rdr;
static uint64_t rd_u64(rdr *r) {
    if (r->bad || r->len - r->pos < 8) { r->bad = 1; return 0; }
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v |= (uint64_t)r->p[r->pos + i] << (8 * i);
    r->pos += 8;
    return v;
}
static int rd_bool(rdr *r) {
    if (r->bad || r->pos >= r->len) { r->bad = 1; return 0; }
    uint8_t v = r->p[r->pos++];
    if (v > 1) r->bad = 1;
    return v;
}
static void rd_bytes(rdr *r, uint8_t *out, size_t n) {
    if (r->bad || r->len - r->pos < n) { r->bad = 1; memset(out, 0, n); return; }
    memcpy(out, r->p + r->pos, n);
    r->pos += n;
}
static uint64_t rd_fe(rdr *r) {
    uint64_t v = rd_u64(r);
    if (v >= GL_P) r->bad = 1;
    return v;
}
static fp6 rd_fp6(rdr *r) {
    fp6 v;
    for (int i = 0; i < 6; i++) v.c[i] = rd_fe(r);
    return v;
}
/* a Vec length whose items take at least min_item bytes each */
static size_t rd_len(rdr *r, size_t min_item) {
    uint64_t n = rd_u64(r);
    if (r->bad) return 0;
    if (min_item && n > (r->len - r->pos) / min_item) { r->bad = 1; return 0; }
    return (size_t)n;
}
static void *rd_arr(rdr *r, size_t n, size_t sz) {
    void *p = ar_alloc(r->ar, n * sz);
    if (!p) r->bad = 1;
    return p;
}

/* ------------------------------------------------------------ proof types */
typedef struct {
    uint8_t leaf[32];
    size_t nlev;
    size_t *cnt;       /* siblings per level */
    uint8_t (**sib)[32];
    uint64_t index;
} mopen;
typedef struct { uint64_t i, child_pos, parent_index, parent_pos; } qref;
typedef struct { fp6 f, s, q; } lpay;
typedef struct {
    size_t nref; qref *refs;
    size_t npay; lpay *pay;
    mopen f0;
    uint64_t final_index;
} fquery;
typedef struct {
    uint64_t base, raw;
    size_t nfi; uint64_t *fidx;
    size_t nfv; uint64_t *fval;
    mopen f0p;
    fp6 fnext;
    int has_l1; mopen l1;
} sprox;
typedef struct {
    uint8_t root_f0[32];
    size_t nroots; uint8_t (*roots)[32];
    size_t nlp; size_t *lp_n; mopen **lp;
    size_t nf0o; mopen *f0o;
    size_t nq; fquery *q;
    size_t nfz; fp6 *fz;
    size_t nfin; fp6 *fin;
    uint64_t n0, omega0;
    int has_ct; size_t nct; size_t *ct_n; fp6 **ct;
    int has_croot; uint8_t croot[32];
    int has_ce; size_t nce; size_t *ce_n; fp6 **ce;
    int has_sp; size_t nsp; sprox *sp;
} dproof;

static void rd_open(rdr *r, mopen *o) {
    rd_bytes(r, o->leaf, 32);
    o->nlev = rd_len(r, 8);
    o->cnt = (size_t *)rd_arr(r, o->nlev, sizeof(size_t));
    o->sib = (uint8_t(**)[32])rd_arr(r, o->nlev, sizeof(void *));
    for (size_t l = 0; l < o->nlev && !r->bad; l++) {
        o->cnt[l] = rd_len(r, 32);
        o->sib[l] = (uint8_t(*)[32])rd_arr(r, o->cnt[l], 32);
        for (size_t j = 0; j < o->cnt[l] && !r->bad; j++) rd_bytes(r, o->sib[l][j], 32);
    }
    o->index = rd_u64(r);
}

static int parse_proof(const uint8_t *bytes, size_t len, arena *ar, dproof *d) {
    rdr r = {bytes, len, 0, 0, ar};
    memset(d, 0, sizeof *d);
    rd_bytes(&r, d->root_f0, 32);
    d->nroots = rd_len(&r, 32);
    d->roots = (uint8_t(*)[32])rd_arr(&r, d->nroots, 32);
    for (size_t i = 0; i < d->nroots && !r.bad; i++) rd_bytes(&r, d->roots[i], 32);
    d->nlp = rd_len(&r, 8);
    d->lp_n = (size_t *)rd_arr(&r, d->nlp, sizeof(size_t));
    d->lp = (mopen **)rd_arr(&r, d->nlp, sizeof(mopen *));
    for (size_t l = 0; l < d->nlp && !r.bad; l++) {
        d->lp_n[l] = rd_len(&r, 48);
        d->lp[l] = (mopen *)rd_arr(&r, d->lp_n[l], sizeof(mopen));
        for (size_t j = 0; j < d->lp_n[l] && !r.bad; j++) rd_open(&r, &d->lp[l][j]);
    }
    d->nf0o = rd_len(&r, 48);
    d->f0o = (mopen *)rd_arr(&r, d->nf0o, sizeof(mopen));
    for (size_t j = 0; j < d->nf0o && !r.bad; j++) rd_open(&r, &d->f0o[j]);
    d->nq = rd_len(&r, 72);
    d->q = (fquery *)rd_arr(&r, d->nq, sizeof(fquery));
    for (size_t j = 0; j < d->nq && !r.bad; j++) {
        fquery *q = &d->q[j];
        q->nref = rd_len(&r, 32);
        q->refs = (qref *)rd_arr(&r, q->nref, sizeof(qref));
        for (size_t k = 0; k < q->nref && !r.bad; k++) {
            q->refs[k].i = rd_u64(&r);
            q->refs[k].child_pos = rd_u64(&r);
            q->refs[k].parent_index = rd_u64(&r);
            q->refs[k].parent_pos = rd_u64(&r);
        }
        q->npay = rd_len(&r, 144);
        q->pay = (lpay *)rd_arr(&r, q->npay, sizeof(lpay));
        for (size_t k = 0; k < q->npay && !r.bad; k++) {
            q->pay[k].f = rd_fp6(&r);
            q->pay[k].s = rd_fp6(&r);
            q->pay[k].q = rd_fp6(&r);
        }
        rd_open(&r, &q->f0);
        q->final_index = rd_u64(&r);
    }
    d->nfz = rd_len(&r, 48);
    d->fz = (fp6 *)rd_arr(&r, d->nfz, sizeof(fp6));
    for (size_t j = 0; j < d->nfz && !r.bad; j++) d->fz[j] = rd_fp6(&r);
    d->nfin = rd_len(&r, 48);
    d->fin = (fp6 *)rd_arr(&r, d->nfin, sizeof(fp6));
    for (size_t j = 0; j < d->nfin && !r.bad; j++) d->fin[j] = rd_fp6(&r);
    d->n0 = rd_u64(&r);
    d->omega0 = rd_fe(&r);
    d->has_ct = rd_bool(&r);
    if (d->has_ct) {
        d->nct = rd_len(&r, 8);
        d->ct_n = (size_t *)rd_arr(&r, d->nct, sizeof(size_t));
        d->ct = (fp6 **)rd_arr(&r, d->nct, sizeof(fp6 *));
        for (size_t j = 0; j < d->nct && !r.bad; j++) {
            d->ct_n[j] = rd_len(&r, 48);
            d->ct[j] = (fp6 *)rd_arr(&r, d->ct_n[j], sizeof(fp6));
            for (size_t k = 0; k < d->ct_n[j] && !r.bad; k++) d->ct[j][k] = rd_fp6(&r);
        }
    }
    d->has_croot = rd_bool(&r);
    if (d->has_croot) rd_bytes(&r, d->croot, 32);
    d->has_ce = rd_bool(&r);
    if (d->has_ce) {
        d->nce = rd_len(&r, 8);
        d->ce_n = (size_t *)rd_arr(&r, d->nce, sizeof(size_t));
        d->ce = (fp6 **)rd_arr(&r, d->nce, sizeof(fp6 *));
        for (size_t j = 0; j < d->nce && !r.bad; j++) {
            d->ce_n[j] = rd_len(&r, 48);
            d->ce[j] = (fp6 *)rd_arr(&r, d->ce_n[j], sizeof(fp6));
            for (size_t k = 0; k < d->ce_n[j] && !r.bad; k++) d->ce[j][k] = rd_fp6(&r);
        }
    }
    d->has_sp = rd_bool(&r);
    if (d->has_sp) {
        d->nsp = rd_len(&r, 8);
        d->sp = (sprox *)rd_arr(&r, d->nsp, sizeof(sprox));
        for (size_t j = 0; j < d->nsp && !r.bad; j++) {
            sprox *s = &d->sp[j];
            s->base = rd_u64(&r);
            s->raw = rd_u64(&r);
            s->nfi = rd_len(&r, 8);
            s->fidx = (uint64_t *)rd_arr(&r, s->nfi, 8);
            for (size_t k = 0; k < s->nfi && !r.bad; k++) s->fidx[k] = rd_u64(&r);
            s->nfv = rd_len(&r, 8);
            s->fval = (uint64_t *)rd_arr(&r, s->nfv, 8);
            for (size_t k = 0; k < s->nfv && !r.bad; k++) s->fval[k] = rd_fe(&r);
            rd_open(&r, &s->f0p);
            s->fnext = rd_fp6(&r);
            s->has_l1 = rd_bool(&r);
            if (s->has_l1) rd_open(&r, &s->l1);
        }
    }
    return r.bad ? -1 : 0;
}

/* ------------------------------------------------------------ transcript (fri.rs:73-130, :1372-1393) */
typedef struct { sha3_ctx h; } vts;
static void vt_init(vts *t) {
    sha3_init(&t->h);
    sha3_update(&t->h, "FSv1-TRANSCRIPT-INIT", 20);
    sha3_update(&t->h, "FRI/FS", 6);
}
static void vt_absorb(vts *t, const void *b, size_t n) {
    sha3_update(&t->h, "FSv1-ABSORB-BYTES", 17);
    sha3_update(&t->h, b, n);
}
static void vt_field(vts *t, uint64_t x) {
    uint8_t le[8];
    for (int i = 0; i < 8; i++) le[i] = (uint8_t)(x >> (8 * i));
    vt_absorb(t, le, 8);
}
static void vt_fp6(vts *t, fp6 v) { for (int i = 0; i < 6; i++) vt_field(t, v.c[i]); }
static uint64_t vt_challenge(const vts *t, const char *label, size_t ln) {
    sha3_ctx c = t->h;
    uint8_t d[32];
    sha3_update(&c, "FSv1-CHALLENGE", 14);
    sha3_update(&c, label, ln);
    sha3_final(&c, d);
    uint64_t acc = 0;
    int end = 32;
    while (end > 0) {
        int start = end - 7 < 0 ? 0 : end - 7, l = end - start;
        uint64_t val = 0;
        for (int i = 0; i < l; i++) val |= (uint64_t)d[start + i] << (8 * i);
        acc = gl_add(gl_mul(acc, (uint64_t)1 << (8 * l)), val);
        end = start;
    }
    return acc;
}
static fp6 vt_ext(const vts *t, const char *tag) {
    fp6 r;
    char sub[64];
    for (int i = 0; i < 6; i++) {
        int n = snprintf(sub, sizeof sub, "%s/c%d", tag, i);
        r.c[i] = vt_challenge(t, sub, (size_t)n);
    }
    return r;
}
static uint64_t index_raw(uint64_t seed, size_t q, size_t n0) {
    vts t;
    vt_init(&t);
    vt_absorb(&t, "FRI/index", 9);
    vt_field(&t, seed);
    vt_field(&t, 0);
    vt_field(&t, gl_from_u64(q));
    uint64_t s = vt_challenge(&t, "out", 3);
    uint64_t npow2 = 1;
    while (npow2 < n0) npow2 <<= 1;
    return (ref_stdrng_first_u64(s) & (npow2 - 1)) % n0;
}

/* ------------------------------------------------------------ Merkle (merkle lib.rs:754-812) */
static void dsl(uint8_t out[32], uint64_t arity, uint64_t level, uint64_t pos, uint64_t label) {
    uint64_t v[4] = {arity, level, pos, label};
    for (int k = 0; k < 4; k++)
        for (int i = 0; i < 8; i++) out[8 * k + i] = (uint8_t)(v[k] >> (8 * i));
}
static void leaf_hash_v(size_t arity, uint64_t label, uint64_t pos, const uint64_t *vals, size_t nv, uint8_t out[32]) {
    sha3_ctx c;
    uint8_t ds[32];
    dsl(ds, arity, 0xFFFFFFFFULL, pos, label);
    sha3_init(&c);
    sha3_update(&c, ds, 32);
    for (size_t i = 0; i < nv; i++) {
        uint8_t le[8];
        for (int b = 0; b < 8; b++) le[b] = (uint8_t)(vals[i] >> (8 * b));
        sha3_update(&c, le, 8);
    }
    sha3_final(&c, out);
}
/* constant-arity config: layer_arities = [arity; depth] */
static int verify_open(size_t arity, size_t depth, uint64_t label, const uint8_t root[32], const mopen *o) {
    uint8_t cur[32];
    memcpy(cur, o->leaf, 32);
    uint64_t idx = o->index;
    for (size_t level = 0; level < o->nlev; level++) {
        if (level >= depth) return 0; /* layer_arities[level] out of range: the reference panics */
        uint64_t pos = idx % arity;
        sha3_ctx c;
        uint8_t ds[32];
        dsl(ds, arity, level + 1, idx / arity, label);
        sha3_init(&c);
        sha3_update(&c, ds, 32);
        size_t used = 0;
        for (size_t i = 0; i < arity; i++) {
            if (i == pos) {
                sha3_update(&c, cur, 32);
            } else {
                if (used >= o->cnt[level]) return 0;
                sha3_update(&c, o->sib[level][used++], 32);
            }
        }
        sha3_final(&c, cur);
        idx /= arity;
    }
    return memcmp(cur, root, 32) == 0;
}

int ref_merkle_verify_opening(size_t arity, size_t depth, uint64_t label, const uint8_t root[32],
                              const uint8_t *opening, size_t len) {
    arena ar = {0};
    rdr r = {opening, len, 0, 0, &ar};
    mopen o;
    rd_open(&r, &o);
    int ok = r.bad ? -1 : verify_open(arity, depth, label, root, &o);
    ar_free(&ar);
    return ok;
}

static size_t pick_ar(size_t n, size_t m) {
    static const size_t as[6] = {128, 64, 32, 16, 8, 4};
    for (int i = 0; i < 6; i++)
        if (m >= as[i] && n % as[i] == 0) return as[i];
    return 2; /* pick_arity_for_layer(..).max(2) */
}
static size_t mdepth(size_t leaves, size_t arity) {
    size_t d = 1, cur = leaves;
    while (cur > arity) { cur = (cur + arity - 1) / arity; d++; }
    return d;
}

/* ------------------------------------------------------------ field helpers */
/* Domain::<F>::new(n).group_gen: the generator of the 2^ceil(log2 n) subgroup */
static int group_gen(uint64_t n, uint64_t *g) {
    uint64_t np = 1;
    int lg = 0;
    while (np < n) { np <<= 1; lg++; if (lg > 32) return -1; }
    *g = gl_root_of_unity((size_t)np);
    return 0;
}
static fp6 ext_pows_dot(const fp6 *c, size_t m, fp6 alpha) {
    fp6 acc = fp6_zero(), ap = fp6_one();
    for (size_t k = 0; k < m; k++) {
        acc = fp6_add(acc, fp6_mul(c[k], ap));
        ap = fp6_mul(ap, alpha);
    }
    return acc;
}
static fp6 eval_final(const fp6 *c, size_t n, fp6 x) {
    fp6 r = fp6_zero();
    for (size_t k = n; k-- > 0;) r = fp6_add(fp6_mul(r, x), c[k]);
    return r;
}
static int fp6_zero_p(fp6 a) { return fp6_is_zero(a); }
/* (1/m) sum_j v_j zeta^(-jk), then coefficient k times shift_inv^k: the
 * monomial coefficients of the interpolant on {s zeta^j} (fri.rs:468-502 with
 * s in Fp6, :714-752 with s = omega^b) */
static void interp_coset(const fp6 *v, size_t m, uint64_t zeta, fp6 sinv, fp6 *out) {
    uint64_t zi = gl_inv(zeta), minv = gl_inv(m);
    for (size_t k = 0; k < m; k++) {
        fp6 acc = fp6_zero();
        uint64_t w = 1, wk = gl_pow(zi, k);
        for (size_t j = 0; j < m; j++) {
            acc = fp6_add(acc, fp6_scale(v[j], w));
            w = gl_mul(w, wk);
        }
        out[k] = fp6_scale(acc, minv);
    }
    fp6 sp = fp6_one();
    for (size_t k = 0; k < m; k++) {
        out[k] = fp6_mul(out[k], sp);
        sp = fp6_mul(sp, sinv);
    }
}

/* batched_degree_check_ext (fri.rs:811-858) */
static int degree_check(fp6 **ct, size_t nf, size_t m, fp6 beta, size_t d_final) {
    if (nf == 0) return 1;
    uint64_t *comp = (uint64_t *)malloc(8 * nf * 6);
    for (size_t b = 0; b < nf; b++) {
        fp6 g = ext_pows_dot(ct[b], m, beta);
        for (int j = 0; j < 6; j++) comp[j * nf + b] = g.c[j];
    }
    int ok = 1;
    for (int j = 0; j < 6 && ok; j++) {
        ref_ntt_inplace(comp + j * nf, nf, 1); /* Domain::new(n_final).ifft */
        for (size_t k = d_final; k < nf; k++)
            if (comp[j * nf + k]) { ok = 0; break; }
    }
    free(comp);
    return ok;
}

/* ------------------------------------------------------------ deep_fri_verify (fri.rs:2136-2710) */
static int verify(const ref_fri_params *p, const dproof *d) {
    const size_t L = p->n_layers;
    const size_t *sch = p->schedule;
    /* layer sizes (fri.rs:1023-1033; a non-dividing schedule panics there) */
    size_t *sz = (size_t *)malloc(sizeof(size_t) * (L + 1));
    int ok = 0;
    fp6 *alpha = (fp6 *)calloc(L + 1, sizeof(fp6));
    fp6 **interp = (fp6 **)calloc(L + 1, sizeof(fp6 *));
    uint64_t *omg = (uint64_t *)calloc(L + 1, sizeof(uint64_t));
    sz[0] = d->n0;
    for (size_t l = 0; l < L; l++) {
        if (sch[l] == 0 || sz[l] % sch[l]) goto out;
        sz[l + 1] = sz[l] / sch[l];
    }
    const int use_cc = p->coeff_commit_final && L > 0, stir = p->stir != 0;
    const size_t normal = use_cc ? L - 1 : L;
    vts tr;
    vt_init(&tr);
    if (p->public_inputs_hash) {
        vt_absorb(&tr, "CAIRO-PUBLIC-INPUTS-V1", 22);
        vt_absorb(&tr, p->public_inputs_hash, 32);
    }
    vt_absorb(&tr, "DEEP-FRI-STATEMENT-V2", 21);
    vt_field(&tr, gl_from_u64(d->n0));
    vt_field(&tr, L);
    for (size_t l = 0; l < L; l++) vt_field(&tr, gl_from_u64(sch[l]));
    vt_field(&tr, gl_from_u64(p->seed_z));
    vt_field(&tr, 6);
    vt_field(&tr, p->coeff_commit_final ? 1 : 0);
    vt_field(&tr, stir ? 1 : 0);
    vt_absorb(&tr, d->root_f0, 32);
    const fp6 z_ext = vt_ext(&tr, "z_fp3");
    /* trace_hash = challenge "FRI/seed": unused by the Merkle checks */
    fp6 zc = z_ext, beta = fp6_zero();
    for (size_t ell = 0; ell < L; ell++) {
        const int fin_cc = use_cc && ell == L - 1;
        if (ell < normal) alpha[ell] = vt_ext(&tr, "alpha");
        if (stir) {
            if (!d->has_ce || ell >= d->nce || d->ce_n[ell] != sch[ell]) goto out;
            size_t m = sch[ell];
            uint64_t om, zeta;
            if (group_gen(sz[ell], &om)) goto out;
            zeta = gl_pow(om, sz[ell] / m);
            if (fp6_zero_p(zc)) goto out;
            interp[ell] = (fp6 *)malloc(sizeof(fp6) * m);
            interp_coset(d->ce[ell], m, zeta, fp6_inv(zc), interp[ell]);
            zc = fp6_pow(zc, m);
            for (size_t j = 0; j < m; j++) vt_fp6(&tr, d->ce[ell][j]);
        } else {
            if (ell >= d->nfz) goto out;
            vt_fp6(&tr, d->fz[ell]);
        }
        if (ell >= d->nroots) goto out;
        vt_absorb(&tr, d->roots[ell], 32);
        if (fin_cc) {
            if (!d->has_croot) goto out;
            vt_absorb(&tr, d->croot, 32);
            alpha[ell] = vt_ext(&tr, "alpha");
            beta = vt_ext(&tr, "beta_deg");
        }
    }
    for (size_t i = 0; i < d->nfin; i++) vt_fp6(&tr, d->fin[i]);
    const uint64_t qseed = vt_challenge(&tr, "query_seed", 10);
    if (d->nfin != p->d_final) goto out;

    if (stir) {
        for (size_t ell = 0; ell < L; ell++) {
            fp6 fold = ext_pows_dot(interp[ell], sch[ell], alpha[ell]);
            fp6 expect;
            if (ell + 1 < L) {
                if (d->ce_n[ell + 1] < 1) goto out;
                expect = d->ce[ell + 1][0];
            } else {
                fp6 zf = z_ext;
                for (size_t i = 0; i < L; i++) zf = fp6_pow(zf, sch[i]);
                expect = eval_final(d->fin, d->nfin, zf);
            }
            if (!fp6_eq(fold, expect)) goto out;
        }
    }
    if (use_cc) {
        size_t nf = sz[L], mf = sch[L - 1];
        if (!d->has_ct || d->nct != nf) goto out;
        for (size_t b = 0; b < nf; b++)
            if (d->ct_n[b] != mf) goto out;
        size_t a = pick_ar(nf, 16), dep = mdepth(nf, a);
        uint8_t root[32];
        uint64_t *vals = (uint64_t *)malloc(8 * 6 * mf * (nf ? nf : 1));
        for (size_t b = 0; b < nf; b++) memcpy(vals + 6 * mf * b, d->ct[b], 48 * mf);
        ref_merkle_root(vals, nf, 6 * mf, a, dep, 0xFE, root);
        free(vals);
        if (memcmp(root, d->croot, 32) != 0) goto out;
        if (!degree_check(d->ct, nf, mf, beta, p->d_final)) goto out;
    }
    for (size_t l = 0; l < L; l++)
        if (group_gen(sz[l], &omg[l])) goto out;
    uint64_t om_final = 1;
    if (sz[L] >= 2 && group_gen(sz[L], &om_final)) goto out;

    if (stir) {
        if (!d->has_sp || d->nsp != p->s0) goto out;
        const size_t n0 = d->n0, m0 = sch[0], nn = sz[0] / m0;
        const uint64_t zeta0 = gl_pow(omg[0], nn), zi0 = gl_inv(omg[0]);
        const size_t fa = pick_ar(n0 / m0, 16), fd = mdepth(n0 / m0, fa);
        size_t l1a = 0, l1d = 0;
        if (L >= 2) { l1a = pick_ar(sz[1], sch[1]); l1d = mdepth(sz[1], l1a); }
        fp6 *fib = (fp6 *)malloc(sizeof(fp6) * m0), *co = (fp6 *)malloc(sizeof(fp6) * m0);
        for (size_t qi = 0; qi < p->s0; qi++) {
            const sprox *s = &d->sp[qi];
            uint64_t raw = index_raw(qseed, qi, n0);
            int good = s->raw == raw && s->base == raw % nn && s->nfi == m0 && s->nfv == m0;
            for (size_t j = 0; good && j < m0; j++) good = s->fidx[j] == s->base + j * nn;
            good = good && verify_open(fa, fd, 0xFF, d->root_f0, &s->f0p) && s->f0p.index == s->base;
            if (good) {
                uint8_t lh[32];
                leaf_hash_v(fa, 0xFF, s->f0p.index, s->fval, m0, lh);
                good = memcmp(lh, s->f0p.leaf, 32) == 0;
            }
            if (good) {
                for (size_t j = 0; j < m0; j++) fib[j] = fp6_from_base(s->fval[j]);
                /* omega_b = omega_0^b (b = 0 gives 1); its inverse un-shifts */
                interp_coset(fib, m0, zeta0, fp6_from_base(gl_pow(zi0, s->base)), co);
                fp6 fold = ext_pows_dot(co, m0, alpha[0]);
                if (L >= 2) {
                    good = s->has_l1 && verify_open(l1a, l1d, 1, d->roots[1], &s->l1) && s->l1.index == s->base;
                    if (good) {
                        uint8_t lh[32];
                        leaf_hash_v(l1a, 1, s->l1.index, s->fnext.c, 6, lh);
                        good = memcmp(lh, s->l1.leaf, 32) == 0 && fp6_eq(fold, s->fnext);
                    }
                } else {
                    fp6 xf = fp6_from_base(gl_pow(om_final, s->base));
                    good = fp6_eq(fold, eval_final(d->fin, d->nfin, xf));
                }
            }
            if (!good) { free(fib); free(co); goto out; }
        }
        free(fib); free(co);
        ok = 1;
        goto out;
    }

    /* classic FRI: r fold queries through every layer */
    {
        const size_t n0 = d->n0;
        const size_t fa = pick_ar(n0, 16), fd = mdepth(n0, fa);
        if (d->nq < p->r) goto out;
        for (size_t q = 0; q < p->r; q++) {
            const fquery *qp = &d->q[q];
            const uint64_t i0 = index_raw(qseed, q, n0);
            uint64_t ei = i0;
            if (qp->nref < L || qp->npay < L) goto out;
            for (size_t ell = 0; ell < L; ell++) {
                if (qp->refs[ell].i != ei) goto out;
                ei %= sz[ell] / sch[ell];
            }
            if (qp->final_index != ei) goto out;
            if (!verify_open(fa, fd, 0xFF, d->root_f0, &qp->f0)) goto out;
            const fp6 f0v = qp->pay[0].f;
            for (int c = 1; c < 6; c++)
                if (f0v.c[c]) goto out;
            {
                uint8_t lh[32];
                leaf_hash_v(fa, 0xFF, qp->f0.index, f0v.c, 1, lh);
                if (memcmp(lh, qp->f0.leaf, 32) != 0) goto out;
            }
            if (qp->f0.index != i0) goto out;
            for (size_t ell = 0; ell < L; ell++) {
                if (ell >= d->nlp || q >= d->lp_n[ell]) goto out;
                const mopen *o = &d->lp[ell][q];
                const qref *rr = &qp->refs[ell];
                const lpay *pay = &qp->pay[ell];
                size_t a = pick_ar(sz[ell], sch[ell]), dep = mdepth(sz[ell], a);
                if (!verify_open(a, dep, ell, d->roots[ell], o)) goto out;
                if (o->index != rr->i) goto out;
                {
                    uint64_t lv[18];
                    uint8_t lh[32];
                    memcpy(lv, pay->f.c, 48);
                    memcpy(lv + 6, pay->s.c, 48);
                    memcpy(lv + 12, pay->q.c, 48);
                    leaf_hash_v(a, ell, o->index, lv, 18, lh);
                    if (memcmp(lh, o->leaf, 32) != 0) goto out;
                }
                const uint64_t xb = gl_pow(omg[ell], rr->i);
                /* DEEP identity q (x - z) = f - f(z) */
                fp6 lhs = fp6_mul(pay->q, fp6_sub(fp6_from_base(xb), z_ext));
                if (!fp6_eq(lhs, fp6_sub(pay->f, d->fz[ell]))) goto out;
                if (ell == L - 1 && use_cc) {
                    const size_t m = sch[ell];
                    if (qp->final_index >= d->nct) goto out;
                    const fp6 *t = d->ct[qp->final_index];
                    fp6 h = fp6_zero();
                    uint64_t xp = 1;
                    for (size_t k = 0; k < m; k++) {
                        h = fp6_add(h, fp6_scale(t[k], xp));
                        xp = gl_mul(xp, xb);
                    }
                    if (!fp6_eq(pay->f, h)) goto out;
                    fp6 fold = ext_pows_dot(t, m, alpha[L - 1]);
                    fp6 xf = fp6_from_base(gl_pow(om_final, qp->final_index));
                    if (!fp6_eq(fold, eval_final(d->fin, d->nfin, xf))) goto out;
                } else {
                    fp6 nxt;
                    if (ell + 1 < L) {
                        nxt = qp->pay[ell + 1].f;
                    } else {
                        fp6 xf = fp6_from_base(gl_pow(om_final, qp->final_index));
                        nxt = eval_final(d->fin, d->nfin, xf);
                    }
                    if (!fp6_eq(pay->s, nxt)) goto out;
                }
            }
        }
        ok = 1;
    }
out:
    for (size_t l = 0; l <= L; l++) free(interp[l]);
    free(interp); free(alpha); free(omg); free(sz);
    return ok;
}

int ref_deep_fri_verify(const ref_fri_params *p, const uint8_t *proof, size_t len) {
    arena ar = {0};
    dproof d;
    int rc = parse_proof(proof, len, &ar, &d);
    if (rc == 0) rc = verify(p, &d);
    ar_free(&ar);
    return rc;
}
