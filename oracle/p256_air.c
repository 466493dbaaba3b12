/* p256_air.c -- CPU restatement of the double-chain ECDSA-P256 AIR (two
 * scalar-multiplication chains, one double-and-conditional-add step of each
 * per trace row), its streaming merge and the per-record ECDSA prover of the
 * mixed-algorithm zone example (test infrastructure: only tests/ load it).
 *
 *   field element, limbs      crates/deep_ali/src/p256_field.rs:202-241 (10 x 26-bit limbs, P_LIMBS_TIGHT)
 *   gadgets (layout, fill,    crates/deep_ali/src/p256_field_air.rs:125-227 (range check),
 *   constraints, this order)    :309-455 (add), :545-698 (sub), :810-1123 (mul),
 *                               :1147-1171 (square), :1209-1276 (select), :1345-1531 (freeze)
 *   group double / add        crates/deep_ali/src/p256_group_air.rs:111-215 (cell allocation),
 *                               :302-447 (double layout, RCB-2016 alg. 6), :471-635 (fill),
 *                               :645-680 (constraints), :705-870 (add layout, alg. 4),
 *                               :884-1046 (fill), :1048-1068 (constraints)
 *   scalar-mul step           crates/deep_ali/src/p256_scalar_mul_air.rs:74-255
 *   double-chain rows         crates/deep_ali/src/p256_ecdsa_double_multirow_air.rs:62-307
 *   merge                     crates/deep_ali/src/lib.rs:1421-1540
 *   per-record prover         crates/swarm-dns/examples/mixed_algorithm_zone_rollup.rs:193-322
 *
 * Big integers are little-endian arrays of 32-bit words; quotients come from
 * plain shift-and-subtract long division (simple and slow, which suits a
 * checker). */
#include <stdlib.h>
#include <string.h>

#include "refcpu.h"
#include "refcpu_int.h"

enum { NL = 10, LB = 26, EC = 270, MCB = 36, MCP = 18, MSP = 19 };
enum { GM, GA, GS, GF, GSEL };
static const int64_t MC_OFF = (int64_t)1 << 35;
static const int64_t PL[NL] = {0x3FFFFFF, 0x3FFFFFF, 0x3FFFFFF, 0x003FFFF, 0, 0, 0, 0x0000400, 0x3FF0000, 0x03FFFFF};

/* ---- 256-bit integers (8 LE words) ---- */
static const uint32_t P_W[8] = {0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0, 0, 0, 1, 0xFFFFFFFF};
static const uint32_t N_W[8] = {0xFC632551, 0xF3B9CAC2, 0xA7179E84, 0xBCE6FAAD, 0xFFFFFFFF, 0xFFFFFFFF, 0, 0xFFFFFFFF};
static const uint8_t B_BE[32] = {0x5a, 0xc6, 0x35, 0xd8, 0xaa, 0x3a, 0x93, 0xe7, 0xb3, 0xeb, 0xbd,
                                 0x55, 0x76, 0x98, 0x86, 0xbc, 0x65, 0x1d, 0x06, 0xb0, 0xcc, 0x53,
                                 0xb0, 0xf6, 0x3b, 0xce, 0x3c, 0x3e, 0x27, 0xd2, 0x60, 0x4b};
static const uint8_t GX_BE[32] = {0x6b, 0x17, 0xd1, 0xf2, 0xe1, 0x2c, 0x42, 0x47, 0xf8, 0xbc, 0xe6,
                                  0xe5, 0x63, 0xa4, 0x40, 0xf2, 0x77, 0x03, 0x7d, 0x81, 0x2d, 0xeb,
                                  0x33, 0xa0, 0xf4, 0xa1, 0x39, 0x45, 0xd8, 0x98, 0xc2, 0x96};
static const uint8_t GY_BE[32] = {0x4f, 0xe3, 0x42, 0xe2, 0xfe, 0x1a, 0x7f, 0x9b, 0x8e, 0xe7, 0xeb,
                                  0x4a, 0x7c, 0x0f, 0x9e, 0x16, 0x2b, 0xce, 0x33, 0x57, 0x6b, 0x31,
                                  0x5e, 0xce, 0xcb, 0xb6, 0x40, 0x68, 0x37, 0xbf, 0x51, 0xf5};

static void w_from_be(const uint8_t be[32], uint32_t w[8]) {
    for (int i = 0; i < 8; i++)
        w[i] = (uint32_t)be[31 - 4 * i] | (uint32_t)be[30 - 4 * i] << 8 | (uint32_t)be[29 - 4 * i] << 16 |
               (uint32_t)be[28 - 4 * i] << 24;
}
static void w_to_be(const uint32_t w[8], uint8_t be[32]) {
    for (int i = 0; i < 8; i++)
        for (int b = 0; b < 4; b++) be[31 - 4 * i - b] = (uint8_t)(w[i] >> (8 * b));
}
static int w_cmp(const uint32_t *a, const uint32_t *b, size_t nw) {
    for (size_t i = nw; i-- > 0;)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}
static int w_is_zero(const uint32_t *a, size_t nw) {
    for (size_t i = 0; i < nw; i++)
        if (a[i]) return 0;
    return 1;
}
/* x (16 words) = q * m + r, r < m (8 words): bit-serial long division */
static void w_divmod(const uint32_t x[16], const uint32_t m[8], uint32_t q[16], uint32_t r[8]) {
    uint32_t rr[9] = {0}, mm[9];
    memcpy(mm, m, 32);
    mm[8] = 0;
    memset(q, 0, 64);
    for (int bit = 511; bit >= 0; bit--) {
        for (int i = 8; i > 0; i--) rr[i] = (rr[i] << 1) | (rr[i - 1] >> 31);
        rr[0] = (rr[0] << 1) | ((x[bit / 32] >> (bit % 32)) & 1);
        if (w_cmp(rr, mm, 9) >= 0) {
            uint64_t borrow = 0;
            for (int i = 0; i < 9; i++) {
                const uint64_t t = (uint64_t)rr[i] - mm[i] - borrow;
                rr[i] = (uint32_t)t;
                borrow = (t >> 63) & 1;
            }
            q[bit / 32] |= 1u << (bit % 32);
        }
    }
    memcpy(r, rr, 32);
}
static void w_mul(const uint32_t a[8], const uint32_t b[8], uint32_t prod[16]) {
    memset(prod, 0, 64);
    for (int i = 0; i < 8; i++) {
        uint64_t carry = 0;
        for (int j = 0; j < 8; j++) {
            const uint64_t t = (uint64_t)a[i] * b[j] + prod[i + j] + carry;
            prod[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        prod[i + 8] = (uint32_t)carry;
    }
}
/* r = a * b mod m */
static void w_mulmod(const uint32_t a[8], const uint32_t b[8], const uint32_t m[8], uint32_t r[8]) {
    uint32_t prod[16], q[16];
    w_mul(a, b, prod);
    w_divmod(prod, m, q, r);
}
/* r = a^(m-2) mod m (m prime) */
static void w_invmod(const uint32_t a[8], const uint32_t m[8], uint32_t r[8]) {
    uint32_t e[8], acc[8] = {1, 0, 0, 0, 0, 0, 0, 0};
    memcpy(e, m, 32);
    e[0] -= 2; /* m is odd and > 2 */
    for (int bit = 255; bit >= 0; bit--) {
        w_mulmod(acc, acc, m, acc);
        if ((e[bit / 32] >> (bit % 32)) & 1) w_mulmod(acc, a, m, acc);
    }
    memcpy(r, acc, 32);
}

/* 26-bit limbs of an 8-word value (from_be_bytes, p256_field.rs:264-272) */
static void w_to_limbs(const uint32_t w[8], int64_t l[NL]) {
    for (int i = 0; i < NL; i++) {
        uint64_t v = 0;
        for (int b = 0; b < LB; b++) {
            const int bit = LB * i + b;
            if (bit < 256) v |= (uint64_t)((w[bit / 32] >> (bit % 32)) & 1) << b;
        }
        l[i] = (int64_t)v;
    }
}
/* limbs -> integer (< 2^260: words 0..8) */
static void limbs_to_w(const int64_t l[NL], uint32_t w[9]) {
    memset(w, 0, 36);
    for (int i = 0; i < NL; i++)
        for (int b = 0; b < LB; b++)
            if ((l[i] >> b) & 1) {
                const int bit = LB * i + b;
                w[bit / 32] |= 1u << (bit % 32);
            }
}

/* ---- gadget layouts ---- */
typedef struct {
    int kind;
    size_t a, b, c, cb, q, qb, cy, car, pos, neg, dl, db, sel;
} gad;

/* RCB-2016 programs as written in p256_group_air.rs: op (M/A/S), destination
 * register, two source registers. Every A and S is followed by a freeze whose
 * output the destination names. */
typedef struct { int op, d, a, b; } gop;
/* double (alg. 6, p256_group_air.rs:364-438): X=0 Y=1 Z=2 b=3 t0..t3=4..7 x3=8 y3=9 z3=10 */
static const gop DBL[34] = {
    {GM, 4, 0, 0},  {GM, 5, 1, 1},  {GM, 6, 2, 2},  {GM, 7, 0, 1},  {GA, 7, 7, 7},  {GM, 10, 0, 2}, {GA, 10, 10, 10},
    {GM, 9, 3, 6},  {GS, 9, 9, 10}, {GA, 8, 9, 9},  {GA, 9, 8, 9},  {GS, 8, 5, 9},  {GA, 9, 5, 9},  {GM, 9, 8, 9},
    {GM, 8, 8, 7},  {GA, 7, 6, 6},  {GA, 6, 6, 7},  {GM, 10, 3, 10}, {GS, 10, 10, 6}, {GS, 10, 10, 4}, {GA, 7, 10, 10},
    {GA, 10, 10, 7}, {GA, 7, 4, 4}, {GA, 4, 7, 4},  {GS, 4, 4, 6},  {GM, 4, 4, 10}, {GA, 9, 9, 4},  {GM, 4, 1, 2},
    {GA, 4, 4, 4},  {GM, 10, 4, 10}, {GS, 8, 8, 10}, {GM, 10, 4, 5}, {GA, 10, 10, 10}, {GA, 10, 10, 10}};
/* add (alg. 4, p256_group_air.rs:745-850): X1 Y1 Z1 = 0..2, X2 Y2 Z2 = 3..5,
 * b = 6, t0..t4 = 7..11, x3 = 12, y3 = 13, z3 = 14 */
static const gop ADD[43] = {
    {GM, 7, 0, 3},   {GM, 8, 1, 4},    {GM, 9, 2, 5},    {GA, 10, 0, 1},   {GA, 11, 3, 4},  {GM, 10, 10, 11},
    {GA, 11, 7, 8},  {GS, 10, 10, 11}, {GA, 11, 1, 2},   {GA, 12, 4, 5},   {GM, 11, 11, 12}, {GA, 12, 8, 9},
    {GS, 11, 11, 12}, {GA, 12, 0, 2},  {GA, 13, 3, 5},   {GM, 12, 12, 13}, {GA, 13, 7, 9},  {GS, 13, 12, 13},
    {GM, 14, 6, 9},  {GS, 12, 13, 14}, {GA, 14, 12, 12}, {GA, 12, 12, 14}, {GS, 14, 8, 12}, {GA, 12, 8, 12},
    {GM, 13, 6, 13}, {GA, 8, 9, 9},    {GA, 9, 8, 9},    {GS, 13, 13, 9},  {GS, 13, 13, 7}, {GA, 8, 13, 13},
    {GA, 13, 8, 13}, {GA, 8, 7, 7},    {GA, 7, 8, 7},    {GS, 7, 7, 9},    {GM, 8, 11, 13}, {GM, 9, 7, 13},
    {GM, 13, 12, 14}, {GA, 13, 13, 9}, {GM, 12, 10, 12}, {GS, 12, 12, 8},  {GM, 14, 11, 14}, {GM, 8, 10, 7},
    {GA, 14, 14, 8}};

/* One group gadget: its b block, its sub-gadgets in allocation order (an A
 * or S is followed by its freeze) and the result registers. */
typedef struct {
    const gop *prog;
    int n_ops, n_in;
    size_t b_base;
    gad g[128];      /* allocation order */
    int op_gad[64];  /* first sub-gadget of each op */
    size_t reg[16];  /* column bases of the registers after the layout pass */
    int res[3];      /* result registers (x3, y3, z3) */
} ggad;

static gad alloc_mul(size_t *cur, size_t a, size_t b) {
    gad g;
    memset(&g, 0, sizeof g);
    g.kind = GM;
    g.a = a;
    g.b = b;
    g.c = *cur;
    g.cb = g.c + NL;
    g.q = g.cb + NL * LB;
    g.qb = g.q + NL;
    g.cy = g.qb + NL * LB;
    *cur = g.cy + MCP * MCB;
    return g;
}
static gad alloc_add(size_t *cur, size_t a, size_t b) {
    gad g;
    memset(&g, 0, sizeof g);
    g.kind = GA;
    g.a = a;
    g.b = b;
    g.c = *cur;
    g.cb = g.c + NL;
    g.car = g.cb + NL * LB;
    *cur = g.car + NL;
    return g;
}
static gad alloc_sub(size_t *cur, size_t a, size_t b) {
    gad g;
    memset(&g, 0, sizeof g);
    g.kind = GS;
    g.a = a;
    g.b = b;
    g.c = *cur;
    g.cb = g.c + NL;
    g.pos = g.cb + NL * LB;
    g.neg = g.pos + NL;
    *cur = g.neg + NL;
    return g;
}
static gad alloc_freeze(size_t *cur, size_t a) {
    gad g;
    memset(&g, 0, sizeof g);
    g.kind = GF;
    g.a = a;
    g.dl = *cur;
    g.db = g.dl + NL;
    g.c = g.db + NL * LB;
    g.cb = g.c + NL;
    g.pos = g.cb + NL * LB;
    g.neg = g.pos + NL;
    *cur = g.neg + NL;
    return g;
}

static void build_group(ggad *G, const gop *prog, int n_ops, const size_t *inputs, int n_in, size_t *cur) {
    G->prog = prog;
    G->n_ops = n_ops;
    G->n_in = n_in;
    for (int i = 0; i < n_in; i++) G->reg[i] = inputs[i];
    G->b_base = *cur;
    G->reg[n_in] = G->b_base; /* b is the register after the inputs */
    *cur += EC;
    int ng = 0;
    for (int k = 0; k < n_ops; k++) {
        const gop o = prog[k];
        G->op_gad[k] = ng;
        if (o.op == GM) {
            G->g[ng] = alloc_mul(cur, G->reg[o.a], G->reg[o.b]);
            G->reg[o.d] = G->g[ng++].c;
        } else {
            G->g[ng] = o.op == GA ? alloc_add(cur, G->reg[o.a], G->reg[o.b]) : alloc_sub(cur, G->reg[o.a], G->reg[o.b]);
            const size_t raw = G->g[ng++].c;
            G->g[ng] = alloc_freeze(cur, raw);
            G->reg[o.d] = G->g[ng++].c;
        }
    }
}

/* sub-gadgets in constraint order (eval_group_*_gadget): muls, adds, subs,
 * freezes after adds, freezes after subs */
static int group_eval_order(const ggad *G, const gad **out) {
    int n = 0;
    const int ng = G->op_gad[G->n_ops - 1] + (G->prog[G->n_ops - 1].op == GM ? 1 : 2);
    for (int kind = 0; kind < 5; kind++)
        for (int i = 0; i < ng; i++) {
            const gad *g = &G->g[i];
            int sel;
            if (kind < 3) sel = g->kind == (kind == 0 ? GM : kind == 1 ? GA : GS);
            else sel = g->kind == GF && G->g[i - 1].kind == (kind == 3 ? GA : GS);
            if (sel) out[n++] = g;
        }
    return n;
}

typedef struct {
    size_t acc[3], base[3], bit;
    ggad dbl, add;
    gad sel[3];
    const gad *order[260];
    int n_order;
} step_layout;

typedef struct {
    step_layout ch[2];
    size_t rproj[6];
    size_t width;
} p256_layout;

/* build_scalar_mul_step_layout (p256_scalar_mul_air.rs:74-155) after the
 * chain's acc / base / bit cells (p256_ecdsa_double_multirow_air.rs:62-100) */
static void build_step(step_layout *S, size_t *cur) {
    for (int i = 0; i < 3; i++) {
        S->acc[i] = *cur;
        *cur += NL;
    }
    for (int i = 0; i < 3; i++) {
        S->base[i] = *cur;
        *cur += NL;
    }
    S->bit = (*cur)++;
    build_group(&S->dbl, DBL, 34, S->acc, 3, cur);
    const size_t two[3] = {S->dbl.reg[8], S->dbl.reg[9], S->dbl.reg[10]};
    const size_t in[6] = {two[0], two[1], two[2], S->base[0], S->base[1], S->base[2]};
    build_group(&S->add, ADD, 43, in, 6, cur);
    const size_t added[3] = {S->add.reg[12], S->add.reg[13], S->add.reg[14]};
    for (int i = 0; i < 3; i++) {
        gad g;
        memset(&g, 0, sizeof g);
        g.kind = GSEL;
        g.a = added[i];
        g.b = two[i];
        g.c = *cur;
        g.cb = g.c + NL;
        g.sel = S->bit;
        *cur = g.cb + NL * LB;
        S->sel[i] = g;
    }
    int n = group_eval_order(&S->dbl, S->order);
    n += group_eval_order(&S->add, S->order + n);
    for (int i = 0; i < 3; i++) S->order[n++] = &S->sel[i];
    S->n_order = n;
}

static p256_layout *g_layout;

static const p256_layout *layout(void) {
    if (!g_layout) {
        p256_layout *L = (p256_layout *)calloc(1, sizeof *L);
        size_t cur = 0;
        build_step(&L->ch[0], &cur);
        build_step(&L->ch[1], &cur);
        for (int i = 0; i < 6; i++) {
            L->rproj[i] = cur;
            cur += NL;
        }
        L->width = cur;
        g_layout = L;
    }
    return g_layout;
}

static size_t gad_constraints(int kind) {
    switch (kind) {
    case GM: return 2 * EC + MSP + MCP * MCB;                  /* 1207 */
    case GA: return EC + 2 * NL;                               /* 290 */
    case GS: return EC + NL + 2 * NL + NL + 1;                 /* 311 */
    case GF: return 2 * EC + NL + 2 * NL + NL + 1 + NL;        /* 591 */
    default: return EC + NL + 1;                               /* 281 */
    }
}

size_t ref_p256_width(void) { return layout()->width; }
size_t ref_p256_n_constraints(void) {
    const p256_layout *L = layout();
    size_t k = 0;
    for (int c = 0; c < 2; c++)
        for (int i = 0; i < L->ch[c].n_order; i++) k += gad_constraints(L->ch[c].order[i]->kind);
    return k + 6 * NL + 12 * NL;
}
/* the column bases the tests name: [0] bit A, [1] bit B, [2..4] select
 * outputs of A, [5..7] of B, [8..13] r_proj, [14..16] acc of A, [17..19] acc
 * of B, [20..22] base of A, [23..25] base of B */
void ref_p256_columns(size_t out[26]) {
    const p256_layout *L = layout();
    out[0] = L->ch[0].bit;
    out[1] = L->ch[1].bit;
    for (int c = 0; c < 2; c++)
        for (int i = 0; i < 3; i++) {
            out[2 + 3 * c + i] = L->ch[c].sel[i].c;
            out[14 + 3 * c + i] = L->ch[c].acc[i];
            out[20 + 3 * c + i] = L->ch[c].base[i];
        }
    for (int i = 0; i < 6; i++) out[8 + i] = L->rproj[i];
}
/* column bases of every sub-gadget of chain `ch` in constraint order: per
 * gadget kind, c, cb, q/dl, qb/db, cy/car/pos, neg; returns the count */
int ref_p256_gadgets(int ch, int *kinds, size_t *cols /* 6 per gadget */) {
    const step_layout *S = &layout()->ch[ch];
    for (int i = 0; i < S->n_order; i++) {
        const gad *g = S->order[i];
        kinds[i] = g->kind;
        size_t *o = cols + 6 * i;
        o[0] = g->c;
        o[1] = g->cb;
        o[2] = g->kind == GM ? g->q : g->dl;
        o[3] = g->kind == GM ? g->qb : g->db;
        o[4] = g->kind == GM ? g->cy : g->kind == GA ? g->car : g->pos;
        o[5] = g->neg;
    }
    return S->n_order;
}

/* ---- fill ---- */
typedef struct {
    uint64_t *tr;
    size_t ht, row;
} tw;
#define CELL(t, col) ((t)->tr[(size_t)(col) * (t)->ht + (t)->row])

static void put_elem(tw *t, size_t lb, size_t bb, const int64_t l[NL]) {
    for (int i = 0; i < NL; i++) {
        CELL(t, lb + i) = (uint64_t)l[i];
        for (int b = 0; b < LB; b++) CELL(t, bb + LB * i + b) = (uint64_t)((l[i] >> b) & 1);
    }
}
static void put_limbs(tw *t, size_t lb, const int64_t l[NL]) {
    for (int i = 0; i < NL; i++) CELL(t, lb + i) = (uint64_t)l[i];
}
static void put_signed(tw *t, size_t pos, size_t neg, const int64_t cy[NL]) {
    for (int k = 0; k < NL; k++) {
        CELL(t, pos + k) = cy[k] == 1;
        CELL(t, neg + k) = cy[k] == -1;
    }
}

/* fill_mul_gadget (p256_field_air.rs:953-1042): c = a b mod p, q = (a b - c) / p */
static int fill_mul(tw *t, const gad *g, const int64_t a[NL], const int64_t b[NL], int64_t c[NL]) {
    uint32_t aw[9], bw[9], prod[16], q[16], r[8];
    limbs_to_w(a, aw);
    limbs_to_w(b, bw);
    if (aw[8] || bw[8]) return -20; /* operands above 2^256 never occur */
    w_mul(aw, bw, prod);
    w_divmod(prod, P_W, q, r);
    if (!w_is_zero(q + 8, 8)) return -21;
    int64_t ql[NL];
    w_to_limbs(r, c);
    w_to_limbs(q, ql);
    int64_t cy[MCP], prev = 0;
    for (int k = 0; k < MSP; k++) {
        int64_t pab = 0, pqp = 0;
        const int lo = k > NL - 1 ? k - (NL - 1) : 0, hi = k < NL - 1 ? k : NL - 1;
        for (int i = lo; i <= hi; i++) {
            pab += a[i] * b[k - i];
            pqp += ql[i] * PL[k - i];
        }
        const int64_t lhs = pab - pqp - (k < NL ? c[k] : 0) + prev;
        if (k < MSP - 1) {
            if (lhs % ((int64_t)1 << LB)) return -22;
            cy[k] = lhs / ((int64_t)1 << LB);
            if (cy[k] >= MC_OFF || cy[k] < -MC_OFF) return -23;
            prev = cy[k];
        } else if (lhs != 0) {
            return -24;
        }
    }
    put_elem(t, g->c, g->cb, c);
    put_elem(t, g->q, g->qb, ql);
    for (int k = 0; k < MCP; k++) {
        const uint64_t biased = (uint64_t)(cy[k] + MC_OFF);
        for (int bb = 0; bb < MCB; bb++) CELL(t, g->cy + MCB * k + bb) = (biased >> bb) & 1;
    }
    return 0;
}
/* fill_add_gadget (p256_field_air.rs:344-402): c = a + b limb-wise, carries 0/1 */
static void fill_add(tw *t, const gad *g, const int64_t a[NL], const int64_t b[NL], int64_t c[NL]) {
    int64_t prev = 0;
    for (int k = 0; k < NL; k++) {
        const int64_t s = a[k] + b[k] + prev;
        c[k] = s & (((int64_t)1 << LB) - 1);
        prev = s >> LB;
        CELL(t, g->car + k) = (uint64_t)prev;
    }
    put_elem(t, g->c, g->cb, c);
}
static int64_t floor_div(int64_t x, int64_t *rem) { /* by 2^26, Euclidean */
    const int64_t R = (int64_t)1 << LB;
    int64_t q = x / R, r = x % R;
    if (r < 0) {
        r += R;
        q -= 1;
    }
    *rem = r;
    return q;
}
/* fill_sub_gadget (p256_field_air.rs:570-634): c = a + p - b, carries -1/0/1 */
static void fill_sub(tw *t, const gad *g, const int64_t a[NL], const int64_t b[NL], int64_t c[NL]) {
    int64_t cy[NL], prev = 0;
    for (int k = 0; k < NL; k++) {
        cy[k] = floor_div(a[k] + PL[k] - b[k] + prev, &c[k]);
        prev = cy[k];
    }
    put_elem(t, g->c, g->cb, c);
    put_signed(t, g->pos, g->neg, cy);
}
/* fill_freeze_gadget (p256_field_air.rs:1374-1447): diff = a - p; c = a if
 * the last carry is -1 (a < p), else diff */
static void fill_freeze(tw *t, const gad *g, const int64_t a[NL], int64_t c[NL]) {
    int64_t cy[NL], d[NL], prev = 0;
    for (int k = 0; k < NL; k++) {
        cy[k] = floor_div(a[k] - PL[k] + prev, &d[k]);
        prev = cy[k];
    }
    for (int k = 0; k < NL; k++) c[k] = cy[NL - 1] == -1 ? a[k] : d[k];
    put_elem(t, g->dl, g->db, d);
    put_elem(t, g->c, g->cb, c);
    put_signed(t, g->pos, g->neg, cy);
}

/* fill_group_double_gadget / fill_group_add_gadget: run the program */
static int fill_group(tw *t, const ggad *G, const int64_t (*in)[NL]) {
    int64_t reg[16][NL], bl[NL];
    uint32_t bw[8];
    w_from_be(B_BE, bw);
    w_to_limbs(bw, bl);
    put_elem(t, G->b_base, G->b_base + NL, bl);
    for (int i = 0; i < G->n_in; i++) memcpy(reg[i], in[i], sizeof reg[i]);
    memcpy(reg[G->n_in], bl, sizeof bl);
    for (int k = 0; k < G->n_ops; k++) {
        const gop o = G->prog[k];
        const gad *g = &G->g[G->op_gad[k]];
        int64_t c[NL];
        if (o.op == GM) {
            const int rc = fill_mul(t, g, reg[o.a], reg[o.b], c);
            if (rc) return rc;
        } else {
            int64_t raw[NL];
            if (o.op == GA) fill_add(t, g, reg[o.a], reg[o.b], raw);
            else fill_sub(t, g, reg[o.a], reg[o.b], raw);
            fill_freeze(t, g + 1, raw, c);
        }
        memcpy(reg[o.d], c, sizeof c);
    }
    return 0;
}

/* fill_scalar_mul_step_gadget (p256_scalar_mul_air.rs:174-220) on one row:
 * acc' = bit ? 2 acc + base : 2 acc */
static int fill_step(tw *t, const step_layout *S, const int64_t acc[3][NL], const int64_t base[3][NL], int bit,
                     int64_t out[3][NL]) {
    for (int i = 0; i < 3; i++) {
        put_limbs(t, S->acc[i], acc[i]);
        put_limbs(t, S->base[i], base[i]);
    }
    CELL(t, S->bit) = (uint64_t)bit;
    int rc = fill_group(t, &S->dbl, acc);
    if (rc) return rc;
    int64_t two[3][NL], in[6][NL], added[3][NL];
    for (int i = 0; i < 3; i++) {
        for (int l = 0; l < NL; l++) {
            two[i][l] = (int64_t)CELL(t, S->dbl.reg[8 + i] + l);
        }
        memcpy(in[i], two[i], sizeof two[i]);
        memcpy(in[3 + i], base[i], sizeof base[i]);
    }
    rc = fill_group(t, &S->add, in);
    if (rc) return rc;
    for (int i = 0; i < 3; i++) {
        for (int l = 0; l < NL; l++) added[i][l] = (int64_t)CELL(t, S->add.reg[12 + i] + l);
        memcpy(out[i], bit ? added[i] : two[i], sizeof out[i]);
        put_elem(t, S->sel[i].c, S->sel[i].cb, out[i]);
    }
    return 0;
}

/* fill_ecdsa_double_multirow (p256_ecdsa_double_multirow_air.rs:134-222).
 * pts: 12 x 32-byte big-endian field elements in the order a_initial x,y,z,
 * a_base x,y,z, b_initial x,y,z, b_base x,y,z; rproj: 6 x 32 bytes (r_a x,y,z,
 * r_b x,y,z); bits: k_a then k_b bytes of 0/1. Values must be below 2^256
 * (their limbs are taken as they are). Trace: width x ht column-major. */
int ref_fill_p256(const uint8_t *pts, const uint8_t *rproj, const uint8_t *a_bits, size_t k_a, const uint8_t *b_bits,
                  size_t k_b, size_t ht, uint64_t *tr) {
    const p256_layout *L = layout();
    if (!ht || (ht & (ht - 1)) || k_a > ht || k_b > ht) return -10;
    int64_t acc[2][3][NL], base[2][3][NL], rp[6][NL];
    for (int c = 0; c < 2; c++)
        for (int i = 0; i < 3; i++) {
            uint32_t w[8];
            w_from_be(pts + 32 * (6 * c + i), w);
            w_to_limbs(w, acc[c][i]);
            w_from_be(pts + 32 * (6 * c + 3 + i), w);
            w_to_limbs(w, base[c][i]);
        }
    for (int i = 0; i < 6; i++) {
        uint32_t w[8];
        w_from_be(rproj + 32 * i, w);
        w_to_limbs(w, rp[i]);
    }
    for (size_t r = 0; r < ht; r++) {
        tw t = {tr, ht, r};
        for (int c = 0; c < 2; c++) {
            const size_t k = c ? k_b : k_a;
            const uint8_t *bits = c ? b_bits : a_bits;
            int64_t out[3][NL];
            const int rc = fill_step(&t, &L->ch[c], (const int64_t(*)[NL])acc[c], (const int64_t(*)[NL])base[c],
                                     r < k ? bits[r] != 0 : 0, out);
            if (rc) return rc;
            memcpy(acc[c], out, sizeof out);
        }
        for (int i = 0; i < 6; i++) put_limbs(&t, L->rproj[i], rp[i]);
    }
    return 0;
}

/* ---- constraints ---- */
static uint64_t pw2(int b) { return (uint64_t)1 << b; }

/* range check of one element block (booleanity of the 260 bits, then the 10
 * limb packings), as every gadget writes it */
static size_t ev_range(const uint64_t *cur, size_t lb, size_t bb, uint64_t *out) {
    size_t o = 0;
    for (int i = 0; i < NL; i++)
        for (int b = 0; b < LB; b++) {
            const uint64_t x = cur[bb + LB * i + b];
            out[o++] = gl_mul(x, gl_sub(1, x));
        }
    for (int i = 0; i < NL; i++) {
        uint64_t s = 0;
        for (int b = 0; b < LB; b++) s = gl_add(s, gl_mul(pw2(b), cur[bb + LB * i + b]));
        out[o++] = gl_sub(cur[lb + i], s);
    }
    return o;
}
/* booleanity, exclusion and the last-limb rule of signed carries */
static size_t ev_signed(const uint64_t *cur, const gad *g, uint64_t *out) {
    size_t o = 0;
    for (int k = 0; k < NL; k++) out[o++] = gl_mul(cur[g->pos + k], gl_sub(1, cur[g->pos + k]));
    for (int k = 0; k < NL; k++) out[o++] = gl_mul(cur[g->neg + k], gl_sub(1, cur[g->neg + k]));
    for (int k = 0; k < NL; k++) out[o++] = gl_mul(cur[g->pos + k], cur[g->neg + k]);
    return o;
}

static size_t ev_gadget(const uint64_t *cur, const gad *g, uint64_t *out) {
    const uint64_t radix = pw2(LB);
    size_t o = 0;
    switch (g->kind) {
    case GM: { /* eval_mul_gadget (p256_field_air.rs:1044-1123) */
        o += ev_range(cur, g->c, g->cb, out + o);
        o += ev_range(cur, g->q, g->qb, out + o);
        uint64_t sc[MCP];
        for (int k = 0; k < MCP; k++) {
            uint64_t v = 0;
            for (int b = 0; b < MCB; b++) v = gl_add(v, gl_mul(pw2(b), cur[g->cy + MCB * k + b]));
            sc[k] = gl_sub(v, (uint64_t)MC_OFF);
        }
        for (int k = 0; k < MSP; k++) {
            uint64_t pab = 0, pqp = 0;
            const int lo = k > NL - 1 ? k - (NL - 1) : 0, hi = k < NL - 1 ? k : NL - 1;
            for (int i = lo; i <= hi; i++) {
                pab = gl_add(pab, gl_mul(cur[g->a + i], cur[g->b + k - i]));
                pqp = gl_add(pqp, gl_mul(cur[g->q + i], (uint64_t)PL[k - i]));
            }
            const uint64_t ck = k < NL ? cur[g->c + k] : 0;
            const uint64_t cin = k ? sc[k - 1] : 0, cout = k < MCP ? sc[k] : 0;
            out[o++] = gl_sub(gl_add(gl_sub(gl_sub(pab, pqp), ck), cin), gl_mul(radix, cout));
        }
        for (int k = 0; k < MCP; k++)
            for (int b = 0; b < MCB; b++) {
                const uint64_t x = cur[g->cy + MCB * k + b];
                out[o++] = gl_mul(x, gl_sub(1, x));
            }
        break;
    }
    case GA: /* eval_add_gadget (p256_field_air.rs:404-455) */
        o += ev_range(cur, g->c, g->cb, out + o);
        for (int k = 0; k < NL; k++) {
            const uint64_t cin = k ? cur[g->car + k - 1] : 0;
            out[o++] = gl_sub(gl_sub(gl_add(gl_add(cur[g->a + k], cur[g->b + k]), cin), cur[g->c + k]),
                              gl_mul(radix, cur[g->car + k]));
        }
        for (int k = 0; k < NL - 1; k++) out[o++] = gl_mul(cur[g->car + k], gl_sub(1, cur[g->car + k]));
        out[o++] = cur[g->car + NL - 1];
        break;
    case GS: /* eval_sub_gadget (p256_field_air.rs:636-698) */
        o += ev_range(cur, g->c, g->cb, out + o);
        for (int k = 0; k < NL; k++) {
            const uint64_t net = gl_sub(cur[g->pos + k], cur[g->neg + k]);
            const uint64_t cin = k ? gl_sub(cur[g->pos + k - 1], cur[g->neg + k - 1]) : 0;
            out[o++] = gl_sub(gl_sub(gl_add(gl_sub(gl_add(cur[g->a + k], (uint64_t)PL[k]), cur[g->b + k]), cin),
                                     cur[g->c + k]),
                              gl_mul(radix, net));
        }
        o += ev_signed(cur, g, out + o);
        out[o++] = gl_add(cur[g->pos + NL - 1], cur[g->neg + NL - 1]);
        break;
    case GF: { /* eval_freeze_gadget (p256_field_air.rs:1449-1531) */
        o += ev_range(cur, g->dl, g->db, out + o);
        o += ev_range(cur, g->c, g->cb, out + o);
        for (int k = 0; k < NL; k++) {
            const uint64_t net = gl_sub(cur[g->pos + k], cur[g->neg + k]);
            const uint64_t cin = k ? gl_sub(cur[g->pos + k - 1], cur[g->neg + k - 1]) : 0;
            out[o++] = gl_sub(gl_add(gl_sub(gl_sub(cur[g->a + k], (uint64_t)PL[k]), cur[g->dl + k]), cin),
                              gl_mul(radix, net));
        }
        o += ev_signed(cur, g, out + o);
        out[o++] = cur[g->pos + NL - 1];
        const uint64_t n9 = cur[g->neg + NL - 1];
        for (int k = 0; k < NL; k++) {
            const uint64_t d = cur[g->dl + k];
            out[o++] = gl_sub(gl_sub(cur[g->c + k], d), gl_mul(n9, gl_sub(cur[g->a + k], d)));
        }
        break;
    }
    default: { /* eval_select_gadget (p256_field_air.rs:1243-1276) */
        o += ev_range(cur, g->c, g->cb, out + o);
        const uint64_t s = cur[g->sel];
        for (int k = 0; k < NL; k++) {
            const uint64_t bk = cur[g->b + k];
            out[o++] = gl_sub(gl_sub(cur[g->c + k], bk), gl_mul(s, gl_sub(cur[g->a + k], bk)));
        }
        out[o++] = gl_mul(s, gl_sub(1, s));
    }
    }
    return o;
}

/* eval_ecdsa_double_multirow_per_row (p256_ecdsa_double_multirow_air.rs:
 * 225-307): both steps, the boundary at the last row, the transitions */
void ref_p256_constraints(const uint64_t *cur, const uint64_t *nxt, size_t trace_row, size_t n_trace, uint64_t *out) {
    const p256_layout *L = layout();
    size_t o = 0;
    for (int c = 0; c < 2; c++)
        for (int i = 0; i < L->ch[c].n_order; i++) o += ev_gadget(cur, L->ch[c].order[i], out + o);
    const int last = trace_row + 1 == n_trace;
    for (int c = 0; c < 2; c++)
        for (int i = 0; i < 3; i++)
            for (int l = 0; l < NL; l++)
                out[o++] = last ? gl_sub(cur[L->ch[c].sel[i].c + l], cur[L->rproj[3 * c + i] + l]) : 0;
    const int tr = trace_row + 1 < n_trace;
    for (int c = 0; c < 2; c++)
        for (int i = 0; i < 3; i++)
            for (int l = 0; l < NL; l++)
                out[o++] = tr ? gl_sub(nxt[L->ch[c].acc[i] + l], cur[L->ch[c].sel[i].c + l]) : 0;
    for (int i = 0; i < 6; i++)
        for (int l = 0; l < NL; l++) out[o++] = tr ? gl_sub(nxt[L->rproj[i] + l], cur[L->rproj[i] + l]) : 0;
}

/* deep_ali_merge_ecdsa_double_multirow_streaming (lib.rs:1421-1540):
 * phi at every point (no thread-count quirk), lambda_j = coeffs[j] (NULL:
 * j + 1), then the Z_H quotient. Rows are gathered 64 points at a time. */
int ref_merge_p256(const uint64_t *lde, size_t n_trace, size_t blowup, const uint64_t *coeffs, uint64_t *c_eval) {
    const size_t n = n_trace * blowup, w = ref_p256_width(), k = ref_p256_n_constraints();
    enum { CH = 64 };
    uint64_t *cur = (uint64_t *)malloc(8 * w * CH), *nxt = (uint64_t *)malloc(8 * w * CH);
    uint64_t *cv = (uint64_t *)malloc(8 * k), *row = (uint64_t *)malloc(8 * w), *rown = (uint64_t *)malloc(8 * w);
    for (size_t i0 = 0; i0 < n; i0 += CH) {
        const size_t m = n - i0 < CH ? n - i0 : CH;
        for (size_t c = 0; c < w; c++)
            for (size_t j = 0; j < m; j++) {
                cur[j * w + c] = lde[c * n + i0 + j];
                nxt[j * w + c] = lde[c * n + (i0 + j + blowup) % n];
            }
        for (size_t j = 0; j < m; j++) {
            const size_t i = i0 + j;
            memcpy(row, cur + j * w, 8 * w);
            memcpy(rown, nxt + j * w, 8 * w);
            ref_p256_constraints(row, rown, i / blowup, n_trace, cv);
            uint64_t acc = 0;
            for (size_t q = 0; q < k; q++) acc = gl_add(acc, gl_mul(coeffs ? coeffs[q] : (uint64_t)(q + 1), cv[q]));
            c_eval[i] = acc;
        }
    }
    free(cur);
    free(nxt);
    free(cv);
    free(row);
    free(rown);
    ref_quotient_from_phi(c_eval, n, n_trace);
    return 0;
}

/* ---- native ECDSA pieces of the per-record prover ---- */

/* projective point (X : Y : Z) over F_p, canonical 8-word values */
typedef struct { uint32_t x[8], y[8], z[8]; } ppt;

static void fp_add(const uint32_t *a, const uint32_t *b, uint32_t *r) {
    uint32_t t[9];
    uint64_t c = 0;
    for (int i = 0; i < 8; i++) {
        c += (uint64_t)a[i] + b[i];
        t[i] = (uint32_t)c;
        c >>= 32;
    }
    t[8] = (uint32_t)c;
    uint32_t pp[9];
    memcpy(pp, P_W, 32);
    pp[8] = 0;
    if (w_cmp(t, pp, 9) >= 0) {
        uint64_t br = 0;
        for (int i = 0; i < 9; i++) {
            const uint64_t d = (uint64_t)t[i] - pp[i] - br;
            t[i] = (uint32_t)d;
            br = (d >> 63) & 1;
        }
    }
    memcpy(r, t, 32);
}
static void fp_sub(const uint32_t *a, const uint32_t *b, uint32_t *r) {
    uint32_t nb[8];
    uint64_t br = 0;
    for (int i = 0; i < 8; i++) { /* p - b */
        const uint64_t d = (uint64_t)P_W[i] - b[i] - br;
        nb[i] = (uint32_t)d;
        br = (d >> 63) & 1;
    }
    fp_add(a, nb, r);
}
static void fp_mul(const uint32_t *a, const uint32_t *b, uint32_t *r) { w_mulmod(a, b, P_W, r); }

/* complete projective addition on y^2 = x^3 - 3x + b (Renes-Costello-Batina
 * 2016, alg. 4), the same formula the AIR's add gadget proves */
static void pt_add(const ppt *P, const ppt *Q, const uint32_t *bc, ppt *R) {
    uint32_t t0[8], t1[8], t2[8], t3[8], t4[8], x3[8], y3[8], z3[8];
    fp_mul(P->x, Q->x, t0);
    fp_mul(P->y, Q->y, t1);
    fp_mul(P->z, Q->z, t2);
    fp_add(P->x, P->y, t3);
    fp_add(Q->x, Q->y, t4);
    fp_mul(t3, t4, t3);
    fp_add(t0, t1, t4);
    fp_sub(t3, t4, t3);
    fp_add(P->y, P->z, t4);
    fp_add(Q->y, Q->z, x3);
    fp_mul(t4, x3, t4);
    fp_add(t1, t2, x3);
    fp_sub(t4, x3, t4);
    fp_add(P->x, P->z, x3);
    fp_add(Q->x, Q->z, y3);
    fp_mul(x3, y3, x3);
    fp_add(t0, t2, y3);
    fp_sub(x3, y3, y3);
    fp_mul(bc, t2, z3);
    fp_sub(y3, z3, x3);
    fp_add(x3, x3, z3);
    fp_add(x3, z3, x3);
    fp_sub(t1, x3, z3);
    fp_add(t1, x3, x3);
    fp_mul(bc, y3, y3);
    fp_add(t2, t2, t1);
    fp_add(t1, t2, t2);
    fp_sub(y3, t2, y3);
    fp_sub(y3, t0, y3);
    fp_add(y3, y3, t1);
    fp_add(t1, y3, y3);
    fp_add(t0, t0, t1);
    fp_add(t1, t0, t0);
    fp_sub(t0, t2, t0);
    fp_mul(t4, y3, t1);
    fp_mul(t0, y3, t2);
    fp_mul(x3, z3, y3);
    fp_add(y3, t2, y3);
    fp_mul(t3, x3, x3);
    fp_sub(x3, t1, x3);
    fp_mul(t4, z3, z3);
    fp_mul(t3, t0, t1);
    fp_add(z3, t1, z3);
    memcpy(R->x, x3, 32);
    memcpy(R->y, y3, 32);
    memcpy(R->z, z3, 32);
}
/* k P by double-and-add from the identity (0 : 1 : 0); the complete formula
 * doubles too */
static void pt_scalar(const ppt *P, const uint32_t k[8], const uint32_t *bc, ppt *R) {
    ppt acc;
    memset(&acc, 0, sizeof acc);
    acc.y[0] = 1;
    for (int bit = 255; bit >= 0; bit--) {
        pt_add(&acc, &acc, bc, &acc);
        if ((k[bit / 32] >> (bit % 32)) & 1) pt_add(&acc, P, bc, &acc);
    }
    *R = acc;
}

/* prove_record_ecdsa (mixed_algorithm_zone_rollup.rs:193-322) from the
 * signature on: u1 = e/s, u2 = r/s mod n (e = digest mod n), the two chains
 * from G and Q with the MSB-first bits of u1 and u2, n_trace = 256,
 * blowup 32, lambda_j = j + 1, STIR with 54 queries bound to
 * pk_hash_pre = SHA3-256("ECDSAP256-MIXEDZONE" || qx || qy || r || s || digest);
 * pi_hash = SHA3-256(pk_hash_pre || root_f0). Returns -30 if the signature
 * does not verify (the example asserts native_ecdsa_verify first). */
int ref_p256_ecdsa_bits(const uint8_t qx[32], const uint8_t qy[32], const uint8_t r_be[32], const uint8_t s_be[32],
                        const uint8_t digest[32], uint8_t u1_bits[256], uint8_t u2_bits[256]) {
    uint32_t r[8], s[8], e[8], w[8], u1[8], u2[8], qxw[8], qyw[8];
    w_from_be(r_be, r);
    w_from_be(s_be, s);
    w_from_be(qx, qxw);
    w_from_be(qy, qyw);
    if (w_is_zero(r, 8) || w_is_zero(s, 8) || w_cmp(r, N_W, 8) >= 0 || w_cmp(s, N_W, 8) >= 0) return -30;
    if (w_cmp(qxw, P_W, 8) >= 0 || w_cmp(qyw, P_W, 8) >= 0) return -31;
    {
        uint32_t dw[16] = {0}, q[16];
        w_from_be(digest, dw);
        w_divmod(dw, N_W, q, e);
    }
    w_invmod(s, N_W, w);
    w_mulmod(e, w, N_W, u1);
    w_mulmod(r, w, N_W, u2);
    /* native verify: x(u1 G + u2 Q) mod n == r */
    uint32_t bc[8];
    w_from_be(B_BE, bc);
    ppt G, Q, A, B, R;
    memset(&G, 0, sizeof G);
    memset(&Q, 0, sizeof Q);
    w_from_be(GX_BE, G.x);
    w_from_be(GY_BE, G.y);
    G.z[0] = 1;
    memcpy(Q.x, qxw, 32);
    memcpy(Q.y, qyw, 32);
    Q.z[0] = 1;
    { /* Q on the curve: y^2 = x^3 - 3x + b */
        uint32_t l[8], x2[8], rr[8];
        fp_mul(qyw, qyw, l);
        fp_mul(qxw, qxw, x2);
        fp_mul(x2, qxw, rr);
        fp_sub(rr, qxw, rr);
        fp_sub(rr, qxw, rr);
        fp_sub(rr, qxw, rr);
        fp_add(rr, bc, rr);
        if (w_cmp(l, rr, 8)) return -31;
    }
    pt_scalar(&G, u1, bc, &A);
    pt_scalar(&Q, u2, bc, &B);
    pt_add(&A, &B, bc, &R);
    if (w_is_zero(R.z, 8)) return -30;
    uint32_t zi[8], x[8];
    w_invmod(R.z, P_W, zi);
    fp_mul(R.x, zi, x);
    {
        uint32_t xx[16] = {0}, q[16], xm[8];
        memcpy(xx, x, 32);
        w_divmod(xx, N_W, q, xm);
        if (w_cmp(xm, r, 8)) return -30;
    }
    uint8_t b1[32], b2[32];
    w_to_be(u1, b1);
    w_to_be(u2, b2);
    for (int i = 0; i < 256; i++) {
        u1_bits[i] = (b1[i / 8] >> (7 - i % 8)) & 1;
        u2_bits[i] = (b2[i / 8] >> (7 - i % 8)) & 1;
    }
    return 0;
}

/* The chain outputs at the last row (the example's first fill pass), as
 * 6 x 32-byte big-endian values r_a x,y,z, r_b x,y,z: run the steps natively
 * without a trace. */
static int chain_outputs(const uint8_t *pts, const uint8_t *a_bits, const uint8_t *b_bits, size_t k, uint8_t *rproj) {
    /* one row of scratch is enough: fill each step into a 1-row trace */
    const p256_layout *L = layout();
    uint64_t *tr = (uint64_t *)calloc(L->width, 8);
    int64_t acc[2][3][NL], base[2][3][NL];
    for (int c = 0; c < 2; c++)
        for (int i = 0; i < 3; i++) {
            uint32_t w[8];
            w_from_be(pts + 32 * (6 * c + i), w);
            w_to_limbs(w, acc[c][i]);
            w_from_be(pts + 32 * (6 * c + 3 + i), w);
            w_to_limbs(w, base[c][i]);
        }
    for (size_t r = 0; r < k; r++) {
        tw t = {tr, 1, 0};
        for (int c = 0; c < 2; c++) {
            int64_t out[3][NL];
            const int rc = fill_step(&t, &L->ch[c], (const int64_t(*)[NL])acc[c], (const int64_t(*)[NL])base[c],
                                     (c ? b_bits : a_bits)[r] != 0, out);
            if (rc) {
                free(tr);
                return rc;
            }
            memcpy(acc[c], out, sizeof out);
        }
    }
    free(tr);
    for (int c = 0; c < 2; c++)
        for (int i = 0; i < 3; i++) {
            uint32_t w[9];
            limbs_to_w(acc[c][i], w);
            w_to_be(w, rproj + 32 * (3 * c + i));
        }
    return 0;
}

/* the 12 chain points of an ECDSA record: (G, G) for chain A, (Q, Q) for
 * chain B, Z = 1 */
static void ecdsa_points(const uint8_t qx[32], const uint8_t qy[32], uint8_t pts[12 * 32]) {
    uint8_t one[32] = {0};
    one[31] = 1;
    for (int c = 0; c < 2; c++)
        for (int h = 0; h < 2; h++) {
            uint8_t *p = pts + 32 * (6 * c + 3 * h);
            memcpy(p, c ? qx : GX_BE, 32);
            memcpy(p + 32, c ? qy : GY_BE, 32);
            memcpy(p + 64, one, 32);
        }
}

int ref_p256_ecdsa_trace(const uint8_t qx[32], const uint8_t qy[32], const uint8_t r_be[32], const uint8_t s_be[32],
                         const uint8_t digest[32], size_t ht, uint64_t *tr, uint8_t rproj[6 * 32]) {
    uint8_t u1[256], u2[256], pts[12 * 32];
    int rc = ref_p256_ecdsa_bits(qx, qy, r_be, s_be, digest, u1, u2);
    if (rc) return rc;
    if (ht != 256) return -10;
    ecdsa_points(qx, qy, pts);
    rc = chain_outputs(pts, u1, u2, 256, rproj);
    if (rc) return rc;
    return ref_fill_p256(pts, rproj, u1, 256, u2, 256, ht, tr);
}

int ref_prove_p256_ecdsa(const uint8_t qx[32], const uint8_t qy[32], const uint8_t r_be[32], const uint8_t s_be[32],
                         const uint8_t digest[32], uint8_t **proof, size_t *proof_len, uint8_t root_f0[32],
                         uint8_t pi_hash[32]) {
    const size_t h = 256, n0 = h * 32, w = ref_p256_width();
    uint64_t *tr = (uint64_t *)calloc(w * h, 8);
    uint8_t rproj[6 * 32];
    int rc = ref_p256_ecdsa_trace(qx, qy, r_be, s_be, digest, h, tr, rproj);
    if (rc) {
        free(tr);
        return rc;
    }
    uint64_t *lde = (uint64_t *)malloc(8 * w * n0), *c = (uint64_t *)malloc(8 * n0);
    ref_lde_columns(tr, w, h, 32, 1, lde);
    free(tr);
    ref_merge_p256(lde, h, 32, NULL, c);
    free(lde);
    uint8_t pis[32];
    sha3_ctx s;
    sha3_init(&s);
    sha3_update(&s, "ECDSAP256-MIXEDZONE", 19);
    sha3_update(&s, qx, 32);
    sha3_update(&s, qy, 32);
    sha3_update(&s, r_be, 32);
    sha3_update(&s, s_be, 32);
    sha3_update(&s, digest, 32);
    sha3_final(&s, pis);
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
    sha3_init(&s);
    sha3_update(&s, pis, 32);
    sha3_update(&s, root_f0, 32);
    sha3_final(&s, pi_hash);
    return 0;
}
