// p256_air.hip -- double-chain ECDSA-P256 AIR for gfx950.
//
// Trace (crates/deep_ali/src/p256_ecdsa_double_multirow_air.rs:62-222 over
// p256_scalar_mul_air.rs:74-220, p256_group_air.rs:111-1046 and
// p256_field_air.rs:125-1531): each row holds, for each of two chains, one
// step acc' = bit ? 2 acc + base : 2 acc as an RCB-2016 doubling (13 mul,
// 15 add, 6 sub gadgets) and addition (14 mul, 20 add, 9 sub), every add and
// sub followed by a freeze to [0, p), and three selects. A field element is
// 10 limbs of 26 bits; most cells are the bits of those limbs and of the
// 36-bit biased carries of the mul gadgets' limb identities
// a b - q p - c = 0. The host computes the limbs, quotients and carries of
// every gadget (the "witness words", about 8.6 k per row against 151 k
// cells); `expand_p256_k` turns them into cells through a per-column
// descriptor (word, and which bit / value / carry sign of it).
//
// Composition (eval_ecdsa_double_multirow_per_row, :225-307) with
// lambda_j = j + 1: one wave per (64 LDE points, sub-gadget); the lane is the
// point, so every column read is a coalesced 512-byte run. Each wave sums its
// gadget's weighted constraints lazily in 128 bits: linear terms as
// lambda * cell, booleanity as lambda * (x - x^2); the 19 position identities
// of a mul gadget are regrouped with lambda affine in the position k,
//   sum_k lambda_k sum_{i+j=k} a_i b_j = sum_i a_i ((L0 + i) B + B1),
//   B = sum_j b_j, B1 = sum_j j b_j,
// the q p products likewise with the constant limbs of p, and the carry
// terms as sum_m sc_m (lambda_{m+1} - 2^26 lambda_m). Per-gadget partial
// sums are added per point by `p256_sum_k`, then the Z_H quotient.
#include <string.h>

#include "air.hpp"
#include "bigint.hpp"
#include "gl.hpp"
#include "keccak.hpp"
#include "ntt.hpp"
#include "ntt_dev.hpp"
#include "p256_air.hpp"

namespace sg {

namespace p256 {

constexpr int NL = 10, LB = 26, EC = 270, MCB = 36, MCP = 18, MSP = 19;
constexpr int64_t MC_OFF = 1LL << 35;
// p in 26-bit limbs (P_LIMBS_TIGHT, p256_field.rs:224-235)
constexpr int64_t PL[NL] = {0x3FFFFFF, 0x3FFFFFF, 0x3FFFFFF, 0x003FFFF, 0, 0, 0, 0x0000400, 0x3FF0000, 0x03FFFFF};
constexpr uint64_t SP = 0x3FFFFFF * 3 + 0x003FFFF + 0x0000400 + 0x3FF0000 + 0x03FFFFF;  // sum_j p_j
constexpr uint64_t SP1 = 0x3FFFFFF * (0 + 1 + 2) + 0x003FFFF * 3 + 0x0000400 * 7 + 0x3FF0000 * 8 +
                         0x03FFFFF * 9;  // sum_j j p_j
constexpr uint32_t P_W[8] = {0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0, 0, 0, 1, 0xFFFFFFFF};
constexpr uint32_t N_W[8] = {0xFC632551, 0xF3B9CAC2, 0xA7179E84, 0xBCE6FAAD, 0xFFFFFFFF, 0xFFFFFFFF, 0, 0xFFFFFFFF};
constexpr uint8_t B_BE[32] = {0x5a, 0xc6, 0x35, 0xd8, 0xaa, 0x3a, 0x93, 0xe7, 0xb3, 0xeb, 0xbd,
                              0x55, 0x76, 0x98, 0x86, 0xbc, 0x65, 0x1d, 0x06, 0xb0, 0xcc, 0x53,
                              0xb0, 0xf6, 0x3b, 0xce, 0x3c, 0x3e, 0x27, 0xd2, 0x60, 0x4b};
constexpr uint8_t GX_BE[32] = {0x6b, 0x17, 0xd1, 0xf2, 0xe1, 0x2c, 0x42, 0x47, 0xf8, 0xbc, 0xe6,
                               0xe5, 0x63, 0xa4, 0x40, 0xf2, 0x77, 0x03, 0x7d, 0x81, 0x2d, 0xeb,
                               0x33, 0xa0, 0xf4, 0xa1, 0x39, 0x45, 0xd8, 0x98, 0xc2, 0x96};
constexpr uint8_t GY_BE[32] = {0x4f, 0xe3, 0x42, 0xe2, 0xfe, 0x1a, 0x7f, 0x9b, 0x8e, 0xe7, 0xeb,
                               0x4a, 0x7c, 0x0f, 0x9e, 0x16, 0x2b, 0xce, 0x33, 0x57, 0x6b, 0x31,
                               0x5e, 0xce, 0xcb, 0xb6, 0x40, 0x68, 0x37, 0xbf, 0x51, 0xf5};

enum : uint32_t { GM, GA, GS, GF, GSEL, GTAIL };
// cell descriptor: word << 7 | code; code < 64 is a bit of the word
enum : uint32_t { D_VAL = 64, D_POS = 65, D_NEG = 66 };

// device gadget table entry (constraint order); for a freeze q / qb are the
// diff limbs / bits, for an add pos holds the carries
struct GDesc {
    uint32_t kind, s, a, b, c, cb, q, qb, cy, pos, neg, sel;
};
// boundary and transitions (p256_ecdsa_double_multirow_air.rs:243-304)
struct TailDesc {
    uint32_t s, selc[6], rproj[6], acc[6];
};

// ---------------------------------------------------------------- layout

// RCB-2016 programs as written in p256_group_air.rs: op, destination
// register, two source registers; every add and sub is followed by a freeze
// whose output the destination names.
struct Op {
    uint8_t op, d, a, b;
};
// doubling (alg. 6, p256_group_air.rs:364-438): X Y Z = 0..2, b = 3,
// t0..t3 = 4..7, x3 = 8, y3 = 9, z3 = 10
constexpr Op DBL[34] = {
    {GM, 4, 0, 0},   {GM, 5, 1, 1},    {GM, 6, 2, 2},    {GM, 7, 0, 1},    {GA, 7, 7, 7},    {GM, 10, 0, 2},
    {GA, 10, 10, 10}, {GM, 9, 3, 6},   {GS, 9, 9, 10},   {GA, 8, 9, 9},    {GA, 9, 8, 9},    {GS, 8, 5, 9},
    {GA, 9, 5, 9},   {GM, 9, 8, 9},    {GM, 8, 8, 7},    {GA, 7, 6, 6},    {GA, 6, 6, 7},    {GM, 10, 3, 10},
    {GS, 10, 10, 6}, {GS, 10, 10, 4},  {GA, 7, 10, 10},  {GA, 10, 10, 7},  {GA, 7, 4, 4},    {GA, 4, 7, 4},
    {GS, 4, 4, 6},   {GM, 4, 4, 10},   {GA, 9, 9, 4},    {GM, 4, 1, 2},    {GA, 4, 4, 4},    {GM, 10, 4, 10},
    {GS, 8, 8, 10},  {GM, 10, 4, 5},   {GA, 10, 10, 10}, {GA, 10, 10, 10}};
// addition (alg. 4, p256_group_air.rs:745-850): X1 Y1 Z1 = 0..2,
// X2 Y2 Z2 = 3..5, b = 6, t0..t4 = 7..11, x3 = 12, y3 = 13, z3 = 14
constexpr Op ADD[43] = {
    {GM, 7, 0, 3},    {GM, 8, 1, 4},    {GM, 9, 2, 5},    {GA, 10, 0, 1},   {GA, 11, 3, 4},   {GM, 10, 10, 11},
    {GA, 11, 7, 8},   {GS, 10, 10, 11}, {GA, 11, 1, 2},   {GA, 12, 4, 5},   {GM, 11, 11, 12}, {GA, 12, 8, 9},
    {GS, 11, 11, 12}, {GA, 12, 0, 2},   {GA, 13, 3, 5},   {GM, 12, 12, 13}, {GA, 13, 7, 9},   {GS, 13, 12, 13},
    {GM, 14, 6, 9},   {GS, 12, 13, 14}, {GA, 14, 12, 12}, {GA, 12, 12, 14}, {GS, 14, 8, 12},  {GA, 12, 8, 12},
    {GM, 13, 6, 13},  {GA, 8, 9, 9},    {GA, 9, 8, 9},    {GS, 13, 13, 9},  {GS, 13, 13, 7},  {GA, 8, 13, 13},
    {GA, 13, 8, 13},  {GA, 8, 7, 7},    {GA, 7, 8, 7},    {GS, 7, 7, 9},    {GM, 8, 11, 13},  {GM, 9, 7, 13},
    {GM, 13, 12, 14}, {GA, 13, 13, 9},  {GM, 12, 10, 12}, {GS, 12, 12, 8},  {GM, 14, 11, 14}, {GM, 8, 10, 7},
    {GA, 14, 14, 8}};

struct Gad {
    uint32_t kind, a, b, c, cb, q, qb, cy, pos, neg, sel, w0;  // columns; w0 = first witness word
};
struct Group {
    const Op *prog = nullptr;
    int n_ops = 0, n_in = 0;
    std::vector<Gad> g;       // allocation order (an add / sub, then its freeze)
    std::vector<int> op_gad;  // first gadget of each op
    uint32_t reg[16] = {};    // register columns after the layout pass
};
struct Step {
    uint32_t acc[3], base[3], bit, w_acc, w_base, w_bit;
    Group dbl, add;
    Gad sel[3];
};
struct Layout {
    Step ch[2];
    uint32_t rproj[6], w_rproj, w_b, n_words;
    std::vector<uint32_t> desc;  // one per column
    std::vector<GDesc> gd;       // both chains, constraint order
    TailDesc tail;
};

struct Builder {
    Layout &L;
    uint32_t words = 0;
    uint32_t col() const { return (uint32_t)L.desc.size(); }
    void val(uint32_t w) { L.desc.push_back(w << 7 | D_VAL); }
    // an element block: 10 limb cells, then the 26 bits of each limb
    uint32_t elem(uint32_t w) {
        const uint32_t c0 = col();
        for (int i = 0; i < NL; i++) val(w + i);
        for (int i = 0; i < NL; i++)
            for (int b = 0; b < LB; b++) L.desc.push_back((w + i) << 7 | b);
        return c0;
    }
    void signed_cells(uint32_t w, Gad &g) {
        g.pos = col();
        for (int k = 0; k < NL; k++) L.desc.push_back((w + k) << 7 | D_POS);
        g.neg = col();
        for (int k = 0; k < NL; k++) L.desc.push_back((w + k) << 7 | D_NEG);
    }
    // alloc_mul_layout (p256_group_air.rs:111-137): c, q, carry bits
    Gad mul(uint32_t a, uint32_t b) {
        Gad g{GM, a, b};
        g.w0 = words;
        words += 2 * NL + MCP;
        g.c = elem(g.w0);
        g.cb = g.c + NL;
        g.q = elem(g.w0 + NL);
        g.qb = g.q + NL;
        g.cy = col();
        for (int k = 0; k < MCP; k++)
            for (int b = 0; b < MCB; b++) L.desc.push_back((g.w0 + 2 * NL + k) << 7 | b);
        return g;
    }
    // alloc_add_layout (:139-161): c, carries
    Gad addg(uint32_t a, uint32_t b) {
        Gad g{GA, a, b};
        g.w0 = words;
        words += 2 * NL;
        g.c = elem(g.w0);
        g.cb = g.c + NL;
        g.pos = col();
        for (int k = 0; k < NL; k++) val(g.w0 + NL + k);
        return g;
    }
    // alloc_sub_layout (:163-187): c, positive and negative carries
    Gad subg(uint32_t a, uint32_t b) {
        Gad g{GS, a, b};
        g.w0 = words;
        words += 2 * NL;
        g.c = elem(g.w0);
        g.cb = g.c + NL;
        signed_cells(g.w0 + NL, g);
        return g;
    }
    // alloc_freeze_layout (:189-216): diff, c, positive and negative carries
    Gad freeze(uint32_t a) {
        Gad g{GF, a, 0};
        g.w0 = words;
        words += 3 * NL;
        g.q = elem(g.w0);
        g.qb = g.q + NL;
        g.c = elem(g.w0 + NL);
        g.cb = g.c + NL;
        signed_cells(g.w0 + 2 * NL, g);
        return g;
    }
    // build_group_double_layout / build_group_add_layout: the b block, then
    // the program's gadgets in order
    void group(Group &G, const Op *prog, int n_ops, const uint32_t *in, int n_in) {
        G.prog = prog;
        G.n_ops = n_ops;
        G.n_in = n_in;
        for (int i = 0; i < n_in; i++) G.reg[i] = in[i];
        G.reg[n_in] = elem(L.w_b);
        for (int k = 0; k < n_ops; k++) {
            const Op o = prog[k];
            G.op_gad.push_back((int)G.g.size());
            if (o.op == GM) {
                G.g.push_back(mul(G.reg[o.a], G.reg[o.b]));
                G.reg[o.d] = G.g.back().c;
            } else {
                G.g.push_back(o.op == GA ? addg(G.reg[o.a], G.reg[o.b]) : subg(G.reg[o.a], G.reg[o.b]));
                G.g.push_back(freeze(G.g.back().c));
                G.reg[o.d] = G.g.back().c;
            }
        }
    }
    // build_ecdsa_double_multirow_layout (p256_ecdsa_double_multirow_air.rs:
    // 62-117) / build_scalar_mul_step_layout (p256_scalar_mul_air.rs:74-155)
    void step(Step &S) {
        S.w_acc = words;
        words += 3 * NL;
        S.w_base = words;
        words += 3 * NL;
        S.w_bit = words++;
        for (int i = 0; i < 3; i++) {
            S.acc[i] = col();
            for (int l = 0; l < NL; l++) val(S.w_acc + NL * i + l);
        }
        for (int i = 0; i < 3; i++) {
            S.base[i] = col();
            for (int l = 0; l < NL; l++) val(S.w_base + NL * i + l);
        }
        S.bit = col();
        val(S.w_bit);
        group(S.dbl, DBL, 34, S.acc, 3);
        const uint32_t two[3] = {S.dbl.reg[8], S.dbl.reg[9], S.dbl.reg[10]};
        const uint32_t in[6] = {two[0], two[1], two[2], S.base[0], S.base[1], S.base[2]};
        group(S.add, ADD, 43, in, 6);
        for (int i = 0; i < 3; i++) {
            Gad g{GSEL, S.add.reg[12 + i], two[i]};
            g.w0 = words;
            words += NL;
            g.c = elem(g.w0);
            g.cb = g.c + NL;
            g.sel = S.bit;
            S.sel[i] = g;
        }
    }
};

uint32_t gad_constraints(uint32_t kind) {
    switch (kind) {
    case GM: return 2 * EC + MSP + MCP * MCB;
    case GA: return EC + 2 * NL;
    case GS: return EC + 4 * NL + 1;
    case GF: return 2 * EC + 5 * NL + 1;
    default: return EC + NL + 1;
    }
}

// eval_group_*_gadget order: muls, adds, subs, freezes after adds, freezes
// after subs
void group_order(const Group &G, std::vector<const Gad *> &out) {
    for (int kind = 0; kind < 5; kind++)
        for (size_t i = 0; i < G.g.size(); i++) {
            const Gad &g = G.g[i];
            const bool take = kind < 3 ? g.kind == (kind == 0 ? GM : kind == 1 ? GA : GS)
                                       : g.kind == GF && G.g[i - 1].kind == (kind == 3 ? GA : GS);
            if (take) out.push_back(&g);
        }
}

const Layout &layout() {
    static const Layout *L = [] {
        Layout *l = new Layout();
        l->desc.reserve(P256_W);
        Builder b{*l};
        l->w_b = b.words;
        b.words += NL;
        b.step(l->ch[0]);
        b.step(l->ch[1]);
        l->w_rproj = b.words;
        b.words += 6 * NL;
        for (int i = 0; i < 6; i++) {
            l->rproj[i] = b.col();
            for (int k = 0; k < NL; k++) b.val(l->w_rproj + NL * i + k);
        }
        l->n_words = b.words;
        uint32_t s = 0;
        for (int c = 0; c < 2; c++) {
            std::vector<const Gad *> ord;
            group_order(l->ch[c].dbl, ord);
            group_order(l->ch[c].add, ord);
            for (int i = 0; i < 3; i++) ord.push_back(&l->ch[c].sel[i]);
            for (const Gad *g : ord) {
                l->gd.push_back(GDesc{g->kind, s, g->a, g->b, g->c, g->cb, g->q, g->qb, g->cy, g->pos, g->neg, g->sel});
                s += gad_constraints(g->kind);
            }
        }
        l->tail.s = s;
        for (int c = 0; c < 2; c++)
            for (int i = 0; i < 3; i++) {
                l->tail.selc[3 * c + i] = l->ch[c].sel[i].c;
                l->tail.acc[3 * c + i] = l->ch[c].acc[i];
            }
        for (int i = 0; i < 6; i++) l->tail.rproj[i] = l->rproj[i];
        if (l->desc.size() != P256_W || s + 180 != P256_K || l->gd.size() != 260)
            throw Error(E_INTERNAL, "p256 layout: sizes differ from the reference");
        return l;
    }();
    return *L;
}

// ---------------------------------------------------------------- host rows

using Limbs = std::array<int64_t, NL>;

void be_to_words(const uint8_t *be, uint32_t w[8]) {
    for (int i = 0; i < 8; i++)
        w[i] = (uint32_t)be[31 - 4 * i] | (uint32_t)be[30 - 4 * i] << 8 | (uint32_t)be[29 - 4 * i] << 16 |
               (uint32_t)be[28 - 4 * i] << 24;
}
void words_to_be(const uint32_t w[8], uint8_t *be) {
    for (int i = 0; i < 8; i++)
        for (int b = 0; b < 4; b++) be[31 - 4 * i - b] = (uint8_t)(w[i] >> (8 * b));
}
// 26-bit limbs of a value held in nw words (bits above 260 dropped)
Limbs to_limbs(const uint32_t *w, int nw) {
    Limbs l;
    for (int i = 0; i < NL; i++) {
        const int bit = LB * i, k = bit / 32, sh = bit % 32;
        const uint64_t lo = k < nw ? w[k] : 0, hi = k + 1 < nw ? w[k + 1] : 0;
        l[i] = (int64_t)(((lo | hi << 32) >> sh) & ((1u << LB) - 1));
    }
    return l;
}
void from_limbs(const Limbs &l, uint32_t w[9]) {
    memset(w, 0, 36);
    for (int i = 0; i < NL; i++) {
        const int bit = LB * i, k = bit / 32, sh = bit % 32;
        const uint64_t v = (uint64_t)l[i] << sh;
        w[k] |= (uint32_t)v;
        if (k + 1 < 9) w[k + 1] |= (uint32_t)(v >> 32);
    }
}
Limbs be_limbs(const uint8_t *be) {
    uint32_t w[8];
    be_to_words(be, w);
    return to_limbs(w, 8);
}

int cmp8(const uint32_t *a, const uint32_t *b);

// r = c mod p for c < 2^512 (16 little-endian words): the NIST fast
// reduction for p = 2^256 - 2^224 + 2^192 + 2^96 - 1 (FIPS 186-4 D.2.3,
// Solinas 1999), s1 + 2 s2 + 2 s3 + s4 + s5 - d1 - d2 - d3 - d4 summed word
// by word in signed 64 bits, then p added or subtracted until r is in [0, p).
void p256_reduce(const uint32_t *c, uint32_t r[8]) {
    const int64_t c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3], c4 = c[4], c5 = c[5], c6 = c[6], c7 = c[7];
    const int64_t c8 = c[8], c9 = c[9], c10 = c[10], c11 = c[11], c12 = c[12], c13 = c[13], c14 = c[14],
                  c15 = c[15];
    int64_t t[8];
    t[0] = c0 + c8 + c9 - c11 - c12 - c13 - c14;
    t[1] = c1 + c9 + c10 - c12 - c13 - c14 - c15;
    t[2] = c2 + c10 + c11 - c13 - c14 - c15;
    t[3] = c3 + 2 * c11 + 2 * c12 + c13 - c15 - c8 - c9;
    t[4] = c4 + 2 * c12 + 2 * c13 + c14 - c9 - c10;
    t[5] = c5 + 2 * c13 + 2 * c14 + c15 - c10 - c11;
    t[6] = c6 + 3 * c14 + 2 * c15 + c13 - c8 - c9;
    t[7] = c7 + 3 * c15 + c8 - c10 - c11 - c12 - c13;
    // value = sum t_j 2^(32 j): carry-propagate to 8 words plus a signed top k
    int64_t k = 0;
    for (int j = 0; j < 8; j++) {
        const int64_t v = t[j] + k;
        r[j] = (uint32_t)v;
        k = v >> 32;  // arithmetic shift: floor division by 2^32
    }
    // (r, k) is a signed multiple-word value congruent to c; |k| <= 6
    while (k != 0 || cmp8(r, P_W) >= 0) {
        int64_t br = 0;
        if (k < 0) {  // add p
            for (int j = 0; j < 8; j++) {
                const int64_t v = (int64_t)r[j] + P_W[j] + br;
                r[j] = (uint32_t)v;
                br = v >> 32;
            }
        } else {  // subtract p
            for (int j = 0; j < 8; j++) {
                const int64_t v = (int64_t)r[j] - (int64_t)P_W[j] + br;
                r[j] = (uint32_t)v;
                br = v >> 32;
            }
        }
        k += br;
    }
}

// x * y mod 2^256 (8 words each)
void mul_lo8(const uint32_t *x, const uint32_t *y, uint32_t out[8]) {
    uint32_t o[8] = {};
    for (int i = 0; i < 8; i++) {
        uint64_t carry = 0;
        for (int j = 0; i + j < 8; j++) {
            const uint64_t t = (uint64_t)x[i] * y[j] + o[i + j] + carry;
            o[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
    }
    memcpy(out, o, 32);
}

// p^-1 mod 2^256 by Newton's iteration x <- x (2 - p x), from p^-1 = -1 mod 2^32
const uint32_t *p_inv256() {
    static const std::array<uint32_t, 8> inv = [] {
        std::array<uint32_t, 8> x{};
        x[0] = 0xFFFFFFFF;
        for (int it = 0; it < 3; it++) {
            uint32_t px[8], two_m[8];
            mul_lo8(P_W, x.data(), px);
            // 2 - p x mod 2^256
            uint64_t br = 0;
            for (int j = 0; j < 8; j++) {
                const uint64_t v = (uint64_t)(j == 0 ? 2 : 0) - px[j] - br;
                two_m[j] = (uint32_t)v;
                br = (v >> 63) & 1;
            }
            uint32_t nx[8];
            mul_lo8(x.data(), two_m, nx);
            memcpy(x.data(), nx, 32);
        }
        return x;
    }();
    return inv.data();
}

// c = prod mod p and q = (prod - c) / p for prod = a b < p^2: the division
// is exact, so q = (prod - c) p^-1 mod 2^256 (q < p < 2^256)
void p256_divmod(const uint32_t prod[16], uint32_t q[8], uint32_t c[8]) {
    p256_reduce(prod, c);
    uint32_t d[8];
    uint64_t br = 0;
    for (int j = 0; j < 8; j++) {
        const uint64_t v = (uint64_t)prod[j] - c[j] - br;
        d[j] = (uint32_t)v;
        br = (v >> 63) & 1;
    }
    mul_lo8(d, p_inv256(), q);
}

// fill_mul_gadget (p256_field_air.rs:953-1042): c = a b mod p, q = (a b - c) / p
// and the biased carries of a b - q p - c = 0; false if the identity fails
bool row_mul(const Gad &g, const Limbs &a, const Limbs &b, uint64_t *W, Limbs &c) {
    uint32_t aw[9], bw[9];
    from_limbs(a, aw);
    from_limbs(b, bw);
    if (aw[8] || bw[8]) return false;  // operands are below p in every program
    uint32_t prod[16] = {};
    for (int i = 0; i < 8; i++) {
        uint64_t carry = 0;
        for (int j = 0; j < 8; j++) {
            const uint64_t t = (uint64_t)aw[i] * bw[j] + prod[i + j] + carry;
            prod[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        prod[i + 8] = (uint32_t)carry;
    }
    uint32_t q[8], r[8];
    p256_divmod(prod, q, r);
    c = to_limbs(r, 8);
    const Limbs ql = to_limbs(q, 8);
    for (int i = 0; i < NL; i++) {
        W[g.w0 + i] = (uint64_t)c[i];
        W[g.w0 + NL + i] = (uint64_t)ql[i];
    }
    int64_t prev = 0;
    for (int k = 0; k < MSP; k++) {
        int64_t acc = prev - (k < NL ? c[k] : 0);
        const int lo = k > NL - 1 ? k - (NL - 1) : 0, hi = k < NL - 1 ? k : NL - 1;
        for (int i = lo; i <= hi; i++) acc += a[i] * b[k - i] - ql[i] * PL[k - i];
        if (k == MSP - 1) return acc == 0;
        if (acc & ((1 << LB) - 1)) return false;
        const int64_t cy = acc >> LB;  // exact multiple: the shift is the division
        if (cy >= MC_OFF || cy < -MC_OFF) return false;
        W[g.w0 + 2 * NL + k] = (uint64_t)(cy + MC_OFF);
        prev = cy;
    }
    return true;
}
// fill_add_gadget (:344-402): c = a + b limb by limb, carries 0 / 1
void row_add(const Gad &g, const Limbs &a, const Limbs &b, uint64_t *W, Limbs &c) {
    int64_t prev = 0;
    for (int k = 0; k < NL; k++) {
        const int64_t s = a[k] + b[k] + prev;
        c[k] = s & ((1 << LB) - 1);
        prev = s >> LB;
        W[g.w0 + k] = (uint64_t)c[k];
        W[g.w0 + NL + k] = (uint64_t)prev;
    }
}
// Euclidean division by 2^26
int64_t split26(int64_t x, int64_t &rem) {
    rem = x & ((1 << LB) - 1);
    return x >> LB;  // arithmetic shift: the floor
}
// fill_sub_gadget (:570-634): c = a + p - b, carries -1 / 0 / 1
void row_sub(const Gad &g, const Limbs &a, const Limbs &b, uint64_t *W, Limbs &c) {
    int64_t prev = 0;
    for (int k = 0; k < NL; k++) {
        prev = split26(a[k] + PL[k] - b[k] + prev, c[k]);
        W[g.w0 + k] = (uint64_t)c[k];
        W[g.w0 + NL + k] = (uint64_t)prev;
    }
}
// fill_freeze_gadget (:1374-1447): diff = a - p; c = a when the last carry is
// -1 (a < p), else diff
void row_freeze(const Gad &g, const Limbs &a, uint64_t *W, Limbs &c) {
    int64_t prev = 0, cy[NL];
    Limbs d;
    for (int k = 0; k < NL; k++) {
        cy[k] = prev = split26(a[k] - PL[k] + prev, d[k]);
    }
    for (int k = 0; k < NL; k++) {
        c[k] = cy[NL - 1] == -1 ? a[k] : d[k];
        W[g.w0 + k] = (uint64_t)d[k];
        W[g.w0 + NL + k] = (uint64_t)c[k];
        W[g.w0 + 2 * NL + k] = (uint64_t)cy[k];
    }
}
// fill_group_double_gadget / fill_group_add_gadget: run the program; reg
// holds every register afterwards (results x3, y3, z3 at 8..10 for the
// doubling, 12..14 for the addition)
bool row_group(const Group &G, const Limbs *in, const Limbs &bconst, uint64_t *W, Limbs *reg) {
    for (int i = 0; i < G.n_in; i++) reg[i] = in[i];
    reg[G.n_in] = bconst;
    for (int k = 0; k < G.n_ops; k++) {
        const Op o = G.prog[k];
        const Gad &g = G.g[G.op_gad[k]];
        Limbs c;
        if (o.op == GM) {
            if (!row_mul(g, reg[o.a], reg[o.b], W, c)) return false;
        } else {
            Limbs raw;
            if (o.op == GA) row_add(g, reg[o.a], reg[o.b], W, raw);
            else row_sub(g, reg[o.a], reg[o.b], W, raw);
            row_freeze(G.g[G.op_gad[k] + 1], raw, W, c);
        }
        reg[o.d] = c;
    }
    return true;
}
// fill_scalar_mul_step_gadget (p256_scalar_mul_air.rs:174-220); acc <- acc'
bool row_step(const Step &S, Limbs *acc, const Limbs *base, int bit, const Limbs &bconst, uint64_t *W) {
    for (int i = 0; i < 3; i++)
        for (int l = 0; l < NL; l++) {
            W[S.w_acc + NL * i + l] = (uint64_t)acc[i][l];
            W[S.w_base + NL * i + l] = (uint64_t)base[i][l];
        }
    W[S.w_bit] = (uint64_t)bit;
    Limbs rd[16], ra[16], in[6];
    if (!row_group(S.dbl, acc, bconst, W, rd)) return false;
    for (int i = 0; i < 3; i++) {
        in[i] = rd[8 + i];
        in[3 + i] = base[i];
    }
    if (!row_group(S.add, in, bconst, W, ra)) return false;
    for (int i = 0; i < 3; i++) {
        acc[i] = bit ? ra[12 + i] : rd[8 + i];
        for (int l = 0; l < NL; l++) W[S.sel[i].w0 + l] = (uint64_t)acc[i][l];
    }
    return true;
}

// ---- the point sequence of a chain on native F_p arithmetic, so that the
// rows' witness words (the gadget quotients and carries, about 10 us a row)
// can be filled in parallel once each row's accumulator is known. Montgomery
// form on 4 x 64-bit words, R = 2^256; p = 2^256 - 2^224 + 2^192 + 2^96 - 1
// is -1 mod 2^64, so -p^-1 mod 2^64 = 1. The programs are the gadgets' own
// (DBL / ADD) and every result is the canonical residue, as in row_group.
struct M4 {
    uint64_t w[4];
};
constexpr uint64_t PM4[4] = {~0ULL, 0x00000000FFFFFFFFULL, 0, 0xFFFFFFFF00000001ULL};
constexpr M4 R2M4 = {{0x3ULL, 0xfffffffbffffffffULL, 0xfffffffffffffffeULL, 0x4fffffffdULL}};  // R^2 mod p

static bool m4_ge_p(const uint64_t *t) {
    for (int i = 3; i >= 0; i--)
        if (t[i] != PM4[i]) return t[i] > PM4[i];
    return true;
}
static void m4_sub_p(uint64_t *t) {
    unsigned __int128 br = 0;
    for (int i = 0; i < 4; i++) {
        const unsigned __int128 d = (unsigned __int128)t[i] - PM4[i] - br;
        t[i] = (uint64_t)d;
        br = (d >> 64) ? 1 : 0;
    }
}
static M4 m4_add(const M4 &a, const M4 &b) {
    M4 r;
    unsigned __int128 c = 0;
    for (int i = 0; i < 4; i++) {
        c += (unsigned __int128)a.w[i] + b.w[i];
        r.w[i] = (uint64_t)c;
        c >>= 64;
    }
    if (c || m4_ge_p(r.w)) m4_sub_p(r.w);
    return r;
}
static M4 m4_sub(const M4 &a, const M4 &b) {
    M4 r;
    unsigned __int128 br = 0;
    for (int i = 0; i < 4; i++) {
        const unsigned __int128 d = (unsigned __int128)a.w[i] - b.w[i] - br;
        r.w[i] = (uint64_t)d;
        br = (d >> 64) ? 1 : 0;
    }
    if (br) {  // + p
        unsigned __int128 c = 0;
        for (int i = 0; i < 4; i++) {
            c += (unsigned __int128)r.w[i] + PM4[i];
            r.w[i] = (uint64_t)c;
            c >>= 64;
        }
    }
    return r;
}
// a b R^-1 mod p (CIOS), a, b < p
static M4 m4_mul(const M4 &a, const M4 &b) {
    uint64_t t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
        unsigned __int128 c = 0;
        for (int j = 0; j < 4; j++) {
            c += (unsigned __int128)a.w[j] * b.w[i] + t[j];
            t[j] = (uint64_t)c;
            c >>= 64;
        }
        c += t[4];
        t[4] = (uint64_t)c;
        t[5] = (uint64_t)(c >> 64);
        const uint64_t m = t[0];
        c = ((unsigned __int128)m * PM4[0] + t[0]) >> 64;
        for (int j = 1; j < 4; j++) {
            c += (unsigned __int128)m * PM4[j] + t[j];
            t[j - 1] = (uint64_t)c;
            c >>= 64;
        }
        c += t[4];
        t[3] = (uint64_t)c;
        t[4] = t[5] + (uint64_t)(c >> 64);
    }
    M4 r = {{t[0], t[1], t[2], t[3]}};
    if (t[4] || m4_ge_p(r.w)) m4_sub_p(r.w);
    return r;
}
static M4 m4_of(const Limbs &l) {
    uint32_t w[9];
    from_limbs(l, w);
    M4 x;
    for (int i = 0; i < 4; i++) x.w[i] = (uint64_t)w[2 * i] | (uint64_t)w[2 * i + 1] << 32;
    return m4_mul(x, R2M4);
}
static Limbs limbs_of(const M4 &x) {
    const M4 one = {{1, 0, 0, 0}};
    const M4 v = m4_mul(x, one);
    uint32_t w[8];
    for (int i = 0; i < 4; i++) {
        w[2 * i] = (uint32_t)v.w[i];
        w[2 * i + 1] = (uint32_t)(v.w[i] >> 32);
    }
    return to_limbs(w, 8);
}
static void run_native(const Op *prog, int n_ops, M4 *reg) {
    for (int k = 0; k < n_ops; k++) {
        const Op o = prog[k];
        reg[o.d] = o.op == GM ? m4_mul(reg[o.a], reg[o.b]) : o.op == GA ? m4_add(reg[o.a], reg[o.b])
                                                                         : m4_sub(reg[o.a], reg[o.b]);
    }
}
// acc_seq[r] = the accumulator entering row r, r <= h (inputs below p)
static void chain_points(const P256Chains &in, int c, uint64_t h, std::array<Limbs, 3> *acc_seq) {
    const M4 bm = m4_of(be_limbs(B_BE));
    M4 acc[3], base[3];
    for (int i = 0; i < 3; i++) {
        acc[i] = m4_of(be_limbs(in.pts[6 * c + i].data()));
        base[i] = m4_of(be_limbs(in.pts[6 * c + 3 + i].data()));
    }
    const std::vector<uint8_t> &bits = c ? in.b_bits : in.a_bits;
    for (uint64_t r = 0;; r++) {
        for (int i = 0; i < 3; i++) acc_seq[r][i] = limbs_of(acc[i]);
        if (r == h) break;
        M4 rd[16], ra[16];
        for (int i = 0; i < 3; i++) rd[i] = acc[i];
        rd[3] = bm;
        run_native(DBL, 34, rd);
        for (int i = 0; i < 3; i++) {
            ra[i] = rd[8 + i];
            ra[3 + i] = base[i];
        }
        ra[6] = bm;
        run_native(ADD, 43, ra);
        const int bit = r < bits.size() ? (bits[r] != 0) : 0;
        for (int i = 0; i < 3; i++) acc[i] = bit ? ra[12 + i] : rd[8 + i];
    }
}
// rows [r0, r1) of chain c from the known accumulators: row_step's own
// accumulator out of each row must equal the next native one
static bool chain_rows_range(const P256Chains &in, int c, uint64_t r0, uint64_t r1,
                             const std::array<Limbs, 3> *acc_seq, uint64_t *words) {
    const Layout &L = layout();
    const uint64_t nw = L.n_words;
    const Limbs bconst = be_limbs(B_BE);
    Limbs base[3];
    for (int i = 0; i < 3; i++) base[i] = be_limbs(in.pts[6 * c + 3 + i].data());
    const std::vector<uint8_t> &bits = c ? in.b_bits : in.a_bits;
    for (uint64_t r = r0; r < r1; r++) {
        Limbs acc[3] = {acc_seq[r][0], acc_seq[r][1], acc_seq[r][2]};
        const int bit = r < bits.size() ? (bits[r] != 0) : 0;
        if (!row_step(L.ch[c], acc, base, bit, bconst, words + r * nw)) return false;
        for (int i = 0; i < 3; i++)
            if (acc[i] != acc_seq[r + 1][i]) return false;
    }
    return true;
}

// The words shared by both chains (the curve constant b and r_proj, bound to
// the chain outputs or given) in every row.
void shared_rows(const P256Chains &in, uint64_t h, const Limbs (*out)[3], uint64_t *words, std::array<Be32, 6> &rproj) {
    const Layout &L = layout();
    const uint64_t nw = L.n_words;
    const Limbs bconst = be_limbs(B_BE);
    Limbs rp[6];
    for (int i = 0; i < 6; i++) {
        if (in.bind_outputs) {
            rp[i] = out[i / 3][i % 3];
            uint32_t w[9];
            from_limbs(rp[i], w);
            words_to_be(w, rproj[i].data());
        } else {
            rp[i] = be_limbs(in.rproj[i].data());
            rproj[i] = in.rproj[i];
        }
    }
    for (uint64_t r = 0; r < h; r++) {
        uint64_t *W = words + r * nw;
        for (int l = 0; l < NL; l++) W[L.w_b + l] = (uint64_t)bconst[l];
        for (int i = 0; i < 6; i++)
            for (int l = 0; l < NL; l++) W[L.w_rproj + NL * i + l] = (uint64_t)rp[i][l];
    }
}

// All h rows of K proofs (words + k * h * n_words): the accumulators of every
// row natively, chain by chain, then the rows' words in parallel chunks.
void batch_words(const std::vector<P256Chains> &in, uint64_t h, uint64_t *words,
                 std::vector<std::array<Be32, 6>> &rp) {
    const size_t K = in.size();
    const uint64_t per = h * layout().n_words;
    std::vector<std::array<Limbs, 3>> seq(2 * K * (h + 1));
    auto seq_of = [&](size_t t) { return seq.data() + t * (h + 1); };
    parallel_for(2 * K, [&](size_t t) { chain_points(in[t / 2], (int)(t % 2), h, seq_of(t)); });
    constexpr uint64_t RCH = 16;  // rows per task
    const uint64_t chunks = (h + RCH - 1) / RCH;
    std::vector<uint8_t> bad(2 * K * chunks, 0);
    parallel_for(2 * K * chunks, [&](size_t u) {
        const size_t t = u / chunks, k = t / 2;
        const uint64_t r0 = (u % chunks) * RCH, r1 = std::min(h, r0 + RCH);
        bad[u] = !chain_rows_range(in[k], (int)(t % 2), r0, r1, seq_of(t), words + k * per);
    });
    for (size_t u = 0; u < bad.size(); u++)
        SG_REQUIRE(!bad[u], "p256 trace: a field product does not close (inputs must be below p)");
    rp.resize(K);
    parallel_for(K, [&](size_t k) {
        Limbs out[2][3];
        for (int c = 0; c < 2; c++) {
            // each chain's output: the accumulator after its last bit
            const std::vector<uint8_t> &bits = c ? in[k].b_bits : in[k].a_bits;
            for (int i = 0; i < 3; i++) out[c][i] = seq_of(2 * k + c)[bits.size()][i];
        }
        shared_rows(in[k], h, out, words + k * per, rp[k]);
    });
}

void check_inputs(const std::vector<P256Chains> &in, uint64_t h) {
    SG_REQUIRE(is_pow2(h) && h >= 2 && h <= (1u << 16), "p256 trace: height must be a power of two in [2, 2^16]");
    uint32_t pw[8];
    for (const P256Chains &x : in) {
        SG_REQUIRE(x.a_bits.size() <= h && x.b_bits.size() <= h, "p256 trace: more bits than rows");
        SG_REQUIRE(!x.bind_outputs || (!x.a_bits.empty() && !x.b_bits.empty()),
                   "p256 trace: binding r_proj to the outputs needs at least one bit per chain");
        for (const Be32 &v : x.pts) {
            be_to_words(v.data(), pw);
            bool below = false;
            for (int i = 7; i >= 0; i--)
                if (pw[i] != P_W[i]) {
                    below = pw[i] < P_W[i];
                    break;
                }
            SG_REQUIRE(below, "p256 trace: point coordinates must be below p");
        }
    }
}

uint64_t cell(uint32_t d, const uint64_t *W) {
    const uint64_t v = W[d >> 7];
    const uint32_t code = d & 127;
    if (code < 64) return (v >> code) & 1;
    if (code == D_VAL) return v;
    return code == D_POS ? (uint64_t)(v == 1) : (uint64_t)(v == ~0ULL);
}

// ---------------------------------------------------------------- device

// one cell per lane, rows fastest: tr_g[col * h + row], g = blockIdx.y
__global__ __launch_bounds__(256) void expand_p256_k(const uint64_t *words, const uint32_t *desc, uint32_t nw,
                                                     uint32_t log_h, uint64_t *tr) {
    const uint64_t g = blockIdx.y, h = 1ULL << log_h;
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= h * P256_W) return;
    const uint64_t row = t & (h - 1), col = t >> log_h;
    const uint32_t d = desc[col];
    const uint64_t v = words[(g * h + row) * nw + (d >> 7)];
    const uint32_t code = d & 127;
    uint64_t x;
    if (code < 64) x = (v >> code) & 1;
    else if (code == D_VAL) x = v;
    else x = code == D_POS ? (uint64_t)(v == 1) : (uint64_t)(v == ~0ULL);
    tr[g * P256_W * h + t] = x;
}

// expand_p256_k and the odd coset in one pass: a workgroup expands the cells
// of 8192 / h whole columns of proof blockIdx.y (h = 2^LOG_H rows), writes
// them to the trace and keeps them in LDS, then runs the odd-coset transform
// of lde_odd_coset on them (inverse steps, twist[i] = w_{2h}^i / h, forward
// steps) and writes the odd half of the blowup-2 LDE; the trace is not read
// back from HBM.
template <int LOG_H>
__global__ __launch_bounds__(NT, 4) void expand_odd_p256_k(const uint64_t *words, const uint32_t *desc, uint32_t nw,
                                                           uint64_t *tr, uint64_t *odd, const uint64_t *twist,
                                                           const uint64_t *tw_f, const uint64_t *tw_i) {
    __shared__ uint64_t lds[LDS_ELEMS];
    constexpr uint32_t S = 1u << LOG_H;
    constexpr uint32_t LOG_G = TILE_LOG - LOG_H;
    const uint64_t k = blockIdx.y;
    uint64_t blk = blockIdx.x;
    if ((gridDim.x & 7) == 0) blk = (blk & 7) * (gridDim.x >> 3) + (blk >> 3);
    const uint64_t col0 = blk << LOG_G;
    const uint64_t *wk = words + k * S * nw;
    uint64_t *trk = tr + k * P256_W * S, *oddk = odd + k * P256_W * S;
    {
        // the lane's descriptor reads, then its word reads, each all in flight
        // before the first is used (the word address needs the descriptor).
        // Columns past the last read the last one's cells (in bounds, not
        // stored), so no branch splits the reads: a conditional read each
        // gets a wait for everything before it.
        uint32_t dv[EPT];
        uint64_t wv[EPT];
#pragma unroll
        for (uint32_t u = 0; u < EPT; u++) {
            const uint64_t col = col0 + ((threadIdx.x + u * NT) >> LOG_H);
            dv[u] = desc[col < P256_W ? col : P256_W - 1];
        }
#pragma unroll
        for (uint32_t u = 0; u < EPT; u++) {
            const uint64_t row = (threadIdx.x + u * NT) & (S - 1);
            wv[u] = wk[row * nw + (dv[u] >> 7)];
        }
#pragma unroll
        for (uint32_t u = 0; u < EPT; u++) {
            const uint32_t e = threadIdx.x + u * NT;
            const uint64_t col = col0 + (e >> LOG_H), row = e & (S - 1);
            uint64_t x = 0;
            if (col < P256_W) {
                const uint64_t v = wv[u];
                const uint32_t code = dv[u] & 127;
                if (code < 64) x = (v >> code) & 1;
                else if (code == D_VAL) x = v;
                else x = code == D_POS ? (uint64_t)(v == 1) : (uint64_t)(v == ~0ULL);
                trk[col * S + row] = x;
            }
            lds[padi<LOG_H>(e)] = x;
        }
    }
    __syncthreads();
    stockham_steps<LOG_H, 0, true>(lds, TILE, tw_i);
    {
        uint64_t tv[EPT];
#pragma unroll
        for (uint32_t u = 0; u < EPT; u++) tv[u] = twist[(threadIdx.x + u * NT) & (S - 1)];
#pragma unroll
        for (uint32_t u = 0; u < EPT; u++) {
            const uint32_t e = threadIdx.x + u * NT;
            lds[padi<LOG_H>(e)] = mul_lazy(lds[padi<LOG_H>(e)], tv[u]);  // the forward steps take any u64
        }
    }
    __syncthreads();
    stockham_steps<LOG_H, 0, false>(lds, TILE, tw_f);
#pragma unroll 2
    for (uint32_t u = 0; u < EPT; u++) {
        const uint32_t e = threadIdx.x + u * NT;
        const uint64_t col = col0 + (e >> LOG_H);
        if (col < P256_W) oddk[col * S + (e & (S - 1))] = canon(lds[padi<LOG_H>(e)]);
    }
}

struct Sums {
    Acc pos{0, 0}, neg{0, 0}, rpos{0, 0}, rneg{0, 0};  // value = pos - neg - 2^26 (rpos - rneg)
    __device__ uint64_t value() const {
        return sub(sub(acc_val(pos), acc_val(neg)), mul(sub(acc_val(rpos), acc_val(rneg)), 1ULL << LB));
    }
};

// a += lambda_slot x for the 1-based slot index: slot itself (GEN = false,
// lambda_j = j + 1) or T[slot - 1] (GEN: the caller's coefficients)
template <bool GEN>
__device__ __forceinline__ void wmad(Acc &a, uint64_t x, uint32_t slot, const uint64_t *T) {
    if constexpr (GEN)
        acc_add(a, mul(x, T[slot - 1]));
    else
        acc_mad(a, x, slot);
}

// lambda (x - x^2) for one boolean-constrained cell
template <bool GEN>
__device__ __forceinline__ void boolean(Sums &S, uint64_t x, uint32_t lam, const uint64_t *T) {
    wmad<GEN>(S.pos, x, lam, T);
    wmad<GEN>(S.neg, mul(x, x), lam, T);
}

// sum_k lambda_{s0 + k + 1} p_k over the 10 limbs of p (a constant term of
// the sub and freeze chains)
template <bool GEN>
__device__ __forceinline__ uint64_t p_limbs_weighted(uint32_t s0, const uint64_t *T) {
    if constexpr (GEN) {
        uint64_t v = 0;
#pragma unroll
        for (int k = 0; k < NL; k++) v = add(v, mul((uint64_t)PL[k], T[s0 + k]));
        return v;
    } else {
        return (uint64_t)(s0 + 1) * SP + SP1;
    }
}

// range check of one element block at slot s: 260 booleanity slots, then
// 10 limb packings limb_i - sum_b 2^b bit_ib
template <bool GEN>
__device__ __forceinline__ void range(Sums &S, const uint64_t *col, uint64_t n0, uint64_t pt, uint32_t lb,
                                      uint32_t bb, uint32_t s, const uint64_t *T) {
#pragma unroll 1
    for (int i = 0; i < NL; i++) {
        Acc pk{0, 0};
#pragma unroll 8
        for (int b = 0; b < LB; b++) {
            const uint64_t x = col[(uint64_t)(bb + LB * i + b) * n0 + pt];
            boolean<GEN>(S, x, s + LB * i + b + 1, T);
            acc_shl(pk, x, b);
        }
        wmad<GEN>(S.pos, col[(uint64_t)(lb + i) * n0 + pt], s + EC - NL + i + 1, T);
        wmad<GEN>(S.neg, acc_val(pk), s + EC - NL + i + 1, T);
    }
}

// signed carries (pos_k, neg_k) of a sub / freeze chain whose identities sit
// at slots s0 + k: net_m = pos_m - neg_m enters as carry-in of m + 1 and as
// -2^26 carry-out of m; then booleanity of pos and neg, exclusion, at s1
template <bool GEN>
__device__ __forceinline__ void signed_chain(Sums &S, const uint64_t *col, uint64_t n0, uint64_t pt, uint32_t pos,
                                             uint32_t neg, uint32_t s0, uint32_t s1, const uint64_t *T) {
#pragma unroll 1
    for (int m = 0; m < NL; m++) {
        const uint64_t cp = col[(uint64_t)(pos + m) * n0 + pt], cn = col[(uint64_t)(neg + m) * n0 + pt];
        if (m + 1 < NL) {
            wmad<GEN>(S.pos, cp, s0 + m + 2, T);
            wmad<GEN>(S.neg, cn, s0 + m + 2, T);
        }
        wmad<GEN>(S.rpos, cp, s0 + m + 1, T);
        wmad<GEN>(S.rneg, cn, s0 + m + 1, T);
        boolean<GEN>(S, cp, s1 + m + 1, T);
        boolean<GEN>(S, cn, s1 + NL + m + 1, T);
        wmad<GEN>(S.pos, mul(cp, cn), s1 + 2 * NL + m + 1, T);
    }
}

// phi partial sums of one sub-gadget (blockIdx.y < ng) or of the boundary and
// transitions (blockIdx.y == ng, launched when grid.y = ng + 1) at 64 points,
// proof blockIdx.z whose wstride columns start at lde + k * wstride * n0:
// part[(k * gridDim.y + g) * n0 + pt]. With odd set (the gadget sums on a
// blowup-2 LDE) the even points are rows pt / 2 of lde (the traces, n0 / 2
// rows) and the odd points rows of odd.
// GEN: slot j weighted by T[j] (the caller's combination coefficients)
// instead of j + 1, with the position identities and the p-limb constants
// summed term by term (their affine-in-slot regrouping needs lambda_j = j + 1).
template <bool GEN>
__global__ __launch_bounds__(64) void compose_p256_k(const uint64_t *lde, const uint64_t *odd, uint64_t wstride,
                                                    uint64_t n0, uint64_t blowup, const GDesc *gd, uint32_t ng,
                                                    TailDesc td, uint64_t *part, const uint64_t *T) {
    const uint32_t g = blockIdx.y;
    const uint64_t k = blockIdx.z;
    const uint64_t pt = (uint64_t)blockIdx.x * 64 + threadIdx.x;
    const uint64_t rows = odd ? n0 / 2 : n0, at = odd ? pt >> 1 : pt;
    const uint64_t *col = (odd && (pt & 1) ? odd : lde) + k * wstride * rows;
#define AT(c) col[(uint64_t)(c) * rows + at]
    Sums S;
    if (g == ng) {
        const uint64_t np = (pt + blowup) & (n0 - 1), n_trace = n0 / blowup, trow = pt / blowup;
        if (trow + 1 == n_trace) {
#pragma unroll 1
            for (int j = 0; j < 6; j++)
                for (int l = 0; l < NL; l++) {
                    const uint32_t lam = td.s + NL * j + l + 1;
                    wmad<GEN>(S.pos, AT(td.selc[j] + l), lam, T);
                    wmad<GEN>(S.neg, AT(td.rproj[j] + l), lam, T);
                }
        } else {
#pragma unroll 1
            for (int j = 0; j < 6; j++)
                for (int l = 0; l < NL; l++) {
                    const uint32_t lam = td.s + 6 * NL + NL * j + l + 1, lam2 = lam + 6 * NL;
                    wmad<GEN>(S.pos, col[(uint64_t)(td.acc[j] + l) * n0 + np], lam, T);
                    wmad<GEN>(S.neg, AT(td.selc[j] + l), lam, T);
                    wmad<GEN>(S.pos, col[(uint64_t)(td.rproj[j] + l) * n0 + np], lam2, T);
                    wmad<GEN>(S.neg, AT(td.rproj[j] + l), lam2, T);
                }
        }
    } else {
        const GDesc d = gd[g];
        const uint32_t s = d.s;
        switch (d.kind) {
        case GM: {  // eval_mul_gadget (p256_field_air.rs:1044-1123)
            range<GEN>(S, col, rows, at, d.c, d.cb, s, T);
            range<GEN>(S, col, rows, at, d.q, d.qb, s + EC, T);
            const uint32_t L0 = s + 2 * EC + 1;  // lambda of position 0
            if constexpr (GEN) {
                // position k: lambda_k (sum_{i+j=k} (a_i b_j - q_i p_j) - c_k)
#pragma unroll 1
                for (int kk = 0; kk < MSP; kk++) {
                    uint64_t ab = 0, qp = 0;
#pragma unroll
                    for (int j = 0; j < NL; j++) {
                        const int i = kk - j;
                        if (i < 0 || i >= NL) continue;
                        ab = add(ab, mul(AT(d.a + i), AT(d.b + j)));
                        qp = add(qp, mul(AT(d.q + i), (uint64_t)PL[j]));  // j unrolled: PL[j] a constant
                    }
                    wmad<GEN>(S.pos, ab, L0 + kk, T);
                    wmad<GEN>(S.neg, qp, L0 + kk, T);
                    if (kk < NL) wmad<GEN>(S.neg, AT(d.c + kk), L0 + kk, T);
                }
            } else {
                uint64_t B = 0, B1 = 0;
#pragma unroll
                for (int j = 0; j < NL; j++) {
                    const uint64_t bj = AT(d.b + j);
                    B = add(B, bj);
                    B1 = add(B1, mul_small(bj, j));
                }
#pragma unroll 1
                for (int i = 0; i < NL; i++) {
                    acc_add(S.pos, mul(AT(d.a + i), add(mul_small(B, L0 + i), B1)));
                    // sum_j lambda_{i+j} q_i p_j = q_i ((L0 + i) SP + SP1), below 2^50
                    acc_add(S.neg, mul(AT(d.q + i), (uint64_t)(L0 + i) * SP + SP1));
                    acc_mad(S.neg, AT(d.c + i), L0 + i);
                }
            }
#pragma unroll 1
            for (int m = 0; m < MCP; m++) {
                Acc sc{0, 0};
#pragma unroll 8
                for (int b = 0; b < MCB; b++) {
                    const uint64_t y = AT(d.cy + MCB * m + b);
                    boolean<GEN>(S, y, s + 2 * EC + MSP + MCB * m + b + 1, T);
                    acc_shl(sc, y, b);
                }
                // the signed carry of position m: carry-in of m + 1, -2^26 carry-out of m
                uint64_t gamma;
                if constexpr (GEN)
                    gamma = sub(T[L0 + m], mul(T[L0 + m - 1], 1ULL << LB));
                else
                    gamma = sub(L0 + m + 1, mul((uint64_t)(L0 + m), 1ULL << LB));
                acc_add(S.pos, mul(sub(acc_val(sc), (uint64_t)MC_OFF), gamma));
            }
            break;
        }
        case GA: {  // eval_add_gadget (:404-455)
            range<GEN>(S, col, rows, at, d.c, d.cb, s, T);
#pragma unroll 1
            for (int kk = 0; kk < NL; kk++) {
                const uint32_t lam = s + EC + kk + 1;
                const uint64_t cy = AT(d.pos + kk);
                wmad<GEN>(S.pos, AT(d.a + kk), lam, T);
                wmad<GEN>(S.pos, AT(d.b + kk), lam, T);
                wmad<GEN>(S.neg, AT(d.c + kk), lam, T);
                if (kk + 1 < NL) {
                    wmad<GEN>(S.pos, cy, lam + 1, T);             // carry-in of kk + 1
                    boolean<GEN>(S, cy, s + EC + NL + kk + 1, T);     // booleanity, kk < 9
                } else {
                    wmad<GEN>(S.pos, cy, s + EC + 2 * NL, T);      // carry[9] = 0
                }
                wmad<GEN>(S.rpos, cy, lam, T);
            }
            break;
        }
        case GS: {  // eval_sub_gadget (:636-698)
            range<GEN>(S, col, rows, at, d.c, d.cb, s, T);
            const uint32_t s0 = s + EC;
            acc_add(S.pos, p_limbs_weighted<GEN>(s0, T));
#pragma unroll 1
            for (int kk = 0; kk < NL; kk++) {
                wmad<GEN>(S.pos, AT(d.a + kk), s0 + kk + 1, T);
                wmad<GEN>(S.neg, AT(d.b + kk), s0 + kk + 1, T);
                wmad<GEN>(S.neg, AT(d.c + kk), s0 + kk + 1, T);
            }
            signed_chain<GEN>(S, col, rows, at, d.pos, d.neg, s0, s0 + NL, T);
            const uint32_t lam = s + EC + 4 * NL + 1;
            wmad<GEN>(S.pos, AT(d.pos + NL - 1), lam, T);
            wmad<GEN>(S.pos, AT(d.neg + NL - 1), lam, T);
            break;
        }
        case GF: {  // eval_freeze_gadget (:1449-1531)
            range<GEN>(S, col, rows, at, d.q, d.qb, s, T);
            range<GEN>(S, col, rows, at, d.c, d.cb, s + EC, T);
            const uint32_t s0 = s + 2 * EC;
            acc_add(S.neg, p_limbs_weighted<GEN>(s0, T));
            Acc M{0, 0};
            const uint32_t sm = s0 + 4 * NL + 1;  // mux slots
#pragma unroll 1
            for (int kk = 0; kk < NL; kk++) {
                const uint64_t a = AT(d.a + kk), dk = AT(d.q + kk);
                wmad<GEN>(S.pos, a, s0 + kk + 1, T);
                wmad<GEN>(S.neg, dk, s0 + kk + 1, T);
                wmad<GEN>(S.pos, AT(d.c + kk), sm + kk + 1, T);
                wmad<GEN>(S.neg, dk, sm + kk + 1, T);
                wmad<GEN>(M, sub(a, dk), sm + kk + 1, T);
            }
            signed_chain<GEN>(S, col, rows, at, d.pos, d.neg, s0, s0 + NL, T);
            wmad<GEN>(S.pos, AT(d.pos + NL - 1), s0 + 4 * NL + 1, T);
            acc_add(S.neg, mul(AT(d.neg + NL - 1), acc_val(M)));
            break;
        }
        default: {  // eval_select_gadget (:1243-1276)
            range<GEN>(S, col, rows, at, d.c, d.cb, s, T);
            Acc M{0, 0};
#pragma unroll 1
            for (int kk = 0; kk < NL; kk++) {
                const uint32_t lam = s + EC + kk + 1;
                const uint64_t bk = AT(d.b + kk);
                wmad<GEN>(S.pos, AT(d.c + kk), lam, T);
                wmad<GEN>(S.neg, bk, lam, T);
                wmad<GEN>(M, sub(AT(d.a + kk), bk), lam, T);
            }
            const uint64_t sel = AT(d.sel);
            acc_add(S.neg, mul(sel, acc_val(M)));
            boolean<GEN>(S, sel, s + EC + NL + 1, T);
        }
        }
    }
#undef AT
    part[(k * gridDim.y + g) * n0 + pt] = S.value();
}

// phi[k][pt] = sum over the ng + 1 partial sums
__global__ __launch_bounds__(256) void p256_sum_k(const uint64_t *part, uint32_t nparts, uint64_t n0,
                                                 uint64_t *phi) {
    const uint64_t k = blockIdx.y, pt = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (pt >= n0) return;
    const uint64_t *p = part + k * nparts * n0 + pt;
    Acc a{0, 0};
    for (uint32_t g = 0; g < nparts; g++) acc_add(a, p[(uint64_t)g * n0]);
    phi[k * n0 + pt] = acc_val(a);
}

// ---------------------------------------------------------------- ECDSA (host)

// F_p / Z_n arithmetic on 8 little-endian words
void mulmod(const uint32_t *a, const uint32_t *b, const uint32_t *m, uint32_t *r) {
    uint32_t prod[16] = {};
    for (int i = 0; i < 8; i++) {
        uint64_t carry = 0;
        for (int j = 0; j < 8; j++) {
            const uint64_t t = (uint64_t)a[i] * b[j] + prod[i + j] + carry;
            prod[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        prod[i + 8] = (uint32_t)carry;
    }
    if (m == P_W) {
        p256_reduce(prod, r);
        return;
    }
    int pl = 16;
    while (pl > 1 && prod[pl - 1] == 0) pl--;
    uint32_t q[17], rr[8] = {};
    divmod_words<8>(prod, pl, m, 8, q, rr);
    memcpy(r, rr, 32);
}
int cmp8(const uint32_t *a, const uint32_t *b) {
    for (int i = 7; i >= 0; i--)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}
bool zero8(const uint32_t *a) {
    for (int i = 0; i < 8; i++)
        if (a[i]) return false;
    return true;
}
void addmod(const uint32_t *a, const uint32_t *b, const uint32_t *m, uint32_t *r) {
    uint32_t t[8];
    uint64_t c = 0;
    for (int i = 0; i < 8; i++) {
        c += (uint64_t)a[i] + b[i];
        t[i] = (uint32_t)c;
        c >>= 32;
    }
    if (c || cmp8(t, m) >= 0) {
        uint64_t br = 0;
        for (int i = 0; i < 8; i++) {
            const uint64_t d = (uint64_t)t[i] - m[i] - br;
            t[i] = (uint32_t)d;
            br = (d >> 63) & 1;
        }
    }
    memcpy(r, t, 32);
}
void submod(const uint32_t *a, const uint32_t *b, const uint32_t *m, uint32_t *r) {
    uint32_t nb[8];
    uint64_t br = 0;
    for (int i = 0; i < 8; i++) {
        const uint64_t d = (uint64_t)m[i] - b[i] - br;
        nb[i] = (uint32_t)d;
        br = (d >> 63) & 1;
    }
    if (zero8(b)) memset(nb, 0, 32);
    addmod(a, nb, m, r);
}
// a^(m-2) mod m, m prime
void invmod(const uint32_t *a, const uint32_t *m, uint32_t *r) {
    uint32_t e[8], acc[8] = {1};
    memcpy(e, m, 32);
    e[0] -= 2;
    for (int bit = 255; bit >= 0; bit--) {
        mulmod(acc, acc, m, acc);
        if ((e[bit / 32] >> (bit % 32)) & 1) mulmod(acc, a, m, acc);
    }
    memcpy(r, acc, 32);
}

struct Pt {
    uint32_t x[8], y[8], z[8];
};
// complete projective addition (Renes-Costello-Batina 2016, alg. 4, a = -3);
// also doubles
void pt_add(const Pt &P, const Pt &Q, const uint32_t *bc, Pt &R) {
    const uint32_t *m = P_W;
    uint32_t t0[8], t1[8], t2[8], t3[8], t4[8], x3[8], y3[8], z3[8];
    mulmod(P.x, Q.x, m, t0);
    mulmod(P.y, Q.y, m, t1);
    mulmod(P.z, Q.z, m, t2);
    addmod(P.x, P.y, m, t3);
    addmod(Q.x, Q.y, m, t4);
    mulmod(t3, t4, m, t3);
    addmod(t0, t1, m, t4);
    submod(t3, t4, m, t3);
    addmod(P.y, P.z, m, t4);
    addmod(Q.y, Q.z, m, x3);
    mulmod(t4, x3, m, t4);
    addmod(t1, t2, m, x3);
    submod(t4, x3, m, t4);
    addmod(P.x, P.z, m, x3);
    addmod(Q.x, Q.z, m, y3);
    mulmod(x3, y3, m, x3);
    addmod(t0, t2, m, y3);
    submod(x3, y3, m, y3);
    mulmod(bc, t2, m, z3);
    submod(y3, z3, m, x3);
    addmod(x3, x3, m, z3);
    addmod(x3, z3, m, x3);
    submod(t1, x3, m, z3);
    addmod(t1, x3, m, x3);
    mulmod(bc, y3, m, y3);
    addmod(t2, t2, m, t1);
    addmod(t1, t2, m, t2);
    submod(y3, t2, m, y3);
    submod(y3, t0, m, y3);
    addmod(y3, y3, m, t1);
    addmod(t1, y3, m, y3);
    addmod(t0, t0, m, t1);
    addmod(t1, t0, m, t0);
    submod(t0, t2, m, t0);
    mulmod(t4, y3, m, t1);
    mulmod(t0, y3, m, t2);
    mulmod(x3, z3, m, y3);
    addmod(y3, t2, m, y3);
    mulmod(t3, x3, m, x3);
    submod(x3, t1, m, x3);
    mulmod(t4, z3, m, z3);
    mulmod(t3, t0, m, t1);
    addmod(z3, t1, m, z3);
    memcpy(R.x, x3, 32);
    memcpy(R.y, y3, 32);
    memcpy(R.z, z3, 32);
}
// R = k1 P1 + k2 P2 by one double-and-add ladder over both scalars
// (Shamir's trick: 256 doublings and an addition of P1, P2 or P1 + P2 per
// nonzero bit pair); the complete formula covers every case
void pt_mul2(const Pt &P1, const uint32_t *k1, const Pt &P2, const uint32_t *k2, const uint32_t *bc, Pt &R) {
    Pt both;
    pt_add(P1, P2, bc, both);
    Pt acc{};
    acc.y[0] = 1;  // the identity (0 : 1 : 0)
    for (int bit = 255; bit >= 0; bit--) {
        pt_add(acc, acc, bc, acc);
        const int b1 = (k1[bit / 32] >> (bit % 32)) & 1, b2 = (k2[bit / 32] >> (bit % 32)) & 1;
        if (b1 && b2)
            pt_add(acc, both, bc, acc);
        else if (b1)
            pt_add(acc, P1, bc, acc);
        else if (b2)
            pt_add(acc, P2, bc, acc);
    }
    R = acc;
}

}  // namespace p256

using namespace p256;

void p256_pk_hash_pre(const P256Record &rec, uint8_t out[32]) {
    Sha3 h;
    h.update("ECDSAP256-MIXEDZONE", 19);
    h.update(rec.qx.data(), 32);
    h.update(rec.qy.data(), 32);
    h.update(rec.r.data(), 32);
    h.update(rec.s.data(), 32);
    h.update(rec.digest.data(), 32);
    h.final(out);
}

P256Chains p256_ecdsa_chains(const P256Record &rec, bool verify) {
    uint32_t r[8], s[8], qx[8], qy[8], e[8], bc[8];
    be_to_words(rec.r.data(), r);
    be_to_words(rec.s.data(), s);
    be_to_words(rec.qx.data(), qx);
    be_to_words(rec.qy.data(), qy);
    be_to_words(B_BE, bc);
    SG_REQUIRE(!zero8(r) && !zero8(s) && cmp8(r, N_W) < 0 && cmp8(s, N_W) < 0, "ecdsa: r and s must be in [1, n)");
    SG_REQUIRE(cmp8(qx, P_W) < 0 && cmp8(qy, P_W) < 0, "ecdsa: public key coordinates must be below p");
    {  // Q on y^2 = x^3 - 3x + b
        uint32_t l[8], t[8];
        mulmod(qy, qy, P_W, l);
        mulmod(qx, qx, P_W, t);
        mulmod(t, qx, P_W, t);
        for (int i = 0; i < 3; i++) submod(t, qx, P_W, t);
        addmod(t, bc, P_W, t);
        SG_REQUIRE(cmp8(l, t) == 0, "ecdsa: public key is not a curve point");
    }
    {  // e = digest mod n (one subtraction suffices: 2^256 < 2n)
        be_to_words(rec.digest.data(), e);
        if (cmp8(e, N_W) >= 0) {
            uint64_t br = 0;
            for (int i = 0; i < 8; i++) {
                const uint64_t d = (uint64_t)e[i] - N_W[i] - br;
                e[i] = (uint32_t)d;
                br = (d >> 63) & 1;
            }
        }
    }
    uint32_t w[8], u1[8], u2[8];
    invmod(s, N_W, w);
    mulmod(e, w, N_W, u1);
    mulmod(r, w, N_W, u2);
    if (verify) {
        Pt G{}, Q{}, R;
        be_to_words(GX_BE, G.x);
        be_to_words(GY_BE, G.y);
        G.z[0] = 1;
        memcpy(Q.x, qx, 32);
        memcpy(Q.y, qy, 32);
        Q.z[0] = 1;
        pt_mul2(G, u1, Q, u2, bc, R);
        SG_REQUIRE(!zero8(R.z), "ecdsa: signature does not verify");
        uint32_t zi[8], x[8];
        invmod(R.z, P_W, zi);
        mulmod(R.x, zi, P_W, x);
        if (cmp8(x, N_W) >= 0) submod(x, N_W, P_W, x);  // x < p < 2n
        SG_REQUIRE(cmp8(x, r) == 0, "ecdsa: signature does not verify");
    }
    P256Chains c;
    Be32 one{};
    one[31] = 1;
    for (int h = 0; h < 2; h++) {
        memcpy(c.pts[3 * h].data(), GX_BE, 32);
        memcpy(c.pts[3 * h + 1].data(), GY_BE, 32);
        c.pts[3 * h + 2] = one;
        c.pts[6 + 3 * h] = rec.qx;
        c.pts[6 + 3 * h + 1] = rec.qy;
        c.pts[6 + 3 * h + 2] = one;
    }
    uint8_t b1[32], b2[32];
    words_to_be(u1, b1);
    words_to_be(u2, b2);
    c.a_bits.resize(256);
    c.b_bits.resize(256);
    for (int i = 0; i < 256; i++) {
        c.a_bits[i] = (b1[i / 8] >> (7 - i % 8)) & 1;
        c.b_bits[i] = (b2[i / 8] >> (7 - i % 8)) & 1;
    }
    c.bind_outputs = true;
    return c;
}

void build_p256_trace_host(const P256Chains &in, uint64_t h, uint64_t *tr, std::array<Be32, 6> *rproj_out) {
    check_inputs({in}, h);
    const Layout &L = layout();
    std::vector<uint64_t> words((size_t)h * L.n_words, 0);
    std::vector<std::array<Be32, 6>> rp;
    batch_words({in}, h, words.data(), rp);
    if (rproj_out) *rproj_out = rp[0];
    parallel_for(16, [&](size_t part) {
        for (uint64_t col = part; col < P256_W; col += 16)
            for (uint64_t r = 0; r < h; r++) tr[col * h + r] = cell(L.desc[col], words.data() + r * L.n_words);
    });
}

void build_p256_traces_dev(Ctx *c, const std::vector<P256Chains> &in, uint64_t h, uint64_t *d_tr,
                           std::vector<std::array<Be32, 6>> *rproj_out, uint64_t *d_odd) {
    const size_t K = in.size();
    SG_REQUIRE(K >= 1 && K <= 65535, "p256 traces: 1 to 65535 proofs");
    SG_REQUIRE(!d_odd || (h >= 2 && h <= 4096), "p256 traces: the fused odd coset needs 2 to 4096 rows");
    check_inputs(in, h);
    const Layout &L = layout();
    Scoped d_desc(c, 4ULL * P256_W);
    SG_HIP(hipMemcpyAsync(d_desc.p, L.desc.data(), 4ULL * P256_W, hipMemcpyHostToDevice, c->stream));
    const uint64_t nw = L.n_words, per = h * nw;
    // every word is written below, straight into pinned memory
    uint64_t *words = (uint64_t *)c->upload_staging(8 * K * per);
    std::vector<std::array<Be32, 6>> rp;
    {
        HostScope hs(c, "p256_words");
        batch_words(in, h, words, rp);
    }
    if (rproj_out) *rproj_out = rp;
    Scoped d_words(c, 8 * K * per);
    SG_HIP(hipMemcpyAsync(d_words.p, words, 8 * K * per, hipMemcpyHostToDevice, c->stream));
    if (d_odd) {
        const uint32_t lh = ilog2(h);
        const uint64_t *twist = odd_twist_table(c, lh);
        const dim3 grid((uint32_t)((P256_W + (TILE >> lh) - 1) >> (TILE_LOG - lh)), (uint32_t)K);
        ProfScope ps(c, "expand_odd_p256", 24.0 * P256_W * h * K);
        switch (lh) {
#define SG_CASE(LH)                                                                                                  \
    case LH:                                                                                                         \
        hipLaunchKernelGGL((expand_odd_p256_k<LH>), grid, dim3(NT), 0, c->stream, d_words.as<uint64_t>(),            \
                           d_desc.as<uint32_t>(), (uint32_t)nw, d_tr, d_odd, twist, c->tw_f, c->tw_i);               \
        break;
            SG_CASE(1) SG_CASE(2) SG_CASE(3) SG_CASE(4) SG_CASE(5) SG_CASE(6)
            SG_CASE(7) SG_CASE(8) SG_CASE(9) SG_CASE(10) SG_CASE(11) SG_CASE(12)
#undef SG_CASE
        default: throw Error(E_INTERNAL, "p256 traces: height out of range");
        }
        SG_HIP(hipGetLastError());
    } else {
        ProfScope ps(c, "expand_p256", 8.0 * P256_W * h * K);
        hipLaunchKernelGGL(expand_p256_k, dim3((uint32_t)((h * P256_W + 255) / 256), (uint32_t)K), dim3(256), 0,
                           c->stream, d_words.as<uint64_t>(), d_desc.as<uint32_t>(), (uint32_t)nw, ilog2(h), d_tr);
        SG_HIP(hipGetLastError());
    }
    SG_HIP(hipStreamSynchronize(c->stream));  // the staging buffer is reused by the next upload
}

namespace p256 {

// the 180 columns of the boundary and transitions, in the order of a compact
// tail trace: select outputs (6 x 10), r_proj (6 x 10), accumulators (6 x 10)
void tail_columns(const Layout &L, uint32_t *cols, TailDesc &compact) {
    compact.s = L.tail.s;
    for (int j = 0; j < 6; j++) {
        compact.selc[j] = NL * j;
        compact.rproj[j] = 6 * NL + NL * j;
        compact.acc[j] = 12 * NL + NL * j;
        for (int l = 0; l < NL; l++) {
            cols[NL * j + l] = L.tail.selc[j] + l;
            cols[6 * NL + NL * j + l] = L.tail.rproj[j] + l;
            cols[12 * NL + NL * j + l] = L.tail.acc[j] + l;
        }
    }
}

// out[k][t][r] = tr[k][cols[t]][r] for the K traces of h rows
__global__ __launch_bounds__(256) void gather_cols_k(const uint64_t *tr, const uint32_t *cols, uint32_t nc,
                                                     uint32_t log_h, uint64_t *out) {
    const uint64_t k = blockIdx.y, h = 1ULL << log_h;
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (uint64_t)nc * h) return;
    const uint64_t c = t >> log_h, r = t & (h - 1);
    out[k * nc * h + t] = tr[(k * P256_W + cols[c]) * h + r];
}

__global__ __launch_bounds__(256) void add_into_k(uint64_t *a, const uint64_t *b, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] = add(a[i], b[i]);
}

}  // namespace p256

void merge_p256_from_trace_batch(Ctx *c, const uint64_t *d_tr, uint32_t K, uint64_t n_trace, uint64_t blowup,
                                 uint64_t *c_eval, const uint64_t *d_odd) {
    SG_REQUIRE(is_pow2(n_trace) && n_trace >= 32 && is_pow2(blowup) && blowup >= 2,
               "merge_p256: n_trace must be a power of two >= 32 and blowup >= 2");
    SG_REQUIRE(K >= 1 && K <= 65535, "merge_p256: 1 to 65535 proofs");
    const uint64_t n = n_trace * blowup, n2 = 2 * n_trace;
    SG_REQUIRE(n <= (1ULL << 20), "merge_p256: domain above 2^20");
    const Layout &L = layout();
    const uint32_t ng = (uint32_t)L.gd.size(), NT = 18 * NL;
    Scoped phi(c, 8 * n * K);
    {
        // the gadget constraints at the 2 n_trace points of <w_{2 n_trace}>
        // (LDE points i = (blowup / 2) t), then interpolated to all n points
        Scoped d_gd(c, sizeof(GDesc) * ng), phi2(c, 8 * n2 * K);
        SG_HIP(hipMemcpyAsync(d_gd.p, L.gd.data(), sizeof(GDesc) * ng, hipMemcpyHostToDevice, c->stream));
        {
            // the blowup-2 LDE as the traces (its even points) and the odd coset
            Scoped odd(c, d_odd ? 0 : 8ULL * P256_W * n_trace * K), part(c, 8ULL * ng * n2 * K);
            if (!d_odd) lde_odd_coset(c, d_tr, odd.as<uint64_t>(), ilog2(n_trace), (uint64_t)P256_W * K);
            const uint64_t *od = d_odd ? d_odd : odd.as<uint64_t>();
            {
                ProfScope ps(c, "compose_p256", (8.0 * P256_W * n2 + 8.0 * ng * n2) * K);
                hipLaunchKernelGGL(compose_p256_k<false>, dim3((uint32_t)(n2 / 64), ng, K), dim3(64), 0, c->stream,
                                   d_tr, od, (uint64_t)P256_W, n2, (uint64_t)2, d_gd.as<GDesc>(), ng, L.tail,
                                   part.as<uint64_t>(), (const uint64_t *)nullptr);
                SG_HIP(hipGetLastError());
            }
            {
                ProfScope ps(c, "p256_sum", 8.0 * (ng + 1) * n2 * K);
                hipLaunchKernelGGL(p256_sum_k, dim3((uint32_t)((n2 + 255) / 256), K), dim3(256), 0, c->stream,
                                   part.as<uint64_t>(), ng, n2, phi2.as<uint64_t>());
                SG_HIP(hipGetLastError());
            }
        }
        lde_columns(c, phi2.as<uint64_t>(), phi.as<uint64_t>(), ilog2(n2), ilog2(blowup / 2), K, n2, n, 1);
    }
    {
        // boundary and transitions at every point, from the LDE of their 180 columns
        uint32_t cols[18 * NL];
        TailDesc td;
        tail_columns(L, cols, td);
        Scoped d_cols(c, 4 * NT), ttr(c, 8ULL * NT * n_trace * K), tlde(c, 8ULL * NT * n * K), tphi(c, 8 * n * K);
        SG_HIP(hipMemcpyAsync(d_cols.p, cols, 4 * NT, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(gather_cols_k, dim3((uint32_t)((NT * n_trace + 255) / 256), K), dim3(256), 0, c->stream,
                           d_tr, d_cols.as<uint32_t>(), NT, ilog2(n_trace), ttr.as<uint64_t>());
        SG_HIP(hipGetLastError());
        lde_columns(c, ttr.as<uint64_t>(), tlde.as<uint64_t>(), ilog2(n_trace), ilog2(blowup), (uint64_t)NT * K,
                    n_trace, n, 1);
        {
            ProfScope ps(c, "compose_p256_tail", (8.0 * NT * n + 8.0 * n) * K);
            hipLaunchKernelGGL(compose_p256_k<false>, dim3((uint32_t)(n / 64), 1, K), dim3(64), 0, c->stream,
                               tlde.as<uint64_t>(), (const uint64_t *)nullptr, (uint64_t)NT, n, blowup,
                               (const GDesc *)nullptr, 0u, td, tphi.as<uint64_t>(), (const uint64_t *)nullptr);
            SG_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(add_into_k, dim3((uint32_t)((n * K + 255) / 256)), dim3(256), 0, c->stream,
                           phi.as<uint64_t>(), tphi.as<uint64_t>(), n * K);
        SG_HIP(hipGetLastError());
    }
    quotient_batch(c, phi.as<uint64_t>(), K, n_trace, n, c_eval);
}

void merge_p256_batch(Ctx *c, const uint64_t *lde, uint32_t K, uint64_t n_trace, uint64_t blowup, uint64_t *c_eval,
                      const uint64_t *d_lam) {
    SG_REQUIRE(is_pow2(n_trace) && n_trace >= 2 && is_pow2(blowup) && blowup >= 2,
               "merge_p256: n_trace and blowup must be powers of two >= 2");
    SG_REQUIRE(K >= 1 && K <= 65535, "merge_p256: 1 to 65535 proofs");
    const uint64_t n = n_trace * blowup;
    SG_REQUIRE(n >= 64 && n <= (1ULL << 20), "merge_p256: domain must be in [64, 2^20]");
    const Layout &L = layout();
    const uint32_t ng = (uint32_t)L.gd.size();
    Scoped d_gd(c, sizeof(GDesc) * ng);
    SG_HIP(hipMemcpyAsync(d_gd.p, L.gd.data(), sizeof(GDesc) * ng, hipMemcpyHostToDevice, c->stream));
    Scoped part(c, 8ULL * (ng + 1) * n * K), phi(c, 8 * n * K);
    {
        ProfScope ps(c, "compose_p256", (8.0 * P256_W * n + 8.0 * (ng + 1) * n) * K);
        if (d_lam)
            hipLaunchKernelGGL(compose_p256_k<true>, dim3((uint32_t)(n / 64), ng + 1, K), dim3(64), 0, c->stream,
                               lde, (const uint64_t *)nullptr, (uint64_t)P256_W, n, blowup, d_gd.as<GDesc>(), ng,
                               L.tail, part.as<uint64_t>(), d_lam);
        else
            hipLaunchKernelGGL(compose_p256_k<false>, dim3((uint32_t)(n / 64), ng + 1, K), dim3(64), 0, c->stream,
                               lde, (const uint64_t *)nullptr, (uint64_t)P256_W, n, blowup, d_gd.as<GDesc>(), ng,
                               L.tail, part.as<uint64_t>(), (const uint64_t *)nullptr);
        SG_HIP(hipGetLastError());
    }
    {
        ProfScope ps(c, "p256_sum", 8.0 * (ng + 2) * n * K);
        hipLaunchKernelGGL(p256_sum_k, dim3((uint32_t)((n + 255) / 256), K), dim3(256), 0, c->stream,
                           part.as<uint64_t>(), ng + 1, n, phi.as<uint64_t>());
        SG_HIP(hipGetLastError());
    }
    quotient_batch(c, phi.as<uint64_t>(), K, n_trace, n, c_eval);
}

}  // namespace sg
