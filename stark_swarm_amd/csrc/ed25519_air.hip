// ed25519_air.hip -- the Ed25519 cofactored-verify AIR, v16, for gfx950.
//
// Trace (crates/deep_ali/src/ed25519_verify_air.rs:4820-4865 and the gadget
// files it composes): SHA-512 of R || A || M on the first k_hash rows, then
// the scalar reduction mod L, the decompression of R and of A, two
// k_scalar-row double-and-add ladders ([s]B, [k]A), the residual
// [s]B - R - [k]A, three doublings and the identity verdict; 41 519 columns
// and 70 606 constraints per row at k_scalar = 256. A field element of
// F_{2^255-19} is 10 limbs of 26/25 bits; most cells are the bits of gadget
// output limbs and of the 36-bit biased carries of the mul gadgets.
//
// Host: the native values of every row (ladder points, gadget outputs,
// carries) become per-row "witness words" (about 900 per row against 41 519
// cells). A per-row-type column descriptor (word index and how the cell
// derives from it: a bit, the value, a carry's sign, a per-proof constant)
// lets `expand_ed_k` write the cells, rows fastest; the SHA-512 rows come
// from the SHA-512 AIR's own expansion.
//
// Composition (eval_verify_air_v16_per_row): the constraints form a table of
// blocks -- a mul / add / sub / select gadget, a run of pins, bindings,
// booleanities, the scalar-reduce and digest-format blocks -- each with the
// slot of its first constraint and the trace rows where it is live (one row
// or a range). `compose_ed_k` runs one 64-lane wave per (64 points, block);
// the lane is the point, so every column read is a coalesced 512-byte run,
// and each constraint value is weighted by lambda[slot] (a coefficient table,
// or j + 1) into a 192-bit lazy sum. Blocks that the two ladders (or the two
// decompressions) share are evaluated once with two weights. In the prover
// (GROUPS) the points are a blowup-2 LDE and every block is summed into the
// polynomial of its row gate; those polynomials are interpolated to the full
// domain and `combine_ed_k` adds the ones whose gate holds at each point's row.
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>

#include "air.hpp"
#include "ed25519_air.hpp"
#include "gl.hpp"
#include "keccak.hpp"
#include "ntt.hpp"
#include "sha256_air.hpp"

namespace sg {

namespace ed {

constexpr int NL = 10, EB = 255, EC = 265, MCB = 36;
constexpr int LW[NL] = {26, 25, 26, 25, 26, 25, 26, 25, 26, 25};
constexpr int LOFF[NL] = {0, 26, 51, 77, 102, 128, 153, 179, 204, 230};
constexpr uint32_t SHA_W = 1510, SHA_K = 1526, OFF_H0 = 1066;
constexpr uint32_t IN_L = 32, R_L = 16, Q_L = 17, LBITS = 16, P_L = 33, CBITS = 24;

// ---------------------------------------------------------------- F_{2^255-19} (host)

struct Fe {
    uint64_t v[4];
};
constexpr uint64_t PF[4] = {0xFFFFFFFFFFFFFFEDULL, ~0ULL, ~0ULL, 0x7FFFFFFFFFFFFFFFULL};
typedef unsigned __int128 u128;

static bool ge4(const uint64_t *a, const uint64_t *b) {
    for (int i = 3; i >= 0; i--)
        if (a[i] != b[i]) return a[i] > b[i];
    return true;
}
static void sub4(uint64_t *a, const uint64_t *b) {
    uint64_t br = 0;
    for (int i = 0; i < 4; i++) {
        const u128 d = (u128)a[i] - b[i] - br;
        a[i] = (uint64_t)d;
        br = (uint64_t)(d >> 127) & 1;
    }
}
static Fe fmul(const Fe &a, const Fe &b) {
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
    // 2^256 = 38 (mod p)
    u128 c = 0;
    Fe r;
    for (int i = 0; i < 4; i++) {
        c += (u128)w[i] + (u128)w[4 + i] * 38;
        r.v[i] = (uint64_t)c;
        c >>= 64;
    }
    while (c) {
        u128 d = c * 38;
        c = 0;
        for (int i = 0; i < 4; i++) {
            d += r.v[i];
            r.v[i] = (uint64_t)d;
            d >>= 64;
        }
        c = d;
    }
    while (ge4(r.v, PF)) sub4(r.v, PF);
    return r;
}
static Fe fadd(const Fe &a, const Fe &b) {
    Fe r;
    u128 c = 0;
    for (int i = 0; i < 4; i++) {
        c += (u128)a.v[i] + b.v[i];
        r.v[i] = (uint64_t)c;
        c >>= 64;
    }
    if (ge4(r.v, PF)) sub4(r.v, PF);
    return r;
}
static Fe fsub(const Fe &a, const Fe &b) {
    Fe r = a;
    if (!ge4(a.v, b.v)) {
        u128 c = 0;
        for (int i = 0; i < 4; i++) {
            c += (u128)r.v[i] + PF[i];
            r.v[i] = (uint64_t)c;
            c >>= 64;
        }
    }
    sub4(r.v, b.v);
    return r;
}
static Fe fzero() { return Fe{{0, 0, 0, 0}}; }
static Fe fone() { return Fe{{1, 0, 0, 0}}; }
static Fe fneg(const Fe &a) { return fsub(fzero(), a); }
static bool feq(const Fe &a, const Fe &b) { return !memcmp(a.v, b.v, 32); }
static bool fis0(const Fe &a) { return !(a.v[0] | a.v[1] | a.v[2] | a.v[3]); }
static Fe fpow(Fe a, const uint64_t e[4]) {
    Fe r = fone();
    for (int i = 255; i >= 0; i--) {
        r = fmul(r, r);
        if ((e[i / 64] >> (i % 64)) & 1) r = fmul(r, a);
    }
    return r;
}
static Fe fsmall(int64_t k) {
    Fe r = fzero();
    if (k >= 0) {
        r.v[0] = (uint64_t)k;
        return r;
    }
    r.v[0] = (uint64_t)(-k);
    return fneg(r);
}
static Fe fbytes(const uint8_t b[32]) {  // bit 255 masked, reduced mod p
    Fe r;
    for (int k = 0; k < 4; k++) {
        r.v[k] = 0;
        for (int i = 0; i < 8; i++) r.v[k] |= (uint64_t)b[8 * k + i] << (8 * i);
    }
    r.v[3] &= 0x7FFFFFFFFFFFFFFFULL;
    if (ge4(r.v, PF)) sub4(r.v, PF);
    return r;
}
static void ftobytes(const Fe &a, uint8_t b[32]) {
    for (int k = 0; k < 4; k++)
        for (int i = 0; i < 8; i++) b[8 * k + i] = (uint8_t)(a.v[k] >> (8 * i));
}

// 10 limbs as the trace holds them, with the field value they represent
// (kept beside the limbs so no gadget re-derives it)
struct Lb {
    std::array<int64_t, NL> l{};
    Fe v{{0, 0, 0, 0}};
    int64_t &operator[](size_t i) { return l[i]; }
    int64_t operator[](size_t i) const { return l[i]; }
};
// tight limbs of a canonical value (a frozen FieldElement, ed25519_field.rs:154-158)
static Lb limbs(const Fe &a) {
    Lb r;
    r.v = a;
    for (int i = 0; i < NL; i++) {
        const int off = LOFF[i], lo = off / 64, sh = off % 64;
        uint64_t x = a.v[lo] >> sh;
        if (sh + LW[i] > 64 && lo + 1 < 4) x |= a.v[lo + 1] << (64 - sh);
        r.l[i] = (int64_t)(x & ((1ULL << LW[i]) - 1));
    }
    return r;
}
static const Fe &of_limbs(const Lb &l) { return l.v; }

struct Pt {
    Fe X, Y, Z, T;
};
struct Consts {
    Fe d, d2, sqrtm1;
    Lb d_l, d2_l;  // 2d as D.add(D): limbs twice d's (ed25519_field.rs:609)
    Pt B;
    Consts();
};
static bool decompress(const Consts &K, const uint8_t in[32], Pt &out);
static const Consts &consts() {
    static const Consts k;
    return k;
}
Consts::Consts() {
    d = fmul(fsmall(-121665), fpow(fsmall(121666), std::array<uint64_t, 4>{0xFFFFFFFFFFFFFFEBULL, ~0ULL, ~0ULL,
                                                                            0x7FFFFFFFFFFFFFFFULL}
                                                         .data()));
    d2 = fadd(d, d);
    const uint64_t e14[4] = {0xFFFFFFFFFFFFFFFBULL, ~0ULL, ~0ULL, 0x1FFFFFFFFFFFFFFFULL};  // (p - 1) / 4
    sqrtm1 = fpow(fsmall(2), e14);
    d_l = limbs(d);
    d2_l.v = d2;
    for (int i = 0; i < NL; i++) d2_l[i] = 2 * d_l[i];
    uint8_t bc[32];
    bc[0] = 0x58;
    for (int i = 1; i < 32; i++) bc[i] = 0x66;
    decompress(*this, bc, B);
}
// RFC 8032 5.1.3 as ed25519_group.rs:157-179 (sqrt_ratio_i, ed25519_field.rs:538-568)
static bool decompress(const Consts &K, const uint8_t in[32], Pt &out) {
    const int sign = (in[31] >> 7) & 1;
    const Fe y = fbytes(in), one = fone(), y2 = fmul(y, y);
    const Fe u = fsub(y2, one), v = fadd(fmul(K.d, y2), one);
    const Fe v3 = fmul(v, fmul(v, v)), v7 = fmul(v3, fmul(v3, v));
    const uint64_t e58[4] = {0xFFFFFFFFFFFFFFFDULL, ~0ULL, ~0ULL, 0x0FFFFFFFFFFFFFFFULL};  // (p - 5) / 8
    Fe r = fmul(fmul(u, v3), fpow(fmul(u, v7), e58));
    const Fe chk = fmul(v, fmul(r, r));
    const bool pos = feq(chk, u), ng = feq(chk, fneg(u));
    if (!pos && !ng) return false;
    if (ng) r = fmul(r, K.sqrtm1);
    if (r.v[0] & 1) r = fneg(r);
    if ((int)(r.v[0] & 1) != sign) r = fneg(r);
    out = Pt{r, y, one, fmul(r, y)};
    return true;
}
static Pt ident() { return Pt{fzero(), fone(), fone(), fzero()}; }
// extended-coordinate formulas of ed25519_group.rs:92-140
static Pt padd(const Pt &p, const Pt &q) {
    const Consts &K = consts();
    const Fe a = fmul(fsub(p.Y, p.X), fsub(q.Y, q.X)), b = fmul(fadd(p.Y, p.X), fadd(q.Y, q.X));
    const Fe c = fmul(fmul(p.T, q.T), K.d2), zz = fmul(p.Z, q.Z), d = fadd(zz, zz);
    const Fe e = fsub(b, a), f = fsub(d, c), g = fadd(d, c), h = fadd(b, a);
    return Pt{fmul(e, f), fmul(g, h), fmul(f, g), fmul(e, h)};
}
static Pt pdbl(const Pt &p) {
    const Fe a = fmul(p.X, p.X), b = fmul(p.Y, p.Y), zz = fmul(p.Z, p.Z), c = fadd(zz, zz), h = fadd(a, b);
    const Fe xy = fadd(p.X, p.Y), e = fsub(h, fmul(xy, xy)), g = fsub(a, b), f = fadd(c, g);
    return Pt{fmul(e, f), fmul(g, h), fmul(f, g), fmul(e, h)};
}
static Pt pneg(const Pt &p) { return Pt{fneg(p.X), p.Y, p.Z, fneg(p.T)}; }

// ---------------------------------------------------------------- scalars mod L, SHA-512 (host)

constexpr uint64_t LL[4] = {0x5812631a5cf5d3edULL, 0x14def9dea2f79cd6ULL, 0, 0x1000000000000000ULL};

// q = x div L (at most 260 bits), r = x mod L, for a 512-bit x (ed25519_scalar.rs:156-211)
static void divmod_l(const uint64_t x_in[8], uint64_t q[5], uint64_t r[4]) {
    uint64_t x[8], ls[8] = {LL[0], LL[1], LL[2], LL[3], 0, 0, 0, 0};
    memcpy(x, x_in, 64);
    for (int s = 0; s < 259; s++) {
        for (int i = 7; i > 0; i--) ls[i] = ls[i] << 1 | ls[i - 1] >> 63;
        ls[0] <<= 1;
    }
    memset(q, 0, 40);
    for (int step = 259; step >= 0; step--) {
        bool ge = true;
        for (int i = 7; i >= 0; i--)
            if (x[i] != ls[i]) {
                ge = x[i] > ls[i];
                break;
            }
        if (ge) {
            uint64_t br = 0;
            for (int i = 0; i < 8; i++) {
                const u128 d = (u128)x[i] - ls[i] - br;
                x[i] = (uint64_t)d;
                br = (uint64_t)(d >> 127) & 1;
            }
            q[step / 64] |= 1ULL << (step % 64);
        }
        for (int i = 0; i < 7; i++) ls[i] = ls[i] >> 1 | ls[i + 1] << 63;
        ls[7] >>= 1;
    }
    memcpy(r, x, 32);
}
static void le_words(const uint8_t *b, int n, uint64_t *w) {
    for (int k = 0; k < n; k++) {
        w[k] = 0;
        for (int i = 0; i < 8; i++) w[k] |= (uint64_t)b[8 * k + i] << (8 * i);
    }
}
// (a * b + c) mod L for 256-bit a, b, c
static void muladd_l(const uint64_t a[4], const uint64_t b[4], const uint64_t c[4], uint64_t out[4]) {
    uint64_t w[8] = {0};
    for (int i = 0; i < 4; i++) {
        u128 cy = 0;
        for (int j = 0; j < 4; j++) {
            cy += (u128)a[i] * b[j] + w[i + j];
            w[i + j] = (uint64_t)cy;
            cy >>= 64;
        }
        w[i + 4] = (uint64_t)cy;
    }
    u128 cy = 0;
    for (int i = 0; i < 8; i++) {
        cy += (u128)w[i] + (i < 4 ? c[i] : 0);
        w[i] = (uint64_t)cy;
        cy >>= 64;
    }
    uint64_t q[5];
    divmod_l(w, q, out);
}

#define ROTR64(x, n) (((x) >> (n)) | ((x) << (64 - (n))))
static void sha512(const uint8_t *msg, size_t len, uint8_t out[64]) {
    uint64_t h[8] = {0x6a09e667f3bcc908ULL, 0xbb67ae8584caa73bULL, 0x3c6ef372fe94f82bULL, 0xa54ff53a5f1d36f1ULL,
                     0x510e527fade682d1ULL, 0x9b05688c2b3e6c1fULL, 0x1f83d9abfb41bd6bULL, 0x5be0cd19137e2179ULL};
    const size_t nb = (len + 17 + 127) / 128;
    std::vector<uint8_t> m(128 * nb, 0);
    if (len) memcpy(m.data(), msg, len);
    m[len] = 0x80;
    for (int i = 0; i < 8; i++) m[128 * nb - 1 - i] = (uint8_t)(((uint64_t)len * 8) >> (8 * i));
    const uint64_t *K512 = sha512_round_constants();
    for (size_t b = 0; b < nb; b++) {
        uint64_t w[80], s[8];
        for (int t = 0; t < 16; t++) {
            w[t] = 0;
            for (int i = 0; i < 8; i++) w[t] = w[t] << 8 | m[128 * b + 8 * t + i];
        }
        for (int t = 16; t < 80; t++)
            w[t] = w[t - 16] + (ROTR64(w[t - 15], 1) ^ ROTR64(w[t - 15], 8) ^ (w[t - 15] >> 7)) + w[t - 7] +
                   (ROTR64(w[t - 2], 19) ^ ROTR64(w[t - 2], 61) ^ (w[t - 2] >> 6));
        memcpy(s, h, 64);
        for (int t = 0; t < 80; t++) {
            const uint64_t t1 = s[7] + (ROTR64(s[4], 14) ^ ROTR64(s[4], 18) ^ ROTR64(s[4], 41)) +
                                ((s[4] & s[5]) ^ (~s[4] & s[6])) + K512[t] + w[t];
            const uint64_t t2 = (ROTR64(s[0], 28) ^ ROTR64(s[0], 34) ^ ROTR64(s[0], 39)) +
                                ((s[0] & s[1]) ^ (s[0] & s[2]) ^ (s[1] & s[2]));
            memmove(s + 1, s, 56);
            s[4] += t1;
            s[0] = t1 + t2;
        }
        for (int k = 0; k < 8; k++) h[k] += s[k];
    }
    for (int k = 0; k < 8; k++)
        for (int i = 0; i < 8; i++) out[8 * k + i] = (uint8_t)(h[k] >> (56 - 8 * i));
}

// ---------------------------------------------------------------- gadget layouts

struct SubL { uint32_t a, b, c, cb, pos, neg; };
struct AddL { uint32_t a, b, c, cb, cy; };
struct MulL { uint32_t a, b, c, cb, cyb; };
struct SelL { uint32_t a, b, bit, c; };
struct PaddL {
    uint32_t p1[4], p2[4], d2;
    SubL ym1, ym2;
    AddL yp1, yp2;
    MulL A, B, tt, C, zz;
    AddL D;
    SubL E, F;
    AddL G, H;
    MulL X3, Y3, T3, Z3;
    uint32_t end;
};
struct PdblL {
    MulL sA, sB, szz;
    AddL C, H, xy;
    MulL sxy;
    SubL E, G;
    AddL F;
    MulL X3, Y3, T3, Z3;
    uint32_t end;
};
struct CaddL {
    uint32_t bit, out[4];
    PaddL add;
    SelL sel[4];
    uint32_t end;
};
struct DecL {
    uint32_t y, sign, d, x, xb, one;
    MulL sqx, sqy, dx2, dxy;
    SubL y2m1, lhs, eq;
    uint32_t end;
};

// CellAllocator (ed25519_group_air.rs:170-200)
struct Al {
    uint32_t n;
    uint32_t take(uint32_t k) {
        const uint32_t r = n;
        n += k;
        return r;
    }
    SubL sub(uint32_t a, uint32_t b) {
        SubL s{a, b, 0, 0, 0, 0};
        s.c = take(NL);
        s.cb = take(EB);
        s.pos = take(NL);
        s.neg = take(NL);
        return s;
    }
    AddL add(uint32_t a, uint32_t b) {
        AddL s{a, b, 0, 0, 0};
        s.c = take(NL);
        s.cb = take(EB);
        s.cy = take(NL);
        return s;
    }
    MulL mul(uint32_t a, uint32_t b) {
        MulL s{a, b, 0, 0, 0};
        s.c = take(NL);
        s.cb = take(EB);
        s.cyb = take(MCB * NL);
        return s;
    }
};
// point_add_layout_at (ed25519_group_air.rs:202-249)
static PaddL padd_at(uint32_t base, const uint32_t p1[4], const uint32_t p2[4]) {
    PaddL L;
    Al a{base};
    memcpy(L.p1, p1, 16);
    memcpy(L.p2, p2, 16);
    L.d2 = a.take(NL);
    L.ym1 = a.sub(p1[1], p1[0]);
    L.ym2 = a.sub(p2[1], p2[0]);
    L.yp1 = a.add(p1[1], p1[0]);
    L.yp2 = a.add(p2[1], p2[0]);
    L.A = a.mul(L.ym1.c, L.ym2.c);
    L.B = a.mul(L.yp1.c, L.yp2.c);
    L.tt = a.mul(p1[3], p2[3]);
    L.C = a.mul(L.tt.c, L.d2);
    L.zz = a.mul(p1[2], p2[2]);
    L.D = a.add(L.zz.c, L.zz.c);
    L.E = a.sub(L.B.c, L.A.c);
    L.F = a.sub(L.D.c, L.C.c);
    L.G = a.add(L.D.c, L.C.c);
    L.H = a.add(L.B.c, L.A.c);
    L.X3 = a.mul(L.E.c, L.F.c);
    L.Y3 = a.mul(L.G.c, L.H.c);
    L.T3 = a.mul(L.E.c, L.H.c);
    L.Z3 = a.mul(L.F.c, L.G.c);
    L.end = a.n;
    return L;
}
// point_double_layout_at (:476-505)
static PdblL pdbl_at(uint32_t base, uint32_t x, uint32_t y, uint32_t z) {
    PdblL L;
    Al a{base};
    L.sA = a.mul(x, x);
    L.sB = a.mul(y, y);
    L.szz = a.mul(z, z);
    L.C = a.add(L.szz.c, L.szz.c);
    L.H = a.add(L.sA.c, L.sB.c);
    L.xy = a.add(x, y);
    L.sxy = a.mul(L.xy.c, L.xy.c);
    L.E = a.sub(L.H.c, L.sxy.c);
    L.G = a.sub(L.sA.c, L.sB.c);
    L.F = a.add(L.C.c, L.G.c);
    L.X3 = a.mul(L.E.c, L.F.c);
    L.Y3 = a.mul(L.G.c, L.H.c);
    L.T3 = a.mul(L.E.c, L.H.c);
    L.Z3 = a.mul(L.F.c, L.G.c);
    L.end = a.n;
    return L;
}
// cond_add_layout_at (:662-719)
static CaddL cadd_at(uint32_t base, const uint32_t acc[4], const uint32_t bp[4]) {
    CaddL L;
    Al a{base};
    L.bit = a.take(1);
    for (int i = 0; i < 4; i++) L.out[i] = a.take(NL);
    L.add = padd_at(a.n, acc, bp);
    const uint32_t o3[4] = {L.add.X3.c, L.add.Y3.c, L.add.Z3.c, L.add.T3.c};
    for (int i = 0; i < 4; i++) L.sel[i] = SelL{acc[i], o3[i], L.bit, L.out[i]};
    L.end = L.add.end;
    return L;
}
// point_decompress_layout_at (:887-955)
static DecL dec_at(uint32_t base, uint32_t y, uint32_t sign) {
    DecL L;
    Al a{base};
    L.y = y;
    L.sign = sign;
    L.d = a.take(NL);
    L.x = a.take(NL);
    L.xb = a.take(EB);
    L.sqx = a.mul(L.x, L.x);
    L.sqy = a.mul(y, y);
    L.dx2 = a.mul(L.d, L.sqx.c);
    L.dxy = a.mul(L.dx2.c, L.sqy.c);
    L.one = a.take(NL);
    L.y2m1 = a.sub(L.sqy.c, L.one);
    L.lhs = a.sub(L.y2m1.c, L.sqx.c);
    L.eq = a.sub(L.lhs.c, L.dxy.c);
    a.take(1);  // sign_match helper
    L.end = a.n;
    return L;
}

// ---------------------------------------------------------------- constraint blocks

enum : uint32_t { K_MUL, K_ADD, K_SUB, K_SEL, K_ELEM, K_PIN, K_ZC, K_BIND, K_BOOL, K_TRANS4, K_PIN4, K_REDUCE, K_FMT };
enum : uint32_t { BF_XN = 1, BF_YN = 2, BF_YREV = 4 };  // bind: x / y from the next row, y index descending

// a block: kind, count n, columns, and up to two targets (slot of its first
// constraint, the live rows [lo, hi], the gate group, the partial-sum slot)
struct Blk {
    uint32_t kind, n, ntg;
    uint32_t a, b, c, cb, x, y, z;
    uint32_t s[2], lo[2], hi[2], grp[2], ps[2];
};

uint32_t blk_constraints(const Blk &b) {
    switch (b.kind) {
    case K_MUL: return EC + NL + MCB * NL;
    case K_ADD: return EC + 2 * NL;
    case K_SUB: return EC + 4 * NL;
    case K_SEL: return 1 + NL;
    case K_ELEM: return EC;
    case K_TRANS4: case K_PIN4: return 4 * NL;
    case K_REDUCE: return Q_L * LBITS + Q_L + 2 * (R_L * LBITS + R_L) + R_L + P_L + P_L * CBITS;
    case K_FMT: return 560;
    default: return b.n;
    }
}

// per-proof constants (pins): offsets into a proof's constant vector
enum : uint32_t { C_D = 0, C_D2 = 10, C_ONE = 20, C_IDENT = 30, C_BP = 70, C_A = 110, C_RY = 150, C_RSIGN = 160,
                  C_AY = 161, C_ASIGN = 171, C_BITS = 172 };  // s bits (k), then k bits (k)

// ---------------------------------------------------------------- v16 layout

// row types of the expansion
enum : uint8_t { RT_PLAIN, RT_SHA, RT_SHA_LAST, RT_REDUCE, RT_DEC_R, RT_DEC_A, RT_LAD_SB, RT_LAD_KA, RT_RESIDUAL,
                 RT_DBL, RT_RESULT, NRT };
// descriptor: index << 8 | code; code < 64 = bit of a row word, D_VAL its
// value, D_POS / D_NEG a signed carry's flags, D_CVAL a per-proof constant
// word, D_RT + b (b < 64) bit b of a constant word on rows below h7
enum : uint32_t { D_VAL = 64, D_POS = 65, D_NEG = 66, D_CVAL = 67, D_RT = 128 };
constexpr uint32_t D_ZERO = 0xFFFFFFFFu, D_SKIP = 0xFFFFFFFEu;
// per-proof constant words of the expansion: the sB, kA and R threads (120
// limbs), then the r-thread bits (k bits in 64-bit words)
constexpr uint32_t CW_R = 120;

struct Layout {
    uint32_t k, k_hash, h, h7, width, n_cons, nw, ncw, nconst;
    uint32_t reduce_row, dec_r, dec_a, sbf, kaf, sbl, kal, residual, dbl[3], result;
    // columns
    uint32_t red_q, red_qb, red_r, red_rb, red_sl, red_slb, red_cyb, digest_bits;
    DecL dec;
    uint32_t base[4], acc[4];
    PdblL ldbl;
    CaddL ca;
    uint32_t sblk_sb, sblk_ka, r_thread;
    uint32_t res[4], zero_c;
    SubL sub_x, sub_t, sub_yz;
    uint32_t dbl_in[3];
    PdblL dbl2;
    uint32_t th_sb, th_ka, th_r, one_r;
    MulL mul_rt;
    uint32_t zero_r1;
    SubL neg_rx, neg_rt, neg_kax, neg_kat;
    PaddL pa_r1, pa_r2;
    // expansion
    std::vector<uint32_t> desc;  // NRT x width
    std::vector<uint8_t> rtype;  // h
    // composition
    std::vector<Blk> blocks;
    std::vector<std::array<uint32_t, 2>> gates;  // gate group -> [lo, hi]
    uint32_t nps = 0;                             // partial-sum slots
};

static uint32_t next_pow2(uint64_t x) {
    uint64_t h = 1;
    while (h < x) h <<= 1;
    return (uint32_t)h;
}

// word emitter: in describe mode it records each column's descriptor, in
// fill mode it writes the row's words (at w[i * ws]); both walk the same
// sequence of words, so indices agree
struct Em {
    uint32_t *desc = nullptr;
    uint64_t *w = nullptr;
    uint64_t ws = 1;
    uint32_t nw = 0;
    uint32_t put(uint64_t v) {
        if (w) w[(uint64_t)nw * ws] = v;
        return nw++;
    }
    void d(uint32_t col, uint32_t idx, uint32_t code) {
        if (desc) desc[col] = idx << 8 | code;
    }
    void limbs(uint32_t col, const Lb &l) {
        for (int i = 0; i < NL; i++) d(col + i, put((uint64_t)l[i]), D_VAL);
    }
    void elem(uint32_t c, uint32_t cb, const Lb &l) {
        uint32_t off = 0;
        for (int i = 0; i < NL; i++) {
            const uint32_t idx = put((uint64_t)l[i]);
            d(c + i, idx, D_VAL);
            for (int b = 0; b < LW[i]; b++) d(cb + off + b, idx, (uint32_t)b);
            off += LW[i];
        }
    }
    void val(uint32_t col, uint64_t v) { d(col, put(v), D_VAL); }
    void bits(uint32_t col, uint32_t n, const uint64_t *words) {
        for (uint32_t q = 0; q * 64 < n; q++) {
            const uint32_t idx = put(words[q]);
            for (uint32_t b = 0; b < 64 && 64 * q + b < n; b++) d(col + 64 * q + b, idx, b);
        }
    }
};

// ---------------------------------------------------------------- gadget witnesses (host)

// fill_add_gadget (ed25519_field_air.rs:330-429): c = freeze(a + b), carries in {0, 1}
static Lb w_add(Em &e, const AddL &L, const Lb &a, const Lb &b) {
    const Lb c = limbs(fadd(of_limbs(a), of_limbs(b)));
    int64_t cy[NL] = {0};
    for (int cand = 0; cand <= 1; cand++) {
        int64_t prev = 19 * cand;
        bool ok = true;
        for (int k = 0; k < NL && ok; k++) {
            const int64_t dd = a[k] + b[k] + prev - c[k];
            if (dd == 0) cy[k] = prev = 0;
            else if (dd == (int64_t)1 << LW[k]) cy[k] = prev = 1;
            else ok = false;
        }
        if (ok && cy[NL - 1] == cand) break;
        memset(cy, 0, sizeof cy);
    }
    e.elem(L.c, L.cb, c);
    for (int k = 0; k < NL; k++) e.val(L.cy + k, (uint64_t)cy[k]);
    return c;
}
// fill_sub_gadget (:584-670): net carries in {-1, 0, 1}
static Lb w_sub(Em &e, const SubL &L, const Lb &a, const Lb &b) {
    const Lb c = limbs(fsub(of_limbs(a), of_limbs(b)));
    int64_t net[NL] = {0};
    const int64_t tops[3] = {0, 1, -1};
    for (int ti = 0; ti < 3; ti++) {
        int64_t prev = 19 * tops[ti];
        bool ok = true;
        for (int k = 0; k < NL && ok; k++) {
            const int64_t rad = (int64_t)1 << LW[k], dd = a[k] - b[k] - prev - c[k];
            if (dd == 0) net[k] = prev = 0;
            else if (dd == -rad) net[k] = prev = 1;
            else if (dd == rad) net[k] = prev = -1;
            else ok = false;
        }
        if (ok && net[NL - 1] == tops[ti]) break;
        memset(net, 0, sizeof net);
    }
    e.elem(L.c, L.cb, c);
    for (int k = 0; k < NL; k++) {
        const uint32_t idx = e.put((uint64_t)net[k]);
        e.d(L.pos + k, idx, D_POS);
        e.d(L.neg + k, idx, D_NEG);
    }
    return c;
}
typedef __int128 i128;
static i128 floor_div(i128 x, i128 m) {
    i128 q = x / m;
    if ((x % m) && ((x < 0) != (m < 0))) q--;
    return q;
}
// fill_mul_gadget (:874-998): c = a b mod p; the carry chain
// t_k + carry_{k-1} - c_k - 2^w_k carry_k = 0 with carry_{-1} = 19 carry_9 has
// one solution, carry_9 = (sum_k t_k 2^off_k - c) / p; it is found from its
// floating estimate and the limb-0 congruence, and checked by closing the chain
static bool mul_carries(const Lb &a, const Lb &b, const Lb &c, int64_t cy[NL]) {
    i128 tk[NL];
    long double est = 0;
    for (int k = 0; k < NL; k++) {
        tk[k] = 0;
        for (int i = 0; i < NL; i++)
            for (int j = 0; j < NL; j++) {
                const int odd = (i & 1) && (j & 1);
                if (i + j == k) tk[k] += (i128)(odd ? 2 : 1) * a[i] * b[j];
                else if (i + j == k + NL) tk[k] += (i128)(odd ? 38 : 19) * a[i] * b[j];
            }
        est += ((long double)tk[k] - (long double)c[k]) * ldexpl(1.0L, LOFF[k] - 255);
    }
    const i128 r0 = (i128)1 << 26;
    // carry_9 = base (mod 2^26), base = (c_0 - t_0) / 19 mod 2^26
    i128 base = (((i128)c[0] - tk[0]) % r0 + r0) % r0;
    base = base * 7064091 % r0;  // 19^-1 mod 2^26
    const i128 m = floor_div((i128)llroundl(est) - base + r0 / 2, r0);
    for (i128 dm = -1; dm <= 1; dm++) {
        const i128 guess = base + (m + dm) * r0;
        i128 prev = 19 * guess;
        bool ok = true;
        for (int k = 0; k < NL && ok; k++) {
            const i128 rad = (i128)1 << LW[k], dd = tk[k] + prev - c[k];
            if (dd % rad) ok = false;
            else cy[k] = (int64_t)(prev = dd / rad);
        }
        if (ok && cy[NL - 1] == guess) return true;
    }
    return false;
}
static Lb w_mul(Em &e, const MulL &L, const Lb &a, const Lb &b, bool *bad) {
    const Lb c = limbs(fmul(of_limbs(a), of_limbs(b)));
    int64_t cy[NL] = {0};
    if (!mul_carries(a, b, c, cy)) {
        memset(cy, 0, sizeof cy);
        if (bad) *bad = true;
    }
    e.elem(L.c, L.cb, c);
    for (int k = 0; k < NL; k++) {
        const uint32_t idx = e.put((uint64_t)(cy[k] + ((int64_t)1 << (MCB - 1))));
        for (int bb = 0; bb < MCB; bb++) e.d(L.cyb + MCB * k + bb, idx, (uint32_t)bb);
    }
    return c;
}
static void w_sel(Em &e, const SelL &L, const Lb &a, const Lb &b, int bit) {
    e.val(L.bit, (uint64_t)bit);
    e.limbs(L.c, bit ? b : a);
}
struct PL {
    Lb x, y, z, t;
};
static PL pl(const Pt &p) { return PL{limbs(p.X), limbs(p.Y), limbs(p.Z), limbs(p.T)}; }
// fill_point_add_gadget (ed25519_group_air.rs:251-321)
static void w_padd(Em &e, const PaddL &L, const Pt &p1, const Pt &p2, bool *bad) {
    const Consts &K = consts();
    e.limbs(L.d2, K.d2_l);
    const PL q1 = pl(p1), q2 = pl(p2);
    const Lb ym1 = w_sub(e, L.ym1, q1.y, q1.x), ym2 = w_sub(e, L.ym2, q2.y, q2.x);
    const Lb yp1 = w_add(e, L.yp1, q1.y, q1.x), yp2 = w_add(e, L.yp2, q2.y, q2.x);
    const Lb A = w_mul(e, L.A, ym1, ym2, bad), B = w_mul(e, L.B, yp1, yp2, bad);
    const Lb tt = w_mul(e, L.tt, q1.t, q2.t, bad), C = w_mul(e, L.C, tt, K.d2_l, bad);
    const Lb zz = w_mul(e, L.zz, q1.z, q2.z, bad), D = w_add(e, L.D, zz, zz);
    const Lb E = w_sub(e, L.E, B, A), F = w_sub(e, L.F, D, C), G = w_add(e, L.G, D, C), H = w_add(e, L.H, B, A);
    w_mul(e, L.X3, E, F, bad);
    w_mul(e, L.Y3, G, H, bad);
    w_mul(e, L.T3, E, H, bad);
    w_mul(e, L.Z3, F, G, bad);
}
// fill_point_double_gadget (:507-546)
static void w_pdbl(Em &e, const PdblL &L, const Pt &p, bool *bad) {
    const PL q = pl(p);
    const Lb A = w_mul(e, L.sA, q.x, q.x, bad), B = w_mul(e, L.sB, q.y, q.y, bad), zz = w_mul(e, L.szz, q.z, q.z, bad);
    const Lb C = w_add(e, L.C, zz, zz), H = w_add(e, L.H, A, B), xy = w_add(e, L.xy, q.x, q.y);
    const Lb xys = w_mul(e, L.sxy, xy, xy, bad);
    const Lb E = w_sub(e, L.E, H, xys), G = w_sub(e, L.G, A, B), F = w_add(e, L.F, C, G);
    w_mul(e, L.X3, E, F, bad);
    w_mul(e, L.Y3, G, H, bad);
    w_mul(e, L.T3, E, H, bad);
    w_mul(e, L.Z3, F, G, bad);
}

// ---------------------------------------------------------------- rows (host)

// the values one proof's rows need
struct Native {
    uint8_t digest[64], kc[32];
    std::vector<uint8_t> s_bits, k_bits;  // MSB-first, k each
    Pt R, A, sB, kA, res1, res2;
    int r_sign = 0, a_sign = 0;  // the encodings' sign bits
    std::vector<Pt> acc_sb, acc_ka;  // ladder accumulators at each row
    uint64_t cw[CW_R + 4];           // expansion constant words
    std::vector<uint64_t> cst;       // pins
};

static Lb one_l() { return limbs(fone()); }
static Lb zero_l() { return limbs(fzero()); }

static void row_common(Em &e, const Layout &L) {
    if (!e.desc) return;
    for (uint32_t i = 0; i < 40; i++) {
        e.d(L.th_sb + i, i, D_CVAL);
        e.d(L.th_ka + i, 40 + i, D_CVAL);
        e.d(L.th_r + i, 80 + i, D_CVAL);
    }
    for (uint32_t i = 0; i < L.k; i++) e.d(L.r_thread + i, CW_R + i / 64, D_RT + i % 64);
}
static void row_reduce(Em &e, const Layout &L, const uint8_t dg[64]) {
    uint64_t x[8], q[5], r[4];
    le_words(dg, 8, x);
    divmod_l(x, q, r);
    for (uint32_t i = 0; i < IN_L; i++) e.val(i, (uint64_t)dg[2 * i] | (uint64_t)dg[2 * i + 1] << 8);
    int64_t ql[Q_L], rl[R_L], sl[R_L], lf[R_L];
    for (uint32_t i = 0; i < Q_L; i++) ql[i] = (int64_t)((q[i / 4] >> (16 * (i % 4))) & 0xFFFF);
    for (uint32_t i = 0; i < R_L; i++) {
        rl[i] = (int64_t)((r[i / 4] >> (16 * (i % 4))) & 0xFFFF);
        lf[i] = (int64_t)((LL[i / 4] >> (16 * (i % 4))) & 0xFFFF);
    }
    int64_t br = 0;
    for (uint32_t i = 0; i < R_L; i++) {  // slack = (L - 1) - r (ed25519_scalar_air.rs:288-309)
        const int64_t v = (i == 0 ? lf[0] - 1 : lf[i]) - rl[i] - br;
        sl[i] = v < 0 ? v + 65536 : v;
        br = v < 0;
    }
    auto limbs16 = [&](uint32_t lc, uint32_t bc, const int64_t *v, uint32_t n) {
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t idx = e.put((uint64_t)v[i]);
            e.d(lc + i, idx, D_VAL);
            for (uint32_t b = 0; b < LBITS; b++) e.d(bc + LBITS * i + b, idx, b);
        }
    };
    limbs16(L.red_q, L.red_qb, ql, Q_L);
    limbs16(L.red_r, L.red_rb, rl, R_L);
    limbs16(L.red_sl, L.red_slb, sl, R_L);
    // carries of input - q L - r, 16-bit positions (compute_chain_carries, :311-356)
    int64_t prev = 0;
    for (uint32_t k = 0; k < P_L; k++) {
        int64_t prod = 0;
        for (uint32_t i = 0; i < Q_L; i++)
            if (k >= i && k - i < R_L) prod += ql[i] * lf[k - i];
        const int64_t in = k < IN_L ? ((int64_t)dg[2 * k] | (int64_t)dg[2 * k + 1] << 8) : 0;
        const int64_t tot = in - prod - (k < R_L ? rl[k] : 0) + prev;
        const int64_t cout = (int64_t)floor_div(tot, 65536);
        const uint32_t idx = e.put((uint64_t)(cout + (1 << (CBITS - 1))));
        for (uint32_t b = 0; b < CBITS; b++) e.d(L.red_cyb + CBITS * k + b, idx, b);
        prev = cout;
    }
}
static void row_dec(Em &e, const Layout &L, const Pt &p, int sign, bool is_r, bool *bad) {
    const Consts &K = consts();
    const PL q = pl(p);
    e.limbs(L.dec.y, q.y);
    e.val(L.dec.sign, (uint64_t)sign);
    e.limbs(L.dec.d, K.d_l);
    e.limbs(L.dec.one, one_l());
    e.elem(L.dec.x, L.dec.xb, q.x);
    const Lb x2 = w_mul(e, L.dec.sqx, q.x, q.x, bad), y2 = w_mul(e, L.dec.sqy, q.y, q.y, bad);
    const Lb dx2 = w_mul(e, L.dec.dx2, K.d_l, x2, bad), dxy = w_mul(e, L.dec.dxy, dx2, y2, bad);
    const Lb y2m1 = w_sub(e, L.dec.y2m1, y2, one_l()), lhs = w_sub(e, L.dec.lhs, y2m1, x2);
    w_sub(e, L.dec.eq, lhs, dxy);
    if (is_r) {  // v13: the R thread's one constant and T = X Y
        e.limbs(L.one_r, one_l());
        w_mul(e, L.mul_rt, q.x, q.y, bad);
    }
}
static void row_ladder(Em &e, const Layout &L, const Pt &base, const Pt &acc, int bit, uint32_t blk,
                       const uint64_t win[4], bool *bad) {
    const PL b = pl(base), a = pl(acc);
    for (int i = 0; i < 4; i++) e.limbs(L.base[i], i == 0 ? b.x : i == 1 ? b.y : i == 2 ? b.z : b.t);
    for (int i = 0; i < 4; i++) e.limbs(L.acc[i], i == 0 ? a.x : i == 1 ? a.y : i == 2 ? a.z : a.t);
    w_pdbl(e, L.ldbl, acc, bad);
    const Pt d = pdbl(acc);
    w_padd(e, L.ca.add, d, base, bad);
    const PL s = pl(padd(d, base)), dl = pl(d);
    w_sel(e, L.ca.sel[0], dl.x, s.x, bit);
    w_sel(e, L.ca.sel[1], dl.y, s.y, bit);
    w_sel(e, L.ca.sel[2], dl.z, s.z, bit);
    w_sel(e, L.ca.sel[3], dl.t, s.t, bit);
    e.bits(blk, L.k, win);
}
static void row_residual(Em &e, const Layout &L, const Native &N, bool *bad) {
    const PL r = pl(Pt{N.R.X, N.R.Y, fone(), fmul(N.R.X, N.R.Y)}), ka = pl(N.kA);
    w_sub(e, L.neg_rx, zero_l(), r.x);
    w_sub(e, L.neg_rt, zero_l(), r.t);
    w_padd(e, L.pa_r1, N.sB, pneg(N.R), bad);
    w_sub(e, L.neg_kax, zero_l(), ka.x);
    w_sub(e, L.neg_kat, zero_l(), ka.t);
    w_padd(e, L.pa_r2, N.res1, pneg(N.kA), bad);
}
static void row_dbl(Em &e, const Layout &L, const Pt &p, bool *bad) {
    const PL q = pl(p);
    e.limbs(L.dbl_in[0], q.x);
    e.limbs(L.dbl_in[1], q.y);
    e.limbs(L.dbl_in[2], q.z);
    w_pdbl(e, L.dbl2, p, bad);
}
static void row_result(Em &e, const Layout &L, const Pt &p8) {
    const PL q = pl(p8);
    e.limbs(L.res[0], q.x);
    e.limbs(L.res[1], q.y);
    e.limbs(L.res[2], q.z);
    e.limbs(L.res[3], q.t);
    w_sub(e, L.sub_x, q.x, zero_l());
    w_sub(e, L.sub_t, q.t, zero_l());
    w_sub(e, L.sub_yz, q.y, q.z);
}
static void row_sha_last(Em &e, const Layout &L, const uint8_t dg[64]) {
    uint64_t w[8];
    for (int k = 0; k < 8; k++) {
        w[k] = 0;
        for (int i = 0; i < 8; i++) w[k] = w[k] << 8 | dg[8 * k + i];
    }
    e.bits(L.digest_bits, 512, w);
}

// the 256-bit window of MSB-first bits starting at j (cell i = bits[j + i])
static void window(const std::vector<uint8_t> &bits, uint32_t j, uint64_t w[4]) {
    memset(w, 0, 32);
    for (uint32_t i = 0; j + i < bits.size(); i++)
        if (bits[j + i]) w[i / 64] |= 1ULL << (i % 64);
}

// emit row r of one proof (fill) or of a row type (describe, N = dummy)
static void emit_row(Em &e, const Layout &L, uint8_t rt, uint32_t r, const Native &N, bool *bad) {
    row_common(e, L);
    switch (rt) {
    case RT_SHA_LAST: row_sha_last(e, L, N.digest); break;
    case RT_REDUCE: row_reduce(e, L, N.digest); break;
    case RT_DEC_R: row_dec(e, L, N.R, N.r_sign, true, bad); break;
    case RT_DEC_A: row_dec(e, L, N.A, N.a_sign, false, bad); break;
    case RT_LAD_SB:
    case RT_LAD_KA: {
        const bool sb = rt == RT_LAD_SB;
        const uint32_t j = r - (sb ? L.sbf : L.kaf);
        const std::vector<uint8_t> &bits = sb ? N.s_bits : N.k_bits;
        uint64_t win[4];
        window(bits, j, win);
        row_ladder(e, L, sb ? consts().B : N.A, (sb ? N.acc_sb : N.acc_ka)[j], bits[j], sb ? L.sblk_sb : L.sblk_ka,
                   win, bad);
        break;
    }
    case RT_RESIDUAL: row_residual(e, L, N, bad); break;
    case RT_DBL: {
        Pt p = N.res2;
        for (uint32_t i = 0; i < r - L.dbl[0]; i++) p = pdbl(p);
        row_dbl(e, L, p, bad);
        break;
    }
    case RT_RESULT: row_result(e, L, pdbl(pdbl(pdbl(N.res2)))); break;
    default: break;
    }
}
static Native dummy_native(uint32_t k) {
    Native N;
    memset(N.digest, 0, 64);
    memset(N.kc, 0, 32);
    N.s_bits.assign(k, 0);
    N.k_bits.assign(k, 0);
    N.R = N.A = N.sB = N.kA = N.res1 = N.res2 = consts().B;
    N.acc_sb.assign(k, ident());
    N.acc_ka.assign(k, ident());
    memset(N.cw, 0, sizeof N.cw);
    return N;
}

static void add_blocks(Layout &L);

// verify_air_layout_v16 (ed25519_verify_air.rs:162-4818) for a given k_hash and k_scalar
static Layout make_layout(uint32_t k_hash, uint32_t k) {
    Layout L;
    L.k = k;
    L.k_hash = k_hash;
    L.reduce_row = k_hash;
    L.dec_r = k_hash + 1;
    L.dec_a = k_hash + 2;
    L.sbf = k_hash + 3;
    L.kaf = L.sbf + k;
    L.sbl = L.sbf + k - 1;
    L.kal = L.kaf + k - 1;
    Al a{IN_L};
    L.red_q = a.take(Q_L);
    L.red_qb = a.take(Q_L * LBITS);
    L.red_r = a.take(R_L);
    L.red_rb = a.take(R_L * LBITS);
    L.red_sl = a.take(R_L);
    L.red_slb = a.take(R_L * LBITS);
    L.red_cyb = a.take(P_L * CBITS);
    const uint32_t reduce_w = a.n;
    L.digest_bits = SHA_W;
    L.dec = dec_at(NL + 1, 0, NL);
    for (int i = 0; i < 4; i++) {
        L.base[i] = (uint32_t)i * NL;
        L.acc[i] = (uint32_t)(4 + i) * NL;
    }
    L.ldbl = pdbl_at(8 * NL, L.acc[0], L.acc[1], L.acc[2]);
    const uint32_t o[4] = {L.ldbl.X3.c, L.ldbl.Y3.c, L.ldbl.Z3.c, L.ldbl.T3.c};
    L.ca = cadd_at(L.ldbl.end, o, L.base);
    uint32_t w = std::max({SHA_W + 512, reduce_w, L.dec.end, L.ca.end});
    const uint32_t h3 = next_pow2((uint64_t)k_hash + 3 + 2 * k);
    L.h7 = h3;
    L.sblk_sb = w;
    L.sblk_ka = w + k;
    L.r_thread = w + 2 * k;
    a.n = w + 3 * k;
    uint32_t h = h3 >= L.kal + 6 ? h3 : next_pow2((uint64_t)L.kal + 5);  // v10
    for (int i = 0; i < 4; i++) L.res[i] = a.take(NL);
    L.zero_c = a.take(NL);
    L.sub_x = a.sub(L.res[0], L.zero_c);
    L.sub_t = a.sub(L.res[3], L.zero_c);
    L.sub_yz = a.sub(L.res[1], L.res[2]);
    for (int i = 0; i < 3; i++) L.dbl_in[i] = a.take(NL);
    L.dbl2 = pdbl_at(a.n, L.dbl_in[0], L.dbl_in[1], L.dbl_in[2]);
    a.n = L.dbl2.end;
    L.th_sb = a.take(4 * NL);
    L.th_ka = a.take(4 * NL);
    L.th_r = a.take(4 * NL);
    L.one_r = a.take(NL);
    L.mul_rt = a.mul(L.th_r, L.th_r + NL);
    L.residual = L.kal + 1;
    for (int i = 0; i < 3; i++) L.dbl[i] = L.kal + 2 + i;
    L.result = L.kal + 5;
    L.zero_r1 = a.take(NL);
    L.neg_rx = a.sub(L.zero_r1, L.th_r);
    L.neg_rt = a.sub(L.zero_r1, L.th_r + 3 * NL);
    const uint32_t p1a[4] = {L.th_sb, L.th_sb + NL, L.th_sb + 2 * NL, L.th_sb + 3 * NL};
    const uint32_t p2a[4] = {L.neg_rx.c, L.th_r + NL, L.th_r + 2 * NL, L.neg_rt.c};
    L.pa_r1 = padd_at(a.n, p1a, p2a);
    a.n = L.pa_r1.end;
    if (!(h >= L.result + 2)) h = next_pow2((uint64_t)L.result + 1);  // v14
    L.neg_kax = a.sub(L.zero_r1, L.th_ka);
    L.neg_kat = a.sub(L.zero_r1, L.th_ka + 3 * NL);
    const uint32_t p1b[4] = {L.pa_r1.X3.c, L.pa_r1.Y3.c, L.pa_r1.Z3.c, L.pa_r1.T3.c};
    const uint32_t p2b[4] = {L.neg_kax.c, L.th_ka + NL, L.th_ka + 2 * NL, L.neg_kat.c};
    L.pa_r2 = padd_at(a.n, p1b, p2b);
    L.width = L.pa_r2.end;
    L.h = h;
    L.ncw = CW_R + 4;
    L.nconst = C_BITS + 2 * k;
    // row types
    L.rtype.assign(h, RT_PLAIN);
    for (uint32_t r = 0; r + 1 < k_hash; r++) L.rtype[r] = RT_SHA;
    L.rtype[k_hash - 1] = RT_SHA_LAST;
    L.rtype[L.reduce_row] = RT_REDUCE;
    L.rtype[L.dec_r] = RT_DEC_R;
    L.rtype[L.dec_a] = RT_DEC_A;
    for (uint32_t j = 0; j < k; j++) {
        L.rtype[L.sbf + j] = RT_LAD_SB;
        L.rtype[L.kaf + j] = RT_LAD_KA;
    }
    L.rtype[L.residual] = RT_RESIDUAL;
    for (int i = 0; i < 3; i++) L.rtype[L.dbl[i]] = RT_DBL;
    L.rtype[L.result] = RT_RESULT;
    // descriptors: describe each row type once with dummy values
    L.desc.assign((size_t)NRT * L.width, D_ZERO);
    const Native N = dummy_native(k);
    const uint32_t sample[NRT] = {L.h - 1, 0, k_hash - 1, L.reduce_row, L.dec_r, L.dec_a, L.sbf, L.kaf,
                                  L.residual, L.dbl[0], L.result};
    L.nw = 0;
    for (uint32_t t = 0; t < NRT; t++) {
        Em e;
        e.desc = &L.desc[(size_t)t * L.width];
        if (t == RT_SHA || t == RT_SHA_LAST)
            for (uint32_t col = 0; col < SHA_W; col++) e.desc[col] = D_SKIP;
        emit_row(e, L, (uint8_t)t, sample[t], N, nullptr);
        L.nw = std::max(L.nw, e.nw);
    }
    add_blocks(L);
    return L;
}

// the constraint blocks in the order of eval_verify_air_v16_per_row
struct BlockBuilder {
    Layout &L;
    uint32_t s = SHA_K;  // the SHA-512 slots come first and are composed apart
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> gid;
    uint32_t gate(uint32_t lo, uint32_t hi) {
        auto it = gid.find({lo, hi});
        if (it != gid.end()) return it->second;
        const uint32_t g = (uint32_t)L.gates.size();
        L.gates.push_back({lo, hi});
        gid[{lo, hi}] = g;
        return g;
    }
    Blk mk(uint32_t kind, uint32_t n = 0) {
        Blk b;
        memset(&b, 0, sizeof b);
        b.kind = kind;
        b.n = n;
        return b;
    }
    void push(Blk b, uint32_t lo, uint32_t hi) {
        b.ntg = 1;
        b.s[0] = s;
        b.lo[0] = lo;
        b.hi[0] = hi;
        b.grp[0] = gate(lo, hi);
        s += blk_constraints(b);
        L.blocks.push_back(b);
    }
    // a second target for the blocks [first, first + count) pushed earlier
    void share(size_t first, size_t count, uint32_t lo, uint32_t hi) {
        for (size_t i = 0; i < count; i++) {
            Blk &b = L.blocks[first + i];
            b.ntg = 2;
            b.s[1] = s;
            b.lo[1] = lo;
            b.hi[1] = hi;
            b.grp[1] = gate(lo, hi);
            s += blk_constraints(b);
        }
    }
    void mul(const MulL &m, uint32_t lo, uint32_t hi) {
        Blk b = mk(K_MUL);
        b.a = m.a, b.b = m.b, b.c = m.c, b.cb = m.cb, b.x = m.cyb;
        push(b, lo, hi);
    }
    void add(const AddL &m, uint32_t lo, uint32_t hi) {
        Blk b = mk(K_ADD);
        b.a = m.a, b.b = m.b, b.c = m.c, b.cb = m.cb, b.x = m.cy;
        push(b, lo, hi);
    }
    void sub(const SubL &m, uint32_t lo, uint32_t hi) {
        Blk b = mk(K_SUB);
        b.a = m.a, b.b = m.b, b.c = m.c, b.cb = m.cb, b.x = m.pos, b.y = m.neg;
        push(b, lo, hi);
    }
    void sel(const SelL &m, uint32_t lo, uint32_t hi) {
        Blk b = mk(K_SEL);
        b.a = m.a, b.b = m.b, b.c = m.c, b.x = m.bit;
        push(b, lo, hi);
    }
    void pin(uint32_t col, uint32_t n, uint32_t coff, uint32_t lo, uint32_t hi) {
        Blk b = mk(K_PIN, n);
        b.a = col, b.x = coff;
        push(b, lo, hi);
    }
    void zc(uint32_t col, uint32_t n, uint32_t lo, uint32_t hi) {
        Blk b = mk(K_ZC, n);
        b.a = col;
        push(b, lo, hi);
    }
    void bind(uint32_t xc, uint32_t yc, uint32_t n, uint32_t flags, uint32_t lo, uint32_t hi) {
        Blk b = mk(K_BIND, n);
        b.a = xc, b.b = yc, b.x = flags;
        push(b, lo, hi);
    }
    void padd(const PaddL &P, uint32_t lo, uint32_t hi) {
        pin(P.d2, NL, C_D2, lo, hi);
        sub(P.ym1, lo, hi);
        sub(P.ym2, lo, hi);
        add(P.yp1, lo, hi);
        add(P.yp2, lo, hi);
        mul(P.A, lo, hi);
        mul(P.B, lo, hi);
        mul(P.tt, lo, hi);
        mul(P.C, lo, hi);
        mul(P.zz, lo, hi);
        add(P.D, lo, hi);
        sub(P.E, lo, hi);
        sub(P.F, lo, hi);
        add(P.G, lo, hi);
        add(P.H, lo, hi);
        mul(P.X3, lo, hi);
        mul(P.Y3, lo, hi);
        mul(P.T3, lo, hi);
        mul(P.Z3, lo, hi);
    }
    void pdbl(const PdblL &P, uint32_t lo, uint32_t hi) {
        mul(P.sA, lo, hi);
        mul(P.sB, lo, hi);
        mul(P.szz, lo, hi);
        add(P.C, lo, hi);
        add(P.H, lo, hi);
        add(P.xy, lo, hi);
        mul(P.sxy, lo, hi);
        sub(P.E, lo, hi);
        sub(P.G, lo, hi);
        add(P.F, lo, hi);
        mul(P.X3, lo, hi);
        mul(P.Y3, lo, hi);
        mul(P.T3, lo, hi);
        mul(P.Z3, lo, hi);
    }
    void decomp(const DecL &D, uint32_t lo, uint32_t hi) {  // eval_point_decompress_gadget (:1031-1090)
        pin(D.d, NL, C_D, lo, hi);
        Blk b = mk(K_ELEM);
        b.c = D.x, b.cb = D.xb;
        push(b, lo, hi);
        mul(D.sqx, lo, hi);
        mul(D.sqy, lo, hi);
        mul(D.dx2, lo, hi);
        mul(D.dxy, lo, hi);
        sub(D.y2m1, lo, hi);
        sub(D.lhs, lo, hi);
        sub(D.eq, lo, hi);
        zc(D.eq.c, NL, lo, hi);
        Blk bo = mk(K_BOOL, 1);
        bo.a = D.sign;
        push(bo, lo, hi);
        bind(D.xb, D.sign, 1, 0, lo, hi);
    }
};

static void add_blocks(Layout &L) {
    BlockBuilder B{L};
    const uint32_t k = L.k, h = L.h;
    // v0: scalar reduce; v1: digest format
    Blk r = B.mk(K_REDUCE);
    r.a = L.red_q, r.b = L.red_qb, r.c = L.red_r, r.cb = L.red_rb, r.x = L.red_sl, r.y = L.red_slb, r.z = L.red_cyb;
    B.push(r, L.reduce_row, L.reduce_row);
    Blk f = B.mk(K_FMT);
    f.a = L.digest_bits, f.b = OFF_H0, f.c = 0;
    B.push(f, L.k_hash - 1, L.k_hash - 1);
    // v2: decompress R, then the same blocks for A
    size_t first = L.blocks.size();
    B.decomp(L.dec, L.dec_r, L.dec_r);
    B.share(first, L.blocks.size() - first, L.dec_a, L.dec_a);
    // v3: the ladders (per-row gadgets shared by both, then transitions and the initial accumulator)
    const uint32_t fr[2] = {L.sbf, L.kaf}, ls[2] = {L.sbl, L.kal};
    size_t row_first = 0, row_count = 0, tr_first = 0, in_first = 0;
    for (int q = 0; q < 2; q++) {
        if (q == 0) {
            row_first = L.blocks.size();
            B.pdbl(L.ldbl, fr[0], ls[0]);
            B.padd(L.ca.add, fr[0], ls[0]);
            for (int i = 0; i < 4; i++) B.sel(L.ca.sel[i], fr[0], ls[0]);
            row_count = L.blocks.size() - row_first;
            tr_first = L.blocks.size();
            Blk t1 = B.mk(K_TRANS4);
            t1.a = L.base[0], t1.b = L.base[0];
            B.push(t1, fr[0], ls[0] - 1);
            Blk t2 = B.mk(K_TRANS4);
            t2.a = L.acc[0], t2.b = L.ca.out[0];
            B.push(t2, fr[0], ls[0] - 1);
            in_first = L.blocks.size();
            Blk p4 = B.mk(K_PIN4);
            p4.a = L.acc[0], p4.x = C_IDENT;
            B.push(p4, fr[0], fr[0]);
        } else {
            B.share(row_first, row_count, fr[1], ls[1]);
            B.share(tr_first, 2, fr[1], ls[1] - 1);
            B.share(in_first, 1, fr[1], fr[1]);
        }
    }
    // v4: base pins at the ladder starts
    B.pin(L.base[0], 4 * NL, C_BP, L.sbf, L.sbf);
    B.pin(L.base[0], 4 * NL, C_A, L.kaf, L.kaf);
    // v5: y and sign pins at the decompression rows
    B.pin(L.dec.y, NL, C_RY, L.dec_r, L.dec_r);
    B.pin(L.dec.sign, 1, C_RSIGN, L.dec_r, L.dec_r);
    B.pin(L.dec.y, NL, C_AY, L.dec_a, L.dec_a);
    B.pin(L.dec.sign, 1, C_ASIGN, L.dec_a, L.dec_a);
    // v6: the shifting scalar blocks
    const uint32_t blk[2] = {L.sblk_sb, L.sblk_ka};
    for (int q = 0; q < 2; q++) {
        B.pin(blk[q], k, C_BITS + (uint32_t)q * k, fr[q], fr[q]);
        B.bind(L.ca.bit, blk[q], 1, 0, fr[q], ls[q]);
        B.bind(blk[q], blk[q] + 1, k - 1, BF_XN, fr[q], ls[q] - 1);
    }
    // v7: the r thread
    B.bind(L.r_thread, L.red_rb + k - 1, k, BF_YREV, L.reduce_row, L.reduce_row);
    B.bind(L.r_thread, L.r_thread, k, BF_XN, 0, h - 1);
    B.bind(L.sblk_ka, L.r_thread, k, 0, L.kaf, L.kaf);
    // v10: result row, cofactor doublings
    B.zc(L.zero_c, NL, L.result, L.result);
    B.sub(L.sub_x, L.result, L.result);
    B.sub(L.sub_t, L.result, L.result);
    B.sub(L.sub_yz, L.result, L.result);
    B.zc(L.sub_x.c, NL, L.result, L.result);
    B.zc(L.sub_t.c, NL, L.result, L.result);
    B.zc(L.sub_yz.c, NL, L.result, L.result);
    B.pdbl(L.dbl2, L.dbl[0], L.dbl[2]);
    B.bind(L.dbl_in[0], L.dbl2.X3.c, NL, BF_XN, L.dbl[0], L.dbl[1]);
    B.bind(L.dbl_in[1], L.dbl2.Y3.c, NL, BF_XN, L.dbl[0], L.dbl[1]);
    B.bind(L.dbl_in[2], L.dbl2.Z3.c, NL, BF_XN, L.dbl[0], L.dbl[1]);
    B.bind(L.res[0], L.dbl2.X3.c, NL, BF_XN, L.dbl[2], L.dbl[2]);
    B.bind(L.res[1], L.dbl2.Y3.c, NL, BF_XN, L.dbl[2], L.dbl[2]);
    B.bind(L.res[2], L.dbl2.Z3.c, NL, BF_XN, L.dbl[2], L.dbl[2]);
    B.bind(L.res[3], L.dbl2.T3.c, NL, BF_XN, L.dbl[2], L.dbl[2]);
    // v11 / v12: ladder result threads (out_x..out_t and the threads are 40 contiguous columns)
    const uint32_t th[2] = {L.th_sb, L.th_ka};
    for (int q = 0; q < 2; q++) {
        B.bind(th[q], L.ca.out[0], 4 * NL, 0, ls[q], ls[q]);
        B.bind(th[q], th[q], 4 * NL, BF_XN, 0, h - 1);
    }
    // v13: the R thread
    B.bind(L.th_r, L.dec.x, NL, 0, L.dec_r, L.dec_r);
    B.bind(L.th_r + NL, L.dec.y, NL, 0, L.dec_r, L.dec_r);
    B.bind(L.th_r + 2 * NL, L.one_r, NL, 0, L.dec_r, L.dec_r);
    B.bind(L.th_r + 3 * NL, L.mul_rt.c, NL, 0, L.dec_r, L.dec_r);
    B.pin(L.one_r, NL, C_ONE, L.dec_r, L.dec_r);
    B.mul(L.mul_rt, L.dec_r, L.dec_r);
    B.bind(L.th_r, L.th_r, 4 * NL, BF_XN, 0, h - 1);
    // v14 / v15 / v16: the residual row
    const uint32_t rr = L.residual;
    B.zc(L.zero_r1, NL, rr, rr);
    B.sub(L.neg_rx, rr, rr);
    B.sub(L.neg_rt, rr, rr);
    B.padd(L.pa_r1, rr, rr);
    B.sub(L.neg_kax, rr, rr);
    B.sub(L.neg_kat, rr, rr);
    B.padd(L.pa_r2, rr, rr);
    B.bind(L.dbl_in[0], L.pa_r2.X3.c, NL, BF_XN, rr, rr);
    B.bind(L.dbl_in[1], L.pa_r2.Y3.c, NL, BF_XN, rr, rr);
    B.bind(L.dbl_in[2], L.pa_r2.Z3.c, NL, BF_XN, rr, rr);
    L.n_cons = B.s;
    uint32_t ps = 0;
    for (Blk &b : L.blocks)
        for (uint32_t t = 0; t < b.ntg; t++) b.ps[t] = ps++;
    L.nps = ps;
}

static std::mutex layouts_mu;
static const Layout &layout(uint32_t k_hash, uint32_t k) {
    static std::map<std::pair<uint32_t, uint32_t>, std::unique_ptr<Layout>> cache;
    std::lock_guard<std::mutex> g(layouts_mu);
    auto &p = cache[{k_hash, k}];
    if (!p) p.reset(new Layout(make_layout(k_hash, k)));
    return *p;
}

static uint32_t k_hash_of(size_t msg_len) {
    const uint64_t nb = (msg_len + 64 + 17 + 127) / 128;
    return std::max<uint32_t>(128, next_pow2(nb * 128));
}

// native values of one signature (fill_verify_air_v16, :4827-4862; prover.rs:884-896)
static Native native_of(const EdSig &s, const Layout &L) {
    const Consts &K = consts();
    Native N;
    const uint32_t k = L.k;
    std::vector<uint8_t> in(64 + s.len);
    memcpy(in.data(), s.sig, 32);
    memcpy(in.data() + 32, s.pk, 32);
    if (s.len) memcpy(in.data() + 64, s.msg, s.len);
    sha512(in.data(), in.size(), N.digest);
    uint64_t x[8], q[5], r[4];
    le_words(N.digest, 8, x);
    divmod_l(x, q, r);
    for (int kk = 0; kk < 4; kk++)
        for (int i = 0; i < 8; i++) N.kc[8 * kk + i] = (uint8_t)(r[kk] >> (8 * i));
    N.s_bits.resize(k);
    N.k_bits.resize(k);
    for (uint32_t i = 0; i < k; i++) {
        const uint32_t b = k - 1 - i;
        N.s_bits[i] = (s.sig[32 + b / 8] >> (b % 8)) & 1;
        N.k_bits[i] = (N.kc[b / 8] >> (b % 8)) & 1;
    }
    SG_REQUIRE(decompress(K, s.sig, N.R), "ed25519: R does not decompress");
    SG_REQUIRE(decompress(K, s.pk, N.A), "ed25519: the public key does not decompress");
    N.r_sign = s.sig[31] >> 7;
    N.a_sign = s.pk[31] >> 7;
    N.acc_sb.resize(k);
    N.acc_ka.resize(k);
    Pt a1 = ident(), a2 = ident();
    for (uint32_t j = 0; j < k; j++) {
        N.acc_sb[j] = a1;
        N.acc_ka[j] = a2;
        a1 = pdbl(a1);
        if (N.s_bits[j]) a1 = padd(a1, K.B);
        a2 = pdbl(a2);
        if (N.k_bits[j]) a2 = padd(a2, N.A);
    }
    N.sB = a1;
    N.kA = a2;
    N.res1 = padd(N.sB, pneg(N.R));
    N.res2 = padd(N.res1, pneg(N.kA));
    // expansion constants: threads (v14) and the r-thread words
    const PL t[3] = {pl(N.sB), pl(N.kA), pl(Pt{N.R.X, N.R.Y, fone(), fmul(N.R.X, N.R.Y)})};
    for (int q3 = 0; q3 < 3; q3++)
        for (int i = 0; i < NL; i++) {
            N.cw[40 * q3 + i] = (uint64_t)t[q3].x[i];
            N.cw[40 * q3 + 10 + i] = (uint64_t)t[q3].y[i];
            N.cw[40 * q3 + 20 + i] = (uint64_t)t[q3].z[i];
            N.cw[40 * q3 + 30 + i] = (uint64_t)t[q3].t[i];
        }
    memset(N.cw + CW_R, 0, 32);
    for (uint32_t i = 0; i < k; i++)
        if (N.k_bits[i]) N.cw[CW_R + i / 64] |= 1ULL << (i % 64);
    // pins
    N.cst.assign(L.nconst, 0);
    for (int i = 0; i < NL; i++) {
        N.cst[C_D + i] = (uint64_t)K.d_l[i];
        N.cst[C_D2 + i] = (uint64_t)K.d2_l[i];
    }
    N.cst[C_ONE] = 1;
    for (int kk = 0; kk < NL; kk++) {  // identity (0, 1, 1, 0), interleaved 4 kk + s
        N.cst[C_IDENT + 4 * kk + 1] = kk == 0;
        N.cst[C_IDENT + 4 * kk + 2] = kk == 0;
    }
    const PL bp = pl(K.B), ap = pl(N.A);
    const Lb *bl[4] = {&bp.x, &bp.y, &bp.z, &bp.t}, *al[4] = {&ap.x, &ap.y, &ap.z, &ap.t};
    for (int q4 = 0; q4 < 4; q4++)
        for (int i = 0; i < NL; i++) {
            N.cst[C_BP + 10 * q4 + i] = (uint64_t)(*bl[q4])[i];
            N.cst[C_A + 10 * q4 + i] = (uint64_t)(*al[q4])[i];
        }
    const Lb ry = limbs(N.R.Y), ay = limbs(N.A.Y);
    for (int i = 0; i < NL; i++) {
        N.cst[C_RY + i] = (uint64_t)ry[i];
        N.cst[C_AY + i] = (uint64_t)ay[i];
    }
    N.cst[C_RSIGN] = s.sig[31] >> 7;
    N.cst[C_ASIGN] = s.pk[31] >> 7;
    for (uint32_t i = 0; i < k; i++) {
        N.cst[C_BITS + i] = N.s_bits[i];
        N.cst[C_BITS + k + i] = N.k_bits[i];
    }
    return N;
}

// the words of every row of one proof, word-major (w[i * h + row]); false if
// a mul gadget's carry chain did not close (a value outside the field's range)
static bool proof_words(const EdSig &s, const Layout &L, const Native &N, uint64_t *w, uint32_t r0, uint32_t r1) {
    // Word i of row r lives at w[i * h + r] (rows fastest, as the expansion
    // kernel reads them). A row's words are emitted into a tile of 64 rows
    // held in cache (the same layout with stride 64) and each word's run of
    // rows is then copied out in one piece, instead of every row writing one
    // u64 into each of its nw cache lines of the destination.
    constexpr uint32_t TR = 64;
    thread_local std::vector<uint64_t> tile;
    tile.resize((size_t)L.nw * TR);
    bool bad = false;
    for (uint32_t rb = r0; rb < r1; rb += TR) {
        const uint32_t nr = std::min(TR, r1 - rb);
        uint32_t nw = 0;
        for (uint32_t r = rb; r < rb + nr; r++) {
            Em e;
            e.w = tile.data() + (r - rb);
            e.ws = TR;
            emit_row(e, L, L.rtype[r], r, N, &bad);
            nw = std::max(nw, e.nw);
        }
        for (uint32_t i = 0; i < nw; i++) memcpy(w + (size_t)i * L.h + rb, tile.data() + (size_t)i * TR, 8 * nr);
    }
    (void)s;
    return !bad;
}

__host__ __device__ inline uint64_t cell_of(uint32_t d, const uint64_t *wv, uint64_t wstride, const uint64_t *cw,
                                            uint32_t row, uint32_t h7) {
    if (d == D_ZERO) return 0;
    const uint32_t idx = d >> 8, code = d & 255;
    if (code >= D_RT) return row < h7 ? (cw[idx] >> (code - D_RT)) & 1 : 0;
    if (code == D_CVAL) return cw[idx];
    const uint64_t v = wv[(uint64_t)idx * wstride];
    if (code < 64) return (v >> code) & 1;
    if (code == D_VAL) return v;
    return code == D_POS ? (uint64_t)(v == 1) : (uint64_t)(v == ~0ULL);
}

// one cell per lane, rows fastest: tr[(g * W + col) * h + row]
__global__ __launch_bounds__(256) void expand_ed_k(const uint64_t *words, const uint64_t *cw, const uint32_t *desc,
                                                   const uint8_t *rtype, uint32_t nw, uint32_t ncw, uint32_t W,
                                                   uint32_t log_h, uint32_t h7, uint64_t *tr) {
    const uint64_t g = blockIdx.y, h = 1ULL << log_h;
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= h * W) return;
    const uint32_t row = (uint32_t)(t & (h - 1)), col = (uint32_t)(t >> log_h);
    const uint32_t d = desc[(uint64_t)rtype[row] * W + col];
    if (d == D_SKIP) return;
    tr[g * W * h + t] = cell_of(d, words + g * nw * h + row, h, cw + g * ncw, row, h7);
}

// ---------------------------------------------------------------- composition (device)

// 192-bit lazy sum of lambda * v products
struct A3 {
    uint64_t lo = 0, hi = 0, top = 0;
};
__device__ __forceinline__ void a3_mad(A3 &a, uint64_t v, uint64_t lam) {
    uint64_t l, h;
    mul128(v, lam, l, h);
    a.lo += l;
    const uint64_t c1 = a.lo < l;
    a.hi += h;
    uint64_t c2 = a.hi < h;
    a.hi += c1;
    c2 += a.hi < c1;
    a.top += c2;
}
__device__ __forceinline__ uint64_t a3_val(const A3 &a) {
    // 2^128 = -2^32 (mod p)
    return sub(reduce128(a.lo, a.hi), mul(canon(a.top), 1ULL << 32));
}
__device__ __forceinline__ uint64_t boolc(uint64_t b) { return sub(b, mul(b, b)); }

// L in 16-bit limbs (L_LIMBS_16, ed25519_scalar_air.rs:84-92)
__constant__ uint16_t L16[16] = {0xd3ed, 0x5cf5, 0x631a, 0x5812, 0x9cd6, 0xa2f7, 0xf9de, 0x14de,
                                 0, 0, 0, 0, 0, 0, 0, 0x1000};
// limb widths 26, 25, 26, ...
__device__ __forceinline__ int lw(int i) { return 26 - (i & 1); }

// GROUPS: part[(k * nparts + ps_t) * n + pt] for each target t of block
// blockIdx.y, the points being a blowup-2 LDE (lde: the traces, odd: the odd
// coset, n / 2 rows each). Direct: the LDE of n points; a lane adds the
// constraints only of the target whose rows hold its point's trace row, into
// part[(k * gridDim.y + blockIdx.y) * n + pt].
template <bool GROUPS>
__global__ __launch_bounds__(64) void compose_ed_k(const uint64_t *lde, const uint64_t *odd, uint64_t wcols,
                                                   uint64_t n, uint64_t blowup, const Blk *blocks,
                                                   const uint64_t *lam, const uint64_t *cst, uint32_t nconst,
                                                   uint32_t nparts, uint64_t *part) {
    const uint64_t k = blockIdx.z;
    const uint64_t pt = (uint64_t)blockIdx.x * 64 + threadIdx.x;
    const Blk B = blocks[blockIdx.y];
    const uint64_t rows = GROUPS ? n / 2 : n, at = GROUPS ? pt >> 1 : pt;
    const uint64_t *col = (GROUPS && (pt & 1) ? odd : lde) + k * wcols * rows;
    const uint64_t ni = GROUPS ? (at + 1) & (rows - 1) : (pt + blowup) & (n - 1);
#define CUR(c) col[(uint64_t)(c) * rows + at]
#define NXT(c) col[(uint64_t)(c) * rows + ni]
    const uint64_t *cs = cst + k * nconst;
    A3 acc[2];
    uint32_t sbase = 0;
    bool live = true;
    if (!GROUPS) {
        const uint64_t row = pt / blowup;
        int sel = -1;
        for (uint32_t t = 0; t < B.ntg; t++)
            if (row >= B.lo[t] && row <= B.hi[t]) sel = (int)t;
        live = sel >= 0;
        if (!__any(live)) {  // no lane of the wave on a live row
            part[(k * gridDim.y + blockIdx.y) * n + pt] = 0;
            return;
        }
        sbase = sel == 1 ? B.s[1] : B.s[0];
    }
    uint32_t j = 0;
    auto put = [&](uint64_t v) __attribute__((always_inline)) {
        if (GROUPS) {
            const uint32_t s0 = B.s[0] + j;
            a3_mad(acc[0], v, lam ? lam[s0] : (uint64_t)s0 + 1);
            if (B.ntg > 1) {
                const uint32_t s1 = B.s[1] + j;
                a3_mad(acc[1], v, lam ? lam[s1] : (uint64_t)s1 + 1);
            }
        } else if (live) {
            const uint32_t sl = sbase + j;
            a3_mad(acc[0], v, lam ? lam[sl] : (uint64_t)sl + 1);
        }
        j++;
    };
    // The runs of cells below are read 16 at a time into registers before
    // they are used, so a wave has 16 loads in flight instead of one (the
    // slot bookkeeping of put between loads kept the compiler from batching
    // them): put(boolc(cell)) for n consecutive columns from c0 ...
    auto bools = [&](uint32_t c0, uint32_t n) __attribute__((always_inline)) {
#pragma unroll 1
        for (uint32_t i0 = 0; i0 < n; i0 += 16) {
            uint64_t v[16];
#pragma unroll
            for (int r = 0; r < 16; r++) v[r] = i0 + r < n ? CUR(c0 + i0 + r) : 0;
#pragma unroll
            for (int r = 0; r < 16; r++)
                if (i0 + r < n) put(boolc(v[r]));
        }
    };
    // ... and sum_b 2^b cell(c0 + b) for b < nb <= 64
    auto packed = [&](uint32_t c0, uint32_t nb) __attribute__((always_inline)) -> uint64_t {
        Acc sacc{0, 0};
#pragma unroll 1
        for (uint32_t b0 = 0; b0 < nb; b0 += 16) {
            uint64_t v[16];
#pragma unroll
            for (int r = 0; r < 16; r++) v[r] = b0 + r < nb ? CUR(c0 + b0 + r) : 0;
#pragma unroll
            for (int r = 0; r < 16; r++) acc_shl(sacc, v[r], (int)((b0 + r) & 63));  // a zero past nb adds nothing
        }
        return acc_val(sacc);
    };
    auto elem = [&](uint32_t c, uint32_t cb) __attribute__((always_inline)) {
        bools(cb, EB);
        uint32_t off = 0;
#pragma unroll 1
        for (int i = 0; i < NL; i++) {
            put(sub(CUR(c + i), packed(cb + off, (uint32_t)lw(i))));
            off += lw(i);
        }
    };
    switch (B.kind) {
    case K_MUL: {  // eval_mul_gadget (ed25519_field_air.rs:1000-1076)
        elem(B.c, B.cb);
        uint64_t a[NL], b[NL], cy[NL];
#pragma unroll
        for (int i = 0; i < NL; i++) {
            a[i] = CUR(B.a + i);
            b[i] = CUR(B.b + i);
        }
#pragma unroll 1
        for (int kk = 0; kk < NL; kk++) cy[kk] = sub(packed(B.x + MCB * kk, MCB), 1ULL << (MCB - 1));
#pragma unroll 1
        for (int kk = 0; kk < NL; kk++) {
            Acc t{0, 0};
#pragma unroll
            for (int i = 0; i < NL; i++) {
                const int jj = kk - i >= 0 ? kk - i : kk + NL - i;
                const bool odd = (i & 1) && (jj & 1);
                const uint32_t f = kk - i >= 0 ? (odd ? 2 : 1) : (odd ? 38 : 19);
                acc_mad(t, mul(a[i], b[jj]), f);
            }
            const uint64_t cin = kk == 0 ? mul_small(cy[NL - 1], 19) : cy[kk - 1];
            put(sub(sub(add(acc_val(t), cin), CUR(B.c + kk)), mul(1ULL << lw(kk), cy[kk])));
        }
        bools(B.x, MCB * NL);
        break;
    }
    case K_ADD: {  // eval_add_gadget (:431-484)
        elem(B.c, B.cb);
#pragma unroll 1
        for (int kk = 0; kk < NL; kk++) {
            const uint64_t cin = kk == 0 ? mul_small(CUR(B.x + NL - 1), 19) : CUR(B.x + kk - 1);
            put(sub(sub(add(add(CUR(B.a + kk), CUR(B.b + kk)), cin), CUR(B.c + kk)),
                    mul(1ULL << lw(kk), CUR(B.x + kk))));
        }
#pragma unroll 1
        for (int kk = 0; kk < NL; kk++) put(boolc(CUR(B.x + kk)));
        break;
    }
    case K_SUB: {  // eval_sub_gadget (:672-748)
        elem(B.c, B.cb);
        auto net = [&](int m) { return sub(CUR(B.x + m), CUR(B.y + m)); };
#pragma unroll 1
        for (int kk = 0; kk < NL; kk++) {
            const uint64_t nin = kk == 0 ? mul_small(net(NL - 1), 19) : net(kk - 1);
            put(add(sub(sub(sub(CUR(B.a + kk), CUR(B.b + kk)), nin), CUR(B.c + kk)), mul(1ULL << lw(kk), net(kk))));
        }
#pragma unroll 1
        for (int kk = 0; kk < NL; kk++) put(boolc(CUR(B.x + kk)));
#pragma unroll 1
        for (int kk = 0; kk < NL; kk++) put(boolc(CUR(B.y + kk)));
#pragma unroll 1
        for (int kk = 0; kk < NL; kk++) put(mul(CUR(B.x + kk), CUR(B.y + kk)));
        break;
    }
    case K_SEL: {  // eval_select_gadget (:1457-1477)
        const uint64_t bit = CUR(B.x);
        put(boolc(bit));
#pragma unroll 1
        for (int kk = 0; kk < NL; kk++) {
            const uint64_t av = CUR(B.a + kk);
            put(sub(sub(CUR(B.c + kk), av), mul(bit, sub(CUR(B.b + kk), av))));
        }
        break;
    }
    case K_ELEM: elem(B.c, B.cb); break;
    case K_PIN:
#pragma unroll 1
        for (uint32_t i = 0; i < B.n; i++) put(sub(CUR(B.a + i), cs[B.x + i]));
        break;
    case K_ZC:
#pragma unroll 1
        for (uint32_t i = 0; i < B.n; i++) put(CUR(B.a + i));
        break;
    case K_BIND:
#pragma unroll 1
        for (uint32_t i = 0; i < B.n; i++) {
            const uint64_t x = (B.x & BF_XN) ? NXT(B.a + i) : CUR(B.a + i);
            const uint32_t yc = (B.x & BF_YREV) ? B.b - i : B.b + i;
            const uint64_t y = (B.x & BF_YN) ? NXT(yc) : CUR(yc);
            put(sub(x, y));
        }
        break;
    case K_BOOL:
        bools(B.a, B.n);
        break;
    case K_TRANS4:  // eval_scalar_mult_transition (ed25519_scalar_mult_air.rs:208-237)
#pragma unroll 1
        for (int kk = 0; kk < NL; kk++)
#pragma unroll
            for (int q = 0; q < 4; q++) put(sub(NXT(B.a + NL * q + kk), CUR(B.b + NL * q + kk)));
        break;
    case K_PIN4:  // eval_scalar_mult_initial_acc (:239-257)
#pragma unroll 1
        for (int kk = 0; kk < NL; kk++)
#pragma unroll
            for (int q = 0; q < 4; q++) put(sub(CUR(B.a + NL * q + kk), cs[B.x + 4 * kk + q]));
        break;
    case K_REDUCE: {  // eval_scalar_reduce_gadget (ed25519_scalar_air.rs:412-584)
        const uint32_t lc[3] = {B.a, B.c, B.x}, bc[3] = {B.b, B.cb, B.y}, nn[3] = {Q_L, R_L, R_L};
#pragma unroll 1
        for (int g = 0; g < 3; g++) {
            bools(bc[g], nn[g] * LBITS);
#pragma unroll 1
            for (uint32_t i = 0; i < nn[g]; i++) put(sub(CUR(lc[g] + i), packed(bc[g] + LBITS * i, LBITS)));
        }
        j += R_L;  // the r + slack = L - 1 slots are zero (:475-477)
        uint64_t prev = 0;
#pragma unroll 1
        for (uint32_t kk = 0; kk < P_L; kk++) {
            const uint64_t cout = sub(packed(B.z + CBITS * kk, CBITS), 1ULL << (CBITS - 1));
            Acc qs{0, 0};
            for (uint32_t i = 0; i < Q_L; i++)
                if (kk >= i && kk - i < R_L) acc_mad(qs, CUR(B.a + i), L16[kk - i]);
            const uint64_t in = kk < IN_L ? CUR(kk) : 0, rk = kk < R_L ? CUR(B.c + kk) : 0;
            put(sub(add(sub(sub(in, acc_val(qs)), rk), prev), mul(1ULL << LBITS, cout)));
            prev = cout;
        }
        bools(B.z, P_L * CBITS);
        break;
    }
    case K_FMT: {  // the v1 digest-format block (ed25519_verify_air.rs:464-560)
        bools(B.a, 512);
#pragma unroll 1
        for (uint32_t w = 0; w < 8; w++) {
            put(sub(CUR(B.b + 2 * w), packed(B.a + 64 * w, 32)));
            put(sub(CUR(B.b + 2 * w + 1), packed(B.a + 64 * w + 32, 32)));
        }
#pragma unroll 1
        for (uint32_t w = 0; w < 8; w++)
            for (uint32_t m = 0; m < 4; m++) {
                const uint32_t lob = 56 - 16 * m, hib = lob - 8;
                Acc s{0, 0};
                for (int jj = 0; jj < 8; jj++) {
                    acc_shl(s, CUR(B.a + 64 * w + lob + jj), jj);
                    acc_shl(s, CUR(B.a + 64 * w + hib + jj), 8 + jj);
                }
                put(sub(NXT(B.c + 4 * w + m), acc_val(s)));
            }
        break;
    }
    default: break;
    }
#undef CUR
#undef NXT
    if (GROUPS) {
        part[(k * nparts + B.ps[0]) * n + pt] = a3_val(acc[0]);
        if (B.ntg > 1) part[(k * nparts + B.ps[1]) * n + pt] = a3_val(acc[1]);
    } else {
        part[(k * gridDim.y + blockIdx.y) * n + pt] = live ? a3_val(acc[0]) : 0;
    }
}

// out[(k * ng + g) * n + pt] = sum of the parts listed for group g
__global__ __launch_bounds__(256) void ed_gsum_k(const uint64_t *part, uint32_t nparts, const uint32_t *goff,
                                                 const uint32_t *gidx, uint32_t ng, uint64_t n, uint64_t *out) {
    const uint64_t pt = (uint64_t)blockIdx.x * 256 + threadIdx.x, k = blockIdx.z;
    const uint32_t g = blockIdx.y;
    if (pt >= n) return;
    Acc a{0, 0};
    for (uint32_t i = goff[g]; i < goff[g + 1]; i++) acc_add(a, part[(k * nparts + gidx[i]) * n + pt]);
    out[(k * ng + g) * n + pt] = acc_val(a);
}

// phi at every point: the Ed25519 gate-group polynomials whose rows hold the
// point's trace row, plus the SHA-512 groups on rows below k_hash - 1 (less
// the round constants' T1 terms)
__global__ __launch_bounds__(256) void combine_ed_k(const uint64_t *pe, const uint32_t *gates, uint32_t ng,
                                                    const uint64_t *ps, uint64_t n, uint64_t blowup, uint32_t k_hash,
                                                    const uint32_t *n_blocks, const uint64_t *t1_lam,
                                                    const uint64_t *kd, uint64_t *phi) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x, k = blockIdx.y;
    if (i >= n) return;
    const uint64_t row = i / blowup;
    Acc a{0, 0};
    const uint64_t *q = pe + k * ng * n + i;
    for (uint32_t g = 0; g < ng; g++)
        if (row >= gates[2 * g] && row <= gates[2 * g + 1]) acc_add(a, q[(uint64_t)g * n]);
    uint64_t v = acc_val(a);
    if (row + 1 < k_hash) {  // sha512 groups: ungated, compression, finalisation, other rows, schedule
        const uint64_t blk = row >> 7, br = row & 127;
        const bool past = blk >= n_blocks[k];
        const bool comp = !past && br < 80, fin = !past && br == 80, sched = !past && br >= 15 && br <= 78;
        const uint64_t *s = ps + k * 5 * n + i;
        v = add(v, add(s[0], fin ? s[2 * n] : s[3 * n]));
        if (comp) {
            const uint64_t kt = kd[br];
            v = sub(sub(add(v, s[n]), mul(t1_lam[0], kt & 0xFFFFFFFFu)), mul(t1_lam[1], kt >> 32));
        }
        if (sched) v = add(v, s[4 * n]);
    }
    phi[k * n + i] = v;
}

// phi += x
__global__ __launch_bounds__(256) void ed_add_k(uint64_t *phi, const uint64_t *x, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) phi[i] = add(phi[i], x[i]);
}

}  // namespace ed

using namespace ed;

EdShape ed25519_shape(size_t msg_len, uint32_t k) {
    SG_REQUIRE(k >= 2 && k <= 256, "ed25519: k_scalar must be in [2, 256]");
    const Layout &L = layout(k_hash_of(msg_len), k);
    EdShape s;
    s.width = L.width;
    s.height = L.h;
    s.n_constraints = L.n_cons;
    s.k_hash = L.k_hash;
    return s;
}

bool ed25519_verify(const EdSig &s) {
    const Consts &K = consts();
    Pt R, A;
    if (!decompress(K, s.sig, R) || !decompress(K, s.pk, A)) return false;
    uint64_t sv[4];
    le_words(s.sig + 32, 4, sv);
    if (ge4(sv, LL)) return false;
    std::vector<uint8_t> in(64 + s.len);
    memcpy(in.data(), s.sig, 32);
    memcpy(in.data() + 32, s.pk, 32);
    if (s.len) memcpy(in.data() + 64, s.msg, s.len);
    uint8_t dg[64];
    sha512(in.data(), in.size(), dg);
    uint64_t x[8], q[5], kv[4];
    le_words(dg, 8, x);
    divmod_l(x, q, kv);
    Pt a1 = ident(), a2 = ident();
    for (int i = 255; i >= 0; i--) {
        a1 = pdbl(a1);
        if ((sv[i / 64] >> (i % 64)) & 1) a1 = padd(a1, K.B);
        a2 = pdbl(a2);
        if ((kv[i / 64] >> (i % 64)) & 1) a2 = padd(a2, A);
    }
    Pt r = padd(padd(a1, pneg(R)), pneg(a2));
    r = pdbl(pdbl(pdbl(r)));
    return fis0(r.X) && feq(r.Y, r.Z);
}

void ed25519_sign(const uint8_t seed[32], const uint8_t *msg, size_t len, uint8_t pk[32], uint8_t sig[64]) {
    const Consts &K = consts();
    uint8_t h[64];
    sha512(seed, 32, h);
    h[0] &= 248;
    h[31] &= 127;
    h[31] |= 64;
    uint64_t a[4];
    le_words(h, 4, a);
    auto smul = [&](const uint64_t e[4], const Pt &b) {
        Pt acc = ident();
        for (int i = 255; i >= 0; i--) {
            acc = pdbl(acc);
            if ((e[i / 64] >> (i % 64)) & 1) acc = padd(acc, b);
        }
        return acc;
    };
    auto compress = [&](const Pt &p, uint8_t out[32]) {
        const uint64_t em2[4] = {0xFFFFFFFFFFFFFFEBULL, ~0ULL, ~0ULL, 0x7FFFFFFFFFFFFFFFULL};
        const Fe zi = fpow(p.Z, em2), x = fmul(p.X, zi), y = fmul(p.Y, zi);
        ftobytes(y, out);
        out[31] |= (uint8_t)((x.v[0] & 1) << 7);
    };
    compress(smul(a, K.B), pk);
    std::vector<uint8_t> buf(64 + len);
    memcpy(buf.data(), h + 32, 32);
    if (len) memcpy(buf.data() + 32, msg, len);
    uint8_t rh[64];
    sha512(buf.data(), 32 + len, rh);
    uint64_t rw[8], q[5], r[4];
    le_words(rh, 8, rw);
    divmod_l(rw, q, r);
    compress(smul(r, K.B), sig);
    memcpy(buf.data(), sig, 32);
    memcpy(buf.data() + 32, pk, 32);
    if (len) memcpy(buf.data() + 64, msg, len);
    uint8_t kh[64];
    sha512(buf.data(), 64 + len, kh);
    uint64_t kw[8], kr[4];
    le_words(kh, 8, kw);
    divmod_l(kw, q, kr);
    uint64_t s[4];
    muladd_l(kr, a, r, s);
    for (int kk = 0; kk < 4; kk++)
        for (int i = 0; i < 8; i++) sig[32 + 8 * kk + i] = (uint8_t)(s[kk] >> (8 * i));
}

void build_ed25519_trace_host(const EdSig &s, uint32_t k, uint64_t *tr) {
    SG_REQUIRE(k >= 2 && k <= 256, "ed25519: k_scalar must be in [2, 256]");
    const Layout &L = layout(k_hash_of(s.len), k);
    const Native N = native_of(s, L);
    std::vector<uint64_t> w((size_t)L.nw * L.h, 0);
    SG_REQUIRE(proof_words(s, L, N, w.data(), 0, L.h), "ed25519 trace: a field product does not close");
    // SHA-512 rows of R || A || M
    std::vector<uint8_t> in(64 + s.len);
    memcpy(in.data(), s.sig, 32);
    memcpy(in.data() + 32, s.pk, 32);
    if (s.len) memcpy(in.data() + 64, s.msg, s.len);
    std::vector<uint64_t> sha((size_t)SHA512_W * L.k_hash);
    uint8_t dg[64];
    build_sha512_trace(in.data(), in.size(), sha.data(), dg, L.k_hash);
    parallel_for(L.width, [&](size_t col) {
        for (uint32_t r = 0; r < L.h; r++) {
            const uint32_t d = L.desc[(size_t)L.rtype[r] * L.width + col];
            tr[col * L.h + r] = d == D_SKIP ? sha[col * L.k_hash + r] : cell_of(d, w.data() + r, L.h, N.cw, r, L.h7);
        }
    });
}

void build_ed25519_traces_dev(Ctx *c, const std::vector<EdSig> &sigs, uint32_t k, uint64_t *d_tr, EdConsts *consts) {
    const size_t G = sigs.size();
    SG_REQUIRE(G >= 1 && G <= 65535, "ed25519 traces: 1 to 65535 proofs");
    SG_REQUIRE(k >= 2 && k <= 256, "ed25519: k_scalar must be in [2, 256]");
    const uint32_t kh = k_hash_of(sigs[0].len);
    for (const EdSig &s : sigs) SG_REQUIRE(k_hash_of(s.len) == kh, "ed25519 traces: one SHA-512 row count per batch");
    const Layout &L = layout(kh, k);
    const uint64_t per = (uint64_t)L.nw * L.h;
    uint64_t *words = (uint64_t *)c->upload_staging(8 * (G * per + G * L.ncw));
    uint64_t *cw = words + G * per;
    std::vector<Native> nat(G);
    std::vector<std::string> err(G);
    {
        HostScope hs(c, "ed25519_native");
        parallel_for(G, [&](size_t g) {
            try {
                nat[g] = native_of(sigs[g], L);
            } catch (const std::exception &e) {
                err[g] = e.what();
            }
        });
    }
    for (size_t g = 0; g < G; g++)
        if (!err[g].empty()) throw Error(E_INVALID, "ed25519 traces: signature " + std::to_string(g) + ": " + err[g]);
    if (consts) {
        consts->cst.resize(G * L.nconst);
        consts->nb.resize(G);
        for (size_t g = 0; g < G; g++) {
            std::copy(nat[g].cst.begin(), nat[g].cst.end(), consts->cst.begin() + g * L.nconst);
            consts->nb[g] = (uint32_t)((sigs[g].len + 64 + 17 + 127) / 128);
        }
    }
    // rows in chunks of 64 across the batch
    const uint32_t chunks = (L.h + 63) / 64;
    std::vector<uint8_t> bad(G * chunks, 0);
    {
        HostScope hs(c, "ed25519_words");
        parallel_for(G * chunks, [&](size_t t) {
            const size_t g = t / chunks;
            const uint32_t r0 = (uint32_t)(t % chunks) * 64, r1 = std::min(L.h, r0 + 64);
            bad[t] = !proof_words(sigs[g], L, nat[g], words + g * per, r0, r1);
        });
    }
    for (size_t t = 0; t < bad.size(); t++)
        SG_REQUIRE(!bad[t], "ed25519 trace: a field product does not close");
    for (size_t g = 0; g < G; g++) memcpy(cw + g * L.ncw, nat[g].cw, 8 * L.ncw);
    Scoped d_w(c, 8 * (G * per + G * L.ncw)), d_desc(c, 4ULL * NRT * L.width), d_rt(c, L.h);
    SG_HIP(hipMemcpyAsync(d_w.p, words, 8 * (G * per + G * L.ncw), hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(d_desc.p, L.desc.data(), 4ULL * NRT * L.width, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(d_rt.p, L.rtype.data(), L.h, hipMemcpyHostToDevice, c->stream));
    {
        ProfScope ps(c, "expand_ed25519", 8.0 * L.width * L.h * G);
        hipLaunchKernelGGL(expand_ed_k, dim3((uint32_t)(((uint64_t)L.h * L.width + 255) / 256), (uint32_t)G),
                           dim3(256), 0, c->stream, d_w.as<uint64_t>(), d_w.as<uint64_t>() + G * per,
                           d_desc.as<uint32_t>(), d_rt.as<uint8_t>(), L.nw, L.ncw, L.width, ilog2(L.h), L.h7, d_tr);
        SG_HIP(hipGetLastError());
    }
    // the SHA-512 rows of R || A || M (columns 0..1509 of rows below k_hash)
    std::vector<std::vector<uint8_t>> ins(G);
    std::vector<ByteSpan> spans(G);
    for (size_t g = 0; g < G; g++) {
        ins[g].resize(64 + sigs[g].len);
        memcpy(ins[g].data(), sigs[g].sig, 32);
        memcpy(ins[g].data() + 32, sigs[g].pk, 32);
        if (sigs[g].len) memcpy(ins[g].data() + 64, sigs[g].msg, sigs[g].len);
        spans[g] = ByteSpan{ins[g].data(), ins[g].size()};
    }
    SG_HIP(hipStreamSynchronize(c->stream));  // the staging buffer is reused by the SHA upload
    std::vector<uint32_t> nb;
    std::vector<std::array<uint8_t, 64>> dg;
    build_sha512_traces_into(c, spans, L.k_hash, L.h, (uint64_t)L.width * L.h, d_tr, nb, dg);
}

// per-proof pins and SHA-512 block counts of a batch
static void batch_consts(Ctx *c, const std::vector<EdSig> &sigs, const Layout &L, Scoped &d_cst, Scoped &d_nb,
                         const EdConsts *pre = nullptr) {
    const size_t G = sigs.size();
    if (pre) {
        SG_REQUIRE(pre->cst.size() == G * L.nconst && pre->nb.size() == G, "ed25519: pins of another batch");
        SG_HIP(hipMemcpyAsync(d_cst.p, pre->cst.data(), 8 * pre->cst.size(), hipMemcpyHostToDevice, c->stream));
        SG_HIP(hipMemcpyAsync(d_nb.p, pre->nb.data(), 4 * G, hipMemcpyHostToDevice, c->stream));
        SG_HIP(hipStreamSynchronize(c->stream));
        return;
    }
    std::vector<uint64_t> cst(G * L.nconst);
    std::vector<uint32_t> nb(G);
    std::vector<std::string> err(G);
    parallel_for(G, [&](size_t g) {
        try {
            const Native N = native_of(sigs[g], L);
            std::copy(N.cst.begin(), N.cst.end(), cst.begin() + g * L.nconst);
        } catch (const std::exception &e) {
            err[g] = e.what();
        }
        nb[g] = (uint32_t)((sigs[g].len + 64 + 17 + 127) / 128);
    });
    for (size_t g = 0; g < G; g++)
        if (!err[g].empty()) throw Error(E_INVALID, "ed25519: signature " + std::to_string(g) + ": " + err[g]);
    SG_HIP(hipMemcpyAsync(d_cst.p, cst.data(), 8 * cst.size(), hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(d_nb.p, nb.data(), 4 * G, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
}

void merge_ed25519_batch(Ctx *c, const std::vector<EdSig> &sigs, uint32_t k, const uint64_t *lde, uint64_t blowup,
                         const uint64_t *d_coeffs, uint64_t *c_eval) {
    const size_t G = sigs.size();
    SG_REQUIRE(G >= 1 && G <= 65535, "merge_ed25519: 1 to 65535 proofs");
    SG_REQUIRE(k >= 2 && k <= 256, "ed25519: k_scalar must be in [2, 256]");
    const uint32_t kh = k_hash_of(sigs[0].len);
    for (const EdSig &s : sigs) SG_REQUIRE(k_hash_of(s.len) == kh, "merge_ed25519: one SHA-512 row count per batch");
    const Layout &L = layout(kh, k);
    SG_REQUIRE(is_pow2(blowup) && blowup >= 2, "merge_ed25519: blowup must be a power of two >= 2");
    const uint64_t n = (uint64_t)L.h * blowup;
    SG_REQUIRE(n <= (1ULL << 22), "merge_ed25519: domain above 2^22");
    const uint32_t nb = (uint32_t)L.blocks.size();
    Scoped d_cst(c, 8ULL * L.nconst * G), d_nb(c, 4 * G), d_blk(c, sizeof(Blk) * nb);
    batch_consts(c, sigs, L, d_cst, d_nb);
    SG_HIP(hipMemcpyAsync(d_blk.p, L.blocks.data(), sizeof(Blk) * nb, hipMemcpyHostToDevice, c->stream));
    Scoped phi(c, 8 * n * G);
    {
        Scoped part(c, 8ULL * nb * n * G);
        {
            ProfScope ps(c, "compose_ed25519", (8.0 * L.width * n + 8.0 * nb * n) * G);
            hipLaunchKernelGGL(compose_ed_k<false>, dim3((uint32_t)(n / 64), nb, (uint32_t)G), dim3(64), 0, c->stream,
                               lde, (const uint64_t *)nullptr, (uint64_t)L.width, n, blowup, d_blk.as<Blk>(), d_coeffs,
                               d_cst.as<uint64_t>(), L.nconst, 0u, part.as<uint64_t>());
            SG_HIP(hipGetLastError());
        }
        std::vector<uint32_t> goff = {0, nb}, gidx(nb);
        for (uint32_t i = 0; i < nb; i++) gidx[i] = i;
        Scoped d_go(c, 8), d_gi(c, 4 * nb);
        SG_HIP(hipMemcpyAsync(d_go.p, goff.data(), 8, hipMemcpyHostToDevice, c->stream));
        SG_HIP(hipMemcpyAsync(d_gi.p, gidx.data(), 4 * nb, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(ed_gsum_k, dim3((uint32_t)((n + 255) / 256), 1, (uint32_t)G), dim3(256), 0, c->stream,
                           part.as<uint64_t>(), nb, d_go.as<uint32_t>(), d_gi.as<uint32_t>(), 1u, n, phi.as<uint64_t>());
        SG_HIP(hipGetLastError());
        SG_HIP(hipStreamSynchronize(c->stream));
    }
    {
        // the SHA-512 slots on rows below k_hash - 1
        Scoped sp(c, 8 * n * G);
        compose_sha512_rows(c, lde, (uint64_t)L.width, (uint32_t)G, n, blowup, d_nb.as<uint32_t>(), L.k_hash,
                            d_coeffs, sp.as<uint64_t>());
        hipLaunchKernelGGL(ed_add_k, dim3((uint32_t)((n * G + 255) / 256)), dim3(256), 0, c->stream,
                           phi.as<uint64_t>(), sp.as<uint64_t>(), n * G);
        SG_HIP(hipGetLastError());
        SG_HIP(hipStreamSynchronize(c->stream));
    }
    quotient_batch(c, phi.as<uint64_t>(), (uint32_t)G, L.h, n, c_eval);
}

void merge_ed25519_from_trace_batch(Ctx *c, const std::vector<EdSig> &sigs, uint32_t k, const uint64_t *d_tr,
                                    uint64_t blowup, const uint64_t *d_coeffs, uint64_t *c_eval,
                                    const EdConsts *consts) {
    const size_t G = sigs.size();
    SG_REQUIRE(G >= 1 && G <= 65535, "merge_ed25519: 1 to 65535 proofs");
    SG_REQUIRE(k >= 2 && k <= 256, "ed25519: k_scalar must be in [2, 256]");
    const uint32_t kh = k_hash_of(sigs[0].len);
    for (const EdSig &s : sigs) SG_REQUIRE(k_hash_of(s.len) == kh, "merge_ed25519: one SHA-512 row count per batch");
    const Layout &L = layout(kh, k);
    SG_REQUIRE(is_pow2(blowup) && blowup >= 4, "merge_ed25519: blowup must be a power of two >= 4");
    const uint64_t h = L.h, n = h * blowup, n2 = 2 * h;
    SG_REQUIRE(n <= (1ULL << 22), "merge_ed25519: domain above 2^22");
    const uint32_t nb = (uint32_t)L.blocks.size(), ng = (uint32_t)L.gates.size();
    Scoped d_cst(c, 8ULL * L.nconst * G), d_nb(c, 4 * G), d_blk(c, sizeof(Blk) * nb), t1(c, 16);
    batch_consts(c, sigs, L, d_cst, d_nb, consts);
    SG_HIP(hipMemcpyAsync(d_blk.p, L.blocks.data(), sizeof(Blk) * nb, hipMemcpyHostToDevice, c->stream));
    // gate groups: CSR of the partial-sum slots, and each group's rows
    std::vector<uint32_t> goff(ng + 1, 0), gidx, gates(2 * ng);
    {
        std::vector<std::vector<uint32_t>> lists(ng);
        for (const Blk &b : L.blocks)
            for (uint32_t t = 0; t < b.ntg; t++) lists[b.grp[t]].push_back(b.ps[t]);
        for (uint32_t g = 0; g < ng; g++) {
            goff[g + 1] = goff[g] + (uint32_t)lists[g].size();
            gidx.insert(gidx.end(), lists[g].begin(), lists[g].end());
            gates[2 * g] = L.gates[g][0];
            gates[2 * g + 1] = L.gates[g][1];
        }
    }
    Scoped d_go(c, 4 * (ng + 1)), d_gi(c, 4 * gidx.size()), d_gates(c, 8 * ng);
    SG_HIP(hipMemcpyAsync(d_go.p, goff.data(), 4 * (ng + 1), hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(d_gi.p, gidx.data(), 4 * gidx.size(), hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(d_gates.p, gates.data(), 8 * ng, hipMemcpyHostToDevice, c->stream));
    Scoped pe(c, 8ULL * ng * n * G), ps(c, 8ULL * 5 * n * G);
    {
        Scoped ge2(c, 8ULL * ng * n2 * G), gs2(c, 8ULL * 5 * n2 * G);
        {
            Scoped odd(c, 8ULL * L.width * h * G);
            lde_odd_coset(c, d_tr, odd.as<uint64_t>(), ilog2(h), (uint64_t)L.width * G);
            {
                Scoped part(c, 8ULL * L.nps * n2 * G);
                {
                    ProfScope pz(c, "compose_ed25519", (8.0 * L.width * n2 + 8.0 * L.nps * n2) * G);
                    hipLaunchKernelGGL(compose_ed_k<true>, dim3((uint32_t)(n2 / 64), nb, (uint32_t)G), dim3(64), 0,
                                       c->stream, d_tr, odd.as<uint64_t>(), (uint64_t)L.width, n2, (uint64_t)2,
                                       d_blk.as<Blk>(), d_coeffs, d_cst.as<uint64_t>(), L.nconst, L.nps,
                                       part.as<uint64_t>());
                    SG_HIP(hipGetLastError());
                }
                hipLaunchKernelGGL(ed_gsum_k, dim3((uint32_t)((n2 + 255) / 256), ng, (uint32_t)G), dim3(256), 0,
                                   c->stream, part.as<uint64_t>(), L.nps, d_go.as<uint32_t>(), d_gi.as<uint32_t>(), ng,
                                   n2, ge2.as<uint64_t>());
                SG_HIP(hipGetLastError());
                SG_HIP(hipStreamSynchronize(c->stream));
            }
            compose_sha512_groups(c, d_tr, odd.as<uint64_t>(), (uint64_t)L.width, (uint32_t)G, n2,
                                  d_nb.as<uint32_t>(), d_coeffs, gs2.as<uint64_t>(), t1.as<uint64_t>());
            SG_HIP(hipStreamSynchronize(c->stream));
        }
        lde_columns(c, ge2.as<uint64_t>(), pe.as<uint64_t>(), ilog2(n2), ilog2(blowup / 2), (uint64_t)ng * G, n2, n,
                    1);
        lde_columns(c, gs2.as<uint64_t>(), ps.as<uint64_t>(), ilog2(n2), ilog2(blowup / 2), 5ULL * G, n2, n, 1);
    }
    Scoped phi(c, 8 * n * G), d_kd(c, 8 * 80);
    SG_HIP(hipMemcpyAsync(d_kd.p, sha512_round_constants(), 8 * 80, hipMemcpyHostToDevice, c->stream));
    {
        ProfScope pz(c, "combine_ed25519", 8.0 * (ng + 6) * n * G);
        hipLaunchKernelGGL(combine_ed_k, dim3((uint32_t)((n + 255) / 256), (uint32_t)G), dim3(256), 0, c->stream,
                           pe.as<uint64_t>(), d_gates.as<uint32_t>(), ng, ps.as<uint64_t>(), n, blowup, L.k_hash,
                           d_nb.as<uint32_t>(), t1.as<uint64_t>(), d_kd.as<uint64_t>(), phi.as<uint64_t>());
        SG_HIP(hipGetLastError());
    }
    quotient_batch(c, phi.as<uint64_t>(), (uint32_t)G, h, n, c_eval);
}

}  // namespace sg
