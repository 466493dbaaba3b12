// rsa_air.hip -- RSA-2048 exponentiation AIR for gfx950.
//
// Trace (crates/deep_ali/src/rsa2048_exp_air.rs:97-202,
// rsa2048_field_air.rs:129-162,244-344): each row holds one product
// c = acc * op2 mod n with its witness quotient q, 26-bit limbs (80 per
// value), the bits of c and q and the 40-bit biased carries of the
// limb-by-limb identity a*b - q*n - c = 0. Row 16 multiplies by s, every
// other row squares, so row 16 outputs s^65537 mod n. The host computes each
// row's limbs, quotient and carries (big-integer division is sequential
// work); `expand_rsa_k` turns them into the 11 041 cells per row.
//
// Composition (eval_rsa_exp_multirow_per_row, rsa2048_exp_air.rs:210-285,
// eval_rsa_mul_gadget, rsa2048_field_air.rs:356-432) with lambda_j = j + 1:
// phi is linear in the slot weights, and field arithmetic is exact, so the
// 11 280 weighted constraints are summed in regrouped form. The 159
// position identities weigh sum_{i+j=k} a_i b_j by lambda_k, and lambda is
// affine in k, so
//   sum_k lambda_k sum_{i+j=k} a_i b_j = sum_i a_i ((L0 + i) B + B1),
//   B = sum_j b_j, B1 = sum_j j b_j,  L0 = lambda of position slot 0,
// and the carry terms collapse to sum_m sc_m (lambda_{m+1} - 2^26 lambda_m).
// That leaves one square per bit cell (booleanity) as the bulk of the work.
// Weighted terms with small weights are summed lazily in 128 bits and
// reduced once.
//
// Launch: a workgroup of 8 waves covers 64 consecutive LDE points; the lane
// is the point (every column read is a coalesced 512-byte run) and the wave
// is a slice of the limbs (i = slice mod 8) and carry positions; the slices
// exchange B, B1, N, N1 and their phi parts through LDS. 8 waves rather than
// 16 because 16-wave workgroups cap a lane at 128 VGPRs, which spills.
#include <string.h>

#include "air.hpp"
#include "bigint.hpp"
#include "gl.hpp"
#include "ntt.hpp"
#include "rsa_air.hpp"

namespace sg {

namespace rsa {

constexpr int NL = 80, LB = 26, CARRY_BITS = 40, CARRY_POS = 158, CROSS = 159, NW = 64;
constexpr uint64_t CARRY_OFF = 1ULL << 39;
constexpr uint32_t OUT_ROW = 16;
// columns of one record
enum : uint32_t {
    C_ACC = 0, C_S = 80, C_N = 160, C_PHASE = 240, C_OP2 = 241, C_EM = 321,
    C_C = 401, C_Q = 481, C_QBITS = 561, C_CBITS = 2641, C_CARRY = 4721
};
// constraint slots of one record
enum : uint32_t {
    K_PHASE = 0, K_SEL = 1, K_CBOOL = 81, K_CPACK = 2161, K_QBOOL = 2241, K_QPACK = 4321,
    K_POS = 4401, K_CARRY = 4560, K_EM = 10880, K_TRANS = 10960
};
static_assert(C_CARRY + CARRY_BITS * CARRY_POS == RSA_W, "column layout");
static_assert(K_TRANS + 4 * NL == RSA_K, "slot layout");

// Everything one row's cells are made of.
struct RowWords {
    uint32_t acc[NL], op2[NL], c[NL], q[NL];  // 26-bit limbs
    uint64_t carry[CARRY_POS];                 // carries biased by 2^39 (< 2^40)
    uint32_t phase, pad;
};
struct RecWords {
    uint32_t s[NL], n[NL], em[NL];
};

SG_HD uint64_t cell(uint32_t col, const RowWords &w, const RecWords &r) {
    if (col < C_S) return w.acc[col];
    if (col < C_N) return r.s[col - C_S];
    if (col < C_PHASE) return r.n[col - C_N];
    if (col == C_PHASE) return w.phase;
    if (col < C_EM) return w.op2[col - C_OP2];
    if (col < C_C) return r.em[col - C_EM];
    if (col < C_Q) return w.c[col - C_C];
    if (col < C_QBITS) return w.q[col - C_Q];
    if (col < C_CBITS) {
        const uint32_t j = col - C_QBITS;
        return (w.q[j / LB] >> (j % LB)) & 1;
    }
    if (col < C_CARRY) {
        const uint32_t j = col - C_CBITS;
        return (w.c[j / LB] >> (j % LB)) & 1;
    }
    const uint32_t j = col - C_CARRY;
    return (w.carry[j / CARRY_BITS] >> (j % CARRY_BITS)) & 1;
}

// ---------------------------------------------------------------- host rows

// little-endian 32-bit words of a 256-byte big-endian value
static void words_from_be(const uint8_t *be, uint32_t *w) {
    for (int i = 0; i < NW; i++)
        w[i] = (uint32_t)be[255 - 4 * i] | (uint32_t)be[254 - 4 * i] << 8 | (uint32_t)be[253 - 4 * i] << 16 |
               (uint32_t)be[252 - 4 * i] << 24;
}

// 26-bit limbs of a value held in nw words (bits above 2080 are not
// representable and must be zero, which a < n < 2^2048 guarantees)
static void to_limbs(const uint32_t *w, int nw, uint32_t *limbs) {
    for (int i = 0; i < NL; i++) {
        const int bit = LB * i, k = bit / 32, sh = bit % 32;
        const uint64_t lo = k < nw ? w[k] : 0, hi = k + 1 < nw ? w[k + 1] : 0;
        limbs[i] = (uint32_t)(((lo | hi << 32) >> sh) & ((1u << LB) - 1));
    }
}

static int top_words(const uint32_t *w, int nw) {
    while (nw > 0 && w[nw - 1] == 0) nw--;
    return nw;
}

static void divmod(const uint32_t *u, int ul, const uint32_t *v, int vl, uint32_t *q, uint32_t *r) {
    divmod_words<NW>(u, ul, v, vl, q, r);
}

// The words of one product row: c = a*b mod n, q = (a*b - c)/n, carries.
// Returns false if the limb identity does not close (cannot happen for
// a, b < n; checked so a bad input never yields an unsatisfiable trace).
static bool product_row(const uint32_t *a, const uint32_t *b, const uint32_t *n, int nl, RowWords &w,
                        uint32_t *c_words) {
    uint32_t prod[2 * NW] = {};
    for (int i = 0; i < NW; i++) {
        uint64_t carry = 0;
        for (int j = 0; j < NW; j++) {
            const uint64_t t = (uint64_t)a[i] * b[j] + prod[i + j] + carry;
            prod[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        prod[i + NW] = (uint32_t)carry;
    }
    uint32_t q[2 * NW + 1] = {}, r[NW] = {};
    const int pl = top_words(prod, 2 * NW);
    divmod(prod, pl ? pl : 1, n, nl, q, r);
    for (int i = pl - nl + 1; i < 2 * NW; i++)
        if (i >= NW && q[i]) return false;
    memcpy(c_words, r, 4 * nl);
    for (int i = nl; i < NW; i++) c_words[i] = 0;
    to_limbs(a, NW, w.acc);
    to_limbs(b, NW, w.op2);
    to_limbs(c_words, NW, w.c);
    to_limbs(q, NW, w.q);
    uint32_t nlimb[NL];
    to_limbs(n, NW, nlimb);
    int64_t prev = 0;
    for (int k = 0; k < CROSS; k++) {
        int64_t acc = prev - (k < NL ? (int64_t)w.c[k] : 0);
        const int lo = k > NL - 1 ? k - (NL - 1) : 0, hi = k < NL - 1 ? k : NL - 1;
        for (int i = lo; i <= hi; i++)
            acc += (int64_t)w.acc[i] * w.op2[k - i] - (int64_t)w.q[i] * nlimb[k - i];
        if (k == CROSS - 1) return acc == 0;
        if (acc & ((1 << LB) - 1)) return false;
        const int64_t cy = acc >> LB;  // exact multiple: the shift is the division
        if (cy >= (int64_t)CARRY_OFF || cy < -(int64_t)CARRY_OFF) return false;
        w.carry[k] = (uint64_t)(cy + (int64_t)CARRY_OFF);
        prev = cy;
    }
    return true;
}

// All h rows of one record; 0, or E_INVALID for s >= n or n = 0, or
// E_INTERNAL if a product row does not close.
static int record_rows(const RsaRecord &in, uint64_t h, RecWords &rec, RowWords *rows) {
    uint32_t n[NW], s[NW], em[NW], acc[NW], c[NW];
    words_from_be(in.n.data(), n);
    words_from_be(in.s.data(), s);
    words_from_be(in.em.data(), em);
    const int nl = top_words(n, NW);
    if (nl == 0) return E_INVALID;
    bool s_below_n = false;
    for (int i = NW - 1; i >= 0; i--)
        if (s[i] != n[i]) {
            s_below_n = s[i] < n[i];
            break;
        }
    if (!s_below_n) return E_INVALID;
    to_limbs(s, NW, rec.s);
    to_limbs(n, NW, rec.n);
    to_limbs(em, NW, rec.em);
    memcpy(acc, s, sizeof acc);
    for (uint64_t r = 0; r < h; r++) {
        RowWords &w = rows[r];
        w.phase = r == OUT_ROW;
        w.pad = 0;
        if (!product_row(acc, w.phase ? s : acc, n, nl, w, c)) return E_INTERNAL;
        memcpy(acc, c, sizeof acc);
    }
    return 0;
}

static void all_rows(const std::vector<RsaRecord> &records, uint64_t h, std::vector<RecWords> &recs,
                     std::vector<RowWords> &rows) {
    SG_REQUIRE(is_pow2(h) && h >= 32 && h <= (1u << 20), "rsa trace: height must be a power of two >= 32");
    const size_t G = records.size();
    recs.resize(G);
    rows.resize(G * h);
    std::vector<int> rc(G, 0);
    parallel_for(G, [&](size_t g) { rc[g] = record_rows(records[g], h, recs[g], rows.data() + g * h); });
    for (size_t g = 0; g < G; g++) {
        SG_REQUIRE(rc[g] != E_INVALID, "rsa trace: need 0 < n and s < n");
        if (rc[g]) throw Error(E_INTERNAL, "rsa trace: product identity does not close");
    }
}

// ---------------------------------------------------------------- device

// one cell per lane, rows fastest: tr_g[col * h + row], g = blockIdx.y
__global__ __launch_bounds__(256) void expand_rsa_k(const RowWords *rows, const RecWords *recs, uint64_t h,
                                                    uint64_t *tr) {
    const uint64_t g = blockIdx.y;
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= h * RSA_W) return;
    const uint64_t row = t & (h - 1), col = t / h;
    tr[g * RSA_W * h + t] = cell((uint32_t)col, rows[g * h + row], recs[g]);
}

constexpr int SL = 8;  // slices (waves) per workgroup
constexpr int PT = 64;  // points per workgroup

// phi of K proofs (grid.y) of R stacked records at n0 points
// gate groups: ungated, the output row's EM check, the transitions
enum { G_ALL, G_OUT, G_TRANS, NG };

// a += lambda_slot x for the 1-based slot index: slot itself (GEN = false,
// lambda_j = j + 1) or T[slot - 1] (GEN: the caller's coefficients)
template <bool GEN>
__device__ __forceinline__ void wmad(Acc &a, uint64_t x, uint32_t slot, const uint64_t *T) {
    if constexpr (GEN)
        acc_add(a, mul(x, T[slot - 1]));
    else
        acc_mad(a, x, slot);
}

// GROUPS = false: phi at every point. GROUPS = true: the three gate groups'
// weighted sums with the gates ignored, phi[(k * NG + g) * n0 + pt], on the
// n0 points of a blowup-2 LDE held as its even points (lde: the traces,
// n0 / 2 rows) and its odd points (odd, the same layout).
// GEN: slot j weighted by T[j] (the caller's combination coefficients)
// instead of j + 1, the position identities summed term by term.
// grid.z workgroups share a group of 64 points: wave s of workgroup z takes
// the limbs and carry positions = z * SL + s (mod SL * gridDim.z), so a
// small proof's few points still spread over many CUs; each workgroup writes
// its partial sums to its own plane of phi (plane stride gridDim.y * n0, or
// gridDim.y * NG * n0 with GROUPS) for sum_planes_k to add.
template <bool GROUPS, bool GEN>
__global__ __launch_bounds__(SL * PT) void compose_rsa_k(const uint64_t *lde, const uint64_t *odd, uint64_t n0,
                                                        uint64_t blowup, uint32_t R, uint64_t *phi,
                                                        const uint64_t *T) {
    __shared__ uint64_t red[SL][4][PT];
    const uint32_t lane = threadIdx.x & (PT - 1), s = threadIdx.x / PT;
    const uint32_t sg = blockIdx.z * SL + s, ST = SL * gridDim.z;  // this wave's slice of ST
    phi += (uint64_t)blockIdx.z * gridDim.y * (GROUPS ? NG : 1) * n0;
    const uint64_t pt = (uint64_t)blockIdx.x * PT + lane;
    // GROUPS: point pt is row pt / 2 of the traces (even pt) or of the odd coset
    const uint64_t rows = GROUPS ? n0 / 2 : n0, at = GROUPS ? pt >> 1 : pt;
    const uint64_t np = GROUPS ? (at + 1) & (rows - 1) : (pt + blowup) & (n0 - 1);
    const uint64_t *base = GROUPS && (pt & 1) ? odd : lde;
    const uint64_t n_trace = n0 / blowup, trow = pt / blowup;
    const bool out_row = GROUPS || trow == OUT_ROW, trans = GROUPS || trow + 1 < n_trace;
    uint64_t out_lane = 0, trans_lane = 0;  // GROUPS: the EM and transition sums
    uint64_t phi_lane = 0;
    for (uint32_t rec = 0; rec < R; rec++) {
        const uint64_t *col = base + ((uint64_t)blockIdx.y * R + rec) * RSA_W * rows;
        const uint32_t O = rec * RSA_K + 1;  // lambda(slot) = O + slot
#define AT(c) col[(uint64_t)(c) * rows + at]
#define NX(c) col[(uint64_t)(c) * rows + np]
        // B, B1 (op2) and N, N1 (modulus) over every limb
        {
            uint64_t b0 = 0, b1 = 0, m0 = 0, m1 = 0;
#pragma unroll
            for (int u = 0; u < NL / SL; u++) {
                const uint32_t i = s + SL * u;
                const uint64_t op = AT(C_OP2 + i), nn = AT(C_N + i);
                b0 = add(b0, op);
                b1 = add(b1, mul_small(op, i));
                m0 = add(m0, nn);
                m1 = add(m1, mul_small(nn, i));
            }
            red[s][0][lane] = b0;
            red[s][1][lane] = b1;
            red[s][2][lane] = m0;
            red[s][3][lane] = m1;
        }
        __syncthreads();
        uint64_t B = 0, B1 = 0, N = 0, N1 = 0;
#pragma unroll
        for (int t = 0; t < SL; t++) {
            B = add(B, red[t][0][lane]);
            B1 = add(B1, red[t][1][lane]);
            N = add(N, red[t][2][lane]);
            N1 = add(N1, red[t][3][lane]);
        }
        __syncthreads();  // red is written again for the next record
        Acc pos{0, 0}, neg{0, 0}, sel{0, 0};
        Acc epos{0, 0}, eneg{0, 0}, tpos{0, 0}, tneg{0, 0};  // GROUPS: EM and transition sums
        Acc &ep = GROUPS ? epos : pos, &en = GROUPS ? eneg : neg, &tp = GROUPS ? tpos : pos, &tn = GROUPS ? tneg : neg;
        const uint32_t L0 = O + K_POS;
#pragma unroll 1
        for (uint32_t i = sg; i < NL; i += ST) {
            const uint64_t a = AT(C_ACC + i), op = AT(C_OP2 + i), sv = AT(C_S + i);
            const uint64_t cc = AT(C_C + i), qq = AT(C_Q + i);
            // select: op - (ph (s - a) + a); the ph part is applied once below
            wmad<GEN>(pos, op, O + K_SEL + i, T);
            wmad<GEN>(neg, a, O + K_SEL + i, T);
            wmad<GEN>(sel, sub(sv, a), O + K_SEL + i, T);
            // position identities, regrouped by the a and q limbs
            if constexpr (GEN) {
                // sum_j lambda_{L0 + i + j} op_j and n_j, term by term
                uint64_t ab = 0, qn = 0;
#pragma unroll 4
                for (int j = 0; j < NL; j++) {
                    const uint64_t lam = T[L0 + i + j - 1];
                    ab = add(ab, mul(AT(C_OP2 + j), lam));
                    qn = add(qn, mul(AT(C_N + j), lam));
                }
                acc_add(pos, mul(a, ab));
                acc_add(neg, mul(qq, qn));
            } else {
                acc_add(pos, mul(a, add(mul_small(B, L0 + i), B1)));
                acc_add(neg, mul(qq, add(mul_small(N, L0 + i), N1)));
            }
            wmad<GEN>(neg, cc, L0 + i, T);
            // range checks of c and q: booleanity and limb packing
            Acc pk{0, 0};
#pragma unroll
            for (int b = 0; b < LB; b++) {
                const uint64_t x = AT(C_CBITS + LB * i + b);
                wmad<GEN>(pos, x, O + K_CBOOL + LB * i + b, T);
                wmad<GEN>(neg, mul(x, x), O + K_CBOOL + LB * i + b, T);
                acc_shl(pk, x, b);
            }
            wmad<GEN>(pos, cc, O + K_CPACK + i, T);
            wmad<GEN>(neg, acc_val(pk), O + K_CPACK + i, T);
            pk = Acc{0, 0};
#pragma unroll
            for (int b = 0; b < LB; b++) {
                const uint64_t x = AT(C_QBITS + LB * i + b);
                wmad<GEN>(pos, x, O + K_QBOOL + LB * i + b, T);
                wmad<GEN>(neg, mul(x, x), O + K_QBOOL + LB * i + b, T);
                acc_shl(pk, x, b);
            }
            wmad<GEN>(pos, qq, O + K_QPACK + i, T);
            wmad<GEN>(neg, acc_val(pk), O + K_QPACK + i, T);
            if (out_row) {
                wmad<GEN>(ep, cc, O + K_EM + i, T);
                wmad<GEN>(en, AT(C_EM + i), O + K_EM + i, T);
            }
        }
        // transitions into the next row (a loop of its own keeps the bit
        // loops' registers free)
        if (trans) {
#pragma unroll 1
            for (uint32_t i = sg; i < NL; i += ST) {
                wmad<GEN>(tp, NX(C_ACC + i), O + K_TRANS + i, T);
                wmad<GEN>(tn, AT(C_C + i), O + K_TRANS + i, T);
                wmad<GEN>(tp, NX(C_S + i), O + K_TRANS + NL + i, T);
                wmad<GEN>(tn, AT(C_S + i), O + K_TRANS + NL + i, T);
                wmad<GEN>(tp, NX(C_N + i), O + K_TRANS + 2 * NL + i, T);
                wmad<GEN>(tn, AT(C_N + i), O + K_TRANS + 2 * NL + i, T);
                wmad<GEN>(tp, NX(C_EM + i), O + K_TRANS + 3 * NL + i, T);
                wmad<GEN>(tn, AT(C_EM + i), O + K_TRANS + 3 * NL + i, T);
            }
        }
        // carries: booleanity, and sc_m (lambda_{m+1} - 2^26 lambda_m) from the
        // position identities
#pragma unroll 1
        for (uint32_t m = sg; m < CARRY_POS; m += ST) {
            Acc sc{0, 0};
#pragma unroll
            for (int b = 0; b < CARRY_BITS; b++) {
                const uint64_t y = AT(C_CARRY + CARRY_BITS * m + b);
                wmad<GEN>(pos, y, O + K_CARRY + CARRY_BITS * m + b, T);
                wmad<GEN>(neg, mul(y, y), O + K_CARRY + CARRY_BITS * m + b, T);
                acc_shl(sc, y, b);
            }
            uint64_t gamma;  // the signed carry of position m: carry-in of m + 1, -2^26 carry-out of m
            if constexpr (GEN)
                gamma = sub(T[L0 + m], mul(T[L0 + m - 1], 1ULL << LB));
            else
                gamma = sub(L0 + m + 1, (uint64_t)(L0 + m) << LB);
            acc_add(pos, mul(sub(acc_val(sc), CARRY_OFF), gamma));
        }
        const uint64_t ph = AT(C_PHASE);
        if (sg == 0) {
            wmad<GEN>(pos, ph, O + K_PHASE, T);
            wmad<GEN>(neg, mul(ph, ph), O + K_PHASE, T);
        }
#undef AT
#undef NX
        phi_lane = add(phi_lane, sub(sub(acc_val(pos), acc_val(neg)), mul(ph, acc_val(sel))));
        if (GROUPS) {
            out_lane = add(out_lane, sub(acc_val(epos), acc_val(eneg)));
            trans_lane = add(trans_lane, sub(acc_val(tpos), acc_val(tneg)));
        }
    }
    red[s][0][lane] = phi_lane;
    red[s][1][lane] = out_lane;
    red[s][2][lane] = trans_lane;
    __syncthreads();
    if (s == 0) {
        uint64_t v = 0, vo = 0, vt = 0;
#pragma unroll
        for (int t = 0; t < SL; t++) {
            v = add(v, red[t][0][lane]);
            vo = add(vo, red[t][1][lane]);
            vt = add(vt, red[t][2][lane]);
        }
        if (GROUPS) {
            uint64_t *o = phi + (uint64_t)blockIdx.y * NG * n0 + pt;
            o[G_ALL * n0] = v;
            o[G_OUT * n0] = vo;
            o[G_TRANS * n0] = vt;
        } else {
            phi[(uint64_t)blockIdx.y * n0 + pt] = v;
        }
    }
}

// phi at every point from the interpolated gate-group polynomials
// pg[(k * NG + g) * n + i]: the EM check on the output row, the transitions
// on every row but the last
__global__ __launch_bounds__(256) void combine_rsa_k(const uint64_t *pg, uint64_t n, uint64_t blowup,
                                                     uint64_t *phi) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint64_t trow = i / blowup, n_trace = n / blowup;
    const uint64_t *q = pg + (uint64_t)blockIdx.y * NG * n + i;
    uint64_t v = q[G_ALL * n];
    if (trow == OUT_ROW) v = add(v, q[G_OUT * n]);
    if (trow + 1 < n_trace) v = add(v, q[G_TRANS * n]);
    phi[n * blockIdx.y + i] = v;
}

// dst[j] = sum of the nz planes src[z * len + j]
__global__ __launch_bounds__(256) void sum_planes_k(const uint64_t *src, uint32_t nz, uint64_t len, uint64_t *dst) {
    const uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= len) return;
    uint64_t v = src[j];
    for (uint32_t z = 1; z < nz; z++) v = add(v, src[(uint64_t)z * len + j]);
    dst[j] = v;
}

// workgroups per group of 64 points (grid.z of compose_rsa_k): enough for a
// small batch to reach many CUs, one when the batch fills the chip itself
inline uint32_t slice_blocks(uint64_t groups) { return groups >= 512 ? 1 : groups >= 64 ? 2 : 10; }

}  // namespace rsa

void build_rsa_trace_host(const std::vector<RsaRecord> &records, uint64_t h, uint64_t *tr) {
    std::vector<rsa::RecWords> recs;
    std::vector<rsa::RowWords> rows;
    rsa::all_rows(records, h, recs, rows);
    const size_t G = records.size();
    parallel_for(G, [&](size_t g) {
        uint64_t *t = tr + g * RSA_W * h;
        for (uint32_t col = 0; col < RSA_W; col++)
            for (uint64_t r = 0; r < h; r++) t[col * h + r] = rsa::cell(col, rows[g * h + r], recs[g]);
    });
}

void build_rsa_traces_dev(Ctx *c, const std::vector<RsaRecord> &records, uint32_t R, uint64_t h, uint64_t *d_tr) {
    const size_t G = records.size();
    SG_REQUIRE(G >= 1 && R >= 1 && G % R == 0 && G <= 65535, "rsa traces: records must fill whole proofs");
    std::vector<rsa::RecWords> recs;
    std::vector<rsa::RowWords> rows;
    rsa::all_rows(records, h, recs, rows);
    Scoped d_rows(c, sizeof(rsa::RowWords) * rows.size()), d_recs(c, sizeof(rsa::RecWords) * G);
    SG_HIP(hipMemcpyAsync(d_rows.p, rows.data(), sizeof(rsa::RowWords) * rows.size(), hipMemcpyHostToDevice,
                          c->stream));
    SG_HIP(hipMemcpyAsync(d_recs.p, recs.data(), sizeof(rsa::RecWords) * G, hipMemcpyHostToDevice, c->stream));
    {
        ProfScope ps(c, "expand_rsa", 8.0 * RSA_W * h * G);
        hipLaunchKernelGGL(rsa::expand_rsa_k, dim3((uint32_t)((h * RSA_W + 255) / 256), (uint32_t)G), dim3(256), 0,
                           c->stream, d_rows.as<rsa::RowWords>(), d_recs.as<rsa::RecWords>(), h, d_tr);
        SG_HIP(hipGetLastError());
    }
    SG_HIP(hipStreamSynchronize(c->stream));  // the host word buffers go out of scope
}

void merge_rsa_batch(Ctx *c, const uint64_t *lde, uint32_t K, uint32_t R, uint64_t n_trace, uint64_t blowup,
                     uint64_t *c_eval, const uint64_t *d_lam) {
    SG_REQUIRE(is_pow2(n_trace) && n_trace >= 32 && is_pow2(blowup) && blowup >= 2,
               "merge_rsa: n_trace must be a power of two >= 32 and blowup >= 2");
    SG_REQUIRE(K >= 1 && K <= 65535, "merge_rsa: 1 to 65535 proofs");
    SG_REQUIRE(R >= 1 && R <= RSA_MAX_RECORDS, "merge_rsa: 1 to 4096 records per proof");
    const uint64_t n = n_trace * blowup;
    SG_REQUIRE(n <= (1ULL << 24), "merge_rsa: domain above 2^24");
    Scoped phi(c, 8 * n * K);
    {
        ProfScope ps(c, "compose_rsa", (8.0 * RSA_W * n * R + 8.0 * n) * K);
        const uint32_t nz = rsa::slice_blocks(n / rsa::PT * K);
        Scoped parts(c, nz > 1 ? 8 * n * K * nz : 8);
        uint64_t *dst = nz > 1 ? parts.as<uint64_t>() : phi.as<uint64_t>();
        const dim3 grid((uint32_t)(n / rsa::PT), K, nz);
        if (d_lam)
            hipLaunchKernelGGL((rsa::compose_rsa_k<false, true>), grid, dim3(rsa::SL * rsa::PT), 0, c->stream, lde,
                               (const uint64_t *)nullptr, n, blowup, R, dst, d_lam);
        else
            hipLaunchKernelGGL((rsa::compose_rsa_k<false, false>), grid, dim3(rsa::SL * rsa::PT), 0, c->stream, lde,
                               (const uint64_t *)nullptr, n, blowup, R, dst, (const uint64_t *)nullptr);
        SG_HIP(hipGetLastError());
        if (nz > 1) {
            hipLaunchKernelGGL(rsa::sum_planes_k, dim3((uint32_t)((n * K + 255) / 256)), dim3(256), 0, c->stream,
                               parts.as<uint64_t>(), nz, n * K, phi.as<uint64_t>());
            SG_HIP(hipGetLastError());
        }
    }
    quotient_batch(c, phi.as<uint64_t>(), K, n_trace, n, c_eval);
}

void merge_rsa_from_trace_batch(Ctx *c, const uint64_t *d_tr, uint32_t K, uint32_t R, uint64_t n_trace,
                                uint64_t blowup, uint64_t *c_eval) {
    SG_REQUIRE(is_pow2(n_trace) && n_trace >= 32 && is_pow2(blowup) && blowup >= 4,
               "merge_rsa: n_trace must be a power of two >= 32 and blowup >= 4");
    SG_REQUIRE(K >= 1 && K <= 65535, "merge_rsa: 1 to 65535 proofs");
    SG_REQUIRE(R >= 1 && R <= RSA_MAX_RECORDS, "merge_rsa: 1 to 4096 records per proof");
    const uint64_t n = n_trace * blowup, n2 = 2 * n_trace;
    SG_REQUIRE(n <= (1ULL << 24), "merge_rsa: domain above 2^24");
    Scoped phi(c, 8 * n * K);
    {
        Scoped groups(c, 8ULL * rsa::NG * n2 * K);
        {
            // the blowup-2 LDE as the traces (its even points) and the odd coset
            Scoped odd(c, 8ULL * RSA_W * R * n_trace * K);
            lde_odd_coset(c, d_tr, odd.as<uint64_t>(), ilog2(n_trace), (uint64_t)RSA_W * R * K);
            ProfScope ps(c, "compose_rsa", (8.0 * RSA_W * n2 * R + 8.0 * rsa::NG * n2) * K);
            const uint32_t nz = rsa::slice_blocks(n2 / rsa::PT * K);
            Scoped parts(c, nz > 1 ? 8ULL * rsa::NG * n2 * K * nz : 8);
            hipLaunchKernelGGL((rsa::compose_rsa_k<true, false>), dim3((uint32_t)(n2 / rsa::PT), K, nz),
                               dim3(rsa::SL * rsa::PT), 0, c->stream, d_tr, odd.as<uint64_t>(), n2, (uint64_t)2, R,
                               nz > 1 ? parts.as<uint64_t>() : groups.as<uint64_t>(), (const uint64_t *)nullptr);
            SG_HIP(hipGetLastError());
            if (nz > 1) {
                const uint64_t len = (uint64_t)rsa::NG * n2 * K;
                hipLaunchKernelGGL(rsa::sum_planes_k, dim3((uint32_t)((len + 255) / 256)), dim3(256), 0, c->stream,
                                   parts.as<uint64_t>(), nz, len, groups.as<uint64_t>());
                SG_HIP(hipGetLastError());
            }
        }
        Scoped full(c, 8ULL * rsa::NG * n * K);
        lde_columns(c, groups.as<uint64_t>(), full.as<uint64_t>(), ilog2(n2), ilog2(blowup / 2),
                    (uint64_t)rsa::NG * K, n2, n, 1);
        ProfScope ps(c, "combine_rsa", 8.0 * (rsa::NG + 1) * n * K);
        hipLaunchKernelGGL(rsa::combine_rsa_k, dim3((uint32_t)((n + 255) / 256), K), dim3(256), 0, c->stream,
                           full.as<uint64_t>(), n, blowup, phi.as<uint64_t>());
        SG_HIP(hipGetLastError());
    }
    quotient_batch(c, phi.as<uint64_t>(), K, n_trace, n, c_eval);
}

void rsa_pkcs1_em_sha256(const uint8_t digest[32], uint8_t em[256]) {
    // RFC 8017 9.2: 00 01 PS(FF..) 00 || DigestInfo(SHA-256) || H
    static const uint8_t info[19] = {0x30, 0x31, 0x30, 0x0D, 0x06, 0x09, 0x60, 0x86, 0x48, 0x01,
                                     0x65, 0x03, 0x04, 0x02, 0x01, 0x05, 0x00, 0x04, 0x20};
    const size_t ps = 256 - 3 - 19 - 32;
    em[0] = 0x00;
    em[1] = 0x01;
    memset(em + 2, 0xFF, ps);
    em[2 + ps] = 0x00;
    memcpy(em + 3 + ps, info, 19);
    memcpy(em + 256 - 32, digest, 32);
}

bool rsa_record_verifies(const RsaRecord &in) {
    std::vector<rsa::RowWords> rows(32);
    rsa::RecWords rec;
    if (rsa::record_rows(in, 32, rec, rows.data()) != 0) return false;
    return memcmp(rows[rsa::OUT_ROW].c, rec.em, sizeof rec.em) == 0;
}

}  // namespace sg
