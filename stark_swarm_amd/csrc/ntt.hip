// ntt.hip -- batched Goldilocks NTT / iNTT and (coset) low-degree extension
// for gfx950.
//
// Replaces ark-poly's Radix2EvaluationDomain::{fft, ifft} and
// lde_trace_columns (crates/deep_ali/src/trace_import.rs:251-281): natural
// order in and out, forward transform X[k] = sum_j x[j] w^(jk) with
// w = 7^((p-1)/n), inverse scaled by 1/n.
//
// A DFT of size n = n1*n2 is two strided passes (four-step):
//   pass A  Z[k1*n2+i2] = w_n^(i2*k1) * DFT_n1( x[i1*n2+i2] )     (columns)
//   pass B  X[k1+n1*k2] = DFT_n2( Z[k1*n2+i2] )                    (rows)
// Each pass is one launch of `ntt_tile<LOG_S, INV>`: a 512-lane workgroup owns
// a tile of G sub-DFTs of size S = 2^LOG_S (S*G = 8192 u64, 70 KiB of padded
// LDS, two workgroups per CU). The tile is loaded with its contiguous
// dimension on consecutive lanes (coalesced runs of G u64 or whole rows), the
// sub-DFT runs as Stockham steps of radix <= 16 held in VGPRs (16 elements
// per lane), and the result is stored the same way, fusing input
// pre-twiddles, the four-step twiddle and output scaling.
//
// Inside a register DFT every twiddle is a power of w_32 = 2^78 (mod p), so
// it is a shift plus one 128->64 reduction instead of a full 64x64 product.
// The LDS index is padded by one u64 per 16 and skewed per sub-DFT (padi) to
// break the bank conflicts of the strided Stockham writes and the g-fast
// loads and stores.
//
// The LDE of a column with blowup B = 2^LB uses
//   LDE[B*k + r] = DFT_n( c_j * w_{Bn}^(r*j) )[k],  r < B,
// so pass A reads each coefficient tile once and emits the B pre-twiddled
// column DFTs (`nr` repetitions), and pass B scatters row DFTs of every r
// into the interleaved output. (Taking the four-step twiddle at pass B's
// input instead, with pass A's pre-twiddle reduced to w_{B n1}^(r*i1),
// measured slower on the 2^20 x 128 LDE: pass B's extra products and
// register spills cost more than pass A saved.)
//
// Passes without any twiddle or scale (the LDE's and the forward NTT's
// second pass) run a PLAIN instantiation with those paths compiled out:
// fewer live registers than the general kernel at the same 128-VGPR budget.
#include <algorithm>
#include <map>
#include <mutex>

#include "gl.hpp"
#include "ctx.hpp"
#include "ntt.hpp"
#include "ntt_dev.hpp"

namespace sg {

namespace {

// Compile-time forms of the pass: MODE_PLAIN has no pre-twiddle, post
// twiddle or scaling, so it needs fewer live registers than the general
// kernel at the same 128-VGPR budget (the 1024-point pass spills 16 bytes a
// lane instead of 72: lde_passB 10.3 -> 9.7 ms on the 2^20 x 128 LDE). A
// third form with the gather / scatter paths compiled out for passes that
// read and write g-fast spilled 64 bytes and ran no faster, so it is not
// kept.
constexpr int MODE_GENERAL = 0, MODE_PLAIN = 1;
// MODE_LDEA: the LDE's first pass, whose shape launch_pass checks: sub-DFTs
// g-fast on both sides with one lane per (g, i) group, a pre-twiddle read
// from the w_{2^13} table, a post twiddle, lazily reduced output and no
// scaling. Every other path is compiled out.
constexpr int MODE_LDEA = 2;
// MODE_POST: the first pass of a two-pass NTT or iNTT: g-fast on both sides,
// no pre-twiddle, a post twiddle (the iNTT's 1/n rides on it), lazily reduced
// output and no shift powers.
constexpr int MODE_POST = 3;

template <int LOG_S, bool INV, int MODE, int TLOG>
// 16 elements a lane, at most 128 VGPRs: 512 lanes with two blocks per CU or
// 1024 lanes with one, 4 waves per SIMD either way, which hides more of the
// 64-bit multiply latency than 256 x 32 at 2 waves per SIMD (about 5 %
// faster on the 2^20 x 128 LDE)
__global__ __launch_bounds__(TileCfg<TLOG>::LANES, 4) void ntt_tile(NttPass p) {
    constexpr int LANES = TileCfg<TLOG>::LANES, PER_LANE = TileCfg<TLOG>::PER_LANE;
    if constexpr (MODE == MODE_PLAIN) {
        // no pre-twiddle, post twiddle or scaling: constants the compiler
        // folds, so those paths and their registers drop out
        p.pre_log = 0;
        p.post_log = 0;
        p.scale = 1;
        p.sp_lo = nullptr;
        p.lazy_out = 0;
    }
    if constexpr (MODE == MODE_LDEA) {
        p.pre_direct = 1;
        p.scale = 1;
        p.sp_lo = nullptr;
        p.lazy_out = 1;
        p.inv = 0;
    }
    if constexpr (MODE == MODE_POST) {
        p.pre_log = 0;
        p.sp_lo = nullptr;
        p.lazy_out = 1;
        p.inv = INV;
    }
    __shared__ uint64_t lds[TileCfg<TLOG>::LDS];
    constexpr uint32_t S = 1u << LOG_S;
    const uint32_t log_g = p.log_g, G = 1u << log_g;
    const uint32_t tile = S << log_g;
    const uint64_t tiles_per_batch = (p.ng + G - 1) >> log_g;
    // Workgroups are dealt round-robin to the 8 XCDs; give each XCD a
    // contiguous range of tiles so tiles that share 128-byte lines (runs of
    // G u64 per row) meet in the same L2.
    uint64_t blk = blockIdx.x;
    if ((gridDim.x & 7) == 0) blk = (blk & 7) * (gridDim.x >> 3) + (blk >> 3);
    // the repetition chunks of one tile are neighbours on one XCD, so all but
    // the first read the tile's input from L2
    const uint32_t rbase = (uint32_t)(blk & ((1u << p.rs_log) - 1)) * p.nr;
    blk >>= p.rs_log;
    const uint64_t b = blk / tiles_per_batch;
    const uint64_t g0 = (blk - b * tiles_per_batch) << log_g;
    const uint64_t b_hi = b / p.nb2, b_lo = b - b_hi * p.nb2;
    const uint64_t *in = p.in + b_hi * p.ib + b_lo * p.ib2 + g0 * p.ig;
    const bool in_gfast = log_g > 0 && p.ig < p.ii;
    const bool out_gfast = log_g > 0 && p.og < p.ok;
    // g-fast lane mapping: element u of a lane is (g = gl, i = il + u*di), so
    // the lane's twiddles form geometric sequences in u (no table gathers)
    const bool inc_in = MODE == MODE_LDEA || MODE == MODE_POST || (in_gfast && (1u << log_g) <= (uint32_t)LANES);
    const bool inc_out = MODE == MODE_LDEA || MODE == MODE_POST || (out_gfast && (1u << log_g) <= (uint32_t)LANES);
    const uint32_t gl = threadIdx.x & (G - 1);
    const uint32_t il = threadIdx.x >> log_g;
    const uint32_t di = LANES >> log_g;
    const bool gok = g0 + gl < p.ng;

    uint64_t *out0 = p.out + b_hi * p.ob + b_lo * p.ob2 + g0 * p.og + (uint64_t)rbase * p.out_rs;
    for (uint32_t rep = 0; rep < p.nr; rep++) {
        // pre-twiddle w^(r * (g*pre_g + i*pre_i + b*pre_b)), r = r0 + rbase + rep
        const uint64_t rr = p.r0 + rbase + rep;
        const bool pre = (MODE == MODE_LDEA || p.pre_log) && rr;
        // the lane index through an empty asm: the compiler cannot see that
        // it is loop-invariant, so the per-element addresses of the fully
        // unrolled load loops below are formed inside the loop rather than
        // hoisted out of it as 16 live 64-bit values that spill to scratch
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        if (inc_in) {
            // pre factors along the lane's elements: a geometric sequence
            // (one product per step), or with pre_direct one table read per
            // element at exponent ed (mod 2^pre_log, kept in 32 bits)
            const bool pre_inc = pre && !p.pre_direct;
            uint64_t w = 1, st = 1;
            uint32_t ed = 0, es = 0;
            if (pre) {
                const uint64_t e0 = (g0 + gl) * p.pre_g + (uint64_t)il * p.pre_i + b * p.pre_b;
                if (p.pre_direct) {
                    ed = (uint32_t)(rr * e0);
                    es = (uint32_t)(rr * di * p.pre_i);
                } else {
                    w = root_pow(p.rt_lo, p.rt_hi, p.pre_log, rr * e0, false);
                    st = root_pow(p.rt_lo, p.rt_hi, p.pre_log, rr * di * p.pre_i, false);
                }
            }
            // every load of the lane (and its table twiddle) is issued before
            // the first is used, so the tile pays one memory latency
            uint64_t xv[PER_LANE], wv[PER_LANE];
            const uint32_t glr = tid & (G - 1), ilr = tid >> log_g;
            const uint64_t *src = in + (int64_t)glr * p.ig + (int64_t)ilr * p.ii;
            const int64_t sstep = (int64_t)di * p.ii;
#pragma unroll
            for (int u = 0; u < PER_LANE; u++) {
                const uint32_t i = ilr + u * di;
                xv[u] = (i < S && gok) ? src[u * sstep] : 0;
                if (pre && p.pre_direct) wv[u] = direct_root(p, ed + (uint32_t)u * es);
            }
#pragma unroll
            for (int u = 0; u < PER_LANE; u++) {
                const uint32_t i = ilr + u * di;
                if (i < S) {
                    uint64_t x = xv[u];
                    // the Stockham steps take lazily reduced values
                    if (pre) x = mul_lazy(x, p.pre_direct ? wv[u] : w);
                    lds[padi<LOG_S>((glr << LOG_S) + i)] = x;
                }
                if (pre_inc) w = mul_lazy(w, st);
            }
        } else {
            // the loads of CH elements and their twiddle table reads are
            // issued together, then used
            constexpr int CH = 8;
#pragma unroll
            for (int u0 = 0; u0 < PER_LANE; u0 += CH) {
                uint64_t xv[CH], ta[CH], tb[CH];
#pragma unroll
                for (int c = 0; c < CH; c++) {
                    const uint32_t e = tid + (u0 + c) * LANES;
                    const uint32_t g = in_gfast ? (e & (G - 1)) : (e >> LOG_S);
                    const uint32_t i = in_gfast ? (e >> log_g) : (e & (S - 1));
                    const bool ok = e < tile && g0 + g < p.ng;
                    xv[c] = ok ? in[(int64_t)g * p.ig + (int64_t)i * p.ii] : 0;
                    ta[c] = tb[c] = 1;
                    if (pre && ok) {
                        const uint64_t ee = (g0 + g) * p.pre_g + (uint64_t)i * p.pre_i + b * p.pre_b;
                        if (p.pre_direct) {
                            ta[c] = direct_root(p, (uint32_t)(ee * rr));
                        } else {
                            const uint32_t e26 = root_index(p.pre_log, ee * rr, false);
                            ta[c] = p.rt_lo[e26 & 8191];
                            tb[c] = p.rt_hi[e26 >> 13];
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < CH; c++) {
                    const uint32_t e = tid + (u0 + c) * LANES;
                    if (e >= tile) continue;
                    const uint32_t g = in_gfast ? (e & (G - 1)) : (e >> LOG_S);
                    const uint32_t i = in_gfast ? (e >> log_g) : (e & (S - 1));
                    uint64_t x = xv[c];
                    if (pre && g0 + g < p.ng) x = mul_lazy(x, p.pre_direct ? ta[c] : mul(ta[c], tb[c]));
                    lds[padi<LOG_S>((g << LOG_S) + i)] = x;
                }
            }
        }
        __syncthreads();
        stockham_steps<LOG_S, 0, INV, TLOG>(lds, tile, p.tw);
        uint64_t *out = out0 + (uint64_t)rep * p.out_rs;
        if (inc_out) {
            // post twiddle w^(+-g*k) and output scale * s^(g*sp_g + k*sp_k);
            // with a post twiddle the scale is folded into its first factor
            uint64_t pw = 1, pst = 1, sw = p.scale, sst = 1;
            if (p.post_log) {
                const uint64_t gg = g0 + gl;
                pw = root_pow(p.rt_lo, p.rt_hi, p.post_log, gg * (il * p.post_mul + rr * p.post_rep), p.inv);
                pst = root_pow(p.rt_lo, p.rt_hi, p.post_log, gg * di * p.post_mul, p.inv);
                if (p.scale != 1) pw = mul(pw, p.scale);
                sw = 1;
            }
            if (p.sp_lo) {
                const uint64_t e0 = (g0 + gl) * p.sp_g + (uint64_t)il * p.sp_k, ed = (uint64_t)di * p.sp_k;
                sw = mul(sw, mul(p.sp_lo[e0 & 8191], p.sp_hi[e0 >> 13]));
                sst = mul(p.sp_lo[ed & 8191], p.sp_hi[ed >> 13]);
            }
            const bool sc = p.sp_lo || (p.scale != 1 && !p.post_log);
#pragma unroll 2
            for (int u = 0; u < PER_LANE; u++) {
                const uint32_t k = il + u * di;
                if (k < S && gok) {
                    uint64_t x = lds[padi<LOG_S>((gl << LOG_S) + k)];
                    if (p.lazy_out) {
                        if (p.post_log) x = mul_lazy(x, pw);
                        if (sc) x = mul_lazy(x, sw);
                    } else {
                        if (p.post_log) x = mul(x, pw);
                        if (sc) x = mul(x, sw);
                        if (!p.post_log && !sc) x = canon(x);
                    }
                    out[(int64_t)gl * p.og + (int64_t)k * p.ok] = x;
                }
                if (p.post_log) pw = mul_lazy(pw, pst);
                if (p.sp_lo) sw = mul_lazy(sw, sst);
            }
        } else {
            // the table reads of CH elements are issued together, then used
            constexpr int CH = 4;
#pragma unroll
            for (int u0 = 0; u0 < PER_LANE; u0 += CH) {
                uint64_t ra[CH], rb[CH], sa[CH], sb[CH];
#pragma unroll
                for (int c = 0; c < CH; c++) {
                    const uint32_t e = tid + (u0 + c) * LANES;
                    const uint32_t g = out_gfast ? (e & (G - 1)) : (e >> LOG_S);
                    const uint32_t k = out_gfast ? (e >> log_g) : (e & (S - 1));
                    const bool ok = e < tile && g0 + g < p.ng;
                    ra[c] = rb[c] = sa[c] = sb[c] = 1;
                    if (p.post_log && ok) {
                        const uint32_t e26 = root_index(p.post_log, (g0 + g) * (k * p.post_mul + rr * p.post_rep),
                                                        p.inv);
                        ra[c] = p.rt_lo[e26 & 8191];
                        rb[c] = p.rt_hi[e26 >> 13];
                    }
                    if (p.sp_lo && ok) {
                        const uint64_t ee = (g0 + g) * p.sp_g + (uint64_t)k * p.sp_k;
                        sa[c] = p.sp_lo[ee & 8191];
                        sb[c] = p.sp_hi[ee >> 13];
                    }
                }
#pragma unroll
                for (int c = 0; c < CH; c++) {
                    const uint32_t e = tid + (u0 + c) * LANES;
                    const uint32_t g = out_gfast ? (e & (G - 1)) : (e >> LOG_S);
                    const uint32_t k = out_gfast ? (e >> log_g) : (e & (S - 1));
                    if (e >= tile || g0 + g >= p.ng) continue;
                    uint64_t x = lds[padi<LOG_S>((g << LOG_S) + k)];
                    if (p.post_log) x = mul(x, mul(ra[c], rb[c]));
                    if (p.scale != 1) x = mul(x, p.scale);
                    if (p.sp_lo) x = mul(x, mul(sa[c], sb[c]));
                    if (!p.lazy_out && !p.post_log && p.scale == 1 && !p.sp_lo) x = canon(x);
                    out[(int64_t)g * p.og + (int64_t)k * p.ok] = x;
                }
            }
        }
        __syncthreads();
    }
}

// Small LDE (B * 2^LOG_S <= 8192 points per column): one workgroup holds
// all B cosets of 8192 / (B * 2^LOG_S) columns, sub-DFT g = (column, r) with
// r fastest, so the interleaved output LDE[B*k + r] of a column leaves in
// runs of B consecutive u64 (the repetition form of ntt_tile writes every
// B-th element instead). Input c_i of coset r is pre-twiddled by
// w_{B n}^(r i); coefficients already carry 1/n and any shift^i.
template <int LOG_S>
__global__ __launch_bounds__(NT, 4) void lde_small_tile(const uint64_t *coef, uint64_t *out, uint64_t ncols,
                                                       uint32_t log_b, uint64_t in_cs, uint64_t out_cs,
                                                       const uint64_t *rt_lo, const uint64_t *rt_hi,
                                                       const uint64_t *tw) {
    __shared__ uint64_t lds[LDS_ELEMS];
    constexpr uint32_t S = 1u << LOG_S;
    constexpr uint32_t LOG_G = TILE_LOG - LOG_S, G = 1u << LOG_G;
    static_assert(G <= (uint32_t)NT, "one sub-DFT per lane at least");
    constexpr uint32_t DI = NT >> LOG_G;  // lanes per sub-DFT
    uint64_t blk = blockIdx.x;
    if ((gridDim.x & 7) == 0) blk = (blk & 7) * (gridDim.x >> 3) + (blk >> 3);
    const uint32_t gl = threadIdx.x & (G - 1), il = threadIdx.x >> LOG_G;
    const uint64_t col = (blk << (LOG_G - log_b)) + (gl >> log_b);
    const uint32_t r = gl & ((1u << log_b) - 1);
    const bool ok = col < ncols;
    const uint32_t log_n0 = LOG_S + log_b;
    {
        uint64_t w = 1, st = 1;
        if (r) {
            w = root_pow(rt_lo, rt_hi, log_n0, (uint64_t)r * il, false);
            st = root_pow(rt_lo, rt_hi, log_n0, (uint64_t)r * DI, false);
        }
        const uint64_t *src = coef + col * in_cs;
        // all of the lane's loads are in flight before the first is used
        uint64_t xv[EPT];
#pragma unroll
        for (uint32_t u = 0; u < EPT; u++) {
            const uint32_t i = il + u * DI;
            xv[u] = (i < S && ok) ? src[i] : 0;
        }
#pragma unroll
        for (uint32_t u = 0; u < EPT; u++) {
            const uint32_t i = il + u * DI;
            if (i < S) {
                uint64_t x = xv[u];
                if (r) x = mul(x, w);
                lds[padi<LOG_S>((gl << LOG_S) + i)] = x;
            }
            if (r) w = mul(w, st);
        }
    }
    __syncthreads();
    stockham_steps<LOG_S, 0, false>(lds, TILE, tw);
    if (!ok) return;
    uint64_t *dst = out + col * out_cs + r;
#pragma unroll 2
    for (uint32_t u = 0; u < EPT; u++) {
        const uint32_t k = il + u * DI;
        if (k < S) dst[(uint64_t)k << log_b] = canon(lds[padi<LOG_S>((gl << LOG_S) + k)]);
    }
}

// Odd coset of G = 8192 / 2^LOG_S whole columns per workgroup in one pass:
// out[c][j] = T_c(w_{2n} w_n^j) = DFT_n( iDFT_n(col_c)[i] * w_{2n}^i )[j].
// The column is read once, inverse-transformed in LDS, twisted by
// twist[i] = w_{2n}^i / n, transformed forward in LDS and written once: half
// the HBM traffic and launches of an iNTT pass followed by a one-pass LDE.
template <int LOG_S>
__global__ __launch_bounds__(NT, 4) void odd_coset_tile(const uint64_t *in, uint64_t *out, uint64_t ncols,
                                                       uint64_t in_cs, uint64_t out_cs, const uint64_t *twist,
                                                       const uint64_t *tw_f, const uint64_t *tw_i) {
    __shared__ uint64_t lds[LDS_ELEMS];
    constexpr uint32_t S = 1u << LOG_S;
    constexpr uint32_t LOG_G = TILE_LOG - LOG_S;
    uint64_t blk = blockIdx.x;
    if ((gridDim.x & 7) == 0) blk = (blk & 7) * (gridDim.x >> 3) + (blk >> 3);
    const uint64_t col0 = blk << LOG_G;
    {
        // all of the lane's loads (and twist factors) are in flight before
        // the first is used
        uint64_t xv[EPT];
#pragma unroll
        for (uint32_t u = 0; u < EPT; u++) {
            const uint32_t e = threadIdx.x + u * NT;
            const uint64_t col = col0 + (e >> LOG_S);
            xv[u] = col < ncols ? in[col * in_cs + (e & (S - 1))] : 0;
        }
#pragma unroll
        for (uint32_t u = 0; u < EPT; u++) lds[padi<LOG_S>(threadIdx.x + u * NT)] = xv[u];
    }
    __syncthreads();
    stockham_steps<LOG_S, 0, true>(lds, TILE, tw_i);
    {
        uint64_t tv[EPT];
#pragma unroll
        for (uint32_t u = 0; u < EPT; u++) tv[u] = twist[(threadIdx.x + u * NT) & (S - 1)];
#pragma unroll
        for (uint32_t u = 0; u < EPT; u++) {
            const uint32_t e = threadIdx.x + u * NT;
            lds[padi<LOG_S>(e)] = mul_lazy(lds[padi<LOG_S>(e)], tv[u]);  // the forward steps take any u64
        }
    }
    __syncthreads();
    stockham_steps<LOG_S, 0, false>(lds, TILE, tw_f);
#pragma unroll 2
    for (uint32_t u = 0; u < EPT; u++) {
        const uint32_t e = threadIdx.x + u * NT;
        const uint64_t col = col0 + (e >> LOG_S);
        if (col < ncols) out[col * out_cs + (e & (S - 1))] = canon(lds[padi<LOG_S>(e)]);
    }
}

template <bool INV, int MODE>
void launch_tile(uint32_t log_s, uint32_t tile_log, dim3 grid, hipStream_t st, const NttPass &p) {
    if (tile_log == 14) {
        switch (log_s) {
#define SG_CASE(L) \
    case L: hipLaunchKernelGGL((ntt_tile<L, INV, MODE, 14>), grid, dim3(TileCfg<14>::LANES), 0, st, p); break;
            SG_CASE(8) SG_CASE(9) SG_CASE(10) SG_CASE(11) SG_CASE(12)
#undef SG_CASE
        default: throw Error(E_INVALID, "ntt pass: 16K tiles hold sub-DFTs of 256 to 4096 points");
        }
        return;
    }
    switch (log_s) {
#define SG_CASE(L) \
    case L: hipLaunchKernelGGL((ntt_tile<L, INV, MODE, 13>), grid, dim3(NT), 0, st, p); break;
        SG_CASE(0) SG_CASE(1) SG_CASE(2) SG_CASE(3) SG_CASE(4) SG_CASE(5) SG_CASE(6)
        SG_CASE(7) SG_CASE(8) SG_CASE(9) SG_CASE(10) SG_CASE(11) SG_CASE(12)
#undef SG_CASE
    default: throw Error(E_INVALID, "ntt pass: sub-DFT larger than 4096");
    }
}

}  // namespace

void launch_pass(Ctx *c, NttPass p, const char *name) {
    SG_REQUIRE(p.log_s <= 12, "ntt pass: sub-DFT larger than 4096");
    if (p.tile_log == 0) p.tile_log = TILE_LOG;
    SG_REQUIRE(p.tile_log == TILE_LOG || (p.tile_log == 14 && p.log_s >= 8),
               "ntt pass: tiles of 2^13, or of 2^14 for sub-DFTs of 256 points and more");
    // as many sub-DFTs per tile as fit and exist
    uint32_t max_g = p.tile_log - p.log_s;
    while (max_g > 0 && (1ULL << max_g) > p.ng) max_g--;
    p.log_g = max_g;
    if (p.nr == 0) p.nr = 1;
    if (p.nb2 == 0) p.nb2 = 1;
    SG_REQUIRE(p.pre_log <= 26 && p.post_log <= 26, "ntt pass: twiddle order above 2^26");
    p.pre_direct = p.pre_log != 0 && p.pre_log <= 13;
    p.rt_lo = c->rt_lo;
    p.rt_hi = c->rt_hi;
    p.tw = p.inv ? c->tw_i : c->tw_f;
    const uint64_t G = 1ULL << p.log_g;
    const uint64_t tiles = ((p.ng + G - 1) / G) * p.nb;
    SG_REQUIRE(tiles < (1ULL << 31), "ntt pass: grid too large");
    const double bytes = 8.0 * (double)p.ng * (double)(1u << p.log_s) * p.nb * (1.0 + p.nr);
    ProfScope ps(c, name, bytes);
    // one repetition per workgroup: the B coset DFTs of a tile run side by
    // side on one XCD instead of one after another, which fills the chip for
    // small LDEs and reads the tile from HBM about once
    uint32_t rs_log = 0;
    while (p.nr % (2u << rs_log) == 0 && (tiles << (rs_log + 1)) < (1ULL << 31)) rs_log++;
    p.rs_log = rs_log;
    p.nr >>= rs_log;
    const dim3 grid((uint32_t)(tiles << rs_log));
    const bool plain = p.pre_log == 0 && p.post_log == 0 && p.scale == 1 && p.sp_lo == nullptr;
    const uint32_t lanes = p.tile_log == 14 ? 1024 : NT;
    const bool ldea = !p.inv && p.pre_direct && p.post_log && p.lazy_out && p.scale == 1 && !p.sp_lo &&
                      p.log_g > 0 && p.ig < p.ii && p.og < p.ok && (1u << p.log_g) <= lanes;
    const bool post = !p.pre_log && p.post_log && p.lazy_out && !p.sp_lo && p.log_g > 0 && p.ig < p.ii &&
                      p.og < p.ok && (1u << p.log_g) <= lanes;
    if (p.inv && plain)
        launch_tile<true, MODE_PLAIN>(p.log_s, p.tile_log, grid, c->stream, p);
    else if (p.inv && post)
        launch_tile<true, MODE_POST>(p.log_s, p.tile_log, grid, c->stream, p);
    else if (p.inv)
        launch_tile<true, MODE_GENERAL>(p.log_s, p.tile_log, grid, c->stream, p);
    else if (post)
        launch_tile<false, MODE_POST>(p.log_s, p.tile_log, grid, c->stream, p);
    else if (plain)
        launch_tile<false, MODE_PLAIN>(p.log_s, p.tile_log, grid, c->stream, p);
    else if (ldea)
        launch_tile<false, MODE_LDEA>(p.log_s, p.tile_log, grid, c->stream, p);
    else
        launch_tile<false, MODE_GENERAL>(p.log_s, p.tile_log, grid, c->stream, p);
    SG_HIP(hipGetLastError());
}

// Power table s^e, e < 2^26, two levels of 2^13 entries, built on the host.
PowTable make_pow_table(Ctx *c, uint64_t s) {
    PowTable t;
    auto it = c->pow_tables.find(s);
    if (it != c->pow_tables.end()) {
        t.buf = it->second;
        t.lo = t.buf;
        t.hi = t.buf + 8192;
        return t;
    }
    std::vector<uint64_t> tab(16384);
    const uint64_t s13 = pow(s, 8192);
    tab[0] = tab[8192] = 1;
    for (int i = 1; i < 8192; i++) {
        tab[i] = mul(tab[i - 1], s);
        tab[8192 + i] = mul(tab[8192 + i - 1], s13);
    }
    t.buf = c->alloc_n<uint64_t>(16384);
    SG_HIP(hipMemcpyAsync(t.buf, tab.data(), 8 * 16384, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    t.lo = t.buf;
    t.hi = t.buf + 8192;
    c->pow_tables[s] = t.buf;  // kept (as a live block) until the context is destroyed
    return t;
}

static NttPass base_pass() {
    NttPass p{};
    p.scale = 1;
    p.nr = 1;
    p.nb2 = 1;
    p.post_mul = 1;
    return p;
}

// ---- launch plans ----

namespace {

std::mutex plan_mu;
struct PlanEntry {
    uint32_t best = 0;  // 0 while a context is still timing the candidates
    std::vector<std::pair<uint32_t, float>> ms;
};
std::map<uint32_t, PlanEntry> plan_table;

uint32_t plan_key(uint32_t kind, uint32_t L, uint32_t LB, uint64_t ncols) {
    uint32_t lc = 0;
    while (lc < 40 && (2ULL << lc) <= ncols) lc++;
    return (kind << 24) | (L << 16) | (LB << 8) | lc;
}

// two-pass L1 window: within 2 of ceil(L/2), both passes of 2^3..2^12
// points, each with the 2^13 tile and, when both passes have 2^8 points or
// more, the 2^14 tile
void two_pass_window(uint32_t L, std::vector<uint32_t> &out) {
    const int mid = (int)(L + 1) / 2;
    const int lo = std::max({mid - 2, (int)L - 12, 3}), hi = std::min({mid + 2, 12, (int)L - 3});
    for (int s = lo; s <= hi; s++) {
        out.push_back(plan_code((uint32_t)s, 13));
        if (s >= 8 && (int)L - s >= 8) out.push_back(plan_code((uint32_t)s, 14));
    }
}

// The split for this call: the forced one if it is admissible, else the
// process's tuned one for the shape, else (with tuning on) the fastest
// candidate timed now with `run`, else the default. Contexts with tuning
// off still take plans other contexts have tuned, but never time any.
template <class Run>
uint32_t choose_plan(Ctx *c, uint32_t kind, uint32_t L, uint32_t LB, uint64_t ncols, const PlanSet &ps, Run run) {
    if (c->force_split || c->force_tile) {
        const uint32_t want = plan_code(c->force_split ? c->force_split : plan_split(ps.dflt),
                                        c->force_tile ? c->force_tile : plan_tile(ps.dflt));
        for (uint32_t s : ps.cand)
            if (plan_split(s) == plan_split(want) && plan_tile(s) == plan_tile(want)) return s;
        return ps.dflt;
    }
    if (ps.cand.size() < 2) return ps.dflt;
    const uint32_t key = plan_key(kind, L, LB, ncols);
    {
        std::lock_guard<std::mutex> g(plan_mu);
        auto it = plan_table.find(key);
        if (it != plan_table.end()) return it->second.best ? it->second.best : ps.dflt;
        if (!c->autotune) return ps.dflt;
        plan_table[key] = PlanEntry{};  // claimed: concurrent callers take the default meanwhile
    }
    PlanEntry e;
    hipEvent_t a, b;
    SG_HIP(hipEventCreate(&a));
    SG_HIP(hipEventCreate(&b));
    float best_ms = 0;
    for (uint32_t s : ps.cand) {
        run(s);  // first launch of a kernel shape (code object, caches) stays out of the timing
        SG_HIP(hipEventRecord(a, c->stream));
        run(s);
        SG_HIP(hipEventRecord(b, c->stream));
        SG_HIP(hipEventSynchronize(b));
        float ms = 0;
        SG_HIP(hipEventElapsedTime(&ms, a, b));
        e.ms.push_back({s, ms});
        if (!e.best || ms < best_ms) {
            e.best = s;
            best_ms = ms;
        }
    }
    SG_HIP(hipEventDestroy(a));
    SG_HIP(hipEventDestroy(b));
    std::lock_guard<std::mutex> g(plan_mu);
    plan_table[key] = e;
    return e.best;
}

}  // namespace

PlanSet ntt_plans(uint32_t L) {
    PlanSet ps;
    ps.dflt = plan_code(L <= 12 ? L : (L + 1) / 2, 13);
    if (L > 12) two_pass_window(L, ps.cand);
    return ps;
}

PlanSet lde_plans(uint32_t L, uint32_t LB) {
    PlanSet ps;
    if (L >= 4 && L <= 9 && L + LB <= (uint32_t)TILE_LOG) {
        ps.dflt = plan_code(L, 13);  // all cosets of a column in one workgroup: nothing to choose
        return ps;
    }
    ps.dflt = plan_code(L <= 12 ? L : (L + 1) / 2, 13);
    if (L <= 12) {
        ps.cand.push_back(plan_code(L, 13));
        if (L >= 8) ps.cand.push_back(plan_code(L, 14));
    }
    two_pass_window(L, ps.cand);
    return ps;
}

std::string autotune_report() {
    std::lock_guard<std::mutex> g(plan_mu);
    static const char *kinds[] = {"ntt", "intt", "lde", "odd"};
    std::string out;
    char line[96];
    for (auto &kv : plan_table) {
        if (!kv.second.best) continue;
        const uint32_t k = kv.first;
        snprintf(line, sizeof line, "%s %u %u %u -> %u.%u |", kinds[(k >> 24) % 4], (k >> 16) & 255,
                 (k >> 8) & 255, k & 255, plan_split(kv.second.best), plan_tile(kv.second.best));
        out += line;
        for (auto &m : kv.second.ms) {
            snprintf(line, sizeof line, " %u.%u:%.4f", plan_split(m.first), plan_tile(m.first), m.second);
            out += line;
        }
        out += "\n";
    }
    return out;
}

namespace {

// DFT of 2^L points per column as two strided passes of 2^L1 and 2^(L - L1)
// points (four-step), through `scratch`
void ntt_two_pass(Ctx *c, const uint64_t *in, uint64_t *out, uint64_t *scratch, uint32_t L, uint32_t plan,
                  uint64_t ncols, uint64_t in_cs, uint64_t out_cs, bool inv, uint64_t scale, const PowTable *spow) {
    const uint32_t L1 = plan_split(plan), L2 = L - L1;
    const uint64_t n = 1ULL << L, n1 = 1ULL << L1, n2 = 1ULL << L2;
    NttPass a = base_pass();
    a.in = in;
    a.out = scratch;
    a.nb = (uint32_t)ncols;
    a.ib = in_cs;
    a.ob = n;
    a.ng = n2;
    a.ig = 1;
    a.ii = n2;
    a.og = 1;
    a.ok = n2;
    a.log_s = L1;
    a.tile_log = plan_tile(plan);
    a.inv = inv;
    a.post_log = L;
    a.lazy_out = 1;  // pass B reads lazily reduced values
    // the iNTT's 1/n rides on pass A's post twiddle, so pass B without shift
    // powers is a plain pass
    a.scale = scale;
    launch_pass(c, a, inv ? "intt_passA" : "ntt_passA");
    NttPass b = base_pass();
    b.in = scratch;
    b.out = out;
    b.nb = (uint32_t)ncols;
    b.ib = n;
    b.ob = out_cs;
    b.ng = n1;
    b.ig = n2;
    b.ii = 1;
    b.og = 1;
    b.ok = n1;
    b.log_s = L2;
    b.tile_log = plan_tile(plan);
    b.inv = inv;
    if (spow) {
        b.sp_lo = spow->lo;
        b.sp_hi = spow->hi;
        b.sp_g = 1;
        b.sp_k = n1;
    }
    launch_pass(c, b, inv ? "intt_passB" : "ntt_passB");
}

}  // namespace

// out[col][k] = DFT(in[col])[k] (natural order), optional 1/n and s^k output scaling.
// A two-pass transform with in == out runs exactly once (no tuning), since a
// second run would read its own output.
void ntt_columns(Ctx *c, const uint64_t *in, uint64_t *out, uint64_t *scratch, uint32_t L, uint64_t ncols,
                 uint64_t in_cs, uint64_t out_cs, bool inv, const PowTable *spow) {
    SG_REQUIRE(L <= 24, "ntt: size above 2^24");
    const uint64_t n = 1ULL << L;
    const uint64_t scale = inv ? sg::inv(n % P) : 1;
    if (L <= 12) {
        NttPass p = base_pass();
        p.in = in;
        p.out = out;
        p.nb = 1;
        p.ng = ncols;
        p.ig = in_cs;
        p.ii = 1;
        p.og = out_cs;
        p.ok = 1;
        p.log_s = L;
        p.inv = inv;
        p.scale = scale;
        if (spow) {
            p.sp_lo = spow->lo;
            p.sp_hi = spow->hi;
            p.sp_k = 1;
        }
        launch_pass(c, p, inv ? "intt_small" : "ntt_small");
        return;
    }
    PlanSet ps = ntt_plans(L);
    if (in == out) ps.cand.clear();
    const uint32_t plan = choose_plan(c, inv ? 1 : 0, L, 0, ncols, ps, [&](uint32_t s) {
        ntt_two_pass(c, in, out, scratch, L, s, ncols, in_cs, out_cs, inv, scale, spow);
    });
    ntt_two_pass(c, in, out, scratch, L, plan, ncols, in_cs, out_cs, inv, scale, spow);
}

namespace {
void lde_forward(Ctx *c, const uint64_t *coef, uint64_t *out, uint32_t L, uint32_t LB, uint32_t plan,
                 uint64_t ncols, uint64_t out_cs);
}  // namespace

// Coset LDE of `ncols` columns: coefficients c = iDFT(col) scaled by shift^j,
// then out[col][blowup*k + r] = DFT_n( c_j * w_{blowup*n}^(r j) )[k] for
// r < blowup -- the evaluations on shift * <w_{blowup*n}> in natural order.
void lde_columns(Ctx *c, const uint64_t *cols, uint64_t *out, uint32_t L, uint32_t LB, uint64_t ncols,
                 uint64_t in_cs, uint64_t out_cs, uint64_t shift) {
    SG_REQUIRE(L + LB <= 26, "lde: domain above 2^26");
    SG_REQUIRE(L <= 24, "lde: trace above 2^24");
    const uint64_t n = 1ULL << L;
    Scoped coef(c, 8 * n * ncols);
    PowTable sp{};
    const bool use_shift = shift != 1;
    if (use_shift) sp = make_pow_table(c, shift);
    {
        Scoped scr(c, 8 * n * ncols);
        ntt_columns(c, cols, coef.as<uint64_t>(), scr.as<uint64_t>(), L, ncols, in_cs, n, true,
                    use_shift ? &sp : nullptr);
    }
    const uint32_t LN = L + LB;
    if (L >= 4 && L <= 9 && LN <= (uint32_t)TILE_LOG) {
        const uint64_t cpt = (uint64_t)TILE >> LN;  // columns per workgroup
        const uint64_t blocks = (ncols + cpt - 1) / cpt;
        SG_REQUIRE(blocks < (1ULL << 31), "lde: grid too large");
        ProfScope ps(c, "lde_small", 8.0 * ncols * (n + (n << LB)));
        const dim3 grid((uint32_t)blocks);
        switch (L) {
#define SG_CASE(LS)                                                                                          \
    case LS:                                                                                                 \
        hipLaunchKernelGGL((lde_small_tile<LS>), grid, dim3(NT), 0, c->stream, coef.as<uint64_t>(), out, ncols, \
                           LB, n, out_cs, c->rt_lo, c->rt_hi, c->tw_f);                                       \
        break;
            SG_CASE(4) SG_CASE(5) SG_CASE(6) SG_CASE(7) SG_CASE(8) SG_CASE(9)
#undef SG_CASE
        default: throw Error(E_INTERNAL, "lde: small-LDE case out of range");
        }
        SG_HIP(hipGetLastError());
    } else {
        const PlanSet ps = lde_plans(L, LB);
        const uint64_t *cf = coef.as<uint64_t>();
        const uint32_t plan = choose_plan(c, 2, L, LB, ncols, ps, [&](uint32_t s) {
            lde_forward(c, cf, out, L, LB, s, ncols, out_cs);
        });
        lde_forward(c, cf, out, L, LB, plan, ncols, out_cs);
    }
}

namespace {

// The forward half of a coset LDE from its (shift-scaled) coefficients coef
// (ncols x 2^L): L1 = L is one pass per column tile with the B cosets as
// repetitions (L <= 12); otherwise two passes of 2^L1 and 2^(L - L1) points.
void lde_forward(Ctx *c, const uint64_t *coef, uint64_t *out, uint32_t L, uint32_t LB, uint32_t plan,
                 uint64_t ncols, uint64_t out_cs) {
    const uint32_t L1 = plan_split(plan), tl = plan_tile(plan);
    const uint64_t n = 1ULL << L, B = 1ULL << LB;
    const uint32_t LN = L + LB;
    if (L1 == L) {
        SG_REQUIRE(L <= 12, "lde: one-pass form above 2^12 points");
        NttPass p = base_pass();
        p.in = coef;
        p.out = out;
        p.nb = 1;
        p.ng = ncols;
        p.ig = n;
        p.ii = 1;
        p.og = out_cs;
        p.ok = B;
        p.log_s = L;
        p.tile_log = tl;
        p.pre_log = LN;
        p.pre_i = 1;
        p.nr = (uint32_t)B;
        p.out_rs = 1;
        launch_pass(c, p, "lde_small");
    } else {
        SG_REQUIRE(L1 < L && L - L1 <= 12 && L1 <= 12, "lde: pass split out of range");
        const uint32_t L2 = L - L1;
        const uint64_t n1 = 1ULL << L1, n2 = 1ULL << L2;
        // scratch [col][k1][r][i2]: row q = k1*B + r of pass B is the k1-th
        // column DFT of coset r, and LDE[B*(k1 + n1*k2) + r] = out[q + B*n1*k2],
        // so pass B writes runs of consecutive q
        Scoped scr(c, 8 * n * B * ncols);
        NttPass a = base_pass();
        a.in = coef;
        a.out = scr.as<uint64_t>();
        a.nb = (uint32_t)ncols;
        a.ib = n;
        a.ob = B * n;
        a.ng = n2;
        a.ig = 1;
        a.ii = n2;
        a.og = 1;
        a.ok = B * n2;
        a.log_s = L1;
        a.tile_log = tl;
        // repetition r multiplies element j = i*n2 + g by w_{Bn}^(r j) =
        // w_{B n1}^(r i) w_{Bn}^(r g): the first factor before the sub-DFT
        // over i (order B n1, one table read), the second, constant over i,
        // after it together with the four-step twiddle w_n^(g k) = w_{Bn}^(B g k)
        a.pre_log = L1 + LB;
        a.pre_i = 1;
        a.post_log = LN;
        a.post_mul = B;
        a.post_rep = 1;
        a.lazy_out = 1;  // pass B reads lazily reduced values
        a.nr = (uint32_t)B;
        a.out_rs = n2;
        launch_pass(c, a, "lde_passA");
        NttPass b = base_pass();
        b.in = scr.as<uint64_t>();
        b.out = out;
        b.nb = (uint32_t)ncols;
        b.ib = B * n;
        b.ob = out_cs;
        b.ng = B * n1;
        b.ig = n2;
        b.ii = 1;
        b.og = 1;
        b.ok = B * n1;
        b.log_s = L2;
        b.tile_log = tl;
        launch_pass(c, b, "lde_passB");
    }
}

// The odd coset's forward half above 2^12 points: coset r = 1 of the
// blowup-2 LDE form, DFT_n(c_j w_{2n}^j), as two passes of 2^L1 and
// 2^(L - L1) points. Element j = i n2 + g takes w_{2 n1}^i before pass A's
// sub-DFT over i (one table read) and w_{2n}^g after it with the four-step
// twiddle: w_n^(g k) w_{2n}^g = w_{2n}^(g (2k + 1)). The coefficients carry
// 1/n and no shift powers, so the iNTT before this has no output product.
void odd_forward(Ctx *c, const uint64_t *coef, uint64_t *out, uint32_t L, uint32_t plan, uint64_t ncols) {
    const uint32_t L1 = plan_split(plan), tl = plan_tile(plan);
    SG_REQUIRE(L1 < L && L - L1 <= 12 && L1 <= 12, "odd coset: pass split out of range");
    const uint32_t L2 = L - L1;
    const uint64_t n = 1ULL << L, n1 = 1ULL << L1, n2 = 1ULL << L2;
    Scoped scr(c, 8 * n * ncols);
    NttPass a = base_pass();
    a.in = coef;
    a.out = scr.as<uint64_t>();
    a.nb = (uint32_t)ncols;
    a.ib = n;
    a.ob = n;
    a.ng = n2;
    a.ig = 1;
    a.ii = n2;
    a.og = 1;
    a.ok = n2;
    a.log_s = L1;
    a.tile_log = tl;
    a.r0 = 1;  // the one repetition, r = 1
    a.pre_log = L1 + 1;
    a.pre_i = 1;
    a.post_log = L + 1;
    a.post_mul = 2;
    a.post_rep = 1;
    a.lazy_out = 1;
    launch_pass(c, a, "odd_passA");
    NttPass b = base_pass();
    b.in = scr.as<uint64_t>();
    b.out = out;
    b.nb = (uint32_t)ncols;
    b.ib = n;
    b.ob = n;
    b.ng = n1;
    b.ig = n2;
    b.ii = 1;
    b.og = 1;
    b.ok = n1;
    b.log_s = L2;
    b.tile_log = tl;
    launch_pass(c, b, "odd_passB");
}

}  // namespace

const uint64_t *odd_twist_table(Ctx *c, uint32_t L) {
    SG_REQUIRE(L >= 1 && L <= 12, "odd coset: twist table size out of range");
    uint64_t *&twist = c->odd_twists[L];
    if (!twist) {
        const uint64_t n = 1ULL << L;
        std::vector<uint64_t> tw(n);
        const uint64_t w = root_of_unity(2 * n);
        tw[0] = sg::inv(n);
        for (uint64_t i = 1; i < n; i++) tw[i] = mul(tw[i - 1], w);
        uint64_t *buf = c->alloc_n<uint64_t>(n);  // kept (as a live block) until the context is destroyed
        SG_HIP(hipMemcpyAsync(buf, tw.data(), 8 * n, hipMemcpyHostToDevice, c->stream));
        SG_HIP(hipStreamSynchronize(c->stream));
        twist = buf;
    }
    return twist;
}

void lde_odd_coset(Ctx *c, const uint64_t *cols, uint64_t *out, uint32_t L, uint64_t ncols) {
    const uint64_t n = 1ULL << L;
    if (L < 1) {
        lde_columns(c, cols, out, L, 0, ncols, n, n, root_of_unity(2 * n));
        return;
    }
    if (L > 12) {
        SG_REQUIRE(L <= 24, "odd coset: trace above 2^24");
        Scoped coef(c, 8 * n * ncols);
        {
            Scoped scr(c, 8 * n * ncols);
            ntt_columns(c, cols, coef.as<uint64_t>(), scr.as<uint64_t>(), L, ncols, n, n, true, nullptr);
        }
        const uint64_t *cf = coef.as<uint64_t>();
        const uint32_t plan = choose_plan(c, 3, L, 0, ncols, lde_plans(L, 0),
                                          [&](uint32_t s) { odd_forward(c, cf, out, L, s, ncols); });
        odd_forward(c, cf, out, L, plan, ncols);
        return;
    }
    const uint64_t blocks = (ncols + (TILE >> L) - 1) >> (TILE_LOG - L);
    SG_REQUIRE(blocks < (1ULL << 31), "odd coset: grid too large");
    const uint64_t *twist = odd_twist_table(c, L);
    ProfScope ps(c, "odd_coset", 16.0 * n * ncols);
    const dim3 grid((uint32_t)blocks);
    switch (L) {
#define SG_CASE(LS)                                                                                                \
    case LS:                                                                                                       \
        hipLaunchKernelGGL((odd_coset_tile<LS>), grid, dim3(NT), 0, c->stream, cols, out, ncols, n, n, twist,    \
                           c->tw_f, c->tw_i);                                                                      \
        break;
        SG_CASE(1) SG_CASE(2) SG_CASE(3) SG_CASE(4) SG_CASE(5) SG_CASE(6)
        SG_CASE(7) SG_CASE(8) SG_CASE(9) SG_CASE(10) SG_CASE(11) SG_CASE(12)
#undef SG_CASE
    default: throw Error(E_INTERNAL, "odd coset: size out of range");
    }
    SG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- allocation-free transforms

// Scratch of the two-pass transforms for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops that forbid hipMalloc):
// one fixed array per device, held until the process ends. Every call on a
// device shares that device's array, so calls that overlap on one device race.
constexpr uint64_t NOALLOC_SCRATCH_WORDS = 1ULL << 22;  // 32 MiB
static __device__ __align__(256) uint64_t scratch[NOALLOC_SCRATCH_WORDS];

static std::mutex scratch_mu;
static std::map<int, uint64_t *> scratch_ptr;  // device -> address of its copy of `scratch`

// the device address of `scratch` on device dev (current), looked up once per device
static uint64_t *noalloc_scratch(int dev) {
    std::lock_guard<std::mutex> g(scratch_mu);
    auto it = scratch_ptr.find(dev);
    if (it != scratch_ptr.end()) return it->second;
    void *ptr = nullptr;
    SG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    SG_REQUIRE(ptr != nullptr, "ntt (no alloc): scratch symbol has no device address");
    scratch_ptr[dev] = (uint64_t *)ptr;
    return (uint64_t *)ptr;
}

uint64_t ntt_noalloc_max_elems() { return NOALLOC_SCRATCH_WORDS; }

void ntt_columns_noalloc(Ctx *c, const uint64_t *in, uint64_t *out, uint32_t L, uint64_t ncols, bool inv) {
    SG_REQUIRE(L <= 24 && ncols >= 1, "ntt (no alloc): size above 2^24 or no columns");
    SG_REQUIRE((ncols << L) <= NOALLOC_SCRATCH_WORDS,
               "ntt (no alloc): n * ncols above the 2^22-element scratch of the allocation-free entry point");
    uint64_t *scr = noalloc_scratch(c->device);
    // no plan timing (it creates events) and no profiling events: the table's
    // plan or the default split, the same transform either way
    const bool tune = c->autotune, prof = c->profiling;
    c->autotune = false;
    c->profiling = false;
    try {
        ntt_columns(c, in, out, scr, L, ncols, 1ULL << L, 1ULL << L, inv, nullptr);
    } catch (...) {
        c->autotune = tune;
        c->profiling = prof;
        throw;
    }
    c->autotune = tune;
    c->profiling = prof;
}

}  // namespace sg
