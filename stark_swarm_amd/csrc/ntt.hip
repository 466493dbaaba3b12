// ntt.hip -- batched Goldilocks NTT / iNTT and (coset) low-degree extension
// for gfx950.
//
// Replaces ark-poly's Radix2EvaluationDomain::{fft, ifft} and
// lde_trace_columns (crates/deep_ali/src/trace_import.rs:251-281): natural
// order in and out, forward transform X[k] = sum_j x[j] w^(jk) with
// w = 7^((p-1)/n), inverse scaled by 1/n.
//
// Structure: a DFT of size n = n1*n2 is two strided passes (four-step):
//   pass A  Z[k1*n2+i2] = w_n^(i2*k1) * DFT_n1( x[i1*n2+i2] )     (columns)
//   pass B  X[k1+n1*k2] = DFT_n2( Z[k1*n2+i2] )                    (rows)
// Each pass is one launch of `ntt_pass`: a workgroup owns a tile of G
// sub-DFTs of size S (S*G = 8192 u64 = 64 KiB of LDS), loads it with the
// tile's contiguous dimension on consecutive lanes, runs radix-16/8/4/2
// Stockham steps between registers and LDS, and writes it back, fusing
// input pre-twiddles, the four-step twiddle and output scaling.
#include "gl.hpp"
#include "ctx.hpp"
#include "ntt.hpp"

namespace sg {

namespace {

constexpr int TILE_LOG = 13;
constexpr int TILE = 1 << TILE_LOG;
constexpr int NT = 256;
constexpr int EPT = TILE / NT;  // elements held per thread during a step

constexpr uint64_t cmul(uint64_t a, uint64_t b) { return (uint64_t)(((unsigned __int128)a * b) % P); }
constexpr uint64_t cpow(uint64_t a, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = cmul(r, a);
        a = cmul(a, a);
        e >>= 1;
    }
    return r;
}
struct Root16 {
    uint64_t f[16], i[16];
};
constexpr Root16 make_root16() {
    Root16 t{};
    uint64_t w = cpow(7, (P - 1) / 16), wi = cpow(w, 15);
    uint64_t a = 1, b = 1;
    for (int k = 0; k < 16; k++) {
        t.f[k] = a;
        t.i[k] = b;
        a = cmul(a, w);
        b = cmul(b, wi);
    }
    return t;
}
__device__ constexpr Root16 ROOT16 = make_root16();

template <int LR>
__device__ __forceinline__ constexpr int brev_c(int x) {
    int r = 0;
    for (int i = 0; i < LR; i++) r |= ((x >> i) & 1) << (LR - 1 - i);
    return r;
}

// in-register DFT of size R = 2^LR: v[k] <- sum_r v[r] w_R^(rk)
template <int LR, bool INV>
__device__ __forceinline__ void dft_reg(uint64_t *v) {
    constexpr int R = 1 << LR;
#pragma unroll
    for (int i = 0; i < R; i++) {
        int j = brev_c<LR>(i);
        if (i < j) {
            uint64_t t = v[i];
            v[i] = v[j];
            v[j] = t;
        }
    }
#pragma unroll
    for (int lh = 0; lh < LR; lh++) {
        const int half = 1 << lh;
#pragma unroll
        for (int base = 0; base < R; base += 2 * half) {
#pragma unroll
            for (int j = 0; j < half; j++) {
                // w_{2 half}^j = w_16^(j * 16 / (2 half))
                const int e = j * (16 >> (lh + 1));
                uint64_t t = v[base + j + half];
                if (j != 0) t = mul(t, INV ? ROOT16.i[e] : ROOT16.f[e]);
                uint64_t u = v[base + j];
                v[base + j] = add(u, t);
                v[base + j + half] = sub(u, t);
            }
        }
    }
}

__device__ __forceinline__ uint64_t root_pow(const NttPass &p, uint32_t lg, uint64_t e, bool inv) {
    uint64_t e26 = (e & ((1ULL << lg) - 1)) << (26 - lg);
    if (inv) e26 = (0 - e26) & ((1ULL << 26) - 1);
    return mul(p.rt_lo[e26 & 8191], p.rt_hi[e26 >> 13]);
}

template <int LR, bool INV>
__device__ __forceinline__ void stockham_step(uint64_t *lds, uint32_t log_s, uint32_t tile, uint32_t log_ns,
                                              const uint64_t *tw) {
    constexpr int R = 1 << LR;
    constexpr int MAXG = EPT / R;
    const uint32_t sub_log = log_s - LR;
    const uint32_t sub = 1u << sub_log;
    const uint32_t groups = tile >> LR;
    const uint32_t Ns = 1u << log_ns;
    uint64_t v[MAXG][R];
#pragma unroll
    for (int u = 0; u < MAXG; u++) {
        uint32_t q = threadIdx.x + u * NT;
        if (q < groups) {
            uint32_t g = q >> sub_log, j = q & (sub - 1);
            const uint64_t *src = lds + (g << log_s) + j;
#pragma unroll
            for (int r = 0; r < R; r++) v[u][r] = src[r * sub];
            if (log_ns) {
                uint32_t jm = j & (Ns - 1);
                const uint64_t *t = tw + (Ns << LR);
#pragma unroll
                for (int r = 1; r < R; r++) v[u][r] = mul(v[u][r], t[jm * r]);
            }
            dft_reg<LR, INV>(v[u]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < MAXG; u++) {
        uint32_t q = threadIdx.x + u * NT;
        if (q < groups) {
            uint32_t g = q >> sub_log, j = q & (sub - 1);
            uint32_t jm = j & (Ns - 1);
            uint32_t idx = ((j >> log_ns) << (log_ns + LR)) + jm;
            uint64_t *dst = lds + (g << log_s) + idx;
#pragma unroll
            for (int r = 0; r < R; r++) dst[r * Ns] = v[u][r];
        }
    }
    __syncthreads();
}

template <bool INV>
__device__ void dft_tile(uint64_t *lds, uint32_t log_s, uint32_t tile, const uint64_t *tw) {
    uint32_t log_ns = 0;
    while (log_ns < log_s) {
        uint32_t left = log_s - log_ns;
        if (left >= 4) {
            stockham_step<4, INV>(lds, log_s, tile, log_ns, tw);
            log_ns += 4;
        } else if (left == 3) {
            stockham_step<3, INV>(lds, log_s, tile, log_ns, tw);
            log_ns += 3;
        } else if (left == 2) {
            stockham_step<2, INV>(lds, log_s, tile, log_ns, tw);
            log_ns += 2;
        } else {
            stockham_step<1, INV>(lds, log_s, tile, log_ns, tw);
            log_ns += 1;
        }
    }
}

__global__ __launch_bounds__(NT) void ntt_pass(NttPass p) {
    __shared__ uint64_t lds[TILE];
    const uint32_t log_s = p.log_s, log_g = p.log_g;
    const uint32_t S = 1u << log_s, G = 1u << log_g;
    const uint32_t tile = S << log_g;
    const uint64_t tiles_per_batch = (p.ng + G - 1) >> log_g;
    const uint64_t blk = blockIdx.x;
    const uint64_t b = blk / tiles_per_batch;
    const uint64_t g0 = (blk - b * tiles_per_batch) << log_g;
    const uint64_t *in = p.in + b * p.ib + g0 * p.ig;
    const bool in_gfast = log_g > 0 && p.ig < p.ii;
    for (uint32_t e = threadIdx.x; e < tile; e += NT) {
        uint32_t g, i;
        if (in_gfast) {
            g = e & (G - 1);
            i = e >> log_g;
        } else {
            i = e & (S - 1);
            g = e >> log_s;
        }
        uint64_t x = 0;
        if (g0 + g < p.ng) {
            x = in[(int64_t)g * p.ig + (int64_t)i * p.ii];
            if (p.pre_log) x = mul(x, root_pow(p, p.pre_log, (g0 + g) * p.pre_g + i * p.pre_i + b * p.pre_b, false));
        }
        lds[(g << log_s) + i] = x;
    }
    __syncthreads();
    if (p.inv)
        dft_tile<true>(lds, log_s, tile, p.tw);
    else
        dft_tile<false>(lds, log_s, tile, p.tw);
    uint64_t *out = p.out + b * p.ob + g0 * p.og;
    const bool out_gfast = log_g > 0 && p.og < p.ok;
    for (uint32_t e = threadIdx.x; e < tile; e += NT) {
        uint32_t g, k;
        if (out_gfast) {
            g = e & (G - 1);
            k = e >> log_g;
        } else {
            k = e & (S - 1);
            g = e >> log_s;
        }
        if (g0 + g >= p.ng) continue;
        uint64_t x = lds[(g << log_s) + k];
        if (p.post_log) x = mul(x, root_pow(p, p.post_log, (g0 + g) * k, p.inv));
        if (p.scale != 1) x = mul(x, p.scale);
        if (p.sp_lo) {
            uint64_t ee = (g0 + g) * p.sp_g + (uint64_t)k * p.sp_k;
            x = mul(x, mul(p.sp_lo[ee & 8191], p.sp_hi[ee >> 13]));
        }
        out[(int64_t)g * p.og + (int64_t)k * p.ok] = x;
    }
}

}  // namespace

void launch_pass(Ctx *c, NttPass p, const char *name) {
    SG_REQUIRE(p.log_s <= 12, "ntt pass: sub-DFT larger than 4096");
    uint32_t max_g = TILE_LOG - p.log_s;
    if (p.log_g > max_g) p.log_g = max_g;
    p.rt_lo = c->rt_lo;
    p.rt_hi = c->rt_hi;
    p.tw = p.inv ? c->tw_i : c->tw_f;
    uint64_t G = 1ULL << p.log_g;
    uint64_t tiles = ((p.ng + G - 1) / G) * p.nb;
    SG_REQUIRE(tiles < (1ULL << 31), "ntt pass: grid too large");
    double bytes = 16.0 * (double)p.ng * (double)(1u << p.log_s) * p.nb;
    ProfScope ps(c, name, bytes);
    hipLaunchKernelGGL(ntt_pass, dim3((uint32_t)tiles), dim3(NT), 0, c->stream, p);
    SG_HIP(hipGetLastError());
}

static uint32_t pick_log_g(uint64_t ng, uint32_t log_s) {
    uint32_t lg = TILE_LOG - log_s;
    while (lg > 0 && (1ULL << lg) > ng) lg--;
    return lg;
}

// Power table s^e, e < 2^26, two levels of 2^13 entries, built on the host.
PowTable make_pow_table(Ctx *c, uint64_t s) {
    PowTable t;
    std::vector<uint64_t> lo(8192), hi(8192);
    uint64_t s13 = pow(s, 8192);
    lo[0] = hi[0] = 1;
    for (int i = 1; i < 8192; i++) {
        lo[i] = mul(lo[i - 1], s);
        hi[i] = mul(hi[i - 1], s13);
    }
    t.buf = c->alloc_n<uint64_t>(16384);
    SG_HIP(hipMemcpyAsync(t.buf, lo.data(), 8 * 8192, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(t.buf + 8192, hi.data(), 8 * 8192, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    t.lo = t.buf;
    t.hi = t.buf + 8192;
    return t;
}

static NttPass base_pass() {
    NttPass p{};
    p.scale = 1;
    return p;
}

// out[col][k] = DFT(in[col])[k] (natural order), optional 1/n and s^k output scaling.
void ntt_columns(Ctx *c, const uint64_t *in, uint64_t *out, uint64_t *scratch, uint32_t L, uint64_t ncols,
                 uint64_t in_cs, uint64_t out_cs, bool inv, const PowTable *spow) {
    SG_REQUIRE(L <= 24, "ntt: size above 2^24");
    const uint64_t n = 1ULL << L;
    uint64_t scale = inv ? sg::inv(n % P) : 1;
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
        p.log_g = pick_log_g(ncols, L);
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
    const uint32_t L1 = (L + 1) / 2, L2 = L - L1;
    const uint64_t n1 = 1ULL << L1, n2 = 1ULL << L2;
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
    a.log_g = pick_log_g(n2, L1);
    a.inv = inv;
    a.post_log = L;
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
    b.log_g = pick_log_g(n1, L2);
    b.inv = inv;
    b.scale = scale;
    if (spow) {
        b.sp_lo = spow->lo;
        b.sp_hi = spow->hi;
        b.sp_g = 1;
        b.sp_k = n1;
    }
    launch_pass(c, b, inv ? "intt_passB" : "ntt_passB");
}

// Coset LDE of `ncols` columns: coefficients c = iDFT(col) scaled by shift^j,
// then out[col][blowup*k + r] = DFT_n( c_j * w_{blowup*n}^(r j) )[k] for
// r < blowup -- the evaluations on shift * <w_{blowup*n}> in natural order.
void lde_columns(Ctx *c, const uint64_t *cols, uint64_t *out, uint32_t L, uint32_t LB, uint64_t ncols,
                 uint64_t in_cs, uint64_t out_cs, uint64_t shift) {
    SG_REQUIRE(L + LB <= 26, "lde: domain above 2^26");
    SG_REQUIRE(L <= 24, "lde: trace above 2^24");
    const uint64_t n = 1ULL << L, B = 1ULL << LB;
    Scoped coef(c, 8 * n * ncols), scr(c, 8 * n * ncols);
    PowTable sp{};
    bool use_shift = shift != 1;
    if (use_shift) sp = make_pow_table(c, shift);
    ntt_columns(c, cols, coef.as<uint64_t>(), scr.as<uint64_t>(), L, ncols, in_cs, n, true, use_shift ? &sp : nullptr);
    const uint32_t LN = L + LB;
    for (uint64_t r = 0; r < B; r++) {
        if (L <= 12) {
            NttPass p = base_pass();
            p.in = coef.as<uint64_t>();
            p.out = out + r;
            p.nb = 1;
            p.ng = ncols;
            p.ig = n;
            p.ii = 1;
            p.og = out_cs;
            p.ok = B;
            p.log_s = L;
            p.log_g = pick_log_g(ncols, L);
            p.pre_log = LN;
            p.pre_i = r;
            launch_pass(c, p, "lde_small");
            continue;
        }
        const uint32_t L1 = (L + 1) / 2, L2 = L - L1;
        const uint64_t n1 = 1ULL << L1, n2 = 1ULL << L2;
        NttPass a = base_pass();
        a.in = coef.as<uint64_t>();
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
        a.log_g = pick_log_g(n2, L1);
        a.post_log = L;
        a.pre_log = LN;
        a.pre_g = r;
        a.pre_i = r * n2;
        launch_pass(c, a, "lde_passA");
        NttPass b = base_pass();
        b.in = scr.as<uint64_t>();
        b.out = out + r;
        b.nb = (uint32_t)ncols;
        b.ib = n;
        b.ob = out_cs;
        b.ng = n1;
        b.ig = n2;
        b.ii = 1;
        b.og = B;
        b.ok = B * n1;
        b.log_s = L2;
        b.log_g = pick_log_g(n1, L2);
        launch_pass(c, b, "lde_passB");
    }
    if (use_shift) c->release(sp.buf);
}

}  // namespace sg
