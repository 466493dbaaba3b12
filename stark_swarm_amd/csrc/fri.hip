// fri.hip -- STIR DEEP-FRI prover for gfx950, batched over independent proofs.
//
// Device work per layer l (size n, fold arity m), crates/deep_ali/src/fri.rs:
//   OOD coset   f_l(zeta^j z) for j < m (fri.rs:338-391). With c = iNTT(f_l)
//               and zeta = w_n^(n/m) a primitive m-th root,
//               f_l(zeta^j z) = sum_rho zeta^(j rho) S_rho,
//               S_rho = sum_{k = rho mod m} c_k z^k,
//               so one streaming pass forms the m residue sums and an m-point
//               DFT finishes (`ood_sums`, `ood_finish`).
//   commit      SHA3 leaves of the 6 components, arity pick_arity(n, m).
//   fold        per coset b: interpolant coefficients of the fibre
//               f[b + j n/m] on w^b zeta^j (m-point inverse DFT, 1/m,
//               w^(-bk)), then f_{l+1}[b] = sum_k coeff_k alpha^k
//               (fri.rs:670-771, `coset_fold`).
// The transcript, challenges and query indices are host work (fri.rs:1398-1718,
// 1813-1893); values cross PCIe only as roots, m coset evaluations, the final
// layer and the query gathers.
//
// Batching: K proofs of the same size run in lockstep. Every launch covers
// all K (grid.y = proof) and each stage synchronizes once for the batch; the
// K transcripts are independent host objects.
//
// Layout: an Fp6 layer is 6 planes of n u64 (component-major), proofs
// back to back ([proof][component][n]), so every kernel reads unit-stride
// streams and the NTT engine sees K*6 columns.
#include "fri.hpp"
#include "gl.hpp"
#include "keccak.hpp"
#include "ntt.hpp"

#include <string.h>

#include <array>

namespace sg {

namespace {

constexpr int FT = 256;

// Residue sums S_rho = sum_{k = rho mod M} c_k z^k of proof blockIdx.y.
// T: per proof, T[i] = z^(2^i) (24 entries). Block b covers [k0, k0 + OT*ch),
// k0 = b*OT*ch; thread tid reads k = k0 + t*OT + tid (coalesced), all of
// residue tid mod M, and keeps one accumulator A = sum_t c_k z^(t*OT). The
// z^tid factors come in as a Horner tree: lanes at distance s combine as
// v_l + z^s v_{l+s} for s = M..32 inside a wave, then across the OT/64 waves
// with z^64; lane rho < M is left with z^-(k0+rho) S_rho(block).
// One wave per block: a small layer's block time is its chain of Fp6
// products (the Horner tree, then the scale by z^(k0+rho), which runs
// lane-parallel below), and one-wave blocks need no cross-wave Horner step.
constexpr int OT = 64, LOG_OT = 6;
template <int M>
__global__ __launch_bounds__(OT) void ood_sums_k(const uint64_t *coef, uint64_t n, int ncomp, const uint64_t *T,
                                                 uint32_t tstride, uint32_t ch, uint64_t *partials) {
    constexpr int NW = OT / 64;
    __shared__ uint64_t zt[64 * 6];  // z^(t*OT), t < ch <= 64
    __shared__ uint64_t red[NW][M * 6];
    __shared__ uint64_t tl[24 * 6];  // z^(2^i), i < 24: read once per block, not along each product chain
    const uint64_t pk = blockIdx.y;
    coef += pk * ncomp * n;
    T += pk * 6 * tstride;
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < 24 * 6; i += OT) tl[i] = T[i];
    __syncthreads();
    auto tget = [&](int i) {
        fp6 r;
#pragma unroll
        for (int q = 0; q < 6; q++) r.c[q] = tl[6 * i + q];
        return r;
    };
    if (tid < ch) {
        fp6 w = f6_one();
        for (int bit = 0; (1u << bit) <= tid; bit++)
            if (tid >> bit & 1) w = f6_mul(w, tget(LOG_OT + bit));
#pragma unroll
        for (int q = 0; q < 6; q++) zt[6 * tid + q] = w.c[q];
    }
    __syncthreads();
    const uint64_t k0 = (uint64_t)blockIdx.x * OT * ch;
    fp6 acc = f6_zero();
    // the first element's factor is z^0 = 1: no product
    if (k0 + tid < n) {
        const uint64_t k = k0 + tid;
        if (ncomp == 1) {
            acc.c[0] = canon(coef[k]);
        } else {
#pragma unroll
            for (int q = 0; q < 6; q++) acc.c[q] = canon(coef[q * n + k]);
        }
    }
    for (uint32_t t = 1; t < ch; t++) {
        const uint64_t k = k0 + (uint64_t)t * OT + tid;
        if (k >= n) break;
        fp6 zp;
#pragma unroll
        for (int q = 0; q < 6; q++) zp.c[q] = zt[6 * t + q];
        if (ncomp == 1) {
            acc = f6_add(acc, f6_scale(zp, coef[k]));
        } else {
            fp6 cv;
#pragma unroll
            for (int q = 0; q < 6; q++) cv.c[q] = coef[q * n + k];
            acc = f6_add(acc, f6_mul(cv, zp));
        }
    }
    const uint32_t lane = tid & 63, wv = tid >> 6;
#pragma unroll
    for (int s = M, bit = __builtin_ctz(M); s < 64; s <<= 1, bit++) {
        fp6 o;
#pragma unroll
        for (int q = 0; q < 6; q++) o.c[q] = __shfl_down(acc.c[q], s, 64);
        if ((lane & s) == 0) acc = f6_add(acc, f6_mul(o, tget(bit)));
    }
    if constexpr (NW == 1 && M <= 8) {
        // Lane rho < M now holds z^-(k0+rho) S_rho(block); the scale factor
        // z^(k0+rho) is applied lane-parallel. Group gi = lane / 8 takes
        // residue gi; its lane q < 6 keeps slot q of the group's value and
        // forms that slot of each Fp6 product (six base products, about a
        // sixth of a whole product's latency), and the group regathers the
        // six slots by shuffles. Fp6 = Fp[y]/(y^6 - 7); slots 0..5 hold the
        // coefficients of y^0, y^2, y^4, y^1, y^3, y^5.
        const uint32_t gi = lane >> 3, q = lane & 7, qs = q < 6 ? q : 5;
        const uint32_t g = qs < 3 ? 2 * qs : 2 * (qs - 3) + 1;  // exponent of slot qs
        const uint32_t src = gi < (uint32_t)M ? gi : 0;
        fp6 a;
        uint64_t mine = 0;
#pragma unroll
        for (int j = 0; j < 6; j++) {
            a.c[j] = __shfl(acc.c[j], (int)src, 64);
            if ((uint32_t)j == qs) mine = a.c[j];
        }
        // k0 is a multiple of OT > rho: the bits of k0 + rho are those of k0 and rho
        const uint64_t e = k0 + src, steps = k0 | (uint64_t)(M - 1);
        for (int bit = 0; bit < 24; bit++) {
            if (!(steps >> bit & 1)) continue;  // the same for the whole wave
            uint64_t sn = 0, sw = 0;
#pragma unroll
            for (int ex = 0; ex < 6; ex++) {
                const int sa = (ex & 1) ? 3 + (ex >> 1) : (ex >> 1);  // slot of y^ex
                const uint32_t f = g >= (uint32_t)ex ? g - ex : g + 6 - ex;
                const uint32_t fs = (f & 1) ? 3 + (f >> 1) : (f >> 1);
                const uint64_t pr = mul(a.c[sa], tl[6 * bit + fs]);
                // y^ex y^f = 7 y^g when ex + f = g + 6
                const bool wrap = (uint32_t)ex > g;
                sn = add(sn, wrap ? 0 : pr);
                sw = add(sw, wrap ? pr : 0);
            }
            const uint64_t r = add(sn, mul7(sw));
            if (e >> bit & 1) mine = r;
#pragma unroll
            for (int j = 0; j < 6; j++) a.c[j] = __shfl(mine, (int)((lane & ~7u) + j), 64);
        }
        if (gi < (uint32_t)M && q < 6) partials[(pk * gridDim.x + blockIdx.x) * M * 6 + gi * 6 + q] = mine;
        return;
    }
    if (lane < (uint32_t)M) {
#pragma unroll
        for (int q = 0; q < 6; q++) red[wv][lane * 6 + q] = acc.c[q];
    }
    __syncthreads();
    if (tid < (uint32_t)M) {
        fp6 v;
#pragma unroll
        for (int q = 0; q < 6; q++) v.c[q] = red[NW - 1][tid * 6 + q];
        const fp6 z64 = tget(6);
#pragma unroll
        for (int w = NW - 2; w >= 0; w--) {
            fp6 x;
#pragma unroll
            for (int q = 0; q < 6; q++) x.c[q] = red[w][tid * 6 + q];
            v = f6_add(f6_mul(v, z64), x);
        }
        // times z^(k0 + rho)
        const uint64_t e = k0 + tid;
        for (int bit = 0; bit < 24; bit++)
            if (e >> bit & 1) v = f6_mul(v, tget(bit));
        uint64_t *part = partials + (pk * gridDim.x + blockIdx.x) * M * 6;
#pragma unroll
        for (int q = 0; q < 6; q++) part[tid * 6 + q] = v.c[q];
    }
}

// One block per proof: sum the block partials, then evals_j = sum_rho zeta^(j rho) S_rho.
// One block of FIN_T threads per proof: sum the block partials (thread
// (c, r) adds blocks b = r mod R of column c = (rho, comp), then the R row
// sums meet in LDS), then evals_j = sum_rho zeta^(j rho) S_rho.
constexpr int FIN_T = 1024;
__global__ __launch_bounds__(FIN_T) void ood_finish_k(const uint64_t *partials, uint32_t nblocks, uint32_t m,
                                                      uint64_t zeta, uint64_t *evals) {
    __shared__ uint64_t red[FIN_T];
    __shared__ uint64_t S[16 * 6];
    const uint64_t pk = blockIdx.x;
    partials += pk * nblocks * m * 6;
    evals += pk * m * 6;
    const uint32_t C = 6 * m, R = FIN_T / C, t = threadIdx.x;
    const uint32_t c = t % C, r = t / C;
    uint64_t s = 0;
    if (r < R) {
        // eight partials in flight per step (clamped reads, masked sums)
        constexpr uint32_t U = 8;
        for (uint32_t b = r; b < nblocks; b += U * R) {
            uint64_t x[U];
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                const uint32_t bb = b + u * R;
                x[u] = partials[(uint64_t)(bb < nblocks ? bb : nblocks - 1) * C + c];
            }
#pragma unroll
            for (uint32_t u = 0; u < U; u++)
                if (b + u * R < nblocks) s = add(s, x[u]);
        }
    }
    red[t] = s;
    __syncthreads();
    if (t < C) {
        uint64_t v = 0;
        for (uint32_t i = 0; i < R; i++) v = add(v, red[i * C + t]);
        S[t] = v;
    }
    __syncthreads();
    if (t < C) {
        const uint32_t j = t / 6, q = t % 6;
        const uint64_t zj = pow(zeta, j);
        uint64_t zp = 1, acc = 0;
        for (uint32_t rho = 0; rho < m; rho++) {
            acc = add(acc, mul(S[rho * 6 + q], zp));
            zp = mul(zp, zj);
        }
        evals[j * 6 + q] = acc;
    }
}

constexpr uint64_t cmulc(uint64_t a, uint64_t b) { return (uint64_t)(((unsigned __int128)a * b) % P); }
constexpr uint64_t cpowc(uint64_t a, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = cmulc(r, a);
        a = cmulc(a, a);
        e >>= 1;
    }
    return r;
}
struct RootI16 {
    uint64_t i[16];
};
constexpr RootI16 make_rooti16() {
    RootI16 t{};
    uint64_t wi = cpowc(cpowc(7, (P - 1) / 16), 15), a = 1;
    for (int k = 0; k < 16; k++) {
        t.i[k] = a;
        a = cmulc(a, wi);
    }
    return t;
}
__device__ constexpr RootI16 ROOTI16 = make_rooti16();

// inverse DFT (unscaled) of size M = 2^LM on one component: v[k] <- sum_j v[j] zeta^(-jk)
template <int LM>
__device__ __forceinline__ void idft_small(uint64_t *v) {
    constexpr int M = 1 << LM;
#pragma unroll
    for (int i = 0; i < M; i++) {
        int j = 0;
#pragma unroll
        for (int b = 0; b < LM; b++) j |= ((i >> b) & 1) << (LM - 1 - b);
        if (i < j) {
            uint64_t t = v[i];
            v[i] = v[j];
            v[j] = t;
        }
    }
#pragma unroll
    for (int lh = 0; lh < LM; lh++) {
        const int half = 1 << lh;
#pragma unroll
        for (int base = 0; base < M; base += 2 * half) {
#pragma unroll
            for (int j = 0; j < half; j++) {
                uint64_t t = v[base + j + half];
                if (j) t = mul(t, ROOTI16.i[j * (16 >> (lh + 1))]);
                uint64_t u = v[base + j];
                v[base + j] = add(u, t);
                v[base + j + half] = sub(u, t);
            }
        }
    }
}

// x * u^E in Fp6 = Fp[u]/(u^6 - 7), components in memory order
// (c0.0, c0.1, c0.2, c1.0, c1.1, c1.2) = coefficients of (1, u^2, u^4, u, u^3, u^5)
template <int E>
__device__ __forceinline__ void f6_add_mul_upow(fp6 &acc, const fp6 &x) {
    constexpr int upow[6] = {0, 2, 4, 1, 3, 5};
    constexpr int slot[6] = {0, 3, 1, 4, 2, 5};  // memory slot of u^i
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const int e = upow[q] + E;
        const uint64_t v = e >= 6 ? mul(x.c[q], 7) : x.c[q];
        const int d = slot[e % 6];
        acc.c[d] = add(acc.c[d], v);
    }
}

// f_next[b] = sum_k alpha^k * (1/m) w^(-bk) * IDFT_m(f[b + j nn])_k, proof blockIdx.y.
// One component at a time: with y_k = sum_q y_{k,q} E_q (E_q the basis
// element of memory slot q), sum_k alpha^k y_k = sum_q E_q sum_k y_{k,q} alpha^k,
// so each component costs a scalar times Fp6 per k (apow: alpha^k, k < M, per
// proof) and only one component's M values are live at a time.
// Optional coefficient-tuple planes: tuples[(k*6 + q) * nn + b] = y_{k,q}.
// A block is 6 waves over 64 outputs b, wave q taking component q (one
// output per lane, no divergence), so a small layer's per-output chain is a
// sixth as long; the six Fp6 partial sums meet in LDS.
// apow null: the powers of the single proof come in pa (a kernel argument, so
// a one-proof fold needs no upload and waits on no copy).
constexpr int FOLD_T = 6 * 64;
struct FoldPow {
    uint64_t v[6 * 16];
};
template <int LM>
__global__ __launch_bounds__(FOLD_T) void coset_fold_k(const uint64_t *f, uint64_t n, int ncomp, const uint64_t *apow,
                                                       FoldPow pa, uint64_t minv, const uint64_t *rt_lo,
                                                       const uint64_t *rt_hi, uint32_t logn, uint64_t *fnext,
                                                       uint64_t *tuples) {
    constexpr int M = 1 << LM;
    __shared__ uint64_t red[6][6][64];
    const uint64_t nn = n >> LM;
    const uint32_t wq = threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint64_t b = (uint64_t)blockIdx.x * 64 + lane;
    const bool live = b < nn;  // no early exit: every wave meets the barrier
    const uint64_t pk = blockIdx.y;
    f += pk * ncomp * n;
    fnext += pk * 6 * nn;
    if (apow) apow += pk * 6 * M;
    if (tuples) tuples += pk * 6 * M * nn;
    // w^(-b) from the 2^26 root table
    uint64_t e26 = ((live ? b : 0) << (26 - logn)) & ((1ULL << 26) - 1);
    e26 = (0 - e26) & ((1ULL << 26) - 1);
    const uint64_t wb = mul(rt_lo[e26 & 8191], rt_hi[e26 >> 13]);
    fp6 acc = f6_zero();
    auto component = [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if (!live) return;
        if (q >= ncomp) {
            if (tuples)
#pragma unroll
                for (int k = 0; k < M; k++) tuples[(uint64_t)(k * 6 + q) * nn + b] = 0;
            return;
        }
        uint64_t v[M];
#pragma unroll
        for (int j = 0; j < M; j++) v[j] = f[q * n + b + j * nn];
        idft_small<LM>(v);
        uint64_t s = minv;  // (1/m) w^(-bk)
        // pq_i = sum_k alpha^k_i y_k: the M products of a slot are summed as
        // 192-bit integers (lo, mid, top < 2^4) and reduced once, about a
        // third of the instructions of a reduced product and add per term
        uint64_t sl[6], sm[6], st[6];
#pragma unroll
        for (int i = 0; i < 6; i++) sl[i] = sm[i] = st[i] = 0;
#pragma unroll
        for (int k = 0; k < M; k++) {
            const uint64_t y = mul(v[k], s);
            if (tuples) tuples[(uint64_t)(k * 6 + q) * nn + b] = y;
#pragma unroll
            for (int i = 0; i < 6; i++) {
                const uint64_t ak = apow ? apow[6 * k + i] : pa.v[6 * k + i];
                uint64_t lo, hi;
                mul128(ak, y, lo, hi);  // hi <= 2^64 - 2, so hi + carry does not wrap
                sl[i] += lo;
                const uint64_t h = hi + (sl[i] < lo);
                sm[i] += h;
                st[i] += sm[i] < h;
            }
            s = mul(s, wb);
        }
        // lo + mid 2^64 + top 2^128 with 2^128 = -2^32 (mod p), top < 16
        fp6 pq;
#pragma unroll
        for (int i = 0; i < 6; i++) pq.c[i] = sub(reduce128(sl[i], sm[i]), st[i] << 32);
        constexpr int upow[6] = {0, 2, 4, 1, 3, 5};
        f6_add_mul_upow<upow[q]>(acc, pq);
    };
    switch (wq) {  // uniform per wave
    case 0: component(std::integral_constant<int, 0>{}); break;
    case 1: component(std::integral_constant<int, 1>{}); break;
    case 2: component(std::integral_constant<int, 2>{}); break;
    case 3: component(std::integral_constant<int, 3>{}); break;
    case 4: component(std::integral_constant<int, 4>{}); break;
    default: component(std::integral_constant<int, 5>{}); break;
    }
#pragma unroll
    for (int q = 0; q < 6; q++) red[wq][q][lane] = acc.c[q];
    __syncthreads();
    if (wq != 0 || !live) return;
#pragma unroll
    for (int w = 1; w < 6; w++)
#pragma unroll
        for (int q = 0; q < 6; q++) acc.c[q] = add(acc.c[q], red[w][q][lane]);
#pragma unroll
    for (int q = 0; q < 6; q++) fnext[q * nn + b] = acc.c[q];
}

template <int LM>
void launch_fold(Ctx *c, const uint64_t *f, uint64_t n, int ncomp, uint32_t K, const uint64_t *d_apow,
                 const FoldPow &pa, uint64_t *fnext, uint64_t *tuples) {
    const uint64_t nn = n >> LM;
    const uint64_t minv = inv(1ULL << LM);
    ProfScope ps(c, "coset_fold", (8.0 * ncomp * n + 48.0 * nn + (tuples ? 48.0 * n : 0)) * K);
    hipLaunchKernelGGL(coset_fold_k<LM>, dim3((uint32_t)((nn + 63) / 64), K), dim3(FOLD_T), 0, c->stream, f, n, ncomp,
                       d_apow, pa, minv, c->rt_lo, c->rt_hi, ilog2(n), fnext, tuples);
    SG_HIP(hipGetLastError());
}

void upload_powers(Ctx *c, const std::vector<fp6> &alpha, uint64_t m, uint64_t *d);

// alpha: one folding challenge per proof; its powers alpha^k (k < m) go to
// the kernel as an argument for one proof, else through d_apow (K m Fp6)
void fold(Ctx *c, const uint64_t *f, uint64_t n, int ncomp, uint64_t m, uint32_t K, const std::vector<fp6> &alpha,
          uint64_t *d_apow, uint64_t *fnext, uint64_t *tuples) {
    SG_REQUIRE(alpha.size() == K && m <= 16, "fold: one alpha per proof, arity at most 16");
    FoldPow pa{};
    const uint64_t *dp = nullptr;
    if (K == 1) {
        fp6 x = f6_one();
        for (uint64_t j = 0; j < m; j++) {
            memcpy(&pa.v[6 * j], x.c, 48);
            x = f6_mul(x, alpha[0]);
        }
    } else {
        upload_powers(c, alpha, m, d_apow);
        dp = d_apow;
    }
    switch (m) {
    case 2: launch_fold<1>(c, f, n, ncomp, K, dp, pa, fnext, tuples); break;
    case 4: launch_fold<2>(c, f, n, ncomp, K, dp, pa, fnext, tuples); break;
    case 8: launch_fold<3>(c, f, n, ncomp, K, dp, pa, fnext, tuples); break;
    case 16: launch_fold<4>(c, f, n, ncomp, K, dp, pa, fnext, tuples); break;
    default: throw Error(E_UNSUPPORTED, "fold arity must be 2, 4, 8 or 16");
    }
}

// coset evaluations of K proofs' layers at z_k zeta^j into d_ev (device,
// K x m x 6), queued on the stream; d_T: per proof tstride entries
// z0^(2^i); this layer's z = z0^(2^toff).
template <int M>
void launch_ood(Ctx *c, const uint64_t *coef, uint64_t n, int ncomp, uint32_t K, const uint64_t *d_T,
                uint32_t tstride, uint32_t toff, uint64_t zeta, uint64_t *d_ev) {
    SG_REQUIRE(n <= (1ULL << 24) && toff + 24 <= tstride, "ood: table too short");
    // elements per thread: enough blocks to fill the chip, at most 64
    uint32_t ch = 64;
    while (ch > 1 && (n / ((uint64_t)OT * ch)) * K < 8192) ch >>= 1;
    const uint32_t nblocks = (uint32_t)((n + (uint64_t)OT * ch - 1) / ((uint64_t)OT * ch));
    Scoped d_part(c, 48ULL * M * nblocks * K);
    {
        ProfScope ps(c, "ood_sums", 8.0 * ncomp * n * K);
        hipLaunchKernelGGL(ood_sums_k<M>, dim3(nblocks, K), dim3(OT), 0, c->stream, coef, n, ncomp, d_T + 6 * toff,
                           tstride, ch, d_part.as<uint64_t>());
        SG_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(ood_finish_k, dim3(K), dim3(FIN_T), 0, c->stream, d_part.as<uint64_t>(), nblocks, (uint32_t)M,
                       zeta, d_ev);
    SG_HIP(hipGetLastError());
}

void ood(Ctx *c, const uint64_t *coef, uint64_t n, int ncomp, uint64_t m, uint32_t K, const uint64_t *d_T,
         uint32_t tstride, uint32_t toff, uint64_t zeta, uint64_t *d_ev) {
    switch (m) {
    case 2: launch_ood<2>(c, coef, n, ncomp, K, d_T, tstride, toff, zeta, d_ev); break;
    case 4: launch_ood<4>(c, coef, n, ncomp, K, d_T, tstride, toff, zeta, d_ev); break;
    case 8: launch_ood<8>(c, coef, n, ncomp, K, d_T, tstride, toff, zeta, d_ev); break;
    case 16: launch_ood<16>(c, coef, n, ncomp, K, d_T, tstride, toff, zeta, d_ev); break;
    default: throw Error(E_UNSUPPORTED, "fold arity must be 2, 4, 8 or 16");
    }
}

// Classic-FRI layer leaves (fri.rs:255-333, 1356-1362): planes [K][18][n] =
// f (6 components), s (6), q (6) per point. This kernel writes f and the DEEP
// quotient q_i = (f_i - f(z)) / (w_n^i - z), with f(z) the first OOD coset
// evaluation of the proof (ev: [K][m][6]) and z its table entry z^(2^0).
__global__ __launch_bounds__(FT) void deep_quotient_k(const uint64_t *f, uint64_t n, int ncomp, const uint64_t *ev,
                                                      uint32_t m, const uint64_t *ztab, uint32_t tstride,
                                                      const uint64_t *rt_lo, const uint64_t *rt_hi, uint32_t logn,
                                                      uint64_t *planes) {
    const uint64_t i = (uint64_t)blockIdx.x * FT + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = blockIdx.y;
    f += k * ncomp * n;
    ev += k * m * 6;
    ztab += k * tstride * 6;
    planes += k * 18 * n;
    fp6 fv, fz, z;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        fv.c[q] = q < ncomp ? f[q * n + i] : 0;
        fz.c[q] = ev[q];
        z.c[q] = ztab[q];
    }
    const uint64_t e26 = (i << (26 - logn)) & ((1ULL << 26) - 1);
    const uint64_t x = mul(rt_lo[e26 & 8191], rt_hi[e26 >> 13]);
    const fp6 qv = f6_mul(f6_sub(fv, fz), f6_inv(f6_sub(f6_base(x), z)));
#pragma unroll
    for (int q = 0; q < 6; q++) {
        planes[q * n + i] = fv.c[q];
        planes[(12 + q) * n + i] = qv.c[q];
    }
}

// s planes: s_i = f_{l+1}[i mod n/m] (the fold, repeated over the fibres,
// fri.rs:773-789), or 0 (next == nullptr) on the coefficient-committed final layer
__global__ __launch_bounds__(FT) void s_planes_k(const uint64_t *next, uint64_t n, uint64_t nn, uint64_t *planes) {
    const uint64_t i = (uint64_t)blockIdx.x * FT + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = blockIdx.y;
    planes += k * 18 * n + 6 * n;
#pragma unroll
    for (int q = 0; q < 6; q++) planes[q * n + i] = next ? next[k * 6 * nn + q * nn + (i & (nn - 1))] : 0;
}

struct Bytes {
    std::vector<uint8_t> v;
    void put(const void *d, size_t n) { v.insert(v.end(), (const uint8_t *)d, (const uint8_t *)d + n); }
    void u8(uint8_t x) { v.push_back(x); }
    void u64(uint64_t x) {
        for (int i = 0; i < 8; i++) v.push_back((uint8_t)(x >> (8 * i)));
    }
    void f6(const uint64_t *c) {
        for (int i = 0; i < 6; i++) u64(c[i]);
    }
};

// alpha_k^j for j < m, proof-major, to d (K * m Fp6)
void upload_powers(Ctx *c, const std::vector<fp6> &alpha, uint64_t m, uint64_t *d) {
    std::vector<fp6> pw(alpha.size() * m);
    for (size_t k = 0; k < alpha.size(); k++) {
        fp6 x = f6_one();
        for (uint64_t j = 0; j < m; j++) {
            pw[k * m + j] = x;
            x = f6_mul(x, alpha[k]);
        }
    }
    SG_HIP(hipMemcpyAsync(d, pw.data(), 48 * pw.size(), hipMemcpyHostToDevice, c->stream));
}


// One Fiat-Shamir transcript per proof with the statement bound
// (bind_statement_to_transcript, fri.rs:1105-1129)
std::vector<Transcript> statement_transcripts(uint32_t K, uint64_t n0, const FriParams &p,
                                              const std::vector<std::array<uint8_t, 32>> *pis) {
    std::vector<Transcript> tr;
    tr.reserve(K);
    for (uint32_t k = 0; k < K; k++) {
        tr.emplace_back("FRI/FS");
        Transcript &t = tr.back();
        if (pis || p.has_pi) {
            t.absorb("CAIRO-PUBLIC-INPUTS-V1");
            t.absorb(pis ? (*pis)[k].data() : p.pi, 32);
        }
        t.absorb("DEEP-FRI-STATEMENT-V2");
        t.absorb_field(n0);
        t.absorb_field(p.schedule.size());
        for (uint64_t m : p.schedule) t.absorb_field(m);
        t.absorb_field(canon(p.seed_z));
        t.absorb_field(6);
        t.absorb_field(p.coeff_commit_final ? 1 : 0);
        t.absorb_field(p.stir ? 1 : 0);
    }
    return tr;
}

// first d_final coefficients of each proof's last layer (fri.rs:1919-1937),
// absorbed, then the query seed (fri.rs:1939-1993)
void final_poly_and_seed(Ctx *c, const uint64_t *f, uint64_t nf, uint32_t K, uint64_t d_final,
                         std::vector<Transcript> &tr, std::vector<std::vector<fp6>> &final_coeffs,
                         std::vector<uint64_t> &query_seed) {
    std::vector<uint64_t> fin(6 * nf * K);
    {
        Scoped coef(c, 8 * 6 * nf * K), scr(c, 8 * 6 * nf * K);
        const uint64_t *src = f;
        if (nf > 1) {
            ntt_columns(c, f, coef.as<uint64_t>(), scr.as<uint64_t>(), ilog2(nf), 6ULL * K, nf, nf, true, nullptr);
            src = coef.as<uint64_t>();
        }
        c->d2h(fin.data(), src, 8 * fin.size());
        SG_HIP(hipStreamSynchronize(c->stream));
    }
    const uint64_t dfin = d_final < nf ? d_final : nf;
    final_coeffs.assign(K, std::vector<fp6>(dfin));
    query_seed.assign(K, 0);
    parallel_for(K, [&](size_t k) {
        for (uint64_t i = 0; i < dfin; i++)
            for (int q = 0; q < 6; q++) final_coeffs[k][i].c[q] = fin[(k * 6 + q) * nf + i];
        for (auto &fc : final_coeffs[k])
            for (int q = 0; q < 6; q++) tr[k].absorb_field(fc.c[q]);
        query_seed[k] = tr[k].challenge_field("query_seed");
    });
}

// Classic DEEP-FRI (LdtMode::Fri): per layer the DEEP quotient q at the one
// out-of-domain point z, the fold s repeated over the fibres, a tree over the
// 18-value leaves (f, s, q), and r queries followed down every layer
// (fri.rs:1461-1600 with use_stir false, fri_prove_queries :1724-1807, the
// payloads of deep_fri_prove :2008-2032). K proofs in lockstep, like the STIR
// prover.
std::vector<std::vector<uint8_t>> deep_fri_prove_classic(Ctx *c, const uint64_t *d_f0, uint32_t K, uint64_t n0,
                                                         const FriParams &p,
                                                         const std::vector<std::array<uint8_t, 32>> *pis) {
    const size_t L = p.schedule.size();
    const bool use_cc = p.coeff_commit_final;
    std::vector<Transcript> tr = statement_transcripts(K, n0, p, pis);
    // f0 tree: one value per leaf (fri.rs:1432-1440, f0_tree_config :1068-1072)
    DevTree t0;
    {
        HostScope hs(c, "fri_f0_trees");
        const uint32_t a = pick_arity(n0, 16);
        LeafSrc src{d_f0, 1, n0, 1, 1, nullptr, n0};
        tree_build(c, t0, src, n0, a, merkle_depth(n0, a), 0xFF, K);
    }
    // the one out-of-domain point z of every layer, as entry 0 of the OOD
    // kernels' table of squares z^(2^i)
    constexpr uint32_t tstride = 24;
    std::vector<fp6> alpha(K);
    std::vector<uint64_t> ztab((size_t)K * tstride * 6);
    parallel_for(K, [&](size_t k) {
        tr[k].absorb(t0.roots[k].data(), 32);
        fp6 w;
        tr[k].challenge_ext("z_fp3", w.c);
        for (uint32_t i = 0; i < tstride; i++) {
            memcpy(&ztab[((size_t)k * tstride + i) * 6], w.c, 48);
            w = f6_mul(w, w);
        }
    });
    Scoped d_ztab(c, 8 * ztab.size());
    SG_HIP(hipMemcpyAsync(d_ztab.p, ztab.data(), 8 * ztab.size(), hipMemcpyHostToDevice, c->stream));
    Scoped d_alpha(c, 48ULL * 16 * K);

    std::vector<uint64_t *> planes(L, nullptr);  // [K][18][n_l] per layer, kept for the queries
    std::vector<DevTree> trees(L);
    std::vector<uint64_t> sz(L + 1);
    std::vector<std::vector<uint64_t>> fz(L, std::vector<uint64_t>(6 * K));
    std::vector<std::array<uint8_t, 32>> roots((size_t)L * K), coeff_roots(K);
    std::vector<uint64_t> tuples_planes;
    uint64_t n_final_tuples = 0, m_final = 0;
    uint64_t cur = n0;
    const uint64_t *f = d_f0;
    int ncomp = 1;
    uint64_t *owned = nullptr;
    for (size_t ell = 0; ell < L; ell++) {
        const uint64_t m = p.schedule[ell], nn = cur / m;
        const bool final_cc = use_cc && ell == L - 1;
        sz[ell] = cur;
        if (!final_cc) parallel_for(K, [&](size_t k) { tr[k].challenge_ext("alpha", alpha[k].c); });
        Scoped d_ev(c, 48ULL * m * K);
        {
            // f(z): the first of the OOD coset evaluations at z (fri.rs:255-299)
            Scoped coef(c, 8ULL * K * ncomp * cur), scr(c, 8ULL * K * ncomp * cur);
            ntt_columns(c, f, coef.as<uint64_t>(), scr.as<uint64_t>(), ilog2(cur), (uint64_t)K * ncomp, cur, cur, true,
                        nullptr);
            ood(c, coef.as<uint64_t>(), cur, ncomp, m, K, d_ztab.as<uint64_t>(), tstride, 0,
                pow(root_of_unity(cur), nn), d_ev.as<uint64_t>());
        }
        planes[ell] = c->alloc_n<uint64_t>(18ULL * cur * K);
        const dim3 grid((uint32_t)((cur + FT - 1) / FT), K);
        {
            ProfScope ps(c, "deep_quotient", (8.0 * ncomp * cur + 96.0 * cur) * K);
            hipLaunchKernelGGL(deep_quotient_k, grid, dim3(FT), 0, c->stream, f, cur, ncomp, d_ev.as<uint64_t>(),
                               (uint32_t)m, d_ztab.as<uint64_t>(), tstride, c->rt_lo, c->rt_hi, ilog2(cur),
                               planes[ell]);
            SG_HIP(hipGetLastError());
        }
        uint64_t *next = c->alloc_n<uint64_t>(6ULL * nn * K);
        if (!final_cc) {
            fold(c, f, cur, ncomp, m, K, alpha, d_alpha.as<uint64_t>(), next, nullptr);
        }
        hipLaunchKernelGGL(s_planes_k, grid, dim3(FT), 0, c->stream, final_cc ? nullptr : next, cur, nn, planes[ell]);
        SG_HIP(hipGetLastError());
        {
            HostScope hs(c, "fri_layer_tree");
            const uint32_t a = pick_arity(cur, m);
            LeafSrc src{planes[ell], 1, cur, 18, 18, nullptr, 18 * cur};
            tree_build(c, trees[ell], src, cur, a, merkle_depth(cur, a), ell, K, false);
            std::vector<uint64_t> ev((size_t)K * m * 6);
            tree_read_roots_with(c, trees[ell], ev.data(), d_ev.p, 8 * ev.size());
            for (uint32_t k = 0; k < K; k++) memcpy(&fz[ell][6 * k], &ev[(size_t)k * m * 6], 48);
        }
        parallel_for(K, [&](size_t k) {
            roots[ell * K + k] = trees[ell].roots[k];
            for (int q = 0; q < 6; q++) tr[k].absorb_field(fz[ell][6 * k + q]);
            tr[k].absorb(trees[ell].roots[k].data(), 32);
        });
        if (final_cc) {
            // coefficient tuples and their tree (fri.rs:1649-1688), then this layer's alpha
            Scoped tup(c, 8ULL * 6 * m * nn * K);
            std::vector<fp6> zero(K, f6_zero());
            fold(c, f, cur, ncomp, m, K, zero, d_alpha.as<uint64_t>(), next, tup.as<uint64_t>());
            const uint32_t a = pick_arity(nn, 16);
            DevTree tc;
            LeafSrc src{tup.as<uint64_t>(), 1, nn, (uint32_t)(6 * m), (uint32_t)(6 * m), nullptr, 6 * m * nn};
            tree_build(c, tc, src, nn, a, merkle_depth(nn, a), 0xFE, K, false);
            tuples_planes.resize((size_t)6 * m * nn * K);
            tree_read_roots_with(c, tc, tuples_planes.data(), tup.p, 8 * tuples_planes.size());
            n_final_tuples = nn;
            m_final = m;
            parallel_for(K, [&](size_t k) {
                coeff_roots[k] = tc.roots[k];
                tr[k].absorb(tc.roots[k].data(), 32);
                tr[k].challenge_ext("alpha", alpha[k].c);
                fp6 beta;
                tr[k].challenge_ext("beta_deg", beta.c);
            });
            tree_free(c, tc);
            fold(c, f, cur, ncomp, m, K, alpha, d_alpha.as<uint64_t>(), next, nullptr);
        }
        if (owned) c->release(owned);
        owned = next;
        f = next;
        ncomp = 6;
        cur = nn;
    }
    sz[L] = cur;
    std::vector<std::vector<fp6>> final_coeffs;
    std::vector<uint64_t> query_seed;
    final_poly_and_seed(c, f, cur, K, p.d_final, tr, final_coeffs, query_seed);
    if (owned) c->release(owned);

    // query positions: i_0 = the raw index, i_{l+1} = i_l mod n_l / m_l
    std::vector<std::vector<uint64_t>> raw;
    fri_query_indices(c, query_seed, p.r, n0, raw);
    const uint64_t r = p.r;
    std::vector<std::vector<std::vector<uint64_t>>> pos(L + 1, std::vector<std::vector<uint64_t>>(K));
    for (uint32_t k = 0; k < K; k++)
        for (uint64_t q = 0; q < r; q++) {
            uint64_t i = raw[k][q];
            for (size_t ell = 0; ell <= L; ell++) {
                pos[ell][k].push_back(i);
                if (ell < L) i %= sz[ell] / p.schedule[ell];
            }
        }
    // payloads (f, s, q) of every query and layer, gathered from the leaf planes
    // every layer's payload gather and every tree's opening digests behind one wait
    std::vector<HostBuf> pay(L);
    std::vector<std::vector<std::vector<std::vector<uint8_t>>>> lop(L);
    std::vector<std::vector<std::vector<uint8_t>>> f0op;
    {
        HostScope hs(c, "fri_gather_openings");
        std::vector<GatherJob> jobs(2 * L + 1);
        for (size_t ell = 0; ell < L; ell++) {
            const uint64_t n = sz[ell];
            GatherJob &g = jobs[ell];
            g.src = planes[ell];
            g.idx.resize((size_t)K * r * 18);
            parallel_for(K, [&](size_t k) {
                uint64_t *w = &g.idx[k * r * 18];
                for (uint64_t q = 0; q < r; q++)
                    for (uint64_t j = 0; j < 18; j++) *w++ = k * 18 * n + j * n + pos[ell][k][q];
            });
            jobs[L + ell] = tree_open_job(trees[ell], pos[ell]);
        }
        jobs[2 * L] = tree_open_job(t0, pos[0]);
        gather_set(c, jobs);
        for (size_t ell = 0; ell < L; ell++) {
            pay[ell] = std::move(jobs[ell].out);
            tree_open_finish(trees[ell], pos[ell], jobs[L + ell], lop[ell]);
        }
        tree_open_finish(t0, pos[0], jobs[2 * L], f0op);
    }
    for (size_t ell = 0; ell < L; ell++) {
        tree_free(c, trees[ell]);
        c->release(planes[ell]);
    }
    tree_free(c, t0);

    HostScope hs(c, "fri_serialize");
    std::vector<std::vector<uint8_t>> proofs(K);
    const uint64_t omega0 = root_of_unity(n0);
    parallel_for(K, [&](size_t k) {
        Bytes b;
        b.put(t0.roots[k].data(), 32);
        b.u64(L);
        for (size_t l = 0; l < L; l++) b.put(roots[l * K + k].data(), 32);
        b.u64(L);  // layer_proofs.layers
        for (size_t l = 0; l < L; l++) {
            b.u64(r);
            for (uint64_t q = 0; q < r; q++) b.put(lop[l][k][q].data(), lop[l][k][q].size());
        }
        b.u64(r);  // f0_openings
        for (uint64_t q = 0; q < r; q++) b.put(f0op[k][q].data(), f0op[k][q].size());
        b.u64(r);  // queries
        for (uint64_t q = 0; q < r; q++) {
            b.u64(L);
            for (size_t l = 0; l < L; l++) {
                const uint64_t i = pos[l][k][q], m = p.schedule[l];
                b.u64(i);
                b.u64(i % m);
                b.u64(i % (sz[l] / m));
                b.u64(0);
            }
            b.u64(L);
            for (size_t l = 0; l < L; l++) b.put(&pay[l][((size_t)k * r + q) * 18], 8 * 18);
            b.put(f0op[k][q].data(), f0op[k][q].size());
            b.u64(pos[L][k][q]);
        }
        b.u64(L);
        for (size_t l = 0; l < L; l++) b.f6(&fz[l][6 * k]);
        b.u64(final_coeffs[k].size());
        for (auto &v : final_coeffs[k]) b.f6(v.c);
        b.u64(n0);
        b.u64(omega0);
        if (use_cc) {
            const uint64_t nt = n_final_tuples, m = m_final;
            const uint64_t *tp = &tuples_planes[(size_t)k * 6 * m * nt];
            b.u8(1);
            b.u64(nt);
            for (uint64_t t = 0; t < nt; t++) {
                b.u64(m);
                for (uint64_t kq = 0; kq < 6 * m; kq++) b.u64(tp[kq * nt + t]);
            }
            b.u8(1);
            b.put(coeff_roots[k].data(), 32);
        } else {
            b.u8(0);
            b.u8(0);
        }
        b.u8(0);  // stir_coset_evals: None
        b.u8(0);  // stir_proximity_queries: None
        proofs[k] = std::move(b.v);
    });
    return proofs;
}
}  // namespace

std::vector<uint64_t> stir_schedule(uint64_t n0) {
    const uint32_t lg = ilog2(n0);
    std::vector<uint64_t> s(lg / 3, 8);
    if (lg % 3) s.push_back(1ULL << (lg % 3));
    return s;
}

std::vector<std::vector<uint8_t>> deep_fri_prove_batch(Ctx *c, const uint64_t *d_f0, uint32_t K, uint64_t n0,
                                                       const FriParams &p, const std::vector<std::array<uint8_t, 32>> *pis,
                                                       std::vector<std::vector<uint64_t>> *raw_out) {
    const size_t L = p.schedule.size();
    SG_REQUIRE(L > 0, "deep_fri_prove: empty fold schedule");
    SG_REQUIRE(K >= 1 && K <= 65535, "deep_fri_prove: batch size out of range");
    SG_REQUIRE(is_pow2(n0) && n0 >= 2 && n0 <= (1ULL << 24), "deep_fri_prove: n0 must be a power of two in [2, 2^24]");
    SG_REQUIRE(!pis || pis->size() == K, "deep_fri_prove: one public-inputs hash per proof");
    {
        uint64_t n = n0;
        for (uint64_t m : p.schedule) {
            SG_REQUIRE(m >= 2 && m <= 16 && is_pow2(m) && n % m == 0, "deep_fri_prove: bad fold schedule");
            n /= m;
        }
    }
    if (!p.stir) {
        SG_REQUIRE(!raw_out, "deep_fri_prove: raw STIR query indices asked of a classic-FRI proof");
        return deep_fri_prove_classic(c, d_f0, K, n0, p, pis);
    }
    const bool use_cc = p.coeff_commit_final;
    std::vector<Transcript> tr = statement_transcripts(K, n0, p, pis);

    // coset-packed f0 trees: leaf b = fibre f0[b + j n0/m0] (fri.rs:1420-1431)
    const uint64_t m0 = p.schedule[0], nn0 = n0 / m0;
    DevTree t0;
    {
        HostScope hs(c, "fri_f0_trees");
        const uint32_t a = pick_arity(nn0, 16);
        LeafSrc src{d_f0, 1, nn0, (uint32_t)m0, (uint32_t)m0, nullptr, n0};
        tree_build(c, t0, src, nn0, a, merkle_depth(nn0, a), 0xFF, K);
    }
    std::vector<fp6> z(K), alpha(K);
    // z_l = z^(m_0 ... m_{l-1}): one table of squares z^(2^i) serves every layer
    uint32_t tstride = 24;
    for (uint64_t m : p.schedule) tstride += ilog2(m);
    std::vector<uint64_t> ztab((size_t)K * tstride * 6);
    parallel_for(K, [&](size_t k) {
        tr[k].absorb(t0.roots[k].data(), 32);
        tr[k].challenge_ext("z_fp3", z[k].c);
        fp6 w = z[k];
        for (uint32_t i = 0; i < tstride; i++) {
            memcpy(&ztab[((size_t)k * tstride + i) * 6], w.c, 48);
            w = f6_mul(w, w);
        }
    });
    Scoped d_ztab(c, 8 * ztab.size());
    SG_HIP(hipMemcpyAsync(d_ztab.p, ztab.data(), 8 * ztab.size(), hipMemcpyHostToDevice, c->stream));
    uint32_t toff = 0;

    // layers
    std::vector<std::vector<uint64_t>> coset_evals(L);  // [l][k * m * 6]
    std::vector<std::array<uint8_t, 32>> roots((size_t)L * K);
    std::vector<uint64_t> tuples_planes;  // final tuples, [k][6m][nn]
    std::vector<std::array<uint8_t, 32>> coeff_roots(K);
    uint64_t n_final_tuples = 0, m_final = 0;
    DevTree t1;
    bool have_t1 = false;
    Scoped d_alpha(c, 48ULL * 16 * K);  // alpha powers, m <= 16

    uint64_t cur = n0;
    const uint64_t *f = d_f0;  // layer planes of all proofs
    int ncomp = 1;
    uint64_t *owned = nullptr;  // current layer planes we allocated
    uint64_t *layer1 = nullptr; // f_1 planes, kept for the queries
    for (size_t ell = 0; ell < L; ell++) {
        const uint64_t m = p.schedule[ell], nn = cur / m;
        const bool final_cc = use_cc && ell == L - 1;
        if (!final_cc) parallel_for(K, [&](size_t k) { tr[k].challenge_ext("alpha", alpha[k].c); });
        // OOD coset evaluations
        coset_evals[ell].resize((size_t)K * m * 6);
        Scoped d_ev(c, 48ULL * m * K);
        {
            Scoped coef(c, 8ULL * K * ncomp * cur), scr(c, 8ULL * K * ncomp * cur);
            ntt_columns(c, f, coef.as<uint64_t>(), scr.as<uint64_t>(), ilog2(cur), (uint64_t)K * ncomp, cur, cur, true,
                        nullptr);
            const uint64_t zeta = pow(root_of_unity(cur), nn);
            ood(c, coef.as<uint64_t>(), cur, ncomp, m, K, d_ztab.as<uint64_t>(), tstride, toff, zeta,
                d_ev.as<uint64_t>());
        }
        toff += ilog2(m);
        // layer commitment: leaf = 6 components (fri.rs:1364-1366); a base-field
        // layer (f0) hashes components 1..5 as zeros. Queued behind the OOD
        // kernels, so one wait returns both the coset evaluations and the roots.
        DevTree tl;
        {
            HostScope hs(c, "fri_ood_and_layer_tree");
            const uint32_t a = pick_arity(cur, m);
            LeafSrc src{f, 1, cur, 6, (uint32_t)ncomp, nullptr, (uint64_t)ncomp * cur};
            tree_build(c, tl, src, cur, a, merkle_depth(cur, a), ell, K, false);
            tree_read_roots_with(c, tl, coset_evals[ell].data(), d_ev.p, 48ULL * m * K);
        }
        parallel_for(K, [&](size_t k) {
            roots[ell * K + k] = tl.roots[k];
            for (uint64_t j = 0; j < 6 * m; j++) tr[k].absorb_field(coset_evals[ell][k * m * 6 + j]);
            tr[k].absorb(tl.roots[k].data(), 32);
        });
        if (ell == 1) {
            t1 = tl;
            have_t1 = true;
        } else {
            tree_free(c, tl);
        }
        uint64_t *next = c->alloc_n<uint64_t>(6ULL * nn * K);
        if (final_cc) {
            Scoped tup(c, 8ULL * 6 * m * nn * K);
            std::vector<fp6> zero(K, f6_zero());
            fold(c, f, cur, ncomp, m, K, zero, d_alpha.as<uint64_t>(), next, tup.as<uint64_t>());
            const uint32_t a = pick_arity(nn, 16);
            DevTree tc;
            LeafSrc src{tup.as<uint64_t>(), 1, nn, (uint32_t)(6 * m), (uint32_t)(6 * m), nullptr, 6 * m * nn};
            tree_build(c, tc, src, nn, a, merkle_depth(nn, a), 0xFE, K, false);
            tuples_planes.resize((size_t)6 * m * nn * K);
            tree_read_roots_with(c, tc, tuples_planes.data(), tup.p, 8 * tuples_planes.size());
            n_final_tuples = nn;
            m_final = m;
            parallel_for(K, [&](size_t k) {
                coeff_roots[k] = tc.roots[k];
                tr[k].absorb(tc.roots[k].data(), 32);
                tr[k].challenge_ext("alpha", alpha[k].c);
                fp6 beta;
                tr[k].challenge_ext("beta_deg", beta.c);
            });
            tree_free(c, tc);
        }
        // fold with this layer's alpha (the final layer's is drawn after its coefficient commitment)
        HostScope hfold(c, "fri_fold");
        fold(c, f, cur, ncomp, m, K, alpha, d_alpha.as<uint64_t>(), next, nullptr);
        if (ell == 0) {
            layer1 = next;
        } else if (owned && owned != layer1) {
            c->release(owned);
        }
        owned = next;
        f = next;
        ncomp = 6;
        cur = nn;
    }

    // final polynomial: first d_final coefficients of the last layer (fri.rs:1919-1937)
    std::vector<std::vector<fp6>> final_coeffs;
    std::vector<uint64_t> query_seed;
    final_poly_and_seed(c, f, cur, K, p.d_final, tr, final_coeffs, query_seed);
    const uint64_t dfin = final_coeffs[0].size();
    std::vector<std::vector<uint64_t>> raw, base(K, std::vector<uint64_t>(p.s0));
    // STIR proximity query indices (fri.rs:1813-1893)
    fri_query_indices(c, query_seed, p.s0, n0, raw);
    for (uint32_t k = 0; k < K; k++)
        for (uint64_t q = 0; q < p.s0; q++) base[k][q] = raw[k][q] % nn0;
    if (raw_out) *raw_out = raw;
    const bool l2 = L >= 2;
    const uint64_t n1 = n0 / m0;
    // gathers: fibre values of f0 and, per component, f_1[base]
    std::vector<uint64_t> gidx;
    for (uint32_t k = 0; k < K; k++)
        for (uint64_t q = 0; q < p.s0; q++)
            for (uint64_t j = 0; j < m0; j++) gidx.push_back(k * n0 + base[k][q] + j * nn0);
    const size_t nfib = gidx.size();
    if (l2)
        for (uint32_t k = 0; k < K; k++)
            for (uint64_t q = 0; q < p.s0; q++)
                for (int q6 = 0; q6 < 6; q6++) gidx.push_back(k * 6 * n1 + q6 * n1 + base[k][q]);
    // the payload gathers and both trees' opening digests behind one wait
    std::vector<uint64_t> gv(gidx.size());
    std::vector<std::vector<std::vector<uint8_t>>> op0, op1;
    {
        HostScope hs(c, "fri_gather_openings");
        std::vector<GatherJob> jobs(l2 ? 4 : 2);
        jobs[0].src = d_f0;
        jobs[0].idx.assign(gidx.begin(), gidx.begin() + nfib);
        jobs[1] = tree_open_job(t0, base);
        if (l2) {
            jobs[2].src = layer1;
            jobs[2].idx.assign(gidx.begin() + nfib, gidx.end());
            jobs[3] = tree_open_job(t1, base);
        }
        gather_set(c, jobs);
        std::copy(jobs[0].out.begin(), jobs[0].out.end(), gv.begin());
        tree_open_finish(t0, base, jobs[1], op0);
        if (l2) {
            std::copy(jobs[2].out.begin(), jobs[2].out.end(), gv.begin() + nfib);
            tree_open_finish(t1, base, jobs[3], op1);
        }
    }
    HostScope hs(c, "fri_serialize");

    // serialize DeepFriProof<SexticExt> (fri.rs:1286-1300, ark-serialize compressed)
    std::vector<std::vector<uint8_t>> proofs(K);
    const uint64_t omega0 = root_of_unity(n0);
    parallel_for(K, [&](size_t k) {
        Bytes b;
        b.put(t0.roots[k].data(), 32);
        b.u64(L);
        for (size_t l = 0; l < L; l++) b.put(roots[l * K + k].data(), 32);
        b.u64(0);  // layer_proofs.layers
        b.u64(0);  // f0_openings
        b.u64(0);  // queries
        b.u64(L);
        for (size_t l = 0; l < L; l++) b.f6(&coset_evals[l][(size_t)k * p.schedule[l] * 6]);  // fz = first coset eval
        b.u64(dfin);
        for (auto &v : final_coeffs[k]) b.f6(v.c);
        b.u64(n0);
        b.u64(omega0);
        if (use_cc) {
            const uint64_t nt = n_final_tuples, m = m_final;
            const uint64_t *tp = &tuples_planes[(size_t)k * 6 * m * nt];
            b.u8(1);
            b.u64(nt);
            for (uint64_t t = 0; t < nt; t++) {
                b.u64(m);
                for (uint64_t kq = 0; kq < 6 * m; kq++) b.u64(tp[kq * nt + t]);
            }
            b.u8(1);
            b.put(coeff_roots[k].data(), 32);
        } else {
            b.u8(0);
            b.u8(0);
        }
        b.u8(1);
        b.u64(L);
        for (size_t l = 0; l < L; l++) {
            const uint64_t m = p.schedule[l];
            b.u64(m);
            for (uint64_t j = 0; j < m; j++) b.f6(&coset_evals[l][((size_t)k * m + j) * 6]);
        }
        b.u8(1);
        b.u64(p.s0);
        for (uint64_t q = 0; q < p.s0; q++) {
            b.u64(base[k][q]);
            b.u64(raw[k][q]);
            b.u64(m0);
            for (uint64_t j = 0; j < m0; j++) b.u64(base[k][q] + j * nn0);
            b.u64(m0);
            for (uint64_t j = 0; j < m0; j++) b.u64(gv[((size_t)k * p.s0 + q) * m0 + j]);
            b.put(op0[k][q].data(), op0[k][q].size());
            if (l2) {
                b.f6(&gv[nfib + ((size_t)k * p.s0 + q) * 6]);
                b.u8(1);
                b.put(op1[k][q].data(), op1[k][q].size());
            } else {
                const uint64_t zero6[6] = {0, 0, 0, 0, 0, 0};
                b.f6(zero6);
                b.u8(0);
            }
        }
        proofs[k] = std::move(b.v);
    });
    tree_free(c, t0);
    if (have_t1) tree_free(c, t1);
    if (owned && owned != layer1) c->release(owned);
    if (layer1) c->release(layer1);
    return proofs;
}

std::vector<uint8_t> deep_fri_prove_dev(Ctx *c, const uint64_t *d_f0, uint64_t n0, const FriParams &p,
                                        std::vector<uint64_t> *raw_out) {
    std::vector<std::vector<uint64_t>> raw;
    std::vector<std::vector<uint8_t>> v = deep_fri_prove_batch(c, d_f0, 1, n0, p, nullptr, raw_out ? &raw : nullptr);
    if (raw_out) *raw_out = raw[0];
    return std::move(v[0]);
}

}  // namespace sg
