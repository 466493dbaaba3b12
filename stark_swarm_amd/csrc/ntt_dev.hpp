// ntt_dev.hpp -- device building blocks of the Goldilocks NTT kernels for
// gfx950: the 8K- or 16K-element LDS tile, radix <= 16 register DFTs whose twiddles
// are powers of 2 (w_32 = 2^78), and the Stockham steps that run S-point
// sub-DFTs of a tile in place. Shared by ntt.hip and the kernels that fuse a
// transform with their own loads (p256_air.hip).
#pragma once
#include "gl.hpp"
#include "ntt.hpp"

namespace sg {
namespace {

constexpr int TILE_LOG = 13;
constexpr int TILE = 1 << TILE_LOG;
constexpr int NT = 512;
constexpr int EPT = TILE / NT;  // elements held per lane during a step
constexpr int MAX_LR = 4;       // register DFTs of at most 2^MAX_LR points
constexpr int PAD_SHIFT = 4;
// padded tile size: one u64 per 16, plus the sub-DFT skew of padi
constexpr int LDS_ELEMS = TILE + (TILE >> PAD_SHIFT) + (TILE >> 5);

// LDS index of tile element e for sub-DFTs of 2^LOG_S points. One u64 of
// padding per 16 breaks the bank conflicts of the strided Stockham accesses;
// for LOG_S >= 6 each sub-DFT also starts 2 u64 after the previous one ends,
// so a sub-DFT's stride is 2 mod 16 and the g-fast loads and stores (the 16
// lanes of a quarter-wave on 8 sub-DFTs at 2 neighbouring offsets) land in
// distinct banks instead of 8 to a bank.
template <int LOG_S>
__device__ __forceinline__ uint32_t padi(uint32_t e) {
    if constexpr (LOG_S >= 6)
        return e + (e >> PAD_SHIFT) + 2 * (e >> LOG_S);
    else
        return e + (e >> PAD_SHIFT);
}

// A tile of 2^TLOG u64: 8K on 512 lanes (two workgroups per CU) or 16K on
// 1024 lanes (one workgroup per CU, 140 KiB of LDS); 16 elements per lane
// either way, so the register budget of the Stockham steps is the same.
template <int TLOG>
struct TileCfg {
    static constexpr int SIZE = 1 << TLOG;
    static constexpr int LANES = TLOG >= 14 ? 1024 : NT;
    static constexpr int PER_LANE = SIZE / LANES;
    static constexpr int LDS = SIZE + (SIZE >> PAD_SHIFT) + (SIZE >> 5);
};

// x * 2^s mod p for a compile-time-foldable s in [0, 96); x may be lazily
// reduced (any u64), the result is canonical
__device__ __forceinline__ uint64_t mul_pow2_pos(uint64_t x, int s);

// x * 2^k mod p for a compile-time-foldable k in [0, 192): 2^96 = -1.
// x may be lazily reduced (any u64); the result is canonical.
__device__ __forceinline__ uint64_t mul_pow2(uint64_t x, int k) {
    const uint64_t y = mul_pow2_pos(x, k % 96);
    return k >= 96 ? neg(y) : y;
}

__device__ __forceinline__ uint64_t mul_pow2_pos(uint64_t x, int s) {
    uint64_t y;
    if (s == 0) {
        y = canon(x);
    } else if (s < 32) {
        // x * 2^s = lo + 2^64 * h, h < 2^32:  lo + h * (2^32 - 1)
        const uint64_t lo = x << s;
        const uint32_t h = (uint32_t)(x >> (64 - s));
        const uint64_t t1 = ((uint64_t)(h ? h - 1 : 0) << 32) | (uint32_t)(0u - h);
        uint64_t r = lo + t1;
        if (r < t1) r += EPS;
        y = canon(r);
    } else if (s < 64) {
        y = reduce128(x << s, x >> (64 - s));
    } else {
        // s in [64, 96): 2^s = -2^(s-96), so x * 2^s = -(x / 2^t), t = 96 - s.
        // With x = a 2^t + b: x / 2^t = a - b 2^s and b 2^s = c (2^32 - 1),
        // c = b << (s - 64) < 2^32; hence x * 2^s = c (2^32 - 1) - a.
        const int t = 96 - s;
        const uint64_t a = x >> t;
        const uint32_t c = (uint32_t)(x << (s - 64));
        const uint64_t ce = ((uint64_t)(c ? c - 1 : 0) << 32) | (uint32_t)(0u - c);
        uint64_t r = ce - a;
        if (ce < a) r -= EPS;  // + p (mod 2^64); ce - a + p > 0 since a < 2^63
        y = canon(r);
    }
    return y;
}

// lazily reduced add / sub: a is any u64, b is canonical; results are any u64
// congruent to a +- b (2^64 = 2^32 - 1 mod p). (32-bit carry chains measured
// slower here: every step reads VCC, which serializes them and costs hazard
// wait states.)
__device__ __forceinline__ uint64_t lz_add(uint64_t a, uint64_t b) {
    uint64_t r;
    const bool c = __builtin_add_overflow(a, b, &r);
    return r + (uint64_t)(c ? 0xFFFFFFFFu : 0u);
}
__device__ __forceinline__ uint64_t lz_sub(uint64_t a, uint64_t b) {
    uint64_t r;
    const bool c = __builtin_sub_overflow(a, b, &r);
    return r - (uint64_t)(c ? 0xFFFFFFFFu : 0u);
}

// w_32^e = 2^(78 e mod 192); inverse: 2^(-78 e mod 192)
__device__ __forceinline__ int w32_exp(int e, bool inv) {
    int k = (78 * e) % 192;
    return inv ? (192 - k) % 192 : k;
}

template <int LR>
__device__ __forceinline__ constexpr int brev_c(int x) {
    int r = 0;
    for (int i = 0; i < LR; i++) r |= ((x >> i) & 1) << (LR - 1 - i);
    return r;
}

// in-register DFT of size R = 2^LR <= 32: v[k] <- sum_r v[r] w_R^(+-rk).
// Values between butterflies are lazily reduced u64 (LDS holds them that way
// too); every product and every store to global memory is canonical.
template <int LR, bool INV>
__device__ __forceinline__ void dft_reg(uint64_t *v) {
    constexpr int R = 1 << LR;
#pragma unroll
    for (int i = 0; i < R; i++) {
        const int j = brev_c<LR>(i);
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
                // w_{2 half}^j = w_32^(j * 32 / (2 half)) = +-2^s (2^96 = -1):
                // t = 2^s x canonical (or x reduced once for j = 0), u lazy,
                // and a negative twiddle swaps the sum and the difference
                const int k = j != 0 ? w32_exp(j * (32 >> (lh + 1)), INV) : 0;
                uint64_t t = v[base + j + half];
                t = (j != 0) ? mul_pow2_pos(t, k % 96) : canon(t);
                const uint64_t u = v[base + j];
                if (k >= 96) {
                    v[base + j] = lz_sub(u, t);
                    v[base + j + half] = lz_add(u, t);
                } else {
                    v[base + j] = lz_add(u, t);
                    v[base + j + half] = lz_sub(u, t);
                }
            }
        }
    }
}

// w_{2^lg}^(+-e) = rt_lo[i & 8191] * rt_hi[i >> 13] with i = root_index(lg, e, inv)
// (rt_lo[j] = w_{2^26}^j, rt_hi[j] = w_{2^26}^(8192 j))
__device__ __forceinline__ uint32_t root_index(uint32_t lg, uint64_t e, bool inv) {
    uint64_t e26 = (e & ((1ULL << lg) - 1)) << (26 - lg);
    if (inv) e26 = (0 - e26) & ((1ULL << 26) - 1);
    return (uint32_t)e26;
}
__device__ __forceinline__ uint64_t root_pow(const uint64_t *rt_lo, const uint64_t *rt_hi, uint32_t lg, uint64_t e,
                                             bool inv) {
    const uint32_t e26 = root_index(lg, e, inv);
    return mul(rt_lo[e26 & 8191], rt_hi[e26 >> 13]);
}

// w_{2^pre_log}^e for pre_log <= 13: rt_hi[j] = w_{2^26}^(8192 j) = w_{2^13}^j
__device__ __forceinline__ uint64_t direct_root(const NttPass &p, uint32_t e) {
    return p.rt_hi[(e & ((1u << p.pre_log) - 1)) << (13 - p.pre_log)];
}

// radix of the next Stockham step: ceil(remaining / steps_left), steps of <= MAX_LR bits
constexpr int next_lr(int left) {
    const int steps = (left + MAX_LR - 1) / MAX_LR;
    return (left + steps - 1) / steps;
}

template <int LOG_S, int LOG_NS, bool INV, int TLOG = TILE_LOG>
__device__ __forceinline__ void stockham_steps(uint64_t *lds, uint32_t tile, const uint64_t *tw) {
    if constexpr (LOG_NS < LOG_S) {
        constexpr int LANES = TileCfg<TLOG>::LANES;
        constexpr int LR = next_lr(LOG_S - LOG_NS);
        constexpr int R = 1 << LR;
        constexpr int MAXG = TileCfg<TLOG>::PER_LANE / R;
        constexpr int SUB_LOG = LOG_S - LR;
        constexpr uint32_t SUB = 1u << SUB_LOG;
        constexpr uint32_t NS = 1u << LOG_NS;
        const uint32_t groups = tile >> LR;
        // the lane index through an empty asm, so a caller's loop around the
        // steps cannot hoist their LDS addresses out as live (spilled) values
        uint32_t lane = threadIdx.x;
        asm volatile("" : "+v"(lane));
        uint64_t v[MAXG][R];
#pragma unroll
        for (int u = 0; u < MAXG; u++) {
            const uint32_t q = lane + u * LANES;
            if (q < groups) {
                const uint32_t g = q >> SUB_LOG, j = q & (SUB - 1);
                const uint32_t base = (g << LOG_S) + j;
                // the group's twiddles (global memory, cache hits) are read
                // before its LDS values so the two latencies overlap
                uint64_t tv[R];
                if constexpr (LOG_NS > 0) {
                    const uint32_t jm = j & (NS - 1);
                    const uint64_t *t = tw + (NS << LR);
#pragma unroll
                    for (int r = 1; r < R; r++) tv[r] = t[jm * r];
                }
#pragma unroll
                for (int r = 0; r < R; r++) v[u][r] = lds[padi<LOG_S>(base + r * SUB)];
                if constexpr (LOG_NS > 0) {
#pragma unroll
                    for (int r = 1; r < R; r++) v[u][r] = mul_lazy(v[u][r], tv[r]);  // dft_reg takes any u64
                }
                dft_reg<LR, INV>(v[u]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < MAXG; u++) {
            const uint32_t q = lane + u * LANES;
            if (q < groups) {
                const uint32_t g = q >> SUB_LOG, j = q & (SUB - 1);
                const uint32_t jm = j & (NS - 1);
                const uint32_t base = (g << LOG_S) + ((j >> LOG_NS) << (LOG_NS + LR)) + jm;
#pragma unroll
                for (int r = 0; r < R; r++) lds[padi<LOG_S>(base + r * NS)] = v[u][r];
            }
        }
        __syncthreads();
        stockham_steps<LOG_S, LOG_NS + LR, INV, TLOG>(lds, tile, tw);
    }
}

}  // namespace
}  // namespace sg
