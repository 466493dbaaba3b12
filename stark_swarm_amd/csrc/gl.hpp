// gl.hpp -- Goldilocks (p = 2^64 - 2^32 + 1) and its Fp3 / Fp6 towers for
// gfx950 device code and host code.
//
// Values are canonical u64 in [0, p) everywhere in memory. The 64x64 -> 128
// product is built by the compiler from 32-bit v_mad_u64_u32 pieces; the
// reduction uses 2^64 = 2^32 - 1 and 2^96 = -1 (mod p).
//
// Tower (crates/deep_ali/src/cubic_ext.rs:83-244, sextic_ext.rs:60-271):
//   Fp3 = Fp[a]/(a^3 - 7), Fp6 = Fp3[u]/(u^2 - a); an Fp6 value is stored as
//   6 u64 (c0.c0, c0.c1, c0.c2, c1.c0, c1.c1, c1.c2), the reference's
//   to_fp_components order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SG_HD __host__ __device__ __forceinline__

namespace sg {

constexpr uint64_t P = 0xFFFFFFFF00000001ULL;
constexpr uint64_t EPS = 0xFFFFFFFFULL;

SG_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

SG_HD uint64_t reduce128(uint64_t lo, uint64_t hi) {
    const uint64_t hh = hi >> 32;
    const uint32_t hl = (uint32_t)hi;
    uint64_t t0 = lo - hh;
    if (lo < hh) t0 -= EPS;
    // hl * (2^32 - 1) = (hl - 1) * 2^32 + (2^32 - hl) for hl > 0, built from
    // 32-bit pieces so it stays off the quarter-rate multiplier
    const uint32_t t1_hi = hl ? hl - 1 : 0;
    const uint32_t t1_lo = 0u - hl;
    const uint64_t t1 = ((uint64_t)t1_hi << 32) | t1_lo;
    uint64_t r = t0 + t1;
    if (r < t1) r += EPS;
    return r >= P ? r - P : r;
}

// full 64x64 -> 128 product from four 32x32 -> 64 multiply-adds
SG_HD void mul128(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi) {
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    const uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    const uint64_t p00 = (uint64_t)a0 * b0;
    const uint64_t mid = (uint64_t)a0 * b1 + (p00 >> 32);
    const uint64_t mid2 = (uint64_t)a1 * b0 + (uint32_t)mid;
    hi = (uint64_t)a1 * b1 + (mid >> 32) + (mid2 >> 32);
    lo = (mid2 << 32) | (uint32_t)p00;
}

SG_HD uint64_t add(uint64_t a, uint64_t b) {
    uint64_t r = a + b;
    if (r < a || r >= P) r -= P;
    return r;
}
SG_HD uint64_t sub(uint64_t a, uint64_t b) { return a >= b ? a - b : a + (P - b); }
SG_HD uint64_t neg(uint64_t a) { return a ? P - a : 0; }
SG_HD uint64_t mul(uint64_t a, uint64_t b) {
    uint64_t lo, hi;
    mul128(a, b, lo, hi);
    return reduce128(lo, hi);
}
SG_HD uint64_t canon(uint64_t x) { return x >= P ? x - P : x; }
// a * b mod p as any u64 congruent to it (reduce128 without its final
// conditional subtraction), for consumers that accept lazy values
SG_HD uint64_t mul_lazy(uint64_t a, uint64_t b) {
    uint64_t lo, hi;
    mul128(a, b, lo, hi);
    const uint64_t hh = hi >> 32;
    const uint32_t hl = (uint32_t)hi;
    uint64_t t0 = lo - hh;
    if (lo < hh) t0 -= EPS;
    const uint64_t t1 = ((uint64_t)(hl ? hl - 1 : 0) << 32) | (uint32_t)(0u - hl);
    uint64_t r = t0 + t1;
    if (r < t1) r += EPS;
    return r;
}

SG_HD uint64_t pow(uint64_t a, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = mul(r, a);
        a = mul(a, a);
        e >>= 1;
    }
    return r;
}
SG_HD uint64_t inv(uint64_t a) { return pow(a, P - 2); }
// omega_n = 7^((p-1)/n), the generator ark-poly's radix-2 domain uses
inline uint64_t root_of_unity(uint64_t n) { return pow(7, (P - 1) / n); }

// 128-bit lazy sum of field terms, reduced once (a sum of up to 2^46
// terms below 2^82 cannot wrap)
struct Acc {
    uint64_t lo, hi;
};
SG_HD void acc_add(Acc &a, uint64_t x) {
    a.lo += x;
    a.hi += a.lo < x;
}
// a += x * k, k < 2^32
SG_HD void acc_mad(Acc &a, uint64_t x, uint32_t k) {
    const uint64_t p0 = (uint64_t)(uint32_t)x * k, p1 = (x >> 32) * (uint64_t)k;
    const uint64_t lo = p0 + (p1 << 32);
    const uint64_t hi = (p1 >> 32) + (lo < p0);
    a.lo += lo;
    a.hi += hi + (a.lo < lo);
}
// a += x << b, b < 64 (compile-time in unrolled loops)
SG_HD void acc_shl(Acc &a, uint64_t x, int b) {
    const uint64_t lo = x << b, hi = b ? x >> (64 - b) : 0;
    a.lo += lo;
    a.hi += hi + (a.lo < lo);
}
SG_HD uint64_t acc_val(const Acc &a) { return reduce128(a.lo, a.hi); }
SG_HD uint64_t mul_small(uint64_t x, uint32_t k) {
    Acc a{0, 0};
    acc_mad(a, x, k);
    return acc_val(a);
}

struct fp3 { uint64_t c0, c1, c2; };
struct fp6 { uint64_t c[6]; };

SG_HD fp3 f3_add(fp3 a, fp3 b) { return {add(a.c0, b.c0), add(a.c1, b.c1), add(a.c2, b.c2)}; }
SG_HD fp3 f3_sub(fp3 a, fp3 b) { return {sub(a.c0, b.c0), sub(a.c1, b.c1), sub(a.c2, b.c2)}; }
SG_HD uint64_t mul7(uint64_t x) { return mul(x, 7); }
SG_HD fp3 f3_mul(fp3 a, fp3 b) {
    uint64_t t = add(mul(a.c1, b.c2), mul(a.c2, b.c1));
    return {add(mul(a.c0, b.c0), mul7(t)),
            add(add(mul(a.c0, b.c1), mul(a.c1, b.c0)), mul7(mul(a.c2, b.c2))),
            add(add(mul(a.c0, b.c2), mul(a.c1, b.c1)), mul(a.c2, b.c0))};
}
SG_HD fp3 f3_mul_a(fp3 x) { return {mul7(x.c2), x.c0, x.c1}; }
SG_HD fp3 f3_scale(fp3 a, uint64_t s) { return {mul(a.c0, s), mul(a.c1, s), mul(a.c2, s)}; }
SG_HD fp3 f3_inv(fp3 a) {
    uint64_t d0 = sub(mul(a.c0, a.c0), mul7(mul(a.c1, a.c2)));
    uint64_t d1 = sub(mul7(mul(a.c2, a.c2)), mul(a.c0, a.c1));
    uint64_t d2 = sub(mul(a.c1, a.c1), mul(a.c0, a.c2));
    uint64_t n = add(mul(a.c0, d0), mul7(add(mul(a.c1, d2), mul(a.c2, d1))));
    uint64_t ni = inv(n);
    return {mul(d0, ni), mul(d1, ni), mul(d2, ni)};
}

SG_HD fp6 f6_zero() { return {{0, 0, 0, 0, 0, 0}}; }
SG_HD fp6 f6_one() { return {{1, 0, 0, 0, 0, 0}}; }
SG_HD fp6 f6_base(uint64_t x) { return {{x, 0, 0, 0, 0, 0}}; }
SG_HD fp3 f6_lo(const fp6 &a) { return {a.c[0], a.c[1], a.c[2]}; }
SG_HD fp3 f6_hi(const fp6 &a) { return {a.c[3], a.c[4], a.c[5]}; }
SG_HD fp6 f6_make(fp3 l, fp3 h) { return {{l.c0, l.c1, l.c2, h.c0, h.c1, h.c2}}; }
SG_HD fp6 f6_add(const fp6 &a, const fp6 &b) {
    fp6 r;
    for (int i = 0; i < 6; i++) r.c[i] = add(a.c[i], b.c[i]);
    return r;
}
SG_HD fp6 f6_sub(const fp6 &a, const fp6 &b) {
    fp6 r;
    for (int i = 0; i < 6; i++) r.c[i] = sub(a.c[i], b.c[i]);
    return r;
}
SG_HD fp6 f6_scale(const fp6 &a, uint64_t s) {
    fp6 r;
    for (int i = 0; i < 6; i++) r.c[i] = mul(a.c[i], s);
    return r;
}
// Karatsuba over Fp3: 3 Fp3 products
SG_HD fp6 f6_mul(const fp6 &a, const fp6 &b) {
    fp3 a0 = f6_lo(a), a1 = f6_hi(a), b0 = f6_lo(b), b1 = f6_hi(b);
    fp3 v0 = f3_mul(a0, b0), v1 = f3_mul(a1, b1);
    fp3 s = f3_mul(f3_add(a0, a1), f3_add(b0, b1));
    return f6_make(f3_add(v0, f3_mul_a(v1)), f3_sub(f3_sub(s, v0), v1));
}
SG_HD fp6 f6_inv(const fp6 &a) {
    fp3 b0 = f6_lo(a), b1 = f6_hi(a);
    fp3 n = f3_sub(f3_mul(b0, b0), f3_mul_a(f3_mul(b1, b1)));
    fp3 ni = f3_inv(n);
    fp3 r1 = f3_mul(b1, ni);
    return f6_make(f3_mul(b0, ni), {neg(r1.c0), neg(r1.c1), neg(r1.c2)});
}
SG_HD fp6 f6_pow(fp6 a, uint64_t e) {
    fp6 r = f6_one();
    while (e) {
        if (e & 1) r = f6_mul(r, a);
        a = f6_mul(a, a);
        e >>= 1;
    }
    return r;
}
SG_HD bool f6_eq(const fp6 &a, const fp6 &b) {
    for (int i = 0; i < 6; i++)
        if (a.c[i] != b.c[i]) return false;
    return true;
}

}  // namespace sg
