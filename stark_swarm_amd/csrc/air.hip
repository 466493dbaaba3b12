// air.hip -- DEEP-ALI composition on gfx950 for the HashRollup AIR and the
// Z_H quotient.
//
// deep_ali_merge_general (crates/deep_ali/src/lib.rs:343-437) with the
// HashRollup evaluator (crates/deep_ali/src/air_workloads.rs:591-599):
//   phi(i) = sum_j (j+1) C_j(row i, row (i+blowup) mod n),  lambda_j = j+1
//            (crates/swarm-dns/src/prover.rs:100-102)
//   C0 = idx' - idx - 1,  C1 = sq - st^2,  C2 = st' - sq - leaf
// then c = NTT(pad(poly_div_zh(iNTT(phi), n_trace))) (lib.rs:190-228: the
// quotient q[j] = sum_{t>=1, j+t*m<n} a[j+t*m], remainder dropped).
// phi_mode reproduces the reference's thread-count quirk (lib.rs:20-33,
// 374-397): 1 = rayon with >1 thread (phi == 0 when n < 4096), 2 = one
// thread (phi always 0), 0 = the non-parallel build (phi always computed).
#include "air.hpp"
#include "gl.hpp"
#include "ntt.hpp"

namespace sg {

namespace {

constexpr int AT = 256;

// grid.y = proof: lde of proof k at lde + 4 k n, phi at phi + k n
// combination coefficients of an AIR with at most 4 constraints, passed by
// value; on = 0: lambda_j = j + 1
struct Lam4 {
    uint64_t v[4];
    uint32_t on;
};
Lam4 lam4(const uint64_t *lam, uint32_t k) {
    Lam4 l{};
    if (lam) {
        l.on = 1;
        for (uint32_t j = 0; j < k && j < 4; j++) l.v[j] = canon(lam[j]);
    }
    return l;
}

__global__ __launch_bounds__(AT) void compose_hash_rollup_k(const uint64_t *lde, uint64_t n, uint64_t blowup,
                                                            uint64_t *phi, Lam4 lam) {
    const uint64_t i = (uint64_t)blockIdx.x * AT + threadIdx.x;
    if (i >= n) return;
    lde += 4 * n * blockIdx.y;
    phi += n * blockIdx.y;
    const uint64_t j = (i + blowup) & (n - 1);
    const uint64_t *idx = lde, *leaf = lde + n, *st = lde + 2 * n, *sq = lde + 3 * n;
    const uint64_t s = st[i], q = sq[i];
    const uint64_t c0 = sub(sub(idx[j], idx[i]), 1);
    const uint64_t c1 = sub(q, mul(s, s));
    const uint64_t c2 = sub(sub(st[j], q), leaf[i]);
    if (lam.on)
        phi[i] = add(add(mul(lam.v[0], c0), mul(lam.v[1], c1)), mul(lam.v[2], c2));
    else  // 1*c0 + 2*c1 + 3*c2
        phi[i] = add(add(c0, add(c1, c1)), add(add(c2, c2), c2));
}

// AirType::Nsec3Chain (air_workloads.rs:789-870): 8 columns (owner limbs,
// next limbs), C_k = nxt[k] - cur[4 + k], phi = sum (k + 1) C_k
__global__ __launch_bounds__(AT) void compose_nsec3_k(const uint64_t *lde, uint64_t n, uint64_t blowup,
                                                     uint64_t *phi, Lam4 lam) {
    const uint64_t i = (uint64_t)blockIdx.x * AT + threadIdx.x;
    if (i >= n) return;
    lde += 8 * n * blockIdx.y;
    phi += n * blockIdx.y;
    const uint64_t j = (i + blowup) & (n - 1);
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint64_t d = sub(lde[k * n + j], lde[(4 + k) * n + i]);
        if (lam.on)
            acc = add(acc, mul(lam.v[k], d));
        else
            for (int t = 0; t <= k; t++) acc = add(acc, d);
    }
    phi[i] = acc;
}

// one residue class rho < m per lane: suffix sums along stride m
__global__ __launch_bounds__(AT) void zh_div_k(const uint64_t *a, uint64_t n, uint64_t m, uint64_t *q) {
    const uint64_t rho = (uint64_t)blockIdx.x * AT + threadIdx.x;
    if (rho >= m) return;
    a += n * blockIdx.y;
    q += n * blockIdx.y;
    const uint64_t B = n / m;
    uint64_t acc = 0;
    q[rho + (B - 1) * m] = 0;  // q has length n - m; the top block is zero padding
    for (uint64_t t = B - 1; t >= 1; t--) {
        acc = add(acc, a[rho + t * m]);
        q[rho + (t - 1) * m] = acc;
    }
}

__global__ void fill_k(uint64_t *p, uint64_t n, uint64_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

void fill_u64(Ctx *c, uint64_t *p, uint64_t n, uint64_t v) {
    if (!n) return;
    hipLaunchKernelGGL(fill_k, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, c->stream, p, n, v);
    SG_HIP(hipGetLastError());
}

bool phi_computed(int phi_mode, uint64_t n) { return phi_mode == 0 || (phi_mode == 1 && n >= 4096); }

void merge_hash_rollup_batch(Ctx *c, const uint64_t *lde, uint32_t K, uint64_t n_trace, uint64_t blowup,
                             int phi_mode, uint64_t *c_eval, const uint64_t *lam) {
    SG_REQUIRE(is_pow2(n_trace) && is_pow2(blowup) && blowup >= 2, "merge: sizes must be powers of two");
    SG_REQUIRE(K >= 1 && K <= 65535, "merge: batch size out of range");
    const uint64_t n = n_trace * blowup;
    SG_REQUIRE(n <= (1ULL << 24), "merge: domain above 2^24");
    if (!phi_computed(phi_mode, n)) {
        // phi == 0: iNTT, division and NTT of zero are zero
        fill_u64(c, c_eval, n * K, 0);
        return;
    }
    Scoped phi(c, 8 * n * K);
    {
        ProfScope ps(c, "compose_hash_rollup", (8.0 * 4 * n + 8.0 * n) * K);
        hipLaunchKernelGGL(compose_hash_rollup_k, dim3((uint32_t)((n + AT - 1) / AT), K), dim3(AT), 0, c->stream, lde,
                           n, blowup, phi.as<uint64_t>(), lam4(lam, 3));
        SG_HIP(hipGetLastError());
    }
    quotient_batch(c, phi.as<uint64_t>(), K, n_trace, n, c_eval);
}

void merge_nsec3_batch(Ctx *c, const uint64_t *lde, uint32_t K, uint64_t n_trace, uint64_t blowup, int phi_mode,
                       uint64_t *c_eval, const uint64_t *lam) {
    SG_REQUIRE(is_pow2(n_trace) && is_pow2(blowup) && blowup >= 2, "merge: sizes must be powers of two");
    SG_REQUIRE(K >= 1 && K <= 65535, "merge: batch size out of range");
    const uint64_t n = n_trace * blowup;
    SG_REQUIRE(n <= (1ULL << 24), "merge: domain above 2^24");
    if (!phi_computed(phi_mode, n)) {
        fill_u64(c, c_eval, n * K, 0);
        return;
    }
    Scoped phi(c, 8 * n * K);
    {
        ProfScope ps(c, "compose_nsec3", (8.0 * 8 * n + 8.0 * n) * K);
        hipLaunchKernelGGL(compose_nsec3_k, dim3((uint32_t)((n + AT - 1) / AT), K), dim3(AT), 0, c->stream, lde, n,
                           blowup, phi.as<uint64_t>(), lam4(lam, 4));
        SG_HIP(hipGetLastError());
    }
    quotient_batch(c, phi.as<uint64_t>(), K, n_trace, n, c_eval);
}

// build_nsec3_chain_trace (air_workloads.rs:811-860): chain = n records of
// 8 raw LE u64 limbs (owner, next), reduced into the field; rows past the
// chain repeat the last next limbs as owner and next, and the last row's
// next becomes row 0's owner.
void build_nsec3_trace(uint64_t n_trace, const uint64_t *chain, uint64_t n, uint64_t *tr) {
    const uint64_t m = n < n_trace ? n : n_trace;
    for (uint64_t i = 0; i < m; i++)
        for (int k = 0; k < 8; k++) tr[k * n_trace + i] = canon(chain[8 * i + k]);
    if (n < n_trace) {
        for (uint64_t i = n; i < n_trace; i++)
            for (int k = 0; k < 4; k++)
                tr[k * n_trace + i] = tr[(4 + k) * n_trace + i] = n ? canon(chain[8 * (n - 1) + 4 + k]) : 0;
        for (int k = 0; k < 4; k++) tr[(4 + k) * n_trace + n_trace - 1] = tr[k * n_trace];
    }
}

void quotient_batch(Ctx *c, uint64_t *phi, uint32_t K, uint64_t n_trace, uint64_t n, uint64_t *c_eval) {
    Scoped coef(c, 8 * n * K), scr(c, 8 * n * K);
    const uint32_t L = ilog2(n);
    ntt_columns(c, phi, coef.as<uint64_t>(), scr.as<uint64_t>(), L, K, n, n, true, nullptr);
    {
        ProfScope ps(c, "zh_div", 16.0 * n * K);
        hipLaunchKernelGGL(zh_div_k, dim3((uint32_t)((n_trace + AT - 1) / AT), K), dim3(AT), 0, c->stream,
                           coef.as<uint64_t>(), n, n_trace, phi);
        SG_HIP(hipGetLastError());
    }
    ntt_columns(c, phi, c_eval, scr.as<uint64_t>(), L, K, n, n, false, nullptr);
}

void merge_hash_rollup_dev(Ctx *c, const uint64_t *lde, uint64_t n_trace, uint64_t blowup, int phi_mode,
                           uint64_t *c_eval, const uint64_t *lam) {
    merge_hash_rollup_batch(c, lde, 1, n_trace, blowup, phi_mode, c_eval, lam);
}

// HashRollup trace (air_workloads.rs:571-589): idx, leaf, state, state^2 with
// state' = state^2 + leaf -- a sequential recurrence, built on the host.
void build_hash_rollup_trace(uint64_t n_trace, const uint64_t *leaves, uint64_t n_leaves, uint64_t *tr) {
    uint64_t s = 0;
    for (uint64_t i = 0; i < n_trace; i++) {
        uint64_t leaf = i < n_leaves ? canon(leaves[i]) : 0;
        tr[i] = canon(i);
        tr[n_trace + i] = leaf;
        tr[2 * n_trace + i] = s;
        tr[3 * n_trace + i] = mul(s, s);
        s = add(tr[3 * n_trace + i], leaf);
    }
}

}  // namespace sg
