// air.hpp -- DEEP-ALI composition kernels (HashRollup) and helpers.
#pragma once
#include <stdint.h>

#include "ctx.hpp"

namespace sg {

void fill_u64(Ctx *c, uint64_t *p, uint64_t n, uint64_t v);
bool phi_computed(int phi_mode, uint64_t n);
// K proofs' phi (device, n values each, overwritten) -> c_eval = NTT(pad(poly_div_zh(iNTT(phi), n_trace)))
void quotient_batch(Ctx *c, uint64_t *phi, uint32_t K, uint64_t n_trace, uint64_t n, uint64_t *c_eval);
// c_eval (device, n_trace*blowup) from the 4-column HashRollup LDE (device,
// column-major); lam: the 3 combination coefficients (host memory), or null
// for lambda_j = j + 1
void merge_hash_rollup_dev(Ctx *c, const uint64_t *lde, uint64_t n_trace, uint64_t blowup, int phi_mode,
                           uint64_t *c_eval, const uint64_t *lam = nullptr);
// K proofs: lde of proof k at lde + 4*k*n0, c_eval at c_eval + k*n0
void merge_hash_rollup_batch(Ctx *c, const uint64_t *lde, uint32_t K, uint64_t n_trace, uint64_t blowup,
                             int phi_mode, uint64_t *c_eval, const uint64_t *lam = nullptr);
void build_hash_rollup_trace(uint64_t n_trace, const uint64_t *leaves, uint64_t n_leaves, uint64_t *tr);
// AirType::Nsec3Chain: K proofs, lde of proof k at lde + 8*k*n0; lam: 4
// coefficients (host memory) or null
void merge_nsec3_batch(Ctx *c, const uint64_t *lde, uint32_t K, uint64_t n_trace, uint64_t blowup, int phi_mode,
                       uint64_t *c_eval, const uint64_t *lam = nullptr);
void build_nsec3_trace(uint64_t n_trace, const uint64_t *chain, uint64_t n, uint64_t *tr);

}  // namespace sg
