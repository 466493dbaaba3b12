// ntt.hpp -- batched natural-order Goldilocks NTT passes and LDE plans.
#pragma once
#include <stdint.h>

#include "ctx.hpp"

namespace sg {

// One strided pass: for batch b < nb and group g < ng, a DFT of size 2^log_s
// over index i of in[b*ib + g*ig + i*ii], written to out[b*ob + g*og + k*ok].
struct NttPass {
    const uint64_t *in;
    uint64_t *out;
    int64_t ib, ig, ii;
    int64_t ob, og, ok;
    uint64_t ng;
    uint32_t nb;
    // batch b splits into (b / nb2, b % nb2) with strides (ib, ib2) / (ob, ob2)
    uint32_t nb2;
    int64_t ib2, ob2;
    // nr repetitions of the pass over the same input, repetition q using
    // pre-twiddle multiplier r = r0 + q and writing at out + q * out_rs
    uint32_t nr, r0;
    int64_t out_rs;
    // repetitions dealt to 2^rs_log workgroups per tile (set by launch_pass):
    // blockIdx.x = tile * 2^rs_log + chunk, chunk c runs repetitions [c nr, (c+1) nr)
    uint32_t rs_log;
    uint32_t log_s, log_g;
    // log2 of the LDS tile: 13 (512 lanes) or 14 (1024 lanes, log_s >= 8); 0 = 13.
    // launch_pass sets log_g to the most sub-DFTs the tile holds.
    uint32_t tile_log;
    uint32_t inv;
    // input pre-multiply by w_{2^pre_log}^(r * (g*pre_g + i*pre_i + b*pre_b)) (0 = off);
    // pre_direct (set by launch_pass when pre_log <= 13): each factor is one
    // read of the w_{2^13} table instead of a product
    uint32_t pre_log, pre_direct;
    uint64_t pre_g, pre_i, pre_b;
    // output multiply by w_{2^post_log}^(g*(k*post_mul + r*post_rep)), inverse
    // root when inv (0 = off); r = r0 + q is the repetition's multiplier
    uint32_t post_log;
    uint64_t post_mul, post_rep;
    // outputs may be any u64 congruent to the value (a scratch that another
    // pass of these kernels reads) instead of canonical
    uint32_t lazy_out;
    uint64_t scale;
    // output multiply by s^(g*sp_g + k*sp_k) from a two-level power table
    const uint64_t *sp_lo, *sp_hi;
    uint64_t sp_g, sp_k;
    const uint64_t *rt_lo, *rt_hi, *tw;
};

struct PowTable {
    uint64_t *buf = nullptr;
    const uint64_t *lo = nullptr, *hi = nullptr;
};

PowTable make_pow_table(Ctx *c, uint64_t s);
void launch_pass(Ctx *c, NttPass p, const char *name);
void ntt_columns(Ctx *c, const uint64_t *in, uint64_t *out, uint64_t *scratch, uint32_t L, uint64_t ncols,
                 uint64_t in_cs, uint64_t out_cs, bool inv, const PowTable *spow);
// The odd half of a blowup-2 LDE: out[c][j] = T_c(w_{2n} w_n^j), the points
// w_{2n}^(2j+1) of <w_{2n}> (the even points w_{2n}^(2j) = w_n^j carry the
// trace values themselves); cols and out are ncols x n, n = 2^L.
void lde_odd_coset(Ctx *c, const uint64_t *cols, uint64_t *out, uint32_t L, uint64_t ncols);
// twist[i] = w_{2n}^i / n for i < n = 2^L (L <= 12) on the device: the
// odd-coset factor between the inverse and the forward transform; built once
// per context and size
const uint64_t *odd_twist_table(Ctx *c, uint32_t L);
void lde_columns(Ctx *c, const uint64_t *cols, uint64_t *out, uint32_t L, uint32_t LB, uint64_t ncols,
                 uint64_t in_cs, uint64_t out_cs, uint64_t shift);

// ntt_columns over contiguous columns (n = 2^L apart) with the two-pass
// scratch in a fixed per-device array instead of an allocation: calls no
// hipMalloc / hipFree and creates no events. n * ncols must not exceed
// ntt_noalloc_max_elems() (2^22); calls that overlap on one device race.
void ntt_columns_noalloc(Ctx *c, const uint64_t *in, uint64_t *out, uint32_t L, uint64_t ncols, bool inv);
uint64_t ntt_noalloc_max_elems();

// Autotuned launch plans. A plan is the log2 size L1 of the first pass of a
// two-pass transform of 2^L points (the second pass has 2^(L - L1)) and the
// LDS tile its passes use (2^13 elements on 512 lanes or 2^14 on 1024 lanes;
// a 2^14 tile holds twice the sub-DFTs, so its global loads and stores run
// twice as long); for an LDE of 2^L <= 4096 points L1 = L is the one-pass
// form with the cosets as repetitions. Two-pass candidates: L1 within 2 of
// ceil(L/2), both passes of 2^3 to 2^12 points, each with the 2^13 tile and,
// when both passes have at least 2^8 points, the 2^14 tile. The fastest
// plan depends on the shape (sub-DFTs per tile, the length of coalesced
// runs, twiddle layers, workgroups per CU), so it is measured rather than
// modelled: on the first call per (kind, L, LB, log2 columns) with autotune
// on, each candidate runs once untimed and once between HIP events on the
// context stream, and the process-wide table keeps the fastest. Every
// candidate computes the same exact transform.
// A plan is coded as L1 | tile_log << 8.
constexpr uint32_t plan_code(uint32_t l1, uint32_t tile_log) { return l1 | tile_log << 8; }
constexpr uint32_t plan_split(uint32_t code) { return code & 255; }
constexpr uint32_t plan_tile(uint32_t code) { return code >> 8 ? code >> 8 : 13; }
struct PlanSet {
    std::vector<uint32_t> cand;  // admissible plans (empty: nothing to choose)
    uint32_t dflt;               // the untuned plan
};
PlanSet ntt_plans(uint32_t L);
PlanSet lde_plans(uint32_t L, uint32_t LB);
// one "kind L LB log_cols -> L1.tile | L1.tile:ms ..." line per tuned shape
std::string autotune_report();

}  // namespace sg
