// verify.cpp -- the consumer-side DEEP-FRI / STIR verifier and the Merkle
// opening check, on the host.
//
// deep_fri_verify (crates/deep_ali/src/fri.rs:2136-2710) replays the prover's
// Fiat-Shamir transcript from the proof, checks the out-of-domain fold chain
// (STIR) or the per-layer DEEP identity and fold values (classic FRI), the
// coefficient commitment with its batched degree check, and every Merkle
// opening the proof carries. It does microseconds-to-milliseconds of scalar
// work per proof (a few hundred SHA3 blocks and Fp6 products), which is why it
// stays on the host, as the reference's does; batches are spread over the
// library's host thread pool.
//
// The proof is read in place: a ProofView holds pointers into the caller's
// bytes (the ark-serialize compressed DeepFriProof<SexticExt> of
// fri.rs:1235-1300: usize as u64 LE, Vec as a u64 length and its items,
// Option as a 0/1 byte, field values as canonical 8-byte LE words). Bytes that
// do not deserialize give E_INVALID; a proof that deserializes but would make
// the reference panic (an index past a vector's end, an unwrap of None, a zero
// inverse) is rejected.
#include <string.h>

#include <array>
#include <vector>

#include "ctx.hpp"
#include "gl.hpp"
#include "keccak.hpp"
#include "verify.hpp"

namespace sg {

namespace {

// ---------------------------------------------------------------- reading
struct Reader {
    const uint8_t *p;
    size_t n, pos = 0;
    bool bad = false;
    const uint8_t *take(size_t k) {
        if (bad || n - pos < k) {
            bad = true;
            return nullptr;
        }
        const uint8_t *r = p + pos;
        pos += k;
        return r;
    }
    uint64_t u64() {
        const uint8_t *b = take(8);
        if (!b) return 0;
        uint64_t v;
        memcpy(&v, b, 8);
        return v;
    }
    uint64_t fe() {
        const uint64_t v = u64();
        if (v >= P) bad = true;
        return v;
    }
    fp6 ext() {
        fp6 v;
        for (int i = 0; i < 6; i++) v.c[i] = fe();
        return v;
    }
    bool flag() {
        const uint8_t *b = take(1);
        if (!b) return false;
        if (*b > 1) bad = true;
        return *b == 1;
    }
    // Vec length; every item takes at least `min_item` bytes
    size_t len(size_t min_item) {
        const uint64_t v = u64();
        if (!bad && min_item && v > (n - pos) / min_item) bad = true;
        return bad ? 0 : (size_t)v;
    }
};

struct Opening {
    const uint8_t *leaf = nullptr;
    std::vector<std::pair<size_t, const uint8_t *>> levels;  // (siblings, first sibling digest)
    uint64_t index = 0;
};

void read_opening(Reader &r, Opening &o) {
    o.leaf = r.take(32);
    const size_t nl = r.len(8);
    o.levels.resize(nl);
    for (size_t l = 0; l < nl && !r.bad; l++) {
        const size_t c = r.len(32);
        o.levels[l] = {c, r.take(32 * c)};
    }
    o.index = r.u64();
}

struct Query {  // FriQueryPayload (classic)
    std::vector<std::array<uint64_t, 4>> refs;  // i, child_pos, parent_index, parent_pos
    std::vector<std::array<fp6, 3>> pay;        // f, s, q
    Opening f0;
    uint64_t final_index = 0;
};
struct Prox {  // StirProximityPayload
    uint64_t base = 0, raw = 0;
    std::vector<uint64_t> fidx, fval;
    Opening f0p;
    fp6 fnext;
    bool has_l1 = false;
    Opening l1;
};
struct ProofView {
    const uint8_t *root_f0 = nullptr;
    std::vector<const uint8_t *> roots;
    std::vector<std::vector<Opening>> layer_openings;
    std::vector<Opening> f0_openings;  // deserialized, never read by the verifier
    std::vector<Query> queries;
    std::vector<fp6> fz, fin;
    uint64_t n0 = 0, omega0 = 0;
    bool has_tuples = false, has_croot = false, has_evals = false, has_prox = false;
    std::vector<std::vector<fp6>> tuples, evals;
    const uint8_t *croot = nullptr;
    std::vector<Prox> prox;
};

bool parse(const uint8_t *b, size_t n, ProofView &v) {
    Reader r{b, n};
    v.root_f0 = r.take(32);
    v.roots.resize(r.len(32));
    for (auto &x : v.roots) x = r.take(32);
    v.layer_openings.resize(r.len(8));
    for (auto &lo : v.layer_openings) {
        lo.resize(r.len(48));
        for (auto &o : lo) read_opening(r, o);
        if (r.bad) return false;
    }
    v.f0_openings.resize(r.len(48));
    for (auto &o : v.f0_openings) read_opening(r, o);
    v.queries.resize(r.len(72));
    for (auto &q : v.queries) {
        q.refs.resize(r.len(32));
        for (auto &x : q.refs)
            for (auto &y : x) y = r.u64();
        q.pay.resize(r.len(144));
        for (auto &x : q.pay)
            for (auto &y : x) y = r.ext();
        read_opening(r, q.f0);
        q.final_index = r.u64();
        if (r.bad) return false;
    }
    v.fz.resize(r.len(48));
    for (auto &x : v.fz) x = r.ext();
    v.fin.resize(r.len(48));
    for (auto &x : v.fin) x = r.ext();
    v.n0 = r.u64();
    v.omega0 = r.fe();
    if ((v.has_tuples = r.flag())) {
        v.tuples.resize(r.len(8));
        for (auto &t : v.tuples) {
            t.resize(r.len(48));
            for (auto &x : t) x = r.ext();
        }
    }
    if ((v.has_croot = r.flag())) v.croot = r.take(32);
    if ((v.has_evals = r.flag())) {
        v.evals.resize(r.len(8));
        for (auto &t : v.evals) {
            t.resize(r.len(48));
            for (auto &x : t) x = r.ext();
        }
    }
    if ((v.has_prox = r.flag())) {
        v.prox.resize(r.len(8));
        for (auto &s : v.prox) {
            s.base = r.u64();
            s.raw = r.u64();
            s.fidx.resize(r.len(8));
            for (auto &x : s.fidx) x = r.u64();
            s.fval.resize(r.len(8));
            for (auto &x : s.fval) x = r.fe();
            read_opening(r, s.f0p);
            s.fnext = r.ext();
            if ((s.has_l1 = r.flag())) read_opening(r, s.l1);
            if (r.bad) return false;
        }
    }
    return !r.bad;
}

// ---------------------------------------------------------------- Merkle
void ds_label(uint8_t out[32], uint64_t arity, uint64_t level, uint64_t pos, uint64_t label) {
    const uint64_t w[4] = {arity, level, pos, label};
    memcpy(out, w, 32);  // little-endian host
}

// MerkleTreeChannel::verify_opening with layer_arities = [arity; depth]
// (merkle lib.rs:754-799): siblings beyond arity - 1 on a level are ignored,
// a path longer than depth is rejected (the reference indexes past its
// arity list there)
bool check_opening(uint32_t arity, uint32_t depth, uint64_t label, const uint8_t *root, const Opening &o) {
    uint8_t cur[32];
    memcpy(cur, o.leaf, 32);
    uint64_t idx = o.index;
    for (size_t level = 0; level < o.levels.size(); level++) {
        if (level >= depth) return false;
        const uint64_t pos = idx % arity;
        const size_t nsib = o.levels[level].first;
        const uint8_t *sib = o.levels[level].second;
        Sha3 h;
        uint8_t ds[32];
        ds_label(ds, arity, level + 1, idx / arity, label);
        h.update(ds, 32);
        size_t used = 0;
        for (uint64_t i = 0; i < arity; i++) {
            if (i == pos) {
                h.update(cur, 32);
            } else {
                if (used >= nsib) return false;
                h.update(sib + 32 * used++, 32);
            }
        }
        h.final(cur);
        idx /= arity;
    }
    return memcmp(cur, root, 32) == 0;
}

uint32_t depth_of(uint64_t leaves, uint32_t arity) { return merkle_depth(leaves, arity); }

// ---------------------------------------------------------------- field helpers
// Domain::<F>::new(n).group_gen: the generator of the 2^ceil(log2 n) subgroup
bool group_gen(uint64_t n, uint64_t &g) {
    if (n == 0 || n > (1ULL << 32)) return false;
    g = root_of_unity(1ULL << ilog2(n));
    return true;
}
fp6 dot_pows(const fp6 *c, size_t m, const fp6 &a) {
    fp6 acc = f6_zero(), ap = f6_one();
    for (size_t k = 0; k < m; k++) {
        acc = f6_add(acc, f6_mul(c[k], ap));
        ap = f6_mul(ap, a);
    }
    return acc;
}
fp6 horner(const std::vector<fp6> &c, const fp6 &x) {
    fp6 r = f6_zero();
    for (size_t k = c.size(); k-- > 0;) r = f6_add(f6_mul(r, x), c[k]);
    return r;
}
// monomial coefficients of the degree < m interpolant of v on the points
// s zeta^j (s given by its inverse): (1/m) sum_j v_j zeta^(-jk) s^(-k)
// (fri.rs:468-502 for s = z_l, :714-752 for s = omega^b)
void coset_interp(const fp6 *v, size_t m, uint64_t zeta, const fp6 &sinv, fp6 *out) {
    const uint64_t zi = inv(zeta), minv = inv(m % P);
    fp6 sp = f6_one();
    for (size_t k = 0; k < m; k++) {
        const uint64_t wk = pow(zi, k);
        uint64_t w = 1;
        fp6 acc = f6_zero();
        for (size_t j = 0; j < m; j++) {
            acc = f6_add(acc, f6_scale(v[j], w));
            w = mul(w, wk);
        }
        out[k] = f6_mul(f6_scale(acc, minv), sp);
        sp = f6_mul(sp, sinv);
    }
}
// in-place inverse DFT with 1/n on <omega_n>, natural order (ark-poly ifft)
void host_intt(std::vector<uint64_t> &a) {
    const size_t n = a.size();
    if (n <= 1) return;
    const uint32_t lg = ilog2(n);
    for (size_t i = 0, j = 0; i < n; i++) {
        if (i < j) std::swap(a[i], a[j]);
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j |= bit;
    }
    const uint64_t wi = inv(root_of_unity(n));
    for (uint32_t s = 1; s <= lg; s++) {
        const size_t half = 1ULL << (s - 1);
        const uint64_t ws = pow(wi, n >> s);
        for (size_t b = 0; b < n; b += 2 * half) {
            uint64_t w = 1;
            for (size_t j = 0; j < half; j++) {
                const uint64_t u = a[b + j], t = mul(a[b + j + half], w);
                a[b + j] = add(u, t);
                a[b + j + half] = sub(u, t);
                w = mul(w, ws);
            }
        }
    }
    const uint64_t ni = inv(n % P);
    for (auto &x : a) x = mul(x, ni);
}
// batched_degree_check_ext (fri.rs:811-858)
bool degree_check(const std::vector<std::vector<fp6>> &ct, size_t m, const fp6 &beta, size_t d_final) {
    const size_t nf = ct.size();
    std::vector<uint64_t> comp(nf);
    std::vector<fp6> g(nf);
    for (size_t b = 0; b < nf; b++) g[b] = dot_pows(ct[b].data(), m, beta);
    for (int j = 0; j < 6; j++) {
        for (size_t b = 0; b < nf; b++) comp[b] = g[b].c[j];
        host_intt(comp);
        for (size_t k = d_final; k < nf; k++)
            if (comp[k]) return false;
    }
    return true;
}
// raw query index q: StdRng(index_seed(seed, 0, q)).next_u64() masked to the
// next power of two, mod n0 (fri.rs:1044-1058)
uint64_t query_index(uint64_t seed, uint64_t q, uint64_t n0) {
    const uint64_t f[3] = {seed, 0, canon(q)};
    const uint64_t s = hash_fields_tagged("FRI/index", f, 3);
    return (stdrng_first_u64(s) & ((1ULL << ilog2(n0)) - 1)) % n0;
}

// ---------------------------------------------------------------- deep_fri_verify
bool verify_view(const FriParams &p, const ProofView &v) {
    const size_t L = p.schedule.size();
    const auto &sch = p.schedule;
    std::vector<uint64_t> sz(L + 1);
    sz[0] = v.n0;
    for (size_t l = 0; l < L; l++) {
        if (sch[l] == 0 || sz[l] % sch[l]) return false;  // layer_sizes_from_schedule asserts
        sz[l + 1] = sz[l] / sch[l];
    }
    const bool use_cc = p.coeff_commit_final && L > 0, stir = p.stir;
    const size_t normal = use_cc ? L - 1 : L;

    Transcript tr("FRI/FS");
    if (p.has_pi) {
        tr.absorb("CAIRO-PUBLIC-INPUTS-V1");
        tr.absorb(p.pi, 32);
    }
    tr.absorb("DEEP-FRI-STATEMENT-V2");
    tr.absorb_field(canon(v.n0));
    tr.absorb_field(L);
    for (uint64_t m : sch) tr.absorb_field(canon(m));
    tr.absorb_field(canon(p.seed_z));
    tr.absorb_field(6);
    tr.absorb_field(p.coeff_commit_final ? 1 : 0);
    tr.absorb_field(stir ? 1 : 0);
    tr.absorb(v.root_f0, 32);
    fp6 z_ext;
    tr.challenge_ext("z_fp3", z_ext.c);

    std::vector<fp6> alpha(L);
    std::vector<std::vector<fp6>> interp(L);
    fp6 zc = z_ext, beta = f6_zero();
    for (size_t ell = 0; ell < L; ell++) {
        if (ell < normal) tr.challenge_ext("alpha", alpha[ell].c);
        if (stir) {
            const uint64_t m = sch[ell];
            if (!v.has_evals || ell >= v.evals.size() || v.evals[ell].size() != m) return false;
            uint64_t om;
            if (!group_gen(sz[ell], om) || f6_eq(zc, f6_zero())) return false;
            interp[ell].resize(m);
            coset_interp(v.evals[ell].data(), m, pow(om, sz[ell] / m), f6_inv(zc), interp[ell].data());
            zc = f6_pow(zc, m);
            for (const fp6 &e : v.evals[ell])
                for (int q = 0; q < 6; q++) tr.absorb_field(e.c[q]);
        } else {
            if (ell >= v.fz.size()) return false;
            for (int q = 0; q < 6; q++) tr.absorb_field(v.fz[ell].c[q]);
        }
        if (ell >= v.roots.size()) return false;
        tr.absorb(v.roots[ell], 32);
        if (use_cc && ell == L - 1) {
            if (!v.has_croot) return false;
            tr.absorb(v.croot, 32);
            tr.challenge_ext("alpha", alpha[ell].c);
            tr.challenge_ext("beta_deg", beta.c);
        }
    }
    for (const fp6 &c : v.fin)
        for (int q = 0; q < 6; q++) tr.absorb_field(c.c[q]);
    const uint64_t qseed = tr.challenge_field("query_seed");
    if (v.fin.size() != p.d_final) return false;

    if (stir) {
        for (size_t ell = 0; ell < L; ell++) {
            const fp6 fold = dot_pows(interp[ell].data(), sch[ell], alpha[ell]);
            fp6 expect;
            if (ell + 1 < L) {
                expect = v.evals[ell + 1][0];
            } else {
                fp6 zf = z_ext;
                for (uint64_t m : sch) zf = f6_pow(zf, m);
                expect = horner(v.fin, zf);
            }
            if (!f6_eq(fold, expect)) return false;
        }
    }
    if (use_cc) {
        const uint64_t nf = sz[L], mf = sch[L - 1];
        if (!v.has_tuples || v.tuples.size() != nf) return false;
        std::vector<uint64_t> vals;
        vals.reserve(6 * mf * nf);
        for (const auto &t : v.tuples) {
            if (t.size() != mf) return false;
            for (const fp6 &x : t) vals.insert(vals.end(), x.c, x.c + 6);
        }
        const uint32_t a = pick_arity(nf, 16);
        uint8_t root[32];
        host_merkle_root(vals.data(), nf, 6 * mf, a, depth_of(nf, a), 0xFE, root);
        if (memcmp(root, v.croot, 32) != 0) return false;
        if (!degree_check(v.tuples, mf, beta, p.d_final)) return false;
    }
    std::vector<uint64_t> om(L);
    for (size_t l = 0; l < L; l++)
        if (!group_gen(sz[l], om[l])) return false;
    uint64_t om_final = 1;
    if (sz[L] >= 2 && !group_gen(sz[L], om_final)) return false;

    if (stir) {
        if (!v.has_prox || v.prox.size() != p.s0) return false;
        const uint64_t n0 = v.n0, m0 = sch[0], nn = sz[0] / m0;
        const uint64_t zeta0 = pow(om[0], nn), om0_inv = inv(om[0]);
        const uint32_t fa = pick_arity(nn, 16), fd = depth_of(nn, fa);
        uint32_t a1 = 0, d1 = 0;
        if (L >= 2) {
            a1 = pick_arity(sz[1], sch[1]);
            d1 = depth_of(sz[1], a1);
        }
        std::vector<fp6> fib(m0), co(m0);
        for (uint64_t qi = 0; qi < p.s0; qi++) {
            const Prox &s = v.prox[qi];
            const uint64_t raw = query_index(qseed, qi, n0);
            if (s.raw != raw || s.base != raw % nn || s.fidx.size() != m0 || s.fval.size() != m0) return false;
            for (uint64_t j = 0; j < m0; j++)
                if (s.fidx[j] != s.base + j * nn) return false;
            if (!check_opening(fa, fd, 0xFF, v.root_f0, s.f0p) || s.f0p.index != s.base) return false;
            uint8_t lh[32];
            host_leaf_hash(fa, 0xFF, s.f0p.index, s.fval.data(), m0, lh);
            if (memcmp(lh, s.f0p.leaf, 32) != 0) return false;
            for (uint64_t j = 0; j < m0; j++) fib[j] = f6_base(s.fval[j]);
            coset_interp(fib.data(), m0, zeta0, f6_base(pow(om0_inv, s.base)), co.data());
            const fp6 fold = dot_pows(co.data(), m0, alpha[0]);
            if (L >= 2) {
                if (!s.has_l1 || !check_opening(a1, d1, 1, v.roots[1], s.l1) || s.l1.index != s.base) return false;
                host_leaf_hash(a1, 1, s.l1.index, s.fnext.c, 6, lh);
                if (memcmp(lh, s.l1.leaf, 32) != 0 || !f6_eq(fold, s.fnext)) return false;
            } else {
                if (!f6_eq(fold, horner(v.fin, f6_base(pow(om_final, s.base))))) return false;
            }
        }
        return true;
    }

    // classic FRI: r queries followed through every layer
    const uint64_t n0 = v.n0;
    const uint32_t fa = pick_arity(n0, 16), fd = depth_of(n0, fa);
    if (v.queries.size() < p.r) return false;
    for (uint64_t q = 0; q < p.r; q++) {
        const Query &qp = v.queries[q];
        const uint64_t i0 = query_index(qseed, q, n0);
        if (qp.refs.size() < L || qp.pay.size() < L) return false;
        uint64_t ei = i0;
        for (size_t ell = 0; ell < L; ell++) {
            if (qp.refs[ell][0] != ei) return false;
            ei %= sz[ell] / sch[ell];
        }
        if (qp.final_index != ei) return false;
        if (!check_opening(fa, fd, 0xFF, v.root_f0, qp.f0)) return false;
        const fp6 &f0v = qp.pay[0][0];
        for (int c = 1; c < 6; c++)
            if (f0v.c[c]) return false;
        uint8_t lh[32];
        host_leaf_hash(fa, 0xFF, qp.f0.index, f0v.c, 1, lh);
        if (memcmp(lh, qp.f0.leaf, 32) != 0 || qp.f0.index != i0) return false;
        for (size_t ell = 0; ell < L; ell++) {
            if (ell >= v.layer_openings.size() || q >= v.layer_openings[ell].size()) return false;
            const Opening &o = v.layer_openings[ell][q];
            const uint64_t i = qp.refs[ell][0];
            const auto &pay = qp.pay[ell];
            const uint32_t a = pick_arity(sz[ell], sch[ell]);
            if (!check_opening(a, depth_of(sz[ell], a), ell, v.roots[ell], o) || o.index != i) return false;
            uint64_t lv[18];
            for (int k = 0; k < 3; k++) memcpy(lv + 6 * k, pay[k].c, 48);
            host_leaf_hash(a, ell, o.index, lv, 18, lh);
            if (memcmp(lh, o.leaf, 32) != 0) return false;
            const uint64_t xb = pow(om[ell], i);
            // DEEP identity q (x - z) = f - f(z)
            if (!f6_eq(f6_mul(pay[2], f6_sub(f6_base(xb), z_ext)), f6_sub(pay[0], v.fz[ell]))) return false;
            const fp6 xf = f6_base(pow(om_final, qp.final_index));
            if (use_cc && ell == L - 1) {
                const uint64_t m = sch[ell];
                if (qp.final_index >= v.tuples.size()) return false;
                const std::vector<fp6> &t = v.tuples[qp.final_index];
                fp6 h = f6_zero();
                uint64_t xp = 1;
                for (uint64_t k = 0; k < m; k++) {
                    h = f6_add(h, f6_scale(t[k], xp));
                    xp = mul(xp, xb);
                }
                if (!f6_eq(pay[0], h)) return false;
                if (!f6_eq(dot_pows(t.data(), m, alpha[L - 1]), horner(v.fin, xf))) return false;
            } else {
                const fp6 nxt = ell + 1 < L ? qp.pay[ell + 1][0] : horner(v.fin, xf);
                if (!f6_eq(pay[1], nxt)) return false;
            }
        }
    }
    return true;
}

}  // namespace

void host_leaf_hash(uint32_t arity, uint64_t label, uint64_t pos, const uint64_t *vals, size_t n, uint8_t out[32]) {
    Sha3 h;
    uint8_t ds[32];
    ds_label(ds, arity, 0xFFFFFFFFULL, pos, label);
    h.update(ds, 32);
    h.update(vals, 8 * n);  // LE64 values on a little-endian host
    h.final(out);
}

void host_merkle_root(const uint64_t *vals, uint64_t n, size_t vpl, uint32_t arity, uint32_t depth, uint64_t label,
                      uint8_t root[32]) {
    std::vector<std::array<uint8_t, 32>> lv(n);
    for (uint64_t i = 0; i < n; i++) host_leaf_hash(arity, label, i, vals + i * vpl, vpl, lv[i].data());
    uint32_t level = 0;
    while (lv.size() > 1) {
        SG_REQUIRE(level < depth, "merkle: more levels than the configured depth");
        while (lv.size() % arity) lv.push_back(lv.back());
        std::vector<std::array<uint8_t, 32>> up(lv.size() / arity);
        for (size_t i = 0; i < up.size(); i++) {
            Sha3 h;
            uint8_t ds[32];
            ds_label(ds, arity, level + 1, i, label);
            h.update(ds, 32);
            h.update(lv[i * arity].data(), 32 * arity);
            h.final(up[i].data());
        }
        lv.swap(up);
        level++;
    }
    memcpy(root, lv[0].data(), 32);
}

int merkle_verify_opening(uint32_t arity, uint32_t depth, uint64_t label, const uint8_t root[32],
                          const uint8_t *opening, size_t len) {
    Reader r{opening, len};
    Opening o;
    read_opening(r, o);
    if (r.bad) return E_INVALID;
    return check_opening(arity, depth, label, root, o) ? 1 : 0;
}

int deep_fri_verify(const FriParams &p, const uint8_t *proof, size_t len) {
    ProofView v;
    if (!parse(proof, len, v)) return E_INVALID;
    return verify_view(p, v) ? 1 : 0;
}

void deep_fri_verify_batch(const FriParams &p, const uint8_t *const *proofs, const size_t *lens, size_t n,
                           const uint8_t *pis, int *results) {
    parallel_for(n, [&](size_t i) {
        FriParams q = p;
        if (pis) {
            q.has_pi = true;
            memcpy(q.pi, pis + 32 * i, 32);
        }
        results[i] = deep_fri_verify(q, proofs[i], lens[i]);
    });
}

}  // namespace sg
