// server.cpp -- request queue, HBM-budget admission and continuous batching
// over worker threads with one HIP stream each (see server.hpp).
#include "server.hpp"

#include <string.h>

#include <algorithm>

#include "sha256_air.hpp"

namespace sg {

static const uint64_t BLOWUP = 32;

uint64_t Server::need_bytes(const ServeRequest &r) {
    if (r.kind == SERVE_ED25519) {  // trace and odd coset, gate-group sums, phi, FRI
        const EdShape sh = ed25519_shape(r.ed_msg.size(), 256);
        const uint64_t h = sh.height, n0 = h * BLOWUP;
        return 16ULL * sh.width * h + 8ULL * 400 * 2 * h + 8ULL * 64 * n0;
    }
    if (r.kind == SERVE_RSA) {
        const uint64_t h = 32, n0 = h * BLOWUP;
        return 8ULL * RSA_W * 4 * h + 8ULL * 3 * 2 * h + 8ULL * 24 * n0;
    }
    if (r.kind == SERVE_P256) {  // trace, its blowup-2 LDE and partial sums, the tail LDE, FRI
        const uint64_t h = P256_ECDSA_ROWS, n0 = h * BLOWUP;
        return 8ULL * P256_W * 4 * h + 8ULL * 300 * 2 * h + 8ULL * 200 * n0 + 8ULL * 16 * n0;
    }
    if (r.kind == SERVE_DS_KSK) {
        const uint64_t h = sha256_trace_height(r.key.size()), n0 = h * BLOWUP;
        return 8ULL * SHA256_W * 4 * h + 8ULL * 5 * 2 * h + 8ULL * 24 * n0;
    }
    uint64_t nt = 8;
    while (nt < 4 * r.domains.size()) nt <<= 1;
    const uint64_t n0 = nt * BLOWUP;
    uint64_t rec = 0;
    for (size_t i = 0; i < r.domains.size(); i++) rec += 64 + r.domains[i].size() + r.rdata[i].size();
    return 8ULL * 4 * (2 * nt + 2 * n0) + 8ULL * 16 * n0 + 2 * rec;
}

Server::Server(int device, int workers, uint64_t hbm_budget, uint32_t max_batch, bool dry_run, bool validate)
    : device_(device), budget_(hbm_budget), max_batch_(std::max<uint32_t>(1, max_batch)), dry_(dry_run),
      validate_(validate), t0_(std::chrono::steady_clock::now()) {
    SG_REQUIRE(workers >= 1 && workers <= 16, "server: 1 to 16 workers");
    SG_REQUIRE(hbm_budget > 0, "server: zero HBM budget");
    if (!dry_) {
        try {
            for (int w = 0; w < workers; w++) ctxs_.push_back(ctx_create(device));
        } catch (...) {
            for (Ctx *c : ctxs_) ctx_destroy(c);
            throw;
        }
    } else {
        ctxs_.assign(workers, nullptr);
    }
    for (int w = 0; w < workers; w++) threads_.emplace_back([this, w] { worker(w); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    for (auto &t : threads_) t.join();
    for (Ctx *c : ctxs_)
        if (c) ctx_destroy(c);
}

double Server::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

uint64_t Server::submit(ServeRequest &&r) {
    SG_REQUIRE(r.kind == SERVE_DS_KSK || r.kind == SERVE_INNER_SHARD || r.kind == SERVE_RSA || r.kind == SERVE_P256 ||
                   r.kind == SERVE_ED25519,
               "server: unknown request kind");
    SG_REQUIRE(r.deadline_s >= 0, "server: negative deadline");
    if (validate_) {  // the checks the provers make, before the request can join a batch
        if (r.kind == SERVE_RSA)
            SG_REQUIRE(rsa_record_verifies(r.rsa), "server: RSA record needs 0 < n, s < n and s^65537 mod n = em");
        if (r.kind == SERVE_P256) (void)p256_ecdsa_chains(r.p256);  // throws E_INVALID for a bad key or signature
        if (r.kind == SERVE_ED25519)
            SG_REQUIRE(ed25519_verify(EdSig{r.ed_pk.data(), r.ed_sig.data(), r.ed_msg.data(), r.ed_msg.size()}),
                       "server: Ed25519 signature does not verify");
    }
    if (r.kind == SERVE_INNER_SHARD) {
        SG_REQUIRE(!r.domains.empty(), "server: an inner shard needs at least one record");
        SG_REQUIRE(r.rdata.size() == r.domains.size() && r.rtype.size() == r.domains.size() &&
                       r.ttl.size() == r.domains.size(),
                   "server: record fields of unequal length");
    }
    Item it;
    it.need = need_bytes(r);
    it.req = std::move(r);
    uint64_t id;
    {
        std::lock_guard<std::mutex> g(mu_);
        SG_REQUIRE(!stop_, "server: stopped");
        id = it.id = next_id_++;
        it.t_submit = now();
        queue_.push_back(std::move(it));
    }
    cv_work_.notify_one();
    return id;
}

bool Server::wait(uint64_t id, ServeResult &out, double timeout_s) {
    std::unique_lock<std::mutex> lk(mu_);
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
    while (true) {
        auto f = done_.find(id);
        if (f != done_.end()) {
            out = std::move(f->second);
            done_.erase(f);
            return true;
        }
        if (cv_done_.wait_until(lk, deadline) == std::cv_status::timeout && done_.find(id) == done_.end())
            return false;
    }
}

size_t Server::completed() const {
    std::lock_guard<std::mutex> g(mu_);
    return completed_;
}

uint64_t Server::peak_in_use() const {
    std::lock_guard<std::mutex> g(mu_);
    return peak_;
}

static bool same_group(const ServeRequest &a, const ServeRequest &b) {
    return a.kind == b.kind && (a.kind != SERVE_INNER_SHARD || a.salt == b.salt);
}

void Server::worker(int w) {
    Ctx *c = ctxs_[w];
    if (c) (void)hipSetDevice(device_);  // worker threads start on device 0
    while (true) {
        std::vector<Item> batch;
        uint64_t taken = 0;
        double t_start;
        uint64_t seq;
        {
            std::unique_lock<std::mutex> lk(mu_);
            // the head request goes first (no overtaking); it is admitted when it
            // fits beside the batches in flight, or alone when nothing runs
            cv_work_.wait(lk, [&] {
                return (stop_ && queue_.empty()) ||
                       (!queue_.empty() && (in_use_ + queue_.front().need <= budget_ || in_use_ == 0));
            });
            if (queue_.empty()) return;  // stop requested and nothing left
            // requests whose deadline passed before admission complete without device work
            const double tn = now();
            std::vector<Item> expired;
            for (auto q = queue_.begin(); q != queue_.end();) {
                if (q->req.deadline_s > 0 && tn - q->t_submit > q->req.deadline_s) {
                    expired.push_back(std::move(*q));
                    q = queue_.erase(q);
                } else {
                    ++q;
                }
            }
            for (Item &x : expired) {
                ServeResult r;
                r.id = x.id;
                r.kind = x.req.kind;
                r.status = E_TIMEOUT;
                r.error = "server: deadline passed before admission";
                r.t_submit = x.t_submit;
                r.t_start = r.t_done = tn;
                done_[r.id] = std::move(r);
                completed_++;
            }
            if (!expired.empty()) cv_done_.notify_all();
            if (queue_.empty()) continue;
            if (!(in_use_ + queue_.front().need <= budget_ || in_use_ == 0)) continue;
            batch.push_back(std::move(queue_.front()));
            queue_.pop_front();
            taken = batch[0].need;
            for (auto q = queue_.begin(); q != queue_.end() && batch.size() < max_batch_;) {
                if (same_group(q->req, batch[0].req) && in_use_ + taken + q->need <= budget_) {
                    taken += q->need;
                    batch.push_back(std::move(*q));
                    q = queue_.erase(q);
                } else {
                    ++q;
                }
            }
            in_use_ += taken;
            peak_ = std::max(peak_, in_use_);
            t_start = now();  // admission time, taken in admission order
            seq = ++batch_seq_;
        }
        cv_work_.notify_one();  // another worker may admit what remains
        std::vector<ServeResult> res(batch.size());
        run_batch(c, batch, res);
        const double t_done = now();
        {
            std::lock_guard<std::mutex> g(mu_);
            in_use_ -= taken;
            for (size_t i = 0; i < batch.size(); i++) {
                ServeResult &r = res[i];
                r.id = batch[i].id;
                r.kind = batch[i].req.kind;
                r.t_submit = batch[i].t_submit;
                r.t_start = t_start;
                r.t_done = t_done;
                r.batch_size = batch.size();
                r.batch_seq = seq;
                done_[r.id] = std::move(r);
            }
            completed_ += batch.size();
        }
        cv_done_.notify_all();
        cv_work_.notify_all();
    }
}

void Server::run_batch(Ctx *c, std::vector<Item> &batch, std::vector<ServeResult> &res) {
    if (dry_) {
        uint64_t need = 0;
        for (auto &b : batch) need += b.need;
        std::this_thread::sleep_for(std::chrono::microseconds(200 + need / (64ULL << 20)));
        return;
    }
    int code = 0;
    std::string msg;
    try {
        prove_batch(c, batch, res);
        return;
    } catch (const Error &e) {
        code = e.code;
        msg = e.what();
    } catch (const std::exception &e) {
        code = E_INTERNAL;
        msg = e.what();
    }
    if (batch.size() == 1) {
        res[0] = ServeResult();
        res[0].status = code;
        res[0].error = msg;
        return;
    }
    // one request of the batch failed it: prove each alone, so that only the
    // failing request carries its error and the others get their proofs
    for (size_t i = 0; i < batch.size(); i++) {
        std::vector<Item> one;
        one.push_back(std::move(batch[i]));
        std::vector<ServeResult> r1(1);
        run_batch(c, one, r1);
        batch[i] = std::move(one[0]);
        res[i] = std::move(r1[0]);
    }
}

void Server::prove_batch(Ctx *c, std::vector<Item> &batch, std::vector<ServeResult> &res) {
    const size_t K = batch.size();
    if (batch[0].req.kind == SERVE_RSA) {
        std::vector<RsaRecord> recs(K);
        std::vector<std::array<uint8_t, 32>> pis(K);
        for (size_t i = 0; i < K; i++) {
            recs[i] = batch[i].req.rsa;
            pis[i] = batch[i].req.fs_binding;
        }
        std::vector<RsaProof> out = prove_rsa_batch(c, recs, 1, 32, pis);
        for (size_t i = 0; i < K; i++) {
            memcpy(res[i].pi_hash, out[i].pi_hash, 32);
            memcpy(res[i].root_f0, out[i].root_f0, 32);
            res[i].n_trace = out[i].n_trace;
            res[i].proof = std::move(out[i].proof);
        }
    } else if (batch[0].req.kind == SERVE_P256) {
        std::vector<P256Record> recs(K);
        for (size_t i = 0; i < K; i++) recs[i] = batch[i].req.p256;
        std::vector<P256Proof> out = prove_ecdsa_p256_batch(c, recs);
        for (size_t i = 0; i < K; i++) {
            memcpy(res[i].pi_hash, out[i].pi_hash, 32);
            memcpy(res[i].root_f0, out[i].root_f0, 32);
            res[i].n_trace = out[i].n_trace;
            res[i].proof = std::move(out[i].proof);
        }
    } else if (batch[0].req.kind == SERVE_ED25519) {
        std::vector<EdSig> sigs(K);
        std::vector<std::array<uint8_t, 32>> fsb(K), roots(K);
        for (size_t i = 0; i < K; i++) {
            const ServeRequest &q = batch[i].req;
            sigs[i] = EdSig{q.ed_pk.data(), q.ed_sig.data(), q.ed_msg.data(), q.ed_msg.size()};
            fsb[i] = q.fs_binding;
            roots[i] = q.ed_root;
        }
        std::vector<ZskKsk> out = prove_zsk_ksk_binding_v2_batch(c, sigs, fsb, roots, 256, BLOWUP, 54);
        for (size_t i = 0; i < K; i++) {
            memcpy(res[i].pi_hash, out[i].pi_hash, 32);
            memcpy(res[i].root_f0, out[i].root_f0, 32);
            res[i].n_trace = out[i].n_trace;
            res[i].proof = std::move(out[i].proof);
        }
    } else if (batch[0].req.kind == SERVE_DS_KSK) {
        std::vector<ByteSpan> keys(K);
        std::vector<std::array<uint8_t, 32>> par(K), fsb(K);
        for (size_t i = 0; i < K; i++) {
            keys[i] = {batch[i].req.key.data(), batch[i].req.key.size()};
            par[i] = batch[i].req.parent;
            fsb[i] = batch[i].req.fs_binding;
        }
        std::vector<DsKsk> out = prove_ds_ksk_binding_batch(c, keys, par, fsb, 1);
        for (size_t i = 0; i < K; i++) {
            memcpy(res[i].pi_hash, out[i].pi_hash, 32);
            memcpy(res[i].root_f0, out[i].root_f0, 32);
            memcpy(res[i].aux, out[i].asserted_digest, 32);
            res[i].n_trace = out[i].n_trace;
            res[i].proof = std::move(out[i].proof);
        }
    } else {
        std::vector<std::vector<DnsRec>> shards(K);
        std::vector<std::array<uint8_t, 32>> fsb(K);
        for (size_t i = 0; i < K; i++) {
            const ServeRequest &q = batch[i].req;
            for (size_t r = 0; r < q.domains.size(); r++)
                shards[i].push_back({q.domains[r].data(), q.domains[r].size(), q.rtype[r], q.ttl[r],
                                     q.rdata[r].data(), q.rdata[r].size()});
            fsb[i] = q.fs_binding;
        }
        std::vector<InnerShard> out = prove_inner_shard_batch(c, batch[0].req.salt.data(), shards, fsb, 1);
        for (size_t i = 0; i < K; i++) {
            memcpy(res[i].pi_hash, out[i].pi_hash, 32);
            memcpy(res[i].root_f0, out[i].root_f0, 32);
            memcpy(res[i].aux, out[i].merkle_root, 32);
            res[i].n_trace = out[i].n_trace;
            res[i].proof = std::move(out[i].proof);
        }
    }
}

}  // namespace sg
