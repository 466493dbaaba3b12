// server.hpp -- a proving server on one GPU: a request queue admitted by an
// HBM budget, worker threads with a context (HIP stream) each, and
// continuous batching -- a free worker takes the head request plus every
// queued request of the same kind that still fits the budget, up to a
// maximum batch, and proves them as one lockstep device batch while the
// other workers' batches run on their own streams.
//
// It stands in for the swarm worker's request loop
// (crates/swarm-worker/src/main.rs:220-335, which calls prove_inner_shard
// from tokio spawn_blocking threads) on the device side; besides inner
// shards it proves DS->KSK bindings, RSA-2048 and ECDSA-P256 signature
// records and Ed25519 ZSK->KSK bindings (prove_zsk_ksk_binding_v2).
//
// Isolation: submit checks that a signature request's signature verifies
// (the provers would assert); a batch that still fails is proved again one
// request at a time, so only the failing request carries the error. A
// request may carry a deadline: one not admitted in time completes with
// E_TIMEOUT without device work.
#pragma once
#include <stdint.h>

#include <array>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ctx.hpp"
#include "prover.hpp"

namespace sg {

enum ServeKind : int { SERVE_DS_KSK = 0, SERVE_INNER_SHARD = 1, SERVE_RSA = 2, SERVE_P256 = 3, SERVE_ED25519 = 4 };
constexpr int E_TIMEOUT = -6;

struct ServeRequest {
    int kind = SERVE_DS_KSK;
    std::vector<uint8_t> key;           // DS-KSK: DNSKEY bytes
    std::array<uint8_t, 32> parent{};   // DS-KSK: parent DS hash
    std::array<uint8_t, 16> salt{};     // inner shard
    std::vector<std::string> domains;   // inner shard records (owned copies)
    std::vector<std::vector<uint8_t>> rdata;
    std::vector<uint16_t> rtype;
    std::vector<uint32_t> ttl;
    RsaRecord rsa{};                    // RSA-2048: n, s, em (fs_binding = its public-input hash)
    P256Record p256{};                  // ECDSA-P256: qx, qy, r, s, digest
    std::array<uint8_t, 32> ed_pk{};    // Ed25519: public key, signature, signed data, merkle root
    std::array<uint8_t, 64> ed_sig{};
    std::vector<uint8_t> ed_msg;
    std::array<uint8_t, 32> ed_root{};
    std::array<uint8_t, 32> fs_binding{};
    double deadline_s = 0;  // > 0: complete with E_TIMEOUT unless admitted within this many seconds
};

struct ServeResult {
    uint64_t id = 0;
    int status = 0;  // 0 or an E_* code
    std::string error;
    int kind = 0;
    uint8_t pi_hash[32] = {}, root_f0[32] = {}, aux[32] = {};  // aux: merkle root / asserted digest / zero
    size_t n_trace = 0, batch_size = 0;
    uint64_t batch_seq = 0;  // admission order of the batch the request ran in
    std::vector<uint8_t> proof;
    double t_submit = 0, t_start = 0, t_done = 0;  // seconds since the server started
};

class Server {
public:
    // dry_run: schedule only (a worker sleeps in proportion to the request's
    // memory estimate and returns no proof) -- for testing admission on hosts
    // without a GPU.
    // validate: check signatures at submit (off only to test the per-request
    // fallback of a failing batch)
    Server(int device, int workers, uint64_t hbm_budget, uint32_t max_batch, bool dry_run, bool validate = true);
    ~Server();
    uint64_t submit(ServeRequest &&r);
    // false on timeout; the result is removed from the server once returned
    bool wait(uint64_t id, ServeResult &out, double timeout_s);
    size_t completed() const;
    uint64_t peak_in_use() const;
    uint64_t budget() const { return budget_; }
    // device bytes a request is admitted with (its lockstep share of trace,
    // LDE, merge and FRI buffers)
    static uint64_t need_bytes(const ServeRequest &r);

private:
    struct Item {
        uint64_t id;
        ServeRequest req;
        uint64_t need;
        double t_submit;
    };
    void worker(int w);
    void run_batch(Ctx *c, std::vector<Item> &batch, std::vector<ServeResult> &res);
    void prove_batch(Ctx *c, std::vector<Item> &batch, std::vector<ServeResult> &res);
    double now() const;

    const int device_;
    const uint64_t budget_;
    const uint32_t max_batch_;
    const bool dry_, validate_;
    mutable std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<Item> queue_;
    std::map<uint64_t, ServeResult> done_;
    uint64_t next_id_ = 1, in_use_ = 0, peak_ = 0, batch_seq_ = 0;
    size_t completed_ = 0;
    bool stop_ = false;
    std::vector<std::thread> threads_;
    std::vector<Ctx *> ctxs_;
    std::chrono::steady_clock::time_point t0_;
};

}  // namespace sg
