// ctx.cpp -- context creation, root tables, caching allocator, event timing.
#include "ctx.hpp"

#include <stdlib.h>
#include <string.h>

#include <condition_variable>
#include <deque>
#include <mutex>

#include "gl.hpp"

namespace sg {

void *Ctx::alloc(size_t bytes) {
    if (bytes == 0) bytes = 256;
    bytes = (bytes + 255) & ~size_t(255);
    auto it = free_blocks.lower_bound(bytes);
    // reuse the smallest cached block that holds the request when it is at
    // most twice as large: hipMalloc of fresh memory costs time in proportion
    // to its size (tens of GB for an uneven batch), reuse costs nothing
    if (it != free_blocks.end() && it->first <= 2 * bytes) {
        void *p = it->second;
        live[p] = it->first;
        cached_bytes -= it->first;
        free_blocks.erase(it);
        return p;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    // out of memory: hand cached blocks back, largest first, until it fits
    while (e == hipErrorOutOfMemory && !free_blocks.empty()) {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(stream);
        auto big = std::prev(free_blocks.end());
        (void)hipFree(big->second);
        cached_bytes -= big->first;
        free_blocks.erase(big);
        e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(e == hipErrorOutOfMemory ? E_OOM : E_DEVICE,
                    std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    live[p] = bytes;
    return p;
}

void Ctx::d2h(void *dst, const void *src, size_t n) {
    constexpr size_t CHUNK = 64ULL << 20;
    if (n == 0) return;
    const size_t want = n < CHUNK ? n : CHUNK;
    if (pin_bytes < want) {
        if (pin) (void)hipHostFree(pin);
        pin = nullptr;
        pin_bytes = 0;
        const size_t sz = want < (1ULL << 20) ? (1ULL << 20) : want;
        SG_HIP(hipHostMalloc(&pin, sz, hipHostMallocDefault));
        pin_bytes = sz;
    }
    for (size_t off = 0; off < n; off += CHUNK) {
        const size_t m = n - off < CHUNK ? n - off : CHUNK;
        SG_HIP(hipMemcpyAsync(pin, (const uint8_t *)src + off, m, hipMemcpyDeviceToHost, stream));
        SG_HIP(hipStreamSynchronize(stream));
        memcpy((uint8_t *)dst + off, pin, m);
    }
}

const void *Ctx::d2h_pinned(const void *src, size_t n) {
    if (pin_bytes < n) {
        if (pin) (void)hipHostFree(pin);
        pin = nullptr;
        pin_bytes = 0;
        const size_t sz = n < (1ULL << 20) ? (1ULL << 20) : n + n / 4;
        SG_HIP(hipHostMalloc(&pin, sz, hipHostMallocDefault));
        pin_bytes = sz;
    }
    if (n) {
        SG_HIP(hipMemcpyAsync(pin, src, n, hipMemcpyDeviceToHost, stream));
        SG_HIP(hipStreamSynchronize(stream));
    }
    return pin;
}

void Ctx::d2h2(void *dst1, const void *src1, size_t n1, void *dst2, const void *src2, size_t n2) {
    constexpr size_t SMALL = 1ULL << 20;  // the pinned buffer's minimum size
    if (n1 + n2 > SMALL) {
        d2h(dst1, src1, n1);
        d2h(dst2, src2, n2);
        return;
    }
    if (pin_bytes < SMALL) {
        if (pin) (void)hipHostFree(pin);
        pin = nullptr;
        pin_bytes = 0;
        SG_HIP(hipHostMalloc(&pin, SMALL, hipHostMallocDefault));
        pin_bytes = SMALL;
    }
    if (n1) SG_HIP(hipMemcpyAsync(pin, src1, n1, hipMemcpyDeviceToHost, stream));
    if (n2) SG_HIP(hipMemcpyAsync((uint8_t *)pin + n1, src2, n2, hipMemcpyDeviceToHost, stream));
    SG_HIP(hipStreamSynchronize(stream));
    if (n1) memcpy(dst1, pin, n1);
    if (n2) memcpy(dst2, (const uint8_t *)pin + n1, n2);
}

void *Ctx::upload_staging(size_t n) {
    if (up_bytes < n) {
        if (up) (void)hipHostFree(up);
        up = nullptr;
        up_bytes = 0;
        const size_t sz = n < (16ULL << 20) ? (16ULL << 20) : n + n / 4;
        SG_HIP(hipHostMalloc(&up, sz, hipHostMallocDefault));
        up_bytes = sz;
    }
    return up;
}

void Ctx::release(void *p) {
    auto it = live.find(p);
    if (it == live.end()) return;
    free_blocks.emplace(it->second, p);
    cached_bytes += it->second;
    live.erase(it);
}

void Ctx::trim() {
    (void)hipStreamSynchronize(stream);
    for (auto &kv : free_blocks) (void)hipFree(kv.second);
    free_blocks.clear();
    cached_bytes = 0;
}

int Ctx::prof_begin(const char *name, double bytes) {
    Pending p;
    p.name = name;
    p.bytes = bytes;
    SG_HIP(hipEventCreate(&p.a));
    SG_HIP(hipEventCreate(&p.b));
    SG_HIP(hipEventRecord(p.a, stream));
    pending.push_back(p);
    return (int)pending.size() - 1;
}

void Ctx::prof_end(int tok) { SG_HIP(hipEventRecord(pending[tok].b, stream)); }

void Ctx::prof_collect() {
    SG_HIP(hipStreamSynchronize(stream));
    for (auto &p : pending) {
        float ms = 0;
        SG_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        auto &s = stats[p.name];
        s.ms += ms;
        s.launches += 1;
        s.bytes += p.bytes;
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    pending.clear();
}

namespace {
std::atomic<int> autotune_dflt{-1};  // -1: not read from the environment yet
}

bool autotune_default() {
    int v = autotune_dflt.load();
    if (v < 0) {
        const char *e = getenv("STARK_AUTOTUNE");
        int want = (e && e[0] == '1') ? 1 : 0;
        autotune_dflt.compare_exchange_strong(v, want);
        v = autotune_dflt.load();
    }
    return v == 1;
}

void set_autotune_default(bool on) { autotune_dflt.store(on ? 1 : 0); }

Ctx *ctx_create(int device) {
    Ctx *c = new Ctx();
    c->device = device;
    c->autotune = autotune_default();
    try {
        SG_HIP(hipSetDevice(device));
        SG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        // omega_{2^26} two-level table
        std::vector<uint64_t> lo(8192), hi(8192);
        uint64_t w = root_of_unity(1ULL << 26), w13 = pow(w, 8192);
        lo[0] = hi[0] = 1;
        for (int i = 1; i < 8192; i++) {
            lo[i] = mul(lo[i - 1], w);
            hi[i] = mul(hi[i - 1], w13);
        }
        std::vector<uint64_t> tf(8192, 0), ti(8192, 0);
        for (int k = 0; k <= 12; k++) {
            uint64_t wk = root_of_unity(1ULL << k), wki = inv(wk);
            uint64_t a = 1, b = 1;
            for (int t = 0; t < (1 << k); t++) {
                tf[(1 << k) + t] = a;
                ti[(1 << k) + t] = b;
                a = mul(a, wk);
                b = mul(b, wki);
            }
        }
        c->rt_lo = c->alloc_n<uint64_t>(8192);
        c->rt_hi = c->alloc_n<uint64_t>(8192);
        c->tw_f = c->alloc_n<uint64_t>(8192);
        c->tw_i = c->alloc_n<uint64_t>(8192);
        SG_HIP(hipMemcpy(c->rt_lo, lo.data(), 8 * 8192, hipMemcpyHostToDevice));
        SG_HIP(hipMemcpy(c->rt_hi, hi.data(), 8 * 8192, hipMemcpyHostToDevice));
        SG_HIP(hipMemcpy(c->tw_f, tf.data(), 8 * 8192, hipMemcpyHostToDevice));
        SG_HIP(hipMemcpy(c->tw_i, ti.data(), 8 * 8192, hipMemcpyHostToDevice));
    } catch (...) {
        ctx_destroy(c);
        throw;
    }
    return c;
}

namespace {

// Worker pool behind parallel_for. A job is a loop over [0, n) that the
// caller and any idle workers share through an atomic index; the caller
// waits until every index has run and no worker still holds the job.
struct PoolJob {
    void (*call)(void *, size_t);
    void *arg;
    size_t n;
    std::atomic<size_t> next{0}, done{0};
    int users = 0;  // workers inside the job (guarded by Pool::mu)
};

struct Pool {
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<PoolJob *> jobs;
    std::vector<std::thread> threads;
    bool stop = false;

    Pool() {
        size_t nt = std::thread::hardware_concurrency();
        if (nt == 0) nt = 4;
        if (nt > 16) nt = 16;
        for (size_t t = 1; t < nt; t++) threads.emplace_back([this] { work(); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv_work.notify_all();
        for (auto &t : threads) t.join();
    }
    static void run_indices(PoolJob *j) {
        for (size_t i; (i = j->next.fetch_add(1)) < j->n;) {
            j->call(j->arg, i);
            j->done.fetch_add(1);
        }
    }
    void retire(PoolJob *j) {  // with mu held: no new worker picks j up
        for (auto it = jobs.begin(); it != jobs.end(); ++it)
            if (*it == j) {
                jobs.erase(it);
                break;
            }
    }
    void work() {
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            cv_work.wait(lk, [&] { return stop || !jobs.empty(); });
            if (stop) return;
            PoolJob *j = jobs.front();
            j->users++;
            lk.unlock();
            run_indices(j);
            lk.lock();
            j->users--;
            retire(j);
            cv_done.notify_all();
        }
    }
    void run(PoolJob &j) {
        {
            std::lock_guard<std::mutex> g(mu);
            jobs.push_back(&j);
        }
        cv_work.notify_all();
        run_indices(&j);
        std::unique_lock<std::mutex> lk(mu);
        retire(&j);
        cv_done.wait(lk, [&] { return j.done.load() == j.n && j.users == 0; });
    }
};

Pool &pool() {
    static Pool p;
    return p;
}

}  // namespace

void parallel_run(size_t n, void (*call)(void *, size_t), void *arg) {
    PoolJob j;
    j.call = call;
    j.arg = arg;
    j.n = n;
    pool().run(j);
}

void ctx_destroy(Ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &p : c->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto &kv : c->live) (void)hipFree(kv.first);
    for (auto &kv : c->free_blocks) (void)hipFree(kv.second);
    if (c->pin) (void)hipHostFree(c->pin);
    if (c->up) (void)hipHostFree(c->up);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // namespace sg
