// ctx.cpp -- context creation, root tables, caching allocator, event timing.
#include "ctx.hpp"

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

Ctx *ctx_create(int device) {
    Ctx *c = new Ctx();
    c->device = device;
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

void ctx_destroy(Ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &p : c->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto &kv : c->live) (void)hipFree(kv.first);
    for (auto &kv : c->free_blocks) (void)hipFree(kv.second);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // namespace sg
