// ctx.hpp -- per-context state of the gfx950 proving library: one HIP stream,
// device-resident twiddle tables, a caching device allocator, optional
// per-kernel event timing, and the error string the C ABI reports.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <chrono>
#include <iterator>
#include <map>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace sg {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

enum Status {
    OK = 0,
    E_INVALID = -1,
    E_OOM = -2,
    E_DEVICE = -3,
    E_UNSUPPORTED = -4,
    E_INTERNAL = -5,
};

#define SG_HIP(expr)                                                                           \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            throw ::sg::Error(_e == hipErrorOutOfMemory ? ::sg::E_OOM : ::sg::E_DEVICE,        \
                              std::string(#expr) + ": " + hipGetErrorString(_e));             \
    } while (0)

#define SG_REQUIRE(cond, msg)                                     \
    do {                                                          \
        if (!(cond)) throw ::sg::Error(::sg::E_INVALID, (msg));  \
    } while (0)

// Device buffer owned by a context's allocator.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct KernelStat {
    double ms = 0;
    uint64_t launches = 0;
    double bytes = 0;  // algorithmic bytes credited to the launches
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;

    // root tables: w26 = omega_{2^26}; rt_lo[e & 8191] * rt_hi[e >> 13] = w26^e
    uint64_t *rt_lo = nullptr, *rt_hi = nullptr;
    // tw_f[(1 << k) + t] = omega_{2^k}^t, tw_i = inverse, k <= 12
    uint64_t *tw_f = nullptr, *tw_i = nullptr;

    // caching allocator: freed blocks are kept by size for reuse
    std::multimap<size_t, void *> free_blocks;
    std::map<void *, size_t> live;
    size_t cached_bytes = 0;

    // profiling
    bool profiling = false;
    struct Pending {
        std::string name;
        hipEvent_t a, b;
        double bytes;
    };
    std::vector<Pending> pending;
    std::map<std::string, KernelStat> stats;

    // pinned staging for device-to-host reads: a copy into pageable memory
    // has cost milliseconds for tens of KB, one into pinned memory does not
    void *pin = nullptr;
    size_t pin_bytes = 0;
    // pinned host memory that host-built inputs are written into before an
    // upload (the SHA / RSA / P-256 witness words); reused across calls, so a
    // caller must have waited for its upload before it returns
    void *up = nullptr;
    size_t up_bytes = 0;
    // shift power tables (s^e for e < 2^26, two levels of 8192) by shift,
    // built once per context
    std::map<uint64_t, uint64_t *> pow_tables;
    // odd-coset twist tables w_{2n}^i / n (i < n) by log2 n, built once per context
    std::map<uint32_t, uint64_t *> odd_twists;
    // NTT / LDE launch plans: with autotune on, the first transform of each
    // shape times the admissible plans (pass split, tile) and the process
    // keeps the fastest; force_split (log2 of the first pass, 0 = off) and
    // force_tile (13 or 14, 0 = off) pin the split and the tile of every
    // transform that admits them.
    bool autotune = false;
    uint32_t force_split = 0, force_tile = 0;
    void *upload_staging(size_t n);
    // copy n bytes of device memory to host memory after the stream's
    // earlier work; returns once dst holds them
    void d2h(void *dst, const void *src, size_t n);
    // n bytes from device src into the context's pinned buffer, behind one
    // wait; the pointer stays valid until the next d2h* call on the context
    const void *d2h_pinned(const void *src, size_t n);
    // two such copies behind one wait (small reads of one stage, e.g. a
    // layer's coset evaluations and its roots)
    void d2h2(void *dst1, const void *src1, size_t n1, void *dst2, const void *src2, size_t n2);

    void *alloc(size_t bytes);
    void release(void *p);
    void trim();
    template <class T>
    T *alloc_n(size_t n) { return (T *)alloc(n * sizeof(T)); }

    // profiling helpers: begin/end around one launch on `stream`
    int prof_begin(const char *name, double bytes);
    void prof_end(int token);
    void prof_collect();  // synchronizes and folds pending events into stats
};

// A borrowed byte string.
struct ByteSpan {
    const uint8_t *p;
    size_t n;
};

// RAII holder for a context allocation
struct Scoped {
    Ctx *c;
    void *p;
    Scoped(Ctx *c_, size_t bytes) : c(c_), p(c_->alloc(bytes)) {}
    ~Scoped() { if (p) c->release(p); }
    Scoped(const Scoped &) = delete;
    Scoped &operator=(const Scoped &) = delete;
    template <class T>
    T *as() const { return (T *)p; }
};

// RAII profiler scope
struct ProfScope {
    Ctx *c;
    int tok;
    ProfScope(Ctx *c_, const char *name, double bytes) : c(c_), tok(c_->profiling ? c_->prof_begin(name, bytes) : -1) {}
    ~ProfScope() { if (tok >= 0) c->prof_end(tok); }
};

// Host wall time of a stage, kept beside the kernel stats as "host:<name>"
// (launches counts the calls). Only for host code on the calling thread.
struct HostScope {
    Ctx *c;
    const char *name;
    std::chrono::steady_clock::time_point t0;
    HostScope(Ctx *c_, const char *n) : c(c_), name(n), t0(std::chrono::steady_clock::now()) {}
    ~HostScope() {
        if (!c->profiling) return;
        auto &s = c->stats[std::string("host:") + name];
        s.ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        s.launches += 1;
    }
};

Ctx *ctx_create(int device);
void ctx_destroy(Ctx *c);
// Whether contexts created from now on tune their NTT / LDE plans (the
// proving server's worker contexts included). Starts as STARK_AUTOTUNE=1 in
// the environment, else off.
bool autotune_default();
void set_autotune_default(bool on);

// Run fn(i) for i < n on a process-wide pool of up to 16 host threads (the
// per-proof host work of a batch: transcripts, record hashing,
// serialization); the caller works on its own loop too. fn must not throw.
void parallel_run(size_t n, void (*call)(void *, size_t), void *arg);
template <class F>
void parallel_for(size_t n, F fn) {
    if (n <= 1) {
        for (size_t i = 0; i < n; i++) fn(i);
        return;
    }
    parallel_run(n, [](void *a, size_t i) { (*static_cast<F *>(a))(i); }, &fn);
}

inline uint32_t ilog2(uint64_t n) {
    uint32_t l = 0;
    while ((1ULL << l) < n) l++;
    return l;
}
inline bool is_pow2(uint64_t n) { return n && !(n & (n - 1)); }

}  // namespace sg
