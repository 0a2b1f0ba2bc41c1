// Shared internals of libnslx: error handling, device buffers, complex
// helpers usable on host and device.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/nslx.h"

namespace nslx {

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define NSLX_HIP(call)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            throw ::nslx::Error(e_ == hipErrorOutOfMemory ? NSLX_E_OOM : NSLX_E_DEVICE,  \
                                std::string(#call) + ": " + hipGetErrorString(e_) +     \
                                    " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
        }                                                                                \
    } while (0)

#define NSLX_REQUIRE(cond, msg)                                                          \
    do {                                                                                 \
        if (!(cond)) throw ::nslx::Error(NSLX_E_ARG, std::string(msg));                  \
    } while (0)

#define NSLX_LAUNCH_CHECK() NSLX_HIP(hipGetLastError())

// Registry of the library's device allocations, kept only when the
// environment sets NSLX_CHECK_PLANS (launch plans are then checked against it
// on the host before they run).
struct AllocRegistry {
    std::mutex mu;
    std::map<uintptr_t, size_t> m;
    bool on = false;
    AllocRegistry() {
        const char* v = std::getenv("NSLX_CHECK_PLANS");
        on = v && *v && *v != '0';
    }
};
inline AllocRegistry& alloc_registry() {
    static AllocRegistry r;
    return r;
}
inline bool alloc_checks_on() { return alloc_registry().on; }
inline void alloc_note(const void* p, size_t bytes, bool add) {
    AllocRegistry& r = alloc_registry();
    if (!r.on || !p) return;
    std::lock_guard<std::mutex> g(r.mu);
    if (add) r.m[(uintptr_t)p] = bytes;
    else r.m.erase((uintptr_t)p);
}
// whether [p, p + bytes) lies inside one registered allocation
inline bool alloc_covers(const void* p, size_t bytes) {
    AllocRegistry& r = alloc_registry();
    std::lock_guard<std::mutex> g(r.mu);
    const uintptr_t a = (uintptr_t)p;
    auto it = r.m.upper_bound(a);
    if (it == r.m.begin()) return false;
    --it;
    return a >= it->first && a + bytes <= it->first + it->second;
}

// With the registry on: throws unless elements [0, last_elem] of size es
// starting at p lie in one registered allocation.
inline void check_extent(const void* p, long long last_elem, size_t es, const char* what) {
    if (!alloc_checks_on() || last_elem < 0) return;
    if (!p || !alloc_covers(p, (size_t)(last_elem + 1) * es))
        throw Error(NSLX_E_INTERNAL, std::string("launch plan: ") + what + " outside any device allocation");
}

// Owning device allocation (hipMalloc / hipFree).
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { reset(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { reset(); }
    void alloc(size_t count) {
        reset();
        if (count == 0) return;
        NSLX_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
        n = count;
        alloc_note(p, count * sizeof(T), true);
    }
    void ensure(size_t count) { if (count > n) alloc(count); }
    void reset() {
        if (p) { alloc_note(p, 0, false); (void)hipFree(p); p = nullptr; n = 0; }
    }
    T* get() const { return p; }
};

// Pinned host allocation (hipHostMalloc).
template <class T>
struct PinBuf {
    T* p = nullptr;
    size_t n = 0;
    PinBuf() = default;
    PinBuf(const PinBuf&) = delete;
    PinBuf& operator=(const PinBuf&) = delete;
    ~PinBuf() { if (p) (void)hipHostFree(p); }
    void ensure(size_t count) {
        if (count <= n) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        NSLX_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), count * sizeof(T), hipHostMallocDefault));
        n = count;
    }
};

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Interleaved complex storage types (match c10::complex<double|float>).
struct alignas(16) c128 { double re, im; };
struct alignas(8) c64 { float re, im; };

template <class R>
struct cplx { R re, im; };

template <class R>
__host__ __device__ inline cplx<R> cmul(cplx<R> a, cplx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class R>
__host__ __device__ inline cplx<R> cmulconj(cplx<R> a, cplx<R> b) {  // conj(a)*b
    return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}

}  // namespace nslx
