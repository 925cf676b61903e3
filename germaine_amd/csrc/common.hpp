// Shared device types and helpers of the gfdfd solve path (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace gfdfd {

// Interleaved complex value (re, im) of real type T.
template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T x, y;
};

// Compact front descriptor as the kernels read it.
struct FrontDev {
    int p, m, pos0, nchild, child_off, level;
    int64_t buf_off, fv_off, idx_off, inv_off;
};

struct TileDev {
    int front, ti, tj;
};

struct hip_error : std::runtime_error {
    hipError_t code;   // what the failing HIP call returned (hipErrorUnknown when no call failed)
    explicit hip_error(const std::string &s, hipError_t c = hipErrorUnknown) : std::runtime_error(s), code(c) {}
};

#define GFDFD_HIP(call)                                                                              \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            throw ::gfdfd::hip_error(std::string(#call) + ": " + hipGetErrorString(e_) + " at " +    \
                                     __FILE__ + ":" + std::to_string(__LINE__), e_);                 \
    } while (0)

// device bytes the library holds now and at most since the last reset
// (every allocation goes through dev_alloc / dev_free)
void mem_track_alloc(void *p, size_t bytes);
void mem_track_free(void *p);

template <typename T>
T *dev_alloc(size_t count) {
    void *p = nullptr;
    if (count == 0) count = 1;
    GFDFD_HIP(hipMalloc(&p, count * sizeof(T)));
    mem_track_alloc(p, count * sizeof(T));
    return static_cast<T *>(p);
}

template <typename T>
T *dev_upload(const std::vector<T> &v) {
    T *p = dev_alloc<T>(v.size());
    if (!v.empty()) GFDFD_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
}

inline void dev_free(void *p) {
    if (p) {
        mem_track_free(p);
        (void)hipFree(p);
    }
}

// ---- device complex arithmetic --------------------------------------------
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) {
    return {a.x + b.x, a.y + b.y};
}
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) {
    return {a.x - b.x, a.y - b.y};
}
// c - a*b
template <typename T>
__device__ __forceinline__ cx<T> cmsub(cx<T> c, cx<T> a, cx<T> b) {
    return {c.x - (a.x * b.x - a.y * b.y), c.y - (a.x * b.y + a.y * b.x)};
}
// a*b and c - a*b with every fused multiply-add written out, so their rounding
// does not depend on how the compiler contracts the expression where it is
// inlined: code that two kernels must run bit-identically uses these
template <typename T>
__device__ __forceinline__ cx<T> cmul_f(cx<T> a, cx<T> b) {
    return {fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x)};
}
template <typename T>
__device__ __forceinline__ cx<T> cmsub_f(cx<T> c, cx<T> a, cx<T> b) {
    return {fma(-a.x, b.x, fma(a.y, b.y, c.x)), fma(-a.x, b.y, fma(-a.y, b.x, c.y))};
}
template <typename T>
__device__ __forceinline__ cx<T> cconjif(cx<T> a, bool c) {
    return c ? cx<T>{a.x, -a.y} : a;
}
template <typename T>
__device__ __forceinline__ T cabs1(cx<T> a) {
    return fabs(a.x) + fabs(a.y);
}
// 1/a by Smith's method (no overflow for large pivots)
template <typename T>
__device__ __forceinline__ cx<T> crecip(cx<T> a) {
    if (fabs(a.x) >= fabs(a.y)) {
        T r = a.y / a.x, d = a.x + a.y * r;
        return {T(1) / d, -r / d};
    } else {
        T r = a.x / a.y, d = a.x * r + a.y;
        return {r / d, T(-1) / d};
    }
}
// a / b by Smith's method
template <typename T>
__device__ __forceinline__ cx<T> cdiv(cx<T> a, cx<T> b) {
    if (fabs(b.x) >= fabs(b.y)) {
        T r = b.y / b.x, d = b.x + b.y * r;
        return {(a.x + a.y * r) / d, (a.y - a.x * r) / d};
    } else {
        T r = b.x / b.y, d = b.x * r + b.y;
        return {(a.x * r + a.y) / d, (a.y * r - a.x) / d};
    }
}

}  // namespace gfdfd
