// Complex products on gfx950 matrix cores: the real MFMA 16x16x4 instructions
// for f32 and f64, four real products per complex product (cmma) or three
// (cmma3 / cacc3, Gauss's form).
//
// Operand layout (one wave): A lane l -> (row l%16, k l/16), B lane l ->
// (k l/16, col l%16).  Accumulator register v of lane l holds column l%16 and
// row 4*(l/16)+v for f32, but row 4*v+l/16 for f64.
#pragma once
#include <hip/hip_runtime.h>

namespace gfdfd {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <typename T>
struct mfma_t;
template <>
struct mfma_t<double> {
    typedef d4v acc;
    static __device__ __forceinline__ int row(int lk, int v) { return 4 * v + lk; }
    static __device__ __forceinline__ acc op(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <>
struct mfma_t<float> {
    typedef f4v acc;
    static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
    static __device__ __forceinline__ acc op(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// (cr, ci) += (ar + i ai) * (br + i bi) on one 16x16 block
template <typename T>
__device__ __forceinline__ void cmma(T ar, T ai, T br, T bi, typename mfma_t<T>::acc &cr,
                                     typename mfma_t<T>::acc &ci) {
    cr = mfma_t<T>::op(ar, br, cr);
    cr = mfma_t<T>::op(-ai, bi, cr);
    ci = mfma_t<T>::op(ar, bi, ci);
    ci = mfma_t<T>::op(ai, br, ci);
}

// Three real products per complex product (Gauss): t1 += ar br, t2 += ai bi,
// t3 += (ar + ai)(br + bi); then re = t1 - t2, im = t3 - t1 - t2.  bs = br + bi
// is passed in, as the caller reuses it across strips.
template <typename T>
__device__ __forceinline__ void cmma3(T ar, T ai, T br, T bi, T bs, typename mfma_t<T>::acc &t1,
                                      typename mfma_t<T>::acc &t2, typename mfma_t<T>::acc &t3) {
    t1 = mfma_t<T>::op(ar, br, t1);
    t2 = mfma_t<T>::op(ai, bi, t2);
    t3 = mfma_t<T>::op(ar + ai, bs, t3);
}
// Accumulators of one 16x16 complex block: the three-product form (three
// accumulators) by default in FP32; four real products into (re, im) by
// default in FP64, where the three-product form measured slower in the
// LDS-tiled products (it stays faster where the operands sit in registers:
// those kernels ask for it with Three = true).
template <typename T, bool Three = sizeof(T) == 4>
struct cacc3 {
    static constexpr bool three = Three;
    typename mfma_t<T>::acc t1, t2, t3;
    __device__ __forceinline__ void zero() {
        t1 = typename mfma_t<T>::acc{0, 0, 0, 0};
        t2 = t1;
        t3 = t1;
    }
    __device__ __forceinline__ void mma(T ar, T ai, T br, T bi, T bs) {
        if constexpr (three) cmma3<T>(ar, ai, br, bi, bs, t1, t2, t3);
        else cmma<T>(ar, ai, br, bi, t1, t2);
    }
    __device__ __forceinline__ T re(int v) const {
        if constexpr (three) return t1[v] - t2[v];
        else return t1[v];
    }
    __device__ __forceinline__ T im(int v) const {
        if constexpr (three) return t3[v] - t1[v] - t2[v];
        else return t2[v];
    }
};
template <typename T>
__device__ __forceinline__ void cmma3_finish(const typename mfma_t<T>::acc &t1, const typename mfma_t<T>::acc &t2,
                                             const typename mfma_t<T>::acc &t3, typename mfma_t<T>::acc &cr,
                                             typename mfma_t<T>::acc &ci) {
    cr = t1 - t2;
    ci = t3 - t1 - t2;
}

}  // namespace gfdfd
