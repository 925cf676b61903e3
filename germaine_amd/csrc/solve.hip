// Multi-right-hand-side forward/backward substitution on the multifrontal
// factors, FP64 residuals and iterative refinement.
//
// The right-hand sides live in elimination order as a row-major panel
// X[pos][r] (all right-hand sides of one unknown contiguous), so every gather
// and scatter of a front moves whole rows.  The forward sweep walks the
// levels upwards: a front's work rows W[f] (m x nrhs) are its own pivot rows of
// X plus the children's contribution rows (pulled through the child->parent
// maps), then y1 = L11^-1 P W[0:p] and W[p:m] -= L21 y1 stay in W for the
// parent.  The backward sweep walks downwards: x1 = U11^-1 (y1 - U12 x2) with x2
// gathered from the (already final) ancestors.  Transposed solves (M^T, M^H)
// use U^T in the forward and L^T in the backward sweep and apply the row
// swaps at the end.  Each thread of the front kernels owns one right-hand
// side; factor entries are wave-uniform (broadcast) loads and the work rows
// are blocked through registers.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "engine.hpp"
#include "runtime.hpp"

namespace gfdfd {

namespace {

constexpr int TB = 16;   // register block of the substitutions

template <typename T>
__device__ __forceinline__ cx<T> ld(const cx<T> *p, bool cj) {
    cx<T> v = *p;
    if (cj) v.y = -v.y;
    return v;
}

template <typename T, typename TIO>
__global__ void __launch_bounds__(256) k_permute_in(const cx<TIO> *__restrict__ B, cx<T> *__restrict__ X,
                                                    const int *__restrict__ perm, int n, int nrhs) {
    const int64_t total = (int64_t)n * nrhs;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int pos = (int)(e / nrhs), r = (int)(e - (int64_t)pos * nrhs);
        cx<TIO> v = B[perm[pos] + (int64_t)r * n];
        X[e] = {(T)v.x, (T)v.y};
    }
}

template <typename T, typename TIO>
__global__ void __launch_bounds__(256) k_permute_out(const cx<T> *__restrict__ X, cx<TIO> *__restrict__ B,
                                                     const int *__restrict__ perm, int n, int nrhs) {
    const int64_t total = (int64_t)n * nrhs;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int pos = (int)(e / nrhs), r = (int)(e - (int64_t)pos * nrhs);
        cx<T> v = X[e];
        B[perm[pos] + (int64_t)r * n] = {(TIO)v.x, (TIO)v.y};
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_fwd_gather(const cx<T> *__restrict__ X, cx<T> *__restrict__ W,
                                                    const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                    const int *__restrict__ children, const int *__restrict__ inv,
                                                    int nrhs) {
    const FrontDev F = fronts[list[blockIdx.x]];
    const int64_t total = (int64_t)F.m * nrhs;
    for (int64_t e = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.y * blockDim.x) {
        int i = (int)(e / nrhs), r = (int)(e - (int64_t)i * nrhs);
        cx<T> v = {T(0), T(0)};
        if (i < F.p) v = X[(int64_t)(F.pos0 + i) * nrhs + r];
        for (int c = 0; c < F.nchild; c++) {
            const FrontDev C = fronts[children[F.child_off + c]];
            int a = inv[C.inv_off + i];
            if (a >= 0) v = cadd(v, W[(C.fv_off + C.p + a) * nrhs + r]);
        }
        W[(F.fv_off + i) * nrhs + r] = v;
    }
}

// forward sweep of one front: lower solve on the pivot rows, update of the contribution rows
template <typename T>
__global__ void __launch_bounds__(64) k_fwd_front(const cx<T> *__restrict__ arena, const FrontDev *__restrict__ fronts,
                                                  const int *__restrict__ list, const int *__restrict__ piv,
                                                  cx<T> *__restrict__ X, cx<T> *__restrict__ W, int nrhs, int trans,
                                                  int cj) {
    const FrontDev F = fronts[list[blockIdx.x]];
    const int r = blockIdx.y * 64 + threadIdx.x;
    if (r >= nrhs) return;
    const int m = F.m, p = F.p;
    const cx<T> *G = arena + F.buf_off;
    cx<T> *w = W + F.fv_off * nrhs + r;
    // element (i,t) of the lower operator: L (unit) or U^T (non-unit)
    auto Lop = [&](int i, int t) { return trans ? ld(G + t + (int64_t)i * m, cj) : ld(G + i + (int64_t)t * m, cj); };
    if (!trans)
        for (int i = 0; i < p; i++) {
            int t = piv[F.pos0 + i];
            if (t != i) {
                cx<T> a = w[(int64_t)i * nrhs], b = w[(int64_t)t * nrhs];
                w[(int64_t)i * nrhs] = b;
                w[(int64_t)t * nrhs] = a;
            }
        }
    for (int tb = 0; tb < p; tb += TB) {
        const int nb = min(TB, p - tb);
        cx<T> y[TB];
#pragma unroll
        for (int k = 0; k < TB; k++) y[k] = k < nb ? w[(int64_t)(tb + k) * nrhs] : cx<T>{T(0), T(0)};
#pragma unroll
        for (int k = 0; k < TB; k++) {
            if (k < nb) {
#pragma unroll
                for (int t = 0; t < k; t++) y[k] = cmsub(y[k], Lop(tb + k, tb + t), y[t]);
                if (trans) y[k] = cdiv(y[k], Lop(tb + k, tb + k));
            }
        }
#pragma unroll
        for (int k = 0; k < TB; k++)
            if (k < nb) w[(int64_t)(tb + k) * nrhs] = y[k];
        for (int i = tb + nb; i < m; i++) {
            cx<T> acc = w[(int64_t)i * nrhs];
#pragma unroll
            for (int k = 0; k < TB; k++)
                if (k < nb) acc = cmsub(acc, Lop(i, tb + k), y[k]);
            w[(int64_t)i * nrhs] = acc;
        }
    }
    for (int i = 0; i < p; i++) X[(int64_t)(F.pos0 + i) * nrhs + r] = w[(int64_t)i * nrhs];
}

// backward sweep of one front: x1 = Uop^-1 (y1 - Uop12 x2)
template <typename T>
__global__ void __launch_bounds__(64) k_bwd_front(const cx<T> *__restrict__ arena, const FrontDev *__restrict__ fronts,
                                                  const int *__restrict__ list, const int *__restrict__ idx,
                                                  cx<T> *__restrict__ X, cx<T> *__restrict__ W, int nrhs, int trans,
                                                  int cj) {
    const FrontDev F = fronts[list[blockIdx.x]];
    const int r = blockIdx.y * 64 + threadIdx.x;
    if (r >= nrhs) return;
    const int m = F.m, p = F.p;
    const cx<T> *G = arena + F.buf_off;
    cx<T> *w = W + F.fv_off * nrhs + r;
    // element (i,t), t >= i, of the upper operator: U (non-unit) or L^T (unit)
    auto Uop = [&](int i, int t) { return trans ? ld(G + t + (int64_t)i * m, cj) : ld(G + i + (int64_t)t * m, cj); };
    for (int i = 0; i < p; i++) w[(int64_t)i * nrhs] = X[(int64_t)(F.pos0 + i) * nrhs + r];
    // contribution of the ancestors
    for (int sb = p; sb < m; sb += TB) {
        const int nb = min(TB, m - sb);
        cx<T> x2[TB];
#pragma unroll
        for (int k = 0; k < TB; k++)
            x2[k] = k < nb ? X[(int64_t)idx[F.idx_off + sb + k] * nrhs + r] : cx<T>{T(0), T(0)};
        for (int i = 0; i < p; i++) {
            cx<T> acc = w[(int64_t)i * nrhs];
#pragma unroll
            for (int k = 0; k < TB; k++)
                if (k < nb) acc = cmsub(acc, Uop(i, sb + k), x2[k]);
            w[(int64_t)i * nrhs] = acc;
        }
    }
    // triangular part, blocks from the bottom
    for (int tb = ((p - 1) / TB) * TB; tb >= 0; tb -= TB) {
        const int nb = min(TB, p - tb);
        cx<T> x[TB];
#pragma unroll
        for (int k = 0; k < TB; k++) x[k] = k < nb ? w[(int64_t)(tb + k) * nrhs] : cx<T>{T(0), T(0)};
#pragma unroll
        for (int k = TB - 1; k >= 0; k--) {
            if (k < nb) {
#pragma unroll
                for (int t = k + 1; t < TB; t++)
                    if (t < nb) x[k] = cmsub(x[k], Uop(tb + k, tb + t), x[t]);
                if (!trans) x[k] = cdiv(x[k], Uop(tb + k, tb + k));
            }
        }
#pragma unroll
        for (int k = 0; k < TB; k++)
            if (k < nb) w[(int64_t)(tb + k) * nrhs] = x[k];
        for (int i = 0; i < tb; i++) {
            cx<T> acc = w[(int64_t)i * nrhs];
#pragma unroll
            for (int k = 0; k < TB; k++)
                if (k < nb) acc = cmsub(acc, Uop(i, tb + k), x[k]);
            w[(int64_t)i * nrhs] = acc;
        }
    }
    for (int i = 0; i < p; i++) X[(int64_t)(F.pos0 + i) * nrhs + r] = w[(int64_t)i * nrhs];
}

// x = P^T w for transposed solves: undo each front's row swaps, last first
template <typename T>
__global__ void __launch_bounds__(64) k_unswap(const FrontDev *__restrict__ fronts, const int *__restrict__ piv,
                                               cx<T> *__restrict__ X, int nrhs) {
    const FrontDev F = fronts[blockIdx.x];
    const int r = blockIdx.y * 64 + threadIdx.x;
    if (r >= nrhs) return;
    for (int i = F.p - 1; i >= 0; i--) {
        int t = piv[F.pos0 + i];
        if (t != i) {
            cx<T> a = X[(int64_t)(F.pos0 + i) * nrhs + r], b = X[(int64_t)(F.pos0 + t) * nrhs + r];
            X[(int64_t)(F.pos0 + i) * nrhs + r] = b;
            X[(int64_t)(F.pos0 + t) * nrhs + r] = a;
        }
    }
}

// r = b - op(M) x, FP64, column-major n x nrhs panels
__global__ void __launch_bounds__(256) k_residual(const cx<double> *__restrict__ vals, const int *__restrict__ colptr,
                                                  const int *__restrict__ rowind, const int64_t *__restrict__ tq,
                                                  int n, int nrhs, int sys, const cx<double> *__restrict__ B,
                                                  const cx<double> *__restrict__ X, cx<double> *__restrict__ R) {
    const int64_t total = (int64_t)n * nrhs;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int rr = (int)(e / n), row = (int)(e - (int64_t)rr * n);
        const cx<double> *x = X + (int64_t)rr * n;
        cx<double> acc = B[e];
        // the pattern is symmetric: row `row` of M has the columns listed in CSC column `row`
        for (int q = colptr[row]; q < colptr[row + 1]; q++) {
            int c = rowind[q];
            cx<double> a = sys == 0 ? vals[tq[q]] : vals[q];
            if (sys == 1) a.y = -a.y;
            acc = cmsub(acc, a, x[c]);
        }
        R[e] = acc;
    }
}

// backward error of one column: residual and max_i |r_i| / (|op(M)||x| + |b|)_i
__global__ void __launch_bounds__(256) k_backerr(const cx<double> *__restrict__ vals, const int *__restrict__ colptr,
                                                 const int *__restrict__ rowind, const int64_t *__restrict__ tq, int n,
                                                 int sys, const cx<double> *__restrict__ B,
                                                 const cx<double> *__restrict__ X, cx<double> *__restrict__ R,
                                                 unsigned long long *__restrict__ omega_bits) {
    __shared__ double s[256];
    double w = 0.0;
    for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < n; row += gridDim.x * blockDim.x) {
        cx<double> acc = B[row];
        double den = hypot(acc.x, acc.y);
        for (int q = colptr[row]; q < colptr[row + 1]; q++) {
            cx<double> a = sys == 0 ? vals[tq[q]] : vals[q];
            if (sys == 1) a.y = -a.y;
            cx<double> x = X[rowind[q]];
            acc = cmsub(acc, a, x);
            den += hypot(a.x, a.y) * hypot(x.x, x.y);
        }
        R[row] = acc;
        double num = hypot(acc.x, acc.y);
        double o = den > 0.0 ? num / den : (num > 0.0 ? INFINITY : 0.0);
        w = fmax(w, o);
    }
    s[threadIdx.x] = w;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(omega_bits, (unsigned long long)__double_as_longlong(s[0]));
}

__global__ void __launch_bounds__(256) k_colnorm2(const cx<double> *__restrict__ V, int n, double *__restrict__ out) {
    __shared__ double s[256];
    const cx<double> *v = V + (int64_t)blockIdx.x * n;
    double acc = 0;
    for (int i = threadIdx.x; i < n; i += 256) acc += v[i].x * v[i].x + v[i].y * v[i].y;
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

__global__ void k_axpy(int64_t count, const cx<double> *__restrict__ D, cx<double> *__restrict__ X) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (int64_t)gridDim.x * blockDim.x)
        X[e] = cadd(X[e], D[e]);
}

template <typename A, typename B>
__global__ void k_cast(int64_t count, const cx<A> *__restrict__ a, cx<B> *__restrict__ b) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (int64_t)gridDim.x * blockDim.x)
        b[e] = {(B)a[e].x, (B)a[e].y};
}

inline int grid_for(int64_t total, int bs = 256, int cap = 8192) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((total + bs - 1) / bs, cap));
}

// per-thread workspace cache (grown on demand, never shrunk)
struct Workspace {
    void *p = nullptr;
    size_t bytes = 0;
    void *get(size_t b) {
        if (b > bytes) {
            dev_free(p);
            p = nullptr;
            bytes = 0;
            GFDFD_HIP(hipMalloc(&p, b));
            bytes = b;
        }
        return p;
    }
    ~Workspace() { dev_free(p); }
};
thread_local Workspace ws_x, ws_w, ws_r, ws_d;
thread_local Workspace ws_n;

}  // namespace

template <typename T, typename TIO>
void solve_panel(const Factors<T> &F, int ifreq, int sys, int nrhs, const cx<TIO> *B, cx<TIO> *X, hipStream_t st) {
    const DevPlan &dp = *F.plan;
    const Plan &pl = *dp.host;
    const int n = pl.n;
    if (ifreq < 0 || ifreq >= F.nfreq) throw std::invalid_argument("frequency index out of range");
    if (sys < 0 || sys > 2) throw std::invalid_argument("invalid system");
    const bool trans = sys != 0, cj = sys == 1;
    const cx<T> *arena = F.arena + ifreq * F.fstride;
    const int *piv = F.piv + (int64_t)ifreq * n;
    cx<T> *Xp = (cx<T> *)ws_x.get(sizeof(cx<T>) * (size_t)n * nrhs);
    cx<T> *Wp = (cx<T> *)ws_w.get(sizeof(cx<T>) * (size_t)pl.fv_rows * nrhs);
    hipLaunchKernelGGL((k_permute_in<T, TIO>), dim3(grid_for((int64_t)n * nrhs)), dim3(256), 0, st, B, Xp, dp.perm, n,
                       nrhs);
    const int rchunks = (nrhs + 63) / 64;
    for (int l = 0; l < pl.nlevels; l++) {
        if (dp.n_all[l] == 0) continue;
        int64_t work = (int64_t)dp.mmax_all[l] * nrhs;
        int chunks = (int)std::min<int64_t>(std::max<int64_t>(1, work / 1024), 256);
        hipLaunchKernelGGL(k_fwd_gather<T>, dim3(dp.n_all[l], chunks), dim3(256), 0, st, Xp, Wp, dp.fronts,
                           dp.lv_all[l], dp.children, dp.inv, nrhs);
        hipLaunchKernelGGL(k_fwd_front<T>, dim3(dp.n_all[l], rchunks), dim3(64), 0, st, arena, dp.fronts, dp.lv_all[l],
                           piv, Xp, Wp, nrhs, (int)trans, (int)cj);
    }
    for (int l = pl.nlevels - 1; l >= 0; l--) {
        if (dp.n_all[l] == 0) continue;
        hipLaunchKernelGGL(k_bwd_front<T>, dim3(dp.n_all[l], rchunks), dim3(64), 0, st, arena, dp.fronts, dp.lv_all[l],
                           dp.idx, Xp, Wp, nrhs, (int)trans, (int)cj);
    }
    if (trans)
        hipLaunchKernelGGL(k_unswap<T>, dim3((int)pl.fronts.size(), rchunks), dim3(64), 0, st, dp.fronts, piv, Xp,
                           nrhs);
    hipLaunchKernelGGL((k_permute_out<T, TIO>), dim3(grid_for((int64_t)n * nrhs)), dim3(256), 0, st, Xp, X, dp.perm,
                       n, nrhs);
    GFDFD_HIP(hipGetLastError());
}

template void solve_panel<double, double>(const Factors<double> &, int, int, int, const cx<double> *, cx<double> *,
                                          hipStream_t);
template void solve_panel<float, double>(const Factors<float> &, int, int, int, const cx<double> *, cx<double> *,
                                         hipStream_t);
template void solve_panel<float, float>(const Factors<float> &, int, int, int, const cx<float> *, cx<float> *,
                                        hipStream_t);

void residual_fp64(const DevPlan &dp, const cx<double> *vals, int sys, int nrhs, const cx<double> *B,
                   const cx<double> *X, cx<double> *R, hipStream_t st) {
    if (!dp.tq) throw std::logic_error("residual needs a grid plan");
    const int n = dp.host->n;
    hipLaunchKernelGGL(k_residual, dim3(grid_for((int64_t)n * nrhs)), dim3(256), 0, st, vals, dp.colptr, dp.rowind,
                       dp.tq, n, nrhs, sys, B, X, R);
}

double backward_error(const DevPlan &dp, const cx<double> *vals, int sys, const cx<double> *B, const cx<double> *X,
                      cx<double> *R, hipStream_t st) {
    if (!dp.tq) throw std::logic_error("backward error needs a grid plan");
    const int n = dp.host->n;
    unsigned long long *d = (unsigned long long *)ws_n.get(sizeof(unsigned long long));
    GFDFD_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), st));
    hipLaunchKernelGGL(k_backerr, dim3(grid_for(n, 256, 2048)), dim3(256), 0, st, vals, dp.colptr, dp.rowind, dp.tq, n,
                       sys, B, X, R, d);
    unsigned long long h = 0;
    GFDFD_HIP(hipMemcpyAsync(&h, d, sizeof h, hipMemcpyDeviceToHost, st));
    GFDFD_HIP(hipStreamSynchronize(st));
    double w;
    std::memcpy(&w, &h, sizeof w);
    return w;
}

void column_norms2(int n, int nrhs, const cx<double> *V, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_colnorm2, dim3(nrhs), dim3(256), 0, st, V, n, out);
}

void axpy_fp64(int64_t count, const cx<double> *D, cx<double> *X, hipStream_t st) {
    hipLaunchKernelGGL(k_axpy, dim3(grid_for(count)), dim3(256), 0, st, count, D, X);
}

void cast_c128_to_c64(int64_t count, const cx<double> *a, cx<float> *b, hipStream_t st) {
    hipLaunchKernelGGL((k_cast<double, float>), dim3(grid_for(count)), dim3(256), 0, st, count, a, b);
}
void cast_c64_to_c128(int64_t count, const cx<float> *a, cx<double> *b, hipStream_t st) {
    hipLaunchKernelGGL((k_cast<float, double>), dim3(grid_for(count)), dim3(256), 0, st, count, a, b);
}

template <typename T>
RefineStats solve_refined(const Factors<T> &F, int ifreq, int sys, int nrhs, const cx<double> *B, cx<double> *X,
                          double tol, int max_iter, hipStream_t st) {
    const DevPlan &dp = *F.plan;
    const int n = dp.host->n;
    const int64_t cnt = (int64_t)n * nrhs;
    const cx<double> *vals = F.vals + (int64_t)ifreq * dp.host->nnz;
    cx<double> *R = (cx<double> *)ws_r.get(sizeof(cx<double>) * cnt);
    cx<double> *D = (cx<double> *)ws_d.get(sizeof(cx<double>) * cnt);
    double *nrm = (double *)ws_n.get(sizeof(double) * 2 * nrhs);
    std::vector<double> h(2 * nrhs);
    column_norms2(n, nrhs, B, nrm + nrhs, st);
    solve_panel<T, double>(F, ifreq, sys, nrhs, B, X, st);
    RefineStats rs;
    for (int it = 0;; it++) {
        residual_fp64(dp, vals, sys, nrhs, B, X, R, st);
        column_norms2(n, nrhs, R, nrm, st);
        GFDFD_HIP(hipMemcpyAsync(h.data(), nrm, sizeof(double) * 2 * nrhs, hipMemcpyDeviceToHost, st));
        GFDFD_HIP(hipStreamSynchronize(st));
        double worst = 0;
        for (int r = 0; r < nrhs; r++) {
            double b = std::sqrt(h[nrhs + r]), rr = std::sqrt(h[r]);
            worst = std::max(worst, b > 0 ? rr / b : rr);
        }
        rs.max_rel_residual = worst;
        rs.iterations = it;
        if (worst <= tol || it >= max_iter) break;
        solve_panel<T, double>(F, ifreq, sys, nrhs, R, D, st);
        axpy_fp64(cnt, D, X, st);
    }
    return rs;
}

template RefineStats solve_refined<double>(const Factors<double> &, int, int, int, const cx<double> *, cx<double> *,
                                           double, int, hipStream_t);
template RefineStats solve_refined<float>(const Factors<float> &, int, int, int, const cx<double> *, cx<double> *,
                                          double, int, hipStream_t);

}  // namespace gfdfd
