// Multi-right-hand-side forward/backward substitution on the multifrontal
// factors, FP64 residuals and iterative refinement.
//
// Right-hand sides live in elimination order as a row-major panel X[pos][r]
// (all right-hand sides of one unknown contiguous), so gathers and scatters
// move whole rows.  The forward sweep walks the levels upwards: a front's work
// rows W[f] (m x nrhs) are its own pivot rows of X (row-permuted by the
// front's pivoting) plus the children's contribution rows, pulled through the
// child->parent maps.  The pivot rows are then solved in 32-row blocks: each
// block kernel solves its 32x32 diagonal triangle in LDS and applies the
// block's columns to every later row of the front as an MFMA product
// (64 rows x 64 right-hand sides per workgroup); the contribution rows left in
// W feed the parent.  The backward sweep walks the levels downwards:
// z = y1 - U12 x2 as one MFMA product over the ancestors' rows of X, then
// 32-row blocks from the bottom.  Transposed solves (M^T, M^H) use U^T in the
// forward and L^T in the backward sweep and undo each front's row swaps as soon
// as its backward step is done, before any descendant reads its unknowns (the
// descendants' L21 rows are stored in the front's original row order).  Up to four right-hand sides take a single-vector path: one
// workgroup per front with the work vector in LDS.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "engine.hpp"
#include "mfma.hpp"
#include "prof.hpp"
#include "runtime.hpp"

namespace gfdfd {

namespace {

constexpr int kVecMaxRhs = 4;   // up to this many right-hand sides use the single-vector kernels
constexpr int BK = 32;          // pivot block of the panel substitutions

// frequencies solved together: launch slot z solves frequency f[z] on buffer slot s[z]
struct Sel {
    int n;
    int f[kMaxBatch];
    int s[kMaxBatch];
};

template <typename T>
__device__ __forceinline__ cx<T> ld(const cx<T> *p, bool cj) {
    cx<T> v = *p;
    if (cj) v.y = -v.y;
    return v;
}

// ---- elimination order <-> user order --------------------------------------
// User panels are column-major n x nrhs (one shot per column); internal panels
// are row-major in elimination order.  The panel kernels move 64 x 32 tiles
// through LDS so both sides are coalesced: columns are read/written along the
// user index k, internal rows are contiguous in the right-hand sides.
constexpr int PT_K = 64, PT_R = 32;

template <typename T, typename TIO>
__global__ void __launch_bounds__(256) k_permute_in_slots(const cx<TIO> *__restrict__ B, cx<T> *__restrict__ Xp,
                                                          const int *__restrict__ iperm, int n, int nrhs, Sel sel) {
    __shared__ cx<T> S[PT_R][PT_K + 1];
    const int ktiles = (n + PT_K - 1) / PT_K;
    const int k0 = (blockIdx.x % ktiles) * PT_K, r0 = (blockIdx.x / ktiles) * PT_R;
    const int64_t panel = (int64_t)n * nrhs;
    B += sel.s[blockIdx.y] * panel;
    Xp += (int64_t)blockIdx.y * panel;
    for (int e = threadIdx.x; e < PT_K * PT_R; e += 256) {
        int kk = e % PT_K, rr = e / PT_K;
        if (k0 + kk < n && r0 + rr < nrhs) {
            cx<TIO> v = B[k0 + kk + (int64_t)(r0 + rr) * n];
            S[rr][kk] = {(T)v.x, (T)v.y};
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < PT_K * PT_R; e += 256) {
        int rr = e % PT_R, kk = e / PT_R;
        if (k0 + kk < n && r0 + rr < nrhs) Xp[(int64_t)iperm[k0 + kk] * nrhs + r0 + rr] = S[rr][kk];
    }
}

// iout: position of user row k (the inverse elimination order)
template <typename T, typename TIO>
__global__ void __launch_bounds__(256) k_permute_out_slots(const cx<T> *__restrict__ Xp, cx<TIO> *__restrict__ X,
                                                           const int *__restrict__ iout, int64_t map_stride, int n,
                                                           int nrhs, Sel sel) {
    __shared__ cx<T> S[PT_R][PT_K + 1];
    const int ktiles = (n + PT_K - 1) / PT_K;
    const int k0 = (blockIdx.x % ktiles) * PT_K, r0 = (blockIdx.x / ktiles) * PT_R;
    const int64_t panel = (int64_t)n * nrhs;
    X += sel.s[blockIdx.y] * panel;
    Xp += (int64_t)blockIdx.y * panel;
    iout += sel.f[blockIdx.y] * map_stride;
    for (int e = threadIdx.x; e < PT_K * PT_R; e += 256) {
        int rr = e % PT_R, kk = e / PT_R;
        if (k0 + kk < n && r0 + rr < nrhs) S[rr][kk] = Xp[(int64_t)iout[k0 + kk] * nrhs + r0 + rr];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < PT_K * PT_R; e += 256) {
        int kk = e % PT_K, rr = e / PT_K;
        if (k0 + kk < n && r0 + rr < nrhs) {
            cx<T> v = S[rr][kk];
            X[k0 + kk + (int64_t)(r0 + rr) * n] = {(TIO)v.x, (TIO)v.y};
        }
    }
}

// single vector: Xp[pos] = B[perm[pos]] and X[k] = Xp[iperm[k]]
template <typename T, typename TIO>
__global__ void __launch_bounds__(256) k_vec_in(const cx<TIO> *__restrict__ B, cx<T> *__restrict__ Xp,
                                                const int *__restrict__ perm, int n) {
    for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < n; pos += gridDim.x * blockDim.x) {
        cx<TIO> v = B[perm[pos]];
        Xp[pos] = {(T)v.x, (T)v.y};
    }
}
template <typename T, typename TIO>
__global__ void __launch_bounds__(256) k_vec_out(const cx<T> *__restrict__ Xp, cx<TIO> *__restrict__ X,
                                                 const int *__restrict__ iout, int n) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        cx<T> v = Xp[iout[k]];
        X[k] = {(TIO)v.x, (TIO)v.y};
    }
}

// value of local front row i before the row swaps: own pivot row of X plus the
// children's contribution rows mapped onto it
template <typename T>
__device__ __forceinline__ cx<T> front_row(const FrontDev &F, int i, int r, int nrhs, const cx<T> *__restrict__ X,
                                           const cx<T> *__restrict__ W, const FrontDev *__restrict__ fronts,
                                           const int *__restrict__ children, const int *__restrict__ inv) {
    cx<T> v = {T(0), T(0)};
    if (i < F.p) v = X[(int64_t)(F.pos0 + i) * nrhs + r];
    for (int c = 0; c < F.nchild; c++) {
        const FrontDev C = fronts[children[F.child_off + c]];
        int a = inv[C.inv_off + i];
        if (a >= 0) v = cadd(v, W[(C.fv_off + C.p + a) * nrhs + r]);
    }
    return v;
}

// forward gather: W[f] rows; for M solves the pivot rows are taken through the
// front's row permutation (P applied on the fly)
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_gather(const cx<T> *__restrict__ X, cx<T> *__restrict__ W,
                                                    const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                    const int *__restrict__ children, const int *__restrict__ inv,
                                                    const int *__restrict__ rowperm, int n, int64_t wstride, int nrhs,
                                                    Sel sel) {
    const FrontDev F = fronts[list[blockIdx.x]];
    const int64_t total = (int64_t)F.m * nrhs;
    X += (int64_t)sel.s[blockIdx.z] * n * nrhs;
    W += (int64_t)sel.s[blockIdx.z] * wstride;
    if (rowperm) rowperm += (int64_t)sel.f[blockIdx.z] * n;
    for (int64_t e = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.y * blockDim.x) {
        int i = (int)(e / nrhs), r = (int)(e - (int64_t)i * nrhs);
        int src = (rowperm && i < F.p) ? rowperm[F.pos0 + i] : i;
        W[(F.fv_off + i) * nrhs + r] = front_row(F, src, r, nrhs, X, W, fronts, children, inv);
    }
}

// ---- panel substitutions on MFMA -------------------------------------------
// Shared-memory layout of the block kernels (dynamic): A tile [BK][64] re/im,
// right-hand-side block [BK][64] re/im, diagonal block [BK][BK+1] complex.
template <typename T>
struct BlkSmem {
    T *Ar, *Ai, *Yr, *Yi;
    cx<T> *D;
    __device__ BlkSmem(unsigned char *raw) {
        Ar = reinterpret_cast<T *>(raw);
        Ai = Ar + BK * 64;
        Yr = Ai + BK * 64;
        Yi = Yr + BK * 64;
        D = reinterpret_cast<cx<T> *>(Yi + BK * 64);
    }
};
template <typename T>
constexpr size_t blk_smem_bytes() {
    return sizeof(T) * 4 * BK * 64 + sizeof(cx<T>) * BK * (BK + 1);
}

// acc = A[64 x kb] * Y[kb x 64] from LDS (4 waves in 2x2, 32x32 each)
template <typename T>
__device__ __forceinline__ void mma_tile(const BlkSmem<T> &s, int kb, typename mfma_t<T>::acc (&cr)[2][2],
                                         typename mfma_t<T>::acc (&ci)[2][2]) {
    typedef typename mfma_t<T>::acc acc_t;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int lr = lane & 15, lk = lane >> 4;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            cr[a][b] = acc_t{0, 0, 0, 0};
            ci[a][b] = acc_t{0, 0, 0, 0};
        }
    const int kend = (kb + 3) & ~3;
    for (int kk = 0; kk < kend; kk += 4) {
        const int k = kk + lk;
        T ar[2], ai[2], br[2], bi[2];
        for (int a = 0; a < 2; a++) {
            ar[a] = s.Ar[k * 64 + wr + a * 16 + lr];
            ai[a] = s.Ai[k * 64 + wr + a * 16 + lr];
        }
        for (int b = 0; b < 2; b++) {
            br[b] = s.Yr[k * 64 + wc + b * 16 + lr];
            bi[b] = s.Yi[k * 64 + wc + b * 16 + lr];
        }
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) cmma<T>(ar[a], ai[a], br[b], bi[b], cr[a][b], ci[a][b]);
    }
}

// rows [row0, row0+64) of w (below rowend) -= acc, right-hand sides [r0, r0+64)
template <typename T>
__device__ __forceinline__ void sub_tile(cx<T> *w, int nrhs, int row0, int rowend, int r0,
                                         typename mfma_t<T>::acc (&cr)[2][2], typename mfma_t<T>::acc (&ci)[2][2]) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
    const int lr = lane & 15, lk = lane >> 4;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int v = 0; v < 4; v++) {
                const int row = row0 + wr + a * 16 + mfma_t<T>::row(lk, v), col = r0 + wc + b * 16 + lr;
                if (row < rowend && col < nrhs) {
                    cx<T> &g = w[(int64_t)row * nrhs + col];
                    g.x -= cr[a][b][v];
                    g.y -= ci[a][b][v];
                }
            }
}

// forward block b0 of every front of a level: y_b = Lop_bb^-1 w_b (written to
// X by the row-tile-0 workgroup), then rows [b0+nb, m) of W -= Lop[:, b] y_b.
// grid: (row tiles, fronts, right-hand-side tiles)
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_blk(const cx<T> *__restrict__ arena, int64_t fstride,
                                                 const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                 cx<T> *__restrict__ X, cx<T> *__restrict__ W, int n, int64_t wstride,
                                                 int nrhs, int rt, int b0, int trans, int cj, Sel sel) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    BlkSmem<T> s(smem_raw);
    const FrontDev F = fronts[list[blockIdx.y]];
    if (F.p <= b0) return;
    const int m = F.m, nb = min(BK, F.p - b0), tid = threadIdx.x;
    const int row0 = b0 + nb + blockIdx.x * 64;
    if (blockIdx.x > 0 && row0 >= m) return;
    const int z = blockIdx.z / rt, r0 = (blockIdx.z - z * rt) * 64;
    const cx<T> *G = arena + sel.f[z] * fstride + F.buf_off;
    X += (int64_t)sel.s[z] * n * nrhs;
    cx<T> *w = W + sel.s[z] * wstride + F.fv_off * nrhs;
    auto Lop = [&](int i, int t) { return trans ? ld(G + t + (int64_t)i * m, cj) : ld(G + i + (int64_t)t * m, cj); };
    for (int e = tid; e < nb * nb; e += 256) {
        int i = e % nb, t = e / nb;
        s.D[i * (BK + 1) + t] = Lop(b0 + i, b0 + t);
    }
    for (int e = tid; e < BK * 64; e += 256) {
        int k = e >> 6, j = e & 63;
        cx<T> v = {T(0), T(0)};
        if (k < nb && r0 + j < nrhs) v = w[(int64_t)(b0 + k) * nrhs + r0 + j];
        s.Yr[e] = v.x;
        s.Yi[e] = v.y;
    }
    __syncthreads();
    if (tid < 64) {   // diagonal triangle, one right-hand side per thread
        const int j = tid;
        for (int k = 0; k < nb; k++) {
            cx<T> acc = {s.Yr[k * 64 + j], s.Yi[k * 64 + j]};
            for (int t = 0; t < k; t++)
                acc = cmsub(acc, s.D[k * (BK + 1) + t], cx<T>{s.Yr[t * 64 + j], s.Yi[t * 64 + j]});
            if (trans) acc = cdiv(acc, s.D[k * (BK + 1) + k]);
            s.Yr[k * 64 + j] = acc.x;
            s.Yi[k * 64 + j] = acc.y;
        }
    }
    __syncthreads();
    if (blockIdx.x == 0)
        for (int e = tid; e < nb * 64; e += 256) {
            int k = e >> 6, j = e & 63;
            if (r0 + j < nrhs) X[(int64_t)(F.pos0 + b0 + k) * nrhs + r0 + j] = {s.Yr[e], s.Yi[e]};
        }
    if (row0 >= m) return;
    for (int e = tid; e < BK * 64; e += 256) {
        int k, i;
        if (!trans) { i = e & 63; k = e >> 6; }        // column-major L: coalesced along rows
        else { k = e % BK; i = e / BK; }               // U^T: coalesced along k
        cx<T> v = {T(0), T(0)};
        if (k < nb && row0 + i < m) v = Lop(row0 + i, b0 + k);
        s.Ar[k * 64 + i] = v.x;
        s.Ai[k * 64 + i] = v.y;
    }
    __syncthreads();
    typename mfma_t<T>::acc cr[2][2], ci[2][2];
    mma_tile<T>(s, nb, cr, ci);
    sub_tile<T>(w, nrhs, row0, m, r0, cr, ci);
}

// backward, ancestors: W[f][0:p] = y1 - Uop12 x2, x2 gathered from X at the
// contribution positions.  grid: (row tiles over p, fronts, rhs tiles)
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_contrib(const cx<T> *__restrict__ arena, int64_t fstride,
                                                     const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                     const int *__restrict__ idx, const cx<T> *__restrict__ X,
                                                     cx<T> *__restrict__ W, int n, int64_t wstride, int nrhs, int rt,
                                                     int trans, int cj, Sel sel) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    BlkSmem<T> s(smem_raw);
    const FrontDev F = fronts[list[blockIdx.y]];
    const int m = F.m, p = F.p, tid = threadIdx.x;
    const int row0 = blockIdx.x * 64;
    if (row0 >= p) return;
    const int z = blockIdx.z / rt, r0 = (blockIdx.z - z * rt) * 64;
    const cx<T> *G = arena + sel.f[z] * fstride + F.buf_off;
    X += (int64_t)sel.s[z] * n * nrhs;
    cx<T> *w = W + sel.s[z] * wstride + F.fv_off * nrhs;
    auto Uop = [&](int i, int t) { return trans ? ld(G + t + (int64_t)i * m, cj) : ld(G + i + (int64_t)t * m, cj); };
    typedef typename mfma_t<T>::acc acc_t;
    acc_t cr[2][2], ci[2][2], tr[2][2], ti[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            cr[a][b] = acc_t{0, 0, 0, 0};
            ci[a][b] = acc_t{0, 0, 0, 0};
        }
    for (int s0 = p; s0 < m; s0 += BK) {
        const int kb = min(BK, m - s0);
        __syncthreads();
        for (int e = tid; e < BK * 64; e += 256) {
            int k, i;
            if (!trans) { i = e & 63; k = e >> 6; }
            else { k = e % BK; i = e / BK; }
            cx<T> v = {T(0), T(0)};
            if (k < kb && row0 + i < p) v = Uop(row0 + i, s0 + k);
            s.Ar[k * 64 + i] = v.x;
            s.Ai[k * 64 + i] = v.y;
        }
        for (int e = tid; e < BK * 64; e += 256) {
            int k = e >> 6, j = e & 63;
            cx<T> v = {T(0), T(0)};
            if (k < kb && r0 + j < nrhs) v = X[(int64_t)idx[F.idx_off + s0 + k] * nrhs + r0 + j];
            s.Yr[e] = v.x;
            s.Yi[e] = v.y;
        }
        __syncthreads();
        mma_tile<T>(s, kb, tr, ti);
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                cr[a][b] += tr[a][b];
                ci[a][b] += ti[a][b];
            }
    }
    // w[0:p] = y1 (from X) - accumulated product
    for (int e = tid; e < 64 * 64; e += 256) {
        int i = e >> 6, j = e & 63;
        if (row0 + i < p && r0 + j < nrhs)
            w[(int64_t)(row0 + i) * nrhs + r0 + j] = X[(int64_t)(F.pos0 + row0 + i) * nrhs + r0 + j];
    }
    __syncthreads();
    sub_tile<T>(w, nrhs, row0, p, r0, cr, ci);
}

// backward block b0 (blocks from the bottom): x_b = Uop_bb^-1 z_b (to X by the
// row-tile-0 workgroup), then rows [0, b0) of W -= Uop[:, b] x_b
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_blk(const cx<T> *__restrict__ arena, int64_t fstride,
                                                 const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                 cx<T> *__restrict__ X, cx<T> *__restrict__ W, int n, int64_t wstride,
                                                 int nrhs, int rt, int b0, int trans, int cj, Sel sel) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    BlkSmem<T> s(smem_raw);
    const FrontDev F = fronts[list[blockIdx.y]];
    if (F.p <= b0) return;
    const int m = F.m, nb = min(BK, F.p - b0), tid = threadIdx.x;
    const int row0 = blockIdx.x * 64;
    if (blockIdx.x > 0 && row0 >= b0) return;
    const int z = blockIdx.z / rt, r0 = (blockIdx.z - z * rt) * 64;
    const cx<T> *G = arena + sel.f[z] * fstride + F.buf_off;
    X += (int64_t)sel.s[z] * n * nrhs;
    cx<T> *w = W + sel.s[z] * wstride + F.fv_off * nrhs;
    auto Uop = [&](int i, int t) { return trans ? ld(G + t + (int64_t)i * m, cj) : ld(G + i + (int64_t)t * m, cj); };
    for (int e = tid; e < nb * nb; e += 256) {
        int i = e % nb, t = e / nb;
        s.D[i * (BK + 1) + t] = Uop(b0 + i, b0 + t);
    }
    for (int e = tid; e < BK * 64; e += 256) {
        int k = e >> 6, j = e & 63;
        cx<T> v = {T(0), T(0)};
        if (k < nb && r0 + j < nrhs) v = w[(int64_t)(b0 + k) * nrhs + r0 + j];
        s.Yr[e] = v.x;
        s.Yi[e] = v.y;
    }
    __syncthreads();
    if (tid < 64) {
        const int j = tid;
        for (int k = nb - 1; k >= 0; k--) {
            cx<T> acc = {s.Yr[k * 64 + j], s.Yi[k * 64 + j]};
            for (int t = k + 1; t < nb; t++)
                acc = cmsub(acc, s.D[k * (BK + 1) + t], cx<T>{s.Yr[t * 64 + j], s.Yi[t * 64 + j]});
            if (!trans) acc = cdiv(acc, s.D[k * (BK + 1) + k]);
            s.Yr[k * 64 + j] = acc.x;
            s.Yi[k * 64 + j] = acc.y;
        }
    }
    __syncthreads();
    if (blockIdx.x == 0)
        for (int e = tid; e < nb * 64; e += 256) {
            int k = e >> 6, j = e & 63;
            if (r0 + j < nrhs) X[(int64_t)(F.pos0 + b0 + k) * nrhs + r0 + j] = {s.Yr[e], s.Yi[e]};
        }
    if (row0 >= b0) return;
    for (int e = tid; e < BK * 64; e += 256) {
        int k, i;
        if (!trans) { i = e & 63; k = e >> 6; }
        else { k = e % BK; i = e / BK; }
        cx<T> v = {T(0), T(0)};
        if (k < nb && row0 + i < b0) v = Uop(row0 + i, b0 + k);
        s.Ar[k * 64 + i] = v.x;
        s.Ai[k * 64 + i] = v.y;
    }
    __syncthreads();
    typename mfma_t<T>::acc cr[2][2], ci[2][2];
    mma_tile<T>(s, nb, cr, ci);
    sub_tile<T>(w, nrhs, row0, b0, r0, cr, ci);
}

// transposed solves, after a level's backward blocks: x = P_f^T w for every
// front of the level, through W rows [0, p) as scratch (copy, then scatter)
template <typename T>
__global__ void __launch_bounds__(256) k_piv_copy(const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                  const cx<T> *__restrict__ X, cx<T> *__restrict__ W, int n,
                                                  int64_t wstride, int nrhs, Sel sel) {
    const FrontDev F = fronts[list[blockIdx.x]];
    X += (int64_t)sel.s[blockIdx.z] * n * nrhs;
    W += sel.s[blockIdx.z] * wstride;
    const int64_t total = (int64_t)F.p * nrhs;
    for (int64_t e = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.y * blockDim.x)
        W[F.fv_off * nrhs + e] = X[(int64_t)F.pos0 * nrhs + e];
}
template <typename T>
__global__ void __launch_bounds__(256) k_piv_unswap(const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                    const int *__restrict__ rowperm, cx<T> *__restrict__ X,
                                                    const cx<T> *__restrict__ W, int n, int64_t wstride, int nrhs,
                                                    Sel sel) {
    const FrontDev F = fronts[list[blockIdx.x]];
    X += (int64_t)sel.s[blockIdx.z] * n * nrhs;
    W += sel.s[blockIdx.z] * wstride;
    rowperm += (int64_t)sel.f[blockIdx.z] * n;
    const int64_t total = (int64_t)F.p * nrhs;
    for (int64_t e = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.y * blockDim.x) {
        int i = (int)(e / nrhs), r = (int)(e - (int64_t)i * nrhs);
        X[(int64_t)(F.pos0 + rowperm[F.pos0 + i]) * nrhs + r] = W[(F.fv_off + i) * nrhs + r];
    }
}

// ---- single right-hand side: one workgroup per front, work vector in LDS ----
template <typename T>
__device__ __forceinline__ cx<T> shfl_c(cx<T> v, int src) {
    return {__shfl(v.x, src), __shfl(v.y, src)};
}

// forward sweep of one front for one right-hand side: gather (row-permuted for
// M solves), blocked lower solve (32-row diagonal blocks solved by wave 0 with
// shuffles, block updates coalesced over rows), pivot rows to X and
// contribution rows to W for the parent
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_vec(const cx<T> *__restrict__ arena, const FrontDev *__restrict__ fronts,
                                                 const int *__restrict__ list, const int *__restrict__ children,
                                                 const int *__restrict__ inv, const int *__restrict__ rowperm,
                                                 cx<T> *__restrict__ X, cx<T> *__restrict__ W, int trans, int cj) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *sw = reinterpret_cast<cx<T> *>(smem_raw);
    const FrontDev F = fronts[list[blockIdx.x]];
    const int m = F.m, p = F.p, tid = threadIdx.x;
    const cx<T> *G = arena + F.buf_off;
    auto Lop = [&](int i, int t) { return trans ? ld(G + t + (int64_t)i * m, cj) : ld(G + i + (int64_t)t * m, cj); };
    for (int i = tid; i < m; i += 256) {
        int src = (!trans && i < p) ? rowperm[F.pos0 + i] : i;
        sw[i] = front_row(F, src, 0, 1, X, W, fronts, children, inv);
    }
    __syncthreads();
    for (int tb = 0; tb < p; tb += 32) {
        const int nb = min(32, p - tb);
        if (tid < 64) {
            const int k = tid;
            cx<T> y = k < nb ? sw[tb + k] : cx<T>{T(0), T(0)};
            for (int t = 0; t < nb; t++) {
                if (trans && k == t) y = cdiv(y, Lop(tb + t, tb + t));
                const cx<T> yt = shfl_c(y, t);
                if (k > t && k < nb) y = cmsub(y, Lop(tb + k, tb + t), yt);
            }
            if (k < nb) sw[tb + k] = y;
        }
        __syncthreads();
        for (int i = tb + nb + tid; i < m; i += 256) {
            cx<T> acc = sw[i];
            for (int t = 0; t < nb; t++) acc = cmsub(acc, Lop(i, tb + t), sw[tb + t]);
            sw[i] = acc;
        }
        __syncthreads();
    }
    for (int i = tid; i < m; i += 256) {
        if (i < p) X[F.pos0 + i] = sw[i];
        else W[F.fv_off + i] = sw[i];
    }
}

// backward sweep of one front for one right-hand side
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_vec(const cx<T> *__restrict__ arena, const FrontDev *__restrict__ fronts,
                                                 const int *__restrict__ list, const int *__restrict__ idx,
                                                 const int *__restrict__ rowperm, cx<T> *__restrict__ X, int trans,
                                                 int cj) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *sw = reinterpret_cast<cx<T> *>(smem_raw);
    const FrontDev F = fronts[list[blockIdx.x]];
    const int m = F.m, p = F.p, tid = threadIdx.x;
    const cx<T> *G = arena + F.buf_off;
    auto Uop = [&](int i, int t) { return trans ? ld(G + t + (int64_t)i * m, cj) : ld(G + i + (int64_t)t * m, cj); };
    for (int i = tid; i < m; i += 256) sw[i] = i < p ? X[F.pos0 + i] : X[idx[F.idx_off + i]];
    __syncthreads();
    for (int i = tid; i < p; i += 256) {
        cx<T> acc = sw[i];
        for (int s = p; s < m; s++) acc = cmsub(acc, Uop(i, s), sw[s]);
        sw[i] = acc;
    }
    __syncthreads();
    for (int tb = ((p - 1) / 32) * 32; tb >= 0; tb -= 32) {
        const int nb = min(32, p - tb);
        if (tid < 64) {
            const int k = tid;
            cx<T> x = k < nb ? sw[tb + k] : cx<T>{T(0), T(0)};
            for (int t = nb - 1; t >= 0; t--) {
                if (!trans && k == t) x = cdiv(x, Uop(tb + t, tb + t));
                const cx<T> xt = shfl_c(x, t);
                if (k < t) x = cmsub(x, Uop(tb + k, tb + t), xt);
            }
            if (k < nb) sw[tb + k] = x;
        }
        __syncthreads();
        for (int i = tid; i < tb; i += 256) {
            cx<T> acc = sw[i];
            for (int t = 0; t < nb; t++) acc = cmsub(acc, Uop(i, tb + t), sw[tb + t]);
            sw[i] = acc;
        }
        __syncthreads();
    }
    // transposed solves: x = P_f^T w before any descendant reads it (descendant
    // L21 rows are in this front's original row order)
    for (int i = tid; i < p; i += 256) X[F.pos0 + (trans ? rowperm[F.pos0 + i] : i)] = sw[i];
}

// ---- FP64 residuals ------------------------------------------------------------
// r = b - op(M) x, column-major n x nrhs panels; one thread per row keeps the
// row's (at most nine) entries in registers and walks the right-hand sides.
// The pattern is symmetric: row `row` of M has the columns listed in CSC
// column `row` (values through tq for M, directly for M^T / M^H).
__global__ void __launch_bounds__(256) k_residual(const cx<double> *__restrict__ vals, const int *__restrict__ colptr,
                                                  const int *__restrict__ rowind, const int64_t *__restrict__ tq,
                                                  int n, int nrhs, int sys, int64_t nnz,
                                                  const cx<double> *__restrict__ B, const cx<double> *__restrict__ X,
                                                  cx<double> *__restrict__ R, Sel sel) {
    vals += sel.f[blockIdx.y] * nnz;
    const int64_t off = (int64_t)sel.s[blockIdx.y] * n * nrhs;
    B += off;
    X += off;
    R += off;
    for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < n; row += gridDim.x * blockDim.x) {
        const int q0 = colptr[row], cnt = colptr[row + 1] - q0;
        cx<double> a[9];
        int c[9];
#pragma unroll
        for (int t = 0; t < 9; t++) {
            if (t < cnt) {
                const int q = q0 + t;
                a[t] = sys == 0 ? vals[tq[q]] : vals[q];
                if (sys == 1) a[t].y = -a[t].y;
                c[t] = rowind[q];
            } else {
                a[t] = {0.0, 0.0};
                c[t] = row;
            }
        }
        for (int r = 0; r < nrhs; r++) {
            const cx<double> *x = X + (int64_t)r * n;
            cx<double> acc = B[(int64_t)r * n + row];
#pragma unroll
            for (int t = 0; t < 9; t++) acc = cmsub(acc, a[t], x[c[t]]);
            R[(int64_t)r * n + row] = acc;
        }
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

// squared 2-norm of every column of the nslot panels: out[slot * nrhs + r]
__global__ void __launch_bounds__(256) k_colnorm2(const cx<double> *__restrict__ V, int n, int nrhs,
                                                  double *__restrict__ out) {
    __shared__ double s[256];
    const cx<double> *v = V + ((int64_t)blockIdx.y * nrhs + blockIdx.x) * n;
    double acc = 0;
    for (int i = threadIdx.x; i < n; i += 256) acc += v[i].x * v[i].x + v[i].y * v[i].y;
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(int64_t)blockIdx.y * nrhs + blockIdx.x] = s[0];
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

template <typename T>
void set_smem_attrs() {
    static bool done = false;
    if (done) return;
    const int big = 160 * 1024;
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_fwd_vec<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, big));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bwd_vec<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, big));
    const int blk = (int)blk_smem_bytes<T>();
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_fwd_blk<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, blk));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bwd_blk<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, blk));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bwd_contrib<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, blk));
    done = true;
}

}  // namespace

Sel make_sel(const std::vector<int> &freqs, const std::vector<int> &slots) {
    if (freqs.empty() || freqs.size() > (size_t)kMaxBatch || slots.size() != freqs.size())
        throw std::invalid_argument("1 to 64 frequencies per solve batch");
    Sel sel;
    sel.n = (int)freqs.size();
    for (int z = 0; z < sel.n; z++) {
        sel.f[z] = freqs[z];
        sel.s[z] = slots[z];
    }
    return sel;
}

template <typename T, typename TIO>
void solve_batch(const Factors<T> &F, const std::vector<int> &freqs, const std::vector<int> &slots, int nslot_buf,
                 int sys, int nrhs, const cx<TIO> *B, cx<TIO> *X, hipStream_t st) {
    const DevPlan &dp = *F.plan;
    const Plan &pl = *dp.host;
    const int n = pl.n;
    for (int f : freqs)
        if (f < 0 || f >= F.nfreq) throw std::invalid_argument("frequency index out of range");
    if (sys < 0 || sys > 2) throw std::invalid_argument("invalid system");
    set_smem_attrs<T>();
    const Sel sel = make_sel(freqs, slots);
    const int ns = sel.n;
    const bool trans = sys != 0, cj = sys == 1;
    const int *iout = dp.iperm;
    const int64_t mstride = 0;
    const double es = sizeof(cx<T>);
    if (nrhs <= kVecMaxRhs) {
        // few right-hand sides: one pass per frequency and column through the single-vector kernels
        cx<T> *Xp = (cx<T> *)ws_x.get(sizeof(cx<T>) * (size_t)n);
        cx<T> *Wp = (cx<T> *)ws_w.get(sizeof(cx<T>) * (size_t)pl.fv_rows);
        for (int z = 0; z < ns; z++) {
            const int f = sel.f[z];
            const cx<T> *arena = F.arena + f * F.fstride;
            const int *rowperm = F.rowperm + (int64_t)f * n;
            for (int r = 0; r < nrhs; r++) {
                const int64_t off = ((int64_t)sel.s[z] * nrhs + r) * n;
                {
                    ProfScope ps(K_PERMUTE, st, 0.0, (double)n * (sizeof(cx<TIO>) + es + 4));
                    hipLaunchKernelGGL((k_vec_in<T, TIO>), dim3(grid_for(n)), dim3(256), 0, st, B + off, Xp, dp.perm,
                                       n);
                }
                for (int l = 0; l < pl.nlevels; l++) {
                    if (dp.n_all[l] == 0) continue;
                    const DevPlan::LevelCounts &c = dp.counts[l];
                    ProfScope ps(K_FWD_VEC, st, 8.0 * c.tri_elems, (c.tri_elems + 2.0 * c.sum_m) * es);
                    size_t lds = sizeof(cx<T>) * (size_t)dp.mmax_all[l];
                    hipLaunchKernelGGL(k_fwd_vec<T>, dim3(dp.n_all[l]), dim3(256), lds, st, arena, dp.fronts,
                                       dp.lv_all[l], dp.children, dp.inv, rowperm, Xp, Wp, (int)trans, (int)cj);
                }
                for (int l = pl.nlevels - 1; l >= 0; l--) {
                    if (dp.n_all[l] == 0) continue;
                    const DevPlan::LevelCounts &c = dp.counts[l];
                    ProfScope ps(K_BWD_VEC, st, 8.0 * c.tri_elems, (c.tri_elems + 2.0 * c.sum_m) * es);
                    size_t lds = sizeof(cx<T>) * (size_t)dp.mmax_all[l];
                    hipLaunchKernelGGL(k_bwd_vec<T>, dim3(dp.n_all[l]), dim3(256), lds, st, arena, dp.fronts,
                                       dp.lv_all[l], dp.idx, rowperm, Xp, (int)trans, (int)cj);
                }
                {
                    ProfScope ps(K_PERMUTE, st, 0.0, (double)n * (sizeof(cx<TIO>) + es + 4));
                    hipLaunchKernelGGL((k_vec_out<T, TIO>), dim3(grid_for(n)), dim3(256), 0, st, Xp, X + off,
                                       iout + f * mstride, n);
                }
            }
        }
        GFDFD_HIP(hipGetLastError());
        return;
    }
    // internal panels are indexed by launch slot: slot z of Xp/Wp serves frequency sel.f[z]
    Sel in = sel;
    for (int z = 0; z < ns; z++) in.s[z] = z;
    const int64_t wstride = pl.fv_rows * (int64_t)nrhs;
    cx<T> *Xp = (cx<T> *)ws_x.get(sizeof(cx<T>) * (size_t)n * nrhs * ns);
    cx<T> *Wp = (cx<T> *)ws_w.get(sizeof(cx<T>) * (size_t)wstride * ns);
    const double nr = nrhs, nz = ns;
    const size_t blk = blk_smem_bytes<T>();
    const int rt = (nrhs + 63) / 64;
    (void)nslot_buf;
    {
        // B slot sel.s[z] -> internal slot z
        ProfScope ps(K_PERMUTE, st, 0.0, nz * ((double)n * nr * (sizeof(cx<TIO>) + es) + 4.0 * n));
        const int tiles = ((n + PT_K - 1) / PT_K) * ((nrhs + PT_R - 1) / PT_R);
        hipLaunchKernelGGL((k_permute_in_slots<T, TIO>), dim3(tiles, ns), dim3(256), 0, st, B, Xp, dp.iperm, n, nrhs,
                           sel);
    }
    for (int l = 0; l < pl.nlevels; l++) {
        if (dp.n_all[l] == 0) continue;
        const DevPlan::LevelCounts &c = dp.counts[l];
        {
            ProfScope ps(K_FWD_GATHER, st, 0.0, nz * 2.0 * c.sum_m * nr * es);
            int64_t work = (int64_t)dp.mmax_all[l] * nrhs;
            int chunks = (int)std::min<int64_t>(std::max<int64_t>(1, work / 1024), 256);
            hipLaunchKernelGGL(k_fwd_gather<T>, dim3(dp.n_all[l], chunks, ns), dim3(256), 0, st, Xp, Wp, dp.fronts,
                               dp.lv_all[l], dp.children, dp.inv, trans ? nullptr : F.rowperm, n, wstride, nrhs, in);
        }
        const int pmax = dp.pmax_all[l];
        ProfScope ps(K_FWD, st, nz * 8.0 * nr * c.tri_elems, nz * (c.tri_elems + 2.0 * c.sum_m * nr) * es);
        for (int b0 = 0; b0 < pmax; b0 += BK) {
            const int tiles = std::max(1, (dp.mmax_all[l] - b0 - 1) / 64 + 1);
            hipLaunchKernelGGL(k_fwd_blk<T>, dim3(tiles, dp.n_all[l], rt * ns), dim3(256), blk, st, F.arena, F.fstride,
                               dp.fronts, dp.lv_all[l], Xp, Wp, n, wstride, nrhs, rt, b0, (int)trans, (int)cj, in);
        }
    }
    for (int l = pl.nlevels - 1; l >= 0; l--) {
        if (dp.n_all[l] == 0) continue;
        const DevPlan::LevelCounts &c = dp.counts[l];
        const int pmax = dp.pmax_all[l];
        ProfScope ps(K_BWD, st, nz * 8.0 * nr * c.tri_elems, nz * (c.tri_elems + 2.0 * c.sum_m * nr) * es);
        hipLaunchKernelGGL(k_bwd_contrib<T>, dim3((pmax + 63) / 64, dp.n_all[l], rt * ns), dim3(256), blk, st, F.arena,
                           F.fstride, dp.fronts, dp.lv_all[l], dp.idx, Xp, Wp, n, wstride, nrhs, rt, (int)trans,
                           (int)cj, in);
        for (int b0 = ((pmax - 1) / BK) * BK; b0 >= 0; b0 -= BK) {
            const int tiles = std::max(1, (b0 + 63) / 64);
            hipLaunchKernelGGL(k_bwd_blk<T>, dim3(tiles, dp.n_all[l], rt * ns), dim3(256), blk, st, F.arena,
                               F.fstride, dp.fronts, dp.lv_all[l], Xp, Wp, n, wstride, nrhs, rt, b0, (int)trans,
                               (int)cj, in);
        }
        if (trans) {
            const int chunks = (int)std::min<int64_t>(std::max<int64_t>(1, (int64_t)pmax * nrhs / 1024), 256);
            hipLaunchKernelGGL(k_piv_copy<T>, dim3(dp.n_all[l], chunks, ns), dim3(256), 0, st, dp.fronts,
                               dp.lv_all[l], Xp, Wp, n, wstride, nrhs, in);
            hipLaunchKernelGGL(k_piv_unswap<T>, dim3(dp.n_all[l], chunks, ns), dim3(256), 0, st, dp.fronts,
                               dp.lv_all[l], F.rowperm, Xp, Wp, n, wstride, nrhs, in);
        }
    }
    {
        ProfScope ps(K_PERMUTE, st, 0.0, nz * ((double)n * nr * (sizeof(cx<TIO>) + es) + 4.0 * n));
        const int tiles = ((n + PT_K - 1) / PT_K) * ((nrhs + PT_R - 1) / PT_R);
        hipLaunchKernelGGL((k_permute_out_slots<T, TIO>), dim3(tiles, ns), dim3(256), 0, st, Xp, X, iout, mstride, n,
                           nrhs, sel);
    }
    GFDFD_HIP(hipGetLastError());
}

template <typename T, typename TIO>
void solve_panel(const Factors<T> &F, int ifreq, int sys, int nrhs, const cx<TIO> *B, cx<TIO> *X, hipStream_t st) {
    solve_batch<T, TIO>(F, {ifreq}, {0}, 1, sys, nrhs, B, X, st);
}

template void solve_panel<double, double>(const Factors<double> &, int, int, int, const cx<double> *, cx<double> *,
                                          hipStream_t);
template void solve_panel<float, double>(const Factors<float> &, int, int, int, const cx<double> *, cx<double> *,
                                         hipStream_t);
template void solve_panel<float, float>(const Factors<float> &, int, int, int, const cx<float> *, cx<float> *,
                                        hipStream_t);
template void solve_batch<double, double>(const Factors<double> &, const std::vector<int> &, const std::vector<int> &,
                                          int, int, int, const cx<double> *, cx<double> *, hipStream_t);
template void solve_batch<float, double>(const Factors<float> &, const std::vector<int> &, const std::vector<int> &,
                                         int, int, int, const cx<double> *, cx<double> *, hipStream_t);

// r_s = b_s - op(M_f) x_s for launch slots z (frequency sel.f[z], buffer slot sel.s[z])
static void residual_sel(const DevPlan &dp, const cx<double> *vals, int sys, int nrhs, const cx<double> *B,
                         const cx<double> *X, cx<double> *R, const Sel &sel, hipStream_t st) {
    if (!dp.tq) throw std::logic_error("residual needs a grid plan");
    const int n = dp.host->n;
    const double nnz = (double)dp.host->nnz;
    ProfScope ps(K_RESIDUAL, st, sel.n * 8.0 * nnz * nrhs, sel.n * (nnz * (16.0 + 12.0) + 3.0 * n * 16.0 * nrhs));
    hipLaunchKernelGGL(k_residual, dim3(grid_for(n, 256, 4096), sel.n), dim3(256), 0, st, vals, dp.colptr, dp.rowind,
                       dp.tq, n, nrhs, sys, dp.host->nnz, B, X, R, sel);
}

void residual_fp64(const DevPlan &dp, const cx<double> *vals, int sys, int nrhs, const cx<double> *B,
                   const cx<double> *X, cx<double> *R, hipStream_t st) {
    residual_sel(dp, vals, sys, nrhs, B, X, R, make_sel({0}, {0}), st);
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
    hipLaunchKernelGGL(k_colnorm2, dim3(nrhs, 1), dim3(256), 0, st, V, n, nrhs, out);
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

// FP64 refinement of nsel frequencies at once; frequencies leave the batch as
// they converge (the correction solves batch the remaining ones).
template <typename T>
std::vector<RefineStats> solve_refined_batch(const Factors<T> &F, const std::vector<int> &freqs, int sys, int nrhs,
                                             const cx<double> *B, cx<double> *X, double tol, int max_iter,
                                             hipStream_t st) {
    const DevPlan &dp = *F.plan;
    const int n = dp.host->n, ns = (int)freqs.size();
    const int64_t panel = (int64_t)n * nrhs;
    cx<double> *R = (cx<double> *)ws_r.get(sizeof(cx<double>) * panel * ns);
    cx<double> *D = (cx<double> *)ws_d.get(sizeof(cx<double>) * panel * ns);
    double *nrm = (double *)ws_n.get(sizeof(double) * 2 * nrhs * ns);
    std::vector<double> h(2 * (size_t)nrhs * ns);
    std::vector<int> slots(ns);
    for (int z = 0; z < ns; z++) slots[z] = z;
    hipLaunchKernelGGL(k_colnorm2, dim3(nrhs, ns), dim3(256), 0, st, B, n, nrhs, nrm + (int64_t)nrhs * ns);
    solve_batch<T, double>(F, freqs, slots, ns, sys, nrhs, B, X, st);
    std::vector<RefineStats> rs(ns);
    std::vector<int> act_f = freqs, act_s = slots;
    for (int it = 0;; it++) {
        const Sel sel = make_sel(act_f, act_s);
        residual_sel(dp, F.vals, sys, nrhs, B, X, R, sel, st);
        hipLaunchKernelGGL(k_colnorm2, dim3(nrhs, ns), dim3(256), 0, st, R, n, nrhs, nrm);
        GFDFD_HIP(hipMemcpyAsync(h.data(), nrm, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
        GFDFD_HIP(hipStreamSynchronize(st));
        std::vector<int> nf, nsl;
        for (size_t z = 0; z < act_s.size(); z++) {
            const int s = act_s[z];
            double worst = 0;
            for (int r = 0; r < nrhs; r++) {
                double b = std::sqrt(h[(size_t)nrhs * ns + (size_t)s * nrhs + r]), rr = std::sqrt(h[(size_t)s * nrhs + r]);
                worst = std::max(worst, b > 0 ? rr / b : rr);
            }
            rs[s].max_rel_residual = worst;
            rs[s].iterations = it;
            if (worst > tol && it < max_iter) {
                nf.push_back(act_f[z]);
                nsl.push_back(s);
            }
        }
        if (nf.empty()) break;
        act_f = nf;
        act_s = nsl;
        solve_batch<T, double>(F, act_f, act_s, ns, sys, nrhs, R, D, st);
        for (int s : act_s) axpy_fp64(panel, D + s * panel, X + s * panel, st);
    }
    return rs;
}

template <typename T>
RefineStats solve_refined(const Factors<T> &F, int ifreq, int sys, int nrhs, const cx<double> *B, cx<double> *X,
                          double tol, int max_iter, hipStream_t st) {
    return solve_refined_batch<T>(F, {ifreq}, sys, nrhs, B, X, tol, max_iter, st)[0];
}

template RefineStats solve_refined<double>(const Factors<double> &, int, int, int, const cx<double> *, cx<double> *,
                                           double, int, hipStream_t);
template RefineStats solve_refined<float>(const Factors<float> &, int, int, int, const cx<double> *, cx<double> *,
                                          double, int, hipStream_t);
template std::vector<RefineStats> solve_refined_batch<double>(const Factors<double> &, const std::vector<int> &, int,
                                                              int, const cx<double> *, cx<double> *, double, int,
                                                              hipStream_t);
template std::vector<RefineStats> solve_refined_batch<float>(const Factors<float> &, const std::vector<int> &, int, int,
                                                             const cx<double> *, cx<double> *, double, int,
                                                             hipStream_t);

}  // namespace gfdfd
