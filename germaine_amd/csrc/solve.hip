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
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <mutex>

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

template <typename T>
__device__ __forceinline__ cx<T> cmac(cx<T> c, cx<T> a, cx<T> b) {   // c + a b
    return cadd(c, cmul(a, b));
}

// ---- elimination order <-> user order --------------------------------------
// User panels are column-major n x nrhs (one shot per column); internal panels
// are row-major in elimination order.  The panel kernels move 64 x 32 tiles
// through LDS so both sides are coalesced: columns are read/written along the
// user index k, internal rows are contiguous in the right-hand sides.
constexpr int PT_K = 64, PT_R = 32;

// Internal panels have a row stride ld >= nrhs (whole 128-byte lines per row);
// the padding columns are zero and stay zero through the sweeps.
template <typename T, typename TIO>
__global__ void __launch_bounds__(256) k_permute_in_slots(const cx<TIO> *__restrict__ B, cx<T> *__restrict__ Xp,
                                                          const int *__restrict__ iperm, int n, int nrhs, int ld,
                                                          Sel sel) {
    __shared__ cx<T> S[PT_R][PT_K + 1];
    const int ktiles = (n + PT_K - 1) / PT_K;
    const int k0 = (blockIdx.x % ktiles) * PT_K, r0 = (blockIdx.x / ktiles) * PT_R;
    B += sel.s[blockIdx.y] * (int64_t)n * nrhs;
    Xp += (int64_t)blockIdx.y * n * ld;
    // all the tile's loads in flight at once (out-of-range entries read B[0] and are dropped)
    constexpr int PER = PT_K * PT_R / 256;
    cx<TIO> b[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int e = threadIdx.x + 256 * u, kk = e % PT_K, rr = e / PT_K;
        b[u] = B[k0 + kk < n && r0 + rr < nrhs ? k0 + kk + (int64_t)(r0 + rr) * n : 0];
    }
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int e = threadIdx.x + 256 * u, kk = e % PT_K, rr = e / PT_K;
        S[rr][kk] = k0 + kk < n && r0 + rr < nrhs ? cx<T>{(T)b[u].x, (T)b[u].y} : cx<T>{T(0), T(0)};
    }
    __syncthreads();
    for (int e = threadIdx.x; e < PT_K * PT_R; e += 256) {
        const int rr = e % PT_R, kk = e / PT_R;
        if (k0 + kk < n && r0 + rr < ld) Xp[(int64_t)iperm[k0 + kk] * ld + r0 + rr] = S[rr][kk];
    }
}

// iout: position of user row k (the inverse elimination order)
template <typename T, typename TIO>
__global__ void __launch_bounds__(256) k_permute_out_slots(const cx<T> *__restrict__ Xp, cx<TIO> *__restrict__ X,
                                                           const int *__restrict__ iout, int64_t map_stride, int n,
                                                           int nrhs, int ld, int accumulate, Sel sel) {
    __shared__ cx<T> S[PT_R][PT_K + 1];
    const int ktiles = (n + PT_K - 1) / PT_K;
    const int k0 = (blockIdx.x % ktiles) * PT_K, r0 = (blockIdx.x / ktiles) * PT_R;
    X += sel.s[blockIdx.y] * (int64_t)n * nrhs;
    Xp += (int64_t)blockIdx.y * n * ld;
    iout += sel.f[blockIdx.y] * map_stride;
    // the rows' positions, then all the tile's loads in flight at once
    // (out-of-range entries read Xp[0] and are dropped)
    constexpr int PER = PT_K * PT_R / 256;
    int64_t q[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int e = threadIdx.x + 256 * u, rr = e % PT_R, kk = e / PT_R;
        q[u] = k0 + kk < n && r0 + rr < nrhs ? (int64_t)iout[k0 + kk] * ld + r0 + rr : 0;
    }
    cx<T> v[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) v[u] = Xp[q[u]];
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int e = threadIdx.x + 256 * u;
        S[e % PT_R][e / PT_R] = v[u];
    }
    __syncthreads();
    cx<TIO> xo[PER];
    if (accumulate)
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int e = threadIdx.x + 256 * u, kk = e % PT_K, rr = e / PT_K;
            xo[u] = X[k0 + kk < n && r0 + rr < nrhs ? k0 + kk + (int64_t)(r0 + rr) * n : 0];
        }
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int e = threadIdx.x + 256 * u, kk = e % PT_K, rr = e / PT_K;
        if (k0 + kk < n && r0 + rr < nrhs) {
            const cx<T> w = S[rr][kk];
            X[k0 + kk + (int64_t)(r0 + rr) * n] =
                accumulate ? cx<TIO>{xo[u].x + (TIO)w.x, xo[u].y + (TIO)w.y} : cx<TIO>{(TIO)w.x, (TIO)w.y};
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
                                                 const int *__restrict__ iout, int n, int accumulate) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const cx<T> v = Xp[iout[k]];
        if (accumulate) X[k] = {X[k].x + (TIO)v.x, X[k].y + (TIO)v.y};
        else X[k] = {(TIO)v.x, (TIO)v.y};
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

// forward gather: W[f] rows from the gather table (per front row: X position
// or -1, then up to two children's contribution rows in W or -1; for M solves
// the table already follows the front's row permutation).  Four entries per
// thread, every load of the four issued before the sums.
// grid (fronts, chunks, slots)
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_gather(const cx<T> *__restrict__ X, cx<T> *__restrict__ W,
                                                    const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                    const int *__restrict__ gsrc, int64_t gstride, int n,
                                                    int64_t wstride, int nrhs, Sel sel) {
    const FrontDev F = fronts[list[blockIdx.x]];
    const int z = blockIdx.z;
    const int64_t total = (int64_t)F.m * nrhs;
    X += (int64_t)sel.s[z] * n * nrhs;
    W += (int64_t)sel.s[z] * wstride;
    const int *gs = gsrc + sel.f[z] * gstride + F.fv_off * 3;
    const int64_t step = (int64_t)gridDim.y * blockDim.x;
    for (int64_t e0 = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; e0 < total; e0 += 4 * step) {
        cx<T> v[4], c1[4], c2[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t e = e0 + u * step;
            v[u] = c1[u] = c2[u] = {T(0), T(0)};
            if (e < total) {
                const int i = (int)(e / nrhs), r = (int)(e - (int64_t)i * nrhs);
                const int a = gs[3 * i], b = gs[3 * i + 1], c = gs[3 * i + 2];
                if (a >= 0) v[u] = X[(int64_t)a * nrhs + r];
                if (b >= 0) c1[u] = W[(int64_t)b * nrhs + r];
                if (c >= 0) c2[u] = W[(int64_t)c * nrhs + r];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t e = e0 + u * step;
            if (e < total) {
                const int i = (int)(e / nrhs), r = (int)(e - (int64_t)i * nrhs);
                W[(F.fv_off + i) * nrhs + r] = cadd(v[u], cadd(c1[u], c2[u]));
            }
        }
    }
}

// ---- panel substitutions on MFMA -------------------------------------------
// Shared-memory layout of the block kernels (dynamic): right-hand-side block
// [BK][64] re/im, then the A tile [BK][64] re/im, whose space the diagonal
// block [BK][BK+1] complex shares: the diagonal product runs (and ends with a
// barrier) before the A tile is stored.
// RT: rows of the A tile (64, or 128 for the forward block kernel's 8 waves).
template <typename T, int RT = 64>
struct BlkSmem {
    T *Ar, *Ai, *Yr, *Yi;
    cx<T> *D;
    __device__ BlkSmem(unsigned char *raw) {
        Yr = reinterpret_cast<T *>(raw);
        Yi = Yr + BK * 64;
        Ar = Yi + BK * 64;
        Ai = Ar + BK * RT;
        D = reinterpret_cast<cx<T> *>(Ar);
    }
};
static_assert(sizeof(cx<double>) * BK * (BK + 1) <= 2 * sizeof(double) * BK * 64 &&
                  sizeof(cx<float>) * BK * (BK + 1) <= 2 * sizeof(float) * BK * 64,
              "the diagonal block fits in the A tile's space");
template <typename T, int RT = 64>
constexpr size_t blk_smem_bytes() {
    return sizeof(T) * (2 * BK * 64 + 2 * BK * RT);
}

// element (i, t) of the operator that solves 32-row diagonal block b0 in the
// solve format of the factors (diagonal blocks inverted in place): forward
// sweeps apply a lower operator -- inv(L_bb) (unit diagonal) or, for
// transposed solves, inv(U_bb)^T -- backward sweeps an upper one -- inv(U_bb)
// or inv(L_bb)^T
template <typename T>
__device__ __forceinline__ cx<T> dop(const cx<T> *G, int m, int b0, int i, int t, bool fwd, bool trans, bool cj) {
    if (fwd ? t > i : t < i) return {T(0), T(0)};
    if (i == t && fwd != trans) return {T(1), T(0)};
    return trans ? ld(G + b0 + t + (int64_t)(b0 + i) * m, cj) : ld(G + b0 + i + (int64_t)(b0 + t) * m, cj);
}
// the same in two steps, so the load can stay in flight: dop_raw loads
// unconditionally (entries outside the nb x nb block read its (0, 0) entry),
// dop_fix applies the masks (zero outside the block)
template <typename T>
__device__ __forceinline__ cx<T> dop_raw(const cx<T> *G, int m, int b0, int i, int t, int nb, bool trans) {
    if (!(i < nb && t < nb)) i = t = 0;
    return G[trans ? b0 + t + (int64_t)(b0 + i) * m : b0 + i + (int64_t)(b0 + t) * m];
}
template <typename T>
__device__ __forceinline__ cx<T> dop_fix(cx<T> v, int i, int t, int nb, bool fwd, bool trans, bool cj) {
    if (!(i < nb && t < nb) || (fwd ? t > i : t < i)) return {T(0), T(0)};
    if (i == t && fwd != trans) return {T(1), T(0)};
    return cj ? cx<T>{v.x, -v.y} : v;
}

// acc = A[64 x kb] * Y[kb x 64] from LDS (4 waves in 2x2, 32x32 each), or
// acc += that when accumulate; column sub-tiles at or beyond ncol (right-hand
// sides past the end) are skipped
template <typename T, int RT = 64>
__device__ __forceinline__ void mma_tile(const BlkSmem<T, RT> &s, int kb, int ncol, cacc3<T> (&c)[2][2],
                                         bool accumulate = false) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int lr = lane & 15, lk = lane >> 4;
    if (!accumulate)
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) c[a][b].zero();
    const int kend = (kb + 3) & ~3;
    for (int kk = 0; kk < kend; kk += 4) {
        const int k = kk + lk;
        T ar[2], ai[2], br[2], bi[2];
        for (int a = 0; a < 2; a++) {
            ar[a] = s.Ar[k * RT + wr + a * 16 + lr];
            ai[a] = s.Ai[k * RT + wr + a * 16 + lr];
        }
        for (int b = 0; b < 2; b++) {
            br[b] = s.Yr[k * 64 + wc + b * 16 + lr];
            bi[b] = s.Yi[k * 64 + wc + b * 16 + lr];
        }
        for (int b = 0; b < 2; b++)
            if (wc + b * 16 < ncol)
                for (int a = 0; a < 2; a++) c[a][b].mma(ar[a], ai[a], br[b], bi[b], br[b] + bi[b]);
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

// the same, with the tile's current values of w loaded beforehand into c
// (load_tile), so the read-modify-write waits on nothing
template <typename T>
__device__ __forceinline__ void load_tile(const cx<T> *w, int nrhs, int row0, int rowend, int r0, cx<T> (&c)[2][2][4]) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
    const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                // unconditional: entries outside read w[0] and are never used
                const int row = row0 + wr + a * 16 + mfma_t<T>::row(lk, v), col = r0 + wc + b * 16 + lr;
                c[a][b][v] = w[row < rowend && col < nrhs ? (int64_t)row * nrhs + col : 0];
            }
}
template <typename T>
__device__ __forceinline__ void sub_tile_from(cx<T> *w, int nrhs, int row0, int rowend, int r0,
                                              const cx<T> (&c)[2][2][4], const cacc3<T> (&acc)[2][2]) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
    const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = row0 + wr + a * 16 + mfma_t<T>::row(lk, v), col = r0 + wc + b * 16 + lr;
                if (row < rowend && col < nrhs)
                    w[(int64_t)row * nrhs + col] = {c[a][b][v].x - acc[a][b].re(v), c[a][b][v].y - acc[a][b].im(v)};
            }
}

// Y[0:nb] = D Y[0:nb] in LDS (D: the block operator, zero outside nb x nb):
// waves 0-3 own 16 rows x 32 right-hand sides each (further waves only meet
// the barriers)
template <typename T, int RT = 64>
__device__ __forceinline__ void diag_apply(const BlkSmem<T, RT> &s, int nb, int ncol) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, rw = (w >> 1) * 16, wc = (w & 1) * 32;
    const int lr = lane & 15, lk = lane >> 4;
    cacc3<T> c[2];
    for (int b = 0; b < 2; b++) c[b].zero();
    const bool act = RT == 64 || w < 4;
    const int kend = act ? (nb + 3) & ~3 : 0;
    for (int kk = 0; kk < kend; kk += 4) {
        const int k = kk + lk;
        const cx<T> a = s.D[(rw + lr) * (BK + 1) + k];
        for (int b = 0; b < 2; b++)
            if (wc + b * 16 < ncol) {
                const T br = s.Yr[k * 64 + wc + b * 16 + lr], bi = s.Yi[k * 64 + wc + b * 16 + lr];
                c[b].mma(a.x, a.y, br, bi, br + bi);
            }
    }
    __syncthreads();
    if (act)
        for (int b = 0; b < 2; b++)
            for (int v = 0; v < 4; v++) {
                const int row = rw + mfma_t<T>::row(lk, v), col = wc + b * 16 + lr;
                if (row < nb) {
                    s.Yr[row * 64 + col] = c[b].re(v);
                    s.Yi[row * 64 + col] = c[b].im(v);
                }
            }
    __syncthreads();
}

// forward block b0 of every front of a level: y_b = Dop_bb w_b (written to X
// by the row-tile-0 workgroup), then rows [b0+nb, m) of W -= Lop[:, b] y_b.
// Everything the workgroup reads -- the diagonal operator, y's rows, the
// operator block below and the rows it updates -- is requested before the
// first barrier, so the workgroup waits on global memory once.
// grid: (row tiles, fronts, right-hand-side tiles)
constexpr int kFwdRT = 128;   // rows per forward block workgroup (8 waves of 32x32 tiles)
template <typename T>
__global__ void __launch_bounds__(4 * kFwdRT) k_fwd_blk(const cx<T> *__restrict__ arena, int64_t fstride,
                                                 const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                 cx<T> *__restrict__ X, cx<T> *__restrict__ W, int n, int64_t wstride,
                                                 int nrhs, int rt, int b0, int trans, int cj, Sel sel) {
    constexpr int RT = kFwdRT, NT = 4 * RT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    BlkSmem<T, RT> s(smem_raw);
    const FrontDev F = fronts[list[blockIdx.y]];
    if (F.p <= b0) return;
    const int m = F.m, nb = min(BK, F.p - b0), tid = threadIdx.x;
    const int row0 = b0 + nb + blockIdx.x * RT;
    if (blockIdx.x > 0 && row0 >= m) return;
    const bool upd = row0 < m;
    const int z = blockIdx.z / rt, r0 = (blockIdx.z - z * rt) * 64, ncol = min(64, nrhs - r0);
    const cx<T> *G = arena + sel.f[z] * fstride + F.buf_off;
    X += (int64_t)sel.s[z] * n * nrhs;
    cx<T> *w = W + sel.s[z] * wstride + F.fv_off * nrhs;
    auto Lop = [&](int i, int t) { return trans ? ld(G + t + (int64_t)i * m, cj) : ld(G + i + (int64_t)t * m, cj); };
    cx<T> wold[2][2][4];
    if (upd) load_tile<T>(w, nrhs, row0, m, r0, wold);
    {
        cx<T> dv[BK * BK / NT], yv[BK * 64 / NT], av[BK * RT / NT];
#pragma unroll
        for (int u = 0; u < BK * BK / NT; u++) {
            const int e = tid + NT * u, i = e % BK, t = e / BK;
            dv[u] = (i < nb && t < nb) ? dop(G, m, b0, i, t, true, trans, cj) : cx<T>{T(0), T(0)};
        }
#pragma unroll
        for (int u = 0; u < BK * 64 / NT; u++) {
            const int e = tid + NT * u, k = e >> 6, j = e & 63;
            yv[u] = (k < nb && j < ncol) ? w[(int64_t)(b0 + k) * nrhs + r0 + j] : cx<T>{T(0), T(0)};
        }
#pragma unroll
        for (int u = 0; u < BK * RT / NT; u++) {
            const int e = tid + NT * u;
            int k, i;
            if (!trans) { i = e % RT; k = e / RT; }        // column-major L: coalesced along rows
            else { k = e % BK; i = e / BK; }               // U^T: coalesced along k
            av[u] = (upd && k < nb && row0 + i < m) ? Lop(row0 + i, b0 + k) : cx<T>{T(0), T(0)};
        }
#pragma unroll
        for (int u = 0; u < BK * BK / NT; u++) {
            const int e = tid + NT * u;
            s.D[(e % BK) * (BK + 1) + e / BK] = dv[u];
        }
#pragma unroll
        for (int u = 0; u < BK * 64 / NT; u++) {
            const int e = tid + NT * u;
            s.Yr[e] = yv[u].x;
            s.Yi[e] = yv[u].y;
        }
        __syncthreads();
        diag_apply<T, RT>(s, nb, ncol);   // ends with a barrier: D's space is free
#pragma unroll
        for (int u = 0; u < BK * RT / NT; u++) {
            const int e = tid + NT * u;
            int k, i;
            if (!trans) { i = e % RT; k = e / RT; }
            else { k = e % BK; i = e / BK; }
            s.Ar[k * RT + i] = av[u].x;
            s.Ai[k * RT + i] = av[u].y;
        }
    }
    __syncthreads();
    if (blockIdx.x == 0)
        for (int e = tid; e < nb * 64; e += NT) {
            int k = e >> 6, j = e & 63;
            if (j < ncol) X[(int64_t)(F.pos0 + b0 + k) * nrhs + r0 + j] = {s.Yr[e], s.Yi[e]};
        }
    if (!upd) return;
    cacc3<T> acc[2][2];
    mma_tile<T, RT>(s, nb, ncol, acc);
    sub_tile_from<T>(w, nrhs, row0, m, r0, wold, acc);
}

// Y1[0:nb1] -= L Y0[0:kb] in LDS (L: nb1 x kb in the diagonal block's layout,
// row i at L[i * (BK + 1)]): waves 0-3 own 16 rows x 32 right-hand sides each,
// as in diag_apply; ends with a barrier
template <typename T>
__device__ __forceinline__ void blk_couple(const cx<T> *L, const T *Y0r, const T *Y0i, T *Y1r, T *Y1i, int kb, int nb1,
                                           int ncol) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, rw = (w >> 1) * 16, wc = (w & 1) * 32;
    const int lr = lane & 15, lk = lane >> 4;
    cacc3<T> c[2];
    for (int b = 0; b < 2; b++) c[b].zero();
    const bool act = w < 4;
    const int kend = act ? (kb + 3) & ~3 : 0;
    for (int kk = 0; kk < kend; kk += 4) {
        const int k = kk + lk;
        const cx<T> a = L[(rw + lr) * (BK + 1) + k];
        for (int b = 0; b < 2; b++)
            if (wc + b * 16 < ncol) {
                const T br = Y0r[k * 64 + wc + b * 16 + lr], bi = Y0i[k * 64 + wc + b * 16 + lr];
                c[b].mma(a.x, a.y, br, bi, br + bi);
            }
    }
    if (act)
        for (int b = 0; b < 2; b++)
            for (int v = 0; v < 4; v++) {
                const int row = rw + mfma_t<T>::row(lk, v), col = wc + b * 16 + lr;
                if (row < nb1) {
                    Y1r[row * 64 + col] -= c[b].re(v);
                    Y1i[row * 64 + col] -= c[b].im(v);
                }
            }
    __syncthreads();
}

// forward blocks b0 and b0 + 32 of every front of a level in one launch:
// y_0 = Dop_0 w_0, w_1 -= Lop[b1, b0] y_0, y_1 = Dop_1 w_1 (both written to X
// by the row-tile-0 workgroup), then rows [b0 + 64, m) of W -= Lop[:, b0:b0+64]
// [y_0; y_1] as one depth-64 update -- the rows below pass through memory once
// per two blocks instead of once per block.  Every operand but the second block
// column of Lop is requested before the first barrier; that one is loaded
// during the first column's products.  A front with a single block left takes
// the first half alone.  LDS: [Y1 | Y0 | A tile], the A tile's space holding
// Dop_0, Lop[b1, b0] and Dop_1 in turn.  grid: (row tiles, fronts, rhs tiles)
template <typename T>
constexpr size_t blk2_smem_bytes() {
    return sizeof(T) * 2 * BK * 64 + blk_smem_bytes<T, kFwdRT>();
}
template <typename T>
__global__ void __launch_bounds__(4 * kFwdRT) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1))) k_fwd_blk2(const cx<T> *__restrict__ arena, int64_t fstride,
                                                  const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                  cx<T> *__restrict__ X, cx<T> *__restrict__ W, int n, int64_t wstride,
                                                  int nrhs, int rt, int b0, int trans, int cj, Sel sel) {
    constexpr int RT = kFwdRT, NT = 4 * RT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    BlkSmem<T, RT> s0(smem_raw + sizeof(T) * 2 * BK * 64);   // Y0 and the A tile
    BlkSmem<T, RT> s1 = s0;                                   // Y1 and the same A tile
    s1.Yr = reinterpret_cast<T *>(smem_raw);
    s1.Yi = s1.Yr + BK * 64;
    const FrontDev F = fronts[list[blockIdx.y]];
    if (F.p <= b0) return;
    const int m = F.m, nb0 = min(BK, F.p - b0), nb1 = max(0, min(BK, F.p - b0 - BK)), tid = threadIdx.x;
    const int row0 = b0 + nb0 + nb1 + blockIdx.x * RT;
    if (blockIdx.x > 0 && row0 >= m) return;
    const bool upd = row0 < m;
    const int z = blockIdx.z / rt, r0 = (blockIdx.z - z * rt) * 64, ncol = min(64, nrhs - r0);
    const cx<T> *G = arena + sel.f[z] * fstride + F.buf_off;
    X += (int64_t)sel.s[z] * n * nrhs;
    cx<T> *w = W + sel.s[z] * wstride + F.fv_off * nrhs;
    cx<T> wold[2][2][4];
    if (upd) load_tile<T>(w, nrhs, row0, m, r0, wold);
    cx<T> d0[BK * BK / NT], d1[BK * BK / NT], l10[BK * BK / NT], y0v[BK * 64 / NT], y1v[BK * 64 / NT],
        av[BK * RT / NT];
    // every load is unconditional (entries outside read a valid address of
    // the front or panel), and its masks are applied when it is staged, so
    // the whole set is in flight at once
    auto lop_at = [&](int i, int t) { return trans ? t + (int64_t)i * m : i + (int64_t)t * m; };
#pragma unroll
    for (int u = 0; u < BK * BK / NT; u++) {
        const int e = tid + NT * u, i = e % BK, t = e / BK;
        d0[u] = dop_raw(G, m, b0, i, t, nb0, trans);
        d1[u] = dop_raw(G, m, b0 + BK, i, t, nb1, trans);
        l10[u] = G[i < nb1 && t < nb0 ? lop_at(b0 + BK + i, b0 + t) : 0];
    }
    const int64_t wc0 = r0;
#pragma unroll
    for (int u = 0; u < BK * 64 / NT; u++) {
        const int e = tid + NT * u, k = e >> 6, j = e & 63;
        y0v[u] = w[k < nb0 && j < ncol ? (int64_t)(b0 + k) * nrhs + wc0 + j : 0];
        y1v[u] = w[k < nb1 && j < ncol ? (int64_t)(b0 + BK + k) * nrhs + wc0 + j : 0];
    }
    int akb = 0;   // columns of the A tile in av
    auto load_a = [&](int c0, int kb) {   // the A tile of block column c0 (kb columns) into av
        akb = kb;
#pragma unroll
        for (int u = 0; u < BK * RT / NT; u++) {
            const int e = tid + NT * u;
            int k, i;
            if (!trans) { i = e % RT; k = e / RT; }        // column-major L: coalesced along rows
            else { k = e % BK; i = e / BK; }               // U^T: coalesced along k
            av[u] = G[upd && k < kb && row0 + i < m ? lop_at(row0 + i, c0 + k) : 0];
        }
    };
    auto stage_a = [&]() {
#pragma unroll
        for (int u = 0; u < BK * RT / NT; u++) {
            const int e = tid + NT * u;
            int k, i;
            if (!trans) { i = e % RT; k = e / RT; }
            else { k = e % BK; i = e / BK; }
            const bool in = upd && k < akb && row0 + i < m;
            s0.Ar[k * RT + i] = in ? av[u].x : T(0);
            s0.Ai[k * RT + i] = in ? (cj ? -av[u].y : av[u].y) : T(0);
        }
    };
    auto stage_d = [&](const cx<T>(&v)[BK * BK / NT], int nb) {   // a diagonal block operator
#pragma unroll
        for (int u = 0; u < BK * BK / NT; u++) {
            const int e = tid + NT * u;
            s0.D[(e % BK) * (BK + 1) + e / BK] = dop_fix(v[u], e % BK, e / BK, nb, true, trans, cj);
        }
    };
    load_a(b0, nb0);
    stage_d(d0, nb0);
#pragma unroll
    for (int u = 0; u < BK * 64 / NT; u++) {
        const int e = tid + NT * u, k = e >> 6, j = e & 63;
        const bool in0 = k < nb0 && j < ncol, in1 = k < nb1 && j < ncol;
        s0.Yr[e] = in0 ? y0v[u].x : T(0);
        s0.Yi[e] = in0 ? y0v[u].y : T(0);
        s1.Yr[e] = in1 ? y1v[u].x : T(0);
        s1.Yi[e] = in1 ? y1v[u].y : T(0);
    }
    __syncthreads();
    diag_apply<T, RT>(s0, nb0, ncol);   // y_0; ends with a barrier
    if (nb1 > 0) {
#pragma unroll
        for (int u = 0; u < BK * BK / NT; u++) {   // Lop[b1, b0]
            const int e = tid + NT * u, i = e % BK, t = e / BK;
            const cx<T> v = l10[u];
            s0.D[i * (BK + 1) + t] = i < nb1 && t < nb0 ? (cj ? cx<T>{v.x, -v.y} : v) : cx<T>{T(0), T(0)};
        }
        __syncthreads();
        blk_couple<T>(s0.D, s0.Yr, s0.Yi, s1.Yr, s1.Yi, nb0, nb1, ncol);   // w_1 -= L10 y_0
        stage_d(d1, nb1);
        __syncthreads();
        diag_apply<T, RT>(s1, nb1, ncol);   // y_1
    }
    if (blockIdx.x == 0)
        for (int e = tid; e < (nb0 + nb1) * 64; e += NT) {
            const int k = e >> 6, j = e & 63;
            if (j < ncol) {
                const int kk = k < nb0 ? k : k - nb0;
                const T *yr = k < nb0 ? s0.Yr : s1.Yr, *yi = k < nb0 ? s0.Yi : s1.Yi;
                X[(int64_t)(F.pos0 + b0 + k) * nrhs + r0 + j] = {yr[kk * 64 + j], yi[kk * 64 + j]};
            }
        }
    if (!upd) return;
    stage_a();   // the diagonal blocks' space is free (diag_apply ended with a barrier)
    __syncthreads();
    if (nb1 > 0) load_a(b0 + BK, nb1);   // in flight during the first column's products
    cacc3<T> acc[2][2];
    mma_tile<T, RT>(s0, nb0, ncol, acc);
    if (nb1 > 0) {
        __syncthreads();   // the first column's tile is no longer read
        stage_a();
        __syncthreads();
        mma_tile<T, RT>(s1, nb1, ncol, acc, true);
    }
    sub_tile_from<T>(w, nrhs, row0, m, r0, wold, acc);
}

// backward, ancestors: W[f][0:p] = y1 - Uop12 x2, x2 gathered from X at the
// contribution positions.  The positions are staged in LDS first (after the
// block tiles, m - p ints), so each 32-column chunk of Uop12 and x2 is one
// memory round trip, loaded into registers while the previous chunk's products
// run.  grid: (row tiles over p, fronts, rhs tiles)
constexpr int kContribRowsMax = 16384;   // m - p of any front on this path
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_contrib(const cx<T> *__restrict__ arena, int64_t fstride,
                                                     const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                     const int *__restrict__ idx, const cx<T> *__restrict__ X,
                                                     cx<T> *__restrict__ W, int n, int64_t wstride, int nrhs, int rt,
                                                     int trans, int cj, Sel sel) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    BlkSmem<T> s(smem_raw);
    int *srow = reinterpret_cast<int *>(smem_raw + blk_smem_bytes<T>());
    const FrontDev F = fronts[list[blockIdx.y]];
    const int m = F.m, p = F.p, tid = threadIdx.x;
    const int row0 = blockIdx.x * 64;
    if (row0 >= p) return;
    for (int e = tid; e < m - p; e += 256) srow[e] = idx[F.idx_off + p + e];
    __syncthreads();
    const int z = blockIdx.z / rt, r0 = (blockIdx.z - z * rt) * 64, ncol = min(64, nrhs - r0);
    const cx<T> *G = arena + sel.f[z] * fstride + F.buf_off;
    X += (int64_t)sel.s[z] * n * nrhs;
    cx<T> *w = W + sel.s[z] * wstride + F.fv_off * nrhs;
    auto uop_at = [&](int i, int t) { return trans ? t + (int64_t)i * m : i + (int64_t)t * m; };
    cacc3<T> acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b].zero();
    constexpr int PER = BK * 64 / 256;
    cx<T> av[PER], yv[PER];
    // a chunk's operands, loaded unconditionally (entries outside read a valid
    // address and are zeroed when staged), so they stay in flight during the
    // previous chunk's products
    auto fetch = [&](int s0) {
        const int kb = min(BK, m - s0);
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int e = tid + 256 * u;
            int k, i;
            if (!trans) { i = e & 63; k = e >> 6; }
            else { k = e % BK; i = e / BK; }
            av[u] = G[k < kb && row0 + i < p ? uop_at(row0 + i, s0 + k) : 0];
        }
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int e = tid + 256 * u, k = e >> 6, j = e & 63;
            const bool in = k < kb && j < ncol;
            yv[u] = X[(int64_t)srow[s0 - p + (in ? k : 0)] * nrhs + (in ? r0 + j : 0)];
        }
    };
    if (p < m) fetch(p);
    for (int s0 = p; s0 < m; s0 += BK) {
        const int kb = min(BK, m - s0);
        __syncthreads();   // the previous chunk's operands are no longer read
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int e = tid + 256 * u;
            int k, i;
            if (!trans) { i = e & 63; k = e >> 6; }
            else { k = e % BK; i = e / BK; }
            const bool ain = k < kb && row0 + i < p;
            s.Ar[k * 64 + i] = ain ? av[u].x : T(0);
            s.Ai[k * 64 + i] = ain ? (cj ? -av[u].y : av[u].y) : T(0);
            const int ky = e >> 6, j = e & 63;
            const bool yin = ky < kb && j < ncol;
            s.Yr[e] = yin ? yv[u].x : T(0);
            s.Yi[e] = yin ? yv[u].y : T(0);
        }
        __syncthreads();
        if (s0 + BK < m) fetch(s0 + BK);   // in flight during this chunk's products
        mma_tile<T>(s, kb, ncol, acc, true);
    }
    // w[0:p] = y1 - accumulated product
    cx<T> y1[2][2][4];
    load_tile<T>(X + (int64_t)F.pos0 * nrhs, nrhs, row0, p, r0, y1);
    sub_tile_from<T>(w, nrhs, row0, p, r0, y1, acc);
}

// backward block b0 (blocks from the bottom): x_b = Dop_bb z_b (to X by the
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
    const int z = blockIdx.z / rt, r0 = (blockIdx.z - z * rt) * 64, ncol = min(64, nrhs - r0);
    const cx<T> *G = arena + sel.f[z] * fstride + F.buf_off;
    X += (int64_t)sel.s[z] * n * nrhs;
    cx<T> *w = W + sel.s[z] * wstride + F.fv_off * nrhs;
    auto uop_at = [&](int i, int t) { return trans ? t + (int64_t)i * m : i + (int64_t)t * m; };
    const bool upd = row0 < b0;
    // every load before the first barrier, unconditional (entries outside read
    // a valid address of the front or panel; masks applied when staged)
    cx<T> wold[2][2][4];
    if (upd) load_tile<T>(w, nrhs, row0, b0, r0, wold);
    {
        cx<T> dv[BK * BK / 256], yv[BK * 64 / 256], av[BK * 64 / 256];
#pragma unroll
        for (int u = 0; u < BK * BK / 256; u++) {
            const int e = tid + 256 * u, i = e % BK, t = e / BK;
            dv[u] = dop_raw(G, m, b0, i, t, nb, trans);
        }
#pragma unroll
        for (int u = 0; u < BK * 64 / 256; u++) {
            const int e = tid + 256 * u, k = e >> 6, j = e & 63;
            yv[u] = w[k < nb && j < ncol ? (int64_t)(b0 + k) * nrhs + r0 + j : 0];
        }
#pragma unroll
        for (int u = 0; u < BK * 64 / 256; u++) {
            const int e = tid + 256 * u;
            int k, i;
            if (!trans) { i = e & 63; k = e >> 6; }
            else { k = e % BK; i = e / BK; }
            av[u] = G[upd && k < nb && row0 + i < b0 ? uop_at(row0 + i, b0 + k) : 0];
        }
#pragma unroll
        for (int u = 0; u < BK * BK / 256; u++) {
            const int e = tid + 256 * u;
            s.D[(e % BK) * (BK + 1) + e / BK] = dop_fix(dv[u], e % BK, e / BK, nb, false, trans, cj);
        }
#pragma unroll
        for (int u = 0; u < BK * 64 / 256; u++) {
            const int e = tid + 256 * u, k = e >> 6, j = e & 63;
            const bool in = k < nb && j < ncol;
            s.Yr[e] = in ? yv[u].x : T(0);
            s.Yi[e] = in ? yv[u].y : T(0);
        }
        __syncthreads();
        diag_apply<T>(s, nb, ncol);   // ends with a barrier: D's space is free
#pragma unroll
        for (int u = 0; u < BK * 64 / 256; u++) {
            const int e = tid + 256 * u;
            int k, i;
            if (!trans) { i = e & 63; k = e >> 6; }
            else { k = e % BK; i = e / BK; }
            const bool in = upd && k < nb && row0 + i < b0;
            s.Ar[k * 64 + i] = in ? av[u].x : T(0);
            s.Ai[k * 64 + i] = in ? (cj ? -av[u].y : av[u].y) : T(0);
        }
    }
    __syncthreads();
    if (blockIdx.x == 0)
        for (int e = tid; e < nb * 64; e += 256) {
            int k = e >> 6, j = e & 63;
            if (j < ncol) X[(int64_t)(F.pos0 + b0 + k) * nrhs + r0 + j] = {s.Yr[e], s.Yi[e]};
        }
    if (!upd) return;
    cacc3<T> acc[2][2];
    mma_tile<T>(s, nb, ncol, acc);
    sub_tile_from<T>(w, nrhs, row0, b0, r0, wold, acc);
}

// ---- fused per-front substitutions (fronts whose panel fits in LDS) ---------
// One workgroup owns one front and RW right-hand sides of one frequency and
// keeps the front's m x RW panel in LDS (re/im planes, row stride RW + 16 so the
// four k-rows of one MFMA operand read fall in different banks).  It gathers
// the panel, runs every 32-pivot block -- the inverted diagonal block, then the
// block column's update of the rows below, both as MFMA products with the
// factor streamed from global memory as the A operand (waves own 16-row
// strips) -- and writes the panel back: one launch per level and sweep.
template <typename T, int RW>
constexpr size_t fused_row_bytes() {
    return 2 * sizeof(T) * (RW + 16) + 4 * sizeof(int);     // panel row + the row's gather sources
}

// The dispatcher hands consecutive workgroup ids to the 8 XCDs round-robin;
// with gridDim.x a multiple of 8, blockIdx.x mod 8 picks the XCD.  Keeping
// front = (group, blockIdx.x mod 8) puts all right-hand-side tiles of a front on
// one XCD, so its factor entries come from HBM once and from that L2 after.
__device__ __forceinline__ void xcd_front_tile(int bx, int rt, int &fi, int &tile) {
    const int grp = bx / (8 * rt), rem = bx - grp * 8 * rt;
    tile = rem >> 3;
    fi = grp * 8 + (rem & 7);
}

// Operand A of a strip product: A(i, k) = A[i * ars + k * acs] (conjugated if
// cj) for i < arows, k < K, zero elsewhere.  Masks in absolute indices
// (row ioff + i, column koff + k): tri 1 keeps columns <= row (lower), tri 2
// columns >= row (upper); unit takes the diagonal as 1; columns >= kneg are
// negated.
struct AOp {
    int64_t ars, acs;
    int arows, K, ioff, koff, tri, kneg;
    bool unit, cj;
};

// The 8 A values of lane (lr, lk) for k = k0 + 4s + lk, s < 8, in two steps:
// load_a_raw issues the loads unconditionally (entries outside the operand
// read A[0], which lies in the same front), so a caller can keep them in
// flight past other work; finish_a applies the operand's masks once they are
// needed (the first use waits for the loads)
template <typename T>
__device__ __forceinline__ void load_a_raw(const cx<T> *__restrict__ A, const AOp &o, int k0, cx<T> (&raw)[8]) {
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
    const bool rok = lr < o.arows;
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int k = k0 + 4 * s + lk;
        raw[s] = A[rok && k < o.K ? lr * o.ars + k * o.acs : 0];
    }
}
template <typename T>
__device__ __forceinline__ void finish_a(const AOp &o, int k0, cx<T> (&av)[8]) {
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4, di = o.ioff + lr;
    const bool rok = lr < o.arows;
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int k = k0 + 4 * s + lk, ka = o.koff + k;
        cx<T> v = av[s];
        if (o.cj) v.y = -v.y;
        if ((o.tri == 1 && ka > di) || (o.tri == 2 && ka < di)) v = {T(0), T(0)};
        if (o.unit && ka == di) v = {T(1), T(0)};
        if (ka >= o.kneg) v = {-v.x, -v.y};
        if (!(rok && k < o.K)) v = {T(0), T(0)};
        av[s] = v;
    }
}
template <typename T>
__device__ __forceinline__ void load_a(const cx<T> *__restrict__ A, const AOp &o, int k0, cx<T> (&av)[8]) {
    load_a_raw<T>(A, o, k0, av);
    finish_a<T>(o, k0, av);
}

// acc[c] += A(16 x [k0, k0+32)) * P[prow0 + k][c*16 .. c*16 + 16) with A from av
template <typename T, int NC>
__device__ __forceinline__ void mma_a(const cx<T> (&av)[8], int k0, int K, const T *Pr, const T *Pi, int prow0, int rs,
                                      typename mfma_t<T>::acc (&cr)[NC], typename mfma_t<T>::acc (&ci)[NC]) {
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int s = 0; s < 8; s++) {
        if (k0 + 4 * s >= K) break;
        const int k = k0 + 4 * s + lk;
        const bool kin = k < K;
        const int off = (prow0 + k) * rs + lr;
        // every column sub-tile, also past the right-hand sides (zero panel
        // columns): a per-sub-tile branch costs far more registers than the MFMAs
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const T br = kin ? Pr[off + c * 16] : T(0), bi = kin ? Pi[off + c * 16] : T(0);
            cmma<T>(av[s].x, av[s].y, br, bi, cr[c], ci[c]);
        }
    }
}

// acc[c] += A(16 x K) * P[prow0 .. prow0 + K)[c*16 .. c*16 + 16), K in chunks of
// 32; av holds chunk 0 on entry (loaded by the caller), and each next chunk is
// requested before the current one's MFMAs
template <typename T, int NC>
__device__ __forceinline__ void strip_mma_from(cx<T> (&av)[8], const cx<T> *__restrict__ A, const AOp &o,
                                               const T *Pr, const T *Pi, int prow0, int rs,
                                               typename mfma_t<T>::acc (&cr)[NC], typename mfma_t<T>::acc (&ci)[NC]) {
    cx<T> an[8];
    for (int k0 = 0; k0 < o.K; k0 += 32) {
        if (k0 + 32 < o.K) load_a_raw<T>(A, o, k0 + 32, an);
        if (k0 > 0) finish_a<T>(o, k0, av);
        mma_a<T, NC>(av, k0, o.K, Pr, Pi, prow0, rs, cr, ci);
#pragma unroll
        for (int s = 0; s < 8; s++) av[s] = an[s];
    }
}

template <typename T, int NC>
__device__ __forceinline__ void strip_mma(const cx<T> *__restrict__ A, const AOp &o, const T *Pr, const T *Pi,
                                          int prow0, int rs, typename mfma_t<T>::acc (&cr)[NC],
                                          typename mfma_t<T>::acc (&ci)[NC]) {
    cx<T> av[8];
    load_a<T>(A, o, 0, av);
    strip_mma_from<T, NC>(av, A, o, Pr, Pi, prow0, rs, cr, ci);
}

template <typename T, int NC>
__device__ __forceinline__ void acc_zero(typename mfma_t<T>::acc (&cr)[NC], typename mfma_t<T>::acc (&ci)[NC]) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
        cr[c] = typename mfma_t<T>::acc{0, 0, 0, 0};
        ci[c] = typename mfma_t<T>::acc{0, 0, 0, 0};
    }
}

// P rows [row0 + 0..16) (those below rowend) op= acc; sub: -=, else =.  All
// columns are stored: those past the right-hand sides hold zeros throughout.
template <typename T, int NC>
__device__ __forceinline__ void acc_store(T *Pr, T *Pi, int rs, int row0, int rowend, bool sub,
                                          typename mfma_t<T>::acc (&cr)[NC], typename mfma_t<T>::acc (&ci)[NC]) {
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int c = 0; c < NC; c++) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = row0 + mfma_t<T>::row(lk, v);
            if (row < rowend) {
                const int o = row * rs + c * 16 + lr;
                if (sub) {
                    Pr[o] -= cr[c][v];
                    Pi[o] -= ci[c][v];
                } else {
                    Pr[o] = cr[c][v];
                    Pi[o] = ci[c][v];
                }
            }
        }
    }
}

// Panel gather of a forward front into LDS: row s of the assembled vector --
// its X row plus the contributions of up to two children (nested-dissection
// fronts have at most two) -- lands at row i = P s.  The rows' sources are
// resolved first (one thread per row), then all element loads of a batch are
// issued before any is stored.  Ends with a barrier.
template <typename T, int RW>
__device__ __forceinline__ void gather_fwd(const FrontDev &F, const int *__restrict__ rp,
                                           const FrontDev *__restrict__ fronts, const int *__restrict__ children,
                                           const int *__restrict__ inv, const cx<T> *__restrict__ X,
                                           const cx<T> *__restrict__ W, int nrhs, int r0, int ncol, T *Pr, T *Pi) {
    constexpr int RS = RW + 16, GB = 8;
    const int m = F.m, p = F.p, tid = threadIdx.x;
    int *sx = reinterpret_cast<int *>(Pi + m * RS), *sw0 = sx + m, *sw1 = sw0 + m;
    for (int i = tid; i < m; i += 256) {
        const int src = (rp && i < p) ? rp[F.pos0 + i] : i;
        int w0 = -1, w1 = -1;
        for (int c = 0; c < F.nchild; c++) {
            const FrontDev C = fronts[children[F.child_off + c]];
            const int a = inv[C.inv_off + src];
            if (a >= 0) {
                const int r = (int)(C.fv_off + C.p + a);
                if (w0 < 0) w0 = r;
                else w1 = r;
            }
        }
        sx[i] = src < p ? F.pos0 + src : -1;
        sw0[i] = w0;
        sw1[i] = w1;
    }
    __syncthreads();
    const int total = m * RW;
    for (int e0 = 0; e0 < total; e0 += 256 * GB) {
        cx<T> v[GB];
#pragma unroll
        for (int u = 0; u < GB; u++) {
            const int e = e0 + u * 256 + tid, i = e / RW, j = e - i * RW;
            v[u] = {T(0), T(0)};
            if (e < total && j < ncol) {
                const int a = sx[i], b = sw0[i], c = sw1[i];
                if (a >= 0) v[u] = X[(int64_t)a * nrhs + r0 + j];
                if (b >= 0) v[u] = cadd(v[u], W[(int64_t)b * nrhs + r0 + j]);
                if (c >= 0) v[u] = cadd(v[u], W[(int64_t)c * nrhs + r0 + j]);
            }
        }
#pragma unroll
        for (int u = 0; u < GB; u++) {
            const int e = e0 + u * 256 + tid, i = e / RW, j = e - i * RW;
            if (e < total) {
                Pr[i * RS + j] = v[u].x;
                Pi[i * RS + j] = v[u].y;
            }
        }
    }
    __syncthreads();
}

// Panel gather of a backward front: pivot rows and the ancestors' rows of X.
// Ends with a barrier.
template <typename T, int RW>
__device__ __forceinline__ void gather_bwd(const FrontDev &F, const int *__restrict__ idx,
                                           const cx<T> *__restrict__ X, int nrhs, int r0, int ncol, T *Pr, T *Pi) {
    constexpr int RS = RW + 16, GB = 8;
    const int m = F.m, p = F.p, tid = threadIdx.x;
    int *srow = reinterpret_cast<int *>(Pi + m * RS);
    for (int i = tid; i < m; i += 256) srow[i] = i < p ? F.pos0 + i : idx[F.idx_off + i];
    __syncthreads();
    const int total = m * RW;
    for (int e0 = 0; e0 < total; e0 += 256 * GB) {
        cx<T> v[GB];
#pragma unroll
        for (int u = 0; u < GB; u++) {
            const int e = e0 + u * 256 + tid, i = e / RW, j = e - i * RW;
            v[u] = {T(0), T(0)};
            if (e < total && j < ncol) v[u] = X[(int64_t)srow[i] * nrhs + r0 + j];
        }
#pragma unroll
        for (int u = 0; u < GB; u++) {
            const int e = e0 + u * 256 + tid, i = e / RW, j = e - i * RW;
            if (e < total) {
                Pr[i * RS + j] = v[u].x;
                Pi[i * RS + j] = v[u].y;
            }
        }
    }
    __syncthreads();
}

// forward sweep of a level's fronts.  grid: (ceil(fronts/8)*8 * rt, frequency slots)
template <typename T, int RW>
__global__ void __launch_bounds__(256) k_fwd_front(const cx<T> *__restrict__ arena, int64_t fstride,
                                                   const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                   int nfr, const int *__restrict__ children,
                                                   const int *__restrict__ inv, const int *__restrict__ rowperm,
                                                   cx<T> *__restrict__ X, cx<T> *__restrict__ W, int n,
                                                   int64_t wstride, int nrhs, int rt, int trans, int cj, Sel sel) {
    constexpr int RS = RW + 16, NC = RW / 16;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int fi, tile;
    xcd_front_tile(blockIdx.x, rt, fi, tile);
    if (fi >= nfr) return;
    const FrontDev F = fronts[list[fi]];
    const int z = blockIdx.y, r0 = tile * RW, ncol = min(RW, nrhs - r0);
    const int m = F.m, p = F.p, tid = threadIdx.x, wave = tid >> 6;
    T *Pr = reinterpret_cast<T *>(smem_raw), *Pi = Pr + m * RS;
    const cx<T> *G = arena + sel.f[z] * fstride + F.buf_off;
    X += (int64_t)sel.s[z] * n * nrhs;
    W += sel.s[z] * wstride;
    gather_fwd<T, RW>(F, rowperm ? rowperm + (int64_t)sel.f[z] * n : nullptr, fronts, children, inv, X, W, nrhs, r0,
                      ncol, Pr, Pi);
    // op(L) entry (i, t) sits at G[i * ars + t * acs]
    const int64_t ars = trans ? m : 1, acs = trans ? 1 : m;
    const bool unit = !trans;
    typename mfma_t<T>::acc cr[NC], ci[NC];
    for (int b0 = 0; b0 < p; b0 += BK) {
        const int kb = min(BK, p - b0), sr = wave * 16;
        const bool act = wave < 2 && sr < kb;
        if (act) {
            acc_zero<T, NC>(cr, ci);
            strip_mma<T, NC>(G + (b0 + sr) * ars + b0 * acs, AOp{ars, acs, kb - sr, kb, sr, 0, 1, 1 << 30, unit, (bool)cj},
                             Pr, Pi, b0, RS, cr, ci);
        }
        __syncthreads();
        if (act) acc_store<T, NC>(Pr, Pi, RS, b0 + sr, b0 + kb, false, cr, ci);
        __syncthreads();
        // trailing strips: each next strip's factor entries are requested
        // before the current strip's MFMAs
        AOp ot{ars, acs, 0, kb, 0, 0, 0, 1 << 30, false, (bool)cj};
        cx<T> at[8], an[8];
        int s0 = b0 + kb + sr;
        if (s0 < m) {
            ot.arows = m - s0;
            load_a_raw<T>(G + s0 * ars + b0 * acs, ot, 0, at);
        }
        for (; s0 < m; s0 += 64) {
            const int s1 = s0 + 64;
            if (s1 < m) {
                AOp on = ot;
                on.arows = m - s1;
                load_a_raw<T>(G + s1 * ars + b0 * acs, on, 0, an);
            }
            ot.arows = m - s0;
            finish_a<T>(ot, 0, at);
            acc_zero<T, NC>(cr, ci);
            mma_a<T, NC>(at, 0, kb, Pr, Pi, b0, RS, cr, ci);
            acc_store<T, NC>(Pr, Pi, RS, s0, m, true, cr, ci);
#pragma unroll
            for (int s = 0; s < 8; s++) at[s] = an[s];
        }
        __syncthreads();
    }
    for (int e = tid; e < m * RW; e += 256) {
        const int i = e / RW, j = e - i * RW;
        if (j >= ncol) continue;
        const cx<T> v = {Pr[i * RS + j], Pi[i * RS + j]};
        if (i < p) X[(int64_t)(F.pos0 + i) * nrhs + r0 + j] = v;
        else W[(F.fv_off + i) * nrhs + r0 + j] = v;
    }
}

// ---- single-block levels: one product per front and sweep, one wave per task --
// Fronts with p <= 32 are stored as [inv(L11); L21 inv(L11)] (m x p) and
// [inv(U11), inv(U11) U12] (p x m).  A task is one front, one frequency and 16
// right-hand sides, and belongs to one wave: no barrier with the other waves
// of the workgroup, which take the next 16-column tiles of the same front (so
// they meet the same factor entries in the caches).  Small LDS and register
// footprints keep many tasks resident per CU, which is what hides the memory
// latency of these small fronts.
constexpr int WT = 16;

// orders the wave's own LDS writes before its later LDS reads by other lanes
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Forward: out = Lop w1 over all m rows (K = p); rows < p are y (to X), rows
// >= p are subtracted from the gathered contribution rows (to W).  LDS per
// wave: the p x 16 pivot rows (B operand) and the m rows' gather sources.
// grid: (ceil(fronts/8)*8 * tile groups, frequency slots), block 64 * waves
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_front1(const cx<T> *__restrict__ arena, int64_t fstride,
                                                    const FrontDev *__restrict__ lvF, int nfr,
                                                    const int *__restrict__ gsrc, int64_t gstride,
                                                    cx<T> *__restrict__ X, cx<T> *__restrict__ W, int n,
                                                    int64_t wstride, int nrhs, int rtg, int wave_lds, int trans, int cj,
                                                    Sel sel) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
    int fi, tg;
    xcd_front_tile(blockIdx.x, rtg, fi, tg);
    const int r0 = (tg * (int)(blockDim.x >> 6) + wave) * WT;
    if (fi >= nfr || r0 >= nrhs) return;
    const int z = blockIdx.y, ncol = min(WT, nrhs - r0);
    const FrontDev F = lvF[fi];
    const int m = F.m, p = F.p;
    T *Br = reinterpret_cast<T *>(smem_raw + wave * wave_lds), *Bi = Br + p * WT;
    int *src = reinterpret_cast<int *>(Bi + p * WT);
    const cx<T> *G = arena + sel.f[z] * fstride + F.buf_off;
    X += (int64_t)sel.s[z] * n * nrhs;
    W += sel.s[z] * wstride;
    const int *gs = gsrc + sel.f[z] * gstride + F.fv_off * 3;
    const int64_t ars = trans ? m : 1, acs = trans ? 1 : m;
    const AOp op{ars, acs, m, p, 0, 0, 1, 1 << 30, !trans, (bool)cj};
    // the rows' gather sources (3m <= 900 entries: up to 15 per lane, issued
    // first, as the memory counter is in order) and strip 0's factor entries in
    // flight together, then the sources stored
    constexpr int SRC_PER = (3 * kSingleMmax + 63) / 64;
    int sv[SRC_PER];
#pragma unroll
    for (int u = 0; u < SRC_PER; u++) {
        if (64 * u >= 3 * m) break;
        const int e = lane + 64 * u;
        sv[u] = gs[e < 3 * m ? e : 0];
    }
    cx<T> av[8];
    load_a_raw<T>(G, op, 0, av);
#pragma unroll
    for (int u = 0; u < SRC_PER; u++) {
        if (64 * u >= 3 * m) break;
        const int e = lane + 64 * u;
        if (e < 3 * m) src[e] = sv[u];
    }
    wave_sync();
    {   // pivot rows (p * 16 <= 512 elements: at most 8 per lane): every source
        // loaded unconditionally (absent ones read row 0, lanes past the
        // right-hand sides column r0), validity in bits, selects once all arrived
        cx<T> xa[8], wb[8], wc[8];
        uint32_t ok = 0;
#pragma unroll
        for (int u = 0; u < 8; u++) {
            if (4 * u >= p) break;
            const int e = lane + 64 * u, i = e >> 4, j = e & 15, ic = i < p ? i : 0;
            const int64_t jc = r0 + (j < ncol ? j : 0);
            const int a = src[3 * ic], b = src[3 * ic + 1], c = src[3 * ic + 2];
            ok |= (uint32_t)((a >= 0) | (b >= 0) << 1 | (c >= 0) << 2) << (3 * u);
            xa[u] = X[(int64_t)max(a, 0) * nrhs + jc];
            wb[u] = W[(int64_t)max(b, 0) * nrhs + jc];
            wc[u] = W[(int64_t)max(c, 0) * nrhs + jc];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            if (4 * u >= p) break;
            const int e = lane + 64 * u, i = e >> 4, j = e & 15;
            cx<T> v = {T(0), T(0)};
            if (i < p && j < ncol) {
                if ((ok >> (3 * u)) & 1) v = xa[u];
                if ((ok >> (3 * u + 1)) & 1) v = cadd(v, wb[u]);
                if ((ok >> (3 * u + 2)) & 1) v = cadd(v, wc[u]);
            }
            if (i < p) {
                Br[i * WT + j] = v.x;
                Bi[i * WT + j] = v.y;
            }
        }
    }
    wave_sync();
    // the lane's four accumulator rows of a strip: both contribution sources,
    // loaded unconditionally, and their validity bits
    const int64_t lc = r0 + (lr < ncol ? lr : 0);
    auto load_w2 = [&](int s0, cx<T> (&wb)[4], cx<T> (&wc)[4], uint32_t &okw) {
        okw = 0;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = s0 + mfma_t<T>::row(lk, v);
            const bool in = row >= p && row < m && lr < ncol;
            const int rc = in ? row : 0, b = src[3 * rc + 1], c = src[3 * rc + 2];
            okw |= (uint32_t)((in && b >= 0) | (in && c >= 0) << 1) << (2 * v);
            wb[v] = W[(int64_t)max(b, 0) * nrhs + lc];
            wc[v] = W[(int64_t)max(c, 0) * nrhs + lc];
        }
    };
    cx<T> wb[4], wc[4], nb[4], nc[4], an[8];
    uint32_t okw, okn = 0;
    load_w2(0, wb, wc, okw);
    for (int s0 = 0; s0 < m; s0 += 16) {
        // the next strip's factor entries and contribution rows are requested
        // before this strip's products (strip 0's before the gather)
        const int s1 = s0 + 16;
        if (s1 < m) {
            AOp o = op;
            o.arows = m - s1;
            o.ioff = s1;
            load_a_raw<T>(G + s1 * ars, o, 0, an);
            load_w2(s1, nb, nc, okn);
        }
        {
            AOp o = op;
            o.arows = m - s0;
            o.ioff = s0;
            finish_a<T>(o, 0, av);
        }
        cx<T> w2[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            w2[v] = {T(0), T(0)};
            if ((okw >> (2 * v)) & 1) w2[v] = wb[v];
            if ((okw >> (2 * v + 1)) & 1) w2[v] = cadd(w2[v], wc[v]);
        }
        cacc3<T> c;
        c.zero();
#pragma unroll
        for (int s = 0; s < 8; s++) {
            if (4 * s >= p) break;
            const int k = 4 * s + lk;
            const bool kin = k < p;
            const T br = kin ? Br[k * WT + lr] : T(0), bi = kin ? Bi[k * WT + lr] : T(0);
            c.mma(av[s].x, av[s].y, br, bi, br + bi);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = s0 + mfma_t<T>::row(lk, v);
            if (row < m && lr < ncol) {
                if (row < p) X[(int64_t)(F.pos0 + row) * nrhs + r0 + lr] = {c.re(v), c.im(v)};
                else W[(F.fv_off + row) * nrhs + r0 + lr] = {w2[v].x - c.re(v), w2[v].y - c.im(v)};
            }
        }
#pragma unroll
        for (int s = 0; s < 8; s++) av[s] = an[s];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            wb[v] = nb[v];
            wc[v] = nc[v];
        }
        okw = okn;
    }
}

// Backward: x1 = [Uop_bb, -Uop12] [y1; x2] for the p <= 32 pivot rows (two
// 16-row strips, K = m in chunks of 32).  No LDS: each lane loads its operand
// rows (k = k0 + 4s + lane/16, column lane%16) straight from X into MFMA layout;
// a chunk's factor entries and operand rows are requested together (the rows'
// sources one chunk ahead), so each chunk waits on memory once; occupancy is
// bounded by registers only.
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_front1(const cx<T> *__restrict__ arena, int64_t fstride,
                                                    const FrontDev *__restrict__ lvF, int nfr,
                                                    const int *__restrict__ idx, const int *__restrict__ rowperm,
                                                    cx<T> *__restrict__ X, int n, int nrhs, int rtg, int trans, int cj,
                                                    Sel sel) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
    int fi, tg;
    xcd_front_tile(blockIdx.x, rtg, fi, tg);
    const int r0 = (tg * (int)(blockDim.x >> 6) + wave) * WT;
    if (fi >= nfr || r0 >= nrhs) return;
    const int z = blockIdx.y, ncol = min(WT, nrhs - r0);
    const FrontDev F = lvF[fi];
    const int m = F.m, p = F.p;
    const cx<T> *G = arena + sel.f[z] * fstride + F.buf_off;
    X += (int64_t)sel.s[z] * n * nrhs;
    const int64_t ars = trans ? m : 1, acs = trans ? 1 : m;
    const bool two = p > 16, cin = lr < ncol;
    const AOp o0{ars, acs, p, m, 0, 0, 2, p, (bool)trans, (bool)cj};
    const AOp o1{ars, acs, p - 16, m, 16, 0, 2, p, (bool)trans, (bool)cj};
    // operand row k: pivot row k of the front, or the ancestors' row idx[k]
    // (the index load is unconditional: rows outside read the front's first)
    auto rows_of = [&](int k0, int (&r)[8]) {
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int k = k0 + 4 * s + lk;
            const int v = idx[F.idx_off + (k < m && k >= p ? k : 0)];
            // (unpredictable: kept a select, so the load is not moved behind a branch)
            r[s] = __builtin_unpredictable(k >= p && k < m) ? v : k < p ? F.pos0 + k : -1;
        }
    };
    cx<T> a0[8], a1[8];
    int rows[8];
    rows_of(0, rows);
    int dst[8];   // output positions of the lane's accumulator rows
    {
        // (transposed solves: the row permutation, loaded unconditionally)
        int rp[8];
        const int *rpf = rowperm + (int64_t)sel.f[z] * n + F.pos0;
#pragma unroll
        for (int v = 0; v < 8; v++) {
            const int row = (v >> 2) * 16 + mfma_t<T>::row(lk, v & 3);
            rp[v] = trans ? rpf[row < p ? row : 0] : row;
        }
#pragma unroll
        for (int v = 0; v < 8; v++) {
            const int row = (v >> 2) * 16 + mfma_t<T>::row(lk, v & 3);
            dst[v] = __builtin_unpredictable(row < p) ? F.pos0 + rp[v] : -1;
        }
    }
    cacc3<T> c0, c1;
    c0.zero();
    c1.zero();
    const int64_t col = r0 + (cin ? lr : 0);
    for (int k0 = 0; k0 < m; k0 += 32) {
        // the chunk's factor entries, operand rows and the next chunk's row
        // indices all requested before the first is used
        cx<T> b[8];
        load_a_raw<T>(G, o0, k0, a0);
        if (two) load_a_raw<T>(G + 16 * ars, o1, k0, a1);
        bool bok[8];
#pragma unroll
        for (int s = 0; s < 8; s++) {
            bok[s] = rows[s] >= 0 && cin;
            b[s] = X[(int64_t)max(rows[s], 0) * nrhs + col];
        }
        if (k0 + 32 < m) rows_of(k0 + 32, rows);
        finish_a<T>(o0, k0, a0);
        if (two) finish_a<T>(o1, k0, a1);
#pragma unroll
        for (int s = 0; s < 8; s++)
            if (!bok[s]) b[s] = {T(0), T(0)};
#pragma unroll
        for (int s = 0; s < 8; s++) {
            if (k0 + 4 * s >= m) break;
            const T bs = b[s].x + b[s].y;
            c0.mma(a0[s].x, a0[s].y, b[s].x, b[s].y, bs);
            if (two) c1.mma(a1[s].x, a1[s].y, b[s].x, b[s].y, bs);
        }
    }
    if (cin) {
#pragma unroll
        for (int v = 0; v < 8; v++) {
            if (dst[v] < 0) continue;
            const T xr = v < 4 ? c0.re(v) : c1.re(v - 4), xi = v < 4 ? c0.im(v) : c1.im(v - 4);
            X[(int64_t)dst[v] * nrhs + r0 + lr] = {xr, xi};
        }
    }
}

// ---- single-block levels with small fronts (m <= 64): staged operator -------
// The workgroup (up to 8 waves = consecutive 16-column tiles of one front)
// stages the front's whole operator and its rows' gather sources in LDS, with
// all loads in flight at once; then every wave loads its right-hand-side
// operand straight from global memory into registers in MFMA layout (k = 4s +
// lane/16, column lane%16: 16 lanes read one 128-byte row segment) and runs all
// its strips with the operator from LDS.  About three memory latencies per
// task, and the only LDS is the operator, so many tasks stay resident.
// Operator element (i, k) sits at A[k * as + i].
constexpr int kSmallM = 64;
// workgroup size limits (waves of 16-column tiles): 8 for the forward, 7 for
// the backward kernel -- at 200 right-hand sides (13 tiles) a front's panel is
// two workgroups, which measured faster than one of 13 to 16 waves (the two
// halves of a front run on two CUs) and than smaller groups
template <typename T>
constexpr int kSmallThreads = 512;
constexpr int kSmallBwdThreads = 448;

// operator row stride: rows rounded to 16, then to 16 or 48 mod 64, so the four
// k-rows of one operand read start in different banks
__host__ __device__ __forceinline__ int small_stride(int rows) {
    int as = (rows + 15) & ~15;
    if ((as & 63) == 0 || (as & 63) == 32) as += 16;
    return as;
}

// Forward.  CH: the level has fronts with children (contribution rows to add);
// the leaves' level has none and keeps no registers for them.  PS: pivot
// slots of four rows a lane gathers (the level's pivot blocks have at most 4 PS
// rows); levels with children run workgroups of kSmallChThreads.
// grid: (ceil(fronts/8)*8 * tile groups, frequency slots)
constexpr int kSmallChThreads = 256;
template <typename T, bool CH, int PS = 8>
__global__ void __launch_bounds__(CH ? kSmallChThreads : kSmallThreads<T>)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && !CH ? 8 : 1))) k_fwd_small1(const cx<T> *__restrict__ arena, int64_t fstride,
                                                    const FrontDev *__restrict__ lvF, int nfr,
                                                    const int *__restrict__ gsrc, int64_t gstride,
                                                    cx<T> *__restrict__ X, cx<T> *__restrict__ W, int n,
                                                    int64_t wstride, int nrhs, int rtg, int trans, int cj, Sel sel) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
    const int nth = blockDim.x;
    int fi, tg;
    xcd_front_tile(blockIdx.x, rtg, fi, tg);
    if (fi >= nfr) return;
    const int z = blockIdx.y, r0 = (tg * (nth >> 6) + wave) * WT, ncol = min(WT, nrhs - r0);
    const FrontDev F = lvF[fi];
    const int m = F.m, p = F.p, as = small_stride(m);
    T *Ar = reinterpret_cast<T *>(smem_raw), *Ai = Ar + p * as;
    int *src = reinterpret_cast<int *>(Ai + p * as);
    const cx<T> *G = arena + sel.f[z] * fstride + F.buf_off;
    X += (int64_t)sel.s[z] * n * nrhs;
    W += sel.s[z] * wstride;
    const int *gs = gsrc + sel.f[z] * gstride + F.fv_off * 3;
    const int64_t ars = trans ? m : 1, acs = trans ? 1 : m;
    const bool cin = lr < ncol;
    cx<T> b[8], w2[CH ? 16 : 1];
    if (!CH) {
        // no children: pivot row k's right-hand side is row pos0 + pi(k) of X, pi
        // the front's row pivoting (src[3k] = pos0 + pi(k)); the operator's
        // column k is staged at pi(k) instead, so the rows pos0.. are read in
        // place, before and alongside the staging loads
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int k = 4 * s + lk;
            b[s] = {T(0), T(0)};
            if (k < p && cin) b[s] = X[(int64_t)(F.pos0 + k) * nrhs + r0 + lr];
        }
    }
    // stage the operator [inv(L11); L21 inv(L11)] (or its transposed-solve form) and the sources
    // m * p <= 2048 entries: one pass at 512 threads or more.  The loads are
    // unconditional (entries past the operator read G[0] and are dropped), so
    // they are all in flight at once
    for (int e0 = 0; e0 < m * p; e0 += 4 * nth) {
        cx<T> v[4];
        int col[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = e0 + u * nth + threadIdx.x, k = e / m, i = e - k * m;
            const bool in = e < m * p;
            v[u] = G[in ? i * ars + k * acs : 0];
            col[u] = k;
            if (!CH) col[u] = gs[in ? 3 * k : 0] - F.pos0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = e0 + u * nth + threadIdx.x, k = e / m, i = e - k * m;
            if (cj) v[u].y = -v[u].y;
            if (i < p && k > i) v[u] = {T(0), T(0)};
            if (i == k && !trans) v[u] = {T(1), T(0)};
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = e0 + u * nth + threadIdx.x, i = e - (e / m) * m;
            if (e < m * p) {
                Ar[col[u] * as + i] = v[u].x;
                Ai[col[u] * as + i] = v[u].y;
            }
        }
    }
    if (CH)
        for (int e = threadIdx.x; e < 3 * m; e += nth) src[e] = gs[e];
    __syncthreads();
    if (r0 >= nrhs) return;
    // pivot rows in operand layout (k = 4s + lk), and the contribution rows the
    // lane's accumulators will meet in the epilogue (rows >= p; m <= 64: 4 strips)
    // The gathers load unconditionally (absent sources read row 0 and are
    // dropped by a select once everything has arrived), so the up to three
    // sources of every row are in flight together; only whole strips past the
    // front (a uniform test) are skipped.  Lanes past the right-hand sides read
    // column r0 (the row stride need not be a multiple of 16).
    if (CH) {
        const int cl = r0 + (cin ? lr : 0);
        // bit 3s + j: source j of pivot slot s exists; bit 24 + 2t + j: source
        // j of contribution slot t exists
        cx<T> xa[PS], wu[PS], ww[PS], cu[16], cw[16];
        uint64_t ok = 0;
#pragma unroll
        for (int s = 0; s < PS; s++) {
            if (4 * s >= p) break;
            const int k = 4 * s + lk, kc = k < p ? k : 0;
            const int a = src[3 * kc], u = src[3 * kc + 1], w = src[3 * kc + 2];
            ok |= (uint64_t)(a >= 0) << (3 * s) | (uint64_t)(u >= 0) << (3 * s + 1) |
                  (uint64_t)(w >= 0) << (3 * s + 2);
            xa[s] = X[(int64_t)max(a, 0) * nrhs + cl];
            wu[s] = W[(int64_t)max(u, 0) * nrhs + cl];
            ww[s] = W[(int64_t)max(w, 0) * nrhs + cl];
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (16 * q >= m) break;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int t = 4 * q + v, row = 16 * q + mfma_t<T>::row(lk, v), rc = row < m ? row : 0;
                const int u = src[3 * rc + 1], w = src[3 * rc + 2];
                ok |= (uint64_t)(u >= 0) << (24 + 2 * t) | (uint64_t)(w >= 0) << (25 + 2 * t);
                cu[t] = W[(int64_t)max(u, 0) * nrhs + cl];
                cw[t] = W[(int64_t)max(w, 0) * nrhs + cl];
            }
        }
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int k = 4 * s + lk;
            b[s] = {T(0), T(0)};
            if (s >= PS || 4 * s >= p) continue;
            if (k < p && cin) {
                if ((ok >> (3 * s)) & 1) b[s] = xa[s];
                if ((ok >> (3 * s + 1)) & 1) b[s] = cadd(b[s], wu[s]);
                if ((ok >> (3 * s + 2)) & 1) b[s] = cadd(b[s], ww[s]);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int t = 4 * q + v, row = 16 * q + mfma_t<T>::row(lk, v);
                cx<T> &o = w2[CH ? t : 0];
                o = {T(0), T(0)};
                if (16 * q >= m) continue;
                if (row >= p && row < m && cin) {
                    if ((ok >> (24 + 2 * t)) & 1) o = cu[t];
                    if ((ok >> (25 + 2 * t)) & 1) o = cadd(o, cw[t]);
                }
            }
    }
    T bs[8];   // br + bi of every operand entry, for the three-product complex MMA
#pragma unroll
    for (int s = 0; s < 8; s++) bs[s] = b[s].x + b[s].y;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (16 * q >= m) break;
        typename mfma_t<T>::acc cr, ci;
        {
            typename mfma_t<T>::acc t1 = {0, 0, 0, 0}, t2 = {0, 0, 0, 0}, t3 = {0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < 8; s++) {
                if (4 * s >= p) break;
                const int k = 4 * s + lk;
                const T ar = k < p ? Ar[k * as + 16 * q + lr] : T(0), ai = k < p ? Ai[k * as + 16 * q + lr] : T(0);
                cmma3<T>(ar, ai, b[s].x, b[s].y, bs[s], t1, t2, t3);
            }
            cmma3_finish<T>(t1, t2, t3, cr, ci);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = 16 * q + mfma_t<T>::row(lk, v);
            if (row < m && cin) {
                if (row < p) X[(int64_t)(F.pos0 + row) * nrhs + r0 + lr] = {cr[v], ci[v]};
                else if (CH)
                    W[(F.fv_off + row) * nrhs + r0 + lr] = {w2[CH ? 4 * q + v : 0].x - cr[v],
                                                            w2[CH ? 4 * q + v : 0].y - ci[v]};
                else W[(F.fv_off + row) * nrhs + r0 + lr] = {-cr[v], -ci[v]};
            }
        }
    }
}

// Backward.  Operator [p][m] staged as A[k * as + i] with as = p rounded to 16;
// right-hand-side operand: all m rows (K = m <= 64, 16 per lane) in registers.
template <typename T>
__global__ void __launch_bounds__(kSmallBwdThreads) k_bwd_small1(const cx<T> *__restrict__ arena, int64_t fstride,
                                                    const FrontDev *__restrict__ lvF, int nfr,
                                                    const int *__restrict__ idx, const int *__restrict__ rowperm,
                                                    cx<T> *__restrict__ X, int n, int nrhs, int rtg, int trans, int cj,
                                                    Sel sel) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
    const int nth = blockDim.x;
    int fi, tg;
    xcd_front_tile(blockIdx.x, rtg, fi, tg);
    if (fi >= nfr) return;
    const int z = blockIdx.y, r0 = (tg * (nth >> 6) + wave) * WT, ncol = min(WT, nrhs - r0);
    const FrontDev F = lvF[fi];
    const int m = F.m, p = F.p, as = (p + 15) & ~15;
    T *Ar = reinterpret_cast<T *>(smem_raw), *Ai = Ar + m * as;
    int *srow = reinterpret_cast<int *>(Ai + m * as);
    int *dstl = srow + kSmallM;
    const cx<T> *G = arena + sel.f[z] * fstride + F.buf_off;
    X += (int64_t)sel.s[z] * n * nrhs;
    const int64_t ars = trans ? m : 1, acs = trans ? 1 : m;
    // stage [inv(U11), -inv(U11) U12] (or its transposed-solve form), rows, output
    // positions; m * p <= 2048 entries: one or two passes at 448 threads
    // (unconditional loads: entries past the operator read G[0] and are dropped,
    // so a thread's loads and its row index (m <= 64 rows: one per thread) are
    // in flight together)
    const int ti = threadIdx.x;
    int srow_v = 0;
    for (int e0 = 0; e0 < m * p; e0 += 4 * nth) {
        cx<T> v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            // entries past the operator re-read its last one (every lane takes
            // the same path, so no load sits behind a branch)
            const int e = min(e0 + u * nth + (int)threadIdx.x, m * p - 1), k = e / p, i = e - k * p;
            v[u] = G[i * ars + k * acs];
        }
        if (e0 == 0) srow_v = idx[F.idx_off + (ti < m && ti >= p ? ti : 0)];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = e0 + u * nth + threadIdx.x, k = e / p, i = e - k * p;
            if (e < m * p) {
                cx<T> w = v[u];
                if (cj) w.y = -w.y;
                if (k < p && k < i) w = {T(0), T(0)};
                if (i == k && trans) w = {T(1), T(0)};
                if (k >= p) w = {-w.x, -w.y};
                Ar[k * as + i] = w.x;
                Ai[k * as + i] = w.y;
            }
        }
    }
    if (m * p == 0) srow_v = idx[F.idx_off + (ti < m && ti >= p ? ti : 0)];
    if (ti < m) srow[ti] = ti < p ? F.pos0 + ti : srow_v;
    for (int i = threadIdx.x; i < p; i += nth)
        dstl[i] = F.pos0 + (trans ? rowperm[(int64_t)sel.f[z] * n + F.pos0 + i] : i);
    __syncthreads();
    if (r0 >= nrhs) return;
    const bool cin = lr < ncol;
    cx<T> b[16];
#pragma unroll
    for (int s = 0; s < 16; s++) {
        const int k = 4 * s + lk;
        b[s] = {T(0), T(0)};
        if (k < m && cin) b[s] = X[(int64_t)srow[k] * nrhs + r0 + lr];
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
        if (16 * q >= p) break;
        cacc3<T, true> c;
        c.zero();
#pragma unroll
        for (int s = 0; s < 16; s++) {
            if (4 * s >= m) break;
            const int k = 4 * s + lk;
            const T ar = k < m ? Ar[k * as + 16 * q + lr] : T(0), ai = k < m ? Ai[k * as + 16 * q + lr] : T(0);
            c.mma(ar, ai, b[s].x, b[s].y, b[s].x + b[s].y);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = 16 * q + mfma_t<T>::row(lk, v);
            if (row < p && cin) X[(int64_t)dstl[row] * nrhs + r0 + lr] = {c.re(v), c.im(v)};
        }
    }
}

// backward sweep of a level's fronts: z = y1 - Uop12 x2 (x2 gathered from the
// ancestors' rows of X), blocks from the bottom, x written to the pivot rows
// (for transposed solves through P_f^T: the descendants' L21 rows are in this
// front's original row order).  grid as k_fwd_front.
template <typename T, int RW>
__global__ void __launch_bounds__(256) k_bwd_front(const cx<T> *__restrict__ arena, int64_t fstride,
                                                   const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                   int nfr, const int *__restrict__ idx,
                                                   const int *__restrict__ rowperm, cx<T> *__restrict__ X, int n,
                                                   int nrhs, int rt, int trans, int cj, Sel sel) {
    constexpr int RS = RW + 16, NC = RW / 16;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int fi, tile;
    xcd_front_tile(blockIdx.x, rt, fi, tile);
    if (fi >= nfr) return;
    const FrontDev F = fronts[list[fi]];
    const int z = blockIdx.y, r0 = tile * RW, ncol = min(RW, nrhs - r0);
    const int m = F.m, p = F.p, tid = threadIdx.x, wave = tid >> 6;
    T *Pr = reinterpret_cast<T *>(smem_raw), *Pi = Pr + m * RS;
    const cx<T> *G = arena + sel.f[z] * fstride + F.buf_off;
    X += (int64_t)sel.s[z] * n * nrhs;
    gather_bwd<T, RW>(F, idx, X, nrhs, r0, ncol, Pr, Pi);
    // op(U) entry (i, t) sits at G[i * ars + t * acs]
    const int64_t ars = trans ? m : 1, acs = trans ? 1 : m;
    const bool unit = trans;
    typename mfma_t<T>::acc cr[NC], ci[NC];
    for (int s0 = wave * 16; s0 < p && m > p; s0 += 64) {
        acc_zero<T, NC>(cr, ci);
        strip_mma<T, NC>(G + s0 * ars + p * acs, AOp{ars, acs, p - s0, m - p, 0, 0, 0, 1 << 30, false, (bool)cj}, Pr,
                         Pi, p, RS, cr, ci);
        acc_store<T, NC>(Pr, Pi, RS, s0, p, true, cr, ci);
    }
    __syncthreads();
    for (int b0 = ((p - 1) / BK) * BK; b0 >= 0; b0 -= BK) {
        const int kb = min(BK, p - b0), sr = wave * 16;
        const bool act = wave < 2 && sr < kb;
        if (act) {
            acc_zero<T, NC>(cr, ci);
            strip_mma<T, NC>(G + (b0 + sr) * ars + b0 * acs, AOp{ars, acs, kb - sr, kb, sr, 0, 2, 1 << 30, unit, (bool)cj},
                             Pr, Pi, b0, RS, cr, ci);
        }
        __syncthreads();
        if (act) acc_store<T, NC>(Pr, Pi, RS, b0 + sr, b0 + kb, false, cr, ci);
        __syncthreads();
        AOp ot{ars, acs, 0, kb, 0, 0, 0, 1 << 30, false, (bool)cj};
        cx<T> at[8], an[8];
        int s0 = sr;
        if (s0 < b0) {
            ot.arows = b0 - s0;
            load_a_raw<T>(G + s0 * ars + b0 * acs, ot, 0, at);
        }
        for (; s0 < b0; s0 += 64) {
            const int s1 = s0 + 64;
            if (s1 < b0) {
                AOp on = ot;
                on.arows = b0 - s1;
                load_a_raw<T>(G + s1 * ars + b0 * acs, on, 0, an);
            }
            ot.arows = b0 - s0;
            finish_a<T>(ot, 0, at);
            acc_zero<T, NC>(cr, ci);
            mma_a<T, NC>(at, 0, kb, Pr, Pi, b0, RS, cr, ci);
            acc_store<T, NC>(Pr, Pi, RS, s0, b0, true, cr, ci);
#pragma unroll
            for (int s = 0; s < 8; s++) at[s] = an[s];
        }
        __syncthreads();
    }
    const int *rp = rowperm + (int64_t)sel.f[z] * n + F.pos0;
    for (int e = tid; e < p * RW; e += 256) {
        const int i = e / RW, j = e - i * RW;
        if (j >= ncol) continue;
        X[(int64_t)(F.pos0 + (trans ? rp[i] : i)) * nrhs + r0 + j] = {Pr[i * RS + j], Pi[i * RS + j]};
    }
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
// forward sweep of one front for one right-hand side: gather (row-permuted for
// M solves), blocked lower solve (each 32-row diagonal block applied as a
// product with its inverse staged in LDS, block updates coalesced over rows),
// pivot rows to X and contribution rows to W for the parent
// (SINGLE: the level's fronts are in the single-block format, one product each;
// compiled apart so that this form carries no registers or LDS of the block loop)
template <typename T, bool SINGLE>
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
    if constexpr (SINGLE) {
        // one product over all rows: out = [Dop; Lop21'] w1.  Four lanes per row
        // (each sums every fourth pivot column, two xor-shuffles add them), 64
        // rows per pass, so every wave works and each lane waits on at most
        // eight loads (p <= 32); the results go to the second half of the
        // dynamic LDS (2 m entries) until every pass has read w1
        cx<T> *so = sw + m;
        const int ri = tid >> 2, part = tid & 3;
#pragma unroll 1
        for (int i0 = 0; i0 < m; i0 += 64) {
            const int i = i0 + ri;
            cx<T> acc = {T(0), T(0)};
            if (i < m)
#pragma unroll 4
                for (int t = part; t < p; t += 4) {
                    const cx<T> a = i < p ? dop(G, m, 0, i, t, true, trans, cj) : Lop(i, t);
                    acc = cadd(acc, cmul(a, sw[t]));
                }
            acc.x += __shfl_xor(acc.x, 1);
            acc.y += __shfl_xor(acc.y, 1);
            acc.x += __shfl_xor(acc.x, 2);
            acc.y += __shfl_xor(acc.y, 2);
            if (part == 0 && i < m) so[i] = acc;
        }
        __syncthreads();
        for (int i = tid; i < m; i += 256) sw[i] = i < p ? so[i] : csub(sw[i], so[i]);
        __syncthreads();
    } else {
        __shared__ cx<T> Dsh[32][33];
        for (int tb = 0; tb < p; tb += 32) {
            const int nb = min(32, p - tb);
            for (int e = tid; e < nb * nb; e += 256)
                Dsh[e % nb][e / nb] = dop(G, m, tb, e % nb, e / nb, true, trans, cj);
            __syncthreads();
            cx<T> y = {T(0), T(0)};
            if (tid < nb)
                for (int t = 0; t <= tid; t++) y = cadd(y, cmul(Dsh[tid][t], sw[tb + t]));
            __syncthreads();
            if (tid < nb) sw[tb + tid] = y;
            __syncthreads();
            // the factor loads of a row are unrolled so that several are in flight at once
            for (int i = tb + nb + tid; i < m; i += 256) {
                cx<T> acc = sw[i];
#pragma unroll 8
                for (int t = 0; t < nb; t++) acc = cmsub(acc, Lop(i, tb + t), sw[tb + t]);
                sw[i] = acc;
            }
            __syncthreads();
        }
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
                                                 int cj, int single) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *sw = reinterpret_cast<cx<T> *>(smem_raw);
    __shared__ cx<T> Dsh[32][33];
    const FrontDev F = fronts[list[blockIdx.x]];
    const int m = F.m, p = F.p, tid = threadIdx.x;
    const cx<T> *G = arena + F.buf_off;
    auto Uop = [&](int i, int t) { return trans ? ld(G + t + (int64_t)i * m, cj) : ld(G + i + (int64_t)t * m, cj); };
    for (int i = tid; i < m; i += 256) sw[i] = i < p ? X[F.pos0 + i] : X[idx[F.idx_off + i]];
    __syncthreads();
    if (single) {
        // x1 = [Dop, -Uop12'] [y1; x2]: 8 threads per pivot row (p <= 32), shuffle sum
        const int i = tid >> 3, part = tid & 7;
        cx<T> acc = {T(0), T(0)};
        if (i < p)
#pragma unroll 4
            for (int t = part; t < m; t += 8) {
                const cx<T> a = t < p ? dop(G, m, 0, i, t, false, trans, cj) : Uop(i, t);
                acc = t < p ? cadd(acc, cmul(a, sw[t])) : csub(acc, cmul(a, sw[t]));
            }
        for (int o = 4; o > 0; o >>= 1) {
            acc.x += __shfl_xor(acc.x, o, 8);
            acc.y += __shfl_xor(acc.y, o, 8);
        }
        if (i < p && part == 0) X[F.pos0 + (trans ? rowperm[F.pos0 + i] : i)] = acc;
        return;
    }
    for (int i = tid; i < p; i += 256) {
        cx<T> acc = sw[i];
#pragma unroll 8
        for (int s = p; s < m; s++) acc = cmsub(acc, Uop(i, s), sw[s]);
        sw[i] = acc;
    }
    __syncthreads();
    for (int tb = ((p - 1) / 32) * 32; tb >= 0; tb -= 32) {
        const int nb = min(32, p - tb);
        for (int e = tid; e < nb * nb; e += 256) Dsh[e % nb][e / nb] = dop(G, m, tb, e % nb, e / nb, false, trans, cj);
        __syncthreads();
        cx<T> x = {T(0), T(0)};
        if (tid < nb)
            for (int t = tid; t < nb; t++) x = cadd(x, cmul(Dsh[tid][t], sw[tb + t]));
        __syncthreads();
        if (tid < nb) sw[tb + tid] = x;
        __syncthreads();
        for (int i = tid; i < tb; i += 256) {
            cx<T> acc = sw[i];
#pragma unroll 8
            for (int t = 0; t < nb; t++) acc = cmsub(acc, Uop(i, tb + t), sw[tb + t]);
            sw[i] = acc;
        }
        __syncthreads();
    }
    // transposed solves: x = P_f^T w before any descendant reads it (descendant
    // L21 rows are in this front's original row order)
    for (int i = tid; i < p; i += 256) X[F.pos0 + (trans ? rowperm[F.pos0 + i] : i)] = sw[i];
}

// value of a front row from its gather-table entry g = (X position, W row,
// W row), -1 for none: X[a] + W[u] + W[w].  The three loads are unconditional
// (row 0 stands in for a missing one and is not added), so they are in flight
// together instead of one after another.
template <typename T>
__device__ __forceinline__ cx<T> gather3(const int *__restrict__ g, const cx<T> *__restrict__ X,
                                         const cx<T> *__restrict__ W) {
    const int a = g[0], u = g[1], w = g[2];
    const cx<T> xa = X[max(a, 0)], wu = W[max(u, 0)], ww = W[max(w, 0)];
    cx<T> v = a >= 0 ? xa : cx<T>{T(0), T(0)};
    if (u >= 0) v = cadd(v, wu);
    if (w >= 0) v = cadd(v, ww);
    return v;
}

// ---- single right-hand side, single-block levels of M solves: one wave per front ----
// In the single-block format a front's forward step is one product,
// [y1; w2 - L21 y1] = [inv(L11); L21 inv(L11)] w1, and its backward step one
// more, x1 = [inv(U11), -inv(U11) U12] [y1; x2].  A wave owns a front, so a
// compute unit holds many fronts at once and a front costs about three memory
// round trips with no workgroup barrier.  Columns of the operators are
// contiguous over rows (M solves; transposed solves keep k_fwd_vec/k_bwd_vec).
// A front of more than 64 rows takes several workgroups (grid y), each of which
// gathers the pivot rows' right-hand side from X, so y1 goes to the front's own
// W rows [0, p) (read by no other front's forward) and the backward reads it there.
template <typename T>
__device__ __forceinline__ cx<T> lane_bcast(cx<T> v, int l) {
    if constexpr (sizeof(T) == 4) {
        return {__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.x), l)),
                __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.y), l))};
    } else {
        auto rd = [l](double d) {
            const long long b = __double_as_longlong(d);
            const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
            const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
            return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
        };
        return {rd(v.x), rd(v.y)};
    }
}

// forward: a wave takes 64 rows of a front (grid.y: the row passes), lane i
// gathers front row i from the gather table -- X position, up to two
// children's contribution rows in W, after the front's row permutation -- and
// sums its row's p products; the pivot values w1[t] (t < p <= 32, gathered by
// lane t of every wave) reach every lane by readlane; PM >= p bounds the
// operator registers (levels of narrow fronts keep more waves resident).
// grid: (fronts, row passes), 64 threads
template <typename T, int PM>
__global__ void __launch_bounds__(64) k_fwd_vec1(const cx<T> *__restrict__ arena, const FrontDev *__restrict__ lvF,
                                                 const int *__restrict__ gsrc, cx<T> *__restrict__ X,
                                                 cx<T> *__restrict__ W) {
    const FrontDev F = lvF[blockIdx.x];
    const int m = F.m, p = F.p, lane = threadIdx.x, i = blockIdx.y * 64 + lane;
    if ((int)blockIdx.y * 64 >= m) return;
    const cx<T> *G = arena + F.buf_off;
    const int *gs = gsrc + (int64_t)F.fv_off * 3;
    auto gather = [&](int r) { return r < m ? gather3(gs + 3 * r, X, W) : cx<T>{T(0), T(0)}; };
    // the operator rows are in flight during the gathers; row i < p of
    // inv(L11) has a unit diagonal and zeros above it, which are not read;
    // rows past m read row m - 1 (in bounds) and are not stored
    cx<T> a[PM];
#pragma unroll
    for (int t = 0; t < PM; t++) {
        a[t] = {T(t == i), T(0)};
        if (t < p && (i >= p || t < i)) a[t] = G[min(i, m - 1) + (int64_t)t * m];
    }
    const cx<T> w0 = gather(lane);
    const cx<T> wi = blockIdx.y == 0 ? w0 : gather(i);
    cx<T> acc = {T(0), T(0)};
#pragma unroll
    for (int t = 0; t < PM; t++)
        if (t < p) acc = cadd(acc, cmul(a[t], lane_bcast(w0, t)));
    if (i < m) W[F.fv_off + i] = i < p ? acc : csub(wi, acc);
}

// backward: the NW waves of a workgroup share a front; its threads are R = (p
// rounded up to a power of two) pivot rows times 64 NW / R column parts, part q
// summing the columns q, q + 64 NW / R, ... in chunks of CH per thread; the
// values [y1; x2] (own pivot rows, then the ancestors' rows at the contribution
// positions) are staged in LDS while the first chunk of the operator is in
// flight.  A row's parts meet by xor-shuffles within a wave, then across the
// waves in LDS.  grid: fronts; LDS: m entries
template <typename T, int CH, int NW>
__global__ void __launch_bounds__(64 * NW) k_bwd_vec1(const cx<T> *__restrict__ arena,
                                                      const FrontDev *__restrict__ lvF, const int *__restrict__ idx,
                                                      const cx<T> *__restrict__ W, cx<T> *__restrict__ X) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *sv = reinterpret_cast<cx<T> *>(smem_raw);
    const FrontDev F = lvF[blockIdx.x];
    const int m = F.m, p = F.p, tid = threadIdx.x;
    if (p == 0) return;
    const cx<T> *G = arena + F.buf_off;
    int lr = 0;
    while ((1 << lr) < p) lr++;
    const int parts = (64 * NW) >> lr, i = tid & ((1 << lr) - 1), q = tid >> lr, step = CH * parts;
    // rows i >= p (up to the power of two) read row p - 1 in bounds, their sums
    // meet only each other; inv(U11)'s zeros left of the diagonal are not read
    const int ic = min(i, p - 1);
    cx<T> a[CH];
    auto fetch = [&](int t0) {
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int t = min(t0 + u * parts, m - 1);
            a[u] = {T(0), T(0)};
            if (t >= ic) a[u] = G[ic + (int64_t)t * m];
        }
    };
    fetch(q);
    // y1 from the front's W rows (k_fwd_vec1), x2 from X at the contribution positions
    for (int t = tid; t < m; t += 64 * NW) sv[t] = t < p ? W[F.fv_off + t] : X[idx[F.idx_off + t]];
    __syncthreads();
    cx<T> acc = {T(0), T(0)};
    for (int t0 = q; t0 < m;) {
#pragma unroll
        for (int u = 0; u < CH; u++) {
            // row i of inv(U11) is zero left of the diagonal; -inv(U11) U12 beyond p
            const int t = t0 + u * parts;
            if (t < m && t >= i) {
                const cx<T> v = sv[t];
                acc = t < p ? cadd(acc, cmul(a[u], v)) : csub(acc, cmul(a[u], v));
            }
        }
        t0 += step;
        if (t0 < m) fetch(t0);
    }
    for (int o = 1 << lr; o < 64; o <<= 1) {
        acc.x += __shfl_xor(acc.x, o);
        acc.y += __shfl_xor(acc.y, o);
    }
    if constexpr (NW > 1) {
        __shared__ cx<T> red[NW][32];
        if ((tid & 63) < (1 << lr)) red[tid >> 6][tid & 63] = acc;
        __syncthreads();
        if (tid < p) {
            cx<T> s = red[0][tid];
            for (int w = 1; w < NW; w++) s = cadd(s, red[w][tid]);
            X[F.pos0 + tid] = s;
        }
    } else {
        if (q == 0 && i < p) X[F.pos0 + i] = acc;
    }
}

// ---- single-vector sweeps of large fronts, split across workgroups ------------
// For levels of few, large multi-block fronts, k_fwd_vec/k_bwd_vec would leave
// the device nearly idle (one workgroup per front).  The triangular part over
// the p pivot rows stays with one workgroup per front; the rectangular part
// -- forward: contribution rows w2 -= L21 y1, backward: pivot rows
// z = y1 - U12 x2 -- runs as chunks of 256 rows on many workgroups.  Each row
// accumulates over t (or s) in the same order as the one-workgroup kernels,
// so the results are the same to the bit.
constexpr int kVecSplitRows = 256;   // rows per workgroup of the rectangular parts

// forward, triangular part: gather the front's rows, solve its p pivot rows
// block by block (updating only pivot rows), write y1 to X and the gathered
// contribution rows to W for k_fwd_vec_rect.  grid: fronts
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_vec_tri(const cx<T> *__restrict__ arena,
                                                     const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                     const int *__restrict__ children, const int *__restrict__ inv,
                                                     const int *__restrict__ rowperm, cx<T> *__restrict__ X,
                                                     cx<T> *__restrict__ W, int trans, int cj) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *sw = reinterpret_cast<cx<T> *>(smem_raw);
    __shared__ cx<T> Dsh[32][33];
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
        for (int e = tid; e < nb * nb; e += 256) Dsh[e % nb][e / nb] = dop(G, m, tb, e % nb, e / nb, true, trans, cj);
        __syncthreads();
        cx<T> y = {T(0), T(0)};
        if (tid < nb)
            for (int t = 0; t <= tid; t++) y = cadd(y, cmul(Dsh[tid][t], sw[tb + t]));
        __syncthreads();
        if (tid < nb) sw[tb + tid] = y;
        __syncthreads();
        for (int i = tb + nb + tid; i < p; i += 256) {
            cx<T> acc = sw[i];
#pragma unroll 8
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

// forward, rectangular part: contribution rows i of chunk blockIdx.x,
// w_i -= sum_t L(i, t) y_t over t = 0..p-1.  grid: (row chunks, fronts)
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_vec_rect(const cx<T> *__restrict__ arena,
                                                      const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                      const cx<T> *__restrict__ X, cx<T> *__restrict__ W, int trans,
                                                      int cj) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *y = reinterpret_cast<cx<T> *>(smem_raw);
    const FrontDev F = fronts[list[blockIdx.y]];
    const int m = F.m, p = F.p, tid = threadIdx.x;
    const int i = p + blockIdx.x * kVecSplitRows + tid;
    if (p + (int)blockIdx.x * kVecSplitRows >= m) return;
    const cx<T> *G = arena + F.buf_off;
    for (int t = tid; t < p; t += 256) y[t] = X[F.pos0 + t];
    __syncthreads();
    if (i >= m) return;
    cx<T> acc = W[F.fv_off + i];
#pragma unroll 8
    for (int t = 0; t < p; t++)
        acc = cmsub(acc, trans ? ld(G + t + (int64_t)i * m, cj) : ld(G + i + (int64_t)t * m, cj), y[t]);
    W[F.fv_off + i] = acc;
}

// backward, rectangular part: pivot rows i of chunk blockIdx.x,
// z_i = x_i - sum_{s >= p} U(i, s) x2_s, written over x_i.  grid: (row chunks, fronts)
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_vec_rect(const cx<T> *__restrict__ arena,
                                                      const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                      const int *__restrict__ idx, cx<T> *__restrict__ X, int trans,
                                                      int cj) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *x2 = reinterpret_cast<cx<T> *>(smem_raw);
    const FrontDev F = fronts[list[blockIdx.y]];
    const int m = F.m, p = F.p, tid = threadIdx.x;
    const int i = blockIdx.x * kVecSplitRows + tid;
    if ((int)blockIdx.x * kVecSplitRows >= p) return;
    const cx<T> *G = arena + F.buf_off;
    for (int s2 = tid; s2 < m - p; s2 += 256) x2[s2] = X[idx[F.idx_off + p + s2]];
    __syncthreads();
    if (i >= p) return;
    cx<T> acc = X[F.pos0 + i];
#pragma unroll 8
    for (int s2 = p; s2 < m; s2++)
        acc = cmsub(acc, trans ? ld(G + s2 + (int64_t)i * m, cj) : ld(G + i + (int64_t)s2 * m, cj), x2[s2 - p]);
    X[F.pos0 + i] = acc;
}

// backward, triangular part: the front's p pivot rows from z, block by block
// from the last; transposed solves apply the row permutation on the way out
// (as k_bwd_vec).  grid: fronts
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_vec_tri(const cx<T> *__restrict__ arena,
                                                     const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                                     const int *__restrict__ rowperm, cx<T> *__restrict__ X,
                                                     int trans, int cj) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *sw = reinterpret_cast<cx<T> *>(smem_raw);
    __shared__ cx<T> Dsh[32][33];
    const FrontDev F = fronts[list[blockIdx.x]];
    const int m = F.m, p = F.p, tid = threadIdx.x;
    const cx<T> *G = arena + F.buf_off;
    auto Uop = [&](int i, int t) { return trans ? ld(G + t + (int64_t)i * m, cj) : ld(G + i + (int64_t)t * m, cj); };
    for (int i = tid; i < p; i += 256) sw[i] = X[F.pos0 + i];
    __syncthreads();
    for (int tb = ((p - 1) / 32) * 32; tb >= 0; tb -= 32) {
        const int nb = min(32, p - tb);
        for (int e = tid; e < nb * nb; e += 256) Dsh[e % nb][e / nb] = dop(G, m, tb, e % nb, e / nb, false, trans, cj);
        __syncthreads();
        cx<T> x = {T(0), T(0)};
        if (tid < nb)
            for (int t = tid; t < nb; t++) x = cadd(x, cmul(Dsh[tid][t], sw[tb + t]));
        __syncthreads();
        if (tid < nb) sw[tb + tid] = x;
        __syncthreads();
        for (int i = tid; i < tb; i += 256) {
            cx<T> acc = sw[i];
#pragma unroll 8
            for (int t = 0; t < nb; t++) acc = cmsub(acc, Uop(i, tb + t), sw[tb + t]);
            sw[i] = acc;
        }
        __syncthreads();
    }
    for (int i = tid; i < p; i += 256) X[F.pos0 + (trans ? rowperm[F.pos0 + i] : i)] = sw[i];
}

// ---- single-vector sweeps through explicit inverses of the pivot triangles ----
// The levels of few multi-block fronts spend the one-workgroup kernels above
// in a chain of 32-row block steps per front.  With the pivot block's
// triangles inverted once (after a factorisation, on the first single-vector
// solve), every part of such a front's sweep is a product with a known matrix,
// split over many workgroups by rows.

// inverse squares: per front, column block of kVinvCb columns, lower (x <
// ncb) or upper triangle; the front's diagonal 32-blocks already hold their
// inverses (strictly lower: inv(L_bb) without its unit diagonal; upper:
// inv(U_bb)) and the blocks off the diagonal hold L11 and U11.  The column
// block of the inverse is solved block by block in LDS, W[p][kVinvCb + 1].
// grid (2 ncb, fronts with squares, nfreq)
constexpr int kVinvCb = 16;
template <typename T>
__global__ void __launch_bounds__(256) k_vinv(const cx<T> *__restrict__ arena, int64_t fstride,
                                              const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                              const int64_t *__restrict__ off, cx<T> *__restrict__ sq,
                                              int64_t sqstride, int ncb) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>(*Wl)[kVinvCb + 1] = reinterpret_cast<cx<T>(*)[kVinvCb + 1]>(smem_raw);
    const FrontDev F = fronts[list[blockIdx.y]];
    const int m = F.m, p = F.p, tid = threadIdx.x;
    const bool upper = (int)blockIdx.x >= ncb;
    const int c0 = ((int)blockIdx.x - (upper ? ncb : 0)) * kVinvCb;
    if (c0 >= p) return;
    const int nc = min(kVinvCb, p - c0);
    const cx<T> *G = arena + blockIdx.z * fstride + F.buf_off;
    cx<T> *S = sq + blockIdx.z * sqstride + off[blockIdx.y];
    const cx<T> one = {T(1), T(0)}, zero = {T(0), T(0)};
    // rows that the column block's inverse can reach: lower [b00, p), upper [0, c0 + nc)
    const int b00 = (c0 / 32) * 32, rlo = upper ? 0 : b00, rhi = upper ? c0 + nc : p;
    for (int e = tid; e < (rhi - rlo) * kVinvCb; e += 256) {
        const int r = rlo + e / kVinvCb, cc = e % kVinvCb;
        Wl[r][cc] = (cc < nc && r == c0 + cc) ? one : zero;
    }
    __syncthreads();
    if (!upper) {
        for (int b0 = b00; b0 < p; b0 += 32) {
            const int kb = min(32, p - b0);
            // y = inv(L_bb) w over the block's rows (unit diagonal), two outputs per thread
            cx<T> y[2];
            for (int u = 0; u < 2; u++) {
                const int e = tid + 256 * u, i = e / kVinvCb, cc = e % kVinvCb;
                y[u] = zero;
                if (i < kb) {
                    cx<T> a = Wl[b0 + i][cc];
                    for (int t = 0; t < i; t++) a = cmac(a, G[(b0 + i) + (int64_t)(b0 + t) * m], Wl[b0 + t][cc]);
                    y[u] = a;
                }
            }
            __syncthreads();
            for (int u = 0; u < 2; u++) {
                const int e = tid + 256 * u, i = e / kVinvCb, cc = e % kVinvCb;
                if (i < kb) Wl[b0 + i][cc] = y[u];
            }
            __syncthreads();
            // rows below the block: w_r -= L(r, block) y
            for (int e = tid; e < (p - b0 - kb) * kVinvCb; e += 256) {
                const int r = b0 + kb + e % (p - b0 - kb), cc = e / (p - b0 - kb);
                cx<T> a = Wl[r][cc];
                for (int t = 0; t < kb; t++) a = cmsub(a, G[r + (int64_t)(b0 + t) * m], Wl[b0 + t][cc]);
                Wl[r][cc] = a;
            }
            __syncthreads();
        }
        // strictly lower part of the square's columns c0.. (the diagonal stays inv(U)'s)
        for (int e = tid; e < (p - b00) * nc; e += 256) {
            const int r = b00 + e % (p - b00), cc = e / (p - b00), c = c0 + cc;
            if (r > c) S[r + (int64_t)c * p] = Wl[r][cc];
        }
    } else {
        for (int b0 = ((c0 + nc - 1) / 32) * 32; b0 >= 0; b0 -= 32) {
            const int kb = min(32, p - b0);
            // x = inv(U_bb) w over the block's rows (diagonal included)
            cx<T> y[2];
            for (int u = 0; u < 2; u++) {
                const int e = tid + 256 * u, i = e / kVinvCb, cc = e % kVinvCb;
                y[u] = zero;
                if (i < kb && b0 + i < rhi) {
                    cx<T> a = zero;
                    for (int t = i; t < kb && b0 + t < rhi; t++)
                        a = cmac(a, G[(b0 + i) + (int64_t)(b0 + t) * m], Wl[b0 + t][cc]);
                    y[u] = a;
                }
            }
            __syncthreads();
            for (int u = 0; u < 2; u++) {
                const int e = tid + 256 * u, i = e / kVinvCb, cc = e % kVinvCb;
                if (i < kb && b0 + i < rhi) Wl[b0 + i][cc] = y[u];
            }
            __syncthreads();
            // rows above the block: w_r -= U(r, block) x
            for (int e = tid; e < b0 * kVinvCb; e += 256) {
                const int r = e % b0, cc = e / b0;
                cx<T> a = Wl[r][cc];
                for (int t = 0; t < kb && b0 + t < rhi; t++)
                    a = cmsub(a, G[r + (int64_t)(b0 + t) * m], Wl[b0 + t][cc]);
                Wl[r][cc] = a;
            }
            __syncthreads();
        }
        for (int e = tid; e < rhi * nc; e += 256) {
            const int r = e % rhi, cc = e / rhi, c = c0 + cc;
            if (r <= c) S[r + (int64_t)c * p] = Wl[r][cc];
        }
    }
}

// the rectangles beside each square: RL = L21 inv(L11) ((m - p) x p) and
// RU = inv(U11) U12 (p x (m - p)), column-major after the square.  A workgroup
// makes one 64 x 64 tile, 4 x 4 entries per thread, the inner dimension in
// steps of 16 through LDS; the triangle's zero part is skipped (RL: from the
// tile's first column on, RU: from its first row on).  The triangles come from
// the square (inv(L11) has a unit diagonal), L21 and U12 from the front.
// grid (RL tiles then RU tiles, fronts with squares, nfreq)
template <typename T>
__global__ void __launch_bounds__(256) k_vrect(const cx<T> *__restrict__ arena, int64_t fstride,
                                               const FrontDev *__restrict__ fronts, const int *__restrict__ list,
                                               const int64_t *__restrict__ off, cx<T> *__restrict__ sq,
                                               int64_t sqstride) {
    __shared__ cx<T> As[16][65];   // [k][row]
    __shared__ cx<T> Bs[16][65];   // [k][column]
    const FrontDev F = fronts[list[blockIdx.y]];
    const int m = F.m, p = F.p, c = m - p, tid = threadIdx.x;
    const int tp = (p + 63) / 64, tc = (c + 63) / 64;
    int tile = blockIdx.x;
    const bool rl = tile < tp * tc;
    if (!rl) tile -= tp * tc;
    if (c == 0 || tile >= tp * tc) return;
    const cx<T> *G = arena + blockIdx.z * fstride + F.buf_off;
    cx<T> *Sq = sq + blockIdx.z * sqstride + off[blockIdx.y];
    const cx<T> *S = Sq;
    cx<T> *O = Sq + (int64_t)p * p + (rl ? 0 : (int64_t)c * p);
    const int I = rl ? c : p, J = rl ? p : c, ldo = I;
    const int ti = rl ? tile % tc : tile % tp, tj = rl ? tile / tc : tile / tp;
    const int i0 = ti * 64, j0 = tj * 64;
    // A(i, k): RL: L21(i, k) = G[p + i + k m]; RU: inv(U11)(i, k), cx<T>{T(0), T(0)} left of the diagonal
    auto A = [&](int i, int k) -> cx<T> {
        if (i >= I || k >= p) return cx<T>{T(0), T(0)};
        if (rl) return G[p + i + (int64_t)k * m];
        return k >= i ? S[i + (int64_t)k * p] : cx<T>{T(0), T(0)};
    };
    // B(k, j): RL: inv(L11)(k, j), unit diagonal, cx<T>{T(0), T(0)} above; RU: U12(k, j) = G[k + (p + j) m]
    auto B = [&](int k, int j) -> cx<T> {
        if (j >= J || k >= p) return cx<T>{T(0), T(0)};
        if (rl) return k > j ? S[k + (int64_t)j * p] : (k == j ? cx<T>{T(1), T(0)} : cx<T>{T(0), T(0)});
        return G[k + (int64_t)(p + j) * m];
    };
    const int ri = tid & 15, cjj = tid >> 4;
    cx<T> acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = cx<T>{T(0), T(0)};
    for (int k0 = rl ? j0 : i0; k0 < p; k0 += 16) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = tid + 256 * u;
            As[e >> 6][e & 63] = A(i0 + (e & 63), k0 + (e >> 6));   // 64 consecutive rows of one column
            Bs[e & 15][e >> 4] = B(k0 + (e & 15), j0 + (e >> 4));   // 16 consecutive k of one column
        }
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < 16; k++) {
            cx<T> av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; a++) av[a] = As[k][ri + 16 * a];
#pragma unroll
            for (int b = 0; b < 4; b++) bv[b] = Bs[k][cjj + 16 * b];
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++) acc[a][b] = cmac(acc[a][b], av[a], bv[b]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int i = i0 + ri + 16 * a, j = j0 + cjj + 16 * b;
            if (i < I && j < J) O[i + (int64_t)j * ldo] = acc[a][b];
        }
}

// a split level's single-vector sweep of a front as one product (VS_ROWS rows
// per workgroup):
//   forward   rows [0, m):  [y1; w2 - T21 y1] = [inv(T1); T21 inv(T1)] w1, to W rows [0, m)
//   backward  rows [0, p):  x1 = [inv(T1'), -inv(T1') T12] [y1; x2], to X (permuted for M^T, M^H)
// T1/T1' are the pivot block's lower/upper triangles of op(M): for M, L11
// (unit) forward and U11 backward, T21 inv(T1) = RL and inv(T1') T12 = RU; for
// M^T and M^H, U11^T forward and L11^T (unit) backward, with RU^T and RL^T,
// read transposed (and conjugated for M^H).  y1 stays in the front's W rows
// [0, p): other workgroups of the front read X's right-hand-side rows while
// this one writes.  The forward gathers its right-hand-side rows through the
// gather table (X position and up to two children's contribution rows): w1 into
// LDS (p entries), the contribution rows it updates itself; the backward stages
// [y1; x2] (m entries).  A workgroup takes RW rows -- one 128-byte line (8 rows
// in FP64, 16 in FP32) or more, for levels with few columns -- and 256 / RW
// column lanes: for M a load instruction reads whole lines of consecutive rows
// of a column (the storage is column-major), for the transposed operators runs
// of consecutive columns of one row.  Each lane's
// first VS_PRE operator entries are requested before the gather, so they
// arrive while it waits; the lanes' sums meet in LDS and are added in lane
// order.  No workgroup depends on another.  grid (row chunks, fronts)
template <typename T>
constexpr int VS_ROWS = 128 / (int)sizeof(cx<T>);
constexpr int VS_PRE = 8;
template <typename T, bool FWD, int RW, int NT>
__global__ void __launch_bounds__(NT) k_vsweep(const cx<T> *__restrict__ sq, const FrontDev *__restrict__ lvF,
                                               const int64_t *__restrict__ off, const int *__restrict__ gsrc,
                                               const int *__restrict__ idx, const int *__restrict__ rowperm,
                                               cx<T> *__restrict__ X, cx<T> *__restrict__ W, int trans, int cj) {
    constexpr int CL = NT / RW;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *wl = reinterpret_cast<cx<T> *>(smem_raw);
    const int fl = blockIdx.y;
    const FrontDev F = lvF[fl];
    const int m = F.m, p = F.p, c = m - p, tid = threadIdx.x;
    const int nrow = FWD ? m : p;
    const int r0 = blockIdx.x * RW;
    if (r0 >= nrow) return;
    const cx<T> *S = sq + off[fl], *RL = S + (int64_t)p * p, *RU = RL + (int64_t)c * p;
    const int *gs = gsrc + (int64_t)F.fv_off * 3;
    auto gather = [&](int i) { return gather3(gs + 3 * i, X, W); };
    // lane -> (row, column lane): rows fastest for M, columns fastest for the transposed operators
    const int lr = trans ? tid / CL : tid % RW, lc = trans ? tid % CL : tid / RW, r = r0 + lr;
    const bool live = r < nrow;
    auto a_at = [&](int t) -> cx<T> {
        if (FWD) {
            if (r < p) {   // inv(L11) strictly lower (M) or inv(U11)^T lower incl. diagonal
                if (!trans) return t < r ? S[r + (int64_t)t * p] : cx<T>{T(0), T(0)};
                return t <= r ? ld(S + t + (int64_t)r * p, cj) : cx<T>{T(0), T(0)};
            }
            const int i = r - p;   // RL row (M) or RU^T row
            return trans ? ld(RU + t + (int64_t)i * p, cj) : RL[i + (int64_t)t * c];
        }
        if (t < p) {   // inv(U11) upper incl. diagonal (M) or inv(L11)^T strictly upper
            if (!trans) return t >= r ? S[r + (int64_t)t * p] : cx<T>{T(0), T(0)};
            return t > r ? ld(S + t + (int64_t)r * p, cj) : cx<T>{T(0), T(0)};
        }
        const int s = t - p;   // -RU (M) or -RL^T
        const cx<T> v = trans ? ld(RL + s + (int64_t)r * c, cj) : RU[r + (int64_t)s * p];
        return {-v.x, -v.y};
    };
    // columns: the forward's triangle rows need only those up to the chunk's
    // last row, the backward's only those from its first
    int t0 = 0, t1 = FWD ? p : m;
    if (FWD && r0 + RW <= p) t1 = min(p, r0 + RW);
    if (!FWD) t0 = r0;
    cx<T> pre[VS_PRE];
#pragma unroll
    for (int u = 0; u < VS_PRE; u++) {
        const int t = t0 + lc + u * CL;
        pre[u] = (live && t < t1) ? a_at(t) : cx<T>{T(0), T(0)};
    }
    // the forward's own contribution row (rows >= p), fetched by the lane that stores it
    cx<T> own = {T(0), T(0)};
    if (FWD && tid < RW && r0 + tid >= p && r0 + tid < m) own = gather(r0 + tid);
    if (FWD) {
        for (int t = tid; t < t1; t += NT) wl[t] = gather(t);
    } else {
        for (int t = t0 + tid; t < m; t += NT) wl[t] = t < p ? W[F.fv_off + t] : X[idx[F.idx_off + t]];
    }
    __syncthreads();
    cx<T> acc = {T(0), T(0)};
#pragma unroll
    for (int u = 0; u < VS_PRE; u++) {
        const int t = t0 + lc + u * CL;
        if (t < t1) acc = cmac(acc, pre[u], wl[t]);
    }
    if (live)
#pragma unroll 8
        for (int t = t0 + lc + VS_PRE * CL; t < t1; t += CL) acc = cmac(acc, a_at(t), wl[t]);
    cx<T> s;
    if constexpr (NT == 256) {
        __shared__ cx<T> red[CL][RW + 1];
        red[lc][lr] = acc;
        __syncthreads();
        if (tid < RW) {
            s = red[0][tid];
            for (int k = 1; k < CL; k++) s = cadd(s, red[k][tid]);
        }
    } else {
        // M solves only (rows fastest): a row's column lanes within a wave are
        // RW apart; the waves' sums meet in LDS
        for (int o = RW; o < 64; o <<= 1) {
            acc.x += __shfl_xor(acc.x, o);
            acc.y += __shfl_xor(acc.y, o);
        }
        __shared__ cx<T> redw[NT / 64][RW];
        if ((tid & 63) < RW) redw[tid >> 6][tid & 63] = acc;
        __syncthreads();
        if (tid < RW) {
            s = redw[0][tid];
            for (int k = 1; k < NT / 64; k++) s = cadd(s, redw[k][tid]);
        }
    }
    if (tid < RW) {
        const int rr = r0 + tid;
        if (rr < nrow) {
            if (FWD) {
                if (rr < p) W[F.fv_off + rr] = trans ? s : cadd(wl[rr], s);
                else W[F.fv_off + rr] = csub(own, s);
            } else {
                X[F.pos0 + (trans ? rowperm[F.pos0 + rr] : rr)] = trans ? cadd(wl[rr], s) : s;
            }
        }
    }
}

// ---- FP64 residuals ------------------------------------------------------------
// r = b - op(M) x, column-major n x nrhs panels; one thread per row keeps the
// row's (at most nine) entries in registers and walks the right-hand sides.
// The pattern is symmetric: row `row` of M has the columns listed in CSC
// column `row` (values through tq for M, directly for M^T / M^H).  With
// `partial`, each workgroup also leaves the sum of |r|^2 over its 256 rows for
// every column in partial[(z * gridDim.x + block) * nrhs + r] (wave sums through
// LDS, no atomics); k_norm_reduce adds them up.  grid: (ceil(n/256), slots).
__global__ void __launch_bounds__(256) k_residual(const cx<double> *__restrict__ vals, const int *__restrict__ colptr,
                                                  const int *__restrict__ rowind, const int64_t *__restrict__ tq,
                                                  int n, int nrhs, int sys, int64_t nnz,
                                                  const cx<double> *__restrict__ B, const cx<double> *__restrict__ X,
                                                  cx<double> *__restrict__ R, double *__restrict__ partial, Sel sel) {
    extern __shared__ double wsum[];   // [4][nrhs] when partial != nullptr
    vals += sel.f[blockIdx.y] * nnz;
    const int64_t off = (int64_t)sel.s[blockIdx.y] * n * nrhs;
    B += off;
    X += off;
    R += off;
    const int row = blockIdx.x * blockDim.x + threadIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool valid = row < n;
    cx<double> a[9];
    int c[9];
    const int q0 = valid ? colptr[row] : 0, cnt = valid ? colptr[row + 1] - q0 : 0;
#pragma unroll
    for (int t = 0; t < 9; t++) {
        if (t < cnt) {
            const int q = q0 + t;
            a[t] = sys == 0 ? vals[tq[q]] : vals[q];
            if (sys == 1) a[t].y = -a[t].y;
            c[t] = rowind[q];
        } else {
            a[t] = {0.0, 0.0};
            c[t] = valid ? row : 0;
        }
    }
    for (int r = 0; r < nrhs; r++) {
        cx<double> acc = {0.0, 0.0};
        if (valid) {
            const cx<double> *x = X + (int64_t)r * n;
            acc = B[(int64_t)r * n + row];
#pragma unroll
            for (int t = 0; t < 9; t++) acc = cmsub(acc, a[t], x[c[t]]);
            R[(int64_t)r * n + row] = acc;
        }
        if (partial) {
            double v = acc.x * acc.x + acc.y * acc.y;
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            if (lane == 0) wsum[wave * nrhs + r] = v;
        }
    }
    if (partial) {
        __syncthreads();
        for (int r = threadIdx.x; r < nrhs; r += blockDim.x)
            partial[((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * nrhs + r] =
                wsum[r] + wsum[nrhs + r] + wsum[2 * nrhs + r] + wsum[3 * nrhs + r];
    }
}

// out[sel.s[z] * nrhs + r] = sum over the nblk workgroups' partial sums. grid (nrhs, slots)
__global__ void __launch_bounds__(256) k_norm_reduce(const double *__restrict__ partial, int nblk, int nrhs,
                                                     double *__restrict__ out, Sel sel) {
    __shared__ double s[256];
    const int r = blockIdx.x, z = blockIdx.y;
    double acc = 0;
    for (int b = threadIdx.x; b < nblk; b += 256) acc += partial[((int64_t)z * nblk + b) * nrhs + r];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(int64_t)sel.s[z] * nrhs + r] = s[0];
}

// ---- row-major panel refinement ----------------------------------------------------
// The panel refinement updates the residual rather than the iterate:
// r_{k+1} = r_k - op(M) d_k in FP64, row-major in grid order (the right-hand
// sides of a point contiguous, row stride ld64), with every correction d_k
// kept in the factor-precision panel the sweeps produced it in (row-major in
// elimination order, row stride ld).  A refinement step is one pass that reads
// r_k at each point and d_k over the point's 3x3 neighbourhood and writes
// r_{k+1} twice: in FP64 for the next step and in the factor precision as the
// next correction solve's right-hand side.  x = d_0 + d_1 + ... is summed in
// FP64 once, by the pass that writes the caller's layout.  r_1 = b - op(M) d_0
// with d_0 exactly as stored, so the sum and the residuals agree to FP64
// rounding as when x itself is carried.

// B (user layout) -> P (factor precision, elimination order: row iperm[k],
// stride ld) on slot sel.s[z]; also the tile's sums of |b|^2 per column in
// bpart[(z * ktiles + ktile) * nrhs + r].  grid (tiles, slots), 64 x 32 tiles.
template <typename T>
__global__ void __launch_bounds__(256) k_ir_in(const cx<double> *__restrict__ B, cx<T> *__restrict__ P,
                                               const int *__restrict__ iperm, int n, int nrhs, int ld,
                                               double *__restrict__ bpart, Sel sel) {
    __shared__ cx<double> S[PT_R][PT_K + 1];
    const int ktiles = (n + PT_K - 1) / PT_K, kt = blockIdx.x % ktiles;
    const int k0 = kt * PT_K, r0 = (blockIdx.x / ktiles) * PT_R;
    const int s = sel.s[blockIdx.y];
    B += (int64_t)s * n * nrhs;
    P += (int64_t)s * n * ld;
    {
        // all the tile's loads in flight at once (out-of-range entries read B[0] and are dropped)
        constexpr int PER = PT_K * PT_R / 256;
        cx<double> v[PER];
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int e = threadIdx.x + 256 * u, kk = e % PT_K, rr = e / PT_K;
            const bool ok = k0 + kk < n && r0 + rr < nrhs;
            v[u] = B[ok ? k0 + kk + (int64_t)(r0 + rr) * n : 0];
            if (!ok) v[u] = {0.0, 0.0};
        }
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int e = threadIdx.x + 256 * u;
            S[e / PT_K][e % PT_K] = v[u];
        }
    }
    __syncthreads();
    {
        // 8 lanes per column
        const int rr = threadIdx.x >> 3, part = threadIdx.x & 7;
        double acc = 0.0;
        for (int kk = part; kk < PT_K; kk += 8) acc += S[rr][kk].x * S[rr][kk].x + S[rr][kk].y * S[rr][kk].y;
        acc += __shfl_xor(acc, 1);
        acc += __shfl_xor(acc, 2);
        acc += __shfl_xor(acc, 4);
        if (part == 0 && r0 + rr < nrhs) bpart[((int64_t)blockIdx.y * ktiles + kt) * nrhs + r0 + rr] = acc;
    }
    for (int e = threadIdx.x; e < PT_K * PT_R; e += 256) {
        const int rr = e % PT_R, kk = e / PT_R;
        if (k0 + kk >= n) continue;
        const cx<double> v = S[rr][kk];
        if (r0 + rr < ld) P[(int64_t)iperm[k0 + kk] * ld + r0 + rr] = {(T)v.x, (T)v.y};
    }
}

// One refinement step of launch slot z (frequency sel.f[z], slot s = sel.s[z]):
// r <- r - op(M) D, with r read from Bu (the caller's b, user layout: point k
// of column c at k + c*n) on the first step and from Rg otherwise;  Rg <- r,
// R <- r in the factor precision (zero padding columns), and the sum of |r|^2
// over the workgroup's points of every column in
// partial[(z * gridDim.x + block) * nrhs + c].  Rg is in grid order (row
// k = i + j*nx, stride ld64; each point read and written by its own thread
// only), D and R in elimination order (row iperm[k], stride ld); D and R must
// be distinct buffers (other workgroups read a point's d while its own
// workgroup writes r).  Row k of M has the entries of CSC column k (the
// pattern is symmetric), all within the 3x3 neighbourhood.  Tiles of kIrTx x
// kIrTy points (8 consecutive points of four grid rows): per 16-column chunk
// the d of the (kIrTy + 2) x (kIrTx + 2) neighbourhood points is staged in LDS
// once (each point's d is used by up to nine rows; an 8 x 4 tile stages 1.9
// points per output), then one thread per (point, column) applies the stencil
// from LDS.  grid (blocks, slots), grid-stride over tiles.
constexpr int kIrTx = 8, kIrTy = 4, kIrCw = 16, kIrPts = kIrTx * kIrTy, kIrStage = (kIrTy + 2) * (kIrTx + 2);
constexpr int kIrThreads = kIrPts * kIrCw, kIrWaves = kIrThreads / 64;
constexpr int kIrCoef0 = (kIrStage + 63) / 64 * 64;   // first thread of the stencil-entry gather
static_assert(kIrStage <= kIrCoef0 && kIrCoef0 + 9 * kIrPts <= kIrThreads,
              "one thread per staged point and per stencil entry of the tile");
constexpr int kIrPer = (kIrStage * kIrCw + kIrThreads - 1) / kIrThreads;   // staging entries per thread

template <typename T>
__global__ void __launch_bounds__(kIrThreads) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 6 : 4))) k_ir_step(const cx<double> *__restrict__ vals, const int *__restrict__ colptr,
                                                 const int *__restrict__ rowind, const int64_t *__restrict__ tq,
                                                 const int *__restrict__ iperm, int nx, int ny, int nrhs, int ld64,
                                                 int ld, int sys, int64_t nnz, const cx<double> *__restrict__ Bu,
                                                 cx<double> *Rg, const cx<T> *__restrict__ D,
                                                 cx<T> *__restrict__ R, double *__restrict__ partial, Sel sel) {
    extern __shared__ double wsum[];   // [kIrWaves][nrhs]
    __shared__ cx<T> sx[kIrStage][kIrCw];
    __shared__ cx<double> scoef[kIrPts][9];
    __shared__ int sidx[kIrPts][9];    // staged row of each stencil entry
    __shared__ int sg[kIrStage];       // grid row of each staged point (-1: outside the grid)
    __shared__ int se[kIrStage];       // its elimination position
    const int z = blockIdx.y, s = sel.s[z];
    const int64_t n = (int64_t)nx * ny;
    vals += sel.f[z] * nnz;
    const int64_t o64 = s * n * ld64, o32 = s * n * ld;
    Rg += o64;
    const bool first = Bu != nullptr;
    if (first) Bu += s * n * nrhs;
    D += o32;
    R += o32;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double *ws = wsum + wave * nrhs;
    for (int c = lane; c < nrhs; c += 64) ws[c] = 0.0;
    // XCD-aware tile order: workgroup b runs on XCD b % 8 (gridDim.x is a
    // multiple of 8); each XCD takes a contiguous range of tiles in column-major
    // order (j fastest), so the tiles above and below a tile -- which stage the
    // same grid rows -- run on the same XCD and meet that XCD's L2.
    const int tpr = (nx + kIrTx - 1) / kIrTx, nyt = (ny + kIrTy - 1) / kIrTy, ntiles = tpr * nyt;
    const int xcd = blockIdx.x % 8, per = (ntiles + 7) / 8, t1 = min(ntiles, (xcd + 1) * per);
    for (int tt = xcd * per + blockIdx.x / 8; tt < t1; tt += gridDim.x / 8) {
        const int j = (tt % nyt) * kIrTy, i0 = (tt / nyt) * kIrTx, w = min(kIrTx, nx - i0), hgt = min(kIrTy, ny - j);
        __syncthreads();   // the previous tile's staging is no longer read
        if (tid < kIrStage) {
            const int jj = j - 1 + tid / (kIrTx + 2), ii = i0 - 1 + tid % (kIrTx + 2);
            const bool in = jj >= 0 && jj < ny && ii >= 0 && ii < nx;
            sg[tid] = in ? jj * nx + ii : -1;
            se[tid] = in ? iperm[jj * nx + ii] : 0;
        } else if (tid >= kIrCoef0 && tid < kIrCoef0 + kIrPts * 9) {
            const int r = (tid - kIrCoef0) / 9, t = (tid - kIrCoef0) % 9, ry = r / kIrTx, rx = r % kIrTx;
            int li = (ry + 1) * (kIrTx + 2) + rx + 1;   // missing entries: zero value on the point itself
            cx<double> a = {0.0, 0.0};
            if (rx < w && ry < hgt) {
                const int k = (j + ry) * nx + i0 + rx, q0 = colptr[k];
                if (t < colptr[k + 1] - q0) {
                    const int q = q0 + t;
                    a = sys == 0 ? vals[tq[q]] : vals[q];
                    if (sys == 1) a.y = -a.y;
                    const int kk = rowind[q], jj = kk / nx, ii = kk - jj * nx;
                    li = (jj - j + 1) * (kIrTx + 2) + (ii - i0 + 1);
                }
            }
            sidx[r][t] = li;
            scoef[r][t] = a;
        }
        __syncthreads();
        const int r = tid >> 4, cc = tid & 15, ry = r / kIrTx, rx = r % kIrTx;
        const bool pin = rx < w && ry < hgt;   // the thread's point is on the grid
        const int own_li = (ry + 1) * (kIrTx + 2) + rx + 1;
        const int64_t k = (int64_t)(j + ry) * nx + i0 + rx;
        const int64_t dst = pin ? (int64_t)se[own_li] * ld : 0;
        // the stencil row is read from LDS per chunk (16 lanes share each
        // address), not held in registers: fewer registers, more waves resident
        for (int c0 = 0; c0 < ld; c0 += kIrCw) {
            const int c = c0 + cc;
            const bool own = pin && c < nrhs;
            // every load of the chunk is issued before the first is used: r, then the
            // neighbourhood's d (a fixed number of staging entries per thread)
            cx<double> acc = {0.0, 0.0};
            if (own) acc = first ? Bu[k + (int64_t)c * n] : Rg[k * ld64 + c];
            cx<T> dv[kIrPer];
#pragma unroll
            for (int q = 0; q < kIrPer; q++) {
                const int e = tid + kIrThreads * q, row = e / kIrCw, cs = c0 + e % kIrCw;
                dv[q] = {T(0), T(0)};
                if (e < kIrStage * kIrCw && sg[row] >= 0 && cs < nrhs) dv[q] = D[(int64_t)se[row] * ld + cs];
            }
#pragma unroll
            for (int q = 0; q < kIrPer; q++) {
                const int e = tid + kIrThreads * q;
                if (e < kIrStage * kIrCw) sx[e / kIrCw][e % kIrCw] = dv[q];
            }
            __syncthreads();
            if (own) {
#pragma unroll
                for (int t = 0; t < 9; t++) {
                    const cx<T> d = sx[sidx[r][t]][cc];
                    acc = cmsub(acc, scoef[r][t], cx<double>{(double)d.x, (double)d.y});
                }
                Rg[k * ld64 + c] = acc;
                R[dst + c] = {(T)acc.x, (T)acc.y};
            } else if (pin && c < ld) {
                R[dst + c] = {T(0), T(0)};
            }
            // the wave's four points of a column: sum, then one LDS update per column
            double v = acc.x * acc.x + acc.y * acc.y;
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            if (lane < 16 && c < nrhs) ws[c] += v;
            __syncthreads();   // sx is restaged for the next chunk
        }
    }
    __syncthreads();
    for (int c = tid; c < nrhs; c += blockDim.x) {
        double v = 0.0;
        for (int q = 0; q < kIrWaves; q++) v += wsum[q * nrhs + c];
        partial[((int64_t)z * gridDim.x + blockIdx.x) * nrhs + c] = v;
    }
}

// corrections kept apart by the panel refinement before they are folded into
// one FP64 sum: buffer j holds correction j, j + kIrBufs, ... of every slot
constexpr int kIrBufs = 8;
struct IrBufs {
    const void *p[kIrBufs];
};
struct IrCounts {
    int c[kMaxBatch];
};

// acc (FP64, elimination order, stride ld) <- (bit z of `had`: acc +) the sum
// of the kIrBufs correction buffers, on the slots of sel.  grid (blocks, slots)
template <typename T>
__global__ void __launch_bounds__(256) k_ir_fold(IrBufs d, cx<double> *__restrict__ acc, uint64_t had, int64_t count,
                                                 Sel sel) {
    const int z = blockIdx.y;
    const int64_t o = sel.s[z] * count;
    for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < count; e += (int64_t)gridDim.x * 256) {
        cx<double> v = ((had >> z) & 1) ? acc[o + e] : cx<double>{0.0, 0.0};
#pragma unroll
        for (int j = 0; j < kIrBufs; j++) {
            const cx<T> w = static_cast<const cx<T> *>(d.p[j])[o + e];
            v.x += (double)w.x;
            v.y += (double)w.y;
        }
        acc[o + e] = v;
    }
}

// X (user layout) on slot sel.s[z] <- the FP64 sum of the first cnt.c[z]
// correction buffers (elimination order, row iperm[k], stride ld), plus acc
// when bit z of `folded` is set.  grid (tiles, slots), 64 x 32 tiles.
template <typename T>
__global__ void __launch_bounds__(256) k_ir_sum_out(IrBufs d, IrCounts cnt, const cx<double> *__restrict__ acc,
                                                    uint64_t folded, cx<double> *__restrict__ X,
                                                    const int *__restrict__ iperm, int n, int nrhs, int ld, Sel sel) {
    __shared__ cx<double> S[PT_R][PT_K + 1];
    constexpr int PER = PT_K * PT_R / 256;
    const int ktiles = (n + PT_K - 1) / PT_K;
    const int k0 = (blockIdx.x % ktiles) * PT_K, r0 = (blockIdx.x / ktiles) * PT_R;
    const int z = blockIdx.y, s = sel.s[z], nb = cnt.c[z];
    const bool fa = (folded >> z) & 1;
    const int64_t o = (int64_t)s * n * ld;
    X += (int64_t)s * n * nrhs;
    // every load of the tile is issued before the first is used: the rows'
    // positions, then all buffers' entries (out-of-range entries read entry o
    // of the slot and are dropped)
    int64_t q[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int e = threadIdx.x + 256 * u, rr = e % PT_R, kk = e / PT_R;
        const bool ok = k0 + kk < n && r0 + rr < nrhs;
        q[u] = ok ? o + (int64_t)iperm[k0 + kk] * ld + r0 + rr : o;
    }
    cx<double> v[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) v[u] = fa ? acc[q[u]] : cx<double>{0.0, 0.0};
#pragma unroll
    for (int j = 0; j < kIrBufs; j++)
        if (j < nb) {
            const cx<T> *dj = static_cast<const cx<T> *>(d.p[j]);
#pragma unroll
            for (int u = 0; u < PER; u++) {
                const cx<T> w = dj[q[u]];
                v[u].x += (double)w.x;
                v[u].y += (double)w.y;
            }
        }
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int e = threadIdx.x + 256 * u;
        S[e % PT_R][e / PT_R] = v[u];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < PT_K * PT_R; e += 256) {
        const int kk = e % PT_K, rr = e / PT_K;
        if (k0 + kk < n && r0 + rr < nrhs) X[k0 + kk + (int64_t)(r0 + rr) * n] = S[rr][kk];
    }
}

// Correction solves on the right-hand sides still above the tolerance only:
// they are packed into a narrower panel (ldc columns) for the solve and
// expanded back, the converged columns receiving a zero correction.
// cmap[z * ldc + j]: source column of packed column j on launch slot z (-1:
// zero padding); imap[z * ld + c]: packed column of column c (-1: zero).
// grid (blocks, slots), grid-stride over the slot's panel entries.
template <typename T>
__global__ void __launch_bounds__(256) k_ir_pack(const cx<T> *__restrict__ P, cx<T> *__restrict__ C,
                                                 const int *__restrict__ cmap, int n, int ld, int ldc, Sel sel) {
    const int z = blockIdx.y, s = sel.s[z];
    const cx<T> *Ps = P + (int64_t)s * n * ld;
    cx<T> *Cs = C + (int64_t)s * n * ldc;
    const int *cm = cmap + z * ldc;
    const int64_t total = (int64_t)n * ldc;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t row = e / ldc;
        const int c = cm[e - row * ldc];
        Cs[e] = c >= 0 ? Ps[row * ld + c] : cx<T>{T(0), T(0)};
    }
}
template <typename T>
__global__ void __launch_bounds__(256) k_ir_unpack(const cx<T> *__restrict__ C, cx<T> *__restrict__ P,
                                                   const int *__restrict__ imap, int n, int ld, int ldc, Sel sel) {
    const int z = blockIdx.y, s = sel.s[z];
    const cx<T> *Cs = C + (int64_t)s * n * ldc;
    cx<T> *Ps = P + (int64_t)s * n * ld;
    const int *im = imap + z * ld;
    const int64_t total = (int64_t)n * ld;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t row = e / ld;
        const int j = im[e - row * ld];
        Ps[e] = j >= 0 ? Cs[row * ldc + j] : cx<T>{T(0), T(0)};
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
            p = dev_alloc<char>(b);
            bytes = b;
        }
        return p;
    }
    void release() {
        dev_free(p);
        p = nullptr;
        bytes = 0;
    }
    ~Workspace() { dev_free(p); }
};
thread_local Workspace ws_x, ws_w, ws_r;
thread_local Workspace ws_n, ws_p;

// fused per-front kernels: widest right-hand-side tile whose panel of the
// level's largest front fits in LDS (0: the level takes the blocked kernels).
// Tuning (set_solve_tuning, or GFDFD_FUSED_RW at start-up): the widest tile
// (64, 32 or 16) and the largest front order a multi-block level may have to
// take the fused kernels.  Single-block levels always take them.
constexpr size_t kFusedLdsMax = 160 * 1024 - 1024;

template <typename T>
size_t fused_bytes(int rw) {
    return rw == 64 ? fused_row_bytes<T, 64>() : rw == 32 ? fused_row_bytes<T, 32>() : fused_row_bytes<T, 16>();
}

struct SolveTuning {
    // multi-block levels take the per-block kernels by default: measured faster
    // than the fused per-front kernels on the 1601x401 workload in FP32
    int rw_cap = 64, fused_mmax = 0;
    int refine_layout = 0;   // panel refinement: 0 row-major (k_ir_step), 1 user layout
    SolveTuning() {
        if (const char *e = std::getenv("GFDFD_FUSED_RW")) {
            const int v = std::atoi(e);
            if (v == 16 || v == 32 || v == 64) rw_cap = v;
        }
    }
};
SolveTuning &tuning() {
    static SolveTuning t;
    return t;
}
int refine_mode() { return tuning().refine_layout; }

template <typename T>
int fused_rw(int mmax, int max_children, bool single) {
    if (max_children > 2 || (!single && mmax > tuning().fused_mmax)) return 0;
    for (int rw = tuning().rw_cap; rw >= 16; rw /= 2)
        if ((size_t)mmax * fused_bytes<T>(rw) <= kFusedLdsMax) return rw;
    return 0;
}

// wave-task launch of a single-block level: one wave (a 16-column tile of one
// front) per workgroup, with that wave's LDS
struct WaveLaunch {
    dim3 grid, block;
    size_t lds;
    int rtg, wave_lds;
};
template <typename T>
WaveLaunch wave_launch(int nfr, int nrhs, int nslots, size_t per_wave) {
    const size_t wl = (per_wave + 15) & ~size_t(15);
    if (wl > kFusedLdsMax) throw std::logic_error("single-block front too large for a wave task");
    const int rt = (nrhs + WT - 1) / WT;
    // one wave per workgroup: measured faster than grouping 2 or 4 tiles of a
    // front (the tasks share nothing but the factor entries, which L2 serves)
    const int wpg = 1;
    const int rtg = rt;
    return {dim3(((nfr + 7) / 8) * 8 * rtg, nslots), dim3(64 * wpg), wpg * wl, rtg, (int)wl};
}

// staged small-front launch: the workgroup's LDS is the staged operator and
// sources only; up to max_waves waves (16-column tiles), as many workgroups as fit
inline WaveLaunch small_launch(int nfr, int nrhs, int nslots, size_t lds, int max_waves) {
    const int rt = (nrhs + WT - 1) / WT;
    const int wpg = std::min(max_waves, rt);
    const int rtg = (rt + wpg - 1) / wpg;
    return {dim3(((nfr + 7) / 8) * 8 * rtg, nslots), dim3(64 * wpg), (lds + 15) & ~size_t(15), rtg, 0};
}

#define FUSED_LAUNCH(kern, T, rw, grid, lds, st, ...)                                                   \
    do {                                                                                              \
        if ((rw) == 64) hipLaunchKernelGGL((kern<T, 64>), grid, dim3(256), lds, st, __VA_ARGS__);     \
        else if ((rw) == 32) hipLaunchKernelGGL((kern<T, 32>), grid, dim3(256), lds, st, __VA_ARGS__); \
        else hipLaunchKernelGGL((kern<T, 16>), grid, dim3(256), lds, st, __VA_ARGS__);                \
    } while (0)

template <typename T>
void set_smem_attrs() {
    static bool done = false;
    if (done) return;
    // the vector kernels also hold a static 32 x 33 diagonal block
    const int big = 160 * 1024 - (int)(32 * 33 * sizeof(cx<T>));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_fwd_vec<T, true>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, big));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_fwd_vec<T, false>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, big));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bwd_vec<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, big));
    const int blk = (int)blk_smem_bytes<T>();
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_fwd_blk<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)blk_smem_bytes<T, kFwdRT>()));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_fwd_blk2<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)blk2_smem_bytes<T>()));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bwd_blk<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, blk));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bwd_contrib<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, blk + 4 * kContribRowsMax));
    // the refinement step's per-wave column sums: up to 64 KB at 1024 right-hand sides
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_ir_step<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, kIrWaves * 8 * 1024));
    const int fl = (int)kFusedLdsMax;
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_fwd_front<T, 64>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, fl));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_fwd_front<T, 32>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, fl));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_fwd_front<T, 16>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, fl));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bwd_front<T, 64>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, fl));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bwd_front<T, 32>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, fl));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bwd_front<T, 16>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, fl));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_fwd_small1<T, true>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, fl));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_fwd_small1<T, true, 3>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, fl));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_fwd_small1<T, false>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, fl));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bwd_small1<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, fl));
    GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_fwd_front1<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, fl));
    done = true;
}

}  // namespace

void set_solve_tuning(int rw_cap, int fused_mmax, int refine_layout) {
    if (rw_cap != 16 && rw_cap != 32 && rw_cap != 64) throw std::invalid_argument("tile width must be 16, 32 or 64");
    if (refine_layout != 0 && refine_layout != 1) throw std::invalid_argument("refinement layout must be 0 or 1");
    tuning().rw_cap = rw_cap;
    tuning().refine_layout = refine_layout;
    tuning().fused_mmax = fused_mmax < 0 ? (1 << 30) : fused_mmax;
}

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

// Forward and backward sweeps of op(M_f) on internal panels (elimination
// order, row-major, right-hand sides contiguous): launch slot z solves
// frequency in.f[z] in place on panel in.s[z] of Xp, with work rows Wp.
// nrhs is the panels' row stride (right-hand sides plus zero padding); the
// algorithmic counts attached to the launches use nreal right-hand sides.
template <typename T>
void sweeps(const Factors<T> &F, const Sel &in, int sys, int nrhs, cx<T> *Xp, cx<T> *Wp, hipStream_t st,
            int nreal) {
    const DevPlan &dp = *F.plan;
    const Plan &pl = *dp.host;
    const int n = pl.n, ns = in.n;
    const bool trans = sys != 0, cj = sys == 1;
    const int64_t wstride = pl.fv_rows * (int64_t)nrhs;
    const double nr = nreal, nz = ns, es = sizeof(cx<T>);
    const size_t blk = blk_smem_bytes<T>();
    const int rt = (nrhs + 63) / 64;
    const bool prof = prof_enabled();
    for (int l = 0; l < pl.nlevels; l++) {
        if (dp.n_all[l] == 0) continue;
        const DevPlan::LevelCounts &c = dp.counts[l];
        const int nfr = dp.n_all[l];
        if (dp.single[l] && dp.mmax_all[l] <= kSmallM) {
            const bool ch = dp.lv_children[l];
            const WaveLaunch wl =
                small_launch(nfr, nrhs, ns, 2 * sizeof(T) * dp.pmax_all[l] * small_stride(dp.mmax_all[l]) + 12 * kSmallM,
                             (ch ? kSmallChThreads : kSmallThreads<T>) / 64);
            // levels without children (the leaves) are a kernel of their own
            ProfScope ps(ch ? K_FWD_SMALL1 : K_FWD_LEAF, st, nz * 8.0 * nr * c.tri_elems,
                         nz * (c.tri_elems + 2.0 * c.sum_m * nr) * es);
            if (ch && dp.pmax_all[l] <= 12)
                hipLaunchKernelGGL((k_fwd_small1<T, true, 3>), wl.grid, wl.block, wl.lds, st, F.arena, F.fstride,
                                   dp.lv_fr[l], nfr, trans ? dp.gsrc_id : F.gsrc, trans ? (int64_t)0 : pl.fv_rows * 3,
                                   Xp, Wp, n, wstride, nrhs, wl.rtg, (int)trans, (int)cj, in);
            else if (ch)
                hipLaunchKernelGGL((k_fwd_small1<T, true>), wl.grid, wl.block, wl.lds, st, F.arena, F.fstride,
                                   dp.lv_fr[l], nfr, trans ? dp.gsrc_id : F.gsrc, trans ? (int64_t)0 : pl.fv_rows * 3,
                                   Xp, Wp, n, wstride, nrhs, wl.rtg, (int)trans, (int)cj, in);
            else
                hipLaunchKernelGGL((k_fwd_small1<T, false>), wl.grid, wl.block, wl.lds, st, F.arena, F.fstride,
                                   dp.lv_fr[l], nfr, trans ? dp.gsrc_id : F.gsrc, trans ? (int64_t)0 : pl.fv_rows * 3,
                                   Xp, Wp, n, wstride, nrhs, wl.rtg, (int)trans, (int)cj, in);
            continue;
        }
        if (dp.single[l]) {
            const WaveLaunch wl = wave_launch<T>(nfr, nrhs, ns, dp.pmax_all[l] * WT * 2 * sizeof(T) + 12 * dp.mmax_all[l]);
            ProfScope ps(K_FWD_FRONT1, st, nz * 8.0 * nr * c.tri_elems, nz * (c.tri_elems + 2.0 * c.sum_m * nr) * es);
            hipLaunchKernelGGL(k_fwd_front1<T>, wl.grid, wl.block, wl.lds, st, F.arena, F.fstride, dp.lv_fr[l], nfr,
                               trans ? dp.gsrc_id : F.gsrc, trans ? (int64_t)0 : pl.fv_rows * 3, Xp, Wp, n, wstride,
                               nrhs, wl.rtg, wl.wave_lds, (int)trans, (int)cj, in);
            continue;
        }
        const int rw = fused_rw<T>(dp.mmax_all[l], dp.max_children, false);
        if (rw) {
            const int rtf = (nrhs + rw - 1) / rw;
            const dim3 grid(((nfr + 7) / 8) * 8 * rtf, ns);
            const size_t lds = (size_t)dp.mmax_all[l] * fused_bytes<T>(rw);
            ProfScope ps(K_FWD_FRONT, st, nz * 8.0 * nr * c.tri_elems, nz * (c.tri_elems + 2.0 * c.sum_m * nr) * es);
            FUSED_LAUNCH(k_fwd_front, T, rw, grid, lds, st, F.arena, F.fstride, dp.fronts, dp.lv_all[l], nfr,
                         dp.children, dp.inv, trans ? nullptr : F.rowperm, Xp, Wp, n, wstride, nrhs, rtf, (int)trans,
                         (int)cj, in);
            continue;
        }
        {
            ProfScope ps(K_FWD_GATHER, st, 0.0, nz * 2.0 * c.sum_m * nr * es);
            int64_t work = (int64_t)dp.mmax_all[l] * nrhs;
            int chunks = (int)std::min<int64_t>(std::max<int64_t>(1, work / 1024), 256);
            hipLaunchKernelGGL(k_fwd_gather<T>, dim3(nfr, chunks, ns), dim3(256), 0, st, Xp, Wp, dp.fronts,
                               dp.lv_all[l], trans ? dp.gsrc_id : F.gsrc, trans ? (int64_t)0 : pl.fv_rows * 3, n,
                               wstride, nrhs, in);
        }
        const int pmax = dp.pmax_all[l];
        static const bool one_block = std::getenv("GFDFD_FWD_BLK1") != nullptr;
        for (int b0 = 0; b0 < pmax; b0 += one_block ? BK : 2 * BK) {
            // factor entries of the block columns, and the panel rows below them read and written
            const int nbk = one_block || b0 + BK >= pmax ? 1 : 2;
            double fe = 0.0;
            if (prof)
                for (int q = 0; q < nbk; q++) fe += c.fwd_blk_elems[b0 / BK + q];
            ProfScope ps(K_FWD_BLK, st, nz * 8.0 * nr * fe, nz * (fe + 2.0 * nr * fe / (nbk * BK)) * es);
            const int tiles = std::max(1, (dp.mmax_all[l] - b0 - 1) / kFwdRT + 1);
            const size_t fblk = blk_smem_bytes<T, kFwdRT>();
            if (one_block)
                hipLaunchKernelGGL(k_fwd_blk<T>, dim3(tiles, nfr, rt * ns), dim3(4 * kFwdRT), fblk, st, F.arena,
                                   F.fstride, dp.fronts, dp.lv_all[l], Xp, Wp, n, wstride, nrhs, rt, b0, (int)trans,
                                   (int)cj, in);
            else
                hipLaunchKernelGGL(k_fwd_blk2<T>, dim3(tiles, nfr, rt * ns), dim3(4 * kFwdRT), blk2_smem_bytes<T>(), st,
                                   F.arena, F.fstride, dp.fronts, dp.lv_all[l], Xp, Wp, n, wstride, nrhs, rt, b0,
                                   (int)trans, (int)cj, in);
        }
    }
    for (int l = pl.nlevels - 1; l >= 0; l--) {
        if (dp.n_all[l] == 0) continue;
        const DevPlan::LevelCounts &c = dp.counts[l];
        const int nfr = dp.n_all[l];
        if (dp.single[l] && dp.mmax_all[l] <= kSmallM) {
            const WaveLaunch wl = small_launch(nfr, nrhs, ns,
                                               2 * sizeof(T) * dp.mmax_all[l] * 32 + 4 * (kSmallM + 32),
                                               kSmallBwdThreads / 64);
            ProfScope ps(K_BWD_SMALL1, st, nz * 8.0 * nr * c.tri_elems,
                         nz * (c.tri_elems + (c.sum_m + c.sum_p) * nr) * es);
            hipLaunchKernelGGL(k_bwd_small1<T>, wl.grid, wl.block, wl.lds, st, F.arena, F.fstride, dp.lv_fr[l], nfr,
                               dp.idx, F.rowperm, Xp, n, nrhs, wl.rtg, (int)trans, (int)cj, in);
            continue;
        }
        if (dp.single[l]) {
            // one wave per workgroup: measured faster than 2 or 4 (no LDS is shared)
            const WaveLaunch wl = small_launch(nfr, nrhs, ns, 0, 1);
            ProfScope ps(K_BWD_FRONT1, st, nz * 8.0 * nr * c.tri_elems,
                         nz * (c.tri_elems + (c.sum_m + c.sum_p) * nr) * es);
            hipLaunchKernelGGL(k_bwd_front1<T>, wl.grid, wl.block, 0, st, F.arena, F.fstride, dp.lv_fr[l], nfr, dp.idx,
                               F.rowperm, Xp, n, nrhs, wl.rtg, (int)trans, (int)cj, in);
            continue;
        }
        const int rw = fused_rw<T>(dp.mmax_all[l], dp.max_children, false);
        if (rw) {
            const int rtf = (nrhs + rw - 1) / rw;
            const dim3 grid(((nfr + 7) / 8) * 8 * rtf, ns);
            const size_t lds = (size_t)dp.mmax_all[l] * fused_bytes<T>(rw);
            ProfScope ps(K_BWD_FRONT, st, nz * 8.0 * nr * c.tri_elems,
                         nz * (c.tri_elems + (c.sum_m + c.sum_p) * nr) * es);
            FUSED_LAUNCH(k_bwd_front, T, rw, grid, lds, st, F.arena, F.fstride, dp.fronts, dp.lv_all[l], nfr, dp.idx,
                         F.rowperm, Xp, n, nrhs, rtf, (int)trans, (int)cj, in);
            continue;
        }
        const int pmax = dp.pmax_all[l];
        {
            ProfScope ps(K_BWD_CONTRIB, st, nz * 8.0 * nr * c.contrib_elems,
                         nz * (c.contrib_elems + (c.sum_m + c.sum_p) * nr) * es);
            if (dp.mmax_all[l] > kContribRowsMax) throw std::logic_error("front too large for the contribution kernel");
            hipLaunchKernelGGL(k_bwd_contrib<T>, dim3((pmax + 63) / 64, nfr, rt * ns), dim3(256),
                               blk + 4 * (size_t)dp.mmax_all[l], st, F.arena,
                               F.fstride, dp.fronts, dp.lv_all[l], dp.idx, Xp, Wp, n, wstride, nrhs, rt, (int)trans,
                               (int)cj, in);
        }
        for (int b0 = ((pmax - 1) / BK) * BK; b0 >= 0; b0 -= BK) {
            const double be = prof ? c.bwd_blk_elems[b0 / BK] : 0.0;
            ProfScope ps(K_BWD_BLK, st, nz * 8.0 * nr * be, nz * (be + 2.0 * nr * be / BK) * es);
            const int tiles = std::max(1, (b0 + 63) / 64);
            hipLaunchKernelGGL(k_bwd_blk<T>, dim3(tiles, nfr, rt * ns), dim3(256), blk, st, F.arena, F.fstride,
                               dp.fronts, dp.lv_all[l], Xp, Wp, n, wstride, nrhs, rt, b0, (int)trans, (int)cj, in);
        }
        if (trans) {
            ProfScope ps(K_PIV_UNSWAP, st, 0.0, nz * 4.0 * c.sum_p * nr * es);
            const int chunks = (int)std::min<int64_t>(std::max<int64_t>(1, (int64_t)pmax * nrhs / 1024), 256);
            hipLaunchKernelGGL(k_piv_copy<T>, dim3(nfr, chunks, ns), dim3(256), 0, st, dp.fronts, dp.lv_all[l], Xp,
                               Wp, n, wstride, nrhs, in);
            hipLaunchKernelGGL(k_piv_unswap<T>, dim3(nfr, chunks, ns), dim3(256), 0, st, dp.fronts, dp.lv_all[l],
                               F.rowperm, Xp, Wp, n, wstride, nrhs, in);
        }
    }
}

// the single-vector sweeps split a level's fronts across workgroups when the
// level is multi-block (some front has more than one 32-pivot block) and has
// too few fronts to fill the device with one workgroup each
static bool vec_split(const DevPlan &dp, int l) {
    static const bool off = std::getenv("GFDFD_VEC_NOSPLIT") != nullptr;
    return !off && !dp.single[l] && dp.n_all[l] < kVecSplitFronts;
}

// the inverse squares and rectangles of the split levels, made once after each
// factorisation (allocated on first use)
template <typename T>
static bool ensure_vinv(const Factors<T> &F, hipStream_t st) {
    static const bool off = std::getenv("GFDFD_VEC_NOINV") != nullptr;
    const DevPlan &dp = *F.plan;
    if (off || dp.vinv_n == 0) return false;
    if (!F.vinv) {
        F.vinv = dev_alloc<cx<T>>((size_t)dp.vinv_elems * F.nfreq);
        F.vinv_ready = false;
    }
    if (!F.vinv_ready) {
        const size_t lds = sizeof(cx<T>) * (size_t)dp.vinv_pmax * (kVinvCb + 1);
        static bool attr = false;
        if (!attr) {
            GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vinv<double>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)(sizeof(cx<double>) * kVinvPmax * (kVinvCb + 1))));
            GFDFD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vinv<float>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)(sizeof(cx<float>) * kVinvPmax * (kVinvCb + 1))));
            attr = true;
        }
        const int ncb = (dp.vinv_pmax + kVinvCb - 1) / kVinvCb;
        hipLaunchKernelGGL(k_vinv<T>, dim3(2 * ncb, dp.vinv_n, F.nfreq), dim3(256), lds, st, F.arena, F.fstride,
                           dp.fronts, dp.vinv_list, dp.vinv_list_off, F.vinv, dp.vinv_elems, ncb);
        if (dp.vinv_rect_tiles > 0)
            hipLaunchKernelGGL(k_vrect<T>, dim3(dp.vinv_rect_tiles, dp.vinv_n, F.nfreq), dim3(256), 0, st, F.arena,
                               F.fstride, dp.fronts, dp.vinv_list, dp.vinv_list_off, F.vinv, dp.vinv_elems);
        GFDFD_HIP(hipGetLastError());
        F.vinv_ready = true;
    }
    return true;
}

// per-level widths of the single-vector kernels: operator registers of the
// one-wave forward (the level's widest pivot block), operator entries per lane
// and round trip of the one-wave backward (8 when that covers every column of
// a lane), and rows per workgroup of the split products (a 128-byte line of
// rows, or more when the columns are few: about VS_PRE columns per lane, so
// each lane requests all its operator entries at once)
template <typename T>
static void launch_fwd_vec1(const DevPlan &dp, int l, const cx<T> *arena, const int *gsrc, cx<T> *Xp, cx<T> *Wp,
                            hipStream_t st) {
    const int pm = dp.pmax_all[l];
    const dim3 g(dp.n_all[l], (dp.mmax_all[l] + 63) / 64), b(64);
    if (pm <= 8)
        hipLaunchKernelGGL((k_fwd_vec1<T, 8>), g, b, 0, st, arena, dp.lv_fr[l], gsrc, Xp, Wp);
    else if (pm <= 16)
        hipLaunchKernelGGL((k_fwd_vec1<T, 16>), g, b, 0, st, arena, dp.lv_fr[l], gsrc, Xp, Wp);
    else
        hipLaunchKernelGGL((k_fwd_vec1<T, 32>), g, b, 0, st, arena, dp.lv_fr[l], gsrc, Xp, Wp);
}
template <typename T>
static void launch_bwd_vec1(const DevPlan &dp, int l, const cx<T> *arena, const cx<T> *Wp, cx<T> *Xp,
                            hipStream_t st) {
    int R = 1;
    while (R < dp.pmax_all[l]) R <<= 1;
    // waves per front: as many as keep a thread's columns within one chunk of 16
    int nw = 1;
    while (nw < 4 && (dp.mmax_all[l] + 64 * nw / R - 1) / (64 * nw / R) > 16) nw *= 2;
    const int per_thread = (dp.mmax_all[l] + 64 * nw / R - 1) / (64 * nw / R);
    const dim3 g(dp.n_all[l]);
    const size_t lds = sizeof(cx<T>) * (size_t)dp.mmax_all[l];
#define BWD1(CH, NW) \
    hipLaunchKernelGGL((k_bwd_vec1<T, CH, NW>), g, dim3(64 * NW), lds, st, arena, dp.lv_fr[l], dp.idx, Wp, Xp)
    if (nw == 1) {
        if (per_thread <= 8) BWD1(8, 1);
        else BWD1(16, 1);
    } else if (nw == 2) {
        BWD1(16, 2);
    } else {
        BWD1(16, 4);
    }
#undef BWD1
}
template <typename T, bool FWD>
static void launch_vsweep(const DevPlan &dp, int l, const cx<T> *sq, const int *gsrc, const int *rowperm, cx<T> *Xp,
                          cx<T> *Wp, bool trans, bool cj, hipStream_t st) {
    const int rows = FWD ? dp.mmax_all[l] : dp.pmax_all[l], cols = FWD ? dp.pmax_all[l] : dp.mmax_all[l];
    int rw = VS_ROWS<T>;
    while (rw < 32 && 2 * rw * cols <= 256 * VS_PRE) rw *= 2;
    const int wgs = (rows + rw - 1) / rw * dp.n_all[l];
    // fewer than two workgroups per compute unit, each lane past its prefetch: 1024 threads (M solves)
    const bool wide = !trans && wgs < 512 && rw * cols > 256 * VS_PRE;
    const dim3 g((rows + rw - 1) / rw, dp.n_all[l]);
    const size_t lds = sizeof(cx<T>) * (size_t)cols;
#define VSWEEP(RW, NT)                                                                                               \
    hipLaunchKernelGGL((k_vsweep<T, FWD, RW, NT>), g, dim3(NT), lds, st, sq, dp.lv_fr[l], dp.vinv_off[l], gsrc,     \
                       dp.idx, rowperm, Xp, Wp, (int)trans, (int)cj)
    if (wide) {
        if (rw == 8) VSWEEP(8, 1024);
        else if (rw == 16) VSWEEP(16, 1024);
        else VSWEEP(32, 1024);
    } else {
        if (rw == 8) VSWEEP(8, 256);
        else if (rw == 16) VSWEEP(16, 256);
        else VSWEEP(32, 256);
    }
#undef VSWEEP
}

// Up to kVecMaxRhs right-hand sides: one pass per frequency and column through
// the single-vector kernels, in the work buffers Xp (n entries, elimination
// order) and Wp (fv_rows entries: every front's contribution rows, each
// written by its front before the parent reads it, so neither needs clearing).
template <typename T, typename TIO>
static void solve_vec(const Factors<T> &F, const Sel &sel, int sys, int nrhs, const cx<TIO> *B, cx<TIO> *X, cx<T> *Xp,
                      cx<T> *Wp, hipStream_t st, bool accumulate, bool inverses = true) {
    const DevPlan &dp = *F.plan;
    const Plan &pl = *dp.host;
    const int n = pl.n;
    const bool trans = sys != 0, cj = sys == 1;
    const double es = sizeof(cx<T>);
    const bool vi = inverses && ensure_vinv(F, st);
    for (int z = 0; z < sel.n; z++) {
        const int f = sel.f[z];
        const cx<T> *arena = F.arena + f * F.fstride;
        const cx<T> *sq = vi ? F.vinv + f * dp.vinv_elems : nullptr;
        const int *rowperm = F.rowperm + (int64_t)f * n;
        // gather table of the front rows: after the row permutation for M solves
        const int *gsrc = trans ? dp.gsrc_id : F.gsrc + (int64_t)f * pl.fv_rows * 3;
        for (int r = 0; r < nrhs; r++) {
            const int64_t off = ((int64_t)sel.s[z] * nrhs + r) * n;
            {
                ProfScope ps(K_PERMUTE_IN, st, 0.0, (double)n * (sizeof(cx<TIO>) + es + 4));
                hipLaunchKernelGGL((k_vec_in<T, TIO>), dim3(grid_for(n)), dim3(256), 0, st, B + off, Xp, dp.perm, n);
            }
            for (int l = 0; l < pl.nlevels; l++) {
                if (dp.n_all[l] == 0) continue;
                const DevPlan::LevelCounts &c = dp.counts[l];
                ProfScope ps(K_FWD_VEC, st, 8.0 * c.tri_elems, (c.tri_elems + 2.0 * c.sum_m) * es);
                size_t lds = sizeof(cx<T>) * (size_t)dp.mmax_all[l];
                if (vec_split(dp, l) && vi && dp.vinv[l]) {
                    launch_vsweep<T, true>(dp, l, sq, gsrc, rowperm, Xp, Wp, trans, cj, st);
                } else if (vec_split(dp, l)) {
                    hipLaunchKernelGGL(k_fwd_vec_tri<T>, dim3(dp.n_all[l]), dim3(256), lds, st, arena, dp.fronts,
                                       dp.lv_all[l], dp.children, dp.inv, rowperm, Xp, Wp, (int)trans, (int)cj);
                    const int chunks = (dp.mmax_all[l] + kVecSplitRows - 1) / kVecSplitRows;
                    hipLaunchKernelGGL(k_fwd_vec_rect<T>, dim3(chunks, dp.n_all[l]), dim3(256),
                                       sizeof(cx<T>) * (size_t)dp.pmax_all[l], st, arena, dp.fronts, dp.lv_all[l], Xp,
                                       Wp, (int)trans, (int)cj);
                } else if (dp.single[l] && !trans) {
                    launch_fwd_vec1<T>(dp, l, arena, gsrc, Xp, Wp, st);
                } else {
                    if (dp.single[l])
                        hipLaunchKernelGGL((k_fwd_vec<T, true>), dim3(dp.n_all[l]), dim3(256), 2 * lds, st, arena,
                                           dp.fronts, dp.lv_all[l], dp.children, dp.inv, rowperm, Xp, Wp, (int)trans,
                                           (int)cj);
                    else
                        hipLaunchKernelGGL((k_fwd_vec<T, false>), dim3(dp.n_all[l]), dim3(256), lds, st, arena,
                                           dp.fronts, dp.lv_all[l], dp.children, dp.inv, rowperm, Xp, Wp, (int)trans,
                                           (int)cj);
                }
            }
            for (int l = pl.nlevels - 1; l >= 0; l--) {
                if (dp.n_all[l] == 0) continue;
                const DevPlan::LevelCounts &c = dp.counts[l];
                ProfScope ps(K_BWD_VEC, st, 8.0 * c.tri_elems, (c.tri_elems + 2.0 * c.sum_m) * es);
                size_t lds = sizeof(cx<T>) * (size_t)dp.mmax_all[l];
                if (vec_split(dp, l) && vi && dp.vinv[l]) {
                    launch_vsweep<T, false>(dp, l, sq, gsrc, rowperm, Xp, Wp, trans, cj, st);
                } else if (vec_split(dp, l)) {
                    const int chunks = (dp.pmax_all[l] + kVecSplitRows - 1) / kVecSplitRows;
                    hipLaunchKernelGGL(k_bwd_vec_rect<T>, dim3(chunks, dp.n_all[l]), dim3(256), lds, st, arena,
                                       dp.fronts, dp.lv_all[l], dp.idx, Xp, (int)trans, (int)cj);
                    hipLaunchKernelGGL(k_bwd_vec_tri<T>, dim3(dp.n_all[l]), dim3(256),
                                       sizeof(cx<T>) * (size_t)dp.pmax_all[l], st, arena, dp.fronts, dp.lv_all[l],
                                       rowperm, Xp, (int)trans, (int)cj);
                } else if (dp.single[l] && !trans) {
                    launch_bwd_vec1<T>(dp, l, arena, Wp, Xp, st);
                } else {
                    hipLaunchKernelGGL(k_bwd_vec<T>, dim3(dp.n_all[l]), dim3(256), lds, st, arena, dp.fronts,
                                       dp.lv_all[l], dp.idx, rowperm, Xp, (int)trans, (int)cj, dp.single[l]);
                }
            }
            {
                ProfScope ps(K_PERMUTE_OUT, st, 0.0, (double)n * (sizeof(cx<TIO>) + es + 4));
                hipLaunchKernelGGL((k_vec_out<T, TIO>), dim3(grid_for(n)), dim3(256), 0, st, Xp, X + off, dp.iperm, n,
                                   (int)accumulate);
            }
        }
    }
    GFDFD_HIP(hipGetLastError());
}

template <typename T, typename TIO>
void solve_batch(const Factors<T> &F, const std::vector<int> &freqs, const std::vector<int> &slots, int nslot_buf,
                 int sys, int nrhs, const cx<TIO> *B, cx<TIO> *X, hipStream_t st, bool accumulate) {
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
        solve_vec<T, TIO>(F, sel, sys, nrhs, B, X, Xp, Wp, st, accumulate);
        return;
    }
    // internal panels are indexed by launch slot: slot z of Xp/Wp serves frequency sel.f[z]
    Sel in = sel;
    for (int z = 0; z < ns; z++) in.s[z] = z;
    // internal row stride: whole 128-byte lines, so every row segment the sweep
    // kernels move is line-aligned (zero padding columns)
    const int ld = (nrhs + (int)(128 / sizeof(cx<T>)) - 1) / (int)(128 / sizeof(cx<T>)) * (int)(128 / sizeof(cx<T>));
    cx<T> *Xp = (cx<T> *)ws_x.get(sizeof(cx<T>) * (size_t)n * ld * ns);
    cx<T> *Wp = (cx<T> *)ws_w.get(sizeof(cx<T>) * (size_t)pl.fv_rows * ld * ns);
    const double nr = nrhs, nz = ns;
    (void)nslot_buf;
    {
        // B slot sel.s[z] -> internal slot z
        ProfScope ps(K_PERMUTE_IN, st, 0.0, nz * ((double)n * nr * (sizeof(cx<TIO>) + es) + 4.0 * n));
        const int tiles = ((n + PT_K - 1) / PT_K) * ((ld + PT_R - 1) / PT_R);
        hipLaunchKernelGGL((k_permute_in_slots<T, TIO>), dim3(tiles, ns), dim3(256), 0, st, B, Xp, dp.iperm, n, nrhs,
                           ld, sel);
    }
    sweeps<T>(F, in, sys, ld, Xp, Wp, st, nrhs);
    {
        ProfScope ps(K_PERMUTE_OUT, st, 0.0, nz * ((double)n * nr * (sizeof(cx<TIO>) + es) + 4.0 * n));
        const int tiles = ((n + PT_K - 1) / PT_K) * ((nrhs + PT_R - 1) / PT_R);
        hipLaunchKernelGGL((k_permute_out_slots<T, TIO>), dim3(tiles, ns), dim3(256), 0, st, Xp, X, iout, mstride, n,
                           nrhs, ld, (int)accumulate, sel);
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
                                          int, int, int, const cx<double> *, cx<double> *, hipStream_t, bool);
template void solve_batch<float, double>(const Factors<float> &, const std::vector<int> &, const std::vector<int> &,
                                         int, int, int, const cx<double> *, cx<double> *, hipStream_t, bool);

// r_s = b_s - op(M_f) x_s for launch slots z (frequency sel.f[z], buffer slot
// sel.s[z]); with nrm, also the squared column norms nrm[sel.s[z] * nrhs + r]
static void residual_sel(const DevPlan &dp, const cx<double> *vals, int sys, int nrhs, const cx<double> *B,
                         const cx<double> *X, cx<double> *R, const Sel &sel, hipStream_t st, double *nrm = nullptr) {
    if (!dp.tq) throw std::logic_error("residual needs a grid plan");
    const int n = dp.host->n;
    const double nnz = (double)dp.host->nnz;
    const int nblk = (n + 255) / 256;
    // column sums in LDS: 4 waves x nrhs doubles; wider panels take the separate norm kernel
    const bool fused = nrm && nrhs <= 1024;
    double *partial = fused ? (double *)ws_p.get(sizeof(double) * (size_t)nblk * nrhs * sel.n) : nullptr;
    {
        ProfScope ps(K_RESIDUAL, st, sel.n * 8.0 * nnz * nrhs, sel.n * (nnz * (16.0 + 12.0) + 3.0 * n * 16.0 * nrhs));
        hipLaunchKernelGGL(k_residual, dim3(nblk, sel.n), dim3(256), fused ? 4 * sizeof(double) * nrhs : 0, st, vals,
                           dp.colptr, dp.rowind, dp.tq, n, nrhs, sys, dp.host->nnz, B, X, R, partial, sel);
    }
    if (fused) {
        ProfScope ps(K_NORM_REDUCE, st, sel.n * (double)nblk * nrhs, sel.n * 8.0 * (nblk + 1.0) * nrhs);
        hipLaunchKernelGGL(k_norm_reduce, dim3(nrhs, sel.n), dim3(256), 0, st, partial, nblk, nrhs, nrm, sel);
    } else if (nrm)
        for (int z = 0; z < sel.n; z++)
            hipLaunchKernelGGL(k_colnorm2, dim3(nrhs, 1), dim3(256), 0, st, R + (int64_t)sel.s[z] * n * nrhs, n, nrhs,
                               nrm + (int64_t)sel.s[z] * nrhs);
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
    ProfScope ps(K_AXPY, st, 2.0 * count, 48.0 * count);
    hipLaunchKernelGGL(k_axpy, dim3(grid_for(count)), dim3(256), 0, st, count, D, X);
}

void cast_c128_to_c64(int64_t count, const cx<double> *a, cx<float> *b, hipStream_t st) {
    hipLaunchKernelGGL((k_cast<double, float>), dim3(grid_for(count)), dim3(256), 0, st, count, a, b);
}
void cast_c64_to_c128(int64_t count, const cx<float> *a, cx<double> *b, hipStream_t st) {
    hipLaunchKernelGGL((k_cast<float, double>), dim3(grid_for(count)), dim3(256), 0, st, count, a, b);
}

// Host side of the stopping test: h holds the squared column norms of the
// residuals (h[s * nrhs + r]) and of b (h[(ns + s) * nrhs + r]).  Records the
// worst relative residual of every active slot and keeps those above tol.
static void keep_unconverged(const std::vector<double> &h, int ns, int nrhs, int it, double tol, int max_iter,
                             std::vector<int> &act_f, std::vector<int> &act_s, std::vector<RefineStats> &rs) {
    static const bool trace = std::getenv("GFDFD_IR_TRACE") != nullptr;
    if (trace)
        for (size_t z = 0; z < act_s.size(); z++) {
            // per step and frequency: columns above tol, and the relative residual's quantiles
            const int s = act_s[z];
            std::vector<double> rel(nrhs);
            for (int r = 0; r < nrhs; r++) {
                const double b = std::sqrt(h[((size_t)ns + s) * nrhs + r]), rr = std::sqrt(h[(size_t)s * nrhs + r]);
                rel[r] = b > 0 ? rr / b : rr;
            }
            std::sort(rel.begin(), rel.end());
            const int above = (int)(rel.end() - std::upper_bound(rel.begin(), rel.end(), tol));
            std::fprintf(stderr, "refine step %d freq %d: %d/%d columns above %.1e; rel residual min %.2e median %.2e "
                         "p90 %.2e max %.2e\n", it, act_f[z], above, nrhs, tol, rel[0], rel[nrhs / 2],
                         rel[(nrhs * 9) / 10], rel[nrhs - 1]);
        }
    std::vector<int> nf, nsl;
    for (size_t z = 0; z < act_s.size(); z++) {
        const int s = act_s[z];
        double worst = 0;
        for (int r = 0; r < nrhs; r++) {
            const double b = std::sqrt(h[((size_t)ns + s) * nrhs + r]), rr = std::sqrt(h[(size_t)s * nrhs + r]);
            worst = std::max(worst, b > 0 ? rr / b : rr);
        }
        rs[s].max_rel_residual = worst;
        rs[s].iterations = it;
        if (worst > tol && it < max_iter) {
            nf.push_back(act_f[z]);
            nsl.push_back(s);
        }
    }
    act_f = nf;
    act_s = nsl;
}

thread_local Workspace ws_r64, ws_acc64, ws_dk[kIrBufs], ws_pack, ws_pmap;
// widest panel of the row-major refinement: its per-wave column sums and the
// staged neighbourhood share the default 64 KB of LDS
constexpr int kIrMaxRhs = 1024;
constexpr int kIrDynLdsMax = kIrWaves * 8 * kIrMaxRhs;   // per-wave column sums at the widest panel
static_assert(kIrDynLdsMax + (kIrStage * kIrCw + 9 * kIrPts) * 16 + (9 * kIrPts + 2 * kIrStage) * 4 <= 160 * 1024,
              "refinement-step LDS");

// Row-major panel refinement (grid plans, kVecMaxRhs < nrhs <= kIrMaxRhs): b
// is transposed once into the first solve's panel (elimination order, factor
// precision), every step is one k_ir_step pass (r -= op(M) d_k) plus, while a
// slot has not converged, one correction solve in place on the step's
// factor-precision right-hand side, which becomes d_{k+1}.  The corrections
// stay in kIrBufs panels (folded into an FP64 sum when those are full) and are
// summed once into the caller's layout.  Slot s of every panel serves
// frequency freqs[s]; converged slots leave the batch.
template <typename T>
std::vector<RefineStats> refine_panels(const Factors<T> &F, const std::vector<int> &freqs, int sys, int nrhs,
                                       const cx<double> *B, cx<double> *X, double tol, int max_iter,
                                       hipStream_t st) {
    const DevPlan &dp = *F.plan;
    const Plan &pl = *dp.host;
    const int n = pl.n, ns = (int)freqs.size();
    set_smem_attrs<T>();
    const int cl = (int)(128 / sizeof(cx<T>));
    const int ld = (nrhs + cl - 1) / cl * cl, ld64 = (nrhs + 7) / 8 * 8;
    const size_t e64 = (size_t)n * ld64 * ns, e32 = (size_t)n * ld * ns;
    cx<double> *Rg = (cx<double> *)ws_r64.get(sizeof(cx<double>) * e64);
    // correction buffer j (allocated on first use)
    auto dbuf = [&](int j) { return (cx<T> *)ws_dk[j].get(sizeof(cx<T>) * e32); };
    cx<T> *Wp = (cx<T> *)ws_w.get(sizeof(cx<T>) * (size_t)pl.fv_rows * ld * ns);
    double *nrm = (double *)ws_n.get(sizeof(double) * 2 * nrhs * ns);
    const int ktiles = (n + PT_K - 1) / PT_K;
    // up to 8192 workgroups per slot (measured faster than 1024-4096), a multiple of 8
    const int nblk = std::min(((pl.nx + kIrTx - 1) / kIrTx) * ((pl.ny + kIrTy - 1) / kIrTy) + 7, 8192) / 8 * 8;
    double *partial = (double *)ws_p.get(sizeof(double) * (size_t)std::max(ktiles, nblk) * nrhs * ns);
    std::vector<int> slots(ns);
    for (int z = 0; z < ns; z++) slots[z] = z;
    const Sel all = make_sel(freqs, slots);
    const double nr = nrhs, es = sizeof(cx<T>), nnz = (double)pl.nnz;
    {
        ProfScope ps(K_PERMUTE_IN, st, 0.0, ns * ((double)n * nr * (16.0 + es) + 4.0 * n));
        const int tiles = ktiles * ((ld + PT_R - 1) / PT_R);
        hipLaunchKernelGGL(k_ir_in<T>, dim3(tiles, ns), dim3(256), 0, st, B, dbuf(0), dp.iperm, n, nrhs, ld, partial,
                           all);
    }
    hipLaunchKernelGGL(k_norm_reduce, dim3(nrhs, ns), dim3(256), 0, st, partial, ktiles, nrhs,
                       nrm + (int64_t)nrhs * ns, all);
    sweeps<T>(F, all, sys, ld, dbuf(0), Wp, st, nrhs);
    std::vector<RefineStats> rs(ns);
    std::vector<double> h(2 * (size_t)nrhs * ns);
    std::vector<int> act_f = freqs, act_s = slots;
    std::vector<int> base(ns, 0);   // per slot: corrections already folded into acc
    uint64_t folded = 0;            // bit s: slot s has a folded sum in acc
    cx<double> *acc = nullptr;
    std::vector<int> hmap;          // column maps of a packed correction solve (host side of ws_pmap)
    for (int it = 0;; it++) {
        const Sel sel = make_sel(act_f, act_s);
        const int jn = (it + 1) % kIrBufs;   // buffer of the next right-hand side
        if (jn == 0) {
            // the buffers hold corrections it + 1 - kIrBufs .. it of every active slot:
            // fold them into acc before the step overwrites buffer 0
            acc = (cx<double> *)ws_acc64.get(sizeof(cx<double>) * e32);
            IrBufs d;
            for (int j = 0; j < kIrBufs; j++) d.p[j] = dbuf(j);
            uint64_t had = 0;
            for (int z = 0; z < sel.n; z++)
                if ((folded >> sel.s[z]) & 1) had |= uint64_t(1) << z;
            const int64_t count = (int64_t)n * ld;
            hipLaunchKernelGGL(k_ir_fold<T>, dim3(grid_for(count), sel.n), dim3(256), 0, st, d, acc, had, count, sel);
            for (int z = 0; z < sel.n; z++) {
                folded |= uint64_t(1) << sel.s[z];
                base[sel.s[z]] = it + 1;
            }
        }
        {
            // per slot: r (b on the first step), d and the neighbours' d from
            // cache, r in FP64 and in the factor precision written; the stencil
            // values and indices
            const double panel = (double)n * nr * (16.0 + es + 16.0 + es);
            ProfScope ps(K_RESIDUAL, st, sel.n * (8.0 * nnz * nr + 2.0 * n * nr), sel.n * (panel + nnz * 28.0));
            hipLaunchKernelGGL(k_ir_step<T>, dim3(nblk, sel.n), dim3(kIrThreads), kIrWaves * sizeof(double) * nrhs, st,
                               F.vals, dp.colptr, dp.rowind, dp.tq, dp.iperm, pl.nx, pl.ny, nrhs, ld64, ld, sys,
                               pl.nnz, it == 0 ? B : nullptr, Rg, dbuf(it % kIrBufs), dbuf(jn), partial, sel);
        }
        {
            ProfScope ps(K_NORM_REDUCE, st, sel.n * (double)nblk * nrhs, sel.n * 8.0 * (nblk + 1.0) * nrhs);
            hipLaunchKernelGGL(k_norm_reduce, dim3(nrhs, sel.n), dim3(256), 0, st, partial, nblk, nrhs, nrm, sel);
        }
        GFDFD_HIP(hipMemcpyAsync(h.data(), nrm, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
        GFDFD_HIP(hipStreamSynchronize(st));
        keep_unconverged(h, ns, nrhs, it, tol, max_iter, act_f, act_s, rs);
        if (act_f.empty()) break;
        // d = op(M)^-1 r in place; when at most half of the columns of every
        // remaining slot are still above the tolerance, only those are solved
        // (packed into ldc columns) and the others get a zero correction
        const Sel nsel = make_sel(act_f, act_s);
        std::vector<std::vector<int>> cols(nsel.n);
        int most = 0;
        for (int z = 0; z < nsel.n; z++) {
            const int s = nsel.s[z];
            for (int c = 0; c < nrhs; c++) {
                const double b = std::sqrt(h[((size_t)ns + s) * nrhs + c]), rr = std::sqrt(h[(size_t)s * nrhs + c]);
                if ((b > 0 ? rr / b : rr) > tol) cols[z].push_back(c);
            }
            most = std::max(most, (int)cols[z].size());
        }
        const int ldc = (most + cl - 1) / cl * cl;
        if (ldc > 0 && 2 * ldc <= ld) {
            hmap.assign((size_t)nsel.n * (ldc + ld), -1);
            for (int z = 0; z < nsel.n; z++)
                for (int j = 0; j < (int)cols[z].size(); j++) {
                    hmap[(size_t)z * ldc + j] = cols[z][j];
                    hmap[(size_t)nsel.n * ldc + (size_t)z * ld + cols[z][j]] = j;
                }
            int *dmap = (int *)ws_pmap.get(sizeof(int) * hmap.size());
            // (hmap outlives the copy: the next use follows this step's stream synchronisation)
            GFDFD_HIP(hipMemcpyAsync(dmap, hmap.data(), sizeof(int) * hmap.size(), hipMemcpyHostToDevice, st));
            cx<T> *C = (cx<T> *)ws_pack.get(sizeof(cx<T>) * (size_t)n * ldc * ns);
            const int64_t ce = (int64_t)n * ldc, pe = (int64_t)n * ld;
            hipLaunchKernelGGL(k_ir_pack<T>, dim3(grid_for(ce), nsel.n), dim3(256), 0, st, dbuf(jn), C, dmap, n, ld, ldc,
                               nsel);
            sweeps<T>(F, nsel, sys, ldc, C, Wp, st, most);
            hipLaunchKernelGGL(k_ir_unpack<T>, dim3(grid_for(pe), nsel.n), dim3(256), 0, st, C, dbuf(jn),
                               dmap + (size_t)nsel.n * ldc, n, ld, ldc, nsel);
        } else {
            sweeps<T>(F, nsel, sys, ld, dbuf(jn), Wp, st, nrhs);
        }
    }
    {
        // slot s holds corrections base[s] .. iterations in buffers 0, 1, ...
        IrBufs d;
        IrCounts cnt;
        int most = 0;
        for (int z = 0; z < ns; z++) {
            cnt.c[z] = rs[z].iterations + 1 - base[z];
            most = std::max(most, cnt.c[z]);
        }
        for (int j = 0; j < kIrBufs; j++) d.p[j] = j < most ? dbuf(j) : nullptr;
        double reads = 0;
        for (int z = 0; z < ns; z++) reads += cnt.c[z] * es + (((folded >> z) & 1) ? 16.0 : 0.0);
        ProfScope ps(K_PERMUTE_OUT, st, 0.0, (double)n * nr * (reads + 16.0 * ns) + 4.0 * n * ns);
        const int tiles = ktiles * ((nrhs + PT_R - 1) / PT_R);
        hipLaunchKernelGGL(k_ir_sum_out<T>, dim3(tiles, ns), dim3(256), 0, st, d, cnt, acc, folded, X, dp.iperm, n,
                           nrhs, ld, all);
    }
    GFDFD_HIP(hipGetLastError());
    return rs;
}

// FP64 refinement of nsel frequencies at once; frequencies leave the batch as
// they converge (the correction solves batch the remaining ones).  Panels take
// the elimination-order refinement above; up to kVecMaxRhs right-hand sides
// (or more than kIrMaxRhs) run the residual in the user layout, with the
// column norms from the residual kernel and each correction added to X by the
// solve's final transpose.
template <typename T>
std::vector<RefineStats> solve_refined_batch(const Factors<T> &F, const std::vector<int> &freqs, int sys, int nrhs,
                                             const cx<double> *B, cx<double> *X, double tol, int max_iter,
                                             hipStream_t st) {
    const DevPlan &dp = *F.plan;
    if (nrhs > kVecMaxRhs && nrhs <= kIrMaxRhs && dp.tq && dp.host->grid && refine_mode() == 0)
        return refine_panels<T>(F, freqs, sys, nrhs, B, X, tol, max_iter, st);
    const int n = dp.host->n, ns = (int)freqs.size();
    const int64_t panel = (int64_t)n * nrhs;
    cx<double> *R = (cx<double> *)ws_r.get(sizeof(cx<double>) * panel * ns);
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
        residual_sel(dp, F.vals, sys, nrhs, B, X, R, sel, st, nrm);
        GFDFD_HIP(hipMemcpyAsync(h.data(), nrm, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
        GFDFD_HIP(hipStreamSynchronize(st));
        keep_unconverged(h, ns, nrhs, it, tol, max_iter, act_f, act_s, rs);
        if (act_f.empty()) break;
        solve_batch<T, double>(F, act_f, act_s, ns, sys, nrhs, R, X, st, true);   // X += op(M)^-1 r
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

// the calling thread's cached solve buffers back to the device (they grow to
// the largest solve and are otherwise kept for the next one)
void release_solve_workspaces() {
    for (Workspace *w : {&ws_x, &ws_w, &ws_r, &ws_n, &ws_p, &ws_r64, &ws_acc64, &ws_pack, &ws_pmap}) w->release();
    for (Workspace &w : ws_dk) w.release();
}

// ---------------------------------------------------------------------------
// Solve without allocation: the work buffers of the few-right-hand-side path
// live in one file-scope device array.  Host code reaches it only through
// the address hipGetSymbolAddress returns for the current device (each GPU
// holds its own copy), looked up once per device.  Every call on a GPU
// shares that GPU's array, so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static void *scratch_on_device() {
    static std::mutex mu;
    static void *ptr[64] = {};
    int dev = 0;
    GFDFD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw std::invalid_argument("solve_noalloc: device index beyond 63");
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw hip_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        ptr[dev] = p;
    }
    return ptr[dev];
}

size_t noalloc_scratch_bytes(const Plan &pl, size_t entry_bytes) {
    const size_t xb = ((size_t)pl.n * entry_bytes + 255) / 256 * 256;
    return xb + (size_t)pl.fv_rows * entry_bytes;
}

template <typename T>
void solve_noalloc(const Factors<T> &F, int ifreq, int sys, int nrhs, const cx<double> *B, cx<double> *X,
                   hipStream_t st) {
    const Plan &pl = *F.plan->host;
    if (ifreq < 0 || ifreq >= F.nfreq) throw std::invalid_argument("frequency index out of range");
    if (sys < 0 || sys > 2) throw std::invalid_argument("invalid system");
    if (nrhs < 1 || nrhs > kVecMaxRhs)
        throw std::invalid_argument("gfdfd_solve_noalloc solves 1 to " + std::to_string(kVecMaxRhs) +
                                    " right-hand sides per call");
    const size_t need = noalloc_scratch_bytes(pl, sizeof(cx<T>));
    if (need > kNoAllocScratchBytes)
        throw std::invalid_argument("gfdfd_solve_noalloc: the grid needs " + std::to_string(need) +
                                    " bytes of scratch (n + work rows entries), the static array holds " +
                                    std::to_string(kNoAllocScratchBytes));
    set_smem_attrs<T>();
    ProfPause no_events;   // profiling would create events
    unsigned char *base = static_cast<unsigned char *>(scratch_on_device());
    cx<T> *Xp = reinterpret_cast<cx<T> *>(base);
    cx<T> *Wp = reinterpret_cast<cx<T> *>(base + ((size_t)pl.n * sizeof(cx<T>) + 255) / 256 * 256);
    solve_vec<T, double>(F, make_sel({ifreq}, {0}), sys, nrhs, B, X, Xp, Wp, st, false, false);
}
template void solve_noalloc<double>(const Factors<double> &, int, int, int, const cx<double> *, cx<double> *,
                                    hipStream_t);
template void solve_noalloc<float>(const Factors<float> &, int, int, int, const cx<double> *, cx<double> *,
                                   hipStream_t);

}  // namespace gfdfd
