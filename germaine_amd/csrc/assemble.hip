// Device assembly of the 9-point mixed-grid PML impedance matrix (AC, SH, TE).
//
// One thread per grid point k = i + j*nx writes the point's (up to) nine
// entries at the CSC column start of M (= the row start of the assembled
// stencil A, M = A^T as UMFPACK sees it after the host's value overwrite in
// src/AC/forward_shot_AC.c:43-48), in the order NW, N, NE, W, C, E, SW, S, SE.
//
// Bit-exact with the reference's host loop: every entry is evaluated with the
// reference's C semantics written out as explicit IEEE operations --
// float-complex sub-expressions in float, the terms that meet the double
// literal 0.25 in double, the sum rounded to complex float, complex products
// as (ac-bd, ad+bc), real*complex componentwise, and libgcc's Smith division
// for the TE conductivity term.  This file is compiled without FP
// contraction.  Per-frequency scalars (Omega^2, Omega, the FD weights) and
// the 1-D PML profiles are computed on the host by gcc-compiled code, as the
// reference does (src/pml_pro.c, src/AC/init_A_AC_9p_pml.c:34-50).
#pragma clang fp contract(off)
#include "assemble.hpp"
#include "common.hpp"
#include "prof.hpp"

namespace gfdfd {

namespace {

struct cf { float r, i; };
struct cd { double r, i; };

__device__ __forceinline__ cd up(cf a) { return {(double)a.r, (double)a.i}; }
__device__ __forceinline__ cf dn(cd a) { return {(float)a.r, (float)a.i}; }
__device__ __forceinline__ cf fmulc(cf a, cf b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ cd dmulc(cd a, cd b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ cf fs(float s, cf a) { return {s * a.r, s * a.i}; }
__device__ __forceinline__ cd ds(double s, cd a) { return {s * a.r, s * a.i}; }
__device__ __forceinline__ cf fadd(cf a, cf b) { return {a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cf fsub(cf a, cf b) { return {a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cd dadd(cd a, cd b) { return {a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cd dsub(cd a, cd b) { return {a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cf fneg(cf a) { return {-a.r, -a.i}; }

// libgcc __divsc3 for finite operands (Smith's method), numerator a + b i
__device__ __forceinline__ cf smith_div(float a, float b, float c, float d) {
    float ratio, denom;
    cf q;
    if (fabsf(c) < fabsf(d)) {
        ratio = c / d;
        denom = (c * ratio) + d;
        q.r = ((a * ratio) + b) / denom;
        q.i = ((b * ratio) - a) / denom;
    } else {
        ratio = d / c;
        denom = (d * ratio) + c;
        q.r = ((b * ratio) + a) / denom;
        q.i = (b - (a * ratio)) / denom;
    }
    return q;
}

struct Lap { float bpm, bmp, bpp, bmm, b00, bp0, bm0, b0p, b0m; };

struct Prof {
    const float *xr, *xi, *xhr, *xhi, *yr, *yi, *yhr, *yhi;
};

// material value at 1-based (i,j), row-major [ny][nx]
__device__ __forceinline__ float at(const float *m, int nx, int i, int j) { return m[(j - 1) * nx + (i - 1)]; }

__device__ __forceinline__ float ac_ivp2(const float *vp, const float *rho, int nx, int i, int j) {
    float r = at(rho, nx, i, j), v = at(vp, nx, i, j);
    return (float)(1.0 / (double)(r * v * v));
}

// SH shear modulus with the one-cell edge-copied halo (src/SH/init_mat_SH.c:22-32)
__device__ __forceinline__ float sh_mu(const float *vs, const float *rho, int nx, int ny, int i, int j) {
    i = i < 1 ? 1 : (i > nx ? nx : i);
    j = j < 1 ? 1 : (j > ny ? ny : j);
    float r = at(rho, nx, i, j), v = at(vs, nx, i, j);
    return r * v * v;
}

template <int PHYS>
__global__ void __launch_bounds__(256) k_assemble(int nx, int ny, AsmScalars sc, const float *__restrict__ m1,
                                                  const float *__restrict__ m2, const float *__restrict__ prof,
                                                  const int *__restrict__ colptr, cx<double> *__restrict__ out,
                                                  int64_t nnz) {
    const int f = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nx * ny) return;
    const AsmFreq fr = sc.freq[f];
    const float *P = prof + (int64_t)f * 4 * (nx + ny + 4);
    Prof pf{P, P + (nx + 2), P + 2 * (nx + 2), P + 3 * (nx + 2), P + 4 * (nx + 2), P + 4 * (nx + 2) + (ny + 2),
            P + 4 * (nx + 2) + 2 * (ny + 2), P + 4 * (nx + 2) + 3 * (ny + 2)};
    const int i = k % nx + 1, j = k / nx + 1;   // reference's 1-based indices
    const float a = sc.a, b = sc.b, c = sc.c, d = sc.d, e = sc.e, idh2 = sc.idh2, damp = 0.0f;
    const cf O2 = {fr.omega2_re, fr.omega2_im};

    // Laplacian weights (src/av_rho.c:32-48 for AC, src/SH/av_mu.c:12-28 for SH, 1 for TE)
    Lap w;
    if (PHYS == 1) {
        float bl = (float)(1.0 / (double)at(m2, nx, i, j));
        w = {bl, bl, bl, bl, bl, bl, bl, bl, bl};
    } else if (PHYS == 2) {
        auto M = [&](int ii, int jj) { return sh_mu(m1, m2, nx, ny, ii, jj); };
        w.bpm = (float)(0.25 * (double)(M(i, j) + M(i + 1, j) + M(i, j - 1) + M(i + 1, j - 1)));
        w.bmp = (float)(0.25 * (double)(M(i, j) + M(i - 1, j) + M(i, j + 1) + M(i - 1, j + 1)));
        w.bpp = (float)(0.25 * (double)(M(i, j) + M(i + 1, j) + M(i, j + 1) + M(i + 1, j + 1)));
        w.bmm = (float)(0.25 * (double)(M(i, j) + M(i - 1, j) + M(i, j - 1) + M(i - 1, j - 1)));
        w.b00 = M(i, j);
        w.bp0 = (float)(0.5 * (double)(M(i, j) + M(i + 1, j)));
        w.bm0 = (float)(0.5 * (double)(M(i, j) + M(i - 1, j)));
        w.b0p = (float)(0.5 * (double)(M(i, j) + M(i, j + 1)));
        w.b0m = (float)(0.5 * (double)(M(i, j) + M(i, j - 1)));
    } else {
        w = {1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f};
    }

    // mass factor of a neighbour
    auto mass = [&](int ii, int jj) -> cf {
        if (PHYS == 1) return fs(ac_ivp2(m1, m2, nx, ii, jj), O2);
        if (PHYS == 2) return fs(at(m2, nx, ii, jj), O2);
        cf q = smith_div(at(m2, nx, ii, jj), 0.0f, fr.omega_re, fr.omega_im);   // sigma / Omega
        cf epse = {at(m1, nx, ii, jj) + (-(q.i * 1.0f)), q.r * 1.0f};            // eps + (.)*I
        return fmulc(fs(sc.mu0, O2), epse);                                       // Omega2 * mu0 * epse
    };

    const cf dy0 = {pf.yr[j], pf.yi[j]}, dx0 = {pf.xr[i], pf.xi[i]};
    const cf dyp = {pf.yhr[j], pf.yhi[j]}, dym = {pf.yhr[j - 1], pf.yhi[j - 1]};
    const cf dxp = {pf.xhr[i], pf.xhi[i]}, dxm = {pf.xhr[i - 1], pf.xhi[i - 1]};
    const double q = 0.25 * (double)idh2, nq = -0.25 * (double)idh2;
    const cd qy = ds(q, up(dy0)), qx = ds(q, up(dx0));

    cx<double> *o = out + (int64_t)f * nnz + colptr[k];
    int n = 0;
    auto emit = [&](cf t) { o[n++] = {(double)t.r, (double)t.i}; };
    auto corner = [&](cf ms, float bw, cf dyh, cf dxh) {
        cd s = dadd(dmulc(ds(bw, qy), up(dyh)), dmulc(ds(bw, qx), up(dxh)));
        return dn(dadd(up(fs(c, ms)), ds(d, s)));
    };
    auto edge = [&](cf ms, cf sA, cf sB, cf et) {
        cd s = dadd(dmulc(qy, up(sA)), dmulc(qx, up(sB)));
        cf t = dn(dadd(dadd(up(fs(b, ms)), ds(d, s)), up(et)));
        return fadd(t, fs(-damp, t));
    };
    if (i > 1 && j > 1) emit(corner(mass(i - 1, j - 1), w.bmm, dym, dxm));                       // NW
    if (j > 1)                                                                                     // N
        emit(edge(mass(i, j - 1), fadd(fs(w.bpm, dym), fs(w.bmm, dym)), fsub(fs(-w.bpm, dxp), fs(w.bmm, dxm)),
                  fs(e, fmulc(fs(w.bm0, fs(idh2, dy0)), dym))));
    if (i < nx && j > 1) emit(corner(mass(i + 1, j - 1), w.bpm, dym, dxp));                      // NE
    if (i > 1)                                                                                     // W
        emit(edge(mass(i - 1, j), fsub(fs(-w.bmp, dyp), fs(w.bmm, dym)), fadd(fs(w.bmp, dxm), fs(w.bmm, dxm)),
                  fs(e, fmulc(fs(w.bm0, fs(idh2, dx0)), dxm))));
    {                                                                                              // C
        cf S1 = fadd(fadd(fadd(fs(w.bmp, dyp), fs(w.bpm, dym)), fs(w.bpp, dyp)), fs(w.bmm, dym));
        cf S2 = fadd(fadd(fadd(fs(w.bmp, dxm), fs(w.bpm, dxp)), fs(w.bpp, dxp)), fs(w.bmm, dxm));
        cd D = ds(d, dsub(dmulc(ds(nq, up(dy0)), up(S1)), dmulc(qx, up(S2))));
        cf E1 = fmulc(fs(idh2, fneg(dy0)), fadd(fs(w.b0p, dyp), fs(w.b0m, dym)));
        cf E2 = fmulc(fs(idh2, dx0), fadd(fs(w.bp0, dxp), fs(w.bm0, dxm)));
        cf t = dn(dadd(dadd(up(fs(a, mass(i, j))), D), up(fs(e, fsub(E1, E2)))));
        t = dn(dadd(up(t), ds((double)damp * 4.0, up(t))));
        emit(t);
    }
    if (i < nx)                                                                                    // E
        emit(edge(mass(i + 1, j), fsub(fs(-w.bpm, dym), fs(w.bpp, dyp)), fadd(fs(w.bpm, dxp), fs(w.bpp, dxp)),
                  fs(e, fmulc(fs(w.bp0, fs(idh2, dx0)), dxp))));
    if (j < ny && i > 1) emit(corner(mass(i - 1, j + 1), w.bmp, dyp, dxm));                      // SW
    if (j < ny)                                                                                    // S
        emit(edge(mass(i, j + 1), fadd(fs(w.bmp, dyp), fs(w.bpp, dyp)), fsub(fs(-w.bmp, dxm), fs(w.bpp, dxp)),
                  fs(e, fmulc(fs(w.b0p, fs(idh2, dy0)), dyp))));
    if (i < nx && j < ny) emit(corner(mass(i + 1, j + 1), w.bpp, dyp, dxp));                     // SE
}

}  // namespace

void assemble_device(int physics, int nx, int ny, const AsmScalars &sc, const float *m1, const float *m2,
                     const float *prof, const int *colptr, cx<double> *out, int64_t nnz, int nfreq, hipStream_t st) {
    dim3 grid((nx * ny + 255) / 256, nfreq);
    // algorithmic traffic: the values written plus the two material grids read
    ProfScope ps(K_ASSEMBLE, st, 0.0, (double)nfreq * (16.0 * nnz + 8.0 * nx * ny));
    switch (physics) {
    case 1:
        hipLaunchKernelGGL(k_assemble<1>, grid, dim3(256), 0, st, nx, ny, sc, m1, m2, prof, colptr, out, nnz);
        break;
    case 2:
        hipLaunchKernelGGL(k_assemble<2>, grid, dim3(256), 0, st, nx, ny, sc, m1, m2, prof, colptr, out, nnz);
        break;
    case 3:
        hipLaunchKernelGGL(k_assemble<3>, grid, dim3(256), 0, st, nx, ny, sc, m1, m2, prof, colptr, out, nnz);
        break;
    default:
        throw std::invalid_argument("physics must be 1 (AC), 2 (SH) or 3 (TE)");
    }
    GFDFD_HIP(hipGetLastError());
}

}  // namespace gfdfd
