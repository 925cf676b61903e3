// Shot kernels on the device: point-source panels, receiver sampling, data
// residuals with the adjoint source panel, and the FWI gradients reduced over
// all shots in FP64.  Panels are column-major n x nshots (one shot per column).
#include <algorithm>

#include "common.hpp"
#include "shot.hpp"

namespace gfdfd {

namespace {

// src/AC/RHS_source_AC.c:46-53 (TE :50-56) for every shot: zero, then the spike
__global__ void k_spikes(int n, int nshots, const int *__restrict__ k, cx<double> amp, cx<double> *__restrict__ B) {
    const int64_t total = (int64_t)n * nshots;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int s = (int)(e / n), row = (int)(e - (int64_t)s * n);
        B[e] = row == k[s] ? amp : cx<double>{0.0, 0.0};
    }
}

// src/AC/calc_seis_AC.c:19-20 after src/vec2mat.c:19-24: the float wavefield
// sample of receiver t is the solution entry k_rec(t)
__global__ void k_sample(int n, int nshots, int ntr, const int *__restrict__ krec, const cx<double> *__restrict__ X,
                         cx<float> *__restrict__ pcal) {
    const int64_t total = (int64_t)ntr * nshots;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int s = (int)(e / ntr), t = (int)(e - (int64_t)s * ntr);
        cx<double> v = X[krec[t] + (int64_t)s * n];
        pcal[e] = {(float)v.x, (float)v.y};
    }
}

// complex float arithmetic as C evaluates it (products as ac-bd, ad+bc; no
// contraction), for the quantities the reference rounds in complex float
struct cf32 { float r, i; };
__device__ __forceinline__ cf32 cf_mul(cf32 a, cf32 b) {
#pragma clang fp contract(off)
    return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
// libgcc __divsc3 for finite operands (Smith's method)
__device__ __forceinline__ cf32 cf_div(cf32 x, cf32 y) {
#pragma clang fp contract(off)
    float ratio, denom;
    cf32 q;
    if (fabsf(y.r) < fabsf(y.i)) {
        ratio = y.r / y.i;
        denom = (y.r * ratio) + y.i;
        q.r = ((x.r * ratio) + x.i) / denom;
        q.i = ((x.i * ratio) - x.r) / denom;
    } else {
        ratio = y.i / y.r;
        denom = (y.i * ratio) + y.r;
        q.r = ((x.i * ratio) + x.r) / denom;
        q.i = (x.i - (x.r * ratio)) / denom;
    }
    return q;
}
// cabsf / cargf / clogf, each rounded once from a double evaluation
__device__ __forceinline__ float cf_abs(cf32 a) { return (float)sqrt((double)a.r * a.r + (double)a.i * a.i); }
__device__ __forceinline__ float cf_arg(cf32 a) { return (float)atan2((double)a.i, (double)a.r); }
__device__ __forceinline__ cf32 cf_log(cf32 a) {
    return {(float)log(sqrt((double)a.r * a.r + (double)a.i * a.i)), cf_arg(a)};
}

// src/AC/calc_res_AC.c:10-95 and src/AC/RHS_source_AC_adj.c:21-36, one
// workgroup per shot.  With stf_inv, lane 0 forms the Wiener quotient
// sum conj(pcal) pobs / sum conj(pobs) pobs in the reference's order and
// complex float rounding (:21-40) and publishes it (waveAC.stfr/stfi) in
// stf[s]; else the estimate is 1.  The residual per receiver: rtm, the observed
// datum (INVMAT=2, :77); misfit 1, (pobs - wien pcal)/|wien| (:62); misfit 2,
// clog(pobs) - clog(pcal) (:66); misfit 3, arg(pcal) - arg(pobs) (:70); muted
// where pobsr<1e-20 && pobsi<1e-20 (:82-85, signed).  The adjoint panel holds
// conj(res)/DH^2 at k_rec (the last of duplicate receivers wins); per-shot
// misfit sum |res|^2 with the muted residuals included (:88).
__global__ void k_adjoint(int n, int nshots, int ntr, const int *__restrict__ krec, const cx<float> *__restrict__ pcal,
                          const cx<float> *__restrict__ pobs, float dh, int misfit, int stf_inv, int rtm,
                          cx<double> *__restrict__ B, double *__restrict__ l2, cx<float> *__restrict__ stf) {
#pragma clang fp contract(off)
    const int s = blockIdx.x;
    cx<double> *b = B + (int64_t)s * n;
    for (int row = threadIdx.x; row < n; row += blockDim.x) b[row] = {0.0, 0.0};
    __shared__ cf32 wien;
    __shared__ float wabs;
    if (threadIdx.x == 0) {
        cf32 w = {1.0f, 0.0f};
        if (stf_inv) {
            // conj() is the double-precision function, so each product is a
            // complex double (float x float products are exact in double; the
            // sums round once) and `wiennom += ...` adds in double before
            // rounding back to the complex float accumulator (:35-36)
            cf32 nom = {0.0f, 0.0f}, den = {0.0f, 0.0f};
            for (int t = 0; t < ntr; t++) {
                const cx<float> o = pobs[(int64_t)s * ntr + t], c = pcal[(int64_t)s * ntr + t];
                const double cr = c.x, ci = -(double)c.y, orr = o.x, oi = o.y;
                nom.r = (float)((double)nom.r + (cr * orr - ci * oi));
                nom.i = (float)((double)nom.i + (cr * oi + ci * orr));
                den.r = (float)((double)den.r + (orr * orr - (-oi) * oi));
                den.i = (float)((double)den.i + (orr * oi + (-oi) * orr));
            }
            w = cf_div(nom, den);
        }
        wien = w;
        wabs = cf_abs(w);
        if (stf) stf[s] = {w.r, w.i};
    }
    __syncthreads();
    __shared__ double part[256];
    double acc = 0.0;
    for (int t = threadIdx.x; t < ntr; t += blockDim.x) {
        const cx<float> o = pobs[(int64_t)s * ntr + t], c = pcal[(int64_t)s * ntr + t];
        cf32 res;
        if (rtm) {
            res = {o.x, o.y};
        } else if (misfit == 2) {
            const cf32 lo = cf_log({o.x, o.y}), lc = cf_log({c.x, c.y});
            res = {lo.r - lc.r, lo.i - lc.i};
        } else if (misfit == 3) {
            res = {cf_arg({c.x, c.y}) - cf_arg({o.x, o.y}), 0.0f};
        } else {
            const cf32 wc = cf_mul(wien, {c.x, c.y});
            res = {(o.x - wc.r) / wabs, (o.y - wc.i) / wabs};
        }
        float rr = res.r, ri = res.i;
        acc += (double)rr * rr + (double)ri * ri;
        if ((double)o.x < 1e-20 && (double)o.y < 1e-20) { rr = 0.0f; ri = 0.0f; }
        bool last = true;
        for (int t2 = t + 1; t2 < ntr; t2++)
            if (krec[t2] == krec[t]) { last = false; break; }
        if (last) b[krec[t]] = {(double)(rr / (dh * dh)), (double)(-ri / (dh * dh))};
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) l2[s] = part[0];
}

// Per-shot weights wien/|wien| of ass_grad_AC/TE (src/AC/ass_grad_AC.c:30,
// src/TE/ass_grad_TE.c:31-35), staged 256 shots at a time in LDS; 1 without
// an estimate.  Each workgroup of 256 cells walks all shots this way.
constexpr int kGradShotChunk = 256;
__device__ __forceinline__ void stage_stf_weights(const cx<float> *__restrict__ stf, int s0, int nshots, double *wr,
                                                  double *wi) {
#pragma clang fp contract(off)
    const int s = s0 + (int)threadIdx.x;
    if (threadIdx.x < kGradShotChunk && s < nshots) {
        if (stf) {
            const cf32 w = {stf[s].x, stf[s].y};
            const float a = cf_abs(w);
            wr[threadIdx.x] = (double)(w.r / a);
            wi[threadIdx.x] = (double)(w.i / a);
        } else {
            wr[threadIdx.x] = 1.0;
            wi[threadIdx.x] = 0.0;
        }
    }
}

// ---- per-shot gradient tapers (ass_grad_AC/TE with SWS_TAPER_CIRCULAR_PER_SHOT) ----
// taper_grad_shot (mode 1, src/taper_grad_shot.c:34-87): the shot's gradient
// is multiplied by f0 + (1 - f0) (0.5 - 0.5 cos(pi r/R))^a within r <= R of
// its source and of every receiver, r measured from (ix DH, iy DH) of cell
// (ix, iy) -- the PML-extended grid's 1-based indices -- to srcpos (m) and to
// recpos * DH, in float as the reference rounds them.  taper_grad_shot1
// (mode 2, src/taper_grad_shot1.c:32-66): the shot's gradient is 1e-5 on the
// (2 FILTSIZE + 1)^2 boxes around iround(srcpos / DH) of every source of the
// survey and around every receiver.  The products here accumulate in double,
// as the rest of the gradient does.
__device__ double taper_at(float x, float y, float cx, float cy, const ShotTaperDev &t) {
    const float dx = x - cx, dy = y - cy;
    const float r = (float)sqrt((double)dx * dx + (double)dy * dy);
    if (r > t.radius) return 1.0;
    const float v = t.f0 + (1.0 - t.f0) * pow(0.5 - 0.5 * cos(3.141592653589793 * r / t.radius), (double)t.a);
    return v;
}

// the taper factors of cell q that do not depend on the shot: the receivers'
// product (mode 1), or whether the cell is in a source or receiver box (mode 2)
__device__ void taper_cell(int q, const ShotTaperDev &t, double &trec, bool &boxed) {
    trec = 1.0;
    boxed = false;
    const int ix = q % t.nx + 1, iy = q / t.nx + 1;
    if (t.mode == 1) {
        const float x = ix * t.dh, y = iy * t.dh;
        for (int r = 0; r < t.nrec; r++) trec *= taper_at(x, y, t.rec_i[r] * t.dh, t.rec_j[r] * t.dh, t);
    } else if (t.mode == 2) {
        for (int k = 0; k < t.nsrc && !boxed; k++)
            boxed = abs(ix - t.src_i[k]) <= t.filtsize && abs(iy - t.src_j[k]) <= t.filtsize;
        for (int r = 0; r < t.nrec && !boxed; r++)
            boxed = abs(ix - t.rec_i[r]) <= t.filtsize && abs(iy - t.rec_j[r]) <= t.filtsize;
    }
}


// src/AC/ass_grad_AC.c:30-31 summed over shots in FP64: cell q uses the
// solutions at q+1 (src/vec2mat.c:19-24); the last cell reads past the
// solution in the reference and gets 0 here.  Shot s is weighted by
// wien_s/|wien_s| (stf NULL: 1), so the sum is Re(Omega^2 sum_s w_s u_s lam_s).
// One thread per cell: the grid covers all n cells.
__global__ void __launch_bounds__(256) k_grad_ac(int n, int nshots, double o2r, double o2i,
                                                 const cx<double> *__restrict__ U, const cx<double> *__restrict__ L,
                                                 const float *__restrict__ vp, const float *__restrict__ rho,
                                                 const cx<float> *__restrict__ stf, ShotTaperDev tp,
                                                 double *__restrict__ grad) {
    __shared__ double wr[kGradShotChunk], wi[kGradShotChunk];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = q + 1 < n;
    double trec = 1.0;
    bool boxed = false;
    if (on && tp.mode) taper_cell(q, tp, trec, boxed);
    const float x = (q % max(tp.nx, 1) + 1) * tp.dh, y = (q / max(tp.nx, 1) + 1) * tp.dh;
    double ar = 0.0, ai = 0.0;
    for (int s0 = 0; s0 < nshots; s0 += kGradShotChunk) {
        stage_stf_weights(stf, s0, nshots, wr, wi);
        __syncthreads();
        const int cnt = min(kGradShotChunk, nshots - s0);
        if (on && !boxed)
            for (int t = 0; t < cnt; t++) {
                const cx<double> u = U[q + 1 + (int64_t)(s0 + t) * n], l = L[q + 1 + (int64_t)(s0 + t) * n];
                const double pr = u.x * l.x - u.y * l.y, pi = u.x * l.y + u.y * l.x;
                const double ts = tp.mode == 1 ? taper_at(x, y, tp.srcx[s0 + t], tp.srcy[s0 + t], tp) : 1.0;
                ar += ts * (wr[t] * pr - wi[t] * pi);
                ai += ts * (wr[t] * pi + wi[t] * pr);
            }
        __syncthreads();
    }
    if (on) {
        const double v = vp[q];
        if (boxed) grad[q] += nshots * (double)1e-5f;   // taper_grad_shot1's value, once per shot
        else grad[q] += trec * (-2.0 * (o2r * ar - o2i * ai) / ((double)rho[q] * v * v * v));
    }
}

// src/AC/hessian_shin_AC.c:117-124 summed over shots in FP64: Shin's
// pseudo-Hessian diagonal, hess += |J|^2 with J = 2 Omega^2 u / (rho vp^3),
// i.e. 4 |Omega^2|^2 |u|^2 / (rho vp^3)^2; cell q uses the solution at q+1
// (vec2mat's shift) and the last cell is left unchanged, as in k_grad_ac.
__global__ void k_hess_shin_ac(int n, int nshots, double o2abs2, const cx<double> *__restrict__ U,
                               const float *__restrict__ vp, const float *__restrict__ rho, double *__restrict__ hess) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        if (q + 1 >= n) continue;
        double acc = 0.0;
        for (int s = 0; s < nshots; s++) {
            const cx<double> u = U[q + 1 + (int64_t)s * n];
            acc += u.x * u.x + u.y * u.y;
        }
        const double v = vp[q], d = (double)rho[q] * v * v * v;
        hess[q] += 4.0 * o2abs2 * acc / (d * d);
    }
}

// src/AC/hessian_AC.c:129-150 summed over shots and receivers in FP64: the
// Gauss-Newton diagonal, hess += |2 Omega^2 u g / (rho vp^3)|^2 for every
// (shot field u, receiver Green's function g) pair, which factorises into
// 4 |Omega^2|^2 (sum_s |u_s|^2) (sum_r |g_r|^2) / (rho vp^3)^2 -- so the
// receivers' Green's functions are one panel solve per frequency, shared by
// all shots, where the reference solves them again for every shot.  Cell q
// uses the solutions at q+1; the last cell is left unchanged.
__global__ void k_hess_gn_ac(int n, int nshots, int nrec, double o2abs2, const cx<double> *__restrict__ U,
                             const cx<double> *__restrict__ Gr, const float *__restrict__ vp,
                             const float *__restrict__ rho, double *__restrict__ hess) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        if (q + 1 >= n) continue;
        double au = 0.0, ag = 0.0;
        for (int s = 0; s < nshots; s++) {
            const cx<double> u = U[q + 1 + (int64_t)s * n];
            au += u.x * u.x + u.y * u.y;
        }
        for (int r = 0; r < nrec; r++) {
            const cx<double> g = Gr[q + 1 + (int64_t)r * n];
            ag += g.x * g.x + g.y * g.y;
        }
        const double v = vp[q], d = (double)rho[q] * v * v * v;
        hess[q] += 4.0 * o2abs2 * au * ag / (d * d);
    }
}

// src/TE/hessian_TE.c:149-157 summed over shots and receivers in FP64: the
// Gauss-Newton diagonals hess_sigma += |(i Omega) u g|^2, hess_eps +=
// |Omega^2 u g|^2, factorised like k_hess_gn_ac (without the reference's
// MAT*_NORM^2 scaling, :187-188).  Cell q uses the solutions at q+1.
__global__ void k_hess_gn_te(int n, int nshots, int nrec, double om_abs2, double o2_abs2,
                             const cx<double> *__restrict__ U, const cx<double> *__restrict__ Gr,
                             double *__restrict__ hsig, double *__restrict__ heps) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        if (q + 1 >= n) continue;
        double au = 0.0, ag = 0.0;
        for (int s = 0; s < nshots; s++) {
            const cx<double> u = U[q + 1 + (int64_t)s * n];
            au += u.x * u.x + u.y * u.y;
        }
        for (int r = 0; r < nrec; r++) {
            const cx<double> g = Gr[q + 1 + (int64_t)r * n];
            ag += g.x * g.x + g.y * g.y;
        }
        hsig[q] += om_abs2 * au * ag;
        heps[q] += o2_abs2 * au * ag;
    }
}

// src/TE/hessian_shin_TE.c:124-128 summed over shots in FP64: hess_sigma +=
// |(i Omega) u|^2 = |Omega|^2 |u|^2, hess_eps += |Omega^2 u|^2 = |Omega^2|^2 |u|^2
// (the reference's MAT1_NORM^2 / MAT2_NORM^2 scaling of the sums, :147-148, is
// the caller's); cell q uses the solution at q+1, the last cell is left unchanged.
__global__ void k_hess_shin_te(int n, int nshots, double om_abs2, double o2_abs2, const cx<double> *__restrict__ U,
                               double *__restrict__ hsig, double *__restrict__ heps) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        if (q + 1 >= n) continue;
        double acc = 0.0;
        for (int s = 0; s < nshots; s++) {
            const cx<double> u = U[q + 1 + (int64_t)s * n];
            acc += u.x * u.x + u.y * u.y;
        }
        hsig[q] += om_abs2 * acc;
        heps[q] += o2_abs2 * acc;
    }
}

// src/AC/apply_hess_AC.c:20-37 (also the TE path's, src/TE/fwi_FD_TE.c:287-288):
// grad *= 1 / (hess + EPS_HESS * max|hess|), in FP64.  One workgroup does the
// whole array -- the maximum first, then the scaling -- so the call needs no
// scratch memory; it runs once per FWI iteration, not on the hot path.
__global__ void __launch_bounds__(1024) k_apply_hess(int n, double eps, const double *__restrict__ hess,
                                                     double *__restrict__ grad) {
    __shared__ double part[1024];
    double mx = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) mx = fmax(mx, fabs(hess[q]));
    part[threadIdx.x] = mx;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) part[threadIdx.x] = fmax(part[threadIdx.x], part[threadIdx.x + o]);
        __syncthreads();
    }
    const double shift = eps * part[0];
    for (int q = threadIdx.x; q < n; q += blockDim.x) grad[q] *= 1.0 / (hess[q] + shift);
}

// src/calc_mat_change_wolfe.c:27-46: trial model m_new = m_old + alpha * dir
// in float, with the reference's hard constraints -- a value below lo, above
// hi or below zero keeps m_old.  dir is the FP64 search direction rounded to
// float, as the reference holds it; the product and the sum round separately
// (no fused multiply-add), so the result is the reference's to the bit.
// m_new may alias m_old.
__global__ void k_model_update(int n, float alpha, const double *__restrict__ dir, const float *m_old,
                               float lo, float hi, float *m_new) {
#pragma clang fp contract(off)
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const float old = m_old[q];
        float v = old + alpha * (float)dir[q];
        if (v < lo) v = old;
        if (v > hi) v = old;
        if (v < 0.0f) v = old;
        m_new[q] = v;
    }
}

// The TE line search's trial model (src/TE/parabolicls_TE.c:30-31,64-69 with
// src/calc_mat_change_wolfe_multi_para.c:22-50 and src/scale_grad.c:16-22):
// the material is normalised, r = (1/norm) * m_old, stepped, r + alpha * dir,
// clamped to [lo/norm, hi/norm], and scaled back, m_new = norm * r -- each
// operation rounded in float as the reference rounds it.  m_new may alias m_old.
__global__ void k_model_update_norm(int n, float alpha, const double *__restrict__ dir, const float *m_old,
                                    float norm, float lo, float hi, float *m_new) {
#pragma clang fp contract(off)
    const float inv = (float)(1.0 / (double)norm), lo_r = lo / norm, hi_r = hi / norm;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        float v = inv * m_old[q] + alpha * (float)dir[q];
        if (v < lo_r) v = lo_r;
        if (v > hi_r) v = hi_r;
        m_new[q] = norm * v;
    }
}

// precond (src/precond.c:17-26) for the filter and the horizontal taper:
// gauss_filt (src/gauss_filt.c:50-145) smooths the [ny, nx] field with the
// normalised Gaussian exp(-(ii^2/sx) - (jj^2/sy)) over a (2 hfsx + 1) x
// (2 hfsy + 1) window, the field extended by its edge values; taper_grad_hor
// (src/taper_grad_hor.c:22-31) then zeroes the first nzero rows.  hfsx < 0
// skips the filter.  Every workgroup builds the weights in LDS: each thread
// rounds exp to float, one thread sums them in the reference's order (x
// outer, y inner) and all divide, so the weights are the reference's; each
// output then sums its window in that order in float without fused
// multiply-adds.  The field is read as float (the reference holds it so).
constexpr int kPrecondMaxWeights = 16000;     // 64000 bytes of dynamic LDS
__global__ void __launch_bounds__(256) k_precond(int nx, int ny, int hfsx, int hfsy, float sx, float sy, int nzero,
                                                  const double *__restrict__ in, double *__restrict__ out) {
#pragma clang fp contract(off)
    extern __shared__ float wts[];             // [2 hfsx + 1][2 hfsy + 1], x outer
    __shared__ float total;
    const int wx = 2 * hfsx + 1, wy = 2 * hfsy + 1, nw = hfsx < 0 ? 0 : wx * wy;
    for (int t = threadIdx.x; t < nw; t += blockDim.x) {
        const int ii = t / wy - hfsx, jj = t % wy - hfsy;
        const float a = (float)(ii * ii) / sx, b = (float)(jj * jj) / sy;
        wts[t] = (float)exp((double)(-a - b));
    }
    __syncthreads();
    if (threadIdx.x == 0 && nw > 0) {
        float sum = 0.0f;
        for (int t = 0; t < nw; t++) sum += wts[t];
        total = sum;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nw; t += blockDim.x) wts[t] /= total;
    __syncthreads();
    const int64_t n = (int64_t)nx * ny;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(q / nx), i = (int)(q % nx);
        float v;
        if (nw == 0) {
            v = (float)in[q];
        } else {
            v = 0.0f;
            for (int ii = -hfsx; ii <= hfsx; ii++) {
                const int ic = min(max(i + ii, 0), nx - 1);
                const float *w = wts + (ii + hfsx) * wy + hfsy;
                for (int jj = -hfsy; jj <= hfsy; jj++) {
                    const int jc = min(max(j + jj, 0), ny - 1);
                    v += (float)in[(int64_t)jc * nx + ic] * w[jj];
                }
            }
        }
        out[q] = j < nzero ? 0.0 : (double)v;
    }
}

// Tikhonov regularisation of the TE materials (src/TE/Tikhonov_cost_TE.c:21-45,
// src/TE/Tikhonov_grad_TE.c:20-37): on the normalised fields r = a * m
// (a = BETA / NORM, in float) the five-point Laplacian of every interior cell,
// (four neighbours summed in float - 4 r) / DH^2 rounded to float as the
// reference rounds it; cost[0..1] = sum 0.5 BETA lap^2 per material (FP64
// here, float in the reference), and, when grad is given, grad += (LAMBDA *
// BETA) * lap.  One workgroup: it runs once per objective evaluation and
// needs no scratch memory.
__global__ void __launch_bounds__(1024) k_tikhonov_te(int nx, int ny, float dh, float a_sig, float a_eps,
                                                      float beta_sig, float beta_eps, float lb_sig, float lb_eps,
                                                      const float *__restrict__ sig, const float *__restrict__ eps,
                                                      double *__restrict__ g_sig, double *__restrict__ g_eps,
                                                      double *__restrict__ cost) {
#pragma clang fp contract(off)
    __shared__ double part[2][1024];
    const double h2 = (double)(dh * dh);
    double cs = 0.0, ce = 0.0;
    const int iw = nx - 2, cells = iw * (ny - 2);
    for (int c = threadIdx.x; c < cells; c += blockDim.x) {
        const int j = c / iw + 1, i = c % iw + 1;
        const int64_t q = (int64_t)j * nx + i;
        const float *m[2] = {sig, eps};
        const float av[2] = {a_sig, a_eps};
        float lap[2];
        for (int k = 0; k < 2; k++) {
            const float *f = m[k];
            const float w = av[k] * f[q - 1], e = av[k] * f[q + 1], s = av[k] * f[q - nx], nn = av[k] * f[q + nx];
            const float sum4 = w + e + s + nn;
            lap[k] = (float)(((double)sum4 - 4.0 * (double)(av[k] * f[q])) / h2);
        }
        cs += 0.5 * (double)beta_sig * (double)lap[0] * (double)lap[0];
        ce += 0.5 * (double)beta_eps * (double)lap[1] * (double)lap[1];
        if (g_sig) g_sig[q] += (double)(lb_sig * lap[0]);
        if (g_eps) g_eps[q] += (double)(lb_eps * lap[1]);
    }
    part[0][threadIdx.x] = cs;
    part[1][threadIdx.x] = ce;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            part[0][threadIdx.x] += part[0][threadIdx.x + o];
            part[1][threadIdx.x] += part[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cost[0] = part[0][0];
        cost[1] = part[1][0];
    }
}

// taper_grad_file (src/taper_grad_file.c:22-66): the model part of the [ny, nx]
// field (rows fsshift .. fsshift+ny0-1, columns npml .. npml+nx0-1) times the
// [ny0, nx0] taper in float, then the PML filled by the reference's copies --
// columns from the nearest model column, the rows below the model from its
// last row and, only with a free surface (the reference's condition), the
// first npml rows from row npml.  Those copies compose to one gather: cell
// (j, i) takes the tapered value at (row map of j, clamped i).
__global__ void k_taper_file(int nx, int ny, int nx0, int ny0, int npml, int fsshift, int free_surf,
                             const float *__restrict__ taper, const double *__restrict__ in,
                             double *__restrict__ out) {
    const int64_t n = (int64_t)nx * ny;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(q / nx), i = (int)(q % nx);
        const int ic = min(max(i, npml), npml + nx0 - 1);
        int jr = j;
        if (free_surf && j < npml) jr = npml;
        if (j >= fsshift + ny0) jr = fsshift + ny0 - 1;
        const double g = in[(int64_t)jr * nx + ic];
        const bool model = jr >= fsshift && jr < fsshift + ny0;
        out[q] = model ? (double)((float)g * taper[(int64_t)(jr - fsshift) * nx0 + (ic - npml)]) : g;
    }
}

// cp_grad_frame (src/cp_grad_frame.c:102-208, the PML_grad = 3 branch it
// always takes): the gradient in the PML is damped by coeff[b] = exp(-(a^2
// (npml - b)^2)) of the band b (1-based, 1 at the outer edge).  The
// reference's coefficient map is written band by band -- left, right, top
// (without a free surface), bottom, each band b spanning b .. len - b + 1 of
// its edge, later ones overwriting earlier ones -- and then every PML column
// and every PML row multiplies by it, so a corner cell is multiplied twice.
// Both compose to a rule per cell: the coefficient of the last band that
// covers it, applied once per PML column/row band it lies in, each product
// rounded to float.  (The reference's replacement of a NaN product by the
// value above it is not reproduced.)
__global__ void k_grad_frame(int nx, int ny, int nx0, int ny0, int npml, int fsshift, int free_surf, float a,
                             double *__restrict__ g) {
    const int64_t n = (int64_t)nx * ny;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int J = (int)(q / nx) + 1, I = (int)(q % nx) + 1;
        const int times = (I <= npml) + (I >= npml + nx0 + 1) + (!free_surf && J <= npml) + (J >= fsshift + ny0 + 1);
        if (times == 0) continue;
        int band = 0;                                   // 0: coefficient 1
        if (I <= npml && J >= I && J <= ny - I + 1) band = I;
        const int ir = nx - I + 1;
        if (ir <= npml && J >= ir && J <= ny - ir + 1) band = ir;
        if (!free_surf && J <= npml && I >= J && I <= nx - J + 1) band = J;
        const int jr = ny - J + 1;
        if (jr <= npml && I >= jr && I <= nx - jr + 1) band = jr;
        float c = 1.0f;
        if (band) {
            const float d = (float)(npml - band);
            c = (float)exp((double)(-(a * a * d * d)));
        }
        float v = (float)g[q];
        for (int t = 0; t < times; t++) v *= c;
        g[q] = v;
    }
}

// src/TE/ass_grad_TE.c:31-35 summed over shots in FP64, shot s weighted by
// wien_s/|wien_s| (stf NULL: 1): with z = sum_s w_s u_s lam_s,
// grad_sigma += Re((i Omega) z), grad_eps += Re(Omega^2 z).  Cell q uses the
// solutions at q+1; the last cell is left unchanged.  One thread per cell.
__global__ void __launch_bounds__(256) k_grad_te(int n, int nshots, double omr, double omi, double o2r, double o2i,
                                                 const cx<double> *__restrict__ U, const cx<double> *__restrict__ L,
                                                 const cx<float> *__restrict__ stf, ShotTaperDev tp,
                                                 double *__restrict__ gsig, double *__restrict__ geps) {
    __shared__ double wr[kGradShotChunk], wi[kGradShotChunk];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = q + 1 < n;
    double trec = 1.0;
    bool boxed = false;
    if (on && tp.mode) taper_cell(q, tp, trec, boxed);
    const float x = (q % max(tp.nx, 1) + 1) * tp.dh, y = (q / max(tp.nx, 1) + 1) * tp.dh;
    double ar = 0.0, ai = 0.0;
    for (int s0 = 0; s0 < nshots; s0 += kGradShotChunk) {
        stage_stf_weights(stf, s0, nshots, wr, wi);
        __syncthreads();
        const int cnt = min(kGradShotChunk, nshots - s0);
        if (on && !boxed)
            for (int t = 0; t < cnt; t++) {
                const cx<double> u = U[q + 1 + (int64_t)(s0 + t) * n], l = L[q + 1 + (int64_t)(s0 + t) * n];
                const double pr = u.x * l.x - u.y * l.y, pi = u.x * l.y + u.y * l.x;
                const double ts = tp.mode == 1 ? taper_at(x, y, tp.srcx[s0 + t], tp.srcy[s0 + t], tp) : 1.0;
                ar += ts * (wr[t] * pr - wi[t] * pi);
                ai += ts * (wr[t] * pi + wi[t] * pr);
            }
        __syncthreads();
    }
    if (on) {
        // Re((i Omega) z) = -(omr*ai + omi*ar);  Re(Omega^2 z) = o2r*ar - o2i*ai
        if (boxed) {
            gsig[q] += nshots * (double)1e-5f;
            geps[q] += nshots * (double)1e-5f;
        } else {
            gsig[q] += trec * -(omr * ai + omi * ar);
            geps[q] += trec * (o2r * ar - o2i * ai);
        }
    }
}

inline int grid_for(int64_t total) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 8192));
}

}  // namespace

void model_update(int n, float alpha, const double *dir, const float *m_old, float lo, float hi, float *m_new,
                  hipStream_t st) {
    hipLaunchKernelGGL(k_model_update, dim3(grid_for(n)), dim3(256), 0, st, n, alpha, dir, m_old, lo, hi, m_new);
    GFDFD_HIP(hipGetLastError());
}

void model_update_norm(int n, float alpha, const double *dir, const float *m_old, float norm, float lo, float hi,
                       float *m_new, hipStream_t st) {
    hipLaunchKernelGGL(k_model_update_norm, dim3(grid_for(n)), dim3(256), 0, st, n, alpha, dir, m_old, norm, lo, hi,
                       m_new);
    GFDFD_HIP(hipGetLastError());
}

void precond(int nx, int ny, int hfsx, int hfsy, float sx, float sy, int nzero, const double *in, double *out,
             hipStream_t st) {
    const int nw = hfsx < 0 ? 0 : (2 * hfsx + 1) * (2 * hfsy + 1);
    if (nw > kPrecondMaxWeights) throw std::invalid_argument("precond: filter window larger than 16000 weights");
    const size_t lds = sizeof(float) * std::max(nw, 1);
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)nx * ny + 255) / 256, 4096));
    hipLaunchKernelGGL(k_precond, dim3(blocks), dim3(256), lds, st, nx, ny, hfsx, hfsy, sx, sy, nzero, in, out);
    GFDFD_HIP(hipGetLastError());
}

void tikhonov_te(int nx, int ny, float dh, float a_sig, float a_eps, float beta_sig, float beta_eps, float lb_sig,
                 float lb_eps, const float *sig, const float *eps, double *g_sig, double *g_eps, double *cost,
                 hipStream_t st) {
    hipLaunchKernelGGL(k_tikhonov_te, dim3(1), dim3(1024), 0, st, nx, ny, dh, a_sig, a_eps, beta_sig, beta_eps, lb_sig,
                       lb_eps, sig, eps, g_sig, g_eps, cost);
    GFDFD_HIP(hipGetLastError());
}

void taper_file(int nx, int ny, int nx0, int ny0, int npml, int fsshift, int free_surf, const float *taper,
                const double *in, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_taper_file, dim3(grid_for((int64_t)nx * ny)), dim3(256), 0, st, nx, ny, nx0, ny0, npml,
                       fsshift, free_surf, taper, in, out);
    GFDFD_HIP(hipGetLastError());
}

void grad_frame(int nx, int ny, int nx0, int ny0, int npml, int fsshift, int free_surf, float a, double *g,
                hipStream_t st) {
    hipLaunchKernelGGL(k_grad_frame, dim3(grid_for((int64_t)nx * ny)), dim3(256), 0, st, nx, ny, nx0, ny0, npml,
                       fsshift, free_surf, a, g);
    GFDFD_HIP(hipGetLastError());
}

void spike_panel(int n, int nshots, const int *k, cx<double> amp, cx<double> *B, hipStream_t st) {
    hipLaunchKernelGGL(k_spikes, dim3(grid_for((int64_t)n * nshots)), dim3(256), 0, st, n, nshots, k, amp, B);
    GFDFD_HIP(hipGetLastError());
}

void sample_receivers(int n, int nshots, int ntr, const int *krec, const cx<double> *X, cx<float> *pcal,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_sample, dim3(grid_for((int64_t)ntr * nshots)), dim3(256), 0, st, n, nshots, ntr, krec, X,
                       pcal);
    GFDFD_HIP(hipGetLastError());
}

void adjoint_sources(int n, int nshots, int ntr, const int *krec, const cx<float> *pcal, const cx<float> *pobs,
                     float dh, int misfit, int stf_inv, int rtm, cx<double> *B, double *l2, cx<float> *stf,
                     hipStream_t st) {
    hipLaunchKernelGGL(k_adjoint, dim3(nshots), dim3(256), 0, st, n, nshots, ntr, krec, pcal, pobs, dh, misfit,
                       stf_inv, rtm, B, l2, stf);
    GFDFD_HIP(hipGetLastError());
}

void grad_ac(int n, int nshots, const float omega2[2], const cx<double> *U, const cx<double> *L, const float *vp,
             const float *rho, const cx<float> *stf, const ShotTaperDev &tp, double *grad, hipStream_t st) {
    hipLaunchKernelGGL(k_grad_ac, dim3((n + 255) / 256), dim3(256), 0, st, n, nshots, (double)omega2[0],
                       (double)omega2[1], U, L, vp, rho, stf, tp, grad);
    GFDFD_HIP(hipGetLastError());
}

void hess_shin_ac(int n, int nshots, const float omega2[2], const cx<double> *U, const float *vp, const float *rho,
                  double *hess, hipStream_t st) {
    const double o2r = omega2[0], o2i = omega2[1];
    hipLaunchKernelGGL(k_hess_shin_ac, dim3(grid_for(n)), dim3(256), 0, st, n, nshots, o2r * o2r + o2i * o2i, U, vp,
                       rho, hess);
    GFDFD_HIP(hipGetLastError());
}

void apply_hess(int n, float eps, const double *hess, double *grad, hipStream_t st) {
    hipLaunchKernelGGL(k_apply_hess, dim3(1), dim3(1024), 0, st, n, (double)eps, hess, grad);
    GFDFD_HIP(hipGetLastError());
}

void hess_gn_ac(int n, int nshots, int nrec, const float omega2[2], const cx<double> *U, const cx<double> *Gr,
                const float *vp, const float *rho, double *hess, hipStream_t st) {
    const double o2r = omega2[0], o2i = omega2[1];
    hipLaunchKernelGGL(k_hess_gn_ac, dim3(grid_for(n)), dim3(256), 0, st, n, nshots, nrec, o2r * o2r + o2i * o2i, U,
                       Gr, vp, rho, hess);
    GFDFD_HIP(hipGetLastError());
}

void hess_gn_te(int n, int nshots, int nrec, const float omega[2], const float omega2[2], const cx<double> *U,
                const cx<double> *Gr, double *hsig, double *heps, hipStream_t st) {
    const double omr = omega[0], omi = omega[1], o2r = omega2[0], o2i = omega2[1];
    hipLaunchKernelGGL(k_hess_gn_te, dim3(grid_for(n)), dim3(256), 0, st, n, nshots, nrec, omr * omr + omi * omi,
                       o2r * o2r + o2i * o2i, U, Gr, hsig, heps);
    GFDFD_HIP(hipGetLastError());
}

void hess_shin_te(int n, int nshots, const float omega[2], const float omega2[2], const cx<double> *U, double *hsig,
                  double *heps, hipStream_t st) {
    const double omr = omega[0], omi = omega[1], o2r = omega2[0], o2i = omega2[1];
    hipLaunchKernelGGL(k_hess_shin_te, dim3(grid_for(n)), dim3(256), 0, st, n, nshots, omr * omr + omi * omi,
                       o2r * o2r + o2i * o2i, U, hsig, heps);
    GFDFD_HIP(hipGetLastError());
}

void grad_te(int n, int nshots, const float omega[2], const float omega2[2], const cx<double> *U,
             const cx<double> *L, const cx<float> *stf, const ShotTaperDev &tp, double *gsig, double *geps,
             hipStream_t st) {
    hipLaunchKernelGGL(k_grad_te, dim3((n + 255) / 256), dim3(256), 0, st, n, nshots, (double)omega[0],
                       (double)omega[1], (double)omega2[0], (double)omega2[1], U, L, stf, tp, gsig, geps);
    GFDFD_HIP(hipGetLastError());
}

}  // namespace gfdfd
