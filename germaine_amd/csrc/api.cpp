// gfdfd extension C-ABI (include/gfdfd.h): thin wrappers that validate
// arguments, catch exceptions, set the status and print a diagnostic.
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <map>
#include <mutex>

#include "../../include/gfdfd.h"
#include "assemble.hpp"
#include "engine.hpp"
#include "prof.hpp"
#include "runtime.hpp"
#include "shot.hpp"

using namespace gfdfd;

struct gfdfd_plan_s {
    std::shared_ptr<const Plan> host;
    std::shared_ptr<DevPlan> dev;   // uploaded on first device use
    DevPlan &device() {
        if (!dev) dev = device_plan(host);
        return *dev;
    }
};

struct gfdfd_factors_s {
    int precision = 64;
    std::unique_ptr<Factors<double>> f64;
    std::unique_ptr<Factors<float>> f32;
};

namespace gfdfd {

namespace {
thread_local std::string g_last_error;
std::mutex g_cache_mu;
std::map<std::tuple<int, int, int, int>, std::shared_ptr<const Plan>> g_host_cache;
std::map<const Plan *, std::shared_ptr<DevPlan>> g_dev_cache;
}  // namespace

int report(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    fprintf(stderr, "gfdfd: %s\n", buf);
    fflush(stderr);
    return code;
}

std::shared_ptr<const Plan> cached_host_plan(int nx, int ny, const PlanOptions &opt) {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto key = std::make_tuple(nx, ny, opt.leaf_area, opt.small_max);
    auto it = g_host_cache.find(key);
    if (it != g_host_cache.end()) return it->second;
    auto pl = std::make_shared<const Plan>(build_grid_plan(nx, ny, opt));
    g_host_cache[key] = pl;
    return pl;
}

std::shared_ptr<DevPlan> device_plan(std::shared_ptr<const Plan> pl) {
    ensure_device();
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto it = g_dev_cache.find(pl.get());
    if (it != g_dev_cache.end()) return it->second;
    auto dp = std::make_shared<DevPlan>(pl);
    g_dev_cache[pl.get()] = dp;
    return dp;
}

std::shared_ptr<DevPlan> cached_grid_plan(int nx, int ny, const PlanOptions &opt) {
    return device_plan(cached_host_plan(nx, ny, opt));
}

template <typename T>
std::unique_ptr<Factors<T>> make_factors(std::shared_ptr<DevPlan> dp, int nfreq, const cx<double> *vals, bool copy_vals,
                                         hipStream_t st) {
    const Plan &pl = *dp->host;
    auto F = std::make_unique<Factors<T>>();
    F->plan = dp;
    F->nfreq = nfreq;
    F->fstride = pl.arena;
    F->arena = dev_alloc<cx<T>>((size_t)pl.arena * nfreq);
    F->piv = dev_alloc<int>((size_t)pl.n * nfreq);
    F->rowperm = dev_alloc<int>((size_t)pl.n * nfreq);
    F->sing = dev_alloc<int>(nfreq);
    F->gsrc = dev_alloc<int>((size_t)pl.fv_rows * 3 * nfreq);
    if (dp->nroots > 0) F->sched = dev_alloc<int>(1);
    if (copy_vals) {
        F->vals = dev_alloc<cx<double>>((size_t)pl.nnz * nfreq);
        F->own_vals = true;
        GFDFD_HIP(hipMemcpyAsync(F->vals, vals, sizeof(cx<double>) * pl.nnz * nfreq, hipMemcpyDeviceToDevice, st));
    } else {
        F->vals = const_cast<cx<double> *>(vals);
    }
    factor_numeric<T>(*dp, *F, F->vals, pl.nnz, st);
    return F;
}
// new values of the same pattern into existing factors (storage reused)
template <typename T>
void refactor(Factors<T> &F, const cx<double> *vals, hipStream_t st) {
    const Plan &pl = *F.plan->host;
    if (!F.own_vals) throw std::invalid_argument("factors do not own their values");
    GFDFD_HIP(hipMemcpyAsync(F.vals, vals, sizeof(cx<double>) * pl.nnz * F.nfreq, hipMemcpyDeviceToDevice, st));
    factor_numeric<T>(*F.plan, F, F.vals, pl.nnz, st);
}

// the same as one graph launch: the first call with a given source buffer
// captures the copy and every factorisation kernel on a private stream; later
// calls with the same buffer replay the graph on st
template <typename T>
void refactor_graph(Factors<T> &F, const cx<double> *vals, hipStream_t st) {
    if (!F.graph || F.graph_src != vals) {
        ProfPause no_events;   // captured launches carry no profiling events
        if (F.graph) {
            GFDFD_HIP(hipGraphExecDestroy(F.graph));
            F.graph = nullptr;
        }
        hipStream_t cs;
        GFDFD_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipGraph_t g = nullptr;
        GFDFD_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        try {
            refactor<T>(F, vals, cs);
        } catch (...) {
            (void)hipStreamEndCapture(cs, &g);
            if (g) (void)hipGraphDestroy(g);
            (void)hipStreamDestroy(cs);
            throw;
        }
        GFDFD_HIP(hipStreamEndCapture(cs, &g));
        const hipError_t e = hipGraphInstantiate(&F.graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        (void)hipStreamDestroy(cs);
        GFDFD_HIP(e);
        F.graph_src = vals;
    }
    F.vinv_ready = false;   // the replay changes the factors
    GFDFD_HIP(hipGraphLaunch(F.graph, st));
}

template std::unique_ptr<Factors<double>> make_factors<double>(std::shared_ptr<DevPlan>, int, const cx<double> *, bool,
                                                               hipStream_t);
template std::unique_ptr<Factors<float>> make_factors<float>(std::shared_ptr<DevPlan>, int, const cx<double> *, bool,
                                                             hipStream_t);

}  // namespace gfdfd

#define GUARD_BEGIN try {
#define GUARD_END                                                                          \
    }                                                                                      \
    catch (const hip_error &e) {                                                           \
        return report(GFDFD_ERR_DEVICE, "%s: %s", __func__, e.what());                     \
    }                                                                                      \
    catch (const std::bad_alloc &e) {                                                      \
        return report(GFDFD_ERR_MEMORY, "%s: out of host memory", __func__);               \
    }                                                                                      \
    catch (const std::invalid_argument &e) {                                               \
        return report(GFDFD_ERR_ARG, "%s: %s", __func__, e.what());                        \
    }                                                                                      \
    catch (const std::exception &e) {                                                      \
        return report(GFDFD_ERR_INTERNAL, "%s: %s", __func__, e.what());                   \
    }

static hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

extern "C" {

const char *gfdfd_last_error(void) { return g_last_error.c_str(); }

int gfdfd_plan_grid_opt(int nx, int ny, int leaf_area, int small_max, gfdfd_plan *out) {
    if (!out) return report(GFDFD_ERR_ARG, "gfdfd_plan_grid: out is NULL");
    *out = nullptr;
    if (nx < 2 || ny < 2) return report(GFDFD_ERR_ARG, "gfdfd_plan_grid: grid %dx%d too small", nx, ny);
    if ((long long)nx * ny > 0x7fffffffLL / 9)
        return report(GFDFD_ERR_ARG, "gfdfd_plan_grid: grid %dx%d exceeds 32-bit indexing", nx, ny);
    if (leaf_area < 1 || small_max < 8 || small_max > 90)
        return report(GFDFD_ERR_ARG, "gfdfd_plan_grid: leaf_area %d / small_max %d out of range", leaf_area, small_max);
    GUARD_BEGIN
    PlanOptions o;
    o.leaf_area = leaf_area;
    o.small_max = small_max;
    *out = new gfdfd_plan_s{cached_host_plan(nx, ny, o), nullptr};
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_plan_grid(int nx, int ny, gfdfd_plan *out) {
    PlanOptions o;
    return gfdfd_plan_grid_opt(nx, ny, o.leaf_area, o.small_max, out);
}

void gfdfd_plan_free(gfdfd_plan *plan) {
    if (plan && *plan) {
        delete *plan;
        *plan = nullptr;
    }
}

int gfdfd_plan_query(gfdfd_plan plan, gfdfd_plan_info *info) {
    if (!plan || !info) return report(GFDFD_ERR_ARG, "gfdfd_plan_query: NULL argument");
    const Plan &pl = *plan->host;
    info->n = pl.n;
    info->nnz = pl.nnz;
    info->fronts = (int)pl.fronts.size();
    info->levels = pl.nlevels;
    info->max_front = pl.max_m;
    info->max_pivots = pl.max_p;
    info->arena_elems = pl.arena;
    info->work_rows = pl.fv_rows;
    info->factor_flops = pl.factor_flops;
    info->nnz_lu = pl.nnz_lu;
    info->plan_device_bytes = plan->dev ? (double)plan->dev->bytes : 0.0;
    return GFDFD_OK;
}

int gfdfd_plan_pattern(gfdfd_plan plan, int *colptr, int *rowind) {
    if (!plan || !colptr || !rowind) return report(GFDFD_ERR_ARG, "gfdfd_plan_pattern: NULL argument");
    const Plan &pl = *plan->host;
    std::memcpy(colptr, pl.colptr.data(), sizeof(int) * pl.colptr.size());
    std::memcpy(rowind, pl.rowind.data(), sizeof(int) * pl.rowind.size());
    return GFDFD_OK;
}

int gfdfd_plan_perm(gfdfd_plan plan, int *perm) {
    if (!plan || !perm) return report(GFDFD_ERR_ARG, "gfdfd_plan_perm: NULL argument");
    const Plan &pl = *plan->host;
    std::memcpy(perm, pl.perm.data(), sizeof(int) * pl.perm.size());
    return GFDFD_OK;
}

int gfdfd_plan_fronts(gfdfd_plan plan, int *p, int *m, int *pos0, int *parent, int *level, long long *idx_off) {
    if (!plan || !p || !m || !pos0 || !parent || !level || !idx_off)
        return report(GFDFD_ERR_ARG, "gfdfd_plan_fronts: NULL argument");
    const Plan &pl = *plan->host;
    for (size_t f = 0; f < pl.fronts.size(); f++) {
        const Front &F = pl.fronts[f];
        p[f] = F.p; m[f] = F.m; pos0[f] = F.pos0; parent[f] = F.parent; level[f] = F.level;
        idx_off[f] = F.idx_off;
    }
    return GFDFD_OK;
}

int gfdfd_plan_index(gfdfd_plan plan, int *idx) {
    if (!plan || !idx) return report(GFDFD_ERR_ARG, "gfdfd_plan_index: NULL argument");
    const Plan &pl = *plan->host;
    std::memcpy(idx, pl.idx.data(), sizeof(int) * pl.idx.size());
    return GFDFD_OK;
}

void gfdfd_fd_weights(float dh, float *out6);
float gfdfd_mu0(void);

int gfdfd_assemble(gfdfd_plan plan, int physics, float dh, int nfreq, const gfdfd_freq *freqs, const float *d_m1,
                   const float *d_m2, double *d_vals, void *stream) {
    if (!plan || !freqs || !d_m1 || !d_m2 || !d_vals || nfreq < 1)
        return report(GFDFD_ERR_ARG, "gfdfd_assemble: missing argument or nfreq < 1");
    if (physics < 1 || physics > 3) return report(GFDFD_ERR_ARG, "gfdfd_assemble: physics %d", physics);
    if (!(dh > 0)) return report(GFDFD_ERR_ARG, "gfdfd_assemble: DH must be positive");
    GUARD_BEGIN
    const Plan &pl = *plan->host;
    if (!pl.grid) return report(GFDFD_ERR_ARG, "gfdfd_assemble: plan is not a grid plan");
    DevPlan &dp = plan->device();
    const int nx = pl.nx, ny = pl.ny;
    const size_t plen = 4 * (size_t)(nx + 2) + 4 * (size_t)(ny + 2);
    std::vector<float> prof(plen * nfreq);
    std::vector<AsmFreq> fr(nfreq);
    for (int f = 0; f < nfreq; f++) {
        if (!freqs[f].prof) return report(GFDFD_ERR_ARG, "gfdfd_assemble: frequency %d has no PML profile", f);
        std::memcpy(&prof[plen * f], freqs[f].prof, plen * sizeof(float));
        fr[f] = {freqs[f].omega2[0], freqs[f].omega2[1], freqs[f].omega[0], freqs[f].omega[1]};
    }
    hipStream_t st = S(stream);
    float *dprof = dev_alloc<float>(prof.size());
    AsmFreq *dfr = dev_alloc<AsmFreq>(nfreq);
    GFDFD_HIP(hipMemcpyAsync(dprof, prof.data(), prof.size() * sizeof(float), hipMemcpyHostToDevice, st));
    GFDFD_HIP(hipMemcpyAsync(dfr, fr.data(), fr.size() * sizeof(AsmFreq), hipMemcpyHostToDevice, st));
    float w[6];
    gfdfd_fd_weights(dh, w);
    AsmScalars sc{w[0], w[1], w[2], w[3], w[4], w[5], gfdfd_mu0(), dfr};
    assemble_device(physics, nx, ny, sc, d_m1, d_m2, dprof, dp.colptr, reinterpret_cast<cx<double> *>(d_vals),
                    pl.nnz, nfreq, st);
    GFDFD_HIP(hipStreamSynchronize(st));
    dev_free(dprof);
    dev_free(dfr);
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_factor(gfdfd_plan plan, int nfreq, const double *d_vals, int precision, void *stream, gfdfd_factors *out) {
    if (!out) return report(GFDFD_ERR_ARG, "gfdfd_factor: out is NULL");
    *out = nullptr;
    if (!plan || !d_vals || nfreq < 1) return report(GFDFD_ERR_ARG, "gfdfd_factor: missing argument or nfreq < 1");
    if (precision != GFDFD_FP64 && precision != GFDFD_FP32)
        return report(GFDFD_ERR_ARG, "gfdfd_factor: precision must be 64 or 32");
    GUARD_BEGIN
    auto h = std::make_unique<gfdfd_factors_s>();
    h->precision = precision;
    const cx<double> *v = reinterpret_cast<const cx<double> *>(d_vals);
    plan->device();
    if (precision == GFDFD_FP64)
        h->f64 = make_factors<double>(plan->dev, nfreq, v, true, S(stream));
    else
        h->f32 = make_factors<float>(plan->dev, nfreq, v, true, S(stream));
    *out = h.release();
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_release_workspace(void) {
    GUARD_BEGIN
    GFDFD_HIP(hipDeviceSynchronize());
    release_solve_workspaces();
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_factor_tuning(int panel_lds_cap, int persistent, int deep_min_m) {
    if (panel_lds_cap < 0) return report(GFDFD_ERR_ARG, "gfdfd_factor_tuning: negative panel LDS cap");
    if (deep_min_m < 0) return report(GFDFD_ERR_ARG, "gfdfd_factor_tuning: negative deep_min_m");
    set_factor_tuning(panel_lds_cap, persistent, deep_min_m);
    return GFDFD_OK;
}

int gfdfd_solve_tuning(int rw_cap, int fused_mmax, int refine_layout) {
    GUARD_BEGIN
    set_solve_tuning(rw_cap, fused_mmax, refine_layout);
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_refactor(gfdfd_factors f, const double *d_vals, void *stream) {
    if (!f || !d_vals) return report(GFDFD_ERR_ARG, "gfdfd_refactor: NULL argument");
    GUARD_BEGIN
    const cx<double> *v = reinterpret_cast<const cx<double> *>(d_vals);
    if (f->f64) refactor<double>(*f->f64, v, S(stream));
    else refactor<float>(*f->f32, v, S(stream));
    return GFDFD_OK;
    GUARD_END
}

void gfdfd_factors_free(gfdfd_factors *f) {
    if (f && *f) {
        delete *f;
        *f = nullptr;
    }
}

int gfdfd_refactor_graph(gfdfd_factors f, const double *d_vals, void *stream) {
    if (!f || !d_vals) return report(GFDFD_ERR_ARG, "gfdfd_refactor_graph: NULL argument");
    GUARD_BEGIN
    const cx<double> *v = reinterpret_cast<const cx<double> *>(d_vals);
    if (f->f64) refactor_graph<double>(*f->f64, v, S(stream));
    else refactor_graph<float>(*f->f32, v, S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_factors_singular(gfdfd_factors f, int *flags) {
    if (!f || !flags) return report(GFDFD_ERR_ARG, "gfdfd_factors_singular: NULL argument");
    GUARD_BEGIN
    int nf = f->f64 ? f->f64->nfreq : f->f32->nfreq;
    const int *d = f->f64 ? f->f64->sing : f->f32->sing;
    GFDFD_HIP(hipMemcpy(flags, d, sizeof(int) * nf, hipMemcpyDeviceToHost));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_solve(gfdfd_factors f, int ifreq, int sys, int nrhs, const double *d_b, double *d_x, void *stream) {
    if (!f || !d_b || !d_x || nrhs < 1) return report(GFDFD_ERR_ARG, "gfdfd_solve: missing argument or nrhs < 1");
    if (sys < 0 || sys > 2) return report(GFDFD_ERR_ARG, "gfdfd_solve: invalid system %d", sys);
    GUARD_BEGIN
    auto B = reinterpret_cast<const cx<double> *>(d_b);
    auto X = reinterpret_cast<cx<double> *>(d_x);
    if (f->f64)
        solve_panel<double, double>(*f->f64, ifreq, sys, nrhs, B, X, S(stream));
    else
        solve_panel<float, double>(*f->f32, ifreq, sys, nrhs, B, X, S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_solve_noalloc(gfdfd_factors f, int ifreq, int sys, int nrhs, const double *d_b, double *d_x,
                        void *stream) {
    if (!f || !d_b || !d_x) return report(GFDFD_ERR_ARG, "gfdfd_solve_noalloc: missing argument");
    GUARD_BEGIN
    auto B = reinterpret_cast<const cx<double> *>(d_b);
    auto X = reinterpret_cast<cx<double> *>(d_x);
    if (f->f64)
        solve_noalloc<double>(*f->f64, ifreq, sys, nrhs, B, X, S(stream));
    else
        solve_noalloc<float>(*f->f32, ifreq, sys, nrhs, B, X, S(stream));
    return GFDFD_OK;
    GUARD_END
}

long long gfdfd_solve_noalloc_limit(void) { return (long long)kNoAllocScratchBytes; }

int gfdfd_solve_refine(gfdfd_factors f, int ifreq, int sys, int nrhs, const double *d_b, double *d_x, double tol,
                       int max_iter, int *iters, double *relres, void *stream) {
    if (!f || !d_b || !d_x || nrhs < 1) return report(GFDFD_ERR_ARG, "gfdfd_solve_refine: missing argument");
    if (d_b == d_x) return report(GFDFD_ERR_ARG, "gfdfd_solve_refine: d_x must not alias d_b");
    if (sys < 0 || sys > 2) return report(GFDFD_ERR_ARG, "gfdfd_solve_refine: invalid system %d", sys);
    GUARD_BEGIN
    auto B = reinterpret_cast<const cx<double> *>(d_b);
    auto X = reinterpret_cast<cx<double> *>(d_x);
    RefineStats rs = f->f64 ? solve_refined<double>(*f->f64, ifreq, sys, nrhs, B, X, tol, max_iter, S(stream))
                            : solve_refined<float>(*f->f32, ifreq, sys, nrhs, B, X, tol, max_iter, S(stream));
    if (iters) *iters = rs.iterations;
    if (relres) *relres = rs.max_rel_residual;
    return GFDFD_OK;
    GUARD_END
}

static int check_sel(gfdfd_factors f, int nsel, const int *freqs, const char *fn) {
    if (!f || !freqs || nsel < 1 || nsel > kMaxBatch)
        return report(GFDFD_ERR_ARG, "%s: need 1..%d frequencies", fn, kMaxBatch);
    const int nf = f->f64 ? f->f64->nfreq : f->f32->nfreq;
    for (int z = 0; z < nsel; z++)
        if (freqs[z] < 0 || freqs[z] >= nf) return report(GFDFD_ERR_ARG, "%s: frequency %d out of range", fn, freqs[z]);
    return GFDFD_OK;
}

int gfdfd_solve_batch(gfdfd_factors f, int nsel, const int *freqs, int sys, int nrhs, const double *d_b, double *d_x,
                      void *stream) {
    if (int e = check_sel(f, nsel, freqs, "gfdfd_solve_batch")) return e;
    if (!d_b || !d_x || nrhs < 1) return report(GFDFD_ERR_ARG, "gfdfd_solve_batch: missing argument or nrhs < 1");
    if (sys < 0 || sys > 2) return report(GFDFD_ERR_ARG, "gfdfd_solve_batch: invalid system %d", sys);
    GUARD_BEGIN
    std::vector<int> fr(freqs, freqs + nsel), sl(nsel);
    for (int z = 0; z < nsel; z++) sl[z] = z;
    auto B = reinterpret_cast<const cx<double> *>(d_b);
    auto X = reinterpret_cast<cx<double> *>(d_x);
    if (f->f64)
        solve_batch<double, double>(*f->f64, fr, sl, nsel, sys, nrhs, B, X, S(stream));
    else
        solve_batch<float, double>(*f->f32, fr, sl, nsel, sys, nrhs, B, X, S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_solve_refine_batch(gfdfd_factors f, int nsel, const int *freqs, int sys, int nrhs, const double *d_b,
                             double *d_x, double tol, int max_iter, int *iters, double *relres, void *stream) {
    if (int e = check_sel(f, nsel, freqs, "gfdfd_solve_refine_batch")) return e;
    if (!d_b || !d_x || nrhs < 1) return report(GFDFD_ERR_ARG, "gfdfd_solve_refine_batch: missing argument");
    if (d_b == d_x) return report(GFDFD_ERR_ARG, "gfdfd_solve_refine_batch: d_x must not alias d_b");
    if (sys < 0 || sys > 2) return report(GFDFD_ERR_ARG, "gfdfd_solve_refine_batch: invalid system %d", sys);
    GUARD_BEGIN
    std::vector<int> fr(freqs, freqs + nsel);
    auto B = reinterpret_cast<const cx<double> *>(d_b);
    auto X = reinterpret_cast<cx<double> *>(d_x);
    std::vector<RefineStats> rs = f->f64 ? solve_refined_batch<double>(*f->f64, fr, sys, nrhs, B, X, tol, max_iter, S(stream))
                                         : solve_refined_batch<float>(*f->f32, fr, sys, nrhs, B, X, tol, max_iter, S(stream));
    for (int z = 0; z < nsel; z++) {
        if (iters) iters[z] = rs[z].iterations;
        if (relres) relres[z] = rs[z].max_rel_residual;
    }
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_residual(gfdfd_plan plan, const double *d_vals, int sys, int nrhs, const double *d_b, const double *d_x,
                   double *d_r, void *stream) {
    if (!plan || !d_vals || !d_b || !d_x || !d_r || nrhs < 1)
        return report(GFDFD_ERR_ARG, "gfdfd_residual: missing argument");
    if (sys < 0 || sys > 2) return report(GFDFD_ERR_ARG, "gfdfd_residual: invalid system %d", sys);
    GUARD_BEGIN
    residual_fp64(plan->device(), reinterpret_cast<const cx<double> *>(d_vals), sys, nrhs,
                  reinterpret_cast<const cx<double> *>(d_b), reinterpret_cast<const cx<double> *>(d_x),
                  reinterpret_cast<cx<double> *>(d_r), S(stream));
    GFDFD_HIP(hipGetLastError());
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_spike_panel(int n, int nshots, const int *d_k, double amp_re, double amp_im, double *d_b, void *stream) {
    if (n < 1 || nshots < 1 || !d_k || !d_b) return report(GFDFD_ERR_ARG, "gfdfd_spike_panel: bad argument");
    GUARD_BEGIN
    spike_panel(n, nshots, d_k, cx<double>{amp_re, amp_im}, reinterpret_cast<cx<double> *>(d_b), S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_sample(int n, int nshots, int ntr, const int *d_krec, const double *d_x, float *d_pcal, void *stream) {
    if (n < 1 || nshots < 1 || ntr < 1 || !d_krec || !d_x || !d_pcal)
        return report(GFDFD_ERR_ARG, "gfdfd_sample: bad argument");
    GUARD_BEGIN
    sample_receivers(n, nshots, ntr, d_krec, reinterpret_cast<const cx<double> *>(d_x),
                     reinterpret_cast<cx<float> *>(d_pcal), S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_adjoint_sources(int n, int nshots, int ntr, const int *d_krec, const float *d_pcal, const float *d_pobs,
                          float dh, double *d_b, double *d_l2, void *stream) {
    if (n < 1 || nshots < 1 || ntr < 1 || !d_krec || !d_pcal || !d_pobs || !d_b || !d_l2)
        return report(GFDFD_ERR_ARG, "gfdfd_adjoint_sources: bad argument");
    GUARD_BEGIN
    adjoint_sources(n, nshots, ntr, d_krec, reinterpret_cast<const cx<float> *>(d_pcal),
                    reinterpret_cast<const cx<float> *>(d_pobs), dh, 1, 0, 0, reinterpret_cast<cx<double> *>(d_b),
                    d_l2, nullptr, S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_residual_sources(int n, int nshots, int ntr, const int *d_krec, const float *d_pcal, const float *d_pobs,
                           float dh, int misfit, int stf_inv, int rtm, double *d_b, double *d_l2, float *d_stf,
                           void *stream) {
    if (n < 1 || nshots < 1 || ntr < 1 || !d_krec || !d_pcal || !d_pobs || !d_b || !d_l2)
        return report(GFDFD_ERR_ARG, "gfdfd_residual_sources: bad argument");
    if (!rtm && (misfit < 1 || misfit > 3))
        return report(GFDFD_ERR_ARG, "gfdfd_residual_sources: misfit must be 1, 2 or 3");
    if (stf_inv != 0 && stf_inv != 1) return report(GFDFD_ERR_ARG, "gfdfd_residual_sources: stf_inv must be 0 or 1");
    GUARD_BEGIN
    adjoint_sources(n, nshots, ntr, d_krec, reinterpret_cast<const cx<float> *>(d_pcal),
                    reinterpret_cast<const cx<float> *>(d_pobs), dh, misfit, stf_inv, rtm ? 1 : 0,
                    reinterpret_cast<cx<double> *>(d_b), d_l2, reinterpret_cast<cx<float> *>(d_stf), S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_grad_ac(int n, int nshots, const float *omega2, const double *d_u, const double *d_lam, const float *d_vp,
                  const float *d_rho, double *d_grad, void *stream) {
    if (n < 1 || nshots < 1 || !omega2 || !d_u || !d_lam || !d_vp || !d_rho || !d_grad)
        return report(GFDFD_ERR_ARG, "gfdfd_grad_ac: bad argument");
    GUARD_BEGIN
    grad_ac(n, nshots, omega2, reinterpret_cast<const cx<double> *>(d_u), reinterpret_cast<const cx<double> *>(d_lam),
            d_vp, d_rho, nullptr, ShotTaperDev{}, d_grad, S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_grad_ac_stf(int n, int nshots, const float *omega2, const double *d_u, const double *d_lam,
                      const float *d_vp, const float *d_rho, const float *d_stf, double *d_grad, void *stream) {
    if (n < 1 || nshots < 1 || !omega2 || !d_u || !d_lam || !d_vp || !d_rho || !d_grad)
        return report(GFDFD_ERR_ARG, "gfdfd_grad_ac_stf: bad argument");
    GUARD_BEGIN
    grad_ac(n, nshots, omega2, reinterpret_cast<const cx<double> *>(d_u), reinterpret_cast<const cx<double> *>(d_lam),
            d_vp, d_rho, reinterpret_cast<const cx<float> *>(d_stf), ShotTaperDev{}, d_grad, S(stream));
    return GFDFD_OK;
    GUARD_END
}

namespace {
// the device form of a gfdfd_shot_taper, with the taper's float exponent
bool taper_dev(const gfdfd_shot_taper *t, int nshots, ShotTaperDev &d) {
    d = ShotTaperDev{};
    if (!t || t->mode == 0) return true;
    if (t->mode != 1 && t->mode != 2) return false;
    if (t->nx < 1 || !(t->dh > 0.0f) || t->nrec < 0 || (t->nrec > 0 && (!t->d_rec_i || !t->d_rec_j))) return false;
    if (t->mode == 1 && (!(t->srtradius > 0.0f) || !t->d_srcx || !t->d_srcy)) return false;
    if (t->mode == 2 && (t->filtsize < 0 || t->nsrc < 0 || (t->nsrc > 0 && (!t->d_src_i || !t->d_src_j))))
        return false;
    (void)nshots;
    d.mode = t->mode;
    d.nx = t->nx;
    d.dh = t->dh;
    d.radius = t->srtradius;
    d.a = gfdfd_shot_taper_exponent();
    d.f0 = 1e-3f;
    d.filtsize = t->filtsize;
    d.srcx = t->d_srcx;
    d.srcy = t->d_srcy;
    d.nsrc = t->nsrc;
    d.src_i = t->d_src_i;
    d.src_j = t->d_src_j;
    d.nrec = t->nrec;
    d.rec_i = t->d_rec_i;
    d.rec_j = t->d_rec_j;
    return true;
}
}  // namespace

int gfdfd_grad_ac_taper(int n, int nshots, const float *omega2, const double *d_u, const double *d_lam,
                        const float *d_vp, const float *d_rho, const float *d_stf, const gfdfd_shot_taper *taper,
                        double *d_grad, void *stream) {
    ShotTaperDev tp;
    if (n < 1 || nshots < 1 || !omega2 || !d_u || !d_lam || !d_vp || !d_rho || !d_grad || !taper_dev(taper, nshots, tp))
        return report(GFDFD_ERR_ARG, "gfdfd_grad_ac_taper: bad argument");
    if (tp.mode && (int64_t)tp.nx * ((n + tp.nx - 1) / tp.nx) != n)
        return report(GFDFD_ERR_ARG, "gfdfd_grad_ac_taper: n = %d is not a multiple of the grid width %d", n, tp.nx);
    GUARD_BEGIN
    grad_ac(n, nshots, omega2, reinterpret_cast<const cx<double> *>(d_u), reinterpret_cast<const cx<double> *>(d_lam),
            d_vp, d_rho, reinterpret_cast<const cx<float> *>(d_stf), tp, d_grad, S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_hess_shin_ac(int n, int nshots, const float *omega2, const double *d_u, const float *d_vp, const float *d_rho,
                       double *d_hess, void *stream) {
    if (n < 1 || nshots < 1 || !omega2 || !d_u || !d_vp || !d_rho || !d_hess)
        return report(GFDFD_ERR_ARG, "gfdfd_hess_shin_ac: bad argument");
    GUARD_BEGIN
    hess_shin_ac(n, nshots, omega2, reinterpret_cast<const cx<double> *>(d_u), d_vp, d_rho, d_hess, S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_apply_hess(int n, float eps_hess, const double *d_hess, double *d_grad, void *stream) {
    if (n < 1 || !d_hess || !d_grad) return report(GFDFD_ERR_ARG, "gfdfd_apply_hess: bad argument");
    GUARD_BEGIN
    apply_hess(n, eps_hess, d_hess, d_grad, S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_model_update(int n, float alpha, const double *d_dir, const float *d_m_old, float lo, float hi,
                       float *d_m_new, void *stream) {
    if (n < 1 || !d_dir || !d_m_old || !d_m_new) return report(GFDFD_ERR_ARG, "gfdfd_model_update: bad argument");
    GUARD_BEGIN
    model_update(n, alpha, d_dir, d_m_old, lo, hi, d_m_new, S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_model_update_norm(int n, float alpha, const double *d_dir, const float *d_m_old, float norm, float lo,
                            float hi, float *d_m_new, void *stream) {
    if (n < 1 || !(norm != 0.0f) || !d_dir || !d_m_old || !d_m_new)
        return report(GFDFD_ERR_ARG, "gfdfd_model_update_norm: bad argument");
    GUARD_BEGIN
    model_update_norm(n, alpha, d_dir, d_m_old, norm, lo, hi, d_m_new, S(stream));
    return GFDFD_OK;
    GUARD_END
}

// gauss_filt's window (src/gauss_filt.c:25-63), evaluated with the reference's
// float and integer arithmetic: the wavelength VREF / FC_low in float, the
// window width rounded half away from zero and made odd, the half width hfs,
// sigma = hfs / 2 in integers and s = 2 sigma^2.
static void gauss_window(float size, float vref, float fc_low, float dh, int *hfs, float *s) {
    const float lam = vref / fc_low;
    int w = (int)std::round((double)((size * lam) / dh));
    if (w == 0) { *hfs = -1; *s = 0.0f; return; }
    if (!(w % 2)) w += w > 0 ? 1 : -1;
    *hfs = std::abs(w) / 2;
    const float sigma = (float)(*hfs / 2);
    *s = (float)(2.0 * sigma * sigma);
}

int gfdfd_precond(int nx, int ny, float dh, float vref, float fc_low, float filt_size_x, float filt_size_y,
                  int spatfilter, int nzero_rows, const double *d_in, double *d_out, void *stream) {
    if (nx < 1 || ny < 1 || !d_in || !d_out || d_in == d_out || nzero_rows < 0)
        return report(GFDFD_ERR_ARG, "gfdfd_precond: bad argument");
    int hfsx = -1, hfsy = -1;
    float sx = 0.0f, sy = 0.0f;
    if (spatfilter) {
        gauss_window(filt_size_x, vref, fc_low, dh, &hfsx, &sx);
        gauss_window(filt_size_y, vref, fc_low, dh, &hfsy, &sy);
        if (hfsx < 0 || hfsy < 0) hfsx = hfsy = -1;        // a zero-width window returns unfiltered
    }
    GUARD_BEGIN
    precond(nx, ny, hfsx, hfsy, sx, sy, nzero_rows, d_in, d_out, S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_tikhonov_te(int nx, int ny, float dh, float beta_sigma, float beta_eps, float norm_sigma, float norm_eps,
                      float lambda_sigma, float lambda_eps, const float *d_sigma, const float *d_eps,
                      double *d_grad_sigma, double *d_grad_eps, double *d_cost, void *stream) {
    if (nx < 3 || ny < 3 || !d_sigma || !d_eps || !d_cost || !(norm_sigma != 0.0f) || !(norm_eps != 0.0f))
        return report(GFDFD_ERR_ARG, "gfdfd_tikhonov_te: bad argument");
    GUARD_BEGIN
    tikhonov_te(nx, ny, dh, beta_sigma / norm_sigma, beta_eps / norm_eps, beta_sigma, beta_eps,
                lambda_sigma * beta_sigma, lambda_eps * beta_eps, d_sigma, d_eps, d_grad_sigma, d_grad_eps, d_cost,
                S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_taper_file(int nx0, int ny0, int npml, int free_surf, const float *d_taper, const double *d_in,
                     double *d_out, void *stream) {
    if (nx0 < 1 || ny0 < 1 || npml < 0 || !d_taper || !d_in || !d_out || d_in == d_out)
        return report(GFDFD_ERR_ARG, "gfdfd_taper_file: bad argument");
    const int fsshift = free_surf ? 0 : npml;
    if (free_surf && npml >= fsshift + ny0)
        return report(GFDFD_ERR_ARG, "gfdfd_taper_file: the model has fewer rows than the PML");
    GUARD_BEGIN
    taper_file(nx0 + 2 * npml, ny0 + npml + fsshift, nx0, ny0, npml, fsshift, free_surf ? 1 : 0, d_taper, d_in, d_out,
               S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_grad_frame(int nx0, int ny0, int npml, int free_surf, double *d_grad, void *stream) {
    if (nx0 < 1 || ny0 < 1 || npml < 2 || !d_grad) return report(GFDFD_ERR_ARG, "gfdfd_grad_frame: bad argument");
    const int fsshift = free_surf ? 0 : npml;
    // src/cp_grad_frame.c:112-116: DAMPING = 60 %, amp = 1 - 0.6 in float, a from the PML width
    const float amp = 1.0 - 60.0 / 100.0;
    const float a = std::sqrt(-std::log((double)amp) / ((npml - 1) * (npml - 1)));
    GUARD_BEGIN
    grad_frame(nx0 + 2 * npml, ny0 + npml + fsshift, nx0, ny0, npml, fsshift, free_surf ? 1 : 0, a, d_grad, S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_hess_gn_ac(int n, int nshots, int nrec, const float *omega2, const double *d_u, const double *d_g,
                     const float *d_vp, const float *d_rho, double *d_hess, void *stream) {
    if (n < 1 || nshots < 1 || nrec < 1 || !omega2 || !d_u || !d_g || !d_vp || !d_rho || !d_hess)
        return report(GFDFD_ERR_ARG, "gfdfd_hess_gn_ac: bad argument");
    GUARD_BEGIN
    hess_gn_ac(n, nshots, nrec, omega2, reinterpret_cast<const cx<double> *>(d_u),
               reinterpret_cast<const cx<double> *>(d_g), d_vp, d_rho, d_hess, S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_hess_gn_te(int n, int nshots, int nrec, const float *omega, const float *omega2, const double *d_u,
                     const double *d_g, double *d_hess_sigma, double *d_hess_eps, void *stream) {
    if (n < 1 || nshots < 1 || nrec < 1 || !omega || !omega2 || !d_u || !d_g || !d_hess_sigma || !d_hess_eps)
        return report(GFDFD_ERR_ARG, "gfdfd_hess_gn_te: bad argument");
    GUARD_BEGIN
    hess_gn_te(n, nshots, nrec, omega, omega2, reinterpret_cast<const cx<double> *>(d_u),
               reinterpret_cast<const cx<double> *>(d_g), d_hess_sigma, d_hess_eps, S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_hess_shin_te(int n, int nshots, const float *omega, const float *omega2, const double *d_u,
                       double *d_hess_sigma, double *d_hess_eps, void *stream) {
    if (n < 1 || nshots < 1 || !omega || !omega2 || !d_u || !d_hess_sigma || !d_hess_eps)
        return report(GFDFD_ERR_ARG, "gfdfd_hess_shin_te: bad argument");
    GUARD_BEGIN
    hess_shin_te(n, nshots, omega, omega2, reinterpret_cast<const cx<double> *>(d_u), d_hess_sigma, d_hess_eps,
                 S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_grad_te(int n, int nshots, const float *omega, const float *omega2, const double *d_u, const double *d_lam,
                  double *d_grad_sigma, double *d_grad_eps, void *stream) {
    if (n < 1 || nshots < 1 || !omega || !omega2 || !d_u || !d_lam || !d_grad_sigma || !d_grad_eps)
        return report(GFDFD_ERR_ARG, "gfdfd_grad_te: bad argument");
    GUARD_BEGIN
    grad_te(n, nshots, omega, omega2, reinterpret_cast<const cx<double> *>(d_u),
            reinterpret_cast<const cx<double> *>(d_lam), nullptr, ShotTaperDev{}, d_grad_sigma, d_grad_eps, S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_grad_te_stf(int n, int nshots, const float *omega, const float *omega2, const double *d_u,
                      const double *d_lam, const float *d_stf, double *d_grad_sigma, double *d_grad_eps,
                      void *stream) {
    if (n < 1 || nshots < 1 || !omega || !omega2 || !d_u || !d_lam || !d_grad_sigma || !d_grad_eps)
        return report(GFDFD_ERR_ARG, "gfdfd_grad_te_stf: bad argument");
    GUARD_BEGIN
    grad_te(n, nshots, omega, omega2, reinterpret_cast<const cx<double> *>(d_u),
            reinterpret_cast<const cx<double> *>(d_lam), reinterpret_cast<const cx<float> *>(d_stf), ShotTaperDev{},
            d_grad_sigma, d_grad_eps, S(stream));
    return GFDFD_OK;
    GUARD_END
}

int gfdfd_grad_te_taper(int n, int nshots, const float *omega, const float *omega2, const double *d_u,
                        const double *d_lam, const float *d_stf, const gfdfd_shot_taper *taper, double *d_grad_sigma,
                        double *d_grad_eps, void *stream) {
    ShotTaperDev tp;
    if (n < 1 || nshots < 1 || !omega || !omega2 || !d_u || !d_lam || !d_grad_sigma || !d_grad_eps ||
        !taper_dev(taper, nshots, tp))
        return report(GFDFD_ERR_ARG, "gfdfd_grad_te_taper: bad argument");
    if (tp.mode && (int64_t)tp.nx * ((n + tp.nx - 1) / tp.nx) != n)
        return report(GFDFD_ERR_ARG, "gfdfd_grad_te_taper: n = %d is not a multiple of the grid width %d", n, tp.nx);
    GUARD_BEGIN
    grad_te(n, nshots, omega, omega2, reinterpret_cast<const cx<double> *>(d_u),
            reinterpret_cast<const cx<double> *>(d_lam), reinterpret_cast<const cx<float> *>(d_stf), tp, d_grad_sigma,
            d_grad_eps, S(stream));
    return GFDFD_OK;
    GUARD_END
}

}  // extern "C"
