// The six UMFPACK entry points GERMAINE links against (include/umfpack.h),
// backed by the GPU multifrontal solver.
//
// Call sequence the host drives (src/AC/forward_shot_AC.c:41-120):
// triplet_to_col -> (value overwrite) -> symbolic -> numeric -> free_symbolic
// -> solve per shot -> free_numeric, one frequency at a time.  The host ignores
// every status (src/AC/forward_shot_AC.c:41,58,61,88), so each non-zero status
// is also printed on stderr; GFDFD_ABORT_ON_ERROR=1 turns errors into abort().
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "../../include/umfpack.h"
#include "engine.hpp"
#include "runtime.hpp"

using namespace gfdfd;

namespace {

constexpr unsigned kSymMagic = 0x53594d46u, kNumMagic = 0x4e554d46u;
constexpr int kDenseMax = 4096;

struct SymbolicObj {
    unsigned magic = kSymMagic;
    std::shared_ptr<const Plan> plan;   // host analysis; uploaded by umfpack_zi_numeric
    int n = 0;
};

struct NumericObj {
    unsigned magic = kNumMagic;
    std::unique_ptr<Factors<double>> F;
    int n = 0;
    bool singular = false;
    cx<double> *b = nullptr, *x = nullptr, *r = nullptr, *d = nullptr, *x0 = nullptr;
    std::vector<cx<double>> hb;
    ~NumericObj() {
        dev_free(b); dev_free(x); dev_free(r); dev_free(d); dev_free(x0);
    }
};

int fail(int code, const char *fn, const char *msg) {
    report(code, "%s: %s (status %d)", fn, msg, code);
    if (const char *s = std::getenv("GFDFD_ABORT_ON_ERROR"))
        if (s[0] == '1') abort();
    return code;
}

// a failed device allocation is UMFPACK's out-of-memory; any other device
// failure (no GPU, a launch or copy error) is an internal error
int hip_status(const hip_error &e) {
    return e.code == hipErrorOutOfMemory ? UMFPACK_ERROR_out_of_memory : UMFPACK_ERROR_internal_error;
}

void set_info(double Info[], int status) {
    if (Info) Info[UMFPACK_STATUS] = status;
}

bool valid_csc(int n, const int Ap[], const int Ai[]) {
    if (Ap[0] != 0) return false;
    for (int c = 0; c < n; c++) {
        if (Ap[c + 1] < Ap[c]) return false;
        for (int q = Ap[c]; q < Ap[c + 1]; q++) {
            if (Ai[q] < 0 || Ai[q] >= n) return false;
            if (q > Ap[c] && Ai[q] <= Ai[q - 1]) return false;
        }
    }
    return true;
}

// complex value k of a split (x, z) or packed (z == NULL) array
inline cx<double> getc(const double *x, const double *z, int64_t k) {
    return z ? cx<double>{x[k], z[k]} : cx<double>{x[2 * k], x[2 * k + 1]};
}

}  // namespace

extern "C" {

int umfpack_zi_triplet_to_col(int n_row, int n_col, int nz, const int Ti[], const int Tj[], const double Tx[],
                              const double Tz[], int Ap[], int Ai[], double Ax[], double Az[], int Map[]) {
    static const char *fn = "umfpack_zi_triplet_to_col";
    if (!Ti || !Tj || !Ap || !Ai) return fail(UMFPACK_ERROR_argument_missing, fn, "argument missing");
    if (n_row <= 0 || n_col <= 0) return fail(UMFPACK_ERROR_n_nonpositive, fn, "n_row or n_col <= 0");
    if (nz < 0) return fail(UMFPACK_ERROR_invalid_matrix, fn, "nz < 0");
    try {
        for (int k = 0; k < nz; k++)
            if (Ti[k] < 0 || Ti[k] >= n_row || Tj[k] < 0 || Tj[k] >= n_col)
                return fail(UMFPACK_ERROR_invalid_matrix, fn, "triplet index out of range");
        const bool values = Tx && Ax;
        // order the triplets by (column, row); stable, so duplicates keep input order
        std::vector<int> order(nz);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return Tj[a] != Tj[b] ? Tj[a] < Tj[b] : Ti[a] < Ti[b];
        });
        std::vector<int> colcount(n_col, 0);
        std::vector<cx<double>> v;
        if (values) v.reserve(nz);
        int out = -1, lastc = -1, lastr = -1;
        std::vector<int> ai;
        ai.reserve(nz);
        for (int t = 0; t < nz; t++) {
            int k = order[t];
            if (Tj[k] != lastc || Ti[k] != lastr) {
                out++;
                ai.push_back(Ti[k]);
                colcount[Tj[k]]++;
                if (values) v.push_back(getc(Tx, Tz, k));
                lastc = Tj[k];
                lastr = Ti[k];
            } else if (values) {
                cx<double> a = getc(Tx, Tz, k);
                v.back().x += a.x;
                v.back().y += a.y;
            }
            if (Map) Map[k] = out;
        }
        Ap[0] = 0;
        for (int c = 0; c < n_col; c++) Ap[c + 1] = Ap[c] + colcount[c];
        std::memcpy(Ai, ai.data(), sizeof(int) * ai.size());
        if (values)
            for (size_t q = 0; q < v.size(); q++) {
                if (Az) { Ax[q] = v[q].x; Az[q] = v[q].y; }
                else { Ax[2 * q] = v[q].x; Ax[2 * q + 1] = v[q].y; }
            }
        return UMFPACK_OK;
    } catch (const std::bad_alloc &) {
        return fail(UMFPACK_ERROR_out_of_memory, fn, "out of host memory");
    }
}

int umfpack_zi_symbolic(int n_row, int n_col, const int Ap[], const int Ai[], const double Ax[], const double Az[],
                        void **Symbolic, const double Control[], double Info[]) {
    static const char *fn = "umfpack_zi_symbolic";
    (void)Ax; (void)Az; (void)Control;
    if (Info) {
        Info[UMFPACK_NROW] = n_row;
        Info[UMFPACK_NCOL] = n_col;
    }
    if (!Symbolic || !Ap || !Ai) { set_info(Info, UMFPACK_ERROR_argument_missing); return fail(UMFPACK_ERROR_argument_missing, fn, "argument missing"); }
    *Symbolic = nullptr;
    if (n_row <= 0 || n_col <= 0) { set_info(Info, UMFPACK_ERROR_n_nonpositive); return fail(UMFPACK_ERROR_n_nonpositive, fn, "n <= 0"); }
    if (n_row != n_col) { set_info(Info, UMFPACK_ERROR_invalid_matrix); return fail(UMFPACK_ERROR_invalid_matrix, fn, "only square matrices are supported"); }
    const int n = n_row;
    if (!valid_csc(n, Ap, Ai)) { set_info(Info, UMFPACK_ERROR_invalid_matrix); return fail(UMFPACK_ERROR_invalid_matrix, fn, "invalid CSC (unsorted, duplicate or out-of-range row indices)"); }
    if (Info) Info[UMFPACK_NZ] = Ap[n];
    try {
        auto S = std::make_unique<SymbolicObj>();
        S->n = n;
        int ny = 0;
        int nx = detect_grid(n, Ap, Ai, &ny);
        if (nx > 0) {
            S->plan = cached_host_plan(nx, ny, PlanOptions());
        } else if (n <= kDenseMax) {
            S->plan = std::make_shared<const Plan>(build_dense_plan(n, Ap, Ai, PlanOptions()));
        } else {
            set_info(Info, UMFPACK_ERROR_invalid_matrix);
            return fail(UMFPACK_ERROR_invalid_matrix, fn, "pattern is not a 9-point grid and too large for one dense front");
        }
        *Symbolic = S.release();
        set_info(Info, UMFPACK_OK);
        return UMFPACK_OK;
    } catch (const hip_error &e) {
        const int st = hip_status(e);
        set_info(Info, st);
        return fail(st, fn, e.what());
    } catch (const std::exception &e) {
        set_info(Info, UMFPACK_ERROR_internal_error);
        return fail(UMFPACK_ERROR_internal_error, fn, e.what());
    }
}

int umfpack_zi_numeric(const int Ap[], const int Ai[], const double Ax[], const double Az[], void *Symbolic,
                       void **Numeric, const double Control[], double Info[]) {
    static const char *fn = "umfpack_zi_numeric";
    (void)Control;
    if (!Numeric || !Ap || !Ai || !Ax) { set_info(Info, UMFPACK_ERROR_argument_missing); return fail(UMFPACK_ERROR_argument_missing, fn, "argument missing"); }
    *Numeric = nullptr;
    auto *S = static_cast<SymbolicObj *>(Symbolic);
    if (!S || S->magic != kSymMagic) { set_info(Info, UMFPACK_ERROR_invalid_Symbolic_object); return fail(UMFPACK_ERROR_invalid_Symbolic_object, fn, "invalid Symbolic object"); }
    const Plan &pl = *S->plan;
    const int n = S->n;
    // the pattern must be the one analysed
    if (Ap[0] != 0 || Ap[n] != pl.nnz || std::memcmp(Ap, pl.colptr.data(), sizeof(int) * (n + 1)) != 0 ||
        std::memcmp(Ai, pl.rowind.data(), sizeof(int) * pl.nnz) != 0) {
        set_info(Info, UMFPACK_ERROR_different_pattern);
        return fail(UMFPACK_ERROR_different_pattern, fn, "pattern differs from the one given to umfpack_zi_symbolic");
    }
    try {
        auto N = std::make_unique<NumericObj>();
        N->n = n;
        std::vector<cx<double>> hv(pl.nnz);
        for (int64_t q = 0; q < pl.nnz; q++) hv[q] = getc(Ax, Az, q);
        std::shared_ptr<DevPlan> dp = device_plan(S->plan);
        cx<double> *dv = dev_upload(hv);
        hipStream_t st = nullptr;
        N->F = make_factors<double>(dp, 1, dv, false, st);
        N->F->own_vals = true;
        int sing = 0;
        GFDFD_HIP(hipMemcpy(&sing, N->F->sing, sizeof(int), hipMemcpyDeviceToHost));
        N->singular = sing != 0;
        *Numeric = N.release();
        if (sing) {
            set_info(Info, UMFPACK_WARNING_singular_matrix);
            return fail(UMFPACK_WARNING_singular_matrix, fn, "matrix is singular (zero pivot)");
        }
        set_info(Info, UMFPACK_OK);
        return UMFPACK_OK;
    } catch (const hip_error &e) {
        const int st = hip_status(e);
        set_info(Info, st);
        return fail(st, fn, e.what());
    } catch (const std::exception &e) {
        set_info(Info, UMFPACK_ERROR_internal_error);
        return fail(UMFPACK_ERROR_internal_error, fn, e.what());
    }
}

int umfpack_zi_solve(int sys, const int Ap[], const int Ai[], const double Ax[], const double Az[], double Xx[],
                     double Xz[], const double Bx[], const double Bz[], void *Numeric, const double Control[],
                     double Info[]) {
    static const char *fn = "umfpack_zi_solve";
    auto *N = static_cast<NumericObj *>(Numeric);
    if (!N || N->magic != kNumMagic) { set_info(Info, UMFPACK_ERROR_invalid_Numeric_object); return fail(UMFPACK_ERROR_invalid_Numeric_object, fn, "invalid Numeric object"); }
    if (!Xx || !Bx) { set_info(Info, UMFPACK_ERROR_argument_missing); return fail(UMFPACK_ERROR_argument_missing, fn, "argument missing"); }
    if (sys != UMFPACK_A && sys != UMFPACK_At && sys != UMFPACK_Aat) { set_info(Info, UMFPACK_ERROR_invalid_system); return fail(UMFPACK_ERROR_invalid_system, fn, "invalid system"); }
    int irstep = UMFPACK_DEFAULT_IRSTEP;
    if (Control && !std::isnan(Control[UMFPACK_IRSTEP])) irstep = std::max(0, (int)Control[UMFPACK_IRSTEP]);
    // refinement needs the matrix (UMFPACK skips it when Ap/Ai/Ax are NULL) and a grid plan
    if (!Ap || !Ai || !Ax || !N->F->plan->tq) irstep = 0;
    try {
        const int n = N->n;
        const DevPlan &dp = *N->F->plan;
        if (!N->b) {
            N->b = dev_alloc<cx<double>>(n);
            N->x = dev_alloc<cx<double>>(n);
            N->r = dev_alloc<cx<double>>(n);
            N->d = dev_alloc<cx<double>>(n);
            N->x0 = dev_alloc<cx<double>>(n);
            N->hb.resize(n);
        }
        hipStream_t st = nullptr;
        for (int i = 0; i < n; i++) N->hb[i] = getc(Bx, Bz, i);
        GFDFD_HIP(hipMemcpyAsync(N->b, N->hb.data(), sizeof(cx<double>) * n, hipMemcpyHostToDevice, st));
        solve_panel<double, double>(*N->F, 0, sys, 1, N->b, N->x, st);
        // iterative refinement as UMFPACK's default: at most irstep corrections,
        // stop at omega <= eps or when a step fails to halve omega (keep the better x)
        int taken = 0;
        double omega = -1;
        if (irstep > 0) {
            const double eps = std::ldexp(1.0, -52);
            omega = backward_error(dp, N->F->vals, sys, N->b, N->x, N->r, st);
            while (omega > eps && taken < irstep) {
                GFDFD_HIP(hipMemcpyAsync(N->x0, N->x, sizeof(cx<double>) * n, hipMemcpyDeviceToDevice, st));
                solve_panel<double, double>(*N->F, 0, sys, 1, N->r, N->d, st);
                axpy_fp64(n, N->d, N->x, st);
                taken++;
                double om2 = backward_error(dp, N->F->vals, sys, N->b, N->x, N->r, st);
                if (!(om2 < omega)) {   // no improvement: restore the previous iterate
                    GFDFD_HIP(hipMemcpyAsync(N->x, N->x0, sizeof(cx<double>) * n, hipMemcpyDeviceToDevice, st));
                    break;
                }
                bool halved = om2 <= 0.5 * omega;
                omega = om2;
                if (!halved) break;
            }
        }
        GFDFD_HIP(hipMemcpyAsync(N->hb.data(), N->x, sizeof(cx<double>) * n, hipMemcpyDeviceToHost, st));
        GFDFD_HIP(hipStreamSynchronize(st));
        for (int i = 0; i < n; i++) {
            if (Xz) { Xx[i] = N->hb[i].x; Xz[i] = N->hb[i].y; }
            else { Xx[2 * i] = N->hb[i].x; Xx[2 * i + 1] = N->hb[i].y; }
        }
        if (Info) {
            Info[UMFPACK_IR_TAKEN] = taken;
            Info[UMFPACK_OMEGA1] = omega;
        }
        if (N->singular) {
            set_info(Info, UMFPACK_WARNING_singular_matrix);
            return fail(UMFPACK_WARNING_singular_matrix, fn, "matrix is singular");
        }
        set_info(Info, UMFPACK_OK);
        return UMFPACK_OK;
    } catch (const std::exception &e) {
        set_info(Info, UMFPACK_ERROR_internal_error);
        return fail(UMFPACK_ERROR_internal_error, fn, e.what());
    }
}

void umfpack_zi_free_symbolic(void **Symbolic) {
    if (!Symbolic || !*Symbolic) return;
    auto *S = static_cast<SymbolicObj *>(*Symbolic);
    if (S->magic == kSymMagic) {
        S->magic = 0;
        delete S;
    }
    *Symbolic = nullptr;
}

void umfpack_zi_free_numeric(void **Numeric) {
    if (!Numeric || !*Numeric) return;
    auto *N = static_cast<NumericObj *>(*Numeric);
    if (N->magic == kNumMagic) {
        N->magic = 0;
        delete N;
    }
    *Numeric = nullptr;
}

}  // extern "C"
