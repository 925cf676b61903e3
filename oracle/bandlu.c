/*
 * ORACLE — test infrastructure only (never linked into germaine_amd/).
 *
 * FP64 complex direct solver used as the checker for the factor/solve path.
 * The reference factors with UMFPACK (SuiteSparse, version not pinned by the
 * reference: include/fd.h:18, src/Makefile:32-35), which is absent here.
 * UMFPACK's published contract for the calls GERMAINE makes is
 *   umfpack_zi_numeric  : LU factorisation with partial (threshold) pivoting,
 *   umfpack_zi_solve(sys): x = M^-1 b (UMFPACK_A), M^-H b (UMFPACK_At),
 *                          M^-T b (UMFPACK_Aat),
 * where M is the CSC matrix handed in (src/AC/forward_shot_AC.c:58,61,88;
 * src/TE/grad_obj_TE.c:149).  Any backward-stable FP64 LU with partial
 * pivoting satisfies the same contract to rounding error, so the oracle
 * restates LAPACK's banded LU with partial pivoting (xGBTF2 / xGBTRS) on the
 * CSC matrix after an optional symmetric bandwidth-reducing permutation.
 * The factor/solve path shares no code or ordering with the product's
 * multifrontal solver.
 *
 * Pinning (tests/test_oracle_solver.py): numpy dense LAPACK solves on small
 * grids, SciPy SuperLU on a 201x101 grid, and the oracle-free backward error
 * ||b - M x|| / (||M|| ||x|| + ||b||).
 */
#include <complex.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

typedef double complex cd;

static inline double cabs1(cd z) { return fabs(creal(z)) + fabs(cimag(z)); }

/*
 * Solve op(M) X = B for nrhs right-hand sides (column j of B at B + j*n).
 * sys: 0 -> M, 1 -> M^H, 2 -> M^T (UMFPACK_A / _At / _Aat numbering).
 * perm: optional, new index -> old index (symmetric permutation applied
 * before banding).  Returns 0, 1 if a zero pivot was met (singular), or
 * -1 on allocation failure.
 */
int orc_band_solve(int n, const int *Ap, const int *Ai, const double *Ax, const double *Az,
                   const int *perm, int sys, int nrhs, const double *Bx, const double *Bz,
                   double *Xx, double *Xz)
{
    int *ip = malloc(sizeof(int) * n);
    if (!ip) return -1;
    for (int i = 0; i < n; i++) ip[perm ? perm[i] : i] = i;

    int kl = 0, ku = 0;
    for (int c = 0; c < n; c++)
        for (int q = Ap[c]; q < Ap[c + 1]; q++) {
            int i = ip[Ai[q]], j = ip[c];
            if (i - j > kl) kl = i - j;
            if (j - i > ku) ku = j - i;
        }
    int kv = ku + kl, ldab = 2 * kl + ku + 1;
    cd *ab = calloc((size_t)ldab * n, sizeof(cd));
    int *ipiv = malloc(sizeof(int) * n);
    cd *w = malloc(sizeof(cd) * n);
    if (!ab || !ipiv || !w) { free(ip); free(ab); free(ipiv); free(w); return -1; }
#define AB(i, j) ab[(size_t)(kv + (i) - (j)) + (size_t)(j) * ldab]
    for (int c = 0; c < n; c++)
        for (int q = Ap[c]; q < Ap[c + 1]; q++)
            AB(ip[Ai[q]], ip[c]) += Ax[q] + Az[q] * I;

    /* factor: column-by-column partial pivoting within the band */
    int info = 0, ju = 0;
    for (int j = 0; j < n; j++) {
        int km = kl < n - 1 - j ? kl : n - 1 - j;
        int p = 0;
        double best = cabs1(AB(j, j));
        for (int r = 1; r <= km; r++) {
            double v = cabs1(AB(j + r, j));
            if (v > best) { best = v; p = r; }
        }
        ipiv[j] = j + p;
        if (best == 0.0) { info = 1; continue; }
        int jhi = j + ku + p < n - 1 ? j + ku + p : n - 1;
        if (jhi > ju) ju = jhi;
        if (p != 0)
            for (int c = j; c <= ju; c++) { cd t = AB(j, c); AB(j, c) = AB(j + p, c); AB(j + p, c) = t; }
        cd rpiv = 1.0 / AB(j, j);
        for (int r = 1; r <= km; r++) AB(j + r, j) *= rpiv;
        for (int c = j + 1; c <= ju; c++) {
            cd u = AB(j, c);
            if (u != 0)
                for (int r = 1; r <= km; r++) AB(j + r, c) -= AB(j + r, j) * u;
        }
    }

    for (int rhs = 0; rhs < nrhs; rhs++) {
        const double *bx = Bx + (size_t)rhs * n, *bz = Bz + (size_t)rhs * n;
        for (int i = 0; i < n; i++) w[ip[i]] = bx[i] + bz[i] * I;
        if (sys == 0) {
            for (int j = 0; j < n - 1; j++) {           /* L, with the row swaps */
                int l = kl < n - 1 - j ? kl : n - 1 - j;
                if (ipiv[j] != j) { cd t = w[j]; w[j] = w[ipiv[j]]; w[ipiv[j]] = t; }
                for (int r = 1; r <= l; r++) w[j + r] -= AB(j + r, j) * w[j];
            }
            for (int j = n - 1; j >= 0; j--) {          /* U */
                w[j] /= AB(j, j);
                int i0 = j - kv > 0 ? j - kv : 0;
                for (int i = i0; i < j; i++) w[i] -= AB(i, j) * w[j];
            }
        } else {
            int cj = (sys == 1);
            for (int j = 0; j < n; j++) {               /* U^T or U^H */
                cd s = w[j];
                int i0 = j - kv > 0 ? j - kv : 0;
                for (int i = i0; i < j; i++) s -= (cj ? conj(AB(i, j)) : AB(i, j)) * w[i];
                w[j] = s / (cj ? conj(AB(j, j)) : AB(j, j));
            }
            for (int j = n - 2; j >= 0; j--) {          /* L^T or L^H, then swaps */
                int l = kl < n - 1 - j ? kl : n - 1 - j;
                cd s = w[j];
                for (int r = 1; r <= l; r++) s -= (cj ? conj(AB(j + r, j)) : AB(j + r, j)) * w[j + r];
                w[j] = s;
                if (ipiv[j] != j) { cd t = w[j]; w[j] = w[ipiv[j]]; w[ipiv[j]] = t; }
            }
        }
        double *xx = Xx + (size_t)rhs * n, *xz = Xz + (size_t)rhs * n;
        for (int i = 0; i < n; i++) { xx[i] = creal(w[ip[i]]); xz[i] = cimag(w[ip[i]]); }
    }
#undef AB
    free(ip); free(ab); free(ipiv); free(w);
    return info;
}

/* y = op(M) x for a CSC matrix (used by tests for residuals) */
void orc_csc_matvec(int n, const int *Ap, const int *Ai, const double *Ax, const double *Az,
                    int sys, const double *xr, const double *xi, double *yr, double *yi)
{
    for (int i = 0; i < n; i++) { yr[i] = 0; yi[i] = 0; }
    for (int c = 0; c < n; c++)
        for (int q = Ap[c]; q < Ap[c + 1]; q++) {
            int r = Ai[q];
            cd a = Ax[q] + Az[q] * I;
            if (sys == 1) a = conj(a);
            if (sys == 0) {
                cd v = a * (xr[c] + xi[c] * I);
                yr[r] += creal(v); yi[r] += cimag(v);
            } else {
                cd v = a * (xr[r] + xi[r] * I);
                yr[c] += creal(v); yi[c] += cimag(v);
            }
        }
}
