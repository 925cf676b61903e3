/*
 * ORACLE — test infrastructure only (never linked into germaine_amd/).
 *
 * FP64 complex multifrontal LU on the CPU with OpenMP: the CPU baseline of the
 * solve path and an independent FP64 checker at sizes the banded oracle
 * cannot reach.  It restates the algorithm class of UMFPACK's numeric
 * factorisation (multifrontal LU with partial pivoting inside each front,
 * dense BLAS-3 style front updates) for the matrices GERMAINE factors
 * (src/AC/forward_shot_AC.c:58-61), with its own geometric nested-dissection
 * ordering and a general symbolic analysis (front structure from the
 * children's contribution sets and the pivots' adjacency, not from grid
 * geometry).  Solves follow GERMAINE's per-shot pattern: one right-hand side
 * at a time (src/AC/forward_shot_AC.c:77-88), shots spread over threads the
 * way MPI ranks split them (src/AC/forward_AC.c:134-147).
 *
 * Pinning (tests/test_oracle_cpu_mf.py): the banded-LU oracle on small grids
 * and SciPy SuperLU on a 201x101 grid, M and M^T.
 */
#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cd;

typedef struct {
    int p, m, pos0, parent, level;
    int *rows;      /* m positions, ascending; first p are the pivots */
    cd *F;          /* m*m column-major after factorisation */
    int *piv;       /* p local row swaps */
    int nkids, kid0;
} mf_front;

typedef struct {
    int n, nx, ny, nf, nlev;
    int *perm, *iperm;
    mf_front *fr;
    int *kids;          /* children lists */
    int *lev_ptr, *lev_fr;
} mf_solver;

/* ---- ordering: recursive bisection of the grid, iterative with a stack ---- */
typedef struct { int i0, i1, j0, j1, parent_slot; } box_t;

static void nd_order(mf_solver *S, int leaf)
{
    int nx = S->nx, ny = S->ny, n = nx * ny;
    /* first pass: build the tree top-down, recording each node's separator */
    int cap = 4 * (n / (leaf > 0 ? leaf : 1) + 16);
    int *sep_i0 = malloc(sizeof(int) * cap), *sep_i1 = malloc(sizeof(int) * cap);
    int *sep_j0 = malloc(sizeof(int) * cap), *sep_j1 = malloc(sizeof(int) * cap);
    int *par = malloc(sizeof(int) * cap), *ch0 = malloc(sizeof(int) * cap), *ch1 = malloc(sizeof(int) * cap);
    box_t *stack = malloc(sizeof(box_t) * cap);
    int top = 0, nn = 0;
    stack[top++] = (box_t){0, nx, 0, ny, -1};
    while (top > 0) {
        box_t b = stack[--top];
        int w = b.i1 - b.i0, h = b.j1 - b.j0;
        if (w <= 0 || h <= 0) continue;
        int id = nn++;
        par[id] = b.parent_slot;
        ch0[id] = ch1[id] = -1;
        if (b.parent_slot >= 0) {
            if (ch0[b.parent_slot] < 0) ch0[b.parent_slot] = id;
            else ch1[b.parent_slot] = id;
        }
        if ((long)w * h <= leaf) {
            sep_i0[id] = b.i0; sep_i1[id] = b.i1; sep_j0[id] = b.j0; sep_j1[id] = b.j1;
            continue;
        }
        if (w >= h) {
            int s = b.i0 + w / 2;
            sep_i0[id] = s; sep_i1[id] = s + 1; sep_j0[id] = b.j0; sep_j1[id] = b.j1;
            stack[top++] = (box_t){b.i0, s, b.j0, b.j1, id};
            stack[top++] = (box_t){s + 1, b.i1, b.j0, b.j1, id};
        } else {
            int s = b.j0 + h / 2;
            sep_i0[id] = b.i0; sep_i1[id] = b.i1; sep_j0[id] = s; sep_j1[id] = s + 1;
            stack[top++] = (box_t){b.i0, b.i1, b.j0, s, id};
            stack[top++] = (box_t){b.i0, b.i1, s + 1, b.j1, id};
        }
    }
    /* postorder numbering of the tree: children before parents */
    int *post = malloc(sizeof(int) * nn), *st = malloc(sizeof(int) * nn), *state = calloc(nn, sizeof(int));
    int np = 0, sp = 0;
    st[sp++] = 0;
    while (sp > 0) {
        int v = st[sp - 1];
        if (state[v] == 0) {
            state[v] = 1;
            if (ch1[v] >= 0) st[sp++] = ch1[v];
            if (ch0[v] >= 0) st[sp++] = ch0[v];
        } else {
            sp--;
            post[np++] = v;
        }
    }
    int *newid = malloc(sizeof(int) * nn);
    for (int k = 0; k < np; k++) newid[post[k]] = k;
    S->nf = np;
    S->fr = calloc(np, sizeof(mf_front));
    int pos = 0;
    for (int k = 0; k < np; k++) {
        int v = post[k];
        mf_front *f = &S->fr[k];
        f->pos0 = pos;
        for (int j = sep_j0[v]; j < sep_j1[v]; j++)
            for (int i = sep_i0[v]; i < sep_i1[v]; i++) {
                S->perm[pos] = i + j * nx;
                S->iperm[i + j * nx] = pos;
                pos++;
            }
        f->p = pos - f->pos0;
        f->parent = par[v] >= 0 ? newid[par[v]] : -1;
    }
    free(sep_i0); free(sep_i1); free(sep_j0); free(sep_j1); free(par); free(ch0); free(ch1);
    free(stack); free(post); free(st); free(state); free(newid);
}

static int cmp_int(const void *a, const void *b) { return (*(const int *)a > *(const int *)b) - (*(const int *)a < *(const int *)b); }

/* ---- symbolic: rows(f) = pivots + ((U children contributions U adj(pivots)) beyond the pivots) ---- */
static void symbolic(mf_solver *S, const int *Ap, const int *Ai)
{
    int nf = S->nf;
    int *cnt = calloc(nf, sizeof(int));
    for (int f = 0; f < nf; f++) if (S->fr[f].parent >= 0) cnt[S->fr[f].parent]++;
    S->kids = malloc(sizeof(int) * (nf + 1));
    int off = 0;
    for (int f = 0; f < nf; f++) { S->fr[f].kid0 = off; S->fr[f].nkids = 0; off += cnt[f]; }
    for (int f = 0; f < nf; f++) {
        int q = S->fr[f].parent;
        if (q >= 0) S->kids[S->fr[q].kid0 + S->fr[q].nkids++] = f;
    }
    free(cnt);
    int *mark = malloc(sizeof(int) * S->n), *buf = malloc(sizeof(int) * S->n);
    for (int i = 0; i < S->n; i++) mark[i] = -1;
    S->nlev = 0;
    for (int f = 0; f < nf; f++) {
        mf_front *F = &S->fr[f];
        int hi = F->pos0 + F->p, nb = 0;
        F->level = 0;
        for (int k = 0; k < F->nkids; k++) {
            mf_front *C = &S->fr[S->kids[F->kid0 + k]];
            if (C->level + 1 > F->level) F->level = C->level + 1;
            for (int r = C->p; r < C->m; r++) {
                int v = C->rows[r];
                if (v >= hi && mark[v] != f) { mark[v] = f; buf[nb++] = v; }
            }
        }
        for (int q = F->pos0; q < hi; q++) {
            int var = S->perm[q];
            /* the pattern is symmetric: column var lists the neighbours */
            for (int t = Ap[var]; t < Ap[var + 1]; t++) {
                int v = S->iperm[Ai[t]];
                if (v >= hi && mark[v] != f) { mark[v] = f; buf[nb++] = v; }
            }
        }
        qsort(buf, nb, sizeof(int), cmp_int);
        F->m = F->p + nb;
        F->rows = malloc(sizeof(int) * F->m);
        for (int i = 0; i < F->p; i++) F->rows[i] = F->pos0 + i;
        memcpy(F->rows + F->p, buf, sizeof(int) * nb);
        if (F->level + 1 > S->nlev) S->nlev = F->level + 1;
    }
    free(mark); free(buf);
    S->lev_ptr = calloc(S->nlev + 1, sizeof(int));
    S->lev_fr = malloc(sizeof(int) * nf);
    for (int f = 0; f < nf; f++) S->lev_ptr[S->fr[f].level + 1]++;
    for (int l = 0; l < S->nlev; l++) S->lev_ptr[l + 1] += S->lev_ptr[l];
    int *fill = calloc(S->nlev, sizeof(int));
    for (int f = 0; f < nf; f++) { int l = S->fr[f].level; S->lev_fr[S->lev_ptr[l] + fill[l]++] = f; }
    free(fill);
}

static int local_index(const mf_front *F, int pos)
{
    int lo = 0, hi = F->m - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        if (F->rows[mid] == pos) return mid;
        if (F->rows[mid] < pos) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

/* partial LU of the first p columns, pivoting among the first p rows;
 * blocks of nb pivots, trailing update parallel over columns when large */
static int partial_lu(cd *F, int m, int p, int *piv, int par)
{
    const int nb = 32;
    int sing = 0;
    for (int k0 = 0; k0 < p; k0 += nb) {
        int kb = p - k0 < nb ? p - k0 : nb;
        for (int j = k0; j < k0 + kb; j++) {
            int r = j;
            double best = -1;
            for (int i = j; i < p; i++) {
                double v = fabs(creal(F[i + (size_t)j * m])) + fabs(cimag(F[i + (size_t)j * m]));
                if (v > best) { best = v; r = i; }
            }
            piv[j] = r;
            if (best == 0) sing = 1;
            if (r != j)
                for (int c = 0; c < m; c++) { cd t = F[j + (size_t)c * m]; F[j + (size_t)c * m] = F[r + (size_t)c * m]; F[r + (size_t)c * m] = t; }
            cd d = 1.0 / F[j + (size_t)j * m];
            for (int i = j + 1; i < m; i++) F[i + (size_t)j * m] *= d;
            for (int c = j + 1; c < k0 + kb; c++) {
                cd u = F[j + (size_t)c * m];
                cd *col = F + (size_t)c * m, *lj = F + (size_t)j * m;
                for (int i = j + 1; i < m; i++) col[i] -= lj[i] * u;
            }
        }
        int c0 = k0 + kb;
        /* U block row: L_kb^-1 * F[k0:k0+kb, c0:m], then trailing update */
        #pragma omp parallel for schedule(static) if (par && (m - c0) > 64)
        for (int c = c0; c < m; c++) {
            cd *col = F + (size_t)c * m;
            for (int j = k0; j < k0 + kb; j++) {
                cd u = col[j];
                const cd *lj = F + (size_t)j * m;
                for (int i = j + 1; i < k0 + kb; i++) col[i] -= lj[i] * u;
            }
            for (int j = k0; j < k0 + kb; j++) {
                cd u = col[j];
                if (u == 0) continue;
                const cd *lj = F + (size_t)j * m;
                for (int i = c0; i < m; i++) col[i] -= lj[i] * u;
            }
        }
    }
    return sing;
}

static void factor_front(mf_solver *S, int f, const int *Ap, const int *Ai, const double *Ax, const double *Az,
                         int par, int *sing)
{
    mf_front *F = &S->fr[f];
    int m = F->m, p = F->p, hi = F->pos0 + p;
    F->F = calloc((size_t)m * m, sizeof(cd));
    F->piv = malloc(sizeof(int) * (p > 0 ? p : 1));
    for (int q = F->pos0; q < hi; q++) {            /* entries owned by this front */
        int var = S->perm[q], lq = q - F->pos0;
        for (int t = Ap[var]; t < Ap[var + 1]; t++) {       /* column var: M[row, var] */
            int pr = S->iperm[Ai[t]];
            if (pr < F->pos0) continue;                       /* owned by a descendant */
            int lr = local_index(F, pr);
            F->F[lr + (size_t)lq * m] += Ax[t] + Az[t] * I;
        }
        /* row var of M (entries M[var, c] with c beyond the pivots): the pattern
         * is symmetric, so column var's rows list the candidate columns c */
        for (int t = Ap[var]; t < Ap[var + 1]; t++) {
            int c = Ai[t], pc = S->iperm[c];
            if (pc < hi) continue;
            int lc = local_index(F, pc);
            for (int u = Ap[c]; u < Ap[c + 1]; u++)
                if (Ai[u] == var) { F->F[lq + (size_t)lc * m] += Ax[u] + Az[u] * I; break; }
        }
    }
    for (int k = 0; k < F->nkids; k++) {             /* extend-add */
        mf_front *C = &S->fr[S->kids[F->kid0 + k]];
        int cm = C->m, cp = C->p;
        int *map = malloc(sizeof(int) * (cm - cp + 1));
        for (int r = cp; r < cm; r++) map[r - cp] = local_index(F, C->rows[r]);
        for (int c = cp; c < cm; c++)
            for (int r = cp; r < cm; r++)
                F->F[map[r - cp] + (size_t)map[c - cp] * m] += C->F[r + (size_t)c * cm];
        free(map);
    }
    if (partial_lu(F->F, m, p, F->piv, par)) *sing = 1;
}

mf_solver *mfc_factor(int nx, int ny, const int *Ap, const int *Ai, const double *Ax, const double *Az, int leaf,
                      int *singular)
{
    mf_solver *S = calloc(1, sizeof(mf_solver));
    S->nx = nx; S->ny = ny; S->n = nx * ny;
    S->perm = malloc(sizeof(int) * S->n);
    S->iperm = malloc(sizeof(int) * S->n);
    nd_order(S, leaf);
    symbolic(S, Ap, Ai);
    int sing = 0;
    for (int l = 0; l < S->nlev; l++) {
        int a = S->lev_ptr[l], b = S->lev_ptr[l + 1];
        if (b - a >= 2 * omp_get_max_threads()) {
            #pragma omp parallel for schedule(dynamic, 1) reduction(| : sing)
            for (int t = a; t < b; t++) factor_front(S, S->lev_fr[t], Ap, Ai, Ax, Az, 0, &sing);
        } else {
            for (int t = a; t < b; t++) factor_front(S, S->lev_fr[t], Ap, Ai, Ax, Az, 1, &sing);
        }
        /* children's update blocks are no longer needed: keep only L and U */
    }
    *singular = sing;
    return S;
}

/* one right-hand side, sys 0 (M) or 2 (M^T); work: n + max m complex */
static void solve_one(const mf_solver *S, int sys, const double *br, const double *bi, double *xr, double *xi, cd *X,
                      cd *w)
{
    int n = S->n;
    for (int q = 0; q < n; q++) X[q] = br[S->perm[q]] + bi[S->perm[q]] * I;
    for (int f = 0; f < S->nf; f++) {            /* forward: L (with P) or U^T */
        const mf_front *F = &S->fr[f];
        int m = F->m, p = F->p;
        for (int i = 0; i < p; i++) w[i] = X[F->pos0 + i];
        if (sys == 0) {
            for (int j = 0; j < p; j++) if (F->piv[j] != j) { cd t = w[j]; w[j] = w[F->piv[j]]; w[F->piv[j]] = t; }
            for (int j = 0; j < p; j++) {
                cd y = w[j];
                const cd *col = F->F + (size_t)j * m;
                for (int i = j + 1; i < p; i++) w[i] -= col[i] * y;
                for (int i = p; i < m; i++) X[F->rows[i]] -= col[i] * y;
            }
        } else {
            for (int j = 0; j < p; j++) {
                const cd *col = F->F + (size_t)j * m;
                cd s = w[j];
                for (int t = 0; t < j; t++) s -= col[t] * w[t];
                w[j] = s / col[j];
            }
            for (int i = p; i < m; i++) {
                const cd *col = F->F + (size_t)i * m;
                cd s = 0;
                for (int t = 0; t < p; t++) s += col[t] * w[t];
                X[F->rows[i]] -= s;
            }
        }
        for (int i = 0; i < p; i++) X[F->pos0 + i] = w[i];
    }
    for (int f = S->nf - 1; f >= 0; f--) {       /* backward: U or L^T (then P^T) */
        const mf_front *F = &S->fr[f];
        int m = F->m, p = F->p;
        for (int i = 0; i < p; i++) w[i] = X[F->pos0 + i];
        if (sys == 0) {
            for (int i = p; i < m; i++) {
                cd x2 = X[F->rows[i]];
                const cd *col = F->F + (size_t)i * m;
                for (int t = 0; t < p; t++) w[t] -= col[t] * x2;
            }
            for (int j = p - 1; j >= 0; j--) {
                const cd *col = F->F + (size_t)j * m;
                w[j] /= col[j];
                cd x = w[j];
                for (int t = 0; t < j; t++) w[t] -= col[t] * x;
            }
        } else {
            for (int j = p - 1; j >= 0; j--) {
                const cd *col = F->F + (size_t)j * m;
                cd s = w[j];
                for (int i = j + 1; i < p; i++) s -= col[i] * w[i];
                for (int i = p; i < m; i++) s -= col[i] * X[F->rows[i]];
                w[j] = s;
            }
            for (int j = p - 1; j >= 0; j--) if (F->piv[j] != j) { cd t = w[j]; w[j] = w[F->piv[j]]; w[F->piv[j]] = t; }
        }
        for (int i = 0; i < p; i++) X[F->pos0 + i] = w[i];
    }
    for (int q = 0; q < n; q++) { xr[S->perm[q]] = creal(X[q]); xi[S->perm[q]] = cimag(X[q]); }
}

/* nrhs right-hand sides (column r at offset r*n), spread over the threads */
void mfc_solve(const mf_solver *S, int sys, int nrhs, const double *br, const double *bi, double *xr, double *xi)
{
    int maxm = 1;
    for (int f = 0; f < S->nf; f++) if (S->fr[f].m > maxm) maxm = S->fr[f].m;
    #pragma omp parallel
    {
        cd *X = malloc(sizeof(cd) * S->n), *w = malloc(sizeof(cd) * maxm);
        #pragma omp for schedule(dynamic, 1)
        for (int r = 0; r < nrhs; r++)
            solve_one(S, sys, br + (size_t)r * S->n, bi + (size_t)r * S->n, xr + (size_t)r * S->n,
                      xi + (size_t)r * S->n, X, w);
        free(X); free(w);
    }
}

double mfc_factor_flops(const mf_solver *S)
{
    double fl = 0;
    for (int f = 0; f < S->nf; f++)
        for (int k = 0; k < S->fr[f].p; k++) { double t = S->fr[f].m - k - 1; fl += 4.0 * (2.0 * t * t + t); }
    return fl;
}

int mfc_threads(void) { return omp_get_max_threads(); }

void mfc_set_threads(int t) { if (t > 0) omp_set_num_threads(t); }

void mfc_free(mf_solver *S)
{
    if (!S) return;
    for (int f = 0; f < S->nf; f++) { free(S->fr[f].rows); free(S->fr[f].F); free(S->fr[f].piv); }
    free(S->fr); free(S->kids); free(S->lev_ptr); free(S->lev_fr); free(S->perm); free(S->iperm); free(S);
}
