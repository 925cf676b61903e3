/*
 * A C caller of the drop-in, compiled against include/umfpack.h and linked
 * with libgfdfd.so, that runs GERMAINE's per-frequency call sequence of
 * src/AC/forward_shot_AC.c:41-120 on a matrix and shot list read from a file:
 *
 *   umfpack_zi_triplet_to_col -> overwrite Ax/Az with the triplet values
 *   (:43-48) -> umfpack_zi_symbolic (:58) -> umfpack_zi_numeric (:61) ->
 *   umfpack_zi_free_symbolic (:62) -> per shot: point source (RHS_source_AC),
 *   umfpack_zi_solve(UMFPACK_A) (:88) -> umfpack_zi_free_numeric (:120)
 *
 * with Control and Info NULL, as GERMAINE passes them.  Unlike GERMAINE it
 * checks every status and stops at the first failing one.
 *
 *   forward_shot_host IN OUT [triplet]
 *
 * IN: int32 n, nz, nshots; int32 Ti[nz], Tj[nz]; float64 Tx[nz], Tz[nz];
 *     int32 ksrc[nshots]; float64 amp_re, amp_im.
 * OUT: int32 Ap[n+1], Ai[nz] (the CSC pattern), then, unless "triplet" is
 *     given, float64 (xr, xi) of every shot, shot after shot (2 n each).
 * Prints one line "status <call> <code>" per call.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "umfpack.h"

static void *xalloc(size_t bytes) {
    void *p = calloc(bytes ? bytes : 1, 1);
    if (!p) {
        fprintf(stderr, "out of host memory\n");
        exit(3);
    }
    return p;
}

static void read_exact(FILE *f, void *dst, size_t bytes) {
    if (fread(dst, 1, bytes, f) != bytes) {
        fprintf(stderr, "short input file\n");
        exit(3);
    }
}

static int report(const char *call, int status) {
    printf("status %s %d\n", call, status);
    fflush(stdout);
    return status;
}

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s IN OUT [triplet]\n", argv[0]);
        return 2;
    }
    const int triplet_only = argc > 3 && strcmp(argv[3], "triplet") == 0;
    FILE *in = fopen(argv[1], "rb");
    if (!in) {
        perror(argv[1]);
        return 2;
    }
    int hdr[3];
    read_exact(in, hdr, sizeof hdr);
    const int n = hdr[0], nz = hdr[1], nshots = hdr[2];
    if (n <= 0 || nz <= 0 || nshots < 0) {
        fprintf(stderr, "bad header\n");
        return 2;
    }
    int *Ti = xalloc(sizeof(int) * nz), *Tj = xalloc(sizeof(int) * nz), *ksrc = xalloc(sizeof(int) * (nshots + 1));
    double *Tx = xalloc(sizeof(double) * nz), *Tz = xalloc(sizeof(double) * nz), amp[2];
    read_exact(in, Ti, sizeof(int) * nz);
    read_exact(in, Tj, sizeof(int) * nz);
    read_exact(in, Tx, sizeof(double) * nz);
    read_exact(in, Tz, sizeof(double) * nz);
    read_exact(in, ksrc, sizeof(int) * nshots);
    read_exact(in, amp, sizeof amp);
    fclose(in);

    /* the host's buffers, sized as src/AC/forward_shot_AC.c:30-35 sizes them */
    int *Ap = xalloc(sizeof(int) * (n + 1)), *Ai = xalloc(sizeof(int) * nz);
    double *Ax = xalloc(sizeof(double) * nz), *Az = xalloc(sizeof(double) * nz);
    double *xr = xalloc(sizeof(double) * nz), *xi = xalloc(sizeof(double) * nz);
    double *br = xalloc(sizeof(double) * n), *bi = xalloc(sizeof(double) * n);
    void *Symbolic = NULL, *Numeric = NULL;
    int rc = 0;

    FILE *out = fopen(argv[2], "wb");
    if (!out) {
        perror(argv[2]);
        return 2;
    }
    if (report("triplet_to_col", umfpack_zi_triplet_to_col(n, n, nz, Ti, Tj, Tx, Tz, Ap, Ai, Ax, Az, NULL)) != UMFPACK_OK)
        return 1;
    fwrite(Ap, sizeof(int), (size_t)n + 1, out);
    fwrite(Ai, sizeof(int), (size_t)nz, out);
    if (triplet_only) {
        fclose(out);
        return 0;
    }
    for (int i = 0; i < nz; i++) {   /* the overwrite of src/AC/forward_shot_AC.c:43-48 */
        Ax[i] = Tx[i];
        Az[i] = Tz[i];
    }
    if (report("symbolic", umfpack_zi_symbolic(n, n, Ap, Ai, Ax, Az, &Symbolic, NULL, NULL)) != UMFPACK_OK) {
        rc = 1;
        goto done;
    }
    if (report("numeric", umfpack_zi_numeric(Ap, Ai, Ax, Az, Symbolic, &Numeric, NULL, NULL)) != UMFPACK_OK) {
        rc = 1;
        goto done;
    }
    umfpack_zi_free_symbolic(&Symbolic);
    report("free_symbolic", Symbolic == NULL ? 0 : -1);
    for (int s = 0; s < nshots; s++) {
        memset(br, 0, sizeof(double) * n);   /* RHS_source_AC: one point source per shot */
        memset(bi, 0, sizeof(double) * n);
        if (ksrc[s] < 0 || ksrc[s] >= n) {
            fprintf(stderr, "source index out of range\n");
            rc = 2;
            goto done;
        }
        br[ksrc[s]] = amp[0];
        bi[ksrc[s]] = amp[1];
        const int st = umfpack_zi_solve(UMFPACK_A, Ap, Ai, Ax, Az, xr, xi, br, bi, Numeric, NULL, NULL);
        if (st != UMFPACK_OK) {
            report("solve", st);
            rc = 1;
            goto done;
        }
        fwrite(xr, sizeof(double), (size_t)n, out);
        fwrite(xi, sizeof(double), (size_t)n, out);
    }
    report("solve", UMFPACK_OK);
done:
    if (Symbolic) umfpack_zi_free_symbolic(&Symbolic);
    if (Numeric) {
        umfpack_zi_free_numeric(&Numeric);
        report("free_numeric", Numeric == NULL ? 0 : -1);
    }
    fclose(out);
    free(Ti); free(Tj); free(Tx); free(Tz); free(ksrc); free(Ap); free(Ai); free(Ax); free(Az);
    free(xr); free(xi); free(br); free(bi);
    return rc;
}
