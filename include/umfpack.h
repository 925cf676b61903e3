/*
 * Drop-in replacement header for the six UMFPACK entry points GERMAINE calls.
 *
 * GERMAINE includes <umfpack.h> (include/fd.h:18) and links -lumfpack
 * (src/Makefile:32-35); building it with -I<this include dir> and linking
 * libgfdfd.so instead resolves exactly these six symbols, which are the only
 * unresolved names of the reference's object set.  The prototypes and the
 * constants below follow SuiteSparse's public UMFPACK API (int-index,
 * split-complex "zi" variant); re-check the numeric constants against an
 * installed umfpack.h if one exists.
 *
 * Scope of the replacement: square matrices with the 9-point FDFD grid
 * pattern (any nx, ny >= 2) are analysed by geometric nested dissection and
 * factored on the GPU; any other pattern is accepted up to n = 4096 and
 * factored as a single dense front.  Larger general patterns return
 * UMFPACK_ERROR_invalid_matrix.
 */
#ifndef GFDFD_UMFPACK_H
#define GFDFD_UMFPACK_H

#ifdef __cplusplus
extern "C" {
#endif

/* sys argument of umfpack_zi_solve */
#define UMFPACK_A   0   /* M x = b   */
#define UMFPACK_At  1   /* M^H x = b (complex conjugate transpose) */
#define UMFPACK_Aat 2   /* M^T x = b (array transpose) */

/* status codes */
#define UMFPACK_OK 0
#define UMFPACK_WARNING_singular_matrix 1
#define UMFPACK_ERROR_out_of_memory (-1)
#define UMFPACK_ERROR_invalid_Numeric_object (-3)
#define UMFPACK_ERROR_invalid_Symbolic_object (-4)
#define UMFPACK_ERROR_argument_missing (-5)
#define UMFPACK_ERROR_n_nonpositive (-6)
#define UMFPACK_ERROR_invalid_matrix (-8)
#define UMFPACK_ERROR_different_pattern (-11)
#define UMFPACK_ERROR_invalid_system (-13)
#define UMFPACK_ERROR_internal_error (-911)

/* Control / Info arrays (may be NULL; GERMAINE always passes NULL,
 * src/AC/forward_shot_AC.c:21,58,61,88) */
#define UMFPACK_CONTROL 20
#define UMFPACK_INFO 90
#define UMFPACK_PRL 0
#define UMFPACK_IRSTEP 7
#define UMFPACK_DEFAULT_IRSTEP 2
#define UMFPACK_STATUS 0
#define UMFPACK_NROW 1
#define UMFPACK_NZ 2
#define UMFPACK_NCOL 16
#define UMFPACK_IR_TAKEN 80
#define UMFPACK_OMEGA1 82

/* replaces the triplet conversion at src/AC/forward_shot_AC.c:41 (11 call sites) */
int umfpack_zi_triplet_to_col(int n_row, int n_col, int nz, const int Ti[], const int Tj[], const double Tx[],
                              const double Tz[], int Ap[], int Ai[], double Ax[], double Az[], int Map[]);

/* replaces the symbolic analysis at src/AC/forward_shot_AC.c:58 (11 call sites) */
int umfpack_zi_symbolic(int n_row, int n_col, const int Ap[], const int Ai[], const double Ax[], const double Az[],
                        void **Symbolic, const double Control[], double Info[]);

/* replaces the numeric factorisation at src/AC/forward_shot_AC.c:61 (11 call sites) */
int umfpack_zi_numeric(const int Ap[], const int Ai[], const double Ax[], const double Az[], void *Symbolic,
                       void **Numeric, const double Control[], double Info[]);

/* replaces the solves at src/AC/forward_shot_AC.c:88, src/AC/grad_obj_AC.c:121,147,
 * src/TE/grad_obj_TE.c:123,149 and the other 11 call sites */
int umfpack_zi_solve(int sys, const int Ap[], const int Ai[], const double Ax[], const double Az[], double Xx[],
                     double Xz[], const double Bx[], const double Bz[], void *Numeric, const double Control[],
                     double Info[]);

/* replaces src/AC/forward_shot_AC.c:62 */
void umfpack_zi_free_symbolic(void **Symbolic);

/* replaces src/AC/forward_shot_AC.c:120 */
void umfpack_zi_free_numeric(void **Numeric);

#ifdef __cplusplus
}
#endif
#endif
