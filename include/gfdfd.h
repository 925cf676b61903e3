/*
 * gfdfd — extension C-ABI of the MI355X solve path.
 *
 * What GERMAINE's one-shot-at-a-time host loop cannot express through the six
 * UMFPACK entry points (include/umfpack.h): device assembly, several
 * frequencies factored in one batch, multi-right-hand-side panel solves,
 * FP32 factors with FP64 refinement, and the shot kernels of the FWI
 * gradient.  All device arrays are plain device pointers (hipMalloc'd or
 * from any allocator); complex arrays are interleaved (re, im) doubles or
 * floats; right-hand-side panels are column-major n x nrhs (one shot per
 * column).  `stream` is a hipStream_t (NULL = default stream).  Functions
 * return GFDFD_OK, a positive warning or a negative error, and print a
 * diagnostic on stderr for every non-zero status; gfdfd_last_error() returns
 * the text.
 *
 * Unknown numbering: k = (i-1) + (j-1)*nx for the reference's 1-based grid
 * point (i,j) (src/AC/init_A_AC_9p_pml.c:52-53); grids include the PML.
 */
#ifndef GFDFD_H
#define GFDFD_H

#ifdef __cplusplus
extern "C" {
#endif

#define GFDFD_OK 0
#define GFDFD_WARN_SINGULAR 1
#define GFDFD_ERR_ARG (-1)
#define GFDFD_ERR_DEVICE (-2)
#define GFDFD_ERR_MEMORY (-3)
#define GFDFD_ERR_INTERNAL (-4)

#define GFDFD_AC 1
#define GFDFD_SH 2
#define GFDFD_TE 3

#define GFDFD_FP64 64   /* complex double fronts */
#define GFDFD_FP32 32   /* complex float fronts (FP32 MFMA), refine in FP64 */

typedef struct gfdfd_plan_s *gfdfd_plan;
typedef struct gfdfd_factors_s *gfdfd_factors;

typedef struct {
    long long n, nnz;           /* unknowns, entries of M */
    int fronts, levels, max_front, max_pivots;
    long long arena_elems;      /* front entries per frequency (sum m^2) */
    long long work_rows;        /* solve work-panel rows (sum m) */
    double factor_flops;        /* real flops of one factorisation (complex ops x4) */
    double nnz_lu;              /* entries of L and U */
    double plan_device_bytes;
} gfdfd_plan_info;

typedef struct {
    float omega2[2];            /* Omega^2 as gfdfd_omega returns it */
    float omega[2];             /* Omega (TE conductivity term) */
    const float *prof;          /* host: PML profiles, 4*(nx+2) + 4*(ny+2) floats (gfdfd_pml_profiles) */
} gfdfd_freq;

const char *gfdfd_last_error(void);

/* ---- symbolic: geometric nested dissection of the nx*ny 9-point grid -------
 * replaces umfpack_zi_symbolic (src/AC/forward_shot_AC.c:58) for the batched
 * path; plans are cached by grid size, so repeated calls are cheap.  Plans
 * are host objects (no device needed) until a device call first uses them.  */
int gfdfd_plan_grid(int nx, int ny, gfdfd_plan *out);
int gfdfd_plan_grid_opt(int nx, int ny, int leaf_area, int small_max, gfdfd_plan *out);
void gfdfd_plan_free(gfdfd_plan *plan);
int gfdfd_plan_query(gfdfd_plan plan, gfdfd_plan_info *info);
/* host copies of the canonical CSC pattern of M (colptr[n+1], rowind[nnz]) */
int gfdfd_plan_pattern(gfdfd_plan plan, int *colptr, int *rowind);
/* elimination order: perm[pos] = unknown (n ints) */
int gfdfd_plan_perm(gfdfd_plan plan, int *perm);
/* front table (nfronts entries each, postorder): pivots, order, first pivot
 * position, parent (-1 root), height, offset of the front's index list */
int gfdfd_plan_fronts(gfdfd_plan plan, int *p, int *m, int *pos0, int *parent, int *level, long long *idx_off);
/* concatenated front index lists (elimination positions, ascending per front) */
int gfdfd_plan_index(gfdfd_plan plan, int *idx);

/* ---- host-side per-frequency parameters (gcc arithmetic of the reference) -- */
/* src/AC/init_A_AC_9p_pml.c:34-42: a, b/4, c/4, d, e, 1/DH^2 */
void gfdfd_fd_weights(float dh, float *out6);
float gfdfd_mu0(void);
/* src/pml_pro.c:20-160 */
void gfdfd_pml_profiles(int nx, int ny, int npml, float dh, float a0_pml, float freq, float s, int free_surf,
                        float *prof);
/* src/AC/init_A_AC_9p_pml.c:50 (AC), src/SH/init_A_SH_9p_pml.c:50, src/TE/init_A_TE_9p_pml.c:54-55 */
void gfdfd_omega(int physics, float freq, float s, float *omega2, float *omega);
/* Omega^2 of the AC gradient, cpowf(2 pi f - i S, 2) (src/AC/ass_grad_AC.c:22; equals the
 * assembly's Omega^2 only for S = 0) */
void gfdfd_omega2_grad_ac(float freq, float s, float *omega2);
/* src/AC/RHS_source_AC.c:37-53, src/TE/RHS_source_TE.c:22-56 */
void gfdfd_source_amp(int physics, float freq, float s, float dh, double *amp);
/* READ_WAVELET = 1 (src/AC/RHS_source_AC.c:41-43, src/TE/RHS_source_TE.c:44-46):
 * the DFT of the SU wavelet at freq (src/read_stf_dft.c:34-70) over DH^2.
 * samples: the trace's ns values; ndt: its header's sample interval
 * (dt = ndt/1e6 for AC/SH, ndt/1e12 for TE).  The reference sums data[1..ns]
 * at t = (j-1) dt after reading data[0..ns-1]; its last term reads an element
 * the read never filled, which counts as zero here.                         */
void gfdfd_wavelet_amp(int physics, float freq, float dh, const float *samples, int ns, int ndt, double *amp);
/* taper_grad_shot's exponent a (src/taper_grad_shot.c:36-39), in float     */
float gfdfd_shot_taper_exponent(void);

/* ---- device assembly --------------------------------------------------------
 * replaces init_mat_AC + init_A_AC_9p_pml (src/AC/init_A_AC_9p_pml.c:17-220),
 * init_A_SH_9p_pml (src/SH/init_A_SH_9p_pml.c:17-222) and init_A_TE_9p_pml
 * (src/TE/init_A_TE_9p_pml.c:21-243) followed by the CSC value overwrite
 * (src/AC/forward_shot_AC.c:41-48).  d_m1/d_m2: AC vp/rho, SH vs/rho, TE
 * absolute permittivity/conductivity, device float [ny][nx].  d_vals receives
 * nfreq x nnz complex doubles: the values of M in canonical CSC order,
 * bit-identical to the reference's (double) triplet values.                 */
int gfdfd_assemble(gfdfd_plan plan, int physics, float dh, int nfreq, const gfdfd_freq *freqs, const float *d_m1,
                   const float *d_m2, double *d_vals, void *stream);

/* ---- numeric: nfreq matrices of one pattern in one batch -------------------
 * replaces umfpack_zi_numeric (src/AC/forward_shot_AC.c:61).  d_vals: nfreq x
 * nnz complex doubles in canonical CSC order (copied; the factors keep an FP64
 * copy for refinement).  precision GFDFD_FP64 or GFDFD_FP32.               */
int gfdfd_factor(gfdfd_plan plan, int nfreq, const double *d_vals, int precision, void *stream, gfdfd_factors *out);
/* refactor in place: new values (same pattern, same nfreq, nfreq x nnz) into
 * existing factors, reusing their device storage -- the next FWI iteration or
 * line-search trial without a new allocation (src/AC/fwi_FD_AC.c:229-409
 * refactors every frequency with umfpack_zi_numeric each time)              */
int gfdfd_refactor(gfdfd_factors f, const double *d_vals, void *stream);
/* gfdfd_refactor as one hipGraph launch: the first call with a given d_vals
 * captures the value copy and every factorisation kernel (all frequencies,
 * all levels) into a graph; later calls with the same d_vals pointer replay
 * it -- write the next values into that buffer (e.g. with gfdfd_assemble)
 * and call again.  A new pointer recaptures.  Graph launches record no
 * gfdfd_profile events.
 * The FWI loop's per-iteration refactorisation (src/AC/fwi_FD_AC.c:229-409)
 * without per-kernel launch cost.                                           */
int gfdfd_refactor_graph(gfdfd_factors f, const double *d_vals, void *stream);
void gfdfd_factors_free(gfdfd_factors *f);
/* per-frequency zero-pivot flags (nfreq ints) */
int gfdfd_factors_singular(gfdfd_factors f, int *flags);

/* ---- solves ------------------------------------------------------------------
 * replaces umfpack_zi_solve (src/AC/forward_shot_AC.c:88, src/AC/grad_obj_AC.c:
 * 121,147, src/TE/grad_obj_TE.c:149) for a whole panel of shots.
 * sys: 0 M x = b, 1 M^H x = b, 2 M^T x = b.  d_b and d_x: n x nrhs complex
 * doubles; gfdfd_solve allows d_x == d_b.                                   */
int gfdfd_solve(gfdfd_factors f, int ifreq, int sys, int nrhs, const double *d_b, double *d_x, void *stream);
/* gfdfd_solve for 1 to 4 right-hand sides with no device allocation (no
 * counterpart in the reference: for callers that cannot allocate -- code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * work buffers (n + work_rows entries of the factor precision, aligned) live
 * in a file-scope device array of gfdfd_solve_noalloc_limit() bytes (64 MiB)
 * that every GPU the library runs on holds until the process ends; a grid
 * needing more is rejected with GFDFD_ERR_ARG (FP64 factors fit up to
 * 1601 x 401).  All calls on one GPU share that GPU's array, so calls that
 * overlap on one GPU race; d_x may equal d_b.                               */
int gfdfd_solve_noalloc(gfdfd_factors f, int ifreq, int sys, int nrhs, const double *d_b, double *d_x,
                        void *stream);
long long gfdfd_solve_noalloc_limit(void);
/* solve with FP64 iterative refinement (residual b - op(M) x from the FP64
 * values) until max over columns of ||r||/||b|| <= tol or max_iter
 * corrections; d_x must not alias d_b.  iters/relres may be NULL.           */
int gfdfd_solve_refine(gfdfd_factors f, int ifreq, int sys, int nrhs, const double *d_b, double *d_x, double tol,
                       int max_iter, int *iters, double *relres, void *stream);
/* the same for nsel frequencies at once: freqs[z] (host array) is solved on
 * panel z of d_b / d_x (nsel consecutive n x nrhs panels), all in the same
 * kernel launches; at most 64 frequencies per call.  The refining variant
 * drops a frequency from the batch once it has converged; iters/relres
 * (nsel entries each) may be NULL.                                          */
int gfdfd_solve_batch(gfdfd_factors f, int nsel, const int *freqs, int sys, int nrhs, const double *d_b, double *d_x,
                      void *stream);
int gfdfd_solve_refine_batch(gfdfd_factors f, int nsel, const int *freqs, int sys, int nrhs, const double *d_b,
                             double *d_x, double tol, int max_iter, int *iters, double *relres, void *stream);
/* r = b - op(M) x in FP64 */
int gfdfd_residual(gfdfd_plan plan, const double *d_vals, int sys, int nrhs, const double *d_b, const double *d_x,
                   double *d_r, void *stream);

/* kernel selection of the panel solves (process-wide; no counterpart in the
 * reference): the widest right-hand-side tile of the fused per-front kernels
 * (16, 32 or 64; default 64 or GFDFD_FUSED_RW) and the largest front order of
 * a level with multi-block fronts that still takes them (negative: as far as
 * LDS allows; 0, the default: such levels take the per-block kernels), and
 * the layout the refinement of 5 to 1024 right-hand sides works in (0, the
 * default: elimination order, one fused residual-and-update pass per step;
 * 1: the caller's column-major layout, two transposes per correction).      */
int gfdfd_solve_tuning(int rw_cap, int fused_mmax, int refine_layout);
/* Factorisation kernel selection (process-wide, for later factorisations and
 * graph captures): a 32-column panel whose candidate rows need more than
 * panel_lds_cap bytes of LDS is factored by the global-memory panel kernel
 * instead of in LDS (0: the hardware limit, 159 KB -- separators above 318
 * rows in FP64, 637 in FP32), which lets small systems exercise that path;
 * persistent != 0 (the default) factors the small subtrees -- fronts of order
 * <= 64 whose descendants are all that small -- in one persistent launch in
 * which a workgroup factors one whole subtree of one frequency per task, 0
 * level by level; levels whose largest blocked front has order >= deep_min_m
 * (default 0: every level) factor two 32-column panels per step and update
 * the trailing block once per 64 pivots, the others once per panel.        */
int gfdfd_factor_tuning(int panel_lds_cap, int persistent, int deep_min_m);

/* ---- shot kernels --------------------------------------------------------------
 * RHS_source_AC / RHS_source_TE for all shots (src/AC/RHS_source_AC.c:46-53):
 * column s of d_b is zero except d_b[k[s]] = amp.                           */
int gfdfd_spike_panel(int n, int nshots, const int *d_k, double amp_re, double amp_im, double *d_b, void *stream);
/* calc_seis_AC after vec2mat (src/AC/calc_seis_AC.c:19-20): d_pcal[s*ntr+t] =
 * (complex float) d_x[d_krec[t] + s*n]                                       */
int gfdfd_sample(int n, int nshots, int ntr, const int *d_krec, const double *d_x, float *d_pcal, void *stream);
/* calc_res_AC (MISFIT=1, STF_INV=0, src/AC/calc_res_AC.c:49-95) and
 * RHS_source_AC_adj (src/AC/RHS_source_AC_adj.c:21-36) for all shots:
 * d_b column s = conj(pobs - pcal)/DH^2 at the receivers (muted where
 * pobsr<1e-20 && pobsi<1e-20; for duplicate receivers the last one wins),
 * d_l2[s] = sum |pobs - pcal|^2 (double).                                   */
int gfdfd_adjoint_sources(int n, int nshots, int ntr, const int *d_krec, const float *d_pcal, const float *d_pobs,
                          float dh, double *d_b, double *d_l2, void *stream);
/* calc_res_AC in full (src/AC/calc_res_AC.c:10-95) with RHS_source_AC_adj,
 * for all shots.  misfit 1 (L2, :62), 2 (logarithmic, clog(pobs) -
 * clog(pcal), :66) or 3 (phase only, arg(pcal) - arg(pobs), :70); rtm != 0
 * is INVMAT=2 (residual = observed data, :77).  stf_inv = 1 estimates each
 * shot's source-time function as the Wiener quotient sum conj(pcal) pobs /
 * sum conj(pobs) pobs in complex float (:21-40; else 1) and writes it to
 * d_stf[2s..2s+1] (complex float, may be NULL) for gfdfd_grad_*_stf.  Mute,
 * duplicate receivers and d_l2 as gfdfd_adjoint_sources.                  */
int gfdfd_residual_sources(int n, int nshots, int ntr, const int *d_krec, const float *d_pcal, const float *d_pobs,
                           float dh, int misfit, int stf_inv, int rtm, double *d_b, double *d_l2, float *d_stf,
                           void *stream);
/* ass_grad_AC (src/AC/ass_grad_AC.c:30-31), STF_INV=0, summed over the nshots
 * columns of d_u (forward) and d_lam (adjoint) in FP64 and added to d_grad
 * (n doubles).  omega2 from gfdfd_omega2_grad_ac.  Cell q combines the
 * solutions at q+1 (vec2mat's shift, src/vec2mat.c:19-24); the last cell,
 * which reads past the solution in the reference, is left unchanged.        */
int gfdfd_grad_ac(int n, int nshots, const float *omega2, const double *d_u, const double *d_lam, const float *d_vp,
                  const float *d_rho, double *d_grad, void *stream);
/* ass_grad_AC with the shots' source-time functions (src/AC/ass_grad_AC.c:21,
 * 30-31): shot s weighted by wien_s/|wien_s|, wien_s = d_stf[2s] + i
 * d_stf[2s+1] from gfdfd_residual_sources (NULL: 1).                       */
int gfdfd_grad_ac_stf(int n, int nshots, const float *omega2, const double *d_u, const double *d_lam,
                      const float *d_vp, const float *d_rho, const float *d_stf, double *d_grad, void *stream);
/* ass_grad_AC's per-shot taper (SWS_TAPER_CIRCULAR_PER_SHOT, src/AC/ass_grad_AC.c:37-43,
 * src/TE/ass_grad_TE.c:40-49).  mode 1, taper_grad_shot (src/taper_grad_shot.c:34-87):
 * each shot's gradient times f0 + (1-f0)(0.5 - 0.5 cos(pi r/R))^a within r <= R of the
 * shot's source and of every receiver, r from (ix DH, iy DH) of the PML-extended grid's
 * 1-based cell (ix, iy) to srcpos (m) and to recpos * DH, as the reference measures it;
 * mode 2, taper_grad_shot1 (src/taper_grad_shot1.c:32-66): each shot's gradient is 1e-5
 * on the (2 FILTSIZE + 1)^2 cell boxes around iround(srcpos/DH) of every source and
 * around every receiver.  Device arrays.                                    */
typedef struct {
    int mode;                       /* 0 none, 1 taper_grad_shot, 2 taper_grad_shot1 */
    int nx;                         /* NX (cells per grid row, PML included); n must be a multiple */
    float dh;                       /* DH */
    float srtradius;                /* SRTRADIUS, mode 1 */
    int filtsize;                   /* FILTSIZE, mode 2 */
    const float *d_srcx, *d_srcy;   /* mode 1: srcpos[1..2] of the panel's nshots shots, metres */
    int nsrc;                       /* mode 2: every source of the survey, */
    const int *d_src_i, *d_src_j;   /*   iround(srcpos[1..2] / DH) */
    int nrec;                       /* receivers: recpos[1..2] (grid indices, receiver()) */
    const int *d_rec_i, *d_rec_j;
} gfdfd_shot_taper;
/* gfdfd_grad_ac_stf with the per-shot taper (taper NULL or mode 0: none)   */
int gfdfd_grad_ac_taper(int n, int nshots, const float *omega2, const double *d_u, const double *d_lam,
                        const float *d_vp, const float *d_rho, const float *d_stf, const gfdfd_shot_taper *taper,
                        double *d_grad, void *stream);
/* hessian_shin_AC (src/AC/hessian_shin_AC.c:117-124): Shin's pseudo-Hessian
 * diagonal, hess += |J|^2 with J = 2 Omega^2 u / (rho vp^3), summed over the
 * nshots forward fields of d_u in FP64 and added to d_hess (n doubles).
 * omega2 from gfdfd_omega2_grad_ac (the same 2 pi f - i S as the gradient);
 * cell q uses the solution at q+1, the last cell is left unchanged.          */
int gfdfd_hess_shin_ac(int n, int nshots, const float *omega2, const double *d_u, const float *d_vp, const float *d_rho,
                       double *d_hess, void *stream);
/* apply_hess_AC (src/AC/apply_hess_AC.c:20-37; the TE path applies it to each
 * material, src/TE/fwi_FD_TE.c:287-288): d_grad *= 1 / (d_hess + eps_hess *
 * max|d_hess|) over n cells, in FP64, without scratch memory.               */
int gfdfd_apply_hess(int n, float eps_hess, const double *d_hess, double *d_grad, void *stream);
/* calc_mat_change_wolfe (src/calc_mat_change_wolfe.c:27-46), the trial
 * model of each line-search step (src/AC/parabolicls_AC.c:58-64): d_m_new =
 * d_m_old + alpha * d_dir in float over n cells; a value below lo (MAT1_LOW),
 * above hi (MAT1_UP) or below zero keeps d_m_old.  d_dir is the FP64 search
 * direction (rounded to float); d_m_new may equal d_m_old.                   */
int gfdfd_model_update(int n, float alpha, const double *d_dir, const float *d_m_old, float lo, float hi,
                       float *d_m_new, void *stream);
/* The TE line search's trial model (src/TE/parabolicls_TE.c:30-31,64-69,
 * calc_mat_change_wolfe_multi_para, src/calc_mat_change_wolfe_multi_para.c:22-50):
 * r = (1/norm) * d_m_old (MAT1_NORM or MAT2_NORM), r += alpha * d_dir, r
 * clamped to [lo/norm, hi/norm] (MATx_LOW, MATx_UP), d_m_new = norm * r, in
 * float over n cells.  norm must be nonzero; d_m_new may equal d_m_old.      */
int gfdfd_model_update_norm(int n, float alpha, const double *d_dir, const float *d_m_old, float norm, float lo,
                            float hi, float *d_m_new, void *stream);
/* precond (src/precond.c:17-26) without the taper file: with spatfilter,
 * gauss_filt (src/gauss_filt.c:25-145) smooths the [ny, nx] field d_in with a
 * normalised Gaussian whose window spans filt_size_x (FILT_SIZE_GRAD) and
 * filt_size_y (FILT_SIZE_GRAD1) wavelengths vref / fc_low (VREF, FC_low) on a
 * grid of spacing dh, the field extended by its edge values; then
 * taper_grad_hor (src/taper_grad_hor.c:22-31) zeroes the first nzero_rows
 * rows (GRADT2 + FSSHIFT when SWS_TAPER_GRAD_HOR, else 0).  Arithmetic in
 * float as the reference's; results in d_out (not d_in).  The window may hold
 * at most 16000 weights.                                                     */
int gfdfd_precond(int nx, int ny, float dh, float vref, float fc_low, float filt_size_x, float filt_size_y,
                  int spatfilter, int nzero_rows, const double *d_in, double *d_out, void *stream);
/* Tikhonov regularisation of the TE materials (src/TE/Tikhonov_cost_TE.c:21-45,
 * src/TE/Tikhonov_grad_TE.c:20-37): with r = (BETA / NORM) * m for sigma and
 * epsilon ([ny, nx] float), the five-point Laplacian lap of each interior cell
 * over spacing dh gives d_cost[0] = sum 0.5 BETA_MAT1 lap_sigma^2 and d_cost[1]
 * = sum 0.5 BETA_MAT2 lap_eps^2 (two device doubles, overwritten), and, for
 * each gradient pointer that is not null, grad += LAMBDA * BETA * lap.  The
 * caller adds LAMBDA_1 cost[0] + LAMBDA_2 cost[1] to the misfit when either
 * LAMBDA is positive, as Tikhonov_cost_TE does.  No scratch memory.          */
int gfdfd_tikhonov_te(int nx, int ny, float dh, float beta_sigma, float beta_eps, float norm_sigma, float norm_eps,
                      float lambda_sigma, float lambda_eps, const float *d_sigma, const float *d_eps,
                      double *d_grad_sigma, double *d_grad_eps, double *d_cost, void *stream);
/* taper_grad_file (src/taper_grad_file.c:22-66), precond's last step when
 * SWS_TAPER_FILE: the field d_in on the grid with PML ([ny, nx], nx = nx0 +
 * 2 npml, ny = ny0 + npml + (free_surf ? 0 : npml)) times d_taper (the
 * taper.bin values, [ny0, nx0] float) on the model part, the PML then filled
 * by the reference's edge copies (its top rows only with a free surface, as
 * the reference does).  Result in d_out (not d_in).                          */
int gfdfd_taper_file(int nx0, int ny0, int npml, int free_surf, const float *d_taper, const double *d_in,
                     double *d_out, void *stream);
/* cp_grad_frame (src/cp_grad_frame.c:102-208, its PML_grad = 3 branch): damps
 * the gradient d_grad (in place, [ny, nx] of the grid with PML, nx = nx0 +
 * 2 npml, ny = ny0 + npml + (free_surf ? 0 : npml)) in the PML with the
 * reference's 60 % Gaussian taper, corners twice as the reference does, in
 * float.  npml >= 2 (the taper divides by npml - 1).                        */
int gfdfd_grad_frame(int nx0, int ny0, int npml, int free_surf, double *d_grad, void *stream);
/* hessian_AC (src/AC/hessian_AC.c:129-150): Gauss-Newton Hessian diagonal,
 * hess += |2 Omega^2 u g / (rho vp^3)|^2 over every pair of the nshots forward
 * fields d_u and the nrec receiver Green's functions d_g (solutions for a
 * 1/DH^2 spike at each receiver, RHS_source_AC_hess.c:20-33), in FP64, added to
 * d_hess (n doubles); omega2 from gfdfd_omega2_grad_ac.  Cell q uses the
 * solutions at q+1; the last cell is left unchanged.                          */
int gfdfd_hess_gn_ac(int n, int nshots, int nrec, const float *omega2, const double *d_u, const double *d_g,
                     const float *d_vp, const float *d_rho, double *d_hess, void *stream);
/* hessian_TE (src/TE/hessian_TE.c:149-157): Gauss-Newton diagonals,
 * hess_sigma += |(i Omega) u g|^2 and hess_eps += |Omega^2 u g|^2 over every
 * pair of the nshots forward fields d_u and the nrec receiver Green's
 * functions d_g (1/DH^2 spikes, RHS_source_TE_hess.c:21-32), in FP64, added to
 * d_hess_sigma / d_hess_eps; omega, omega2 from gfdfd_omega(GFDFD_TE, ...);
 * the MAT*_NORM^2 scaling (:187-188) is the caller's.                        */
int gfdfd_hess_gn_te(int n, int nshots, int nrec, const float *omega, const float *omega2, const double *d_u,
                     const double *d_g, double *d_hess_sigma, double *d_hess_eps, void *stream);
/* hessian_shin_TE (src/TE/hessian_shin_TE.c:124-128): pseudo-Hessian
 * diagonals, hess_sigma += |(i Omega) u|^2 and hess_eps += |Omega^2 u|^2 over
 * the nshots columns of d_u in FP64, added to d_hess_sigma / d_hess_eps (n
 * doubles each); omega, omega2 from gfdfd_omega(GFDFD_TE, ...).  The
 * reference's MAT1_NORM^2 / MAT2_NORM^2 scaling of the sums (:147-148) is left
 * to the caller.  Cell q uses the solution at q+1; the last cell is unchanged. */
int gfdfd_hess_shin_te(int n, int nshots, const float *omega, const float *omega2, const double *d_u,
                       double *d_hess_sigma, double *d_hess_eps, void *stream);
/* ass_grad_TE (src/TE/ass_grad_TE.c:31-35): sigma and epsilon gradients */
int gfdfd_grad_te(int n, int nshots, const float *omega, const float *omega2, const double *d_u, const double *d_lam,
                  double *d_grad_sigma, double *d_grad_eps, void *stream);
/* ass_grad_TE with the shots' source-time functions (src/TE/ass_grad_TE.c:21,
 * 31-35): shot s weighted by wien_s/|wien_s| (d_stf as gfdfd_grad_ac_stf).  */
int gfdfd_grad_te_stf(int n, int nshots, const float *omega, const float *omega2, const double *d_u,
                      const double *d_lam, const float *d_stf, double *d_grad_sigma, double *d_grad_eps,
                      void *stream);
/* gfdfd_grad_te_stf with the per-shot taper, applied to both gradients as
 * src/TE/ass_grad_TE.c:40-49 does                                           */
int gfdfd_grad_te_taper(int n, int nshots, const float *omega, const float *omega2, const double *d_u,
                        const double *d_lam, const float *d_stf, const gfdfd_shot_taper *taper, double *d_grad_sigma,
                        double *d_grad_eps, void *stream);

/* ---- measurement ------------------------------------------------------------
 * Per-kernel-class timing with HIP events on the launch stream, plus the
 * algorithmic flops and HBM bytes of each launch derived from the plan.
 * Classes (one kernel symbol per precision each; piv_unswap two, fwd_front
 * and bwd_front one per tile width): assemble, front_assemble,
 * scatter_orig, small_lu, panel, panel_apply, trsm_u, gemm_update, diag_inv,
 * fwd_gather, fwd_leaf, fwd_small1, fwd_front1, fwd_front, fwd_blk,
 * bwd_small1, bwd_front1, bwd_front, bwd_contrib, bwd_blk, piv_unswap,
 * fwd_vec, bwd_vec, permute_in, permute_out, residual, norm_reduce, axpy,
 * subtree.
 * No counterpart in the reference (it times whole phases with MPI_Wtime,
 * src/AC/forward_shot_AC.c:54-65).                                          */
int gfdfd_profile(int enable);
/* Device memory the library holds (plans, factors, solve work buffers): the
 * bytes allocated now and the most since the last reset; reset_peak != 0
 * restarts the peak from the current value.  No counterpart in the reference
 * (UMFPACK reports host memory in Info[]).                                   */
int gfdfd_device_memory(long long *current_bytes, long long *peak_bytes, int reset_peak);
/* Free the solve work buffers the calling thread has cached (they grow to the
 * largest solve so far and are otherwise kept for the next call); waits for
 * the device first.  No counterpart in the reference.                        */
int gfdfd_release_workspace(void);
/* record only the named class (NULL or "": every class), e.g. to time one
 * kernel inside a timed region at the cost of two events per launch       */
int gfdfd_profile_select(const char *kernel);
void gfdfd_profile_reset(void);
int gfdfd_profile_read(const char *kernel, double *ms, long long *launches, double *flops, double *bytes);

#ifdef __cplusplus
}
#endif
#endif
