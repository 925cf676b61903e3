// Shot kernels (see shot.hip).
#pragma once
#include "common.hpp"

namespace gfdfd {

// ass_grad_AC/TE's per-shot taper (SWS_TAPER_CIRCULAR_PER_SHOT): mode 0 none,
// 1 taper_grad_shot, 2 taper_grad_shot1; device pointers
struct ShotTaperDev {
    int mode = 0, nx = 1;
    float dh = 1.0f, radius = 0.0f, a = 0.0f, f0 = 1e-3f;
    int filtsize = 0;
    const float *srcx = nullptr, *srcy = nullptr;   // mode 1: the panel's shots' sources (m)
    int nsrc = 0;
    const int *src_i = nullptr, *src_j = nullptr;   // mode 2: iround(srcpos / DH) of every source
    int nrec = 0;
    const int *rec_i = nullptr, *rec_j = nullptr;   // recpos (grid indices)
};

void spike_panel(int n, int nshots, const int *k, cx<double> amp, cx<double> *B, hipStream_t st);
void sample_receivers(int n, int nshots, int ntr, const int *krec, const cx<double> *X, cx<float> *pcal,
                      hipStream_t st);
void adjoint_sources(int n, int nshots, int ntr, const int *krec, const cx<float> *pcal, const cx<float> *pobs,
                     float dh, int misfit, int stf_inv, int rtm, cx<double> *B, double *l2, cx<float> *stf,
                     hipStream_t st);
void grad_ac(int n, int nshots, const float omega2[2], const cx<double> *U, const cx<double> *L, const float *vp,
             const float *rho, const cx<float> *stf, const ShotTaperDev &tp, double *grad, hipStream_t st);
void hess_shin_ac(int n, int nshots, const float omega2[2], const cx<double> *U, const float *vp, const float *rho,
                  double *hess, hipStream_t st);
void apply_hess(int n, float eps, const double *hess, double *grad, hipStream_t st);
void model_update(int n, float alpha, const double *dir, const float *m_old, float lo, float hi, float *m_new,
                  hipStream_t st);
void tikhonov_te(int nx, int ny, float dh, float a_sig, float a_eps, float beta_sig, float beta_eps, float lb_sig,
                 float lb_eps, const float *sig, const float *eps, double *g_sig, double *g_eps, double *cost,
                 hipStream_t st);
void taper_file(int nx, int ny, int nx0, int ny0, int npml, int fsshift, int free_surf, const float *taper,
                const double *in, double *out, hipStream_t st);
void grad_frame(int nx, int ny, int nx0, int ny0, int npml, int fsshift, int free_surf, float a, double *g,
                hipStream_t st);
void precond(int nx, int ny, int hfsx, int hfsy, float sx, float sy, int nzero, const double *in, double *out,
             hipStream_t st);
void model_update_norm(int n, float alpha, const double *dir, const float *m_old, float norm, float lo, float hi,
                       float *m_new, hipStream_t st);
void hess_gn_ac(int n, int nshots, int nrec, const float omega2[2], const cx<double> *U, const cx<double> *Gr,
                const float *vp, const float *rho, double *hess, hipStream_t st);
void hess_gn_te(int n, int nshots, int nrec, const float omega[2], const float omega2[2], const cx<double> *U,
                const cx<double> *Gr, double *hsig, double *heps, hipStream_t st);
void hess_shin_te(int n, int nshots, const float omega[2], const float omega2[2], const cx<double> *U, double *hsig,
                  double *heps, hipStream_t st);
void grad_te(int n, int nshots, const float omega[2], const float omega2[2], const cx<double> *U,
             const cx<double> *L, const cx<float> *stf, const ShotTaperDev &tp, double *gsig, double *geps,
             hipStream_t st);

}  // namespace gfdfd
