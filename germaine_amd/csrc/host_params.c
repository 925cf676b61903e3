/*
 * Host-side per-frequency parameters of the assembly, compiled with gcc so
 * that complex float arithmetic, libgcc complex division and glibc cpowf/cos
 * round exactly as in the reference's own host code:
 *   PML profiles            src/pml_pro.c:20-160
 *   FD weights, 1/DH^2      src/AC/init_A_AC_9p_pml.c:34-42 (same in TE, SH)
 *   Omega^2, Omega          src/AC/init_A_AC_9p_pml.c:50, src/SH/init_A_SH_9p_pml.c:50,
 *                           src/TE/init_A_TE_9p_pml.c:54-58
 *   point-source amplitude  src/AC/RHS_source_AC.c:37-53, src/TE/RHS_source_TE.c:22-56
 *   gradient Omega^2        src/AC/ass_grad_AC.c:22
 */
#include <complex.h>
#include <math.h>

typedef float complex fcx;

void gfdfd_fd_weights(float dh, float *out6)
{
    float a = 0.6287326, b = 0.3712667, c, d = 0.4382634, e;
    c = 1 - a - b;
    b = 0.25 * b;
    c = 0.25 * c;
    e = 1 - d;
    out6[0] = a; out6[1] = b; out6[2] = c; out6[3] = d; out6[4] = e;
    out6[5] = 1.0 / (dh * dh);
}

float gfdfd_mu0(void)
{
    float mu0 = 4.0 * M_PI * 1e-7;
    return mu0;
}

/* prof: x profiles (re, im, half re, half im; nx+2 each) then y profiles (ny+2 each) */
void gfdfd_pml_profiles(int nx, int ny, int npml, float dh, float a0_pml, float freq, float s,
                        int free_surf, float *prof)
{
    float *px[4], *py[4];
    for (int q = 0; q < 4; q++) {
        px[q] = prof + q * (nx + 2);
        py[q] = prof + 4 * (nx + 2) + q * (ny + 2);
    }
    for (int q = 0; q < 4; q++) {
        float v = (q & 1) ? 0.0 : 1.0;
        for (int i = 0; i <= nx + 1; i++) px[q][i] = v;
        for (int i = 0; i <= ny + 1; i++) py[q][i] = v;
    }
    float pih = M_PI / 2.0;
    fcx ci = 0.0 + 1.0 * I;
    float lpml = npml * dh;
    fcx om = 2.0 * M_PI * (freq - I * s);
    for (int i = 1; i <= npml; i++) {
        float x = i * dh, xh = x + 0.5 * dh;
        float eps = a0_pml * (1.0 - cos((lpml - x) * pih / lpml));
        float epsh = a0_pml * (1.0 - cos((lpml - xh) * pih / lpml));
        /* one-based point i and its mirror nx-i+1 / ny-i+1 carry the same value */
        float v[4];
        v[0] = creal(1.0 / (1.0 + eps / (ci * om)));
        v[1] = cimag(1.0 / (1.0 + eps / (ci * om)));
        v[2] = creal(1.0 / (1.0 + epsh / (ci * om)));
        v[3] = cimag(1.0 / (1.0 + epsh / (ci * om)));
        for (int q = 0; q < 4; q++) {
            px[q][i] = v[q];
            px[q][nx - i + 1] = v[q];
            py[q][ny - i + 1] = v[q];
            py[q][i] = free_surf == 1 ? ((q & 1) ? 0.0 : 1.0) : v[q];
        }
    }
    for (int q = 0; q < 4; q++) {
        px[q][0] = px[q][1];
        px[q][nx + 1] = px[q][nx];
        py[q][0] = py[q][1];
        py[q][ny + 1] = py[q][ny];
    }
}

/* physics 1 = AC (f - iS), 2 = SH and 3 = TE (f + iS) */
void gfdfd_omega(int physics, float freq, float s, float *omega2, float *omega)
{
    fcx w2, w;
    if (physics == 1) {
        w2 = cpowf(2.0 * M_PI * (freq - s * I), 2.0);
        w = 2.0 * M_PI * (freq - s * I);
    } else {
        w2 = cpowf(2.0 * M_PI * (freq + s * I), 2.0);
        w = 2.0 * M_PI * (freq + s * I);
    }
    omega2[0] = crealf(w2);
    omega2[1] = cimagf(w2);
    omega[0] = crealf(w);
    omega[1] = cimagf(w);
}

void gfdfd_omega2_grad_ac(float freq, float s, float *omega2)
{
    fcx w2 = cpowf(((2.0 * M_PI * freq) - (I * s)), 2.0);
    omega2[0] = crealf(w2);
    omega2[1] = cimagf(w2);
}

void gfdfd_source_amp(int physics, float freq, float s, float dh, double *amp)
{
    float re = 1.0, im = 0.0;
    if (physics == 3) {
        fcx omega = 2.0 * M_PI * (freq + s * I);
        float mu0 = 4.0 * M_PI * 1e-7;
        re = -creal((omega * mu0 * I) * (1.0 + 0.0 * I));
        im = -cimag((omega * mu0 * I) * (1.0 + 0.0 * I));
    }
    amp[0] = re / (dh * dh);
    amp[1] = im / (dh * dh);
}
