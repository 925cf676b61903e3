/*
 * Host-side per-frequency parameters of the assembly, compiled with gcc so
 * that complex float arithmetic, libgcc complex division and glibc cpowf/cos
 * round exactly as in the reference's own host code:
 *   PML profiles            src/pml_pro.c:20-160
 *   FD weights, 1/DH^2      src/AC/init_A_AC_9p_pml.c:34-42 (same in TE, SH)
 *   Omega^2, Omega          src/AC/init_A_AC_9p_pml.c:50, src/SH/init_A_SH_9p_pml.c:50,
 *                           src/TE/init_A_TE_9p_pml.c:54-58
 *   point-source amplitude  src/AC/RHS_source_AC.c:37-53, src/TE/RHS_source_TE.c:22-56
 *   wavelet source amplitude src/read_stf_dft.c:34-70 with RHS_source_AC.c:41-53
 *   per-shot taper exponent src/taper_grad_shot.c:37-40
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

/* READ_WAVELET == 1: the source amplitude is the DFT of the time-domain
 * wavelet at the frequency (src/read_stf_dft.c:64-70), divided by DH^2 as the
 * spike is (src/AC/RHS_source_AC.c:52-53, src/TE/RHS_source_TE.c:55-56; the
 * DFT replaces the TE spike's -i omega mu0 factor).  samples: the ns values
 * of the SU trace; ndt: its header's sample interval, converted as the
 * reference does (dt = ndt / 1e6 for AC/SH, ndt / 1e12 for TE).  The reference reads the ns samples into data[0..ns-1] but
 * sums data[1..ns] at t = (j-1) dt, j = 1..ns: sample k >= 1 enters at time
 * (k-1) dt, sample 0 is skipped, and the last term takes data[ns], an element
 * the read never filled (undefined contents); here that term is zero.  The
 * sum is rounded to float after every term (amp[] is float).               */
void gfdfd_wavelet_amp(int physics, float freq, float dh, const float *samples, int ns, int ndt, double *amp)
{
    float dt;
    if (physics == 3) dt = ndt / 1e12;
    else dt = ndt / 1e6;
    float re = 0.0, im = 0.0;
    for (int j = 1; j <= ns; j++) {
        float t = (j - 1) * dt;
        float w = j < ns ? samples[j] : 0.0f;
        re += w * cos(2.0 * t * freq * M_PI) * dt;
        im += w * sin(2.0 * t * freq * M_PI) * dt;
    }
    amp[0] = re / (dh * dh);
    amp[1] = im / (dh * dh);
}

/* taper_grad_shot's exponent a = -(log(frh - f0) - log(1 - f0)) / log(2) with
 * frh = 0.1, f0 = 1e-3 in float (src/taper_grad_shot.c:36-39): the taper
 * f0 + (1 - f0) (0.5 - 0.5 cos(pi r / R))^a is 0.1 at r = R/2.            */
float gfdfd_shot_taper_exponent(void)
{
    float frh = 0.1, f0 = 1e-3, a;
    a = -(log(frh - f0) - log(1.0 - f0)) / log(2.0);
    return a;
}
