// Device assembly entry (see assemble.hip).
#pragma once
#include "common.hpp"

namespace gfdfd {

// per-frequency scalars computed on the host (cpowf, Omega)
struct AsmFreq {
    float omega2_re, omega2_im, omega_re, omega_im;
};

// FD weights as the reference rounds them (src/AC/init_A_AC_9p_pml.c:34-42), mu0 for TE
struct AsmScalars {
    float a, b, c, d, e, idh2, mu0;
    const AsmFreq *freq;   // device array [nfreq]
};

// prof: device [nfreq][4(nx+2) + 4(ny+2)] PML profiles (x r,i,hr,hi then y r,i,hr,hi)
// m1/m2: AC vp/rho, SH vs/rho, TE eps/sigma, device float [ny][nx]
// out: nfreq x nnz FP64 complex values of M in canonical CSC order
void assemble_device(int physics, int nx, int ny, const AsmScalars &sc, const float *m1, const float *m2,
                     const float *prof, const int *colptr, cx<double> *out, int64_t nnz, int nfreq, hipStream_t st);

}  // namespace gfdfd
