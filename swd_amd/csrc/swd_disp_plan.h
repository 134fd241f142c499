// Batched dispersion plan: device buffers and the three stages of the SWD
// dispersion path for one batch of (shot, side) columns of one geometry.
#pragma once

#include <vector>

#include "swd_disp.h"
#include "swd_host.h"

namespace swd {

struct DispGeom {
    int nt, ng, nshots;
    double df, dt, fmin, fmax, vmin, vmax, ds;
    int offset, m, M;
    // derived by disp_geometry()
    int np, lf, nf, npair, ccn, c0, nt_eff;
};

int disp_geometry(DispGeom& g);
DispCol window_col(int side, int is, int b, int offset, int m, int M, int shot);

struct DispPlan {
    DispGeom g;
    hipStream_t st = nullptr;
    std::vector<DispCol> cols;
    int ncol = 0, Lmax = 1;
    DevBuf W;       // [nt][2 npair] twiddles
    DevBuf D;       // [shot][2 npair][ng] trace spectra at the bins
    DevBuf X;       // [npair][Lmax][ncol] window-gathered spectra (complex)
    DevBuf MM;      // [npair][np][ncol] slant stack mm (complex)
    DevBuf ml;      // [ncol][npair][np] normalised |mm|
    DevBuf mlr;     // [ncol][npair][np] |mm|
    DevBuf GMM;     // [npair][np][ncol] d misfit / d mm (complex)
    DevBuf GD;      // [npair][Lmax][ncol] adjoint slant stack (complex)
    DevBuf Bm;      // [shot][ng][2 npair] inverse-DFT operand
    DevBuf crpre;   // [ncol][npair] picks (velocity)
    DevBuf cr0;     // [ncol][npair] observed picks
    DevBuf res;     // [ncol] mean squared pick misfit
    DevBuf status;  // [ncol] pick status
    DevBuf dcols;

    int init(const DispGeom& geom, const std::vector<DispCol>& columns, hipStream_t stream);
    int forward(const float* traces);             // traces [shot][t][g]
    int pick(int lp, float fg, float lfg);
    int adjoint(double beta, float* adj_out);     // adj_out [shot][t][g]
};

}  // namespace swd
