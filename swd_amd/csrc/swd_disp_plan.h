// Batched dispersion plan: device buffers and the three stages of the SWD
// dispersion path for one batch of (shot, side) columns of one geometry.
#pragma once

#include <vector>

#include "swd_abi.h"
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
    DevBuf peak;       // [shot][g] trace peaks max_t |x| (norm_trace)
    DevBuf peak_part;  // [shot][kPeakParts] partial peaks of the adjoint sources
    // variants of the field-data scripts (all off: SWD.m)
    bool norm_traces = false;    // divide each trace by its peak before the Radon
    bool auto_start = false;     // pick from the column of the largest |mm|
    int res_first = 1, res_last = 0;  // misfit over cr(res_first : end - res_last), 1-based
    bool adj_peak_norm = false;  // divide each shot's adjoint source by its peak |a|
    bool fk = false;             // fk_filter each window before its spectrum
    DevBuf fkUp;     // [t][2 npair] Hilbert table U' of the fk filter
    DevBuf fkH;      // [shot][2 npair][g] U'^T traces
    DevBuf fkPar;    // [shot][g][2] even / odd-t trace sums (double)
    DevBuf fkEO;     // [2][2 npair] even / odd-t sums of W (double)
    DevBuf fkCot;    // [nt] cot(pi tau / 2nt) (double, table scratch)

    // allocate and fill the fk filter tables (once)
    int enable_fk();

    int init(const DispGeom& geom, const std::vector<DispCol>& columns, hipStream_t stream);
    int forward(const float* traces);             // traces [shot][t][g]
    int pick(int lp, float fg, float lfg);
    int adjoint(double beta, float* adj_out);     // adj_out [shot][t][g]
};

inline int set_plan_variant(DispPlan& d, const swd_variant& v)
{
    if (v.fk_filter && !d.fkUp.p) {
        const int rc = d.enable_fk();
        if (rc) return rc;
    }
    d.norm_traces = v.norm_trace != 0;
    d.auto_start = v.auto_start != 0;
    d.res_first = v.res_first > 1 ? v.res_first : 1;
    d.res_last = v.res_last > 0 ? v.res_last : 0;
    d.adj_peak_norm = v.adj_peak_norm != 0;
    d.fk = v.fk_filter != 0;
    return 0;
}

}  // namespace swd
