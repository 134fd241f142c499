// Dispersion-path types and launchers (Radon, picking, soft-argmax backward).
#pragma once

#include <hip/hip_runtime.h>

#include "swd_internal.h"

namespace swd {

constexpr int kMaxPair = 1024;  // frequency columns a pick kernel can hold in LDS
constexpr int kPickMaxNp = 2048;  // velocities of a column the pick kernel stages in LDS

// One (shot, side) spectrum column.
struct DispCol {
    int enabled;  // branch guard of SWD.m:175/181 and ADWDgrad_1.m:13/47
    int shot;     // index into the batch's gathers
    int tr0;      // 0-based trace of window position l = 0
    int dir;      // +1 right window, -1 left window (fliplr)
    int L;        // window length (dataLen)
};

struct GemmArgs {
    const float* A;
    const float* B;
    float* C;
    int M, N, K;
    long long sAb, sAm, sAk;
    long long sBb, sBk, sBn;
    long long sCb, sCm, sCn;
    float alpha;
};

struct SlantArgs {
    const float2* in;
    float2* out;
    int np, Lmax, ncol, npair, c0, m;
    double df, vmin, ds;
};

int launch_dft_table(hipStream_t st, float* W, int nt_eff, int npair, int c0, int ccn);
int launch_gemm_f32(hipStream_t st, const GemmArgs& g, int batch);

// Arithmetic of the Radon DFT / inverse-DFT GEMMs: 0 = FP32 MFMA (default),
// 1 = BF16 split precision (three bf16 MFMA products per FP32 product).
int radon_precision();
int set_radon_precision(int mode);
// a[shot][*] /= max |a[shot][*]| over each shot's n values (NaN skipped);
// part holds nshots x kPeakParts partial maxima
constexpr int kPeakParts = 64;
int launch_shot_peak_norm(hipStream_t st, float* a, size_t n, int nshots, float* part);
// peak[shot][g] = max_t |traces[shot][t][g]| (norm_trace's divisor) and
// par[shot][g][0/1] = even / odd-t sums (fk filter); either may be null
int launch_trace_stats(hipStream_t st, const float* traces, int nt, int ng, int nshots, float* peak, double* par);
// fk filter tables from W (nt rows): cotab (nt doubles, scratch), U' ([s][2 npair]
// like W, s < nt) and EO (even / odd-t sums of W, 2 x 2 npair doubles)
int launch_fk_tables(hipStream_t st, const float* W, int nt_eff, int nt, int npair, double* cotab, float* Up,
                     double* EO);
// window spectra of fk-filtered windows (X as launch_radon_gather writes it)
// from D = W^T traces, H = U'^T traces and the trace sums; peak may be null
int launch_fk_window(hipStream_t st, const float* D, const float* H, const double* par, const float* peak,
                     const double* EO, const DispCol* cols, int ncol, int npair, int ng, int Lmax, int nt, float2* X);
// peak may be null (no trace normalisation)
int launch_radon_gather(hipStream_t st, const float* D, const DispCol* cols, int ncol, int npair, int ng, int Lmax,
                        const float* peak, float2* X);
int launch_slant_stack(hipStream_t st, const SlantArgs& a, bool adjoint);
int launch_radon_abs(hipStream_t st, const float2* MM, int np, int npair, int ncol, float* ml, float* mlr);
// lp = 0: start at the column of the largest |mm| (mlr); the misfit
// averages columns r0 .. npair-1-r1
int launch_pick(hipStream_t st, const float* ml, const float* mlr, const DispCol* cols, int ncol, int np, int npair,
                int lp, float fg, float lfg, double vmin, int r0, int r1, const double* cr0, double* cr_pre,
                double* res, int* status);
int launch_argmax_pick(hipStream_t st, const float* ml, int np, int npair, double vmin, const double* cr0,
                       double* cr_pre);
// the same on FP64 spectra
int launch_pick(hipStream_t st, const double* ml, const double* mlr, const DispCol* cols, int ncol, int np, int npair,
                int lp, float fg, float lfg, double vmin, int r0, int r1, const double* cr0, double* cr_pre,
                double* res, int* status);
int launch_argmax_pick(hipStream_t st, const double* ml, int np, int npair, double vmin, const double* cr0,
                       double* cr_pre);
int launch_softargmax(hipStream_t st, const float* mlr, const float2* MM, const DispCol* cols, int np, int npair,
                      int ncol, double beta, const double* cr_pre, const double* cr0, float2* GMM);
int launch_adj_place(hipStream_t st, const float2* GD, const DispCol* cols, int ncol, int npair, int ng, int Lmax,
                     float* Bm);

}  // namespace swd
