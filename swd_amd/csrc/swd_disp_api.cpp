// Host side of the dispersion path: the batched DispPlan used by the engine,
// and the MATLAB-compatible single-shot entry points RTr/RTl(+AD),
// LHDispPick and ADWDgrad_1 of the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <vector>

#include "swd_abi.h"
#include "swd_disp.h"
#include "swd_host.h"
#include "swd_disp_plan.h"

namespace swd {

static std::atomic<int> g_radon_prec{0};
int radon_precision() { return g_radon_prec.load(); }
int set_radon_precision(int mode) { return g_radon_prec.exchange(mode == 1 ? 1 : 0); }

// MATLAB round (half away from zero) and fix(1./df./dt) (RTr.m:16,19-20)
static int mround(double x) { return (int)(x >= 0 ? std::floor(x + 0.5) : -std::floor(-x + 0.5)); }

int disp_geometry(DispGeom& g)
{
    if (!(g.df > 0) || !(g.dt > 0) || g.nt <= 0 || g.ng <= 0 || g.nshots <= 0) {
        set_error("bad Radon geometry"); return kErrArg;
    }
    g.ccn = (int)std::trunc((1.0 / g.df) / g.dt);
    g.lf = mround(g.fmin / g.df) + 1;
    g.nf = mround(g.fmax / g.df) + 1;
    g.npair = g.nf - g.lf + 1;
    g.c0 = g.lf - 1;
    g.np = (int)std::lround(g.vmax - g.vmin) + 1;
    g.nt_eff = std::min(g.nt, g.ccn);
    if (g.lf < 1 || g.npair < 1 || g.np < 1 || g.ccn < 1 || g.nf > g.ccn) { set_error("bad Radon frequency/velocity axes"); return kErrArg; }
    if (g.npair > kMaxPair) { set_error("npair %d exceeds %d", g.npair, kMaxPair); return kErrArg; }
    if (g.np > kPickMaxNp) { set_error("np %d exceeds %d", g.np, kPickMaxNp); return kErrArg; }
    if (!(g.vmin > 0)) { set_error("vmin must be positive"); return kErrArg; }
    return kOk;
}

// RTr.m:5-12 / RTl.m:5-12 windows -> column descriptor (0-based traces)
DispCol window_col(int side, int is, int b, int offset, int m, int M, int shot)
{
    DispCol c;
    c.shot = shot;
    c.enabled = 1;
    const int src = M * is - M + 1;
    if (side == 0) {
        int nend = m + offset - 1 + src;
        if (nend > b) nend = b;
        c.tr0 = src + m - 1;
        c.dir = 1;
        c.L = nend - (src + m) + 1;
    } else {
        int nend = src - offset - m;
        if (nend < 1) nend = 0;
        const int last = src - m;   // round() of an integer
        c.tr0 = last - 1;
        c.dir = -1;
        c.L = last - nend;
    }
    if (c.L <= 0 || c.tr0 < 0 || c.tr0 >= b || c.tr0 + c.dir * (c.L - 1) < 0 || c.tr0 + c.dir * (c.L - 1) >= b)
        c.enabled = 0;
    return c;
}

int DispPlan::init(const DispGeom& geom, const std::vector<DispCol>& columns, hipStream_t stream)
{
    g = geom;
    st = stream;
    cols = columns;
    ncol = (int)cols.size();
    Lmax = 1;
    for (const auto& c : cols) {
        if (c.enabled && (c.shot < 0 || c.shot >= g.nshots)) { set_error("column shot out of range"); return kErrArg; }
        if (c.enabled) Lmax = std::max(Lmax, c.L);
    }
    int rc;
    const size_t np = g.np, npair = g.npair, nc = ncol, L = Lmax;
    if ((rc = W.alloc((size_t)std::max(g.nt_eff, g.nt) * 2 * npair * sizeof(float)))) return rc;
    SWD_HIP(hipMemsetAsync(W.p, 0, W.bytes, st));
    if ((rc = launch_dft_table(st, W.as<float>(), std::max(g.nt_eff, g.nt), g.npair, g.c0, g.ccn))) return rc;
    if ((rc = D.alloc((size_t)g.nshots * 2 * npair * g.ng * sizeof(float)))) return rc;
    if ((rc = X.alloc(npair * L * nc * sizeof(float2)))) return rc;
    if ((rc = MM.alloc(npair * np * nc * sizeof(float2)))) return rc;
    if ((rc = ml.alloc(nc * npair * np * sizeof(float)))) return rc;
    if ((rc = mlr.alloc(nc * npair * np * sizeof(float)))) return rc;
    if ((rc = GMM.alloc(npair * np * nc * sizeof(float2)))) return rc;
    if ((rc = GD.alloc(npair * L * nc * sizeof(float2)))) return rc;
    if ((rc = Bm.alloc((size_t)g.nshots * g.ng * 2 * npair * sizeof(float)))) return rc;
    if ((rc = crpre.alloc(nc * npair * sizeof(double)))) return rc;
    if ((rc = cr0.alloc(nc * npair * sizeof(double)))) return rc;
    if ((rc = res.alloc(nc * sizeof(double)))) return rc;
    if ((rc = status.alloc(nc * sizeof(int)))) return rc;
    if ((rc = peak.alloc((size_t)g.nshots * g.ng * sizeof(float)))) return rc;
    if ((rc = peak_part.alloc((size_t)g.nshots * kPeakParts * sizeof(float)))) return rc;
    SWD_HIP(hipMemsetAsync(cr0.p, 0, cr0.bytes, st));
    SWD_HIP(hipMemsetAsync(res.p, 0, res.bytes, st));
    SWD_HIP(hipMemsetAsync(status.p, 0, status.bytes, st));
    if ((rc = dcols.upload(cols.data(), cols.size() * sizeof(DispCol), st))) return rc;
    return kOk;
}

// R1-R3: DFT at the bins, window gather, slant stack, |mm| and normalisation.
// With norm_traces the gathers are first divided trace by trace by their peak
// |x| (norm_trace on the predicted data, fieldexamples/outcrop/outcrop.m:164):
// the peaks come from the traces, the division is applied to the spectra.
int DispPlan::forward(const float* traces)
{
    int rc;
    if ((norm_traces || fk) &&
        (rc = launch_trace_stats(st, traces, g.nt, g.ng, g.nshots, norm_traces ? peak.as<float>() : nullptr,
                                 fk ? fkPar.as<double>() : nullptr)))
        return rc;
    GemmArgs a;
    memset(&a, 0, sizeof a);
    a.A = W.as<float>();                 // A(m = k2, k = t) = W[t][k2]
    a.sAm = 1; a.sAk = 2 * g.npair; a.sAb = 0;
    a.B = traces;                        // B(k = t, n = g) = traces[shot][t][g]
    a.sBk = g.ng; a.sBn = 1; a.sBb = (long long)g.nt * g.ng;
    a.C = D.as<float>();                 // C[shot][k2][g]
    a.sCm = g.ng; a.sCn = 1; a.sCb = 2LL * g.npair * g.ng;
    a.M = 2 * g.npair; a.N = g.ng; a.K = g.nt_eff; a.alpha = 1.f;
    if ((rc = launch_gemm_f32(st, a, g.nshots))) return rc;
    if (fk) {
        // H = U'^T traces over all nt samples, then the windows' filtered spectra
        a.A = fkUp.as<float>();
        a.C = fkH.as<float>();
        a.K = g.nt;
        if ((rc = launch_gemm_f32(st, a, g.nshots))) return rc;
        if ((rc = launch_fk_window(st, D.as<float>(), fkH.as<float>(), fkPar.as<double>(),
                                   norm_traces ? peak.as<float>() : nullptr, fkEO.as<double>(), dcols.as<DispCol>(),
                                   ncol, g.npair, g.ng, Lmax, g.nt, X.as<float2>())))
            return rc;
    } else if ((rc = launch_radon_gather(st, D.as<float>(), dcols.as<DispCol>(), ncol, g.npair, g.ng, Lmax,
                                         norm_traces ? peak.as<float>() : nullptr, X.as<float2>()))) {
        return rc;
    }
    SlantArgs s;
    s.in = X.as<float2>(); s.out = MM.as<float2>();
    s.np = g.np; s.Lmax = Lmax; s.ncol = ncol; s.npair = g.npair; s.c0 = g.c0; s.m = g.m;
    s.df = g.df; s.vmin = g.vmin; s.ds = g.ds;
    if ((rc = launch_slant_stack(st, s, false))) return rc;
    return launch_radon_abs(st, MM.as<float2>(), g.np, g.npair, ncol, ml.as<float>(), mlr.as<float>());
}

int DispPlan::enable_fk()
{
    int rc;
    const size_t npair2 = 2 * (size_t)g.npair;
    if ((rc = fkUp.alloc((size_t)g.nt * npair2 * sizeof(float))) ||
        (rc = fkH.alloc((size_t)g.nshots * npair2 * g.ng * sizeof(float))) ||
        (rc = fkPar.alloc((size_t)g.nshots * g.ng * 2 * sizeof(double))) ||
        (rc = fkEO.alloc(2 * npair2 * sizeof(double))) || (rc = fkCot.alloc((size_t)g.nt * sizeof(double))))
        return rc;
    if ((rc = launch_fk_tables(st, W.as<float>(), g.nt_eff, g.nt, g.npair, fkCot.as<double>(), fkUp.as<float>(),
                               fkEO.as<double>())))
        return rc;
    SWD_HIP(hipStreamSynchronize(st));
    return kOk;
}

int DispPlan::pick(int lp, float fg, float lfg)
{
    if (auto_start) lp = 0;
    else if (lp < 1 || lp > g.npair) { set_error("pick start column %d outside 1..%d", lp, g.npair); return kErrArg; }
    const int r0 = std::max(res_first, 1) - 1, r1 = std::max(res_last, 0);
    return launch_pick(st, ml.as<float>(), mlr.as<float>(), dcols.as<DispCol>(), ncol, g.np, g.npair, lp, fg, lfg,
                       g.vmin, r0, r1, cr0.as<double>(), crpre.as<double>(), res.as<double>(), status.as<int>());
}

// G1-G4: soft-argmax backward + phase, adjoint slant stack, placement, inverse DFT
int DispPlan::adjoint(double beta, float* adj_out)
{
    if (g.ccn < g.nt) { set_error("ccn (%d) < nt (%d): ADWDgrad_1 indexes past the inverse FFT", g.ccn, g.nt); return kErrArg; }
    int rc;
    if ((rc = launch_softargmax(st, mlr.as<float>(), MM.as<float2>(), dcols.as<DispCol>(), g.np, g.npair, ncol, beta,
                                crpre.as<double>(), cr0.as<double>(), GMM.as<float2>())))
        return rc;
    SlantArgs s;
    s.in = GMM.as<float2>(); s.out = GD.as<float2>();
    s.np = g.np; s.Lmax = Lmax; s.ncol = ncol; s.npair = g.npair; s.c0 = g.c0; s.m = g.m;
    s.df = g.df; s.vmin = g.vmin; s.ds = g.ds;
    if ((rc = launch_slant_stack(st, s, true))) return rc;
    SWD_HIP(hipMemsetAsync(Bm.p, 0, Bm.bytes, st));
    if ((rc = launch_adj_place(st, GD.as<float2>(), dcols.as<DispCol>(), ncol, g.npair, g.ng, Lmax, Bm.as<float>())))
        return rc;
    GemmArgs a;
    memset(&a, 0, sizeof a);
    a.A = W.as<float>();                 // A(m = t, k = k2) = W[t][k2]
    a.sAm = 2 * g.npair; a.sAk = 1; a.sAb = 0;
    a.B = Bm.as<float>();                // B(k = k2, n = g) = Bm[shot][g][k2]
    a.sBk = 1; a.sBn = 2 * g.npair; a.sBb = 2LL * g.npair * g.ng;
    a.C = adj_out;                       // C[shot][t][g]
    a.sCm = g.ng; a.sCn = 1; a.sCb = (long long)g.nt * g.ng;
    a.M = g.nt; a.N = g.ng; a.K = 2 * g.npair; a.alpha = (float)(1.0 / g.ccn);
    if ((rc = launch_gemm_f32(st, a, g.nshots))) return rc;
    // a_data_res./max(abs(a_data_res(:))) of the field scripts' ADWDgrad_1 (line 79)
    if (adj_peak_norm) return launch_shot_peak_norm(st, adj_out, (size_t)g.nt * g.ng, g.nshots, peak_part.as<float>());
    return kOk;
}

// ---------------------------------------------------------------------------
// single-shot MATLAB-compatible entry points
// ---------------------------------------------------------------------------
template <typename T>
static int radon_impl(int side, const T* seismo, int nt, int ng, int is, double df, double dt, int np, double vmin,
                      double vmax, double fmin, double fmax, double ds, int offset, int m, int M, float* ml_out,
                      int* dataLen, float* mlr_out, float* mm_out, int* info)
{
    if (!seismo || !ml_out || (side != 0 && side != 1)) { set_error("bad Radon arguments"); return kErrArg; }
    DispGeom g;
    memset(&g, 0, sizeof g);
    g.nt = nt; g.ng = ng; g.nshots = 1; g.df = df; g.dt = dt; g.fmin = fmin; g.fmax = fmax;
    g.vmin = vmin; g.vmax = vmax; g.ds = ds; g.offset = offset; g.m = m; g.M = M;
    int rc = disp_geometry(g);
    if (rc) return rc;
    if (g.np != np) { set_error("np (%d) must equal vmax-vmin+1 (%d)", np, g.np); return kErrArg; }
    DispCol c = window_col(side, is, ng, offset, m, M, 0);
    if (!c.enabled) { set_error("empty or out-of-range Radon window"); return kErrArg; }
    Stream st;
    if ((rc = st.create())) return rc;
    DispPlan plan;
    if ((rc = plan.init(g, {c}, st.s))) return rc;
    DevBuf tr;
    {
        std::vector<float> h((size_t)nt * ng);
        for (int t = 0; t < nt; ++t)
            for (int k = 0; k < ng; ++k) h[(size_t)t * ng + k] = (float)seismo[(size_t)k * nt + t];
        if ((rc = tr.upload(h.data(), h.size() * sizeof(float), st.s))) return rc;
        SWD_HIP(hipStreamSynchronize(st.s));
    }
    if ((rc = plan.forward(tr.as<float>()))) return rc;
    const size_t n = (size_t)g.np * g.npair;
    SWD_HIP(hipMemcpyAsync(ml_out, plan.ml.p, n * sizeof(float), hipMemcpyDeviceToHost, st.s));
    if (mlr_out) SWD_HIP(hipMemcpyAsync(mlr_out, plan.mlr.p, n * sizeof(float), hipMemcpyDeviceToHost, st.s));
    if (mm_out) SWD_HIP(hipMemcpyAsync(mm_out, plan.MM.p, n * sizeof(float2), hipMemcpyDeviceToHost, st.s));
    SWD_HIP(hipStreamSynchronize(st.s));
    if (dataLen) *dataLen = c.L;
    if (info) {
        info[0] = g.lf; info[1] = g.nf; info[2] = g.ccn; info[3] = c.tr0 + 1;
    }
    return kOk;
}

template <typename V>
static int lhdisppick_impl(const V* ml, int np, int npair, double vmin, const double* cr_0, int method, int ini,
                           double* res, double* cr_pre)
{
    if (!ml || !cr_0 || !cr_pre || np < 1 || npair < 1) { set_error("bad LHDispPick arguments"); return kErrArg; }
    if (npair > kMaxPair) { set_error("npair %d exceeds %d", npair, kMaxPair); return kErrArg; }
    if (np > kPickMaxNp) { set_error("np %d exceeds %d", np, kPickMaxNp); return kErrArg; }
    int rc;
    Stream st;
    if ((rc = st.create())) return rc;
    DevBuf dml, dcr0, dpre, dres, dstat, dcols;
    DispCol c{1, 0, 0, 1, np};
    if ((rc = dml.upload(ml, (size_t)np * npair * sizeof(V), st.s)) ||
        (rc = dcr0.upload(cr_0, (size_t)npair * sizeof(double), st.s)) || (rc = dpre.alloc(npair * sizeof(double))) ||
        (rc = dres.alloc(sizeof(double))) || (rc = dstat.alloc(sizeof(int))) ||
        (rc = dcols.upload(&c, sizeof c, st.s)))
        return rc;
    if (method == 1) {
        if (ini < 1 || ini > npair) { set_error("ini %d outside 1..%d", ini, npair); return kErrArg; }
        if ((rc = launch_pick(st.s, dml.as<V>(), (const V*)nullptr, dcols.as<DispCol>(), 1, np, npair, ini, 2.f, 4.f,
                              vmin, 0, 0, dcr0.as<double>(), dpre.as<double>(), dres.as<double>(), dstat.as<int>())))
            return rc;
    } else {
        if ((rc = launch_argmax_pick(st.s, dml.as<V>(), np, npair, vmin, dcr0.as<double>(), dpre.as<double>())))
            return rc;
    }
    int stv = 0;
    SWD_HIP(hipMemcpyAsync(cr_pre, dpre.p, npair * sizeof(double), hipMemcpyDeviceToHost, st.s));
    if (method == 1) SWD_HIP(hipMemcpyAsync(&stv, dstat.p, sizeof(int), hipMemcpyDeviceToHost, st.s));
    SWD_HIP(hipStreamSynchronize(st.s));
    if (stv) { set_error("interp1 needs at least two kept pick columns"); return kErrPick; }
    double acc = 0;
    for (int i = 0; i < npair; ++i) { const double d = cr_pre[i] - cr_0[i]; acc += d * d; }
    if (res) *res = acc / npair;
    return kOk;
}

static int adwdgrad1_impl(int nt, int ng, int ns, int npair, int is, double w, int m, int M, double beta,
                          const double* gor, const double* gol, const float* mmr, const float* mlr, int posr, int Lr,
                          const float* mml, const float* mll, int posl, int Ll, int np, double vmin, double df,
                          double ds, int lf, int ccn, float* a_data_res)
{
    if (!a_data_res || nt <= 0 || ng <= 0 || npair <= 0 || np <= 0 || M == 0) { set_error("bad ADWDgrad_1 arguments"); return kErrArg; }
    const bool right = is <= ns - (int)std::floor((double)m / M) - mround(w / M);
    const bool left = is >= mround(w / M) + (int)std::floor((double)m / M) + 1;
    if (right && (!gor || !mmr || !mlr)) { set_error("right-side inputs missing"); return kErrArg; }
    if (left && (!gol || !mml || !mll)) { set_error("left-side inputs missing"); return kErrArg; }
    DispGeom g;
    memset(&g, 0, sizeof g);
    g.nt = nt; g.ng = ng; g.nshots = 1; g.df = df; g.np = np; g.vmin = vmin; g.vmax = vmin + np - 1;
    g.ds = ds; g.m = m; g.M = M; g.lf = lf; g.npair = npair; g.nf = lf + npair - 1; g.c0 = lf - 1; g.ccn = ccn;
    g.nt_eff = std::min(nt, ccn);
    if (ccn < nt) { set_error("ccn (%d) < nt (%d): index exceeds the inverse FFT length", ccn, nt); return kErrArg; }
    if (npair > kMaxPair || lf < 1) { set_error("bad frequency axis"); return kErrArg; }
    DispCol cr{right ? 1 : 0, 0, posr - 1, 1, Lr};
    DispCol cl{left ? 1 : 0, 0, posl - 1, -1, Ll};
    if (right && (Lr < 1 || posr < 1 || posr - 1 + Lr - 1 >= ng)) { set_error("right window outside the gather"); return kErrArg; }
    if (left && (Ll < 1 || posl > ng || posl - 1 - (Ll - 1) < 0)) { set_error("left window outside the gather"); return kErrArg; }
    int rc;
    Stream st;
    if ((rc = st.create())) return rc;
    DispPlan plan;
    if ((rc = plan.init(g, {cr, cl}, st.s))) return rc;
    // MM[k][j][col], mlr[col][k][j], cr_pre = grad_output, cr0 = 0
    const size_t n = (size_t)np * npair;
    std::vector<float2> hmm(2 * n, make_float2(0.f, 0.f));
    std::vector<float> hml(2 * n, 0.f);
    std::vector<double> hpre(2 * (size_t)npair, 0.0);
    for (int s = 0; s < 2; ++s) {
        const float* mmS = s == 0 ? mmr : mml;
        const float* mlS = s == 0 ? mlr : mll;
        const double* gS = s == 0 ? gor : gol;
        if (!(s == 0 ? right : left)) continue;
        for (int k = 0; k < npair; ++k)
            for (int j = 0; j < np; ++j) {
                const size_t src = (size_t)k * np + j;
                hmm[src * 2 + s] = make_float2(mmS[2 * src], mmS[2 * src + 1]);
                hml[(size_t)s * n + src] = mlS[src];
            }
        for (int k = 0; k < npair; ++k) hpre[(size_t)s * npair + k] = gS[k];
    }
    SWD_HIP(hipMemcpyAsync(plan.MM.p, hmm.data(), hmm.size() * sizeof(float2), hipMemcpyHostToDevice, st.s));
    SWD_HIP(hipMemcpyAsync(plan.mlr.p, hml.data(), hml.size() * sizeof(float), hipMemcpyHostToDevice, st.s));
    SWD_HIP(hipMemcpyAsync(plan.crpre.p, hpre.data(), hpre.size() * sizeof(double), hipMemcpyHostToDevice, st.s));
    DevBuf dadj;
    if ((rc = dadj.alloc((size_t)nt * ng * sizeof(float)))) return rc;
    if ((rc = plan.adjoint(beta, dadj.as<float>()))) return rc;
    std::vector<float> h((size_t)nt * ng);
    SWD_HIP(hipMemcpyAsync(h.data(), dadj.p, h.size() * sizeof(float), hipMemcpyDeviceToHost, st.s));
    SWD_HIP(hipStreamSynchronize(st.s));
    for (int t = 0; t < nt; ++t)
        for (int k = 0; k < ng; ++k) a_data_res[(size_t)k * nt + t] = h[(size_t)t * ng + k];
    return kOk;
}

}  // namespace swd

using namespace swd;

extern "C" {

int swd_set_radon_precision(int mode)
{
    if (mode != 0 && mode != 1) { set_error("Radon precision must be 0 (FP32) or 1 (BF16 x 3)"); return kErrArg; }
    return set_radon_precision(mode);
}


int swd_radon(int side, const double* seismo, int nt, int ng, int is, double df, double dt, int np, double vmin,
              double vmax, double fmin, double fmax, double ds, int offset, int m, int M, float* ml, int* dataLen,
              float* mlr, float* mm, int* info)
{
    return radon_impl<double>(side, seismo, nt, ng, is, df, dt, np, vmin, vmax, fmin, fmax, ds, offset, m, M, ml,
                              dataLen, mlr, mm, info);
}

int swd_radon_f32(int side, const float* seismo, int nt, int ng, int is, double df, double dt, int np, double vmin,
                  double vmax, double fmin, double fmax, double ds, int offset, int m, int M, float* ml, int* dataLen,
                  float* mlr, float* mm, int* info)
{
    return radon_impl<float>(side, seismo, nt, ng, is, df, dt, np, vmin, vmax, fmin, fmax, ds, offset, m, M, ml,
                             dataLen, mlr, mm, info);
}

int swd_lhdisppick(const float* ml, int np, int npair, double vmin, const double* cr_0, int method, int ini,
                   double* res, double* cr_pre)
{
    return lhdisppick_impl<float>(ml, np, npair, vmin, cr_0, method, ini, res, cr_pre);
}

int swd_lhdisppick_f64(const double* ml, int np, int npair, double vmin, const double* cr_0, int method, int ini,
                       double* res, double* cr_pre)
{
    return lhdisppick_impl<double>(ml, np, npair, vmin, cr_0, method, ini, res, cr_pre);
}

int swd_adwdgrad1(int nt, int ng, int ns, int npair, int is, double w, int m, int M, double SoftArgNorm,
                  const double* grad_outputr, const double* grad_outputl, const float* mmr, const float* mlr,
                  int posr, int Lr, const float* mml, const float* mll, int posl, int Ll, int np, double vmin,
                  double df, double ds, int lf, int ccn, float* a_data_res)
{
    return adwdgrad1_impl(nt, ng, ns, npair, is, w, m, M, SoftArgNorm, grad_outputr, grad_outputl, mmr, mlr, posr,
                          Lr, mml, mll, posl, Ll, np, vmin, df, ds, lf, ccn, a_data_res);
}

}  // extern "C"
