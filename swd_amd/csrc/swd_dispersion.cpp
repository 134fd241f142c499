// The dispersion path alone, batched over every (shot, side) column of a set
// of gathers: RTr/RTl + LHDispPick (SWD.m:173-185) and RTrAD/RTlAD +
// ADWDgrad_1 (SWD.m:190) on gathers the caller uploads, with the same
// DispPlan the engine runs after its forward FD.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "swd_engine.h"

namespace swd {

struct Dispersion {
    swd_problem p{};
    std::vector<int> is_index;
    Stream st;
    DispPlan disp;
    DevBuf gathers;  // [shot][t][g] float
    DevBuf adj;      // [shot][t][g] float
    bool have_gathers = false, have_picks = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;

    ~Dispersion()
    {
        if (st.s) (void)hipStreamSynchronize(st.s);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }

    int create(const swd_problem* prob)
    {
        if (!prob) { set_error("null problem"); return kErrArg; }
        p = *prob;
        if (p.nt < 2 || p.ng < 1 || p.nshots < 1) { set_error("bad gather sizes"); return kErrArg; }
        is_index.resize(p.nshots);
        for (int s = 0; s < p.nshots; ++s) is_index[s] = p.is_index ? p.is_index[s] : s + 1;
        p.sx = p.sz = p.is_index = nullptr;  // not kept
        int rc;
        if ((rc = st.create())) return rc;
        if ((rc = build_disp_plan(p, is_index, st.s, disp))) return rc;
        const size_t n = (size_t)p.nshots * p.nt * p.ng;
        if ((rc = gathers.alloc(n * sizeof(float))) || (rc = adj.alloc(n * sizeof(float)))) return rc;
        SWD_HIP(hipEventCreate(&e0));
        SWD_HIP(hipEventCreate(&e1));
        SWD_HIP(hipStreamSynchronize(st.s));
        return kOk;
    }

    float elapsed_ms()
    {
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) ms = 0;
        return ms;
    }

    int check_status()
    {
        const size_t S = p.nshots;
        std::vector<int> stat(2 * S);
        SWD_HIP(hipMemcpyAsync(stat.data(), disp.status.p, stat.size() * sizeof(int), hipMemcpyDeviceToHost, st.s));
        SWD_HIP(hipStreamSynchronize(st.s));
        for (size_t c = 0; c < 2 * S; ++c)
            if (disp.cols[c].enabled && stat[c]) {
                set_error("pick of shot %zu (%s side) kept fewer than two columns", c % S, c < S ? "right" : "left");
                return kErrPick;
            }
        return kOk;
    }
};

}  // namespace swd

using namespace swd;

struct swd_dispersion {
    Dispersion d;
};

extern "C" {

int swd_dispersion_create(const swd_problem* p, swd_dispersion** out)
{
    if (!out) { set_error("null output handle"); return kErrArg; }
    *out = nullptr;
    swd_dispersion* h = new (std::nothrow) swd_dispersion;
    if (!h) { set_error("out of host memory"); return kErrNoMem; }
    const int rc = h->d.create(p);
    if (rc) { delete h; return rc; }
    *out = h;
    return kOk;
}

int swd_dispersion_destroy(swd_dispersion* h)
{
    delete h;
    return kOk;
}

int swd_dispersion_info(swd_dispersion* h, long long* vals, int n)
{
    if (!h || !vals) return kErrArg;
    const DispGeom& g = h->d.disp.g;
    long long sumL = 0, ncol = 0;
    for (const auto& c : h->d.disp.cols)
        if (c.enabled) { sumL += c.L; ++ncol; }
    const long long v[8] = {g.npair, g.np, g.lf, g.nf, g.ccn, h->d.disp.Lmax, sumL, ncol};
    for (int i = 0; i < n && i < 8; ++i) vals[i] = v[i];
    return kOk;
}

int swd_dispersion_set_gathers(swd_dispersion* h, const float* gathers)
{
    if (!h || !gathers) { set_error("null argument"); return kErrArg; }
    Dispersion& d = h->d;
    SWD_HIP(hipMemcpyAsync(d.gathers.p, gathers, d.gathers.bytes, hipMemcpyHostToDevice, d.st.s));
    SWD_HIP(hipStreamSynchronize(d.st.s));
    d.have_gathers = true;
    d.have_picks = false;
    return kOk;
}

int swd_dispersion_set_variant(swd_dispersion* h, const swd_variant* v)
{
    if (!h || !v) { set_error("null argument"); return kErrArg; }
    Dispersion& d = h->d;
    if (v->res_first < 0 || v->res_last < 0 || (v->res_first > 1 ? v->res_first - 1 : 0) + v->res_last >= d.disp.g.npair) {
        set_error("bad misfit window cr(%d:end-%d) for %d columns", v->res_first, v->res_last, d.disp.g.npair);
        return kErrArg;
    }
    int rc;
    if ((rc = set_plan_variant(d.disp, *v))) return rc;
    d.have_picks = false;
    return kOk;
}

int swd_dispersion_set_observed_picks(swd_dispersion* h, const double* cr0_r, const double* cr0_l)
{
    if (!h) return kErrArg;
    Dispersion& d = h->d;
    const size_t n = (size_t)d.disp.g.npair * d.p.nshots;
    if (cr0_r) SWD_HIP(hipMemcpyAsync(d.disp.cr0.p, cr0_r, n * sizeof(double), hipMemcpyHostToDevice, d.st.s));
    if (cr0_l)
        SWD_HIP(hipMemcpyAsync((double*)d.disp.cr0.p + n, cr0_l, n * sizeof(double), hipMemcpyHostToDevice, d.st.s));
    SWD_HIP(hipStreamSynchronize(d.st.s));
    return kOk;
}

int swd_dispersion_pick(swd_dispersion* h, double* cr_r, double* cr_l, double* res_r, double* res_l, double* ms)
{
    if (!h) return kErrArg;
    Dispersion& d = h->d;
    if (!d.have_gathers) { set_error("gathers must be set first"); return kErrArg; }
    int rc;
    SWD_HIP(hipEventRecord(d.e0, d.st.s));
    if ((rc = d.disp.forward(d.gathers.as<float>())) || (rc = d.disp.pick(d.p.ini, 2.f, 4.f))) return rc;
    SWD_HIP(hipEventRecord(d.e1, d.st.s));
    const size_t S = d.p.nshots, n = (size_t)d.disp.g.npair * S;
    if (cr_r) SWD_HIP(hipMemcpyAsync(cr_r, d.disp.crpre.p, n * sizeof(double), hipMemcpyDeviceToHost, d.st.s));
    if (cr_l) SWD_HIP(hipMemcpyAsync(cr_l, (double*)d.disp.crpre.p + n, n * sizeof(double), hipMemcpyDeviceToHost, d.st.s));
    if (res_r) SWD_HIP(hipMemcpyAsync(res_r, d.disp.res.p, S * sizeof(double), hipMemcpyDeviceToHost, d.st.s));
    if (res_l) SWD_HIP(hipMemcpyAsync(res_l, (double*)d.disp.res.p + S, S * sizeof(double), hipMemcpyDeviceToHost, d.st.s));
    if ((rc = d.check_status())) return rc;
    if (ms) *ms = d.elapsed_ms();
    d.have_picks = true;
    return kOk;
}

int swd_dispersion_adjoint(swd_dispersion* h, double beta, float* adj, double* ms)
{
    if (!h) return kErrArg;
    Dispersion& d = h->d;
    if (!d.have_picks) { set_error("picks must be computed first"); return kErrArg; }
    int rc;
    SWD_HIP(hipEventRecord(d.e0, d.st.s));
    if ((rc = d.disp.adjoint(beta > 0 ? beta : d.p.beta, d.adj.as<float>()))) return rc;
    SWD_HIP(hipEventRecord(d.e1, d.st.s));
    if (adj) SWD_HIP(hipMemcpyAsync(adj, d.adj.p, d.adj.bytes, hipMemcpyDeviceToHost, d.st.s));
    SWD_HIP(hipStreamSynchronize(d.st.s));
    if (ms) *ms = d.elapsed_ms();
    return kOk;
}

}  // extern "C"
