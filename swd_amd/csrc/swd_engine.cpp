// Batched SWD engine (see swd_engine.h) and its C ABI.
//
// One gradient is the parfor body of SWD.m:171-194 for every shot, then the
// reduction and assembly of SWD.m:193,199-201:
//   forward FD (+ strain snapshots)       all shots, one launch per time step
//   RTrAD/RTlAD -> LHDispPick            all (shot, side) columns batched
//   ADWDgrad_1                           all columns batched
//   adjoint FD + imaging                 all shots, one launch per time step
//   image reduction + dk_vs              one kernel
#include "swd_engine.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <thread>

namespace swd {

Engine::~Engine()
{
    if (st) (void)hipStreamSynchronize(st);
    for (auto& b : batch_fd)
        if (b) {
            (void)hipStreamSynchronize(b);
            (void)hipStreamDestroy(b);
        }
    if (own_stream && st) (void)hipStreamDestroy(st);
}

int problem_geometry(const swd_problem& p, DispGeom& g)
{
    memset(&g, 0, sizeof g);
    g.nt = p.nt; g.ng = p.ng; g.nshots = p.nshots; g.df = p.df; g.dt = p.dt; g.fmin = p.fmin; g.fmax = p.fmax;
    g.vmin = p.vmin; g.vmax = p.vmax; g.ds = p.ds; g.offset = p.offset; g.m = p.m; g.M = p.M;
    return disp_geometry(g);
}

int build_disp_plan(const swd_problem& p, const std::vector<int>& is_index, hipStream_t st, DispPlan& disp)
{
    DispGeom g;
    int rc = problem_geometry(p, g);
    if (rc) return rc;
    if (p.M == 0) { set_error("M must be non-zero"); return kErrArg; }
    if (p.ini < 1 || p.ini > g.npair) { set_error("ini %d outside 1..%d", p.ini, g.npair); return kErrArg; }
    if ((int)is_index.size() != p.nshots) { set_error("is_index has %zu entries for %d shots", is_index.size(), p.nshots); return kErrArg; }
    // branch guards of SWD.m:175,181 (right: is <= ns - floor(m/M) - round(w/M); left: is >= round(w/M) + floor(m/M) + 1)
    const int ns = p.ns > 0 ? p.ns : p.nshots;
    auto mr = [](double x) { return (int)(x >= 0 ? std::floor(x + 0.5) : -std::floor(-x + 0.5)); };
    std::vector<DispCol> cols(2 * p.nshots);
    for (int side = 0; side < 2; ++side)
        for (int s = 0; s < p.nshots; ++s) {
            const int is = is_index[s];
            DispCol c = window_col(side, is, p.ng, p.offset, p.m, p.M, s);
            const bool guard = side == 0 ? is <= ns - (int)std::floor((double)p.m / p.M) - mr(p.w / p.M)
                                         : is >= mr(p.w / p.M) + (int)std::floor((double)p.m / p.M) + 1;
            if (!guard) c.enabled = 0;
            cols[side * p.nshots + s] = c;
        }
    return disp.init(g, cols, st);
}

int Engine::create(const swd_problem* prob, hipStream_t stream)
{
    if (!prob) { set_error("null problem"); return kErrArg; }
    p = *prob;
    if (p.nz < 1 || p.nx < 1 || p.nbc < 2 || p.nt < 2 || p.in_wf < 1 || p.nt % p.in_wf != 0 || p.nshots < 1 ||
        p.ng < 1 || p.dg < 1 || !(p.dx > 0) || !(p.dt > 0)) {
        set_error("bad problem sizes (need nz,nx>=1, nbc>=2, nt%%in_wf==0, shots and receivers)");
        return kErrArg;
    }
    if (p.fd_order != 22) { set_error("the batched engine runs fd_order 22 (the SWD configuration)"); return kErrArg; }
    if (!p.sx || !p.sz) { set_error("null source positions"); return kErrArg; }
    pt_m = (p.fd_order - 20) / 2 + 1;
    nzbc = p.nz + pt_m + p.nbc;
    nxbc = p.nx + 2 * p.nbc;
    pitch = grid_pitch(nxbc);
    plane = (size_t)nzbc * pitch;
    nrec = p.nt / p.in_wf;
    splane = (size_t)p.nz * p.nx;
    sx.assign(p.sx, p.sx + p.nshots);
    sz.assign(p.sz, p.sz + p.nshots);
    is_index.resize(p.nshots);
    for (int s = 0; s < p.nshots; ++s) is_index[s] = p.is_index ? p.is_index[s] : s + 1;
    const int pt = 1;  // C-side pad_top of fd_order 22
    for (int s = 0; s < p.nshots; ++s) {
        const int r = sz[s] + pt_m - 1, c = sx[s] + p.nbc - 1;
        if (r < pt || r > nzbc - 1 - pt || c < pt || c > nxbc - 1 - pt) {
            set_error("shot %d lies outside the update block", s);
            return kErrArg;
        }
    }
    {
        const int r = p.gz + pt_m - 1, c0 = p.gx0 + p.nbc - 1, c1 = c0 + (p.ng - 1) * p.dg;
        if (r < pt || r > nzbc - 1 - pt || c0 < pt || c1 > nxbc - 1 - pt) {
            set_error("receivers lie outside the update block");
            return kErrArg;
        }
    }
    if (stream) st = stream;
    else {
        SWD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        own_stream = true;
    }
    tz = fd_tile_rows();
    int rc = build_disp_plan(p, is_index, st, disp);
    if (rc) return rc;
    npair = disp.g.npair;
    np = disp.g.np;
    // device buffers
    const size_t S = p.nshots, ntr = (size_t)p.nt * p.ng;
    if ((rc = coefs.alloc(10 * plane * sizeof(float))) || (rc = vs_d.alloc(splane * sizeof(double))) ||
        (rc = wavelet.alloc(p.nt * sizeof(float))) || (rc = state.alloc(S * 10 * plane * sizeof(float))) ||
        (rc = traces.alloc(S * ntr * sizeof(float))) ||
        (rc = snaps.alloc(S * (size_t)nrec * 3 * splane * sizeof(float))) ||
        (rc = adjsrc.alloc(S * ntr * sizeof(float))) || (rc = imgs.alloc(S * 3 * splane * sizeof(double))) ||
        (rc = imgp.alloc(S * 3 * splane * sizeof(float))) ||
        (rc = outs.alloc(4 * splane * sizeof(double))) || (rc = vp_d.alloc(splane * sizeof(double))) ||
        (rc = geo.alloc((size_t)nzbc * nxbc * sizeof(double))) || (rc = vs_cur.alloc(splane * sizeof(double))) ||
        (rc = vs_tr.alloc(splane * sizeof(double))) || (rc = dk_sm.alloc(splane * sizeof(double))) ||
        (rc = stats.alloc(4 * sizeof(double))))
        return rc;
    SWD_HIP(hipMemsetAsync(coefs.p, 0, coefs.bytes, st));
    {
        std::vector<double> g((size_t)nzbc * nxbc);
        damp_geometry(p.nz, p.nx, p.nbc, p.dx, pt_m, g.data());
        SWD_HIP(hipMemcpyAsync(geo.p, g.data(), geo.bytes, hipMemcpyHostToDevice, st));
        SWD_HIP(hipStreamSynchronize(st));
    }
    return kOk;
}

size_t Engine::footprint_bytes() const
{
    const DevBuf* b[] = {&coefs, &vs_d, &wavelet, &state, &strain, &traces, &snaps, &adjsrc, &imgs, &imgp, &outs, &djobs, &dbmap,
                         &disp.W, &disp.D, &disp.X, &disp.MM, &disp.ml, &disp.mlr, &disp.GMM, &disp.GD,
                         &disp.Bm, &disp.crpre, &disp.cr0, &disp.res, &disp.status, &vp_d, &geo, &vs_cur,
                         &vs_tr, &dk_sm, &stats, &disp.dcols};
    size_t t = 0;
    for (auto* x : b) t += x->bytes;
    return t;
}

int Engine::set_model(const double* vp, const double* vs)
{
    if (!vp || !vs) { set_error("null model"); return kErrArg; }
    const size_t n = (size_t)nzbc * nxbc;
    std::vector<double> c(7 * n);
    int rc = prepare_model(vp, vs, p.nz, p.nx, p.nbc, p.dx, p.dt, p.fd_order, c.data(), c.data() + n,
                           c.data() + 2 * n, c.data() + 3 * n, c.data() + 4 * n, c.data() + 5 * n, c.data() + 6 * n);
    if (rc) return rc;
    vpmax = 0;
    for (size_t k = 0; k < splane; ++k) vpmax = std::max(vpmax, std::max(vp[k], vs[k]));
    // staggered 2nd-order stability bound in 2-D: vmax dt/dx <= 1/sqrt(2)
    if (vpmax * p.dt / p.dx > 1.0 / std::sqrt(2.0)) {
        set_error("CFL violated: vmax*dt/dx = %.4f > 0.7071", vpmax * p.dt / p.dx);
        return kErrCfl;
    }
    // device order temp, b, b1, ca, cl, cm1 from prepare_model's temp, ca, cl, cm, cm1, b, b1
    const int map[6] = {0, 5, 6, 1, 2, 4};
    std::vector<float> h(6 * plane);
    std::vector<float> tmp(n);
    for (int k = 0; k < 6; ++k) {
        const double* srcp = c.data() + (size_t)map[k] * n;
        for (size_t q = 0; q < n; ++q) tmp[q] = (float)srcp[q];
        pack_grid(tmp.data(), nzbc, nxbc, pitch, h.data() + (size_t)k * plane);
        zero_ring(h.data() + (size_t)k * plane, nzbc, nxbc, pitch);
    }
    b1_same = std::equal(h.begin() + plane, h.begin() + 2 * plane, h.begin() + 2 * plane);
    SWD_HIP(hipMemcpyAsync(coefs.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, st));
    if ((rc = launch_pack_coef4(st, coefs.as<float>(), plane, coefs.as<float>() + 6 * plane))) return rc;
    SWD_HIP(hipMemcpyAsync(vs_d.p, vs, splane * sizeof(double), hipMemcpyHostToDevice, st));
    SWD_HIP(hipMemcpyAsync(vp_d.p, vp, splane * sizeof(double), hipMemcpyHostToDevice, st));
    SWD_HIP(hipStreamSynchronize(st));
    have_model = true;
    return kOk;
}

int Engine::set_vs_device(const double* vs_dev)
{
    if (!have_model) { set_error("set_model must run once first (it fixes vp)"); return kErrArg; }
    int rc;
    double h[4];
    if (vp_ratio > 0 && (rc = launch_model_scale(st, vs_dev, vp_ratio, vp_d.as<double>(), splane))) return rc;
    if ((rc = launch_model_stats(st, vp_d.as<double>(), vs_dev, splane, stats.as<double>()))) return rc;
    SWD_HIP(hipMemcpyAsync(h, stats.p, sizeof h, hipMemcpyDeviceToHost, st));
    SWD_HIP(hipStreamSynchronize(st));
    const double vmin = h[0], vmax = h[1];
    if (!(vmax * p.dt / p.dx <= 1.0 / std::sqrt(2.0))) {
        set_error("CFL violated: vmax*dt/dx = %.4f > 0.7071", vmax * p.dt / p.dx);
        return kErrCfl;
    }
    vpmax = vmax;
    const double kappa = damp_kappa(vmin, p.nbc, p.dx);
    if ((rc = launch_prep_coefs(st, vp_d.as<double>(), vs_dev, geo.as<double>(), kappa, p.nz, p.nx, p.nbc, pt_m, nzbc,
                                nxbc, pitch, plane, p.dt, p.dt / p.dx, coefs.as<float>())) ||
        (rc = launch_pack_coef4(st, coefs.as<float>(), plane, coefs.as<float>() + 6 * plane)))
        return rc;
    b1_same = true;  // prep_coefs_kernel writes b1 = b
    if (vs_dev != vs_d.as<double>())
        SWD_HIP(hipMemcpyAsync(vs_d.p, vs_dev, splane * sizeof(double), hipMemcpyDeviceToDevice, st));
    SWD_HIP(hipStreamSynchronize(st));
    return kOk;
}

// Mean pick misfit of the right side (SWD.m:195 and the line search use
// mean(res_r)) over the shots whose right branch runs.
double Engine::mean_right(const double* res_r) const
{
    double s = 0.0;
    int n = 0;
    for (int k = 0; k < p.nshots; ++k)
        if (disp.cols[k].enabled) { s += res_r[k]; ++n; }
    return n ? s / n : 0.0;
}

int Engine::iterate(swd_iteration* it)
{
    if (!it) { set_error("null iteration record"); return kErrArg; }
    if (!have_model || !have_wavelet) { set_error("model and wavelet must be set first"); return kErrArg; }
    int rc;
    const size_t S = p.nshots;
    std::vector<double> rr(S), rl(S);
    // gradient of the current model (SWD.m:171-200) and its misfit
    if ((rc = gradient(nullptr, rr.data(), rl.data(), nullptr, nullptr))) return rc;
    double res0 = mean_right(rr.data());
    it->residual = res0;
    SWD_HIP(hipMemcpyAsync(vs_cur.p, vs_d.p, splane * sizeof(double), hipMemcpyDeviceToDevice, st));
    // dk_vs = smooth2a(dk_vs, smoothZ, smoothX) (SWD.m:201)
    if ((rc = launch_smooth2a(st, outs.as<double>(), p.nz, p.nx, it->smooth_z, it->smooth_x, dk_sm.as<double>())))
        return rc;
    // alpha = ||vs|| / ||dk|| * pur (SWD.m:203-205)
    double h[4];
    if ((rc = launch_model_stats(st, vs_cur.as<double>(), dk_sm.as<double>(), splane, stats.as<double>()))) return rc;
    SWD_HIP(hipMemcpyAsync(h, stats.p, sizeof h, hipMemcpyDeviceToHost, st));
    SWD_HIP(hipStreamSynchronize(st));
    const double alpha = std::sqrt(h[2]) / std::sqrt(h[3]) * it->pur;
    it->alpha = alpha;
    it->forwards = 0;
    // vs1 = clip(vs + alpha * f * dk) (SWD.m:206-207) and its forward-only misfit (SWD.m:215-229)
    auto step = [&](double f) -> int {
        return launch_model_step(st, vs_cur.as<double>(), dk_sm.as<double>(), alpha * f, it->vsmin, it->vsmax,
                                 vs_tr.as<double>(), splane);
    };
    auto trial = [&](double f, double* res) -> int {
        int r;
        if ((r = step(f)) || (r = set_vs_device(vs_tr.as<double>())) ||
            (r = forward_picks(rr.data(), rl.data(), nullptr, nullptr)))
            return r;
        ++it->forwards;
        *res = mean_right(rr.data());
        return kOk;
    };
    double f1 = it->f1, f2 = it->f2, res1 = 0, res2 = it->res2, res3 = 0;
    if ((rc = trial(f1, &res1))) return rc;
    if (res1 > res0) {  // SWD.m:233-256: halve while the misfit grows
        while (res1 > res0 && f1 > 0.0001) {
            f2 = f1;
            res2 = res1;
            f1 = f1 * 0.5;
            if ((rc = trial(f1, &res1))) return rc;
        }
    } else {  // SWD.m:257-278: double once
        f2 = f1 * 2;
        if ((rc = trial(f2, &res2))) return rc;
    }
    // parabola through (0, res0), (f1, res1), (f2, res2) (SWD.m:279-287)
    double gama = (f1 * f1 * (res0 - res2) + f2 * f2 * (res1 - res0)) /
                  (2 * res0 * (f1 - f2) + 2 * res1 * f2 - 2 * res2 * f1);
    if (std::isinf(gama)) gama = 0;
    if ((rc = trial(gama, &res3))) return rc;
    // keep the better trial when the parabola step is worse (SWD.m:308-322)
    if (res3 > res1 || res3 > res2) {
        if (res1 > res2) {
            res0 = res2;
            gama = f2;
        } else {
            res0 = res1;
            gama = f1;
        }
        if ((rc = step(gama)) || (rc = set_vs_device(vs_tr.as<double>()))) return rc;
    } else {
        res0 = res3;
    }
    // vs = vs1: the accepted model stays in vs_d with its coefficient planes
    it->res0 = res0;
    it->gama = gama;
    it->f1 = f1;
    it->f2 = f2;
    it->res1 = res1;
    it->res2 = res2;
    it->res3 = res3;
    return kOk;
}

int Engine::set_wavelet(const double* s)
{
    if (!s) { set_error("null wavelet"); return kErrArg; }
    std::vector<float> h(p.nt);
    for (int k = 0; k < p.nt; ++k) h[k] = (float)s[k];
    SWD_HIP(hipMemcpyAsync(wavelet.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, st));
    SWD_HIP(hipStreamSynchronize(st));
    have_wavelet = true;
    return kOk;
}

int Engine::set_observed_picks(const double* cr0_r, const double* cr0_l)
{
    const size_t n = (size_t)npair * p.nshots;
    if (cr0_r) SWD_HIP(hipMemcpyAsync(disp.cr0.p, cr0_r, n * sizeof(double), hipMemcpyHostToDevice, st));
    if (cr0_l)
        SWD_HIP(hipMemcpyAsync((double*)disp.cr0.p + n, cr0_l, n * sizeof(double), hipMemcpyHostToDevice, st));
    SWD_HIP(hipStreamSynchronize(st));
    return kOk;
}

void Engine::make_jobs(std::vector<FdJob<float>>& jobs, bool adj, bool snapshots)
{
    float* C = coefs.as<float>();
    for (int s = 0; s < p.nshots; ++s) {
        FdJob<float> J;
        memset(&J, 0, sizeof J);
        float* S = state.as<float>() + (size_t)s * 10 * plane;
        for (int b = 0; b < 2; ++b) {
            J.uu[b] = S + (5 * b + 0) * plane;
            J.ww[b] = S + (5 * b + 1) * plane;
            J.xx[b] = S + (5 * b + 2) * plane;
            J.zz[b] = S + (5 * b + 3) * plane;
            J.xz[b] = S + (5 * b + 4) * plane;
        }
        J.temp = C; J.b = C + plane; J.b1 = b1_same ? J.b : C + 2 * plane; J.ca = C + 3 * plane;
        J.cl = C + 4 * plane; J.cm1 = C + 5 * plane; J.c4 = C + 6 * plane;
        const size_t ntr = (size_t)p.nt * p.ng;
        J.src = adj ? adjsrc.as<float>() + s * ntr : wavelet.as<float>();
        J.trace_w = adj ? nullptr : traces.as<float>() + s * ntr;
        J.trace_u = nullptr;
        J.snap = snaps.as<float>() + (size_t)s * nrec * 3 * splane;
        J.img = adj ? imgs.as<double>() + (size_t)s * 3 * splane : nullptr;
        J.imgp = adj ? imgp.as<float>() + (size_t)s * 3 * splane : nullptr;
        J.nzbc = nzbc; J.nxbc = nxbc; J.pitch = pitch; J.nz = p.nz; J.nx = p.nx; J.nbc = p.nbc; J.pt = 1;
        J.sz = sz[s] + pt_m - 1; J.sx = sx[s] + p.nbc - 1; J.stype = adj ? 0 : 5;
        J.gz = p.gz + pt_m - 1; J.gx = p.gx0 + p.nbc - 1; J.dg = p.dg; J.ng = p.ng;
        J.nt = p.nt; J.in_wf = p.in_wf; J.start = 0; J.nsf = (adj || snapshots) ? 3 : 0;
        J.dtx = (float)(p.dt / p.dx); J.dt = (float)p.dt;
        jobs.push_back(J);
    }
}

int run_fd_batch(const std::vector<Engine*>& engines, bool adj, bool snapshots, hipStream_t st, int tz,
                 MarchStats* stats)
{
    std::vector<FdJob<float>> jobs;
    int steps = 0;
    for (Engine* e : engines) {
        e->make_jobs(jobs, adj, snapshots);
        steps = std::max(steps, adj ? e->p.nt - 1 : e->p.nt);
        SWD_HIP(hipMemsetAsync(e->state.p, 0, e->state.bytes, st));
        if (adj) {
            SWD_HIP(hipMemsetAsync(e->imgs.p, 0, e->imgs.bytes, st));
            SWD_HIP(hipMemsetAsync(e->imgp.p, 0, e->imgp.bytes, st));
        }
        else SWD_HIP(hipMemsetAsync(e->traces.p, 0, e->traces.bytes, st));
    }
    Engine* e0 = engines[0];
    MarchStats ms;
    // the first engine's graph setting holds for the batch; it owns the
    // captured graphs (single-engine and multi-engine batches apart)
    const int rc = e0->use_graph
        ? run_march_graph<float>(st, jobs, adj, steps, e0->nt_block[adj], e0->tb_rows[adj], tz,
                                 e0->graphs[adj][snapshots][engines.size() > 1 ? 1 : 0], &ms)
        : run_march<float>(st, jobs, adj, steps, e0->nt_block[adj], e0->tb_rows[adj], tz, &ms);
    if (rc == kOk)
        for (Engine* e : engines) {
            if (adj) { e->times.adj_spl = ms.nt_block; e->times.adj_rows = ms.rows; }
            else { e->times.fwd_spl = ms.nt_block; e->times.fwd_rows = ms.rows; }
        }
    if (stats) *stats = ms;
    return rc;
}

// ---------------------------------------------------------------------------
// Uneven batch with overlapped phases
// ---------------------------------------------------------------------------
namespace {

// One launch table of a batch phase: the jobs of the models it advances and
// their block map, kept alive until every launch that reads it has run.
struct BatchTable {
    DevBuf djobs, dbmap;
    int nblocks = 0;
    bool b1 = false;
};

// Launches of one phase go in chunks with an event after each; a chunk is
// issued only once the chunk two before it has run, so about two chunks are
// queued ahead of the device, and a model joins or leaves within two chunks
// of the moment it can.
struct Pacer {
    static constexpr int kChunk = 8;
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool used[2] = {false, false};
    int k = 0;
    Pacer() = default;
    Pacer(const Pacer&) = delete;
    Pacer& operator=(const Pacer&) = delete;
    ~Pacer()
    {
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    int create()
    {
        for (auto& e : ev) SWD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        return kOk;
    }
    // 1: a chunk may be issued, 0: not yet, < 0: error
    int ready()
    {
        if (!used[k]) return 1;
        const hipError_t e = hipEventQuery(ev[k]);
        if (e == hipSuccess) return 1;
        if (e == hipErrorNotReady) return 0;
        set_error("event query failed: %s", hipGetErrorString(e));
        return kErrHip;
    }
    int issued(hipStream_t st)
    {
        SWD_HIP(hipEventRecord(ev[k], st));
        used[k] = true;
        k ^= 1;
        return kOk;
    }
};

// 1: the event has completed, 0: not yet, < 0: error
int event_done(hipEvent_t e)
{
    const hipError_t q = hipEventQuery(e);
    if (q == hipSuccess) return 1;
    if (q == hipErrorNotReady) return 0;
    set_error("event query failed: %s", hipGetErrorString(q));
    return kErrHip;
}

}  // namespace

// The FD phases of the models overlap: the forward launches advance every
// model until its last step, and as soon as a model's forward has run, its
// Radon / pick / adjoint-source stage runs on its own stream and its adjoint
// jobs join the adjoint launches at their next chunk, while longer models are
// still in their forward.  Its assembly follows its last adjoint launch.
// Done one phase after the other, the batch leaves the GPU part idle while
// only the longest models remain in a phase: the 8-model uneven batch took
// 380 ms that way, 345 ms through the serving loop, which overlaps requests.
// A dependency between streams is enqueued only once its event has completed
// (the host polls), so no stream holds a barrier that could stall the others
// sharing its hardware queue.  Launch tables change only at chunk
// boundaries; a joining model's job origin (start, launch0) is the adjoint
// launch it joins at, so its steps and ping-pong parity are those it would
// have alone, and results are bitwise those of each model's own gradient.
int batch_gradient_overlapped(const std::vector<Engine*>& v, double** dk_vs, double* kernel_us)
{
    const int n = (int)v.size();
    Engine* e0 = v[0];
    const int NT[2] = {e0->nt_block[0], e0->nt_block[1]};
    if (NT[0] < 2 || NT[1] < 2) { set_error("overlapped batch needs time-blocked launches"); return kErrArg; }
    for (auto& b : e0->batch_fd)
        if (!b) SWD_HIP(hipStreamCreateWithFlags(&b, hipStreamNonBlocking));
    hipStream_t fs[2] = {e0->batch_fd[0], e0->batch_fd[1]};
    int rc;
    std::vector<std::vector<FdJob<float>>> jobs[2];
    std::vector<int> nl[2];   // launches of each model per phase
    for (int ph = 0; ph < 2; ++ph) jobs[ph].resize(n);
    for (int m = 0; m < n; ++m) {
        Engine* x = v[m];
        x->make_jobs(jobs[0][m], false, true);
        x->make_jobs(jobs[1][m], true, true);
        nl[0].push_back(div_up(x->p.nt, NT[0]));
        nl[1].push_back(div_up(std::max(x->p.nt - 1, 0), NT[1]));
        SWD_HIP(hipMemsetAsync(x->state.p, 0, x->state.bytes, fs[0]));
        SWD_HIP(hipMemsetAsync(x->traces.p, 0, x->traces.bytes, fs[0]));
    }
    EventSet ev(3 * (size_t)n);   // per model: forward run, stage run, adjoint run
    EventSet span(4);             // per phase: before its first launch, after its latest
    Pacer pace[2];
    if ((rc = ev.create()) || (rc = span.create()) || (rc = pace[0].create()) || (rc = pace[1].create())) return rc;
    enum { kFwd, kAfterFwd, kStage, kAdj, kAfterAdj, kDone };
    std::vector<int> at(n, kFwd), join(n, 0);
    std::vector<std::unique_ptr<BatchTable>> keep;
    BatchTable* tab[2] = {nullptr, nullptr};
    bool dirty[2] = {true, true};
    int g[2] = {0, 0};   // launches issued per phase
    int left = n;
    // errors only stop the loop: everything queued drains before the tables go
    rc = kOk;
    auto hip = [&](hipError_t e, const char* what) {
        if (e != hipSuccess) {
            set_error("%s failed: %s", what, hipGetErrorString(e));
            rc = kErrHip;
        }
        return rc == kOk;
    };
    while (rc == kOk && left > 0) {
        bool moved = false;
        for (int m = 0; m < n && rc == kOk; ++m) {
            Engine* x = v[m];
            int d = 0;
            if (at[m] == kAfterFwd && (d = event_done(ev.ev[3 * m])) > 0) {
                // the model's forward has run: its stage, and the adjoint's initial state
                if (!hip(hipStreamWaitEvent(x->st, ev.ev[3 * m], 0), "hipStreamWaitEvent") ||
                    !hip(hipMemsetAsync(x->state.p, 0, x->state.bytes, x->st), "hipMemsetAsync") ||
                    !hip(hipMemsetAsync(x->imgs.p, 0, x->imgs.bytes, x->st), "hipMemsetAsync") ||
                    !hip(hipMemsetAsync(x->imgp.p, 0, x->imgp.bytes, x->st), "hipMemsetAsync"))
                    break;
                if ((rc = x->dispersion()) || (rc = x->adjoint_source()) || (rc = ev.record(3 * m + 1, x->st))) break;
                at[m] = kStage;
                moved = true;
            } else if (at[m] == kStage && (d = event_done(ev.ev[3 * m + 1])) > 0) {
                join[m] = g[1];
                for (auto& J : jobs[1][m]) {
                    J.start = g[1] * NT[1];
                    J.launch0 = g[1];
                }
                at[m] = kAdj;
                dirty[1] = true;
                moved = true;
                if (nl[1][m] == 0) at[m] = kAfterAdj, rc = ev.record(3 * m + 2, x->st);
            } else if (at[m] == kAfterAdj && (d = event_done(ev.ev[3 * m + 2])) > 0) {
                if (!hip(hipStreamWaitEvent(x->st, ev.ev[3 * m + 2], 0), "hipStreamWaitEvent") ||
                    (rc = x->assemble()))
                    break;
                if (dk_vs && dk_vs[m] &&
                    !hip(hipMemcpyAsync(dk_vs[m], x->outs.p, x->splane * sizeof(double), hipMemcpyDeviceToHost, x->st),
                         "hipMemcpyAsync"))
                    break;
                at[m] = kDone;
                --left;
                moved = true;
            }
            if (d < 0) rc = d;
        }
        for (int ph = 0; ph < 2 && rc == kOk; ++ph) {
            const int in = ph ? kAdj : kFwd;
            std::vector<int> mem;
            for (int m = 0; m < n; ++m)
                if (at[m] == in) mem.push_back(m);
            if (mem.empty()) continue;
            const int r = pace[ph].ready();
            if (r <= 0) {
                rc = r;
                continue;
            }
            if (dirty[ph]) {
                std::vector<FdJob<float>> js;
                for (int m : mem) js.insert(js.end(), jobs[ph][m].begin(), jobs[ph][m].end());
                const int rows = (ph == 0 && g[0] == 0 && e0->tb_rows[0] > 0) ? e0->tb_rows[0]
                                                                              : auto_band_rows(js, ph == 1, NT[ph]);
                auto t = std::make_unique<BatchTable>();
                MarchPlan plan;
                if ((rc = build_march_plan(js, rows, plan, kMarchLanes - 2 * NT[ph], NT[ph], ph == 1)) ||
                    (rc = t->djobs.upload(js.data(), js.size() * sizeof(FdJob<float>), fs[ph])) ||
                    (rc = t->dbmap.upload(plan.bmap.data(), plan.bmap.size() * sizeof(BlockMap), fs[ph])))
                    break;
                t->nblocks = plan.nblocks;
                t->b1 = jobs_b1(js);
                tab[ph] = t.get();
                keep.push_back(std::move(t));
                dirty[ph] = false;
            }
            int next = INT_MAX;   // the next launch after which a model leaves
            for (int m : mem) next = std::min(next, (ph ? join[m] : 0) + nl[ph][m]);
            const int end = std::min(next, g[ph] + Pacer::kChunk);
            const BatchTable& T = *tab[ph];
            if (g[ph] == 0 && (rc = span.record(2 * ph, fs[ph]))) break;
            for (; g[ph] < end && rc == kOk; ++g[ph])
                rc = ph ? launch_adj_tb<float>(fs[1], T.djobs.as<FdJob<float>>(), T.dbmap.as<BlockMap>(), T.nblocks,
                                               g[1] * NT[1], g[1], NT[1], T.b1)
                        : launch_fwd_tb<float>(fs[0], T.djobs.as<FdJob<float>>(), T.dbmap.as<BlockMap>(), T.nblocks,
                                               g[0] * NT[0], g[0], NT[0], T.b1);
            if (rc || (rc = span.record(2 * ph + 1, fs[ph])) || (rc = pace[ph].issued(fs[ph]))) break;
            for (int m : mem)
                if ((ph ? join[m] : 0) + nl[ph][m] == g[ph]) {
                    if ((rc = ev.record(3 * m + 2 * ph, fs[ph]))) break;
                    at[m] = ph ? kAfterAdj : kAfterFwd;
                    dirty[ph] = true;
                }
            moved = true;
        }
        if (!moved && rc == kOk) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    for (auto s : fs) (void)hipStreamSynchronize(s);
    for (Engine* x : v) (void)hipStreamSynchronize(x->st);
    if (kernel_us)
        for (int ph = 0; ph < 2; ++ph) {
            float ms = 0;
            if (rc != kOk || g[ph] == 0 || hipEventElapsedTime(&ms, span.ev[2 * ph], span.ev[2 * ph + 1]) != hipSuccess)
                ms = 0;
            kernel_us[ph] = g[ph] ? 1000.0 * ms / g[ph] : 0.0;
        }
    return rc;
}

int Engine::run_fd(bool adj, bool snapshots)
{
    if (fd_path == 0) {
        MarchStats ms;
        const int rc = run_fd_batch({this}, adj, snapshots, st, tz, &ms);
        if (adj) times.adj_kernel_us = ms.kernel_us;
        else times.fwd_kernel_us = ms.kernel_us;
        return rc;
    }
    // plain comparator: one shot at a time, one launch per reference statement
    int rc;
    if (!strain.p) {
        if ((rc = strain.alloc(4 * plane * sizeof(float)))) return rc;
    }
    SWD_HIP(hipMemsetAsync(strain.p, 0, strain.bytes, st));
    std::vector<FdJob<float>> jobs;
    make_jobs(jobs, adj, snapshots);
    if (adj) {
        SWD_HIP(hipMemsetAsync(imgs.p, 0, imgs.bytes, st));
        SWD_HIP(hipMemsetAsync(imgp.p, 0, imgp.bytes, st));
    }
    else SWD_HIP(hipMemsetAsync(traces.p, 0, traces.bytes, st));
    float* w[4] = {strain.as<float>(), strain.as<float>() + plane, strain.as<float>() + 2 * plane,
                   strain.as<float>() + 3 * plane};
    for (auto& J : jobs) {
        SWD_HIP(hipMemsetAsync(J.uu[0], 0, 10 * plane * sizeof(float), st));
        if (adj) {
            for (int it = p.nt - 2; it >= 0; --it)
                if ((rc = plain_adjoint_step<float>(st, J, it))) return rc;
        } else {
            for (int it = 0; it < p.nt; ++it)
                if ((rc = plain_forward_step<float>(st, J, w, it))) return rc;
        }
    }
    SWD_HIP(hipStreamSynchronize(st));
    return kOk;
}

int Engine::forward(bool snapshots)
{
    if (!have_model || !have_wavelet) { set_error("model and wavelet must be set first"); return kErrArg; }
    return run_fd(false, snapshots);
}
int Engine::adjoint() { return run_fd(true, true); }

int Engine::dispersion()
{
    int rc = disp.forward(traces.as<float>());
    if (rc) return rc;
    return disp.pick(p.ini, 2.f, 4.f);
}

int Engine::adjoint_source() { return disp.adjoint(p.beta, adjsrc.as<float>()); }

int Engine::assemble()
{
    double* o = outs.as<double>();
    return launch_assemble_gradient(st, imgs.as<double>(), p.nshots, p.nz, p.nx, vs_d.as<double>(), o, o + splane,
                                    o + 2 * splane, o + 3 * splane);
}

int Engine::read_picks(double* res_r, double* res_l, double* cr_r, double* cr_l)
{
    const size_t S = p.nshots, n = (size_t)npair * S;
    std::vector<int> stat(2 * S);
    SWD_HIP(hipMemcpyAsync(stat.data(), disp.status.p, stat.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    if (res_r) SWD_HIP(hipMemcpyAsync(res_r, disp.res.p, S * sizeof(double), hipMemcpyDeviceToHost, st));
    if (res_l) SWD_HIP(hipMemcpyAsync(res_l, (double*)disp.res.p + S, S * sizeof(double), hipMemcpyDeviceToHost, st));
    if (cr_r) SWD_HIP(hipMemcpyAsync(cr_r, disp.crpre.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    if (cr_l) SWD_HIP(hipMemcpyAsync(cr_l, (double*)disp.crpre.p + n, n * sizeof(double), hipMemcpyDeviceToHost, st));
    SWD_HIP(hipStreamSynchronize(st));
    for (size_t c = 0; c < 2 * S; ++c)
        if (disp.cols[c].enabled && stat[c]) {
            set_error("pick of shot %zu (%s side) kept fewer than two columns", c % S, c < S ? "right" : "left");
            return kErrPick;
        }
    return kOk;
}

static float elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) ms = 0;
    return ms;
}

int Engine::gradient(double* dk_vs, double* res_r, double* res_l, double* cr_r, double* cr_l)
{
    EventSet ev(5);
    int rc;
    if ((rc = ev.create()) || (rc = ev.record(0, st)) || (rc = forward(true)) || (rc = ev.record(1, st)) ||
        (rc = dispersion()) || (rc = adjoint_source()) || (rc = ev.record(2, st)) || (rc = adjoint()) ||
        (rc = ev.record(3, st)) || (rc = assemble()) || (rc = ev.record(4, st)))
        return rc;
    if (dk_vs) SWD_HIP(hipMemcpyAsync(dk_vs, outs.p, splane * sizeof(double), hipMemcpyDeviceToHost, st));
    rc = read_picks(res_r, res_l, cr_r, cr_l);
    const hipEvent_t* e = ev.ev.data();
    times.fwd_ms = elapsed(e[0], e[1]);
    times.disp_ms = elapsed(e[1], e[2]);
    times.adj_ms = elapsed(e[2], e[3]);
    times.asm_ms = elapsed(e[3], e[4]);
    return rc;
}

int Engine::forward_picks(double* res_r, double* res_l, double* cr_r, double* cr_l)
{
    int rc;
    if ((rc = forward(false)) || (rc = dispersion())) return rc;
    return read_picks(res_r, res_l, cr_r, cr_l);
}

int Engine::observe(double* cr0_r, double* cr0_l)
{
    int rc;
    if ((rc = forward(false)) || (rc = dispersion())) return rc;
    const size_t n = (size_t)npair * p.nshots;
    SWD_HIP(hipMemcpyAsync(disp.cr0.p, disp.crpre.p, 2 * n * sizeof(double), hipMemcpyDeviceToDevice, st));
    return read_picks(nullptr, nullptr, cr0_r, cr0_l);
}

}  // namespace swd

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace swd;

struct swd_engine {
    Engine e;
};

extern "C" {

int swd_engine_create(const swd_problem* p, swd_engine** out)
{
    if (!out) { set_error("null output handle"); return kErrArg; }
    *out = nullptr;
    swd_engine* h = new (std::nothrow) swd_engine;
    if (!h) { set_error("out of host memory"); return kErrNoMem; }
    const int rc = h->e.create(p, nullptr);
    if (rc) { delete h; return rc; }
    *out = h;
    return kOk;
}

int swd_problem_npair(const swd_problem* p, int* npair)
{
    if (!p || !npair) { set_error("null argument"); return kErrArg; }
    DispGeom g;
    const int rc = problem_geometry(*p, g);
    if (rc) return rc;
    *npair = g.npair;
    return kOk;
}

int swd_engine_destroy(swd_engine* e)
{
    delete e;
    return kOk;
}

int swd_engine_set_model(swd_engine* e, const double* vp, const double* vs)
{
    return e ? e->e.set_model(vp, vs) : kErrArg;
}

int swd_engine_set_wavelet(swd_engine* e, const double* s) { return e ? e->e.set_wavelet(s) : kErrArg; }

int swd_engine_set_observed_picks(swd_engine* e, const double* cr0_r, const double* cr0_l)
{
    return e ? e->e.set_observed_picks(cr0_r, cr0_l) : kErrArg;
}

int swd_engine_observe(swd_engine* e, double* cr0_r, double* cr0_l) { return e ? e->e.observe(cr0_r, cr0_l) : kErrArg; }

int swd_engine_gradient(swd_engine* e, double* dk_vs, double* res_r, double* res_l, double* cr_r, double* cr_l)
{
    return e ? e->e.gradient(dk_vs, res_r, res_l, cr_r, cr_l) : kErrArg;
}

int swd_engine_forward_picks(swd_engine* e, double* res_r, double* res_l, double* cr_r, double* cr_l)
{
    return e ? e->e.forward_picks(res_r, res_l, cr_r, cr_l) : kErrArg;
}

int swd_engine_iterate(swd_engine* e, swd_iteration* it) { return e ? e->e.iterate(it) : kErrArg; }

int swd_smooth2a(const double* in, int nz, int nx, int nr, int nc, double* out)
{
    if (!in || !out || nz < 1 || nx < 1 || nr < 0 || nc < 0) { set_error("bad smooth2a arguments"); return kErrArg; }
    Stream s;
    DevBuf a, b;
    int rc;
    const size_t n = (size_t)nz * nx * sizeof(double);
    if ((rc = s.create()) || (rc = a.alloc(n)) || (rc = b.alloc(n))) return rc;
    SWD_HIP(hipMemcpyAsync(a.p, in, n, hipMemcpyHostToDevice, s.s));
    if ((rc = launch_smooth2a(s.s, a.as<double>(), nz, nx, nr, nc, b.as<double>()))) return rc;
    SWD_HIP(hipMemcpyAsync(out, b.p, n, hipMemcpyDeviceToHost, s.s));
    SWD_HIP(hipStreamSynchronize(s.s));
    return kOk;
}

int swd_engine_read_model(swd_engine* e, double* vs)
{
    if (!e || !vs) { set_error("null argument"); return kErrArg; }
    Engine& g = e->e;
    SWD_HIP(hipMemcpyAsync(vs, g.vs_d.p, g.splane * sizeof(double), hipMemcpyDeviceToHost, g.st));
    SWD_HIP(hipStreamSynchronize(g.st));
    return kOk;
}

int swd_engine_set_variant(swd_engine* e, const swd_variant* v)
{
    if (!e || !v) { set_error("null argument"); return kErrArg; }
    if (!(v->vp_ratio >= 0) || std::isinf(v->vp_ratio)) { set_error("vp_ratio must be finite and >= 0"); return kErrArg; }
    if (v->res_first < 0 || v->res_last < 0) { set_error("negative misfit window"); return kErrArg; }
    Engine& g = e->e;
    if ((v->res_first > 1 ? v->res_first - 1 : 0) + v->res_last >= g.npair) {
        set_error("misfit window cr(%d:end-%d) is empty for %d columns", v->res_first, v->res_last, g.npair);
        return kErrArg;
    }
    int rc;
    if ((rc = set_plan_variant(g.disp, *v))) return rc;
    g.vp_ratio = v->vp_ratio;
    return kOk;
}

int swd_engine_set_graph(swd_engine* e, int on)
{
    if (!e) return kErrArg;
    const int old = e->e.use_graph ? 1 : 0;
    e->e.use_graph = on != 0;
    return old;
}

int swd_engine_set_fd_path(swd_engine* e, int mode)
{
    if (!e) return kErrArg;
    const int old = e->e.fd_path;
    e->e.fd_path = mode ? 1 : 0;
    return old;
}

int swd_engine_set_tile_rows(swd_engine* e, int tz)
{
    if (!e || tz < 1) { set_error("bad tile rows"); return kErrArg; }
    e->e.tz = tz;
    return kOk;
}

int swd_engine_run(swd_engine* e, int stage)
{
    if (!e) return kErrArg;
    Engine& g = e->e;
    int rc;
    switch (stage) {
    case 1: return g.forward(true);
    case 2:
        if ((rc = g.dispersion())) return rc;
        SWD_HIP(hipStreamSynchronize(g.st));
        return kOk;
    case 3:
        if ((rc = g.adjoint_source())) return rc;
        SWD_HIP(hipStreamSynchronize(g.st));
        return kOk;
    case 4: return g.adjoint();
    case 5:
        if ((rc = g.assemble())) return rc;
        SWD_HIP(hipStreamSynchronize(g.st));
        return kOk;
    default: set_error("unknown stage %d", stage); return kErrArg;
    }
}

int swd_engine_set_picks(swd_engine* e, const double* cr_r, const double* cr_l)
{
    if (!e) return kErrArg;
    Engine& g = e->e;
    const size_t n = (size_t)g.npair * g.p.nshots;
    if (cr_r) SWD_HIP(hipMemcpyAsync(g.disp.crpre.p, cr_r, n * sizeof(double), hipMemcpyHostToDevice, g.st));
    if (cr_l)
        SWD_HIP(hipMemcpyAsync((double*)g.disp.crpre.p + n, cr_l, n * sizeof(double), hipMemcpyHostToDevice, g.st));
    SWD_HIP(hipStreamSynchronize(g.st));
    return kOk;
}

int swd_engine_set_time_block(swd_engine* e, int nt_block, int rows)
{
    return swd_engine_set_phase_time_block(e, -1, nt_block, rows);
}

int swd_engine_set_phase_time_block(swd_engine* e, int phase, int nt_block, int rows)
{
    if (!e || nt_block < 1 || nt_block > 6 || rows < 0 || phase < -1 || phase > 1) {
        set_error("time block must be 1..6 steps, rows >= 0 (0: automatic), phase -1..1");
        return kErrArg;
    }
    for (int ph = 0; ph < 2; ++ph)
        if (phase < 0 || phase == ph) {
            e->e.nt_block[ph] = nt_block;
            e->e.tb_rows[ph] = rows;
        }
    return kOk;
}

int swd_batch_autotune(swd_engine** es, int n, int steps, int* choice, double* table, int table_rows, int* n_rows)
{
    if (!es || n < 1) { set_error("empty batch"); return kErrArg; }
    std::vector<Engine*> v;
    for (int i = 0; i < n; ++i) {
        if (!es[i]) { set_error("null engine %d", i); return kErrArg; }
        if (!es[i]->e.have_model || !es[i]->e.have_wavelet) { set_error("engine %d lacks model/wavelet", i); return kErrArg; }
        v.push_back(&es[i]->e);
    }
    std::vector<TuneRow> rows;
    const int rc = autotune_fd_batch(v, steps, rows);
    if (rc) return rc;
    if (choice) {
        choice[0] = v[0]->nt_block[0]; choice[1] = v[0]->tb_rows[0];
        choice[2] = v[0]->nt_block[1]; choice[3] = v[0]->tb_rows[1];
    }
    if (table)
        for (int i = 0; i < (int)rows.size() && i < table_rows; ++i) {
            table[4 * i + 0] = rows[i].nt_block; table[4 * i + 1] = rows[i].rows;
            table[4 * i + 2] = rows[i].fwd_us; table[4 * i + 3] = rows[i].adj_us;
        }
    if (n_rows) *n_rows = (int)rows.size();
    return kOk;
}

int swd_engine_autotune(swd_engine* e, int steps, int* choice, double* table, int table_rows, int* n_rows)
{
    return swd_batch_autotune(&e, 1, steps, choice, table, table_rows, n_rows);
}

int swd_engine_info(swd_engine* e, long long* vals, int n)
{
    if (!e || !vals) return kErrArg;
    const Engine& g = e->e;
    const long long v[10] = {g.nzbc, g.nxbc, g.pitch, g.nrec, g.npair, g.np, (long long)g.footprint_bytes(),
                             g.disp.g.ccn, g.disp.g.lf, g.tz};
    for (int i = 0; i < n && i < 10; ++i) vals[i] = v[i];
    return kOk;
}

int swd_engine_timings(swd_engine* e, double* ms, int n)
{
    if (!e || !ms) return kErrArg;
    const StageTimes& t = e->e.times;
    const Engine& g = e->e;
    const double v[15] = {t.fwd_ms, t.disp_ms, t.adj_ms, t.asm_ms, t.fwd_kernel_us, t.adj_kernel_us,
                          (double)t.fwd_spl, (double)t.adj_spl, (double)t.fwd_rows, (double)t.adj_rows,
                          (double)g.snap_placements, (double)g.snap_choice, g.snap_placement_ms[0],
                          g.snap_placement_ms[1], g.snap_placement_ms[2]};
    for (int i = 0; i < n && i < 15; ++i) ms[i] = v[i];
    return kOk;
}

int swd_engine_read(swd_engine* e, int what, void* dst, long long bytes)
{
    if (!e || !dst) return kErrArg;
    Engine& g = e->e;
    const DevBuf* src = nullptr;
    size_t off = 0, n = 0;
    switch (what) {
    case 0: src = &g.traces; n = g.traces.bytes; break;                       // [shot][t][g] float
    case 1: src = &g.adjsrc; n = g.adjsrc.bytes; break;                       // [shot][t][g] float
    case 2: src = &g.outs; n = 4 * g.splane * sizeof(double); break;          // dk_vs, g_cl, g_cm, g_il
    case 3: src = &g.disp.ml; n = g.disp.ml.bytes; break;                     // [col][npair][np] float
    case 4: src = &g.imgs; n = g.imgs.bytes; break;                           // [shot][3][nz][nx] double
    case 5: src = &g.dk_sm; n = g.dk_sm.bytes; break;                         // smoothed dk_vs, nz x nx double
    default: set_error("unknown buffer %d", what); return kErrArg;
    }
    if ((size_t)bytes < n) { set_error("destination too small (%lld < %zu)", bytes, n); return kErrArg; }
    SWD_HIP(hipMemcpyAsync(dst, (const char*)src->p + off, n, hipMemcpyDeviceToHost, g.st));
    SWD_HIP(hipStreamSynchronize(g.st));
    return kOk;
}

int swd_batch_gradient(swd_engine** es, int n, double** dk_vs, double* kernel_us)
{
    if (!es || n < 1) { set_error("empty batch"); return kErrArg; }
    std::vector<Engine*> v;
    for (int i = 0; i < n; ++i) {
        if (!es[i]) { set_error("null engine %d", i); return kErrArg; }
        if (!es[i]->e.have_model || !es[i]->e.have_wavelet) { set_error("engine %d lacks model/wavelet", i); return kErrArg; }
        v.push_back(&es[i]->e);
    }
    hipStream_t st = v[0]->st;
    // every engine's stream must be idle before the shared launches
    for (Engine* x : v) SWD_HIP(hipStreamSynchronize(x->st));
    bool overlap = v[0]->nt_block[0] >= 2 && v[0]->nt_block[1] >= 2;
    for (Engine* x : v) overlap = overlap && x->fd_path == 0;
    if (overlap) return batch_gradient_overlapped(v, dk_vs, kernel_us);
    MarchStats us[2];
    int rc = run_fd_batch(v, false, true, st, v[0]->tz, &us[0]);
    if (rc) return rc;
    for (Engine* x : v) {
        if ((rc = x->dispersion()) || (rc = x->adjoint_source())) return rc;
        SWD_HIP(hipStreamSynchronize(x->st));
    }
    if ((rc = run_fd_batch(v, true, true, st, v[0]->tz, &us[1]))) return rc;
    for (int i = 0; i < n; ++i) {
        Engine* x = v[i];
        if ((rc = x->assemble())) return rc;
        if (dk_vs && dk_vs[i])
            SWD_HIP(hipMemcpyAsync(dk_vs[i], x->outs.p, x->splane * sizeof(double), hipMemcpyDeviceToHost, x->st));
        SWD_HIP(hipStreamSynchronize(x->st));
    }
    if (kernel_us) { kernel_us[0] = us[0].kernel_us; kernel_us[1] = us[1].kernel_us; }
    return kOk;
}

}  // extern "C"
