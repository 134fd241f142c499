// Host side of the SWD FD entry points (C ABI in include/swd_abi.h):
// the MEX-compatible forward / adjoint calls, model preparation and the
// Ricker wavelet.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>

#include "swd_abi.h"
#include "swd_internal.h"
#include "swd_host.h"

namespace swd {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

const char* last_error() { return g_err; }

static std::atomic<int> g_fd_path{0};
static std::atomic<int> g_fd_tz{16};

int fd_tile_rows() { return g_fd_tz.load(); }
void set_fd_tile_rows(int tz) { g_fd_tz.store(tz < 1 ? 1 : tz); }
int fd_force_plain() { return g_fd_path.load(); }

int order_pt(int fd_order)
{
    switch (fd_order) {
    case 22: return 1;
    case 24: return 2;
    case 26: return 3;
    case 28: return 4;
    default: return -1;
    }
}

int grid_pitch(int nxbc) { return div_up(nxbc, 32) * 32; }

// ---------------------------------------------------------------------------
// layout conversion between MATLAB column-major and the internal pitched rows
// ---------------------------------------------------------------------------
template <typename T>
void pack_grid(const T* cm, int nzbc, int nxbc, int pitch, T* out)
{
    for (int i = 0; i < nzbc; ++i) {
        T* row = out + (size_t)i * pitch;
        for (int j = 0; j < nxbc; ++j) row[j] = cm[(size_t)j * nzbc + i];
        for (int j = nxbc; j < pitch; ++j) row[j] = T(0);
    }
}
template void pack_grid<float>(const float*, int, int, int, float*);
template void pack_grid<double>(const double*, int, int, int, double*);

// The reference never updates the outer ring of the padded grid; with zero
// coefficients there, the fused kernels compute that ring as exact zeros.
template <typename T>
void zero_ring(T* plane, int nzbc, int nxbc, int pitch)
{
    for (int j = 0; j < nxbc; ++j) {
        plane[j] = T(0);
        plane[(size_t)(nzbc - 1) * pitch + j] = T(0);
    }
    for (int i = 0; i < nzbc; ++i) {
        plane[(size_t)i * pitch] = T(0);
        plane[(size_t)i * pitch + nxbc - 1] = T(0);
    }
}
template void zero_ring<float>(float*, int, int, int);
template void zero_ring<double>(double*, int, int, int);

// ---------------------------------------------------------------------------
// decoded FD descriptors
// ---------------------------------------------------------------------------
struct FwdDesc {
    int nt, nzbc, nxbc, ng, sz, sx, gz, gx, dg, stype, order, in_wf, nz, nx, fmt, nbc, pt;
    double dtx;
};

template <typename T>
static int decode_fwd(const T* iv, FwdDesc& d)
{
    d.nt = (int)iv[0]; d.nzbc = (int)iv[1]; d.nxbc = (int)iv[2]; d.dtx = (double)iv[3];
    d.ng = (int)iv[4]; d.sz = (int)iv[5] - 1; d.sx = (int)iv[6] - 1; d.gz = (int)iv[7] - 1;
    d.gx = (int)iv[8] - 1; d.dg = (int)iv[9]; d.stype = (int)iv[10]; d.order = (int)iv[11];
    d.in_wf = (int)iv[12]; d.nz = (int)iv[13]; d.nx = (int)iv[14]; d.fmt = (int)iv[15];
    d.nbc = (d.nxbc - d.nx) / 2;
    d.pt = order_pt(d.order);
    if (d.pt < 0) { set_error("fd_order %d is not one of 22/24/26/28", d.order); return kErrArg; }
    if (d.nt <= 0 || d.nzbc <= 2 * d.pt || d.nxbc <= 2 * d.pt || d.ng <= 0 || d.dg <= 0 || d.in_wf <= 0) {
        set_error("bad grid/time descriptor"); return kErrArg;
    }
    if (d.nt % d.in_wf != 0) { set_error("nt (%d) must be a multiple of in_wf (%d)", d.nt, d.in_wf); return kErrArg; }
    if (d.gz < 0 || d.gz >= d.nzbc || d.gx < 0 || d.gx + (d.ng - 1) * d.dg >= d.nxbc) {
        set_error("receivers outside the padded grid"); return kErrArg;
    }
    if (d.stype >= 1 && d.stype <= 5 && (d.sz < 1 || d.sz >= d.nzbc || d.sx < 0 || d.sx + 1 >= d.nxbc)) {
        set_error("source outside the padded grid"); return kErrArg;
    }
    if (d.pt + d.nz >= d.nzbc || d.nbc < 0 || d.nbc + d.nx > d.nxbc || d.nz <= 0 || d.nx <= 0) {
        set_error("interior nz x nx does not fit the padded grid"); return kErrArg;
    }
    return kOk;
}

static bool in_block(int i, int j, int nzbc, int nxbc, int pt)
{
    return i >= pt && i < nzbc - pt && j >= pt && j < nxbc - pt;
}

static bool fwd_fusable(const FwdDesc& d)
{
    if (d.order != 22) return false;
    const int pt = d.pt;
    switch (d.stype) {
    case 1: case 3: case 4: case 5: return in_block(d.sz, d.sx, d.nzbc, d.nxbc, pt);
    case 2:
        return in_block(d.sz, d.sx, d.nzbc, d.nxbc, pt) && in_block(d.sz - 1, d.sx, d.nzbc, d.nxbc, pt) &&
               in_block(d.sz, d.sx + 1, d.nzbc, d.nxbc, pt);
    default: return true;
    }
}

// ---------------------------------------------------------------------------
// one-shot forward (MEX-compatible)
// ---------------------------------------------------------------------------
// The device coefficient block (planes temp, b, b1, ca, cl, cm1); for float
// it also holds temp, ca, cl, cm1 interleaved per cell (FdJob::c4), which the
// time-blocked kernels read.
template <typename T>
static int upload_coefs(DevBuf& coefs, const std::vector<T>& hc, size_t plane, hipStream_t st, const T*& c4)
{
    int rc;
    c4 = nullptr;
    if constexpr (std::is_same<T, float>::value) {
        if ((rc = coefs.alloc(10 * plane * sizeof(float)))) return rc;
        SWD_HIP(hipMemcpyAsync(coefs.p, hc.data(), hc.size() * sizeof(float), hipMemcpyHostToDevice, st));
        if ((rc = launch_pack_coef4(st, (const float*)coefs.p, plane, (float*)coefs.p + 6 * plane))) return rc;
        c4 = (const T*)coefs.p + 6 * plane;
        return kOk;
    }
    return coefs.upload(hc.data(), hc.size() * sizeof(T), st);
}

template <typename T>
static int staggerfd_impl(const T* iv, const T* const coef_in[7], const T* s, T* seismo_u, T* seismo_w, T* snap_out[4])
{
    if (!iv || !s || !seismo_w) { set_error("null argument"); return kErrArg; }
    for (int k = 0; k < 7; ++k)
        if (k != 3 && !coef_in[k]) { set_error("null coefficient array"); return kErrArg; }
    FwdDesc d;
    int rc = decode_fwd(iv, d);
    if (rc) return rc;
    const bool want_snap = snap_out[0] && snap_out[1] && snap_out[2] && snap_out[3];
    const bool fused = !fd_force_plain() && fwd_fusable(d);
    const int P = grid_pitch(d.nxbc);
    const size_t plane = (size_t)d.nzbc * P;
    const int nrec = d.nt / d.in_wf;
    const size_t splane = (size_t)d.nz * d.nx;
    const size_t ntr = (size_t)d.nt * d.ng;

    Stream st;
    if ((rc = st.create())) return rc;
    DevBuf coefs, state, strain, traces, snaps, src;
    // coefficient order in the device block: temp, b, b1, ca, cl, cm1
    const int map[6] = {0, 5, 6, 1, 2, 4};
    std::vector<T> hc(6 * plane);
    for (int k = 0; k < 6; ++k) {
        pack_grid(coef_in[map[k]], d.nzbc, d.nxbc, P, hc.data() + k * plane);
        if (fused) zero_ring(hc.data() + k * plane, d.nzbc, d.nxbc, P);
    }
    const T* c4 = nullptr;
    if ((rc = upload_coefs(coefs, hc, plane, st.s, c4))) return rc;
    if ((rc = state.alloc(10 * plane * sizeof(T))) || (rc = traces.alloc(2 * ntr * sizeof(T))) ||
        (rc = src.upload(s, (size_t)d.nt * sizeof(T), st.s)))
        return rc;
    SWD_HIP(hipMemsetAsync(state.p, 0, 10 * plane * sizeof(T), st.s));
    SWD_HIP(hipMemsetAsync(traces.p, 0, 2 * ntr * sizeof(T), st.s));
    if (want_snap) {
        if ((rc = snaps.alloc((size_t)nrec * 4 * splane * sizeof(T)))) return rc;
    }
    T* C = (T*)coefs.p;
    T* S = (T*)state.p;
    FdJob<T> J;
    memset(&J, 0, sizeof J);
    for (int b = 0; b < 2; ++b) {
        J.uu[b] = S + (5 * b + 0) * plane;
        J.ww[b] = S + (5 * b + 1) * plane;
        J.xx[b] = S + (5 * b + 2) * plane;
        J.zz[b] = S + (5 * b + 3) * plane;
        J.xz[b] = S + (5 * b + 4) * plane;
    }
    J.temp = C; J.b = C + plane; J.b1 = C + 2 * plane; J.ca = C + 3 * plane; J.cl = C + 4 * plane; J.cm1 = C + 5 * plane;
    J.c4 = c4;
    // b1 equal to b (as the wrappers make it): the fused kernels then read b for both
    if (fused && std::equal(hc.begin() + plane, hc.begin() + 2 * plane, hc.begin() + 2 * plane)) J.b1 = J.b;
    J.src = (const T*)src.p;
    J.trace_w = (T*)traces.p;
    J.trace_u = seismo_u ? (T*)traces.p + ntr : nullptr;
    J.snap = want_snap ? (T*)snaps.p : nullptr;
    J.img = nullptr;
    J.nzbc = d.nzbc; J.nxbc = d.nxbc; J.pitch = P; J.nz = d.nz; J.nx = d.nx; J.nbc = d.nbc; J.pt = d.pt;
    J.sz = d.sz; J.sx = d.sx; J.stype = (d.stype >= 1 && d.stype <= 5) ? d.stype : 0;
    J.gz = d.gz; J.gx = d.gx; J.dg = d.dg; J.ng = d.ng;
    J.nt = d.nt; J.in_wf = d.in_wf; J.start = 0; J.nsf = want_snap ? 4 : 0;
    J.dtx = (T)d.dtx; J.dt = T(0);
    if (fused) {
        std::vector<FdJob<T>> jobs{J};
        if ((rc = run_march<T>(st.s, jobs, false, d.nt, fd_time_block(), fd_tb_rows(), fd_tile_rows(), nullptr)))
            return rc;
    } else {
        if ((rc = strain.alloc(4 * plane * sizeof(T)))) return rc;
        SWD_HIP(hipMemsetAsync(strain.p, 0, 4 * plane * sizeof(T), st.s));
        T* w[4] = {(T*)strain.p, (T*)strain.p + plane, (T*)strain.p + 2 * plane, (T*)strain.p + 3 * plane};
        for (int it = 0; it < d.nt; ++it)
            if ((rc = plain_forward_step<T>(st.s, J, w, it))) return rc;
    }
    // results back to MATLAB layout
    std::vector<T> h(2 * ntr);
    SWD_HIP(hipMemcpyAsync(h.data(), traces.p, 2 * ntr * sizeof(T), hipMemcpyDeviceToHost, st.s));
    SWD_HIP(hipStreamSynchronize(st.s));
    for (int it = 0; it < d.nt; ++it)
        for (int g = 0; g < d.ng; ++g) seismo_w[(size_t)g * d.nt + it] = h[(size_t)it * d.ng + g];
    if (seismo_u)
        for (int it = 0; it < d.nt; ++it)
            for (int g = 0; g < d.ng; ++g) seismo_u[(size_t)g * d.nt + it] = h[ntr + (size_t)it * d.ng + g];
    if (want_snap) {
        std::vector<T> hs((size_t)nrec * 4 * splane);
        SWD_HIP(hipMemcpyAsync(hs.data(), snaps.p, hs.size() * sizeof(T), hipMemcpyDeviceToHost, st.s));
        SWD_HIP(hipStreamSynchronize(st.s));
        for (int k = 0; k < nrec; ++k)
            for (int f = 0; f < 4; ++f) {
                const T* src4 = hs.data() + ((size_t)k * 4 + f) * splane;
                T* dst = snap_out[f] + (size_t)k * splane;
                for (int iz = 0; iz < d.nz; ++iz)
                    for (int ix = 0; ix < d.nx; ++ix) dst[(size_t)ix * d.nz + iz] = src4[(size_t)iz * d.nx + ix];
            }
    }
    return kOk;
}

// ---------------------------------------------------------------------------
// one-shot adjoint (MEX-compatible)
// ---------------------------------------------------------------------------
struct AdjDesc {
    int nt, nzbc, nxbc, ng, nbc, gz, gx, dg, order, in_wf, nz, nx, pt;
    double dtx, dt;
};

template <typename T>
static int decode_adj(const T* iv, AdjDesc& d)
{
    d.nt = (int)iv[0]; d.nzbc = (int)iv[1]; d.nxbc = (int)iv[2]; d.dtx = (double)iv[3];
    d.ng = (int)iv[4]; d.nbc = (int)iv[5]; d.gz = (int)iv[6] - 1; d.gx = (int)iv[7] - 1; d.dg = (int)iv[8];
    d.order = (int)iv[9]; d.in_wf = (int)iv[10]; d.nz = (int)iv[11]; d.nx = (int)iv[12]; d.dt = (double)iv[13];
    d.pt = order_pt(d.order);
    if (d.pt < 0) { set_error("fd_order %d is not one of 22/24/26/28", d.order); return kErrArg; }
    if (d.nt <= 1 || d.nzbc <= 2 * d.pt || d.nxbc <= 2 * d.pt || d.ng <= 0 || d.dg <= 0 || d.in_wf <= 0) {
        set_error("bad grid/time descriptor"); return kErrArg;
    }
    if (d.nt % d.in_wf != 0) { set_error("nt (%d) must be a multiple of in_wf (%d)", d.nt, d.in_wf); return kErrArg; }
    if (d.gz < 0 || d.gz >= d.nzbc || d.gx < 0 || d.gx + (d.ng - 1) * d.dg >= d.nxbc) {
        set_error("receivers outside the padded grid"); return kErrArg;
    }
    if (d.pt + d.nz >= d.nzbc || d.nbc < 0 || d.nbc + d.nx > d.nxbc || d.nz <= 0 || d.nx <= 0) {
        set_error("interior nz x nx does not fit the padded grid"); return kErrArg;
    }
    return kOk;
}

template <typename T>
static int e2drtm_impl(const T* iv, const T* const coef_in[7], const T* adj, const T* const snap_in[4], T* img_out[3])
{
    if (!iv || !adj) { set_error("null argument"); return kErrArg; }
    for (int k = 0; k < 7; ++k)
        if (k != 3 && !coef_in[k]) { set_error("null coefficient array"); return kErrArg; }
    for (int k = 0; k < 4; ++k)
        if (!snap_in[k]) { set_error("null snapshot array"); return kErrArg; }
    for (int k = 0; k < 3; ++k)
        if (!img_out[k]) { set_error("null image array"); return kErrArg; }
    AdjDesc d;
    int rc = decode_adj(iv, d);
    if (rc) return rc;
    bool fused = !fd_force_plain() && d.order == 22 && d.gz >= 1 && d.gz <= d.nzbc - 2 && d.gx >= 1 &&
                 d.gx + (d.ng - 1) * d.dg <= d.nxbc - 2;
    const int P = grid_pitch(d.nxbc);
    const size_t plane = (size_t)d.nzbc * P;
    const int nrec = d.nt / d.in_wf;
    const size_t splane = (size_t)d.nz * d.nx;
    const size_t ntr = (size_t)d.nt * d.ng;

    Stream st;
    if ((rc = st.create())) return rc;
    DevBuf coefs, state, dadj, snaps, img;
    const int map[6] = {0, 5, 6, 1, 2, 4};
    std::vector<T> hc(6 * plane);
    for (int k = 0; k < 6; ++k) {
        pack_grid(coef_in[map[k]], d.nzbc, d.nxbc, P, hc.data() + k * plane);
        if (fused) zero_ring(hc.data() + k * plane, d.nzbc, d.nxbc, P);
    }
    const T* c4 = nullptr;
    if ((rc = upload_coefs(coefs, hc, plane, st.s, c4))) return rc;
    {
        std::vector<T> h(ntr);
        for (int it = 0; it < d.nt; ++it)
            for (int g = 0; g < d.ng; ++g) h[(size_t)it * d.ng + g] = adj[(size_t)g * d.nt + it];
        if ((rc = dadj.upload(h.data(), h.size() * sizeof(T), st.s))) return rc;
        SWD_HIP(hipStreamSynchronize(st.s));
    }
    {
        // the engine's snapshot layout [slot][row][x][fux, bwz, fuz + bwx]: the
        // imaging condition reads only these three (cpp:275-279), and every
        // kernel forms fuz + bwx in T, so the sums are the same as from planes
        std::vector<T> h((size_t)nrec * 3 * splane);
        for (int k = 0; k < nrec; ++k) {
            const size_t so = (size_t)k * splane;
            const T *fux = snap_in[0] + so, *fuz = snap_in[1] + so, *bwx = snap_in[2] + so, *bwz = snap_in[3] + so;
            for (int iz = 0; iz < d.nz; ++iz) {
                T* row = h.data() + ((size_t)k * d.nz + iz) * 3 * d.nx;
                for (int ix = 0; ix < d.nx; ++ix) {
                    const size_t c = (size_t)ix * d.nz + iz;
                    row[3 * ix] = fux[c];
                    row[3 * ix + 1] = bwz[c];
                    row[3 * ix + 2] = fuz[c] + bwx[c];
                }
            }
        }
        if ((rc = snaps.upload(h.data(), h.size() * sizeof(T), st.s))) return rc;
        SWD_HIP(hipStreamSynchronize(st.s));
    }
    DevBuf imgp;
    if ((rc = state.alloc(10 * plane * sizeof(T))) || (rc = img.alloc(3 * splane * sizeof(double))) ||
        (rc = imgp.alloc(3 * splane * sizeof(float))))
        return rc;
    SWD_HIP(hipMemsetAsync(state.p, 0, 10 * plane * sizeof(T), st.s));
    SWD_HIP(hipMemsetAsync(img.p, 0, 3 * splane * sizeof(double), st.s));
    SWD_HIP(hipMemsetAsync(imgp.p, 0, 3 * splane * sizeof(float), st.s));
    T* C = (T*)coefs.p;
    T* S = (T*)state.p;
    FdJob<T> J;
    memset(&J, 0, sizeof J);
    for (int b = 0; b < 2; ++b) {
        J.uu[b] = S + (5 * b + 0) * plane;
        J.ww[b] = S + (5 * b + 1) * plane;
        J.xx[b] = S + (5 * b + 2) * plane;
        J.zz[b] = S + (5 * b + 3) * plane;
        J.xz[b] = S + (5 * b + 4) * plane;
    }
    J.temp = C; J.b = C + plane; J.b1 = C + 2 * plane; J.ca = C + 3 * plane; J.cl = C + 4 * plane; J.cm1 = C + 5 * plane;
    J.c4 = c4;
    // b1 equal to b (as the wrappers make it): the fused kernels then read b for both
    if (fused && std::equal(hc.begin() + plane, hc.begin() + 2 * plane, hc.begin() + 2 * plane)) J.b1 = J.b;
    J.src = (const T*)dadj.p;
    J.snap = (T*)snaps.p;
    J.img = (double*)img.p;
    J.imgp = (float*)imgp.p;
    J.nzbc = d.nzbc; J.nxbc = d.nxbc; J.pitch = P; J.nz = d.nz; J.nx = d.nx; J.nbc = d.nbc; J.pt = d.pt;
    J.gz = d.gz; J.gx = d.gx; J.dg = d.dg; J.ng = d.ng;
    J.nt = d.nt; J.in_wf = d.in_wf; J.start = 0; J.nsf = 3;
    J.dtx = (T)d.dtx; J.dt = (T)d.dt;
    if (fused) {
        std::vector<FdJob<T>> jobs{J};
        if ((rc = run_march<T>(st.s, jobs, true, d.nt - 1, fd_time_block(), fd_tb_rows(), fd_tile_rows(), nullptr)))
            return rc;
    } else {
        for (int it = d.nt - 2; it >= 0; --it)
            if ((rc = plain_adjoint_step<T>(st.s, J, it))) return rc;
    }
    std::vector<double> h(3 * splane);
    SWD_HIP(hipMemcpyAsync(h.data(), img.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st.s));
    SWD_HIP(hipStreamSynchronize(st.s));
    for (int f = 0; f < 3; ++f)
        for (int iz = 0; iz < d.nz; ++iz)
            for (int ix = 0; ix < d.nx; ++ix)
                img_out[f][(size_t)ix * d.nz + iz] = (T)h[f * splane + (size_t)iz * d.nx + ix];
    return kOk;
}

// ---------------------------------------------------------------------------
// model preparation (staggerfd_eigen.m:37-107) and Ricker (ricker.m)
// ---------------------------------------------------------------------------
static void pad_rep(const double* p0, int nz, int nx, int nbc, int top, double* p)
{
    const int nzbc = nz + top + nbc, nxbc = nx + 2 * nbc;
    for (int j = 0; j < nxbc; ++j) {
        const int js = std::min(std::max(j - nbc, 0), nx - 1);
        for (int i = 0; i < nzbc; ++i) {
            const int is = std::min(std::max(i - top, 0), nz - 1);
            p[(size_t)j * nzbc + i] = p0[(size_t)js * nz + is];
        }
    }
}

// damp_circle (core/damp_circle.m:8-67) with a free surface, kappa factored
// out: side ramps (r/a)^2 on the model rows, bottom ramp, lower corners
// (dist/nbc)^2, top rows 0.  g is nzbc x nxbc column-major.
void damp_geometry(int nz, int nx, int nbc, double dx, int top, double* g)
{
    const int nzbc = nz + top + nbc, nxbc = nx + 2 * nbc;
    std::fill(g, g + (size_t)nzbc * nxbc, 0.0);
    const double a = (nbc - 1) * dx;
    std::vector<double> d1(nbc);
    for (int k = 0; k < nbc; ++k) { const double r = k * dx / a; d1[k] = r * r; }
    auto G = [&](int iz, int ix) -> double& { return g[(size_t)(ix - 1) * nzbc + (iz - 1)]; };
    for (int iz = 1 + top; iz <= nz + top; ++iz)
        for (int k = 1; k <= nbc; ++k) { G(iz, k) = d1[nbc - k]; G(iz, nx + nbc + k) = d1[k - 1]; }
    for (int ix = 1 + nbc; ix <= nx + nbc; ++ix)
        for (int k = 1; k <= nbc; ++k) G(nzbc - nbc + k, ix) = d1[k - 1];
    for (int iz = nz + top + 1; iz <= nzbc; ++iz) {
        const double dz = iz - nz - top;
        for (int ix = 1; ix <= nbc; ++ix) {
            const double dxl = ix - nbc - 1;
            const double dist = std::sqrt(dxl * dxl + dz * dz);
            G(iz, ix) = (dist / nbc) * (dist / nbc);
        }
        for (int ix = nx + nbc + 1; ix <= nxbc; ++ix) {
            const double dxr = ix - nbc - nx;
            const double dist = std::sqrt(dxr * dxr + dz * dz);
            G(iz, ix) = (dist / nbc) * (dist / nbc);
        }
    }
}

// kappa = 3 vmin ln(1e7) / (2 a), a = (nbc - 1) dx (core/damp_circle.m)
double damp_kappa(double vmin, int nbc, double dx)
{
    const double a = (nbc - 1) * dx;
    return 3.0 * vmin * std::log(10000000.0) / (2.0 * a);
}

int prepare_model(const double* vp, const double* vs, int nz, int nx, int nbc, double dx, double dt, int fd_order,
                  double* temp, double* ca, double* cl, double* cm, double* cm1, double* b, double* b1)
{
    if (!vp || !vs || !temp || !ca || !cl || !cm || !cm1 || !b || !b1) { set_error("null argument"); return kErrArg; }
    if (nz < 1 || nx < 1 || nbc < 2 || !(dx > 0) || !(dt > 0)) { set_error("bad model size"); return kErrArg; }
    if (order_pt(fd_order) < 0) { set_error("fd_order %d is not one of 22/24/26/28", fd_order); return kErrArg; }
    const int top = (fd_order - 20) / 2 + 1;      // MATLAB pad_top with a free surface
    const int nzbc = nz + top + nbc, nxbc = nx + 2 * nbc;
    const size_t n0 = (size_t)nz * nx, n = (size_t)nzbc * nxbc;
    std::vector<double> ca0(n0), cm0(n0), cl0(n0);
    double vmin = INFINITY;
    for (size_t k = 0; k < n0; ++k) {
        ca0[k] = vp[k] * vp[k];                   // calparam with den = 1
        cm0[k] = vs[k] * vs[k];
        cl0[k] = ca0[k] - 2 * cm0[k];
        vmin = std::min(vmin, std::min(vp[k], vs[k]));
    }
    pad_rep(cm0.data(), nz, nx, nbc, top, cm);
    pad_rep(cl0.data(), nz, nx, nbc, top, cl);
    pad_rep(ca0.data(), nz, nx, nbc, top, ca);
    // damp_circle with isfs = 1: damp = kappa(vmin) * geometry factor
    std::vector<double> geo(n);
    damp_geometry(nz, nx, nbc, dx, top, geo.data());
    const double kappa = damp_kappa(vmin, nbc, dx);
    const double dtx = dt / dx;
    for (size_t k = 0; k < n; ++k) {
        const double damp = kappa * geo[k];
        temp[k] = 1 - damp * dt;
        b[k] = dtx;
    }
    const int fr = top - 1;                       // free-surface row (0-based)
    for (int j = 0; j < nxbc; ++j) {
        const size_t o = (size_t)j * nzbc + fr;
        ca[o] = 2 * cm[o];
        cl[o] = 0.0;
        b[o] = 2 * b[o];
    }
    std::memcpy(b1, b, n * sizeof(double));
    std::memcpy(cm1, cm, n * sizeof(double));
    return kOk;
}

int ricker(double f, double dt, int nt, double* w)
{
    if (!w || !(f > 0) || !(dt > 0)) { set_error("bad Ricker arguments"); return kErrArg; }
    double nw = 2.2 / f / dt;
    nw = 2 * std::floor(nw / 2) + 1;
    const int inw = (int)nw;
    const double nc = std::floor(nw / 2);
    if (nt < inw) { set_error("nt is smaller than condition!"); return kErrArg; }
    for (int k = 0; k < nt; ++k) w[k] = 0;
    for (int k = 1; k <= inw; ++k) {
        const double al = (nc - k + 1) * f * dt * M_PI;
        const double be = al * al;
        w[k - 1] = (1. - be * 2) * std::exp(-be);
    }
    return kOk;
}

}  // namespace swd

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace swd;

extern "C" {

const char* swd_version(void) { return "swd-mi355x 0.1 (gfx950)"; }
const char* swd_last_error(void) { return last_error(); }

int swd_device_info(int* n_devices, char* arch, int arch_len)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    if (n_devices) *n_devices = n;
    if (arch && arch_len > 0) {
        arch[0] = 0;
        if (n > 0) {
            hipDeviceProp_t pr;
            if (hipGetDeviceProperties(&pr, 0) == hipSuccess) snprintf(arch, arch_len, "%s", pr.gcnArchName);
        }
    }
    if (n == 0) { set_error("no HIP device"); return kErrNoDevice; }
    return kOk;
}

int swd_set_fd_path(int mode) { return g_fd_path.exchange(mode ? 1 : 0); }

int swd_set_fd_time_block(int nt_block, int rows)
{
    if (nt_block < 1 || nt_block > 6 || rows < 0) {
        set_error("time block must be 1..6 steps, rows >= 0 (0: automatic)");
        return kErrArg;
    }
    set_fd_time_block(nt_block, rows);
    return kOk;
}

int swd_staggerfd_f32(const float* iv, const float* temp, const float* ca, const float* cl, const float* cm,
                      const float* cm1, const float* b, const float* b1, const float* s, float* seismo_u,
                      float* seismo_w, float* fux, float* fuz, float* bwx, float* bwz)
{
    const float* c[7] = {temp, ca, cl, cm, cm1, b, b1};
    float* sn[4] = {fux, fuz, bwx, bwz};
    return staggerfd_impl<float>(iv, c, s, seismo_u, seismo_w, sn);
}

int swd_staggerfd_f64(const double* iv, const double* temp, const double* ca, const double* cl, const double* cm,
                      const double* cm1, const double* b, const double* b1, const double* s, double* seismo_u,
                      double* seismo_w, double* fux, double* fuz, double* bwx, double* bwz)
{
    const double* c[7] = {temp, ca, cl, cm, cm1, b, b1};
    double* sn[4] = {fux, fuz, bwx, bwz};
    return staggerfd_impl<double>(iv, c, s, seismo_u, seismo_w, sn);
}

int swd_e2drtm_f32(const float* iv, const float* temp, const float* ca, const float* cl, const float* cm,
                   const float* cm1, const float* b, const float* b1, const float* adj, const float* fux,
                   const float* fuz, const float* bwx, const float* bwz, float* cl_img, float* cm_img,
                   float* illum_div)
{
    const float* c[7] = {temp, ca, cl, cm, cm1, b, b1};
    const float* sn[4] = {fux, fuz, bwx, bwz};
    float* im[3] = {cl_img, cm_img, illum_div};
    return e2drtm_impl<float>(iv, c, adj, sn, im);
}

int swd_e2drtm_f64(const double* iv, const double* temp, const double* ca, const double* cl, const double* cm,
                   const double* cm1, const double* b, const double* b1, const double* adj, const double* fux,
                   const double* fuz, const double* bwx, const double* bwz, double* cl_img, double* cm_img,
                   double* illum_div)
{
    const double* c[7] = {temp, ca, cl, cm, cm1, b, b1};
    const double* sn[4] = {fux, fuz, bwx, bwz};
    double* im[3] = {cl_img, cm_img, illum_div};
    return e2drtm_impl<double>(iv, c, adj, sn, im);
}

int swd_prepare_model(const double* vp, const double* vs, int nz, int nx, int nbc, double dx, double dt,
                      int fd_order, double* temp, double* ca, double* cl, double* cm, double* cm1, double* b,
                      double* b1)
{
    return prepare_model(vp, vs, nz, nx, nbc, dx, dt, fd_order, temp, ca, cl, cm, cm1, b, b1);
}

int swd_ricker(double f, double dt, int nt, double* w) { return ricker(f, dt, nt, w); }

}  // extern "C"
