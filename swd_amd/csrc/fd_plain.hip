// "Plain" SWD FD path for gfx950: one kernel launch per statement of the
// reference time loops, for one shot, on the internal x-fastest layout.
//
// It covers every fd_order (22/24/26/28) and source type (1..5) the MEX
// accepts, and sources/receivers anywhere on the grid, so it is the general
// fallback of the fused march kernels.  It is also the unfused "one launch
// per Eigen statement" GPU comparator the batched engine is measured against.
//
// Statement map:
//   forward  core/eigen_staggerfd_single.cpp  :224-255 vel, :286-300 src,
//            :302-329 strain, :334-336 stress, :340 free surface,
//            :342-345 receivers, :346-351 snapshots
//   adjoint  core/eigen_e2drtm_single.cpp     :182-235 vel, :238-242 src,
//            :245-268 stress, :270 free surface, :274-280 imaging
#include "swd_internal.h"

namespace swd {

template <typename T>
struct PlainGrid {
    int nzbc, nxbc, pitch, pt;
    T S[4];  // stencil coefficients (order 22: S[0] = 1)
};

#define PG_IJ                                                     \
    const int j = blockIdx.x * blockDim.x + threadIdx.x + g.pt;   \
    const int i = blockIdx.y * blockDim.y + threadIdx.y + g.pt;   \
    if (j >= g.nxbc - g.pt || i >= g.nzbc - g.pt) return;         \
    const size_t P = (size_t)g.pitch;                             \
    auto A = [&](const T* a, int ii, int jj) -> T { return a[(size_t)ii * P + jj]; };

// velocity statements (uu then ww as separate launches)
template <typename T>
__global__ void plain_vel_uu(PlainGrid<T> g, T* uu, const T* xx, const T* xz, const T* temp, const T* b)
{
    PG_IJ
    T acc;
    if (g.pt == 1) {
        acc = A(xx, i, j + 1) - A(xx, i, j) + A(xz, i, j) - A(xz, i - 1, j);
    } else {
        acc = T(0);
        for (int q = 1; q <= g.pt; ++q) acc += g.S[q - 1] * (A(xx, i, j + q - 1) - A(xx, i, j - q));
        for (int q = 1; q <= g.pt; ++q) acc += g.S[q - 1] * (A(xz, i + q - 1, j) - A(xz, i - q, j));
    }
    const size_t o = (size_t)i * P + j;
    uu[o] = temp[o] * uu[o] + b[o] * acc;
}

template <typename T>
__global__ void plain_vel_ww(PlainGrid<T> g, T* ww, const T* xz, const T* zz, const T* temp, const T* b1)
{
    PG_IJ
    T acc;
    if (g.pt == 1) {
        acc = A(xz, i, j) - A(xz, i, j - 1) + A(zz, i + 1, j) - A(zz, i, j);
    } else {
        acc = T(0);
        for (int q = 1; q <= g.pt; ++q) acc += g.S[q - 1] * (A(xz, i, j + q) - A(xz, i, j - q + 1));
        for (int q = 1; q <= g.pt; ++q) acc += g.S[q - 1] * (A(zz, i + q, j) - A(zz, i - q + 1, j));
    }
    const size_t o = (size_t)i * P + j;
    ww[o] = temp[o] * ww[o] + b1[o] * acc;
}

template <typename T>
__global__ void plain_source(PlainGrid<T> g, int stype, int sz, int sx, const T* s, int it, T* uu, T* ww, T* xx, T* zz)
{
    const size_t P = (size_t)g.pitch;
    const T v = s[it];
    const size_t o = (size_t)sz * P + sx;
    if (stype == 1) { xx[o] += v; zz[o] += v; }
    else if (stype == 2) { uu[o] += v; uu[o - P] -= v; ww[o] += v; ww[o + 1] += v; }
    else if (stype == 3) { uu[o] += v; ww[o] += v; }
    else if (stype == 4) { zz[o] += v; }
    else if (stype == 5) { ww[o] = v; }
}

// which = 0 fux, 1 fuz, 2 bwx, 3 bwz (one launch per strain statement)
template <typename T>
__global__ void plain_strain(PlainGrid<T> g, int which, T* out, const T* uu, const T* ww)
{
    PG_IJ
    T a = T(0);
    if (g.pt == 1) {
        if (which == 0) a = A(uu, i, j) - A(uu, i, j - 1);
        else if (which == 1) a = A(uu, i + 1, j) - A(uu, i, j);
        else if (which == 2) a = A(ww, i, j + 1) - A(ww, i, j);
        else a = A(ww, i, j) - A(ww, i - 1, j);
    } else {
        for (int q = 1; q <= g.pt; ++q) {
            if (which == 0) a += g.S[q - 1] * (A(uu, i, j + q) - A(uu, i, j - q + 1));
            else if (which == 1) a += g.S[q - 1] * (A(uu, i + q, j) - A(uu, i - q + 1, j));
            else if (which == 2) a += g.S[q - 1] * (A(ww, i, j + q - 1) - A(ww, i, j - q));
            else a += g.S[q - 1] * (A(ww, i + q - 1, j) - A(ww, i - q, j));
        }
    }
    out[(size_t)i * P + j] = a;
}

// full-array stress statements: which = 0 xx, 1 zz, 2 xz
template <typename T>
__global__ void plain_stress(PlainGrid<T> g, int which, T* f, const T* temp, const T* ca, const T* cl,
                             const T* cm1, const T* fux, const T* fuz, const T* bwx, const T* bwz, T dtx)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (j >= g.nxbc || i >= g.nzbc) return;
    const size_t o = (size_t)i * g.pitch + j;
    if (which == 0) f[o] = temp[o] * f[o] + (ca[o] * fux[o] + cl[o] * bwz[o]) * dtx;
    else if (which == 1) f[o] = temp[o] * f[o] + (ca[o] * bwz[o] + cl[o] * fux[o]) * dtx;
    else f[o] = temp[o] * f[o] + (cm1[o] * (fuz[o] + bwx[o])) * dtx;
}

template <typename T>
__global__ void plain_zero_row(T* zz, int row, int pitch, int nxbc)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nxbc) zz[(size_t)row * pitch + j] = T(0);
}

template <typename T>
__global__ void plain_receivers(const T* f, int pitch, int gz, int gx, int dg, int ng, int it, T* trace)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ng) trace[(size_t)it * ng + k] = f[(size_t)gz * pitch + gx + k * dg];
}

// snapshot copy of one strain field into slot (it/in_wf), field index fi of nsf
template <typename T>
__global__ void plain_snapshot(const T* f, int pitch, int pt, int nbc, int nz, int nx, T* snap, int slot, int nsf, int fi)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= nx || iz >= nz) return;
    const size_t plane = (size_t)nz * nx;
    const T v = f[(size_t)(pt + 1 + iz) * pitch + nbc + ix];
    if (nsf == 3) snap[(size_t)slot * 3 * plane + ((size_t)iz * nx + ix) * 3 + fi] = v;   // [row][x][3]
    else snap[((size_t)slot * nsf + fi) * plane + (size_t)iz * nx + ix] = v;
}

// 3-field snapshot layout: field 2 holds fuz + bwx
template <typename T>
__global__ void plain_snapshot_sum(const T* fuz, const T* bwx, int pitch, int pt, int nbc, int nz, int nx, T* snap,
                                   int slot)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= nx || iz >= nz) return;
    const size_t plane = (size_t)nz * nx;
    const size_t o = (size_t)(pt + 1 + iz) * pitch + nbc + ix;
    snap[(size_t)slot * 3 * plane + ((size_t)iz * nx + ix) * 3 + 2] = fuz[o] + bwx[o];
}

// adjoint velocity statements
template <typename T>
__global__ void plain_adj_uu(PlainGrid<T> g, T* uu, const T* xx, const T* zz, const T* xz, const T* temp,
                             const T* b, const T* ca, const T* cl, const T* cm1)
{
    PG_IJ
    T acc;
    if (g.pt == 1) {
        acc = (A(ca, i, j) * A(xx, i, j) - A(ca, i, j - 1) * A(xx, i, j - 1)) +
              (A(cl, i, j) * A(zz, i, j) - A(cl, i, j - 1) * A(zz, i, j - 1)) +
              (A(cm1, i, j) * A(xz, i, j) - A(cm1, i - 1, j) * A(xz, i - 1, j));
    } else {
        acc = T(0);
        for (int q = 1; q <= g.pt; ++q)
            acc += g.S[q - 1] * (A(ca, i, j + q - 1) * A(xx, i, j + q - 1) - A(ca, i, j - q) * A(xx, i, j - q));
        for (int q = 1; q <= g.pt; ++q)
            acc += g.S[q - 1] * (A(cl, i, j + q - 1) * A(zz, i, j + q - 1) - A(cl, i, j - q) * A(zz, i, j - q));
        for (int q = 1; q <= g.pt; ++q)
            acc += g.S[q - 1] * (A(cm1, i + q - 1, j) * A(xz, i + q - 1, j) - A(cm1, i - q, j) * A(xz, i - q, j));
    }
    const size_t o = (size_t)i * P + j;
    uu[o] = temp[o] * uu[o] - b[o] * acc;
}

template <typename T>
__global__ void plain_adj_ww(PlainGrid<T> g, T* ww, const T* xx, const T* zz, const T* xz, const T* temp,
                             const T* b1, const T* ca, const T* cl, const T* cm1)
{
    PG_IJ
    T acc;
    if (g.pt == 1) {
        acc = (A(cl, i + 1, j) * A(xx, i + 1, j) - A(cl, i, j) * A(xx, i, j)) +
              (A(ca, i + 1, j) * A(zz, i + 1, j) - A(ca, i, j) * A(zz, i, j)) +
              (A(cm1, i, j + 1) * A(xz, i, j + 1) - A(cm1, i, j) * A(xz, i, j));
    } else {
        acc = T(0);
        for (int q = 1; q <= g.pt; ++q)
            acc += g.S[q - 1] * (A(cl, i + q, j) * A(xx, i + q, j) - A(cl, i - q + 1, j) * A(xx, i - q + 1, j));
        for (int q = 1; q <= g.pt; ++q)
            acc += g.S[q - 1] * (A(ca, i + q, j) * A(zz, i + q, j) - A(ca, i - q + 1, j) * A(zz, i - q + 1, j));
        for (int q = 1; q <= g.pt; ++q)
            acc += g.S[q - 1] * (A(cm1, i, j + q) * A(xz, i, j + q) - A(cm1, i, j - q + 1) * A(xz, i, j - q + 1));
    }
    const size_t o = (size_t)i * P + j;
    ww[o] = temp[o] * ww[o] - b1[o] * acc;
}

template <typename T>
__global__ void plain_adj_inject(T* ww, int pitch, int gz, int gx, int dg, int ng, const T* adj, int it, T dt)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ng) ww[(size_t)gz * pitch + gx + k * dg] += dt * adj[(size_t)it * ng + k];
}

// adjoint stress statements: which = 0 xx, 1 zz, 2 xz (block only)
template <typename T>
__global__ void plain_adj_stress(PlainGrid<T> g, int which, T* f, const T* uu, const T* ww, const T* temp, T dtx)
{
    PG_IJ
    const size_t o = (size_t)i * P + j;
    if (g.pt == 1) {
        if (which == 0) f[o] = temp[o] * f[o] - dtx * (A(uu, i, j + 1) - A(uu, i, j));
        else if (which == 1) f[o] = temp[o] * f[o] - dtx * (A(ww, i, j) - A(ww, i - 1, j));
        else f[o] = temp[o] * f[o] - dtx * (A(uu, i + 1, j) - A(uu, i, j)) - dtx * (A(ww, i, j) - A(ww, i, j - 1));
    } else {
        T a = T(0), c = T(0);
        if (which == 0) {
            for (int q = 1; q <= g.pt; ++q) a += g.S[q - 1] * (A(uu, i, j + q) - A(uu, i, j - q + 1));
            f[o] = temp[o] * f[o] - dtx * a;
        } else if (which == 1) {
            for (int q = 1; q <= g.pt; ++q) a += g.S[q - 1] * (A(ww, i + q - 1, j) - A(ww, i - q, j));
            f[o] = temp[o] * f[o] - dtx * a;
        } else {
            for (int q = 1; q <= g.pt; ++q) a += g.S[q - 1] * (A(uu, i + q, j) - A(uu, i - q + 1, j));
            for (int q = 1; q <= g.pt; ++q) c += g.S[q - 1] * (A(ww, i, j + q - 1) - A(ww, i, j - q));
            f[o] = temp[o] * f[o] - dtx * a - dtx * c;
        }
    }
}

// imaging statements (which = 0 cl_img, 1 cm_img, 2 illum), double accumulators
template <typename T>
__global__ void plain_imaging(int which, const T* xx, const T* zz, const T* xz, int pitch, int pt, int nbc,
                              int nz, int nx, const T* snap, int slot, int nsf, double* img)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= nx || iz >= nz) return;
    const size_t plane = (size_t)nz * nx;
    const size_t io = (size_t)iz * nx + ix;
    const T* sn = snap + (size_t)slot * nsf * plane + io;
    const T* sc = snap + (size_t)slot * 3 * plane + ((size_t)iz * nx + ix) * 3;
    double Fux, Fwz, Fsh;
    if (nsf == 3) { Fux = sc[0]; Fwz = sc[1]; Fsh = sc[2]; }
    else { Fux = sn[0]; Fwz = sn[3 * plane]; Fsh = (double)(sn[plane] + sn[2 * plane]); }
    const size_t o = (size_t)(pt + 1 + iz) * pitch + nbc + ix;
    const double X = xx[o], Z = zz[o], XZ = xz[o];
    if (which == 0) img[io] += (Fux + Fwz) * (X + Z);
    else if (which == 1) img[plane + io] += 2.0 * (X * Fux + Z * Fwz + 0.5 * XZ * Fsh);
    else img[2 * plane + io] += (Fux + Fwz) * (Fux + Fwz);
}

#undef PG_IJ

// ---------------------------------------------------------------------------
// host drivers for one shot (state in J.*[0]; strain scratch w[0..3])
// ---------------------------------------------------------------------------
template <typename T>
static PlainGrid<T> plain_grid(const FdJob<T>& J)
{
    PlainGrid<T> g;
    g.nzbc = J.nzbc;
    g.nxbc = J.nxbc;
    g.pitch = J.pitch;
    g.pt = J.pt;
    const float s24[2] = {1.1250f, -0.0416666667f};
    const float s26[3] = {1.17187f, -6.51042E-2f, 4.68750E-3f};
    const float s28[4] = {1.19629f, -7.97526E-2f, 9.57031E-3f, -6.97545E-4f};
    for (int q = 0; q < 4; ++q) g.S[q] = T(0);
    if (J.pt == 1) g.S[0] = T(1);
    if (J.pt == 2) for (int q = 0; q < 2; ++q) g.S[q] = T(s24[q]);
    if (J.pt == 3) for (int q = 0; q < 3; ++q) g.S[q] = T(s26[q]);
    if (J.pt == 4) for (int q = 0; q < 4; ++q) g.S[q] = T(s28[q]);
    return g;
}

template <typename T>
int plain_forward_step(hipStream_t st, const FdJob<T>& J, T* const w[4], int it)
{
    const PlainGrid<T> g = plain_grid(J);
    const dim3 blk(64, 4);
    const dim3 grdB(div_up(J.nxbc - 2 * J.pt, 64), div_up(J.nzbc - 2 * J.pt, 4));
    const dim3 grdF(div_up(J.nxbc, 64), div_up(J.nzbc, 4));
    T *uu = J.uu[0], *ww = J.ww[0], *xx = J.xx[0], *zz = J.zz[0], *xz = J.xz[0];
    plain_vel_uu<T><<<grdB, blk, 0, st>>>(g, uu, xx, xz, J.temp, J.b);
    plain_vel_ww<T><<<grdB, blk, 0, st>>>(g, ww, xz, zz, J.temp, J.b1);
    if (J.stype >= 1 && J.stype <= 5)
        plain_source<T><<<1, 1, 0, st>>>(g, J.stype, J.sz, J.sx, J.src, it, uu, ww, xx, zz);
    for (int k = 0; k < 4; ++k) plain_strain<T><<<grdB, blk, 0, st>>>(g, k, w[k], uu, ww);
    plain_stress<T><<<grdF, blk, 0, st>>>(g, 0, xx, J.temp, J.ca, J.cl, J.cm1, w[0], w[1], w[2], w[3], J.dtx);
    plain_stress<T><<<grdF, blk, 0, st>>>(g, 1, zz, J.temp, J.ca, J.cl, J.cm1, w[0], w[1], w[2], w[3], J.dtx);
    plain_stress<T><<<grdF, blk, 0, st>>>(g, 2, xz, J.temp, J.ca, J.cl, J.cm1, w[0], w[1], w[2], w[3], J.dtx);
    plain_zero_row<T><<<div_up(J.nxbc, 256), 256, 0, st>>>(zz, J.pt, J.pitch, J.nxbc);
    plain_receivers<T><<<div_up(J.ng, 256), 256, 0, st>>>(ww, J.pitch, J.gz, J.gx, J.dg, J.ng, it, J.trace_w);
    if (J.trace_u)
        plain_receivers<T><<<div_up(J.ng, 256), 256, 0, st>>>(uu, J.pitch, J.gz, J.gx, J.dg, J.ng, it, J.trace_u);
    if (J.nsf > 0 && it % J.in_wf == 0) {
        const dim3 grdI(div_up(J.nx, 64), div_up(J.nz, 4));
        const int slot = it / J.in_wf;
        if (J.nsf == 4) {
            for (int k = 0; k < 4; ++k)
                plain_snapshot<T><<<grdI, blk, 0, st>>>(w[k], J.pitch, J.pt, J.nbc, J.nz, J.nx, J.snap, slot, 4, k);
        } else {
            plain_snapshot<T><<<grdI, blk, 0, st>>>(w[0], J.pitch, J.pt, J.nbc, J.nz, J.nx, J.snap, slot, 3, 0);
            plain_snapshot<T><<<grdI, blk, 0, st>>>(w[3], J.pitch, J.pt, J.nbc, J.nz, J.nx, J.snap, slot, 3, 1);
            plain_snapshot_sum<T><<<grdI, blk, 0, st>>>(w[1], w[2], J.pitch, J.pt, J.nbc, J.nz, J.nx, J.snap, slot);
        }
    }
    SWD_HIP(hipGetLastError());
    return kOk;
}

template <typename T>
int plain_adjoint_step(hipStream_t st, const FdJob<T>& J, int it)
{
    const PlainGrid<T> g = plain_grid(J);
    const dim3 blk(64, 4);
    const dim3 grdB(div_up(J.nxbc - 2 * J.pt, 64), div_up(J.nzbc - 2 * J.pt, 4));
    T *uu = J.uu[0], *ww = J.ww[0], *xx = J.xx[0], *zz = J.zz[0], *xz = J.xz[0];
    plain_adj_uu<T><<<grdB, blk, 0, st>>>(g, uu, xx, zz, xz, J.temp, J.b, J.ca, J.cl, J.cm1);
    plain_adj_ww<T><<<grdB, blk, 0, st>>>(g, ww, xx, zz, xz, J.temp, J.b1, J.ca, J.cl, J.cm1);
    plain_adj_inject<T><<<div_up(J.ng, 256), 256, 0, st>>>(ww, J.pitch, J.gz, J.gx, J.dg, J.ng, J.src, it, J.dt);
    plain_adj_stress<T><<<grdB, blk, 0, st>>>(g, 0, xx, uu, ww, J.temp, J.dtx);
    plain_adj_stress<T><<<grdB, blk, 0, st>>>(g, 1, zz, uu, ww, J.temp, J.dtx);
    plain_adj_stress<T><<<grdB, blk, 0, st>>>(g, 2, xz, uu, ww, J.temp, J.dtx);
    plain_zero_row<T><<<div_up(J.nxbc, 256), 256, 0, st>>>(zz, J.pt, J.pitch, J.nxbc);
    if (it % J.in_wf == 0) {
        const dim3 grdI(div_up(J.nx, 64), div_up(J.nz, 4));
        for (int k = 0; k < 3; ++k)
            plain_imaging<T><<<grdI, blk, 0, st>>>(k, xx, zz, xz, J.pitch, J.pt, J.nbc, J.nz, J.nx, J.snap,
                                                   it / J.in_wf, J.nsf, J.img);
    }
    SWD_HIP(hipGetLastError());
    return kOk;
}

template int plain_forward_step<float>(hipStream_t, const FdJob<float>&, float* const[4], int);
template int plain_forward_step<double>(hipStream_t, const FdJob<double>&, double* const[4], int);
template int plain_adjoint_step<float>(hipStream_t, const FdJob<float>&, int);
template int plain_adjoint_step<double>(hipStream_t, const FdJob<double>&, int);

}  // namespace swd
