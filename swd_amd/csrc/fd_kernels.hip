// SWD elastic staggered-grid FD kernels for gfx950 (MI355X).
//
// Two kernel families implement the same time steps:
//
//  * fused "march" kernels (fd_order 22, the configuration SWD drives):
//    one launch advances every shot of a batch by one full time step
//    (velocity half-step, source, strain, stress, free surface, receivers,
//    snapshot / imaging).  Each wavefront owns 62 consecutive columns
//    (lanes 1..62; lanes 0 and 63 are a one-column halo that recomputes the
//    neighbouring velocities) and marches down a band of rows, keeping the
//    rows it needs in registers and exchanging x-neighbours with lane
//    shuffles.  All loads/stores are row-contiguous (x fastest) and
//    coalesced; state is ping-ponged so bands never race.
//
//  * "plain" kernels: one kernel per statement of the reference time loop,
//    any fd_order in {22,24,26,28} and any source type, one shot per launch.
//    They are the general fallback and the unfused comparator.
//
// Forward step, core/eigen_staggerfd_single.cpp:207-357 (0-based, block B =
// rows/cols pt..n-pt-1):
//   uu = temp*uu + b *(xx[i,j+1]-xx[i,j] + xz[i,j]-xz[i-1,j])          (:225)
//   ww = temp*ww + b1*(xz[i,j]-xz[i,j-1] + zz[i+1,j]-zz[i,j])          (:227)
//   ww[sz,sx] = s[it]  (type 5; types 1-4 add, :286-300)
//   fux = uu[i,j]-uu[i,j-1]  fuz = uu[i+1,j]-uu[i,j]                   (:303-304)
//   bwx = ww[i,j+1]-ww[i,j]  bwz = ww[i,j]-ww[i-1,j]                   (:305-306)
//   xx = temp*xx + (ca*fux + cl*bwz)*dtx; zz = temp*zz + (ca*bwz + cl*fux)*dtx;
//   xz = temp*xz + cm1*(fuz+bwx)*dtx;  zz[pt,:] = 0                    (:334-340)
// Adjoint step, core/eigen_e2drtm_single.cpp:181-283 (it = nt-2 .. 0).
#include <type_traits>

#include "fd_cell.h"
#include "swd_internal.h"

namespace swd {

template <typename T>
__device__ __forceinline__ T ld_row(const T* __restrict__ a, int r, size_t P, int xs, bool pred)
{
    return pred ? a[(size_t)r * P + xs] : T(0);
}

// Snapshots and image sums stream through non-temporal accesses, as in the
// time-blocked kernels: they must not evict the state the next launch reads.
template <typename T>
__device__ __forceinline__ T stream_ld(const T* p) { return __builtin_nontemporal_load(p); }
template <typename T>
__device__ __forceinline__ void stream_st(T* p, T v) { __builtin_nontemporal_store(v, p); }

template <typename T>
__device__ __forceinline__ T shfl_from_left(T v) { return __shfl_up(v, 1, kMarchLanes); }
template <typename T>
__device__ __forceinline__ T shfl_from_right(T v) { return __shfl_down(v, 1, kMarchLanes); }

// ---------------------------------------------------------------------------
// fused forward step
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kMarchLanes* kMarchWaves) void fd_fwd_march(
    const FdJob<T>* __restrict__ jobs, const BlockMap* __restrict__ bmap, int step, int launch)
{
    const BlockMap bm = bmap[blockIdx.x];
    const FdJob<T>& J = jobs[bm.job];
    const int it = step - J.start;
    if (it < 0 || it >= J.nt) return;
    const int rd = (launch - J.launch0) & 1, wr = rd ^ 1;  // ping-pong buffers of this launch
    const int lane = threadIdx.x & (kMarchLanes - 1);
    const int tile = bm.tile0 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar row loop
    if (tile >= J.tiles_x * J.tiles_z) return;
    const int tix = tile % J.tiles_x, tiz = tile / J.tiles_x;
    const int nzbc = J.nzbc, nxbc = J.nxbc;
    const size_t P = (size_t)J.pitch;
    const int x = tix * kMarchOut + lane;
    const bool xload = x < nxbc;
    const bool xin = x >= 1 && x <= nxbc - 2;
    const bool out = xin && lane >= 1 && lane <= kMarchOut;
    const int z0 = 1 + tiz * J.tz;
    const int z1 = min(z0 + J.tz, nzbc - 1);
    const int xs = xload ? x : 0;

    const T* __restrict__ uu0 = J.uu[rd];
    const T* __restrict__ ww0 = J.ww[rd];
    const T* __restrict__ xx0 = J.xx[rd];
    const T* __restrict__ zz0 = J.zz[rd];
    const T* __restrict__ xz0 = J.xz[rd];
    T* __restrict__ uu1 = J.uu[wr];
    T* __restrict__ ww1 = J.ww[wr];
    T* __restrict__ xx1 = J.xx[wr];
    T* __restrict__ zz1 = J.zz[wr];
    T* __restrict__ xz1 = J.xz[wr];
    const T* __restrict__ temp = J.temp;
    const T* __restrict__ bb = J.b;
    const T* __restrict__ bb1 = J.b1;
    const T dtx = J.dtx;
    const T sv = J.src[it];
    const int sz = J.sz, sx = J.sx, stype = J.stype;

    auto ld = [&](const T* __restrict__ a, int r) -> T { return xload ? a[(size_t)r * P + xs] : T(0); };
    auto ldin = [&](const T* __restrict__ a, int r) -> T { return xin ? a[(size_t)r * P + xs] : T(0); };
    // velocity after the half-step and the velocity-type sources (types 2, 3, 5)
    auto fix_u = [&](T u, int r) -> T {
        if (stype == 2 && x == sx) { if (r == sz) u += sv; if (r == sz - 1) u -= sv; }
        if (stype == 3 && x == sx && r == sz) u += sv;
        return u;
    };
    auto fix_w = [&](T w, int r) -> T {
        if (r == sz) {
            if (stype == 5 && x == sx) w = sv;
            if (stype == 2 && (x == sx || x == sx + 1)) w += sv;
            if (stype == 3 && x == sx) w += sv;
        }
        return w;
    };

    // prologue: old stress rows z0-1 and z0, ww_new(z0-1), uu_new(z0)
    int r = z0;
    T xx_c = ld(xx0, r), zz_c = ld(zz0, r), xz_c = ld(xz0, r);
    const T xz_m = ld(xz0, r - 1), zz_m = ld(zz0, r - 1);
    T ww_m = T(0);
    {
        const T xz_ml = shfl_from_left(xz_m);
        if (r - 1 >= 1 && xin)
            ww_m = fix_w(cell_vel(temp[(size_t)(r - 1) * P + xs], ww0[(size_t)(r - 1) * P + xs],
                                  bb1[(size_t)(r - 1) * P + xs], xz_m - xz_ml + zz_c - zz_m),
                         r - 1);
    }
    T uu_c;
    {
        const T xx_cr = shfl_from_right(xx_c);
        uu_c = xin ? fix_u(cell_vel(temp[(size_t)r * P + xs], uu0[(size_t)r * P + xs], bb[(size_t)r * P + xs],
                                    xx_cr - xx_c + xz_c - xz_m),
                           r)
                   : T(0);
    }

    const bool snap_step = J.nsf > 0 && (it % J.in_wf) == 0;
    const bool xint = x >= J.nbc && x < J.nbc + J.nx;
    for (; r < z1; ++r) {
        const size_t o = (size_t)r * P + xs;
        const size_t on = o + P;
        // old stress row r+1 (row nzbc-1 is the zero ring)
        const T xx_n = ld(xx0, r + 1), zz_n = ld(zz0, r + 1), xz_n = ld(xz0, r + 1);
        // uu_new(r+1)
        const T xx_nr = shfl_from_right(xx_n);
        T uu_n = T(0);
        if (r + 1 <= nzbc - 2 && xin)
            uu_n = fix_u(cell_vel(temp[on], uu0[on], bb[on], xx_nr - xx_n + xz_n - xz_c), r + 1);
        // ww_new(r)
        const T xz_cl = shfl_from_left(xz_c);
        const T tc = ldin(temp, r);
        T ww_c = xin ? fix_w(cell_vel(tc, ww0[o], bb1[o], xz_c - xz_cl + zz_n - zz_c), r) : T(0);
        // strain and stress at row r
        const T uu_l = shfl_from_left(uu_c);
        const T ww_r = shfl_from_right(ww_c);
        if (out) {
            const T fux = uu_c - uu_l;
            const T fuz = uu_n - uu_c;
            const T bwx = ww_r - ww_c;
            const T bwz = ww_c - ww_m;
            T xo = xx_c, zo = zz_c;
            if (r == sz && x == sx) {
                if (stype == 1) { xo += sv; zo += sv; }
                if (stype == 4) zo += sv;
            }
            const T ca = J.ca[o], cl = J.cl[o], cm1 = J.cm1[o];
            xx1[o] = cell_stress2(tc, xo, ca, fux, cl, bwz, dtx);
            zz1[o] = (r == J.pt) ? T(0) : cell_stress2(tc, zo, ca, bwz, cl, fux, dtx);
            xz1[o] = cell_shear(tc, xz_c, cm1, fuz + bwx, dtx);
            uu1[o] = uu_c;
            ww1[o] = ww_c;
            if (r == J.gz) {
                const int d = x - J.gx;
                if (d >= 0 && d % J.dg == 0 && d / J.dg < J.ng) {
                    const size_t ti = (size_t)it * J.ng + d / J.dg;
                    J.trace_w[ti] = ww_c;
                    if (J.trace_u) J.trace_u[ti] = uu_c;
                }
            }
            if (snap_step && xint && r > J.pt && r <= J.pt + J.nz) {
                const size_t plane = (size_t)J.nz * J.nx;
                const size_t sb = (size_t)(it / J.in_wf) * J.nsf * plane;
                const size_t so = sb + (size_t)(r - J.pt - 1) * J.nx + (x - J.nbc);
                if (J.nsf == 3) {   // [row][x][fux, bwz, fuz + bwx]
                    const size_t sr = sb + ((size_t)(r - J.pt - 1) * J.nx + (x - J.nbc)) * 3;
                    stream_st(J.snap + sr, fux);
                    stream_st(J.snap + sr + 1, bwz);
                    stream_st(J.snap + sr + 2, fuz + bwx);
                } else {
                    stream_st(J.snap + so, fux);
                    stream_st(J.snap + so + plane, fuz);
                    stream_st(J.snap + so + 2 * plane, bwx);
                    stream_st(J.snap + so + 3 * plane, bwz);
                }
            }
        }
        xx_c = xx_n;
        zz_c = zz_n;
        xz_c = xz_n;
        uu_c = uu_n;
        ww_m = ww_c;
    }
}

// ---------------------------------------------------------------------------
// fused adjoint step + imaging
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kMarchLanes* kMarchWaves) void fd_adj_march(
    const FdJob<T>* __restrict__ jobs, const BlockMap* __restrict__ bmap, int step, int launch)
{
    const BlockMap bm = bmap[blockIdx.x];
    const FdJob<T>& J = jobs[bm.job];
    const int k = step - J.start;
    const int it = J.nt - 2 - k;
    if (k < 0 || it < 0) return;
    const int rd = (launch - J.launch0) & 1, wr = rd ^ 1;
    const int lane = threadIdx.x & (kMarchLanes - 1);
    const int tile = bm.tile0 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar row loop
    if (tile >= J.tiles_x * J.tiles_z) return;
    const int tix = tile % J.tiles_x, tiz = tile / J.tiles_x;
    const int nzbc = J.nzbc, nxbc = J.nxbc;
    const size_t P = (size_t)J.pitch;
    const int x = tix * kMarchOut + lane;
    const bool xload = x < nxbc;
    const bool xin = x >= 1 && x <= nxbc - 2;
    const bool out = xin && lane >= 1 && lane <= kMarchOut;
    const int z0 = 1 + tiz * J.tz;
    const int z1 = min(z0 + J.tz, nzbc - 1);
    const int xs = xload ? x : 0;

    const T* __restrict__ uu0 = J.uu[rd];
    const T* __restrict__ ww0 = J.ww[rd];
    const T* __restrict__ xx0 = J.xx[rd];
    const T* __restrict__ zz0 = J.zz[rd];
    const T* __restrict__ xz0 = J.xz[rd];
    const T* __restrict__ temp = J.temp;
    const T* __restrict__ ca = J.ca;
    const T* __restrict__ cl = J.cl;
    const T* __restrict__ cm1 = J.cm1;
    const T dtx = J.dtx;
    // adjoint-source column of this lane (receiver row gz only)
    int gi = -1;
    {
        const int d = x - J.gx;
        if (d >= 0 && d % J.dg == 0 && d / J.dg < J.ng) gi = d / J.dg;
    }
    const T inj = gi >= 0 ? J.dt * J.src[(size_t)it * J.ng + gi] : T(0);

    // stress-coefficient products of one row: (ca xx, cl zz, cl xx, ca zz, cm1 xz)
    struct Prod { T caxx, clzz, clxx, cazz, cmxz; };
    auto prod = [&](int rr, T xx, T zz, T xz) -> Prod {
        Prod p;
        if (xload) {
            const size_t o = (size_t)rr * P + xs;
            const T a = ca[o], l = cl[o], m = cm1[o];
            p.caxx = a * xx; p.clzz = l * zz; p.clxx = l * xx; p.cazz = a * zz; p.cmxz = m * xz;
        } else {
            p.caxx = p.clzz = p.clxx = p.cazz = p.cmxz = T(0);
        }
        return p;
    };
    // ww_b update of row rr from products of rows rr (pc) and rr+1 (pn)
    auto wwb = [&](int rr, const Prod& pc, const Prod& pn) -> T {
        const T cmxz_r = shfl_from_right(pc.cmxz);
        if (!xin || rr < 1 || rr > nzbc - 2) return T(0);
        const size_t o = (size_t)rr * P + xs;
        T w = cell_adj_vel(temp[o], ww0[o], J.b1[o], (pn.clxx - pc.clxx) + (pn.cazz - pc.cazz) + (cmxz_r - pc.cmxz));
        if (rr == J.gz) w += inj;
        return w;
    };
    // uu_b update of row rr from products of rows rr (pc) and rr-1 (pm)
    auto uub = [&](int rr, const Prod& pc, const Prod& pm) -> T {
        const T caxx_l = shfl_from_left(pc.caxx);
        const T clzz_l = shfl_from_left(pc.clzz);
        if (!xin || rr < 1 || rr > nzbc - 2) return T(0);
        const size_t o = (size_t)rr * P + xs;
        return cell_adj_vel(temp[o], uu0[o], J.b[o], (pc.caxx - caxx_l) + (pc.clzz - clzz_l) + (pc.cmxz - pm.cmxz));
    };

    int r = z0;
    T xx_c = ld_row(xx0, r, P, xs, xload), zz_c = ld_row(zz0, r, P, xs, xload), xz_c = ld_row(xz0, r, P, xs, xload);
    const Prod pm = prod(r - 1, ld_row(xx0, r - 1, P, xs, xload), ld_row(zz0, r - 1, P, xs, xload),
                         ld_row(xz0, r - 1, P, xs, xload));
    Prod pc = prod(r, xx_c, zz_c, xz_c);
    T ww_m = wwb(r - 1, pm, pc);
    T uu_c = uub(r, pc, pm);

    const bool img_step = (it % J.in_wf) == 0;
    const bool xint = x >= J.nbc && x < J.nbc + J.nx;
    T* __restrict__ uu1 = J.uu[wr];
    T* __restrict__ ww1 = J.ww[wr];
    T* __restrict__ xx1 = J.xx[wr];
    T* __restrict__ zz1 = J.zz[wr];
    T* __restrict__ xz1 = J.xz[wr];
    for (; r < z1; ++r) {
        const size_t o = (size_t)r * P + xs;
        const T xx_n = ld_row(xx0, r + 1, P, xs, xload), zz_n = ld_row(zz0, r + 1, P, xs, xload),
                xz_n = ld_row(xz0, r + 1, P, xs, xload);
        const Prod pn = prod(r + 1, xx_n, zz_n, xz_n);
        const T uu_n = uub(r + 1, pn, pc);
        const T ww_c = wwb(r, pc, pn);
        const T uu_r = shfl_from_right(uu_c);
        const T ww_l = shfl_from_left(ww_c);
        if (out) {
            const T tc = temp[o];
            const T xxn = cell_adj_stress(tc, xx_c, dtx, uu_r - uu_c);
            const T zzn = (r == J.pt) ? T(0) : cell_adj_stress(tc, zz_c, dtx, ww_c - ww_m);
            const T xzn = cell_adj_shear(tc, xz_c, dtx, uu_n - uu_c, ww_c - ww_l);
            xx1[o] = xxn;
            zz1[o] = zzn;
            xz1[o] = xzn;
            uu1[o] = uu_c;
            ww1[o] = ww_c;
            if (img_step && xint && r > J.pt && r <= J.pt + J.nz) {
                const size_t plane = (size_t)J.nz * J.nx;
                const size_t io = (size_t)(r - J.pt - 1) * J.nx + (x - J.nbc);
                const T* sb = J.snap + (size_t)(it / J.in_wf) * J.nsf * plane;
                const T* sn = sb + io;
                double Fux, Fwz, Fsh;
                if (J.nsf == 3) {
                    const T* sc = sb + ((size_t)(r - J.pt - 1) * J.nx + (x - J.nbc)) * 3;
                    Fux = stream_ld(sc); Fwz = stream_ld(sc + 1); Fsh = stream_ld(sc + 2);
                } else {
                    Fux = stream_ld(sn); Fwz = stream_ld(sn + 3 * plane);
                    Fsh = (double)(stream_ld(sn + plane) + stream_ld(sn + 2 * plane));
                }
                double c0, c1, c2;
                image_terms(Fux, Fwz, Fsh, xxn, zzn, xzn, c0, c1, c2);
                double* im = J.img + io;
                if constexpr (std::is_same<T, float>::value) {
                    // FP32 partials, flushed into the FP64 images every kImgFlush imaging steps
                    float* pp = J.imgp + 3 * io;
                    if (((it / J.in_wf) % kImgFlush) == 0) {
                        stream_st(im, image_flush(stream_ld(im), stream_ld(pp), c0));
                        stream_st(im + plane, image_flush(stream_ld(im + plane), stream_ld(pp + 1), c1));
                        stream_st(im + 2 * plane, image_flush(stream_ld(im + 2 * plane), stream_ld(pp + 2), c2));
                        stream_st(pp, 0.f);
                        stream_st(pp + 1, 0.f);
                        stream_st(pp + 2, 0.f);
                    } else {
                        stream_st(pp, image_partial(stream_ld(pp), c0, im));
                        stream_st(pp + 1, image_partial(stream_ld(pp + 1), c1, im + plane));
                        stream_st(pp + 2, image_partial(stream_ld(pp + 2), c2, im + 2 * plane));
                    }
                } else {
                    stream_st(im, stream_ld(im) + c0);
                    stream_st(im + plane, stream_ld(im + plane) + c1);
                    stream_st(im + 2 * plane, stream_ld(im + 2 * plane) + c2);
                }
            }
        }
        xx_c = xx_n;
        zz_c = zz_n;
        xz_c = xz_n;
        pc = pn;
        uu_c = uu_n;
        ww_m = ww_c;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T>
int launch_fwd_march(hipStream_t st, const FdJob<T>* d_jobs, const BlockMap* d_bmap, int nblocks, int step,
                     int launch)
{
    fd_fwd_march<T><<<nblocks, kMarchLanes * kMarchWaves, 0, st>>>(d_jobs, d_bmap, step, launch);
    SWD_HIP(hipGetLastError());
    return kOk;
}

template <typename T>
int launch_adj_march(hipStream_t st, const FdJob<T>* d_jobs, const BlockMap* d_bmap, int nblocks, int step,
                     int launch)
{
    fd_adj_march<T><<<nblocks, kMarchLanes * kMarchWaves, 0, st>>>(d_jobs, d_bmap, step, launch);
    SWD_HIP(hipGetLastError());
    return kOk;
}

template int launch_fwd_march<float>(hipStream_t, const FdJob<float>*, const BlockMap*, int, int, int);
template int launch_fwd_march<double>(hipStream_t, const FdJob<double>*, const BlockMap*, int, int, int);
template int launch_adj_march<float>(hipStream_t, const FdJob<float>*, const BlockMap*, int, int, int);
template int launch_adj_march<double>(hipStream_t, const FdJob<double>*, const BlockMap*, int, int, int);

}  // namespace swd
