// Temporally blocked SWD FD march kernels for gfx950 (MI355X).
//
// One launch advances every shot of a batch by NT time steps.  A wavefront
// owns 64 consecutive columns and marches down a band of rows; the NT time
// levels run as a pipeline in registers (the coefficient rows they read sit
// in a per-wave LDS ring): level l handles row i - l while level 0 loads row
// i, so level l always sees rows r-1, r, r+1 of level l-1.
// Each level shrinks the valid lanes by one on each side (the stencil
// radius), so a wavefront writes 64 - 2 NT columns, and a band of H output
// rows loads H + 2 NT + 1 rows.  Intermediate levels never touch memory:
// per output cell and NT steps the kernel moves (H+2NT+1)/H x 11 loads and
// 5 stores instead of NT x 16, which is what lifts the stencil above the
// single-step HBM bound.
//
// Per cell and level the arithmetic is exactly that of the single-step
// march (fd_kernels.hip), in the same order, so results are bitwise equal.
// The kernels rely on the coefficient planes being zero on the outer ring
// (the fused paths pack them so): every ring cell then computes exactly 0,
// as the reference's never-updated ring, without per-lane predicates.
// Only source type 5 (the hard vertical-velocity source SWD uses) is
// handled here; other source types go through the single-step kernels.
#include <type_traits>

#include "fd_cell.h"
#include "swd_internal.h"

namespace swd {

// Global-address-space pointers and 32-bit byte offsets: loads and stores of
// grid rows become global_load/store with a scalar base and a vector offset.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ T gld(gptr<const T> base, unsigned byte_off)
{
    return *(gptr<const T>)((gptr<const char>)base + byte_off);
}
template <typename T>
__device__ __forceinline__ void gst(gptr<T> base, unsigned byte_off, T v)
{
    *(gptr<T>)((gptr<char>)base + byte_off) = v;
}

// Streaming accesses (non-temporal): snapshots and image sums are touched
// once per imaging step, while the next launch reads the state rows this one
// writes.  Marking the streams non-temporal keeps them from evicting that
// state from the L2 and the Infinity Cache (forward phase 113 -> 87 ms,
// adjoint 157 -> 146 ms on the 60-shot production workload).  The state and
// coefficient accesses keep the default policy: made non-temporal too, they
// were slower.
template <typename T>
__device__ __forceinline__ T stream_ld(gptr<const T> p) { return __builtin_nontemporal_load(p); }
template <typename T>
__device__ __forceinline__ void stream_st(gptr<T> p, T v) { __builtin_nontemporal_store(v, p); }

// Neighbour-lane exchange with DPP full-wave shifts (a VALU operand
// modifier, no LDS round trip): wave_shr:1 (0x138) gives lane i the value of
// lane i-1, wave_shl:1 (0x130) that of lane i+1.  Lanes 0 / 63 receive 0;
// they are halo lanes whose results are never written.
__device__ __forceinline__ int dpp_from_left(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ int dpp_from_right(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false); }

__device__ __forceinline__ float tb_left(float v) { return __int_as_float(dpp_from_left(__float_as_int(v))); }
__device__ __forceinline__ float tb_right(float v) { return __int_as_float(dpp_from_right(__float_as_int(v))); }
__device__ __forceinline__ double tb_left(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = dpp_from_left((int)(b & 0xffffffffLL)), hi = dpp_from_left((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double tb_right(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = dpp_from_right((int)(b & 0xffffffffLL)), hi = dpp_from_right((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// a step index clamped into [0, nt) (a level past nt reads a valid source sample it does not use)
__device__ __forceinline__ int itl_clamp(int itl, int nt) { return min(itl, nt - 1); }

// Issue priority from a wave's progress through its band.  The waves sharing
// a SIMD start together with equal work, and the arbiter issues oldest first,
// so without this they finish staggered (production forward: 46, 50, 56 and
// 63 us) and the last one runs alone, its memory latency no longer hidden by
// the others.  A wave drops priority as it passes 60, 80 and 93% of its rows,
// so the ones behind catch up and the SIMD's waves finish within the last
// stretch of rows.
struct ProgressPrio {
    int t1, t2, t3;
    __device__ __forceinline__ ProgressPrio(int i0, int i1)
    {
        const int n = i1 - i0 + 1;
        t1 = i0 + n * 60 / 100;
        t2 = i0 + n * 80 / 100;
        t3 = i0 + n * 93 / 100;
    }
    __device__ __forceinline__ void at(int i) const
    {
        if (i < t1) __builtin_amdgcn_s_setprio(3);
        else if (i < t2) __builtin_amdgcn_s_setprio(2);
        else if (i < t3) __builtin_amdgcn_s_setprio(1);
        else __builtin_amdgcn_s_setprio(0);
    }
};

// Minimum waves per SIMD the float kernels are compiled for (1 for the
// double kernels: no constraint).  Float kernels at NT <= 5 fit 128 VGPRs
// anyway; the hint changes the adjoint's instruction schedule (production
// adjoint 146 -> 138 ms at 5 x 26, results bitwise equal) and brings the
// forward at NT 6 to 4 waves (111 -> 84 ms at 6 x 26, a 52-byte spill).  At
// NT 6 with a separate b1 plane 3 waves is what fits.  The double kernels
// would spill hundreds of bytes under it.
template <typename T, int NT, bool B1>
constexpr int kTbWaves = !std::is_same<T, float>::value ? 1 : (NT >= 6 && B1) ? 3 : 4;

// coefficient record of one row for this lane
template <typename T>
struct CoefRow {
    T temp, b, b1, ca, cl, cm1;
};

// One row of state and coefficients for this lane.  Rows above/below the
// grid are clamped onto the first/last row, which (like the first/last
// column) is the zero ring of every state and coefficient plane.  B1: the
// b1 plane differs from b and is loaded; otherwise (the wrappers set
// b1 = b, core/staggerfd_eigen.m:106) b1 is b: one load and one ring plane
// fewer.
template <typename T, bool B1>
struct RowLoad {
    T xx, zz, xz, uu, ww;
    CoefRow<T> c;
    // float: temp, ca, cl, cm1 come from the interleaved block c4 (16 B per
    // cell, at 4 x the plane offset), one 16-byte load per lane instead of
    // four: the wave touches 9 lines instead of 12 for them and holds 4
    // VGPRs fewer; the production adjoint got 1.5% faster per launch, the
    // forward 0.3%
    __device__ __forceinline__ void load(gptr<const T> xx0, gptr<const T> zz0, gptr<const T> xz0, gptr<const T> uu0,
                                         gptr<const T> ww0, gptr<const T> ct, gptr<const T> cb, gptr<const T> cb1,
                                         gptr<const T> cca, gptr<const T> ccl, gptr<const T> ccm1, gptr<const T> c4,
                                         int row, int nzbc, unsigned rowb, unsigned xb)
    {
        const int rc = min(max(row, 0), nzbc - 1);
        const unsigned o = (unsigned)rc * rowb + xb;
        xx = gld(xx0, o); zz = gld(zz0, o); xz = gld(xz0, o); uu = gld(uu0, o); ww = gld(ww0, o);
        const T b = gld(cb, o);
        if constexpr (std::is_same<T, float>::value) {
            typedef float f4 __attribute__((ext_vector_type(4)));
            const f4 q = *(gptr<const f4>)((gptr<const char>)c4 + 4u * o);
            c = CoefRow<T>{q.x, b, B1 ? gld(cb1, o) : b, q.y, q.z, q.w};
        } else {
            (void)c4;
            c = CoefRow<T>{gld(ct, o), b, B1 ? gld(cb1, o) : b, gld(cca, o), gld(ccl, o), gld(ccm1, o)};
        }
    }
};

// Coefficients of rows i .. i-NT, which the NT levels of row i read (level l
// writes row i - l - 1 and reads the row above it).  Float kernels keep them
// in a per-wave LDS ring of NT+1 rows, row k in slot k mod (NT+1): a wave
// only touches its own slots and its LDS accesses stay in program order, so
// no barrier is needed.  That frees 5 (NT+1) VGPRs; the production forward
// and adjoint at NT 5 reach 4 waves per SIMD, and the gradient got 11 ms
// faster.  Double kernels keep the rows in registers (their LDS ring would
// exceed 64 KB per workgroup).
template <typename T, int NT, bool B1>
struct CoefRing {
    static constexpr bool kLds = std::is_same<T, float>::value;
    static constexpr int NC = B1 ? 6 : 5, NS = NT + 1;
    static constexpr int kLdsElems = kLds ? kMarchWaves * NS * NC * kMarchLanes : 1;
    T* w = nullptr;                     // LDS: this wave's slots, at this lane
    CoefRow<T> reg[kLds ? 1 : NT + 1];  // registers: reg[k] holds row i - k

    __device__ __forceinline__ void init(T* lds, int wave, int lane)
    {
        if constexpr (kLds) {
            w = lds + wave * NS * NC * kMarchLanes + lane;
#pragma unroll
            for (int k = 0; k < NS * NC; ++k) w[k * kMarchLanes] = T(0);
        } else {
#pragma unroll
            for (int k = 0; k <= NT; ++k) reg[k] = CoefRow<T>{T(0), T(0), T(0), T(0), T(0), T(0)};
        }
    }
    // row i enters (its slot last held row i - NT - 1, which no level reads any more)
    __device__ __forceinline__ void push(int i, const CoefRow<T>& c)
    {
        if constexpr (kLds) {
            T* s = w + ((i + 2 * NS) % NS) * NC * kMarchLanes;
            s[0] = c.temp;
            s[kMarchLanes] = c.b;
            if (B1) s[2 * kMarchLanes] = c.b1;
            s[(NC - 3) * kMarchLanes] = c.ca;
            s[(NC - 2) * kMarchLanes] = c.cl;
            s[(NC - 1) * kMarchLanes] = c.cm1;
        } else {
#pragma unroll
            for (int k = NT; k > 0; --k) reg[k] = reg[k - 1];
            reg[0] = c;
        }
    }
    // coefficients of row i - l - 1, the row level l writes (l = 0 .. NT-1)
    __device__ __forceinline__ CoefRow<T> below(int i, int l) const
    {
        if constexpr (kLds) {
            const T* s = w + ((i - l - 1 + 2 * NS) % NS) * NC * kMarchLanes;
            CoefRow<T> c;
            c.temp = s[0];
            c.b = s[kMarchLanes];
            c.b1 = B1 ? s[2 * kMarchLanes] : c.b;
            c.ca = s[(NC - 3) * kMarchLanes];
            c.cl = s[(NC - 2) * kMarchLanes];
            c.cm1 = s[(NC - 1) * kMarchLanes];
            return c;
        } else {
            return reg[l + 1];
        }
    }
};

// Snapshot and image values of one imaging row for this lane, loaded one
// row ahead like the state: raw snapshot fields and the current image sums
// (float kernels: the FP32 partials; double kernels: the FP64 images).
template <typename T>
struct ImgRow {
    T fux, fwz, fsh;  // snapshot fields of the imaging condition (fsh = fuz + bwx)
    T m0, m1, m2;
};

template <typename T, int NT, bool B1>
__global__ __launch_bounds__(kMarchLanes* kMarchWaves) __attribute__((amdgpu_waves_per_eu(kTbWaves<T, NT, B1>))) void fd_fwd_tb(const FdJob<T>* __restrict__ jobs,
                                                                      const BlockMap* __restrict__ bmap, int step,
                                                                      int launch)
{
    constexpr int W = kMarchLanes - 2 * NT;  // output columns per wavefront
    const BlockMap bm = bmap[blockIdx.x];  // the list is in dispatch order
    const FdJob<T>& J = jobs[bm.job];
    const int it0 = step - J.start;
    if (it0 < 0 || it0 >= J.nt) return;
    const int lane = threadIdx.x & (kMarchLanes - 1);
    // the wave index is uniform: read it into a scalar register so the row
    // loop, its bounds and every per-row condition stay scalar
    const int tile = bm.tile0 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (tile >= J.tiles_x * J.tiles_z) return;
    const int rd = (launch - J.launch0) & 1;
    const int tix = tile % J.tiles_x, tiz = tile / J.tiles_x;
    const int nzbc = J.nzbc, nxbc = J.nxbc;
    const size_t P = (size_t)J.pitch;
    const int x = tix * W + 1 - NT + lane;
    const bool xin = x >= 1 && x <= nxbc - 2;
    const bool out = xin && lane >= NT && lane < NT + W;
    // lanes and rows off the grid read the (all-zero) ring instead
    const int xs = min(max(x, 0), nxbc - 1);
    const int z0 = 1 + tiz * J.tz;
    const int z1 = min(z0 + J.tz, nzbc - 1);

    // uniform selections keep the pointers in scalar registers
    const gptr<const T> uu0 = (gptr<const T>)(rd ? J.uu[1] : J.uu[0]);
    const gptr<const T> ww0 = (gptr<const T>)(rd ? J.ww[1] : J.ww[0]);
    const gptr<const T> xx0 = (gptr<const T>)(rd ? J.xx[1] : J.xx[0]);
    const gptr<const T> zz0 = (gptr<const T>)(rd ? J.zz[1] : J.zz[0]);
    const gptr<const T> xz0 = (gptr<const T>)(rd ? J.xz[1] : J.xz[0]);
    const gptr<T> uu1 = (gptr<T>)(rd ? J.uu[0] : J.uu[1]);
    const gptr<T> ww1 = (gptr<T>)(rd ? J.ww[0] : J.ww[1]);
    const gptr<T> xx1 = (gptr<T>)(rd ? J.xx[0] : J.xx[1]);
    const gptr<T> zz1 = (gptr<T>)(rd ? J.zz[0] : J.zz[1]);
    const gptr<T> xz1 = (gptr<T>)(rd ? J.xz[0] : J.xz[1]);
    const gptr<const T> ctemp = (gptr<const T>)J.temp, cb = (gptr<const T>)J.b, cb1 = (gptr<const T>)J.b1;
    const gptr<const T> cca = (gptr<const T>)J.ca, ccl = (gptr<const T>)J.cl, ccm1 = (gptr<const T>)J.cm1;
    const gptr<const T> cc4 = (gptr<const T>)J.c4;
    const unsigned rowb = (unsigned)(P * sizeof(T));
    const unsigned xb = (unsigned)(xs * sizeof(T));
    const T dtx = J.dtx;
    // source column; off the grid when the job has no type-5 source
    const int sz = J.sz, sx = J.stype == 5 ? J.sx : -(1 << 30), pt = J.pt, gz = J.gz;
    const int nt = J.nt;
    const int ng = J.ng, in_wf = J.in_wf, nsf = J.nsf, nzi = J.nz, nxi = J.nx, nbc = J.nbc;
    const gptr<T> trace_w = (gptr<T>)J.trace_w;
    const gptr<T> trace_u = (gptr<T>)J.trace_u;
    const gptr<T> snap = (gptr<T>)J.snap;

    // per-level source value (levels past nt compute but record nothing).
    // Read through a global-address-space pointer: a generic (flat) load
    // here leaves a FLAT access pending at the row loop's header, and the
    // compiler then waits for every outstanding memory access (vmcnt 0) once
    // per loop iteration, in the forward right after the next row's loads
    // are issued.  With 16 waves per CU the other waves hid that stall
    // (production forward 65.9 -> 65.3 us per launch, adjoint unchanged).
    const gptr<const T> src = (gptr<const T>)J.src;
    T sv[NT];
#pragma unroll
    for (int l = 0; l < NT; ++l) sv[l] = it0 + l < nt ? src[itl_clamp(it0 + l, nt)] : T(0);
    // receiver index and snapshot column of this lane's output column, -1
    // where it records nothing (integers, so no lane masks stay live)
    int gi = -1;
    {
        const int d = x - J.gx;
        if (out && d >= 0 && d % J.dg == 0 && d / J.dg < J.ng) gi = d / J.dg;
    }
    const int sxc = (out && x >= J.nbc && x < J.nbc + J.nx) ? x - J.nbc : -1;

    // carried per level l (1..NT, stored at index l-1)
    T cxx[NT], czz[NT], cxz[NT], cwin[NT], cuo[NT], cwo[NT];
#pragma unroll
    for (int l = 0; l < NT; ++l) cxx[l] = czz[l] = cxz[l] = cwin[l] = cuo[l] = cwo[l] = T(0);
    __shared__ T cring_s[CoefRing<T, NT, B1>::kLdsElems];
    CoefRing<T, NT, B1> ring;
    ring.init(cring_s, __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane);

    const int i0 = z0 - NT - 1, i1 = z1 - 1 + NT;
    // row loads are issued one iteration ahead so their latency overlaps the
    // level updates of the current row
    auto row = [&](const RowLoad<T, B1>& cur, const int i) {
        // level 0: row i of the stored state and its coefficients
        T ixx = cur.xx, izz = cur.zz, ixz = cur.xz, iu = cur.uu, iw = cur.ww;
        ring.push(i, cur.c);
        CoefRow<T> cq = cur.c;  // coefficients of level l's input row i - l

#pragma unroll
        for (int l = 0; l < NT; ++l) {
            // level l+1 consumes level-l row q = i - l, produces row r = q - 1
            const int r = i - l - 1;
            const CoefRow<T> cr_ = ring.below(i, l);
            // uu_new(q)
            const T xx_r = tb_right(ixx);
            const T u_q = cell_vel(cq.temp, iu, cq.b, xx_r - ixx + ixz - cxz[l]);
            // ww_new(r)
            const T xz_l = tb_left(cxz[l]);
            T w_r = cell_vel(cr_.temp, cwin[l], cr_.b1, cxz[l] - xz_l + izz - czz[l]);
            if (r == sz && x == sx) w_r = sv[l];
            // stress(r)
            const T u_r = cuo[l];
            const T uu_l = tb_left(u_r);
            const T ww_rt = tb_right(w_r);
            const T fux = u_r - uu_l;
            const T fuz = u_q - u_r;
            const T bwx = ww_rt - w_r;
            const T bwz = w_r - cwo[l];
            const T tc = cr_.temp;
            const T oxx = cell_stress2(tc, cxx[l], cr_.ca, fux, cr_.cl, bwz, dtx);
            const T ozz = (r == pt) ? T(0) : cell_stress2(tc, czz[l], cr_.ca, bwz, cr_.cl, fux, dtx);
            const T oxz = cell_shear(tc, cxz[l], cr_.cm1, fuz + bwx, dtx);
            // records of level l+1 (step it0 + l) owned by this wavefront;
            // the row and step tests are wave-uniform and come first
            if (it0 + l < nt && r >= z0 && r < z1) {
                const int itl = it0 + l;
                if (r == gz && gi >= 0) {
                    const size_t ti = (size_t)itl * ng + gi;
                    trace_w[ti] = w_r;
                    if (trace_u) trace_u[ti] = u_r;
                }
                if (nsf > 0 && (itl % in_wf) == 0 && r > pt && r <= pt + nzi && sxc >= 0) {
                    const size_t plane = (size_t)nzi * nxi;
                    const size_t sb = (size_t)(itl / in_wf) * nsf * plane;
                    const size_t so = sb + (size_t)(r - pt - 1) * nxi + sxc;
                    if (nsf == 3) {   // [row][x][fux, bwz, fuz + bwx]: one 12-byte store
                        const size_t sr = sb + ((size_t)(r - pt - 1) * nxi + sxc) * 3;
                        stream_st(snap + sr, fux);
                        stream_st(snap + sr + 1, bwz);
                        stream_st(snap + sr + 2, fuz + bwx);
                    } else {
                        stream_st(snap + so, fux);
                        stream_st(snap + so + plane, fuz);
                        stream_st(snap + so + 2 * plane, bwx);
                        stream_st(snap + so + 3 * plane, bwz);
                    }
                }
            }
            // carry this level's inputs and outputs to the next iteration
            cxx[l] = ixx; czz[l] = izz; cxz[l] = ixz; cwin[l] = iw;
            cuo[l] = u_q;
            cwo[l] = w_r;
            // level-(l+1) row r becomes the input of the next level
            ixx = oxx; izz = ozz; ixz = oxz; iu = u_r; iw = w_r;
            cq = cr_;
        }
        // final level: row i - NT
        const int rf = i - NT;
        if (rf >= z0 && rf < z1 && out) {
            const unsigned o = (unsigned)rf * rowb + xb;
            gst(xx1, o, ixx);
            gst(zz1, o, izz);
            gst(xz1, o, ixz);
            gst(uu1, o, iu);
            gst(ww1, o, iw);
        }
    };
    // Rows are loaded one step ahead into two alternating buffers, so the
    // loads of row i+1 are in flight while row i is processed.
    RowLoad<T, B1> A, B;
    const ProgressPrio prio(i0, i1);
    A.load(xx0, zz0, xz0, uu0, ww0, ctemp, cb, cb1, cca, ccl, ccm1, cc4, i0, nzbc, rowb, xb);
    for (int i = i0; i <= i1; i += 2) {
        prio.at(i);
        B.load(xx0, zz0, xz0, uu0, ww0, ctemp, cb, cb1, cca, ccl, ccm1, cc4, i + 1, nzbc, rowb, xb);
        row(A, i);
        if (i + 1 > i1) break;
        A.load(xx0, zz0, xz0, uu0, ww0, ctemp, cb, cb1, cca, ccl, ccm1, cc4, i + 2, nzbc, rowb, xb);
        row(B, i + 1);
    }
}

// Adjoint (reverse-time) counterpart with the imaging condition.  Level l
// runs adjoint step it = nt - 2 - (k0 + l).
template <typename T, int NT, bool B1>
__global__ __launch_bounds__(kMarchLanes* kMarchWaves) __attribute__((amdgpu_waves_per_eu(kTbWaves<T, NT, B1>))) void fd_adj_tb(const FdJob<T>* __restrict__ jobs,
                                                                      const BlockMap* __restrict__ bmap, int step,
                                                                      int launch)
{
    constexpr int W = kMarchLanes - 2 * NT;
    const BlockMap bm = bmap[blockIdx.x];  // the list is in dispatch order
    const FdJob<T>& J = jobs[bm.job];
    const int k0 = step - J.start;
    if (k0 < 0 || J.nt - 2 - k0 < 0) return;
    const int lane = threadIdx.x & (kMarchLanes - 1);
    // the wave index is uniform: read it into a scalar register so the row
    // loop, its bounds and every per-row condition stay scalar
    const int tile = bm.tile0 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (tile >= J.tiles_x * J.tiles_z) return;
    const int rd = (launch - J.launch0) & 1;
    const int tix = tile % J.tiles_x, tiz = tile / J.tiles_x;
    const int nzbc = J.nzbc, nxbc = J.nxbc;
    const size_t P = (size_t)J.pitch;
    const int x = tix * W + 1 - NT + lane;
    const bool xin = x >= 1 && x <= nxbc - 2;
    const bool out = xin && lane >= NT && lane < NT + W;
    // lanes and rows off the grid read the (all-zero) ring instead
    const int xs = min(max(x, 0), nxbc - 1);
    const int z0 = 1 + tiz * J.tz;
    const int z1 = min(z0 + J.tz, nzbc - 1);

    // uniform selections keep the pointers in scalar registers
    const gptr<const T> uu0 = (gptr<const T>)(rd ? J.uu[1] : J.uu[0]);
    const gptr<const T> ww0 = (gptr<const T>)(rd ? J.ww[1] : J.ww[0]);
    const gptr<const T> xx0 = (gptr<const T>)(rd ? J.xx[1] : J.xx[0]);
    const gptr<const T> zz0 = (gptr<const T>)(rd ? J.zz[1] : J.zz[0]);
    const gptr<const T> xz0 = (gptr<const T>)(rd ? J.xz[1] : J.xz[0]);
    const gptr<T> uu1 = (gptr<T>)(rd ? J.uu[0] : J.uu[1]);
    const gptr<T> ww1 = (gptr<T>)(rd ? J.ww[0] : J.ww[1]);
    const gptr<T> xx1 = (gptr<T>)(rd ? J.xx[0] : J.xx[1]);
    const gptr<T> zz1 = (gptr<T>)(rd ? J.zz[0] : J.zz[1]);
    const gptr<T> xz1 = (gptr<T>)(rd ? J.xz[0] : J.xz[1]);
    const gptr<const T> ctemp = (gptr<const T>)J.temp, cb = (gptr<const T>)J.b, cb1 = (gptr<const T>)J.b1;
    const gptr<const T> cca = (gptr<const T>)J.ca, ccl = (gptr<const T>)J.cl, ccm1 = (gptr<const T>)J.cm1;
    const gptr<const T> cc4 = (gptr<const T>)J.c4;
    const unsigned rowb = (unsigned)(P * sizeof(T));
    const unsigned xb = (unsigned)(xs * sizeof(T));
    const T dtx = J.dtx;
    const int pt = J.pt, gz = J.gz;
    const int in_wf = J.in_wf, nsf = J.nsf, nzi = J.nz, nxi = J.nx, nbc = J.nbc;
    const gptr<const T> snap = (gptr<const T>)J.snap;
    const gptr<double> img = (gptr<double>)J.img;
    // float kernels sum the images in FP32 partials (J.imgp), flushed into
    // the FP64 images every kImgFlush imaging steps; double kernels add into
    // the FP64 images directly
    constexpr bool kPart = std::is_same<T, float>::value;
    const gptr<T> imgp = (gptr<T>)(kPart ? (T*)J.imgp : (T*)J.img);

    int gi = -1;
    {
        const int d = x - J.gx;
        if (d >= 0 && d % J.dg == 0 && d / J.dg < J.ng) gi = d / J.dg;
    }
    // adjoint step of each level (< 0: past the end, nothing recorded); the
    // adjoint source is read through a global pointer, as in the forward
    const gptr<const T> src = (gptr<const T>)J.src;
    int itl[NT];
    T inj[NT];
#pragma unroll
    for (int l = 0; l < NT; ++l) {
        itl[l] = J.nt - 2 - (k0 + l);
        inj[l] = (itl[l] >= 0 && gi >= 0) ? J.dt * src[(size_t)itl[l] * J.ng + gi] : T(0);
    }
    // image column of this lane's output column, -1 where it images nothing
    // (an integer, so no lane mask stays live)
    const int icol = (out && x >= J.nbc && x < J.nbc + J.nx) ? x - J.nbc : -1;
    // Imaging levels of this launch.  With exactly one (the usual case,
    // in_wf >= NT) its snapshot and image values are prefetched one row
    // ahead with the state; otherwise every imaging level loads inline.
    int limg = -1, nimg = 0;
#pragma unroll
    for (int l = 0; l < NT; ++l)
        if (itl[l] >= 0 && (itl[l] % in_wf) == 0) { limg = l; ++nimg; }
    const bool pf_img = nimg == 1;
    const bool flush = pf_img && ((itl[limg] / in_wf) % kImgFlush) == 0;
    const size_t iplane = (size_t)nzi * nxi;
    const gptr<const T> snb = snap + (pf_img ? (size_t)(itl[limg] / in_wf) * nsf * iplane : 0);
    // Image rows are addressed as a uniform row base (scalar registers) plus
    // this lane's byte offset, like the state rows.  Snapshot slots are
    // [row][x][fux, bwz, fuz + bwx] (the time-blocked adjoint takes only that
    // layout; 11 MB less fetch per production launch than the fields of a row
    // as three segments); the FP32 partials are interleaved [cell][3] too, one
    // stream instead of three planes (production adjoint 139 -> 133 ms mean
    // over seven processes); the double kernels keep their FP64 images as
    // planes.
    // this lane's cell offset in a snapshot / image row
    // (64-bit, so the three interleaved partials of a cell stay one
    // dwordx3 access)
    const size_t lane_c = (size_t)max(icol, 0);
    const size_t lane_p = kPart ? 3 * lane_c : lane_c;
    auto img_load = [&](ImgRow<T>& q, const int i) {
        const int r = i - limg - 1;
        if (pf_img && r >= z0 && r < z1 && r > pt && r <= pt + nzi && icol >= 0) {
            const int rr = r - pt - 1;
            const gptr<const T> sr = snb + ((size_t)rr * nxi + lane_c) * 3;
            q.fux = stream_ld(sr);
            q.fwz = stream_ld(sr + 1);
            q.fsh = stream_ld(sr + 2);
            if constexpr (kPart) {
                const gptr<const T> pr = imgp + (size_t)rr * 3 * nxi + lane_p;
                q.m0 = stream_ld(pr);
                q.m1 = stream_ld(pr + 1);
                q.m2 = stream_ld(pr + 2);
            } else {
                const gptr<const T> ir = imgp + (size_t)rr * nxi + lane_p;
                q.m0 = stream_ld(ir);
                q.m1 = stream_ld(ir + iplane);
                q.m2 = stream_ld(ir + 2 * iplane);
            }
        }
    };

    // carried per level: previous input row (stress, old ww), previous
    // outputs u_l(r), w_l(r-1).  The previous row's coefficient products are
    // formed again from it and its coefficients (ring.below(i, l)): the same
    // products, without carrying 3 NT more registers.
    T cxx[NT], czz[NT], cxz[NT], cwin[NT], cuo[NT], cwo[NT];
#pragma unroll
    for (int l = 0; l < NT; ++l) cxx[l] = czz[l] = cxz[l] = cwin[l] = cuo[l] = cwo[l] = T(0);
    __shared__ T cring_s[CoefRing<T, NT, B1>::kLdsElems];
    CoefRing<T, NT, B1> ring;
    ring.init(cring_s, __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane);

    const int i0 = z0 - NT - 1, i1 = z1 - 1 + NT;
    // lean: at most one imaging level in this launch, imaged after the level
    // chain from prefetched values; otherwise every imaging level images inline
    auto row = [&](const RowLoad<T, B1>& cur, const ImgRow<T>& icur, const int i, auto lean) {
        constexpr bool kLean = decltype(lean)::value;
        T ixx = cur.xx, izz = cur.zz, ixz = cur.xz, iu = cur.uu, iw = cur.ww;
        T sxx = T(0), szz = T(0), sxz = T(0);
        ring.push(i, cur.c);
        CoefRow<T> cq = cur.c;  // coefficients of level l's input row i - l

#pragma unroll
        for (int l = 0; l < NT; ++l) {
            const int r = i - l - 1;
            const CoefRow<T> cr_ = ring.below(i, l);
            // coefficient products of input row q = r + 1
            const T caxx = cq.ca * ixx, clzz = cq.cl * izz, clxx = cq.cl * ixx, cazz = cq.ca * izz,
                    cmxz = cq.cm1 * ixz;
            // uu_b(q)
            const T caxx_l = tb_left(caxx), clzz_l = tb_left(clzz);
            // coefficient products of the previous input row r
            const T pclxx = cr_.cl * cxx[l], pcazz = cr_.ca * czz[l], pcmxz = cr_.cm1 * cxz[l];
            const T u_q = cell_adj_vel(cq.temp, iu, cq.b, (caxx - caxx_l) + (clzz - clzz_l) + (cmxz - pcmxz));
            // ww_b(r) (+ adjoint source at the receiver row)
            const T cmxz_r = tb_right(pcmxz);
            T w_r = cell_adj_vel(cr_.temp, cwin[l], cr_.b1, (clxx - pclxx) + (cazz - pcazz) + (cmxz_r - pcmxz));
            if (r == gz) w_r += inj[l];
            // adjoint stress(r)
            const T u_r = cuo[l];
            const T u_rt = tb_right(u_r);
            const T w_l = tb_left(w_r);
            const T tc = cr_.temp;
            const T oxx = cell_adj_stress(tc, cxx[l], dtx, u_rt - u_r);
            const T ozz = (r == pt) ? T(0) : cell_adj_stress(tc, czz[l], dtx, w_r - cwo[l]);
            const T oxz = cell_adj_shear(tc, cxz[l], dtx, u_q - u_r, w_r - w_l);
            if constexpr (kLean) {
                // the imaging level's adjoint stresses, imaged after the chain
                if (l == limg) { sxx = oxx; szz = ozz; sxz = oxz; }
            } else {
                if (itl[l] >= 0 && (itl[l] % in_wf) == 0 && r >= z0 && r < z1 && r > pt && r <= pt + nzi &&
                           icol >= 0) {
                    const size_t plane = (size_t)nzi * nxi;
                    const size_t io = (size_t)(r - pt - 1) * nxi + icol;
                    const gptr<const T> sb = snap + (size_t)(itl[l] / in_wf) * 3 * plane;
                    const gptr<const T> sc = sb + ((size_t)(r - pt - 1) * nxi + icol) * 3;
                    const double Fux = stream_ld(sc), Fwz = stream_ld(sc + 1), Fsh = stream_ld(sc + 2);
                    double c0, c1, c2;
                    image_terms(Fux, Fwz, Fsh, oxx, ozz, oxz, c0, c1, c2);
                    const gptr<double> im = img + io;
                    const gptr<const double> ic = im;
                    if constexpr (kPart) {
                        const gptr<T> pp = imgp + 3 * io;
                        const gptr<const T> pc = pp;
                        if (((itl[l] / in_wf) % kImgFlush) == 0) {
                            stream_st(im, image_flush(stream_ld(ic), stream_ld(pc), c0));
                            stream_st(im + plane, image_flush(stream_ld(ic + plane), stream_ld(pc + 1), c1));
                            stream_st(im + 2 * plane, image_flush(stream_ld(ic + 2 * plane), stream_ld(pc + 2), c2));
                            stream_st(pp, T(0));
                            stream_st(pp + 1, T(0));
                            stream_st(pp + 2, T(0));
                        } else {
                            stream_st(pp, image_partial(stream_ld(pc), c0, im));
                            stream_st(pp + 1, image_partial(stream_ld(pc + 1), c1, im + plane));
                            stream_st(pp + 2, image_partial(stream_ld(pc + 2), c2, im + 2 * plane));
                        }
                    } else {
                        stream_st(im, stream_ld(ic) + c0);
                        stream_st(im + plane, stream_ld(ic + plane) + c1);
                        stream_st(im + 2 * plane, stream_ld(ic + 2 * plane) + c2);
                    }
                }
            }
            cxx[l] = ixx; czz[l] = izz; cxz[l] = ixz; cwin[l] = iw;
            cuo[l] = u_q;
            cwo[l] = w_r;
            ixx = oxx; izz = ozz; ixz = oxz; iu = u_r; iw = w_r;
            cq = cr_;
        }
        if constexpr (kLean) {
            const int r = i - limg - 1;
            if (pf_img && r >= z0 && r < z1 && r > pt && r <= pt + nzi) {
                const int rr = r - pt - 1;
                const gptr<double> imr = img + (size_t)rr * nxi + lane_c;
                // new FP32 partials of this row (0 after a flush)
                float n0 = 0.f, n1 = 0.f, n2 = 0.f;
                if (icol >= 0) {
                    // prefetched snapshot and image values of this row
                    double c0, c1, c2;
                    image_terms(icur.fux, icur.fwz, icur.fsh, sxx, szz, sxz, c0, c1, c2);
                    double* imd = (double*)imr;
                    if constexpr (kPart) {
                        if (flush) {
                            // image_flush as atomic adds without return: the same sum
                            // (one lane per cell and launch), and no wait on an FP64
                            // load (24 us less per flush launch at the production shape)
                            atomicAdd(imd, (double)icur.m0 + c0);
                            atomicAdd(imd + iplane, (double)icur.m1 + c1);
                            atomicAdd(imd + 2 * iplane, (double)icur.m2 + c2);
                        } else {
                            // image_partial for the three images, with one test on the common path
                            double v0 = (double)icur.m0 + c0, v1 = (double)icur.m1 + c1, v2 = (double)icur.m2 + c2;
                            // a spill is the only access to that FP64 image in this launch (one
                            // imaging level per launch here), so it goes as an atomic add without
                            // return: no wait for it on the path
                            if (image_tiny(v0)) { atomicAdd(imd, v0); v0 = 0.0; }
                            if (image_tiny(v1)) { atomicAdd(imd + iplane, v1); v1 = 0.0; }
                            if (image_tiny(v2)) { atomicAdd(imd + 2 * iplane, v2); v2 = 0.0; }
                            n0 = (float)v0;
                            n1 = (float)v1;
                            n2 = (float)v2;
                        }
                    } else {
                        stream_st(imr, (double)icur.m0 + c0);
                        stream_st(imr + iplane, (double)icur.m1 + c1);
                        stream_st(imr + 2 * iplane, (double)icur.m2 + c2);
                    }
                }
                if constexpr (kPart) {
                    if (icol >= 0) {
                        const gptr<T> pr = imgp + (size_t)rr * 3 * nxi + lane_p;
                        stream_st(pr, (T)n0);
                        stream_st(pr + 1, (T)n1);
                        stream_st(pr + 2, (T)n2);
                    }
                }
            }
        }
        const int rf = i - NT;
        if (rf >= z0 && rf < z1 && out) {
            const unsigned o = (unsigned)rf * rowb + xb;
            gst(xx1, o, ixx);
            gst(zz1, o, izz);
            gst(xz1, o, ixz);
            gst(uu1, o, iu);
            gst(ww1, o, iw);
        }
    };
    // Rows are loaded one step ahead into two alternating buffers, so the
    // loads of row i+1 are in flight while row i is processed.
    auto march = [&](auto lean) {
        constexpr bool kLean = decltype(lean)::value;
        RowLoad<T, B1> A, B;
        ImgRow<T> IA, IB;
        const ProgressPrio prio(i0, i1);
        A.load(xx0, zz0, xz0, uu0, ww0, ctemp, cb, cb1, cca, ccl, ccm1, cc4, i0, nzbc, rowb, xb);
        if constexpr (kLean) img_load(IA, i0);
        for (int i = i0; i <= i1; i += 2) {
            prio.at(i);
            B.load(xx0, zz0, xz0, uu0, ww0, ctemp, cb, cb1, cca, ccl, ccm1, cc4, i + 1, nzbc, rowb, xb);
            if constexpr (kLean) img_load(IB, i + 1);
            row(A, IA, i, lean);
            if (i + 1 > i1) break;
            A.load(xx0, zz0, xz0, uu0, ww0, ctemp, cb, cb1, cca, ccl, ccm1, cc4, i + 2, nzbc, rowb, xb);
            if constexpr (kLean) img_load(IA, i + 2);
            row(B, IB, i + 1, lean);
        }
    };
    if (nimg <= 1) march(std::true_type{});
    else march(std::false_type{});
}

template <typename T, bool B1>
static int fwd_tb_nt(hipStream_t st, const FdJob<T>* d_jobs, const BlockMap* d_bmap, int nblocks, int step, int launch,
                     int nt_block)
{
    const dim3 blk(kMarchLanes * kMarchWaves);
    switch (nt_block) {
    case 2: fd_fwd_tb<T, 2, B1><<<nblocks, blk, 0, st>>>(d_jobs, d_bmap, step, launch); break;
    case 3: fd_fwd_tb<T, 3, B1><<<nblocks, blk, 0, st>>>(d_jobs, d_bmap, step, launch); break;
    case 4: fd_fwd_tb<T, 4, B1><<<nblocks, blk, 0, st>>>(d_jobs, d_bmap, step, launch); break;
    case 5: fd_fwd_tb<T, 5, B1><<<nblocks, blk, 0, st>>>(d_jobs, d_bmap, step, launch); break;
    case 6: fd_fwd_tb<T, 6, B1><<<nblocks, blk, 0, st>>>(d_jobs, d_bmap, step, launch); break;
    default: set_error("unsupported time block %d", nt_block); return kErrArg;
    }
    SWD_HIP(hipGetLastError());
    return kOk;
}

template <typename T, bool B1>
static int adj_tb_nt(hipStream_t st, const FdJob<T>* d_jobs, const BlockMap* d_bmap, int nblocks, int step, int launch,
                     int nt_block)
{
    const dim3 blk(kMarchLanes * kMarchWaves);
    switch (nt_block) {
    case 2: fd_adj_tb<T, 2, B1><<<nblocks, blk, 0, st>>>(d_jobs, d_bmap, step, launch); break;
    case 3: fd_adj_tb<T, 3, B1><<<nblocks, blk, 0, st>>>(d_jobs, d_bmap, step, launch); break;
    case 4: fd_adj_tb<T, 4, B1><<<nblocks, blk, 0, st>>>(d_jobs, d_bmap, step, launch); break;
    case 5: fd_adj_tb<T, 5, B1><<<nblocks, blk, 0, st>>>(d_jobs, d_bmap, step, launch); break;
    case 6: fd_adj_tb<T, 6, B1><<<nblocks, blk, 0, st>>>(d_jobs, d_bmap, step, launch); break;
    default: set_error("unsupported time block %d", nt_block); return kErrArg;
    }
    SWD_HIP(hipGetLastError());
    return kOk;
}

template <typename T>
int launch_fwd_tb(hipStream_t st, const FdJob<T>* d_jobs, const BlockMap* d_bmap, int nblocks, int step, int launch,
                  int nt_block, bool b1)
{
    return b1 ? fwd_tb_nt<T, true>(st, d_jobs, d_bmap, nblocks, step, launch, nt_block)
              : fwd_tb_nt<T, false>(st, d_jobs, d_bmap, nblocks, step, launch, nt_block);
}

template <typename T>
int launch_adj_tb(hipStream_t st, const FdJob<T>* d_jobs, const BlockMap* d_bmap, int nblocks, int step, int launch,
                  int nt_block, bool b1)
{
    return b1 ? adj_tb_nt<T, true>(st, d_jobs, d_bmap, nblocks, step, launch, nt_block)
              : adj_tb_nt<T, false>(st, d_jobs, d_bmap, nblocks, step, launch, nt_block);
}

// Resident workgroups per CU of the time-blocked kernel (occupancy API), 0 if unknown.
template <typename T>
int tb_blocks_per_cu(bool adj, int nt_block, bool b1)
{
    const void* f = nullptr;
#define SWD_TB_FN(N)                                                                                              \
    f = adj ? (b1 ? (const void*)fd_adj_tb<T, N, true> : (const void*)fd_adj_tb<T, N, false>)                    \
            : (b1 ? (const void*)fd_fwd_tb<T, N, true> : (const void*)fd_fwd_tb<T, N, false>)
    switch (nt_block) {
    case 2: SWD_TB_FN(2); break;
    case 3: SWD_TB_FN(3); break;
    case 4: SWD_TB_FN(4); break;
    case 5: SWD_TB_FN(5); break;
    case 6: SWD_TB_FN(6); break;
    default: return 0;
    }
#undef SWD_TB_FN
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, kMarchLanes * kMarchWaves, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return nb;
}
template int tb_blocks_per_cu<float>(bool, int, bool);
template int tb_blocks_per_cu<double>(bool, int, bool);

template int launch_fwd_tb<float>(hipStream_t, const FdJob<float>*, const BlockMap*, int, int, int, int, bool);
template int launch_adj_tb<float>(hipStream_t, const FdJob<float>*, const BlockMap*, int, int, int, int, bool);
template int launch_fwd_tb<double>(hipStream_t, const FdJob<double>*, const BlockMap*, int, int, int, int, bool);
template int launch_adj_tb<double>(hipStream_t, const FdJob<double>*, const BlockMap*, int, int, int, int, bool);

}  // namespace swd
