// Cell updates of the SWD staggered-grid scheme, written once with explicit
// fused multiply-adds and shared by every fused FD kernel (single-step and
// time-blocked), so all of them round identically.  The library is compiled
// with -ffp-contract=off: no other contraction happens.
//
// forward  (core/eigen_staggerfd_single.cpp:224-228, :334-336)
//   v'   = temp*v + b*d
//   xx'  = temp*xx + (ca*fux + cl*bwz)*dtx
//   zz'  = temp*zz + (ca*bwz + cl*fux)*dtx
//   xz'  = temp*xz + (cm1*(fuz + bwx))*dtx
// adjoint  (core/eigen_e2drtm_single.cpp:182-187, :245-248)
//   v'   = temp*v - b*d
//   s'   = temp*s - dtx*d1 (- dtx*d2)
#pragma once

#include <hip/hip_runtime.h>

namespace swd {

template <typename T>
__device__ __forceinline__ T cell_vel(T temp, T v, T b, T d)
{
    return fma(b, d, temp * v);
}

template <typename T>
__device__ __forceinline__ T cell_stress2(T temp, T s, T c1, T e1, T c2, T e2, T dtx)
{
    return fma(fma(c1, e1, c2 * e2), dtx, temp * s);
}

template <typename T>
__device__ __forceinline__ T cell_shear(T temp, T s, T cm1, T e, T dtx)
{
    return fma(cm1 * e, dtx, temp * s);
}

template <typename T>
__device__ __forceinline__ T cell_adj_vel(T temp, T v, T b, T d)
{
    return fma(-b, d, temp * v);
}

template <typename T>
__device__ __forceinline__ T cell_adj_stress(T temp, T s, T dtx, T d)
{
    return fma(-dtx, d, temp * s);
}

template <typename T>
__device__ __forceinline__ T cell_adj_shear(T temp, T s, T dtx, T d1, T d2)
{
    return fma(-dtx, d2, fma(-dtx, d1, temp * s));
}

// Imaging condition (core/eigen_e2drtm_single.cpp:274-280) at one cell in
// FP64: c0 -> cl_img, c1 -> cm_img, c2 -> illum.
__device__ __forceinline__ void image_terms(double Fux, double Fwz, double Fsh, double X, double Z, double XZ,
                                            double& c0, double& c1, double& c2)
{
    c0 = (Fux + Fwz) * (X + Z);
    c1 = 2.0 * (X * Fux + Z * Fwz + 0.5 * XZ * Fsh);
    c2 = (Fux + Fwz) * (Fux + Fwz);
}

// An imaging step into the FP32 partial sum p, and the flush of p plus the
// step into the FP64 total (every kImgFlush imaging steps); every kernel
// that accumulates images uses these, so the sums do not depend on the
// launch shape.  A sum below FP32's normal range (0 < |v| < 2^-126, e.g.
// the squared strain of a cell the wavefield barely reaches) would lose its
// relative precision, or vanish, as a float: it goes to the FP64 total at
// once and the partial restarts from 0.  Otherwise the ratios
// cl_img ./ illum of weakly lit cells would differ from FP64 (0/0 or x/0
// where FP64 gives a finite ratio).
__device__ __forceinline__ bool image_tiny(double v) { return (v != 0.0) & (fabs(v) < 0x1p-126); }
template <typename DP>
__device__ __forceinline__ float image_partial(float p, double c, DP tot)
{
    const double v = (double)p + c;
    if (image_tiny(v)) {
        *tot += v;
        return 0.f;
    }
    return (float)v;
}
__device__ __forceinline__ double image_flush(double tot, float p, double c) { return tot + ((double)p + c); }

}  // namespace swd
