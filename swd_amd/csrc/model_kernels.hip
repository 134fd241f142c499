// Device side of the SWD model update (one inversion iteration without host
// round trips): coefficient planes from device vp / vs, smooth2a, the
// clipped model step and the FP64 reductions of the step length.
#include <hip/hip_runtime.h>

#include "swd_host.h"

namespace swd {

// Coefficient planes of the fused kernels (temp, b, b1, ca, cl, cm1; packed
// row-major with the grid pitch, outer ring and pitch padding zero) from the
// interior model, as the wrappers prepare them (core/staggerfd_eigen.m:37-107
// with den = 1): edge-replicated vp^2 / vs^2, cl = ca - 2 cm, damping kappa *
// geometry (precomputed geometry factors, nzbc x nxbc column-major), free
// surface row: ca = 2 cm, cl = 0, b = 2 dtx.  Arithmetic in FP64 in the
// host's order, then rounded to FP32, so the planes equal the host's.
__global__ void prep_coefs_kernel(const double* __restrict__ vp, const double* __restrict__ vs,
                                  const double* __restrict__ geo, double kappa, int nz, int nx, int nbc, int top,
                                  int nzbc, int nxbc, int pitch, size_t plane, double dt, double dtx,
                                  float* __restrict__ coefs)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    if (ix >= pitch || iz >= nzbc) return;
    const size_t o = (size_t)iz * pitch + ix;
    if (ix >= nxbc || ix == 0 || ix == nxbc - 1 || iz == 0 || iz == nzbc - 1) {
        for (int k = 0; k < 6; ++k) coefs[k * plane + o] = 0.f;
        return;
    }
    const int iz0 = min(max(iz - top, 0), nz - 1), ix0 = min(max(ix - nbc, 0), nx - 1);
    const double p = vp[(size_t)ix0 * nz + iz0], s = vs[(size_t)ix0 * nz + iz0];
    double ca = p * p;
    const double cm = s * s;
    double cl = ca - 2 * cm;
    double b = dtx;
    if (iz == top - 1) {
        ca = 2 * cm;
        cl = 0.0;
        b = 2 * b;
    }
    const double damp = kappa * geo[(size_t)ix * nzbc + iz];
    const double temp = 1 - damp * dt;
    coefs[o] = (float)temp;
    coefs[plane + o] = (float)b;
    coefs[2 * plane + o] = (float)b;
    coefs[3 * plane + o] = (float)ca;
    coefs[4 * plane + o] = (float)cl;
    coefs[5 * plane + o] = (float)cm;
}

// smooth2a (core/smooth2a.m:52-76): mean over the (2Nr+1) x (2Nc+1) box cut
// at the edges, NaNs left out of sum and count, NaN where the input is NaN.
// Column-major nz x nx.
__global__ void smooth2a_kernel(const double* __restrict__ in, int nz, int nx, int Nr, int Nc, double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= nz || j >= nx) return;
    double sum = 0.0, cnt = 0.0;
    for (int c = max(j - Nc, 0); c <= min(j + Nc, nx - 1); ++c)
        for (int r = max(i - Nr, 0); r <= min(i + Nr, nz - 1); ++r) {
            const double v = in[(size_t)c * nz + r];
            if (!isnan(v)) {
                sum += v;
                cnt += 1.0;
            }
        }
    const double self = in[(size_t)j * nz + i];
    out[(size_t)j * nz + i] = isnan(self) ? self : sum / cnt;
}

// vs1 = vs + step * dk, then vs1(vs1 < vsmin) = vsmin, vs1(vs1 > vsmax) = vsmax
// (SWD.m:206-207); NaN passes through as in MATLAB.
__global__ void model_step_kernel(const double* __restrict__ vs, const double* __restrict__ dk, double step,
                                  double vsmin, double vsmax, double* __restrict__ out, size_t n)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double v = vs[k] + step * dk[k];
    if (v < vsmin) v = vsmin;
    if (v > vsmax) v = vsmax;
    out[k] = v;
}

// vp = ratio * vs (the field scripts' vp = vs1*1.732 after every model
// change, fieldexamples/outcrop/outcrop.m:217,245,271,301,331), in FP64.
__global__ void model_scale_kernel(const double* __restrict__ vs, double ratio, double* __restrict__ vp, size_t n)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    vp[k] = vs[k] * ratio;
}

// One workgroup, fixed summation order (deterministic): out[0] = min(a, b),
// out[1] = max(a, b), out[2] = sum a^2, out[3] = sum b^2.  NaNs propagate
// into the sums; min / max skip them.
__global__ __launch_bounds__(1024) void model_stats_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                           size_t n, double* __restrict__ out)
{
    __shared__ double smin[1024], smax[1024], sa[1024], sb[1024];
    const int t = threadIdx.x;
    double mn = INFINITY, mx = -INFINITY, qa = 0.0, qb = 0.0;
    for (size_t k = t; k < n; k += 1024) {
        const double x = a[k], y = b[k];
        mn = fmin(mn, fmin(x, y));
        mx = fmax(mx, fmax(x, y));
        qa += x * x;
        qb += y * y;
    }
    smin[t] = mn; smax[t] = mx; sa[t] = qa; sb[t] = qb;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w) {
            smin[t] = fmin(smin[t], smin[t + w]);
            smax[t] = fmax(smax[t], smax[t + w]);
            sa[t] += sa[t + w];
            sb[t] += sb[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = smin[0];
        out[1] = smax[0];
        out[2] = sa[0];
        out[3] = sb[0];
    }
}

// temp, ca, cl, cm1 of a coefficient block (planes temp, b, b1, ca, cl, cm1)
// interleaved per cell: the same values, one 16-byte record per cell
__global__ void pack_coef4_kernel(const float* __restrict__ c, size_t plane, float4* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= plane) return;
    out[i] = make_float4(c[i], c[3 * plane + i], c[4 * plane + i], c[5 * plane + i]);
}

int launch_pack_coef4(hipStream_t st, const float* coefs, size_t plane, float* out)
{
    if ((reinterpret_cast<uintptr_t>(out) & 15) != 0) { set_error("interleaved coefficients need 16-byte alignment"); return kErrArg; }
    const int blk = 256;
    pack_coef4_kernel<<<(unsigned)((plane + blk - 1) / blk), blk, 0, st>>>(coefs, plane, (float4*)out);
    SWD_HIP(hipGetLastError());
    return kOk;
}

int launch_prep_coefs(hipStream_t st, const double* vp, const double* vs, const double* geo, double kappa, int nz,
                      int nx, int nbc, int top, int nzbc, int nxbc, int pitch, size_t plane, double dt, double dtx,
                      float* coefs)
{
    const dim3 blk(128), grd((pitch + 127) / 128, nzbc);
    prep_coefs_kernel<<<grd, blk, 0, st>>>(vp, vs, geo, kappa, nz, nx, nbc, top, nzbc, nxbc, pitch, plane, dt, dtx,
                                           coefs);
    SWD_HIP(hipGetLastError());
    return kOk;
}

int launch_smooth2a(hipStream_t st, const double* in, int nz, int nx, int Nr, int Nc, double* out)
{
    if (Nr < 0 || Nc < 0) { set_error("smooth2a needs Nr, Nc >= 0"); return kErrArg; }
    const dim3 blk(128), grd((nz + 127) / 128, nx);
    smooth2a_kernel<<<grd, blk, 0, st>>>(in, nz, nx, Nr, Nc, out);
    SWD_HIP(hipGetLastError());
    return kOk;
}

int launch_model_step(hipStream_t st, const double* vs, const double* dk, double step, double vsmin, double vsmax,
                      double* out, size_t n)
{
    model_step_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(vs, dk, step, vsmin, vsmax, out, n);
    SWD_HIP(hipGetLastError());
    return kOk;
}

int launch_model_scale(hipStream_t st, const double* vs, double ratio, double* vp, size_t n)
{
    model_scale_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(vs, ratio, vp, n);
    SWD_HIP(hipGetLastError());
    return kOk;
}

int launch_model_stats(hipStream_t st, const double* a, const double* b, size_t n, double* out)
{
    model_stats_kernel<<<1, 1024, 0, st>>>(a, b, n, out);
    SWD_HIP(hipGetLastError());
    return kOk;
}

}  // namespace swd
