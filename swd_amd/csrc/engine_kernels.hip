// Small kernels of the batched SWD engine: shot reduction of the images and
// the Vs-gradient assembly (SWD.m:193,199-201), in double.
#include "swd_internal.h"

namespace swd {

// imgs: [nshots][3][nz][nx] (cl_img, cm_img, illum) in double.
// Outputs (MATLAB layout, nz x nx column-major): dk_vs, g_cl, g_cm, g_il.
//   g_cl = sum_s cl_img / sum_s illum,  g_cm = sum_s cm_img / sum_s illum,
//   dk_vs = -4 vs g_cl + 2 vs g_cm
__global__ void assemble_gradient_kernel(const double* __restrict__ imgs, int nshots, int nz, int nx,
                                         const double* __restrict__ vs, double* __restrict__ dk,
                                         double* __restrict__ gcl, double* __restrict__ gcm,
                                         double* __restrict__ gil)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    if (ix >= nx || iz >= nz) return;
    const size_t plane = (size_t)nz * nx;
    const size_t i = (size_t)iz * nx + ix;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int s = 0; s < nshots; ++s) {
        const double* p = imgs + (size_t)s * 3 * plane + i;
        a += p[0];
        b += p[plane];
        c += p[2 * plane];
    }
    const size_t o = (size_t)ix * nz + iz;
    const double v = vs[o];
    const double ga = a / c, gb = b / c;
    dk[o] = -4 * v * ga + 2 * v * gb;
    gcl[o] = a;
    gcm[o] = b;
    gil[o] = c;
}

int launch_assemble_gradient(hipStream_t st, const double* imgs, int nshots, int nz, int nx, const double* vs,
                             double* dk, double* gcl, double* gcm, double* gil)
{
    assemble_gradient_kernel<<<dim3(div_up(nx, 128), nz), 128, 0, st>>>(imgs, nshots, nz, nx, vs, dk, gcl, gcm, gil);
    SWD_HIP(hipGetLastError());
    return kOk;
}

}  // namespace swd
