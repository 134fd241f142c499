/*
 * TEST INFRASTRUCTURE ONLY - CPU oracle for the SWD hot path (FD part).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library (oracle/_build/libswd_oracle.so); the product never does.
 *
 * Contents
 *   oracle_staggerfd_{d,f}  forward time loop   core/eigen_staggerfd_single.cpp:98-383
 *   oracle_e2drtm_{d,f}     adjoint + imaging   core/eigen_e2drtm_single.cpp:69-297
 *   oracle_prep             wrapper model prep  core/staggerfd_eigen.m:37-107,
 *                           core/calparam.m:3-5, core/pad.m:3-9, core/damp_circle.m:5-67
 *   oracle_ricker           wavelet             core/ricker.m:18-35
 *   oracle_batch_{d,f}      one shot per thread (the reference runs one shot per
 *                           single-threaded parpool worker, SWD.m:171)
 *
 * The FP64 MEX (eigen_*_double) ships only as binaries, which are not opened:
 * its algorithm is taken to be the single-precision .cpp evaluated in double.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define REAL double
#define SFX(name) name##_d
#include "swd_oracle_fd.inc"
#undef REAL
#undef SFX

#define REAL float
#define SFX(name) name##_f
#include "swd_oracle_fd.inc"
#undef REAL
#undef SFX

/* pad.m:3-9 - edge replicate: nbc columns left/right, pad_top rows on top and
 * nbc rows at the bottom (isfs), or nbc rows on top (no free surface). */
static void pad_cm(const double *p0, int nz, int nx, int nbc, int top, double *p)
{
    const int nzbc = nz + top + nbc, nxbc = nx + 2 * nbc;
    for (int j = 0; j < nxbc; ++j) {
        int js = j - nbc;
        if (js < 0) js = 0;
        if (js > nx - 1) js = nx - 1;
        for (int i = 0; i < nzbc; ++i) {
            int is = i - top;
            if (is < 0) is = 0;
            if (is > nz - 1) is = nz - 1;
            p[(size_t)j * nzbc + i] = p0[(size_t)js * nz + is];
        }
    }
}

/* damp_circle.m:5-67, 1-based loops restated with 0-based storage. */
void oracle_damp_circle(double vmin, int nzbc, int nxbc, int nbc, double dx, int isfs,
                        int pad_top, double *damp)
{
    const int nz = nzbc - nbc - pad_top, nx = nxbc - 2 * nbc;
    const double a = (nbc - 1) * dx;
    const double kappa = 3.0 * vmin * log(10000000.0) / (2.0 * a);
    double *d1 = (double *)malloc(sizeof(double) * (size_t)nbc);
    for (int k = 0; k < nbc; ++k) {
        const double r = k * dx / a;
        d1[k] = kappa * (r * r);
    }
#define D(iz, ix) damp[(size_t)((ix) - 1) * nzbc + ((iz) - 1)]
    memset(damp, 0, sizeof(double) * (size_t)nzbc * nxbc);
    for (int iz = 1 + pad_top; iz <= nz + pad_top; ++iz) {
        for (int k = 1; k <= nbc; ++k) D(iz, k) = d1[nbc - k];
        for (int k = 1; k <= nbc; ++k) D(iz, nx + nbc + k) = d1[k - 1];
    }
    for (int ix = 1 + nbc; ix <= nx + nbc; ++ix) {
        for (int k = 1; k <= pad_top; ++k) D(k, ix) = isfs ? 0.0 : d1[nbc - k];
        for (int k = 1; k <= nbc; ++k) D(nzbc - nbc + k, ix) = d1[k - 1];
    }
    if (!isfs) {
        for (int iz = 1; iz <= pad_top; ++iz)
            for (int ix = 1; ix <= nbc; ++ix) {
                const double dist = sqrt((double)(ix - nbc - 1) * (ix - nbc - 1) + (double)(iz - pad_top - 1) * (iz - pad_top - 1));
                D(iz, ix) = kappa * ((dist / nbc) * (dist / nbc));
            }
        for (int iz = 1; iz <= pad_top; ++iz)
            for (int ix = nx + nbc + 1; ix <= nxbc; ++ix) {
                const double dist = sqrt((double)(ix - nx - nbc) * (ix - nx - nbc) + (double)(iz - pad_top - 1) * (iz - pad_top - 1));
                D(iz, ix) = kappa * ((dist / nbc) * (dist / nbc));
            }
    }
    for (int iz = nz + pad_top + 1; iz <= nzbc; ++iz)
        for (int ix = 1; ix <= nbc; ++ix) {
            const double dist = sqrt((double)(ix - nbc - 1) * (ix - nbc - 1) + (double)(iz - nz - pad_top) * (iz - nz - pad_top));
            D(iz, ix) = kappa * ((dist / nbc) * (dist / nbc));
        }
    for (int iz = nz + pad_top + 1; iz <= nzbc; ++iz)
        for (int ix = nx + nbc + 1; ix <= nxbc; ++ix) {
            const double dist = sqrt((double)(ix - nbc - nx) * (ix - nbc - nx) + (double)(iz - nz - pad_top) * (iz - nz - pad_top));
            D(iz, ix) = kappa * ((dist / nbc) * (dist / nbc));
        }
#undef D
    free(d1);
}

/*
 * staggerfd_eigen.m:37-107 for parameter_type 0 and isfs 1 (the only
 * configuration the wrapper supports: b, b1 and cm1 are defined only inside
 * its `if (isfs)`).  den scales the density (the wrapper fixes den = 1; other
 * values exist for the density-invariance test).  Outputs are nzbc x nxbc with
 * nzbc = nz + pad_top + nbc, nxbc = nx + 2 nbc, pad_top = (fd_order-20)/2+1.
 */
int oracle_prep(const double *vp, const double *vs, int nz, int nx, int nbc, double dx,
                double dt, int fd_order, double den, double *temp, double *ca, double *cl,
                double *cm, double *cm1, double *b, double *b1)
{
    if (nz < 1 || nx < 1 || nbc < 2) return -2;
    if (fd_order != 22 && fd_order != 24 && fd_order != 26 && fd_order != 28) return -2;
    const int pad_top = (fd_order - 20) / 2 + 1;
    const int nzbc = nz + pad_top + nbc, nxbc = nx + 2 * nbc;
    const size_t n0 = (size_t)nz * nx, n = (size_t)nzbc * nxbc;
    const double dtx = dt / dx;
    double *ca0 = (double *)malloc(sizeof(double) * n0), *cm0 = (double *)malloc(sizeof(double) * n0);
    double *cl0 = (double *)malloc(sizeof(double) * n0), *damp = (double *)malloc(sizeof(double) * n);
    double vmin = INFINITY;
    for (size_t k = 0; k < n0; ++k) {
        ca0[k] = (vp[k] * vp[k]) * den;          /* calparam.m:3 */
        cm0[k] = (vs[k] * vs[k]) * den;          /* calparam.m:4 */
        cl0[k] = ca0[k] - 2 * cm0[k];            /* calparam.m:5 */
        const double m = vp[k] < vs[k] ? vp[k] : vs[k];
        if (m < vmin) vmin = m;                  /* staggerfd_eigen.m:89 */
    }
    pad_cm(cm0, nz, nx, nbc, pad_top, cm);
    pad_cm(cl0, nz, nx, nbc, pad_top, cl);
    pad_cm(ca0, nz, nx, nbc, pad_top, ca);
    oracle_damp_circle(vmin, nzbc, nxbc, nbc, dx, 1, pad_top, damp);
    for (size_t k = 0; k < n; ++k) temp[k] = 1 - damp[k] * dt;   /* staggerfd_eigen.m:91 */
    const int pr = pad_top - 1;                                  /* MATLAB row pad_top */
    for (int j = 0; j < nxbc; ++j) {
        const size_t o = (size_t)j * nzbc + pr;
        cm[o] = 1.0 * cm[o];
        ca[o] = 2 * cm[o];
        cl[o] = 0.0;
    }
    for (size_t k = 0; k < n; ++k) b[k] = dtx / den;
    for (int j = 0; j < nxbc; ++j) b[(size_t)j * nzbc + pr] = 2 * b[(size_t)j * nzbc + pr];
    memcpy(b1, b, sizeof(double) * n);
    memcpy(cm1, cm, sizeof(double) * n);
    free(ca0); free(cm0); free(cl0); free(damp);
    return 0;
}

/* ricker.m:18-35 with nargin == 3: nt samples, error when nt < nw. */
int oracle_ricker(double f, double dt, int nt, double *w)
{
    double nw = 2.2 / f / dt;
    nw = 2 * floor(nw / 2) + 1;
    const int inw = (int)nw;
    const double nc = floor(nw / 2);
    if (nt < inw) return -2;
    for (int k = 0; k < nt; ++k) w[k] = 0;
    for (int k = 1; k <= inw; ++k) {
        const double alpha = (nc - k + 1) * f * dt * M_PI;
        const double beta = alpha * alpha;
        w[k - 1] = (1. - beta * 2) * exp(-beta);
    }
    return 0;
}

/*
 * CPU baseline driver: `nshots` independent forward + adjoint solves of one
 * model, one shot per thread (threads <= 0: all OpenMP threads).  Each shot
 * uses its own source column sx[s] (1-based, padded grid) and adjoint source
 * adj + s*nt*ng.  Images are summed over shots into the three nz*nx outputs.
 * Returns the number of shots that failed.
 */
#define BATCH_DRIVER(SUF, T)                                                                 \
    int oracle_batch_##SUF(const T *ivf, const T *ivb, const T *temp, const T *ca,          \
                           const T *cl, const T *cm, const T *cm1, const T *b, const T *b1, \
                           const T *s, const int *sx, int nshots, const T *adj,             \
                           T *seismo_w, T *g_cl, T *g_cm, T *g_il, int threads)             \
    {                                                                                        \
        const int nt = (int)ivf[0], ng = (int)ivf[4], in_wf = (int)ivf[12];                 \
        const int nz = (int)ivf[13], nx = (int)ivf[14];                                      \
        const size_t nsnap = (size_t)nz * nx * (size_t)(nt / in_wf);                          \
        const size_t nimg = (size_t)nz * nx;                                                  \
        int failed = 0;                                                                      \
        for (size_t k = 0; k < nimg; ++k) { g_cl[k] = 0; g_cm[k] = 0; g_il[k] = 0; }         \
        if (threads <= 0) threads = 1;                                                       \
        _Pragma("omp parallel for schedule(dynamic, 1) num_threads(threads) reduction(+ : failed)") \
        for (int sh = 0; sh < nshots; ++sh) {                                                \
            T iv[16];                                                                        \
            memcpy(iv, ivf, sizeof(iv));                                                     \
            iv[6] = (T)sx[sh];                                                               \
            T *su = (T *)malloc(sizeof(T) * (size_t)nt * ng);                                \
            T *sn = (T *)malloc(sizeof(T) * nsnap * 4);                                      \
            T *im = (T *)malloc(sizeof(T) * nimg * 3);                                       \
            int rc = -4;                                                                     \
            if (su && sn && im) {                                                            \
                rc = oracle_staggerfd_##SUF(iv, temp, ca, cl, cm, cm1, b, b1, s, su,         \
                                            seismo_w + (size_t)sh * nt * ng, sn, sn + nsnap, \
                                            sn + 2 * nsnap, sn + 3 * nsnap);                 \
                if (rc == 0)                                                                 \
                    rc = oracle_e2drtm_##SUF(ivb, temp, ca, cl, cm, cm1, b, b1,              \
                                             adj + (size_t)sh * nt * ng, sn, sn + nsnap,     \
                                             sn + 2 * nsnap, sn + 3 * nsnap, im, im + nimg,  \
                                             im + 2 * nimg);                                 \
            }                                                                                \
            if (rc == 0) {                                                                   \
                _Pragma("omp critical")                                                      \
                for (size_t k = 0; k < nimg; ++k) {                                          \
                    g_cl[k] += im[k]; g_cm[k] += im[nimg + k]; g_il[k] += im[2 * nimg + k];   \
                }                                                                            \
            } else {                                                                         \
                failed += 1;                                                                 \
            }                                                                                \
            free(su); free(sn); free(im);                                                    \
        }                                                                                    \
        return failed;                                                                       \
    }

BATCH_DRIVER(d, double)
BATCH_DRIVER(f, float)

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
