// swd_lhdisppick_noalloc: the dispersion pick of swd_lhdisppick for callers
// that cannot allocate device memory (code under stream capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// Every device buffer the pick kernel needs (the staged spectrum, the observed
// picks, the picks, the misfit, the status flag and the column record) lives in
// one file-scope device array, `scratch`, of kNoAllocScratch bytes per GPU,
// held until the process ends.  Each device's address of the array is taken
// once with hipGetSymbolAddress on the first call on that device; the array's
// name itself is a host-side stand-in and is never passed to a kernel or copy.
// The call issues no hipMalloc / hipFree and creates no stream (it runs on the
// per-thread default stream).  Overlapping calls on one GPU would race on the
// array: a per-device lock serialises them within the process.  The array keeps
// the last call's contents; every region a call reads is written by that call
// first (uploads, or kernel outputs written for every column), so nothing needs
// clearing.
#include <hip/hip_runtime.h>

#include <mutex>

#include "swd_disp.h"
#include "swd_host.h"

namespace swd {

constexpr size_t kNoAllocScratch = (size_t)4 << 20;  // 4 MiB per GPU
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

static std::mutex g_noalloc_setup;
static std::mutex g_noalloc_lock[kNoAllocMaxDevices];
static unsigned char* g_noalloc_ptr[kNoAllocMaxDevices];  // per-device address of `scratch`

static size_t round256(size_t n) { return (n + 255) & ~(size_t)255; }

// Device address of the current device's array (taken once per device).
static int noalloc_scratch(int* dev_out, unsigned char** out)
{
    int dev = 0;
    SWD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) {
        set_error("device %d beyond the %d devices of the no-allocation scratch table", dev, kNoAllocMaxDevices);
        return kErrNoDevice;
    }
    std::lock_guard<std::mutex> g(g_noalloc_setup);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress of the pick scratch array failed on device %d: %s", dev,
                      hipGetErrorString(e));
            return kErrHip;
        }
        g_noalloc_ptr[dev] = (unsigned char*)p;
    }
    *dev_out = dev;
    *out = g_noalloc_ptr[dev];
    return kOk;
}

static int lhdisppick_noalloc_impl(const float* ml, int np, int npair, double vmin, const double* cr_0, int method,
                                   int ini, double* res, double* cr_pre)
{
    if (!ml || !cr_0 || !cr_pre || np < 1 || npair < 1) { set_error("bad LHDispPick arguments"); return kErrArg; }
    if (npair > kMaxPair) { set_error("npair %d exceeds %d", npair, kMaxPair); return kErrArg; }
    if (np > kPickMaxNp) { set_error("np %d exceeds %d", np, kPickMaxNp); return kErrArg; }
    if (method == 1 && (ini < 1 || ini > npair)) { set_error("ini %d outside 1..%d", ini, npair); return kErrArg; }
    // scratch layout: column record | misfit | status | cr_0 | cr_pre | ml
    const size_t off_res = 256, off_stat = 512, off_cr0 = 768;
    const size_t off_pre = off_cr0 + round256((size_t)npair * sizeof(double));
    const size_t off_ml = off_pre + round256((size_t)npair * sizeof(double));
    const size_t need = off_ml + (size_t)np * npair * sizeof(float);
    if (need > kNoAllocScratch) {
        set_error("a %d x %d spectrum needs %zu bytes of the no-allocation scratch; it holds %zu", np, npair, need,
                  kNoAllocScratch);
        return kErrArg;
    }
    int rc, dev = 0;
    unsigned char* base = nullptr;
    if ((rc = noalloc_scratch(&dev, &base))) return rc;
    std::lock_guard<std::mutex> lock(g_noalloc_lock[dev]);
    hipStream_t st = hipStreamPerThread;
    const DispCol c{1, 0, 0, 1, np};
    int stv = 0;
    // everything queued below is waited for before the lock is released, on
    // the error paths too: the next call reuses the array, and the copies read
    // the caller's buffers
    rc = [&]() -> int {
        SWD_HIP(hipMemcpyAsync(base, &c, sizeof c, hipMemcpyHostToDevice, st));
        SWD_HIP(hipMemcpyAsync(base + off_cr0, cr_0, (size_t)npair * sizeof(double), hipMemcpyHostToDevice, st));
        SWD_HIP(hipMemcpyAsync(base + off_ml, ml, (size_t)np * npair * sizeof(float), hipMemcpyHostToDevice, st));
        const float* dml = (const float*)(base + off_ml);
        const double* dcr0 = (const double*)(base + off_cr0);
        double* dpre = (double*)(base + off_pre);
        int r;
        if (method == 1) {
            if ((r = launch_pick(st, dml, nullptr, (const DispCol*)base, 1, np, npair, ini, 2.f, 4.f, vmin, 0, 0,
                                 dcr0, dpre, (double*)(base + off_res), (int*)(base + off_stat))))
                return r;
        } else if ((r = launch_argmax_pick(st, dml, np, npair, vmin, dcr0, dpre))) {
            return r;
        }
        SWD_HIP(hipMemcpyAsync(cr_pre, dpre, (size_t)npair * sizeof(double), hipMemcpyDeviceToHost, st));
        if (method == 1) SWD_HIP(hipMemcpyAsync(&stv, base + off_stat, sizeof(int), hipMemcpyDeviceToHost, st));
        return kOk;
    }();
    const hipError_t se = hipStreamSynchronize(st);
    if (rc) return rc;
    if (se != hipSuccess) { set_error("pick stream: %s", hipGetErrorString(se)); return kErrHip; }
    if (stv) { set_error("interp1 needs at least two kept pick columns"); return kErrPick; }
    double acc = 0;
    for (int i = 0; i < npair; ++i) { const double d = cr_pre[i] - cr_0[i]; acc += d * d; }
    if (res) *res = acc / npair;
    return kOk;
}

}  // namespace swd

extern "C" int swd_lhdisppick_noalloc(const float* ml, int np, int npair, double vmin, const double* cr_0,
                                      int method, int ini, double* res, double* cr_pre)
{
    return swd::lhdisppick_noalloc_impl(ml, np, npair, vmin, cr_0, method, ini, res, cr_pre);
}
