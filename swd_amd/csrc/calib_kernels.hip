// Plain device copy of a known byte count.  Its FETCH_SIZE / WRITE_SIZE
// counter readings scale the same counters of the FD kernels to HBM bytes
// (they count L2-to-memory traffic in KiB, aggregated over the XCDs).
#include <hip/hip_runtime.h>

#include "swd_host.h"

namespace swd {

__global__ __launch_bounds__(256) void calib_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

}  // namespace swd

using namespace swd;

extern "C" int swd_calib_copy(long long bytes, int reps, double* ms)
{
    if (bytes < 16 || reps < 1) { set_error("calibration copy needs bytes >= 16 and reps >= 1"); return kErrArg; }
    const size_t n = (size_t)bytes / sizeof(float4);
    DevBuf a, b;
    int rc;
    if ((rc = a.alloc(n * sizeof(float4))) || (rc = b.alloc(n * sizeof(float4)))) return rc;
    SWD_HIP(hipMemset(a.p, 0, a.bytes));
    hipEvent_t e0, e1;
    SWD_HIP(hipEventCreate(&e0));
    SWD_HIP(hipEventCreate(&e1));
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 256 * 64);
    SWD_HIP(hipEventRecord(e0, nullptr));
    for (int r = 0; r < reps; ++r) calib_copy<<<blocks, 256>>>(a.as<float4>(), b.as<float4>(), n);
    SWD_HIP(hipGetLastError());
    SWD_HIP(hipEventRecord(e1, nullptr));
    SWD_HIP(hipEventSynchronize(e1));
    float t = 0;
    SWD_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = t / reps;
    return kOk;
}
