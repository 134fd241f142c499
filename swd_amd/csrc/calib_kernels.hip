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

// The copy that sets the practical ceiling: four 16-byte non-temporal loads
// in flight per lane, then four non-temporal stores.  On one MI355X (1 GiB)
// it moved 6.0 TB/s of read + write against 4.7 for calib_copy; a read-only
// stream reached 6.3 and a write-only one 4.4.
__global__ __launch_bounds__(256) void copy_peak_kernel(const float4* __restrict__ a, float4* __restrict__ b,
                                                        size_t n)
{
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4* __restrict__ src = (const f4*)a;
    f4* __restrict__ dst = (f4*)b;
    constexpr int U = 4;
    const size_t chunk = (size_t)256 * U;
    for (size_t base = (size_t)blockIdx.x * chunk; base < n; base += (size_t)gridDim.x * chunk) {
        f4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t i = base + (size_t)u * 256 + threadIdx.x;
            if (i < n) v[u] = __builtin_nontemporal_load(src + i);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t i = base + (size_t)u * 256 + threadIdx.x;
            if (i < n) __builtin_nontemporal_store(v[u], dst + i);
        }
    }
}

// Source pattern of the copies: every 16-byte element distinct, so an element
// a copy skipped or wrote to the wrong place is seen by copy_check.
__device__ inline uint4 copy_pattern(size_t i)
{
    const unsigned lo = (unsigned)i, hi = (unsigned)(i >> 32);
    return make_uint4(lo, ~lo, lo * 2654435761u + hi, hi ^ 0x5a5a5a5au);
}

__global__ __launch_bounds__(256) void copy_fill(uint4* __restrict__ a, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) a[i] = copy_pattern(i);
}

// bad is set to 1 by any lane whose destination element differs from the
// pattern (a plain store: every writer writes the same value)
__global__ __launch_bounds__(256) void copy_check(const uint4* __restrict__ b, size_t n, int* __restrict__ bad)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint4 v = b[i], w = copy_pattern(i);
        if (v.x != w.x || v.y != w.y || v.z != w.z || v.w != w.w) *bad = 1;
    }
}

}  // namespace swd

using namespace swd;

template <typename F>
static int time_copies(long long bytes, int reps, double* ms, F launch)
{
    if (bytes < 16 || reps < 1) { set_error("copy needs bytes >= 16 and reps >= 1"); return kErrArg; }
    const size_t n = (size_t)bytes / sizeof(float4);
    DevBuf a, b, bad;
    int rc;
    if ((rc = a.alloc(n * sizeof(float4))) || (rc = b.alloc(n * sizeof(float4))) || (rc = bad.alloc(sizeof(int))))
        return rc;
    const int fill_blocks = (int)std::min<size_t>((n + 255) / 256, 256 * 64);
    copy_fill<<<fill_blocks, 256>>>(a.as<uint4>(), n);
    SWD_HIP(hipGetLastError());
    SWD_HIP(hipMemset(b.p, 0xff, b.bytes));
    SWD_HIP(hipMemset(bad.p, 0, sizeof(int)));
    hipEvent_t e0, e1;
    SWD_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        set_error("hipEventCreate failed");
        return kErrHip;
    }
    rc = kOk;
    float t = 0;
    if (hipEventRecord(e0, nullptr) != hipSuccess) rc = kErrHip;
    for (int r = 0; r < reps && rc == kOk; ++r) launch(a.as<float4>(), b.as<float4>(), n);
    if (rc == kOk && (hipGetLastError() != hipSuccess || hipEventRecord(e1, nullptr) != hipSuccess ||
                      hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess))
        rc = kErrHip;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != kOk) { set_error("copy kernel or its timing events failed"); return rc; }
    // the timed copies must have moved every element: a timing of a copy
    // that skipped part of the buffer would overstate the rate
    copy_check<<<fill_blocks, 256>>>(b.as<uint4>(), n, bad.as<int>());
    SWD_HIP(hipGetLastError());
    int nbad = 1;
    SWD_HIP(hipMemcpy(&nbad, bad.p, sizeof(int), hipMemcpyDeviceToHost));
    if (nbad) {
        set_error("the copy kernel left part of its %zu-byte destination unwritten or wrong", n * sizeof(float4));
        return kErrHip;
    }
    if (ms) *ms = t / reps;
    return kOk;
}

extern "C" int swd_calib_copy(long long bytes, int reps, double* ms)
{
    return time_copies(bytes, reps, ms, [](const float4* a, float4* b, size_t n) {
        const int blocks = (int)std::min<size_t>((n + 255) / 256, 256 * 64);
        calib_copy<<<blocks, 256>>>(a, b, n);
    });
}

extern "C" int swd_copy_peak(long long bytes, int reps, double* ms)
{
    return time_copies(bytes, reps, ms, [](const float4* a, float4* b, size_t n) {
        const int blocks = (int)std::min<size_t>((n + 1023) / 1024, 16384);
        copy_peak_kernel<<<blocks, 256>>>(a, b, n);
    });
}
