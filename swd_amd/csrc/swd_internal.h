// Internal types shared by the SWD HIP kernels and the host library.
// Not part of the C-ABI (include/swd_abi.h is).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace swd {

// Status codes (mirrored in include/swd_abi.h).
enum : int {
    kOk = 0,
    kErrArg = -2,
    kErrCfl = -3,
    kErrNoMem = -4,
    kErrHip = -5,
    kErrPick = -6,
    kErrNoDevice = -7,
    kErrBusy = -8,
};

void set_error(const char* fmt, ...);

#define SWD_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::swd::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                             __FILE__, __LINE__);                                       \
            return ::swd::kErrHip;                                                      \
        }                                                                               \
    } while (0)

// One shot (forward) or one adjoint solve in a batched FD launch.
//
// Internal grid layout (all fields of one job share it): row-major, x fastest,
// `pitch` elements per row, nzbc rows; row = depth index (0..nzbc-1),
// col = distance index (0..nxbc-1) of the reference's padded grid.  The outer
// ring (row 0, row nzbc-1, col 0, col nxbc-1) of every state array stays 0.
// State is ping-ponged: kernels read cur[] and write nxt[].
template <typename T>
struct FdJob {
    // state, [0] = read buffer of this step, [1] = write buffer (swapped by the host)
    T* uu[2];
    T* ww[2];
    T* xx[2];
    T* zz[2];
    T* xz[2];
    // model coefficients (read-only, shared by all shots of a model)
    const T* temp;
    const T* b;
    const T* b1;
    const T* ca;
    const T* cl;
    const T* cm1;
    // float jobs: temp, ca, cl, cm1 interleaved per cell ([row][pitch][4], 16 B
    // per cell), which the time-blocked kernels read with one 16-byte load per
    // lane instead of four; null for double jobs
    const T* c4;
    // forward: wavelet s[nt]; adjoint: adjoint source [nt][ng] (g fastest)
    const T* src;
    // forward receivers [nt][ng] (g fastest); trace_u may be null
    T* trace_w;
    T* trace_u;
    // snapshots: nsf = 3, [nt/in_wf][nz][nx][3] (fux, bwz, fuz+bwx of a cell
    // side by side: one 12-byte access per lane); nsf = 4, [nt/in_wf][4][nz][nx]
    // planes (fux, fuz, bwx, bwz)
    T* snap;
    // adjoint images [3][nz][nx] (cl_img, cm_img, illum), accumulated in double
    double* img;
    // FP32 partial sums of the images, [nz][nx][3] (required by the fused
    // float kernels, unused by the others): every imaging step adds into
    // them, and imaging step k with k % kImgFlush == 0 (the last one, it = 0,
    // included) moves them into img and clears them
    float* imgp;
    int nzbc, nxbc, pitch, nz, nx, nbc, pt;
    int sz, sx, stype;      // 0-based source cell and source type (1..5; 0 = none)
    int gz, gx, dg, ng;     // 0-based receiver row / first column, spacing, count
    int nt, in_wf;          // steps and snapshot interval
    int start;              // launch step at which this job's step 0 runs
    int launch0;            // launch index of the job's first launch (ping-pong parity)
    int nsf;                // snapshot fields (3 or 4); 0 = no snapshots
    int tiles_x, tiles_z, tz;  // march tiling (fused kernels)
    T dtx, dt;
};

// Imaging steps per FP64 flush of the FP32 partial images (FdJob::imgp).
// 32: dk_vs moves 3.6e-8 relative L2 against 16.  At the production shape a
// flush launch of the adjoint takes about 30 us more than the others, so
// flushing every 32 imaging steps costs about 1 ms per gradient.
constexpr int kImgFlush = 32;

// Block -> (job, first tile) map entry for fused launches.
struct BlockMap {
    int job;
    int tile0;
};

constexpr int kMarchLanes = 64;          // one wavefront
constexpr int kMarchOut = kMarchLanes - 2;  // output columns per wavefront (1-col halo each side)
constexpr int kMarchWaves = 4;           // wavefronts per workgroup

inline int div_up(int a, int b) { return (a + b - 1) / b; }

// XCD-grouped block order of the time-blocked launches: workgroup i (placed
// on XCD i mod 8 by the dispatcher, each XCD with its own L2) runs entry
// xcd_block(i, n) of the job-major block list.  Within every group of 8 G workgroups,
// XCD x gets G consecutive entries, so neighbouring tiles of a shot, which
// share halo rows and columns, mostly share one L2; a trailing partial group
// keeps the identity order.  A bijection on [0, n).  G = 8: the production
// adjoint ran 10% faster than with the round-robin order, and an uneven
// multi-model batch (jobs ending at different launches) as fast.  Giving
// each XCD one contiguous eighth of the list instead leaves XCDs whose jobs
// have ended idle: that batch ran 30% slower.
__host__ __device__ inline int xcd_block(int i, int n)
{
    constexpr int G = 8, S = 8 * G;
    if (i >= n / S * S) return i;
    const int s = i / S, k = i - s * S;
    return s * S + (k & 7) * G + (k >> 3);
}

}  // namespace swd
