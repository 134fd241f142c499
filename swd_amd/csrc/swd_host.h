// Host-side helpers of libswd: RAII device buffers and streams, and the
// declarations of the kernel launchers defined in the .hip files.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "swd_internal.h"

namespace swd {

const char* last_error();
int order_pt(int fd_order);
int grid_pitch(int nxbc);
int fd_tile_rows();
void set_fd_tile_rows(int tz);
int fd_force_plain();

template <typename T>
void pack_grid(const T* colmajor, int nzbc, int nxbc, int pitch, T* out);
template <typename T>
void zero_ring(T* plane, int nzbc, int nxbc, int pitch);

int prepare_model(const double* vp, const double* vs, int nz, int nx, int nbc, double dx, double dt, int fd_order,
                  double* temp, double* ca, double* cl, double* cm, double* cm1, double* b, double* b1);
int ricker(double f, double dt, int nt, double* w);
void damp_geometry(int nz, int nx, int nbc, double dx, int top, double* g);
double damp_kappa(double vmin, int nbc, double dx);

// model update kernels (model_kernels.hip)
int launch_prep_coefs(hipStream_t st, const double* vp, const double* vs, const double* geo, double kappa, int nz,
                      int nx, int nbc, int top, int nzbc, int nxbc, int pitch, size_t plane, double dt, double dtx,
                      float* coefs);
int launch_smooth2a(hipStream_t st, const double* in, int nz, int nx, int Nr, int Nc, double* out);
int launch_model_step(hipStream_t st, const double* vs, const double* dk, double step, double vsmin, double vsmax,
                      double* out, size_t n);
int launch_model_scale(hipStream_t st, const double* vs, double ratio, double* vp, size_t n);
// out (4 x plane floats, 16-byte aligned) = temp, ca, cl, cm1 of the coefficient
// block interleaved per cell (FdJob::c4)
int launch_pack_coef4(hipStream_t st, const float* coefs, size_t plane, float* out);
int launch_model_stats(hipStream_t st, const double* a, const double* b, size_t n, double* out);

// First-fit sub-allocator over one allocation made once (device memory via
// hipMalloc or pinned host memory via hipHostMalloc), with coalescing frees.
// The gradient server builds every request's buffers from one, so admitting
// and retiring requests never calls hipMalloc or hipFree.  Thread-safe.
class Arena {
public:
    Arena() = default;
    Arena(const Arena&) = delete;
    Arena& operator=(const Arena&) = delete;
    ~Arena();
    int init_device(size_t bytes);
    int init_host(size_t bytes);
    void* take(size_t n);             // nullptr when no free block fits
    void give(void* p, size_t n);
    size_t capacity() const { return cap_; }
    size_t used() const;
    size_t largest_free() const;

private:
    struct Impl;
    Impl* impl_ = nullptr;
    char* base_ = nullptr;
    size_t cap_ = 0;
    bool host_ = false;
};

// DevBuf allocations made by this thread while an ArenaScope is alive come
// from its arena instead of hipMalloc.
Arena* current_arena();
struct ArenaScope {
    Arena* prev;
    explicit ArenaScope(Arena* a);
    ~ArenaScope();
};

// Device allocation owned by the host object (freed on destruction).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    Arena* arena = nullptr;  // set when the memory came from an arena
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), arena(o.arena)
    {
        o.p = nullptr; o.bytes = 0; o.arena = nullptr;
    }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p; bytes = o.bytes; arena = o.arena;
            o.p = nullptr; o.bytes = 0; o.arena = nullptr;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release()
    {
        if (p) {
            if (arena) arena->give(p, bytes);
            else (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
        arena = nullptr;
    }
    int alloc(size_t n)
    {
        release();
        if (n == 0) n = 256;
        if (Arena* a = current_arena()) {
            p = a->take(n);
            if (!p) {
                set_error("arena has no free block of %zu bytes (%zu of %zu in use)", n, a->used(), a->capacity());
                return kErrNoMem;
            }
            arena = a;
            bytes = n;
            return kOk;
        }
        if (hipMalloc(&p, n) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            set_error("hipMalloc of %zu bytes failed", n);
            return kErrNoMem;
        }
        bytes = n;
        return kOk;
    }
    int upload(const void* h, size_t n, hipStream_t st)
    {
        int rc = alloc(n);
        if (rc) return rc;
        SWD_HIP(hipMemcpyAsync(p, h, n, hipMemcpyHostToDevice, st));
        return kOk;
    }
    template <typename U>
    U* as() const { return (U*)p; }
};

// A default-priority HIP stream owned by the host object.
struct Stream {
    hipStream_t s = nullptr;
    int create()
    {
        SWD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        return kOk;
    }
    ~Stream()
    {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
};

// Timing events, destroyed on every return path.
struct EventSet {
    std::vector<hipEvent_t> ev;
    explicit EventSet(size_t n) : ev(n, nullptr) {}
    EventSet(const EventSet&) = delete;
    EventSet& operator=(const EventSet&) = delete;
    ~EventSet()
    {
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    int create()
    {
        for (auto& e : ev) SWD_HIP(hipEventCreate(&e));
        return kOk;
    }
    int record(size_t i, hipStream_t st)
    {
        SWD_HIP(hipEventRecord(ev[i], st));
        return kOk;
    }
};

struct MarchPlan {
    std::vector<BlockMap> bmap;
    int nblocks = 0;
};

// Tile every job of a batch for a march launch: out_cols output columns per
// wavefront (62 single-step, 64 - 2 NT time-blocked) and tz rows per band.
// The block list is in dispatch order (entry i runs as workgroup i); for a
// time-blocked plan (NT > 1, adj: the adjoint kernel) it is arranged for
// the XCDs' L2s and the CUs' load balance.
template <typename T>
int build_march_plan(std::vector<FdJob<T>>& jobs, int tz, MarchPlan& plan, int out_cols = kMarchOut, int NT = 1,
                     bool adj = false);

template <typename T>
int launch_fwd_march(hipStream_t st, const FdJob<T>* d_jobs, const BlockMap* d_bmap, int nblocks, int step,
                     int launch);
template <typename T>
int launch_adj_march(hipStream_t st, const FdJob<T>* d_jobs, const BlockMap* d_bmap, int nblocks, int step,
                     int launch);
// b1: some job's b1 plane differs from its b plane (jobs whose planes are
// equal point b1 at b, see jobs_b1)
template <typename T>
int launch_fwd_tb(hipStream_t st, const FdJob<T>* d_jobs, const BlockMap* d_bmap, int nblocks, int step, int launch,
                  int nt_block, bool b1);
template <typename T>
int launch_adj_tb(hipStream_t st, const FdJob<T>* d_jobs, const BlockMap* d_bmap, int nblocks, int step, int launch,
                  int nt_block, bool b1);

template <typename T>
int tb_blocks_per_cu(bool adjoint, int nt_block, bool b1 = false);

// whether any job reads a b1 plane of its own (J.b1 != J.b)
template <typename T>
inline bool jobs_b1(const std::vector<FdJob<T>>& jobs)
{
    for (const auto& J : jobs)
        if (J.b1 != J.b) return true;
    return false;
}

// Time-blocking settings: steps per launch (1 = single-step kernels) and band
// rows of the time-blocked kernels (0 = chosen from occupancy); process-wide
// defaults for the MEX path.
int fd_time_block();
int fd_tb_rows();
void set_fd_time_block(int nt_block, int rows);

// Band rows for a time-blocked launch over `jobs` when none are given: the
// band height that minimises (waves rounds) x (rows loaded per wave), with the
// resident-wave count taken from the occupancy API.
template <typename T>
int auto_band_rows(const std::vector<FdJob<T>>& jobs, bool adjoint, int nt_block);

// What one run_march call did: mean event-timed duration of one launch of the
// dominant kernel (sampled), event-timed span of the whole launch sequence,
// steps per time-blocked launch (1 if none) and band rows used.
struct MarchStats {
    float kernel_us = 0, span_ms = 0;
    int nt_block = 1, rows = 0;
};

// Advance a batch of jobs (all starting together) over `steps` forward or
// adjoint steps with time-blocked launches plus single-step remainders.
template <typename T>
int run_march(hipStream_t st, std::vector<FdJob<T>>& jobs, bool adjoint, int steps, int nt_block, int tb_rows,
              int ss_rows, MarchStats* stats);

// The same march as captured HIP graphs (the time-blocked launches, then the
// single-step remainder), replayed while the jobs and the launch parameters
// stay the same (model, wavelet and adjoint-source contents may change
// between replays; pointers and sizes may not).  The graphs hold the launches
// only: whatever zeroes the state runs before them on the stream.
// stats->kernel_us is the dominant graph's event-timed span over its launch
// count (so it includes the gaps between its kernels).
struct MarchGraph {
    std::vector<unsigned char> key;  // launch parameters and job tables it was captured for
    std::vector<DevBuf> djt, dbt;    // job tables and block maps of the time-blocked segments
    DevBuf djs, dbs;                 // those of the single-step remainder
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    int launches[2] = {0, 0};
    int nt_block = 1, rows = 0;
    bool ready = false;
    MarchGraph() = default;
    MarchGraph(const MarchGraph&) = delete;
    MarchGraph& operator=(const MarchGraph&) = delete;
    ~MarchGraph()
    {
        reset();
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    void reset()
    {
        for (int i = 0; i < 2; ++i) {
            if (exec[i]) (void)hipGraphExecDestroy(exec[i]);
            if (graph[i]) (void)hipGraphDestroy(graph[i]);
            exec[i] = nullptr;
            graph[i] = nullptr;
            launches[i] = 0;
        }
        ready = false;
        key.clear();
    }
};
template <typename T>
int run_march_graph(hipStream_t st, std::vector<FdJob<T>>& jobs, bool adjoint, int steps, int nt_block, int tb_rows,
                    int ss_rows, MarchGraph& g, MarchStats* stats);
template <typename T>
int plain_forward_step(hipStream_t st, const FdJob<T>& J, T* const w[4], int it);
template <typename T>
int plain_adjoint_step(hipStream_t st, const FdJob<T>& J, int it);

}  // namespace swd
