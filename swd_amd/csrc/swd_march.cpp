// Host driver of the fused FD march kernels: tiling, launch sequence
// (time-blocked launches plus a single-step remainder), ping-pong parity and
// event timing of one launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <vector>

#include "swd_host.h"

namespace swd {

static std::atomic<int> g_nt_block{4};
static std::atomic<int> g_tb_rows{32};

int fd_time_block() { return g_nt_block.load(); }
int fd_tb_rows() { return g_tb_rows.load(); }
void set_fd_time_block(int nt_block, int rows)
{
    g_nt_block.store(std::min(std::max(nt_block, 1), 6));
    g_tb_rows.store(std::max(rows, 1));
}

template <typename T>
int build_march_plan(std::vector<FdJob<T>>& jobs, int tz, MarchPlan& plan, int out_cols)
{
    plan.bmap.clear();
    for (size_t j = 0; j < jobs.size(); ++j) {
        FdJob<T>& J = jobs[j];
        J.tz = tz;
        J.tiles_x = div_up(J.nxbc - 2, out_cols);
        J.tiles_z = div_up(J.nzbc - 2, tz);
        const int tiles = J.tiles_x * J.tiles_z;
        for (int t = 0; t < tiles; t += kMarchWaves) plan.bmap.push_back(BlockMap{(int)j, t});
    }
    plan.nblocks = (int)plan.bmap.size();
    if (plan.nblocks == 0) { set_error("empty FD batch"); return kErrArg; }
    return kOk;
}
template int build_march_plan<float>(std::vector<FdJob<float>>&, int, MarchPlan&, int);
template int build_march_plan<double>(std::vector<FdJob<double>>&, int, MarchPlan&, int);

template <typename T>
int run_march(hipStream_t st, std::vector<FdJob<T>>& jobs, bool adj, int steps, int nt_block, int tb_rows,
              int ss_rows, float* kernel_us, int* launches_tb)
{
    bool tb_ok = nt_block >= 2;
    for (const auto& J : jobs) {
        if (J.stype != 0 && J.stype != 5) tb_ok = false;
        if (J.start != 0 || J.launch0 != 0) { set_error("run_march jobs must start together"); return kErrArg; }
    }
    const int NT = tb_ok ? nt_block : 1;
    const int nL = NT > 1 ? steps / NT : 0;
    const int rem = steps - nL * NT;
    int rc;
    std::vector<FdJob<T>> jt = jobs, js = jobs;
    MarchPlan ptb, pss;
    DevBuf djt, dbt, djs, dbs;
    if (nL > 0) {
        if ((rc = build_march_plan(jt, tb_rows, ptb, kMarchLanes - 2 * NT))) return rc;
        if ((rc = djt.upload(jt.data(), jt.size() * sizeof(FdJob<T>), st)) ||
            (rc = dbt.upload(ptb.bmap.data(), ptb.bmap.size() * sizeof(BlockMap), st)))
            return rc;
    }
    if (rem > 0) {
        if ((rc = build_march_plan(js, ss_rows, pss, kMarchOut))) return rc;
        if ((rc = djs.upload(js.data(), js.size() * sizeof(FdJob<T>), st)) ||
            (rc = dbs.upload(pss.bmap.data(), pss.bmap.size() * sizeof(BlockMap), st)))
            return rc;
    }
    // launches of the dominant kind are event-timed on this stream with a
    // stride coprime to the snapshot interval
    const int ndom = nL > 0 ? nL : rem;
    const int stride = nL > 0 ? 7 : 37;
    const int nsamp = (ndom + stride - 1) / stride;
    std::vector<hipEvent_t> ev(2 * nsamp);
    for (auto& e : ev) SWD_HIP(hipEventCreate(&e));
    rc = kOk;
    int launch = 0;
    for (int L = 0; L < nL && rc == kOk; ++L, ++launch) {
        const bool samp = L % stride == 0;
        if (samp) SWD_HIP(hipEventRecord(ev[2 * (L / stride)], st));
        rc = adj ? launch_adj_tb<T>(st, djt.as<FdJob<T>>(), dbt.as<BlockMap>(), ptb.nblocks, L * NT, launch, NT)
                 : launch_fwd_tb<T>(st, djt.as<FdJob<T>>(), dbt.as<BlockMap>(), ptb.nblocks, L * NT, launch, NT);
        if (samp && rc == kOk) SWD_HIP(hipEventRecord(ev[2 * (L / stride) + 1], st));
    }
    for (int j = 0; j < rem && rc == kOk; ++j, ++launch) {
        const bool samp = nL == 0 && j % stride == 0;
        if (samp) SWD_HIP(hipEventRecord(ev[2 * (j / stride)], st));
        const int step = nL * NT + j;
        rc = adj ? launch_adj_march<T>(st, djs.as<FdJob<T>>(), dbs.as<BlockMap>(), pss.nblocks, step, launch)
                 : launch_fwd_march<T>(st, djs.as<FdJob<T>>(), dbs.as<BlockMap>(), pss.nblocks, step, launch);
        if (samp && rc == kOk) SWD_HIP(hipEventRecord(ev[2 * (j / stride) + 1], st));
    }
    const hipError_t se = hipStreamSynchronize(st);
    float tot = 0;
    if (rc == kOk)
        for (int i = 0; i < nsamp; ++i) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]) == hipSuccess) tot += ms;
        }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (rc) return rc;
    if (se != hipSuccess) { set_error("FD march failed: %s", hipGetErrorString(se)); return kErrHip; }
    if (kernel_us) *kernel_us = nsamp > 0 ? 1000.f * tot / nsamp : 0.f;
    if (launches_tb) *launches_tb = nL > 0 ? NT : 1;
    return kOk;
}
template int run_march<float>(hipStream_t, std::vector<FdJob<float>>&, bool, int, int, int, int, float*, int*);
template int run_march<double>(hipStream_t, std::vector<FdJob<double>>&, bool, int, int, int, int, float*, int*);

}  // namespace swd
