// Host driver of the fused FD march kernels: tiling, launch sequence
// (time-blocked launches plus a single-step remainder), ping-pong parity and
// event timing of one launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <type_traits>
#include <vector>

#include "swd_host.h"

namespace swd {

static std::atomic<int> g_nt_block{4};
static std::atomic<int> g_tb_rows{0};

int fd_time_block() { return g_nt_block.load(); }
int fd_tb_rows() { return g_tb_rows.load(); }
void set_fd_time_block(int nt_block, int rows)
{
    g_nt_block.store(std::min(std::max(nt_block, 1), 6));
    g_tb_rows.store(std::max(rows, 0));
}

// Estimated work of one time-blocked tile, in rows of the level chain: the
// rows the wavefront marches, plus its imaging (adjoint) or snapshot
// (forward) rows weighted by their cost relative to a chain row and by the
// share of the wavefront's columns they cover.  Weights from per-wave
// timestamps of the production workload: with one imaging level per launch,
// adjoint CUs holding only imaging tiles took 1.65 times as long as CUs
// holding only absorbing-border tiles; snapshot stores cost the forward
// about a tenth of that.
template <typename T>
static double tb_tile_work(const FdJob<T>& J, int tile, int out_cols, int NT, bool adj)
{
    if (tile >= J.tiles_x * J.tiles_z) return 0.0;
    const int tix = tile % J.tiles_x, tiz = tile / J.tiles_x;
    const int z0 = 1 + tiz * J.tz, z1 = std::min(z0 + J.tz, J.nzbc - 1);
    const int x0 = 1 + tix * out_cols, x1 = std::min(x0 + out_cols, J.nxbc - 1);
    const double rows = z1 - z0 + 2 * NT + 1;
    if (J.nsf <= 0) return rows;
    const int ir = std::max(0, std::min(z1, J.pt + J.nz + 1) - std::max(z0, J.pt + 1));
    const int ic = std::max(0, std::min(x1, J.nbc + J.nx) - std::max(x0, J.nbc));
    const double per_launch = std::min(1.0, (double)NT / std::max(J.in_wf, 1));  // imaging levels per launch
    return rows + (adj ? 0.65 : 0.1) * per_launch * ir * ((double)ic / out_cols);
}

// Workgroups the device holds at once for a time-blocked kernel: CUs x
// resident workgroups per CU (0 if unknown).
template <typename T>
static long long tb_resident(bool adj, int NT, bool b1, int& cus)
{
    int dev = 0;
    cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        cus = 0;
        return 0;
    }
    return (long long)cus * tb_blocks_per_cu<T>(adj, NT, b1);
}

template <typename T>
int build_march_plan(std::vector<FdJob<T>>& jobs, int tz, MarchPlan& plan, int out_cols, int NT, bool adj)
{
    std::vector<BlockMap> list;
    for (size_t j = 0; j < jobs.size(); ++j) {
        FdJob<T>& J = jobs[j];
        J.tz = tz;
        J.tiles_x = div_up(J.nxbc - 2, out_cols);
        J.tiles_z = div_up(J.nzbc - 2, tz);
        const int tiles = J.tiles_x * J.tiles_z;
        for (int t = 0; t < tiles; t += kMarchWaves) list.push_back(BlockMap{(int)j, t});
    }
    const int n = (int)list.size();
    plan.nblocks = n;
    if (n == 0) { set_error("empty FD batch"); return kErrArg; }
    if (NT <= 1) {   // single-step kernels: job-major
        plan.bmap = std::move(list);
        return kOk;
    }
    plan.bmap.resize(n);
    int cus = 0;
    const long long resident = tb_resident<T>(adj, NT, jobs_b1(jobs), cus);
    if (!adj || cus < 8 || cus % 8 || n > resident) {
        // the XCD grouping alone: for the forward, whose tiles differ little
        // in work (balancing its CUs as below made it 1% slower), and when
        // there are more workgroups than the device holds (the dispatcher
        // refills CUs as workgroups end)
        for (int i = 0; i < n; ++i) plan.bmap[i] = list[xcd_block(i, n)];
        return kOk;
    }
    // Every workgroup resident at once: a CU then finishes with its
    // heaviest workgroups and nothing rebalances them.  The dispatcher gives
    // workgroup i to XCD i mod 8 and fills the CUs in rounds of one workgroup
    // each, so workgroups i, i + cus, i + 2 cus, ... share a CU.  Workgroups
    // sorted by estimated work (heaviest first, ties in list order) fill the
    // rounds in snake order, so the CUs' sums even out; within a round,
    // consecutive ones go to one XCD (one L2), as neighbouring tiles of a shot.
    // Job-major with the XCD grouping only, the production adjoint's CUs
    // finished between 61 and 103 us, in step with their share of imaging
    // tiles.
    std::vector<double> work(n, 0.0);
    for (int k = 0; k < n; ++k)
        for (int w = 0; w < kMarchWaves; ++w)
            work[k] += tb_tile_work(jobs[list[k].job], list[k].tile0 + w, out_cols, NT, adj);
    std::vector<int> order(n);
    for (int k = 0; k < n; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
    const int per_xcd = cus / 8;
    for (int i = 0; i < n; ++i) {
        const int r = i / cus, c = i % cus, m = std::min(cus, n - r * cus);
        // position in the round: XCD-major when the round is whole
        const int p = m == cus ? (c % 8) * per_xcd + c / 8 : c;
        plan.bmap[i] = list[order[r * cus + ((r & 1) ? m - 1 - p : p)]];
    }
    return kOk;
}
template int build_march_plan<float>(std::vector<FdJob<float>>&, int, MarchPlan&, int, int, bool);
template int build_march_plan<double>(std::vector<FdJob<double>>&, int, MarchPlan&, int, int, bool);

template <typename T>
int auto_band_rows(const std::vector<FdJob<T>>& jobs, bool adj, int NT)
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        cus = 0;
    }
    const int nb = tb_blocks_per_cu<T>(adj, NT, jobs_b1(jobs));
    const long long slots = (long long)std::max(cus, 1) * std::max(nb, 1) * kMarchWaves;
    const int out_cols = kMarchLanes - 2 * NT;
    int zmax = 1;
    for (const auto& J : jobs) zmax = std::max(zmax, J.nzbc - 2);
    int best = zmax;
    double best_cost = -1;
    for (int b = 1; b <= 64; ++b) {
        const int H = div_up(zmax, b);
        if (H < 8 && b > 1) break;
        long long waves = 0;
        for (const auto& J : jobs) waves += (long long)div_up(J.nxbc - 2, out_cols) * div_up(J.nzbc - 2, H);
        const double cost = (double)div_up((int)std::min<long long>(waves, 1LL << 30), (int)slots) * (H + 2 * NT + 1);
        if (best_cost < 0 || cost < best_cost) { best_cost = cost; best = H; }
    }
    return best;
}
template int auto_band_rows<float>(const std::vector<FdJob<float>>&, bool, int);
template int auto_band_rows<double>(const std::vector<FdJob<double>>&, bool, int);

// Time-blocked launches 0 .. nL-1 of a batch, in segments of constant job
// set: a job takes part in the launches up to its last step (ceil(steps_j /
// NT), steps_j = nt forward, nt - 1 adjoint) and leaves the block map after
// them.  The first segment (every job) uses tb_rows (from occupancy when <=
// 0); a later segment, with fewer jobs, picks its band height from
// occupancy, so the tail of an uneven batch still fills the GPU.  Tiling
// does not change results: every band height gives the same bits.
template <typename T>
struct TbSegment {
    int L0 = 0, L1 = 0, rows = 0;
    std::vector<FdJob<T>> jobs;
    MarchPlan plan;
};

template <typename T>
static int tb_segments(const std::vector<FdJob<T>>& jobs, bool adj, int NT, int nL, int tb_rows,
                       std::vector<TbSegment<T>>& segs)
{
    segs.clear();
    std::vector<int> last(jobs.size());
    for (size_t j = 0; j < jobs.size(); ++j)
        last[j] = std::min(nL, div_up(std::max(adj ? jobs[j].nt - 1 : jobs[j].nt, 0), NT));
    std::vector<int> cuts = last;
    std::sort(cuts.begin(), cuts.end());
    cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
    int L0 = 0, rc;
    for (int c : cuts) {
        if (c <= L0) continue;
        TbSegment<T> sg;
        sg.L0 = L0;
        sg.L1 = c;
        for (size_t j = 0; j < jobs.size(); ++j)
            if (last[j] > L0) sg.jobs.push_back(jobs[j]);
        sg.rows = (L0 == 0 && tb_rows > 0) ? tb_rows : auto_band_rows(sg.jobs, adj, NT);
        if ((rc = build_march_plan(sg.jobs, sg.rows, sg.plan, kMarchLanes - 2 * NT, NT, adj))) return rc;
        segs.push_back(std::move(sg));
        L0 = c;
    }
    return kOk;
}

template <typename T>
int run_march(hipStream_t st, std::vector<FdJob<T>>& jobs, bool adj, int steps, int nt_block, int tb_rows,
              int ss_rows, MarchStats* stats)
{
    bool tb_ok = nt_block >= 2;
    for (const auto& J : jobs) {
        if (adj && std::is_same<T, float>::value && !J.imgp) { set_error("float adjoint jobs need partial images"); return kErrArg; }
        if (J.stype != 0 && J.stype != 5) tb_ok = false;
        if (adj && J.nsf != 3) tb_ok = false;   // the time-blocked adjoint reads the 3-field snapshots
        // float jobs need the interleaved coefficients, addressed with 32-bit offsets
        if (std::is_same<T, float>::value && (!J.c4 || (size_t)J.nzbc * J.pitch * 16 > 0xffffffffull)) tb_ok = false;
        if (J.start != 0 || J.launch0 != 0) { set_error("run_march jobs must start together"); return kErrArg; }
    }
    const int NT = tb_ok ? nt_block : 1;
    const bool b1 = jobs_b1(jobs);
    const int nL = NT > 1 ? steps / NT : 0;
    const int rem = steps - nL * NT;
    int rc;
    std::vector<TbSegment<T>> segs;
    std::vector<DevBuf> djt, dbt;
    if (nL > 0) {
        if ((rc = tb_segments(jobs, adj, NT, nL, tb_rows, segs))) return rc;
        djt.resize(segs.size());
        dbt.resize(segs.size());
        for (size_t k = 0; k < segs.size(); ++k)
            if ((rc = djt[k].upload(segs[k].jobs.data(), segs[k].jobs.size() * sizeof(FdJob<T>), st)) ||
                (rc = dbt[k].upload(segs[k].plan.bmap.data(), segs[k].plan.bmap.size() * sizeof(BlockMap), st)))
                return rc;
    }
    std::vector<FdJob<T>> js = jobs;
    MarchPlan pss;
    DevBuf djs, dbs;
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
    EventSet ev(2 * nsamp + 2);
    if ((rc = ev.create())) return rc;
    // from the first launch on, errors only stop the loops: the stream is
    // drained before the tables above are released
    int launch = 0;
    rc = ev.record(2 * nsamp, st);
    for (size_t k = 0; k < segs.size() && rc == kOk; ++k) {
        const TbSegment<T>& sg = segs[k];
        for (int L = sg.L0; L < sg.L1 && rc == kOk; ++L, ++launch) {
            const bool samp = L % stride == 0;
            if (samp && (rc = ev.record(2 * (L / stride), st))) break;
            rc = adj ? launch_adj_tb<T>(st, djt[k].as<FdJob<T>>(), dbt[k].as<BlockMap>(), sg.plan.nblocks, L * NT,
                                        launch, NT, b1)
                     : launch_fwd_tb<T>(st, djt[k].as<FdJob<T>>(), dbt[k].as<BlockMap>(), sg.plan.nblocks, L * NT,
                                        launch, NT, b1);
            if (samp && rc == kOk) rc = ev.record(2 * (L / stride) + 1, st);
        }
    }
    launch = nL;
    for (int j = 0; j < rem && rc == kOk; ++j, ++launch) {
        const bool samp = nL == 0 && j % stride == 0;
        if (samp && (rc = ev.record(2 * (j / stride), st))) break;
        const int step = nL * NT + j;
        rc = adj ? launch_adj_march<T>(st, djs.as<FdJob<T>>(), dbs.as<BlockMap>(), pss.nblocks, step, launch)
                 : launch_fwd_march<T>(st, djs.as<FdJob<T>>(), dbs.as<BlockMap>(), pss.nblocks, step, launch);
        if (samp && rc == kOk) rc = ev.record(2 * (j / stride) + 1, st);
    }
    if (rc == kOk) rc = ev.record(2 * nsamp + 1, st);
    const hipError_t se = hipStreamSynchronize(st);
    float tot = 0, span = 0;
    if (rc == kOk) {
        for (int i = 0; i < nsamp; ++i) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, ev.ev[2 * i], ev.ev[2 * i + 1]) == hipSuccess) tot += ms;
        }
        if (hipEventElapsedTime(&span, ev.ev[2 * nsamp], ev.ev[2 * nsamp + 1]) != hipSuccess) span = 0;
    }
    if (rc) return rc;
    if (se != hipSuccess) { set_error("FD march failed: %s", hipGetErrorString(se)); return kErrHip; }
    if (stats) {
        stats->kernel_us = nsamp > 0 ? 1000.f * tot / nsamp : 0.f;
        stats->span_ms = span;
        stats->nt_block = nL > 0 ? NT : 1;
        stats->rows = nL > 0 ? segs[0].rows : ss_rows;
    }
    return kOk;
}
template int run_march<float>(hipStream_t, std::vector<FdJob<float>>&, bool, int, int, int, int, MarchStats*);

template <typename T>
static void key_add(std::vector<unsigned char>& k, const T& v)
{
    const unsigned char* p = (const unsigned char*)&v;
    k.insert(k.end(), p, p + sizeof(T));
}

template <typename T>
int run_march_graph(hipStream_t st, std::vector<FdJob<T>>& jobs, bool adj, int steps, int nt_block, int tb_rows,
                    int ss_rows, MarchGraph& g, MarchStats* stats)
{
    bool tb_ok = nt_block >= 2;
    for (const auto& J : jobs) {
        if (adj && std::is_same<T, float>::value && !J.imgp) { set_error("float adjoint jobs need partial images"); return kErrArg; }
        if (J.stype != 0 && J.stype != 5) tb_ok = false;
        if (adj && J.nsf != 3) tb_ok = false;   // the time-blocked adjoint reads the 3-field snapshots
        // float jobs need the interleaved coefficients, addressed with 32-bit offsets
        if (std::is_same<T, float>::value && (!J.c4 || (size_t)J.nzbc * J.pitch * 16 > 0xffffffffull)) tb_ok = false;
        if (J.start != 0 || J.launch0 != 0) { set_error("run_march jobs must start together"); return kErrArg; }
    }
    const int NT = tb_ok ? nt_block : 1;
    const bool b1 = jobs_b1(jobs);
    const int nL = NT > 1 ? steps / NT : 0;
    const int rem = steps - nL * NT;
    int rc;
    // the key: launch parameters and the jobs as given (the segments and
    // plans are a function of them)
    std::vector<unsigned char> key;
    key_add(key, adj); key_add(key, steps); key_add(key, NT); key_add(key, tb_rows); key_add(key, ss_rows);
    key_add(key, b1);
    const unsigned char* pj = (const unsigned char*)jobs.data();
    key.insert(key.end(), pj, pj + jobs.size() * sizeof(FdJob<T>));
    for (auto& e : g.ev)
        if (!e) SWD_HIP(hipEventCreate(&e));
    if (!g.ready || key != g.key) {
        g.reset();
        std::vector<TbSegment<T>> segs;
        if (nL > 0 && (rc = tb_segments(jobs, adj, NT, nL, tb_rows, segs))) return rc;
        std::vector<FdJob<T>> js = jobs;
        MarchPlan pss;
        if (rem > 0 && (rc = build_march_plan(js, ss_rows, pss, kMarchOut))) return rc;
        // the tables go up before the capture: the graphs' kernels read them at every replay
        g.djt.resize(segs.size());
        g.dbt.resize(segs.size());
        for (size_t k = 0; k < segs.size(); ++k)
            if ((rc = g.djt[k].upload(segs[k].jobs.data(), segs[k].jobs.size() * sizeof(FdJob<T>), st)) ||
                (rc = g.dbt[k].upload(segs[k].plan.bmap.data(), segs[k].plan.bmap.size() * sizeof(BlockMap), st)))
                return rc;
        if (rem > 0 && ((rc = g.djs.upload(js.data(), js.size() * sizeof(FdJob<T>), st)) ||
                        (rc = g.dbs.upload(pss.bmap.data(), pss.bmap.size() * sizeof(BlockMap), st))))
            return rc;
        SWD_HIP(hipStreamSynchronize(st));
        // part 0: the time-blocked launches, part 1: the single-step remainder
        for (int part = 0; part < 2; ++part) {
            const int n = part == 0 ? nL : rem;
            if (n == 0) continue;
            SWD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            rc = kOk;
            if (part == 0) {
                for (size_t k = 0; k < segs.size() && rc == kOk; ++k)
                    for (int L = segs[k].L0; L < segs[k].L1 && rc == kOk; ++L)
                        rc = adj ? launch_adj_tb<T>(st, g.djt[k].as<FdJob<T>>(), g.dbt[k].as<BlockMap>(),
                                                    segs[k].plan.nblocks, L * NT, L, NT, b1)
                                 : launch_fwd_tb<T>(st, g.djt[k].as<FdJob<T>>(), g.dbt[k].as<BlockMap>(),
                                                    segs[k].plan.nblocks, L * NT, L, NT, b1);
            } else {
                for (int i = 0; i < n && rc == kOk; ++i)
                    rc = adj ? launch_adj_march<T>(st, g.djs.as<FdJob<T>>(), g.dbs.as<BlockMap>(), pss.nblocks,
                                                   nL * NT + i, nL + i)
                             : launch_fwd_march<T>(st, g.djs.as<FdJob<T>>(), g.dbs.as<BlockMap>(), pss.nblocks,
                                                   nL * NT + i, nL + i);
            }
            hipGraph_t graph = nullptr;
            const hipError_t ce = hipStreamEndCapture(st, &graph);
            if (rc != kOk || ce != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                if (rc == kOk) { set_error("graph capture failed: %s", hipGetErrorString(ce)); rc = kErrHip; }
                g.reset();
                return rc;
            }
            g.graph[part] = graph;
            SWD_HIP(hipGraphInstantiate(&g.exec[part], graph, nullptr, nullptr, 0));
            g.launches[part] = n;
        }
        g.key.swap(key);
        g.ready = true;
        g.nt_block = nL > 0 ? NT : 1;
        g.rows = nL > 0 ? segs[0].rows : ss_rows;
    }
    // a failed enqueue still drains what went before it
    hipError_t qe = hipEventRecord(g.ev[0], st);
    for (int part = 0; part < 2 && qe == hipSuccess; ++part) {
        if (g.exec[part]) qe = hipGraphLaunch(g.exec[part], st);
        if (qe == hipSuccess) qe = hipEventRecord(g.ev[part + 1], st);
    }
    const hipError_t se = hipStreamSynchronize(st);
    if (qe != hipSuccess || se != hipSuccess) {
        set_error("FD graph replay failed: %s", hipGetErrorString(qe != hipSuccess ? qe : se));
        return kErrHip;
    }
    float span[2] = {0, 0};
    for (int part = 0; part < 2; ++part)
        if (hipEventElapsedTime(&span[part], g.ev[part], g.ev[part + 1]) != hipSuccess) span[part] = 0;
    if (stats) {
        // launches of the dominant kind: their graph's span over its launch count
        const int dom = g.launches[0] > 0 ? 0 : 1;
        stats->span_ms = span[0] + span[1];
        stats->kernel_us = g.launches[dom] > 0 ? 1000.f * span[dom] / g.launches[dom] : 0.f;
        stats->nt_block = g.nt_block;
        stats->rows = g.rows;
    }
    return kOk;
}
template int run_march_graph<float>(hipStream_t, std::vector<FdJob<float>>&, bool, int, int, int, int, MarchGraph&,
                                    MarchStats*);
template int run_march<double>(hipStream_t, std::vector<FdJob<double>>&, bool, int, int, int, int, MarchStats*);

}  // namespace swd
