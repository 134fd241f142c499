// Autotuner of the fused FD march: picks the time block (steps per launch)
// and band height of the time-blocked kernels by timing them on the batch
// they will run, separately for the forward and the adjoint phase (for a
// batch of several models, the time block only: band heights then come from
// occupancy, launch table by launch table).
//
// The best setting depends on the batch: a launch is fastest when its waves
// fill the resident-wave slots of the GPU in whole rounds while each wave
// re-loads few halo rows (band H loads H + 2*NT + 1 rows for H outputs), and
// the number of waves is shots x column tiles x bands.  Results do not depend
// on the setting (every setting is bitwise equal), so the tuner only times.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "swd_engine.h"

namespace swd {

// Band heights that split the tallest update block into b = 1, 2, ... equal
// bands (smallest height 8 rows).
static std::vector<int> band_candidates(int zmax)
{
    std::vector<int> hs;
    for (int b = 1; b <= 32; ++b) {
        const int H = div_up(zmax, b);
        if (H < 8 && b > 1) break;
        if (hs.empty() || hs.back() != H) hs.push_back(H);
    }
    return hs;
}

// Snapshot-buffer placement of one engine.  The adjoint reads the forward's
// snapshots once per imaging step, and where the buffer's pages land in
// physical memory moves its launch time by up to about 10% (85-99 us per
// 5-step launch at 60 shots on 132 x 660, steady for one allocation).  Time
// one full forward + adjoint march on the engine's buffer and on up to
// kSnapPlacements - 1 more, each allocated while the earlier ones are held (so
// it takes other pages), and keep the fastest.  A candidate is only tried
// while kSnapHeadroom of device memory stays free beside it.
constexpr int kSnapPlacements = 3;
constexpr size_t kSnapHeadroom = (size_t)8 << 30;

static int choose_snapshot_placement(Engine* e, int tz)
{
    hipStream_t st = e->st;
    auto full_march = [&](float* ms) -> int {
        MarchStats f, a;
        int rc;
        if ((rc = run_fd_batch({e}, false, true, st, tz, &f)) || (rc = run_fd_batch({e}, true, true, st, tz, &a)))
            return rc;
        *ms = f.span_ms + a.span_ms;
        return kOk;
    };
    float best = 0;
    int rc = full_march(&best);
    if (rc) return rc;
    e->snap_placement_ms[0] = best;
    e->snap_placements = 1;
    e->snap_choice = 0;
    // every candidate stays allocated until the choice is made; the ones not
    // chosen are freed when `held` goes out of scope
    std::vector<DevBuf> held;
    held.reserve(kSnapPlacements);
    for (int k = 1; k < kSnapPlacements; ++k) {
        size_t fr = 0, tot = 0;
        SWD_HIP(hipMemGetInfo(&fr, &tot));
        if (fr < e->snaps.bytes + kSnapHeadroom) break;
        DevBuf nb;
        if (nb.alloc(e->snaps.bytes)) break;
        std::swap(e->snaps, nb);  // the engine marches on the new buffer
        float t = 0;
        if ((rc = full_march(&t))) {
            std::swap(e->snaps, nb);
            held.push_back(std::move(nb));
            return rc;
        }
        e->snap_placement_ms[k] = t;
        e->snap_placements = k + 1;
        if (t < best) {
            best = t;
            e->snap_choice = k;
        } else {
            std::swap(e->snaps, nb);  // back to the faster buffer
        }
        held.push_back(std::move(nb));
    }
    return kOk;
}

int autotune_fd_batch(const std::vector<Engine*>& es, int steps, std::vector<TuneRow>& table)
{
    table.clear();
    if (es.empty()) { set_error("empty batch"); return kErrArg; }
    hipStream_t st = es[0]->st;
    for (Engine* e : es) SWD_HIP(hipStreamSynchronize(e->st));
    int min_nt = INT_MAX, zmax = 1;
    for (Engine* e : es) {
        min_nt = std::min(min_nt, e->p.nt);
        zmax = std::max(zmax, e->nzbc - 2);
    }
    // 60 steps: a whole number of launches for every time block 1..6
    if (steps <= 0) steps = 60;
    steps = std::min(steps, min_nt - 1);
    if (steps < 1) { set_error("nt too small to tune"); return kErrArg; }

    std::vector<FdJob<float>> jf, ja;
    for (Engine* e : es) {
        e->make_jobs(jf, false, true);
        e->make_jobs(ja, true, true);
        SWD_HIP(hipMemsetAsync(e->state.p, 0, e->state.bytes, st));
        SWD_HIP(hipMemsetAsync(e->adjsrc.p, 0, e->adjsrc.bytes, st));
        SWD_HIP(hipMemsetAsync(e->imgs.p, 0, e->imgs.bytes, st));
    }
    const int tz = es[0]->tz;

    std::vector<TuneRow> cand;
    {
        TuneRow r;
        r.nt_block = 1;
        r.rows = tz;
        cand.push_back(r);
    }
    // A batch of several models changes its job set as models finish, and
    // every launch table after the first takes its band height from
    // occupancy (auto_band_rows); the tuner then picks only the time block,
    // each candidate at those band heights (rows 0).  Tuning the first
    // table's height on the first 60 steps picked 18-row forward bands for
    // the 8-model uneven batch, where the occupancy choice (36) ran the whole
    // batch gradient in 324 instead of 340 ms (means of two processes each).
    const bool batch = es.size() > 1;
    for (int NT = 2; NT <= 6; ++NT)
        for (int H : batch ? std::vector<int>{0} : band_candidates(zmax)) {
            TuneRow r;
            r.nt_block = NT;
            r.rows = H;
            cand.push_back(r);
        }

    // best of two event-timed runs, microseconds per time step
    auto measure = [&](std::vector<FdJob<float>>& jobs, bool adj, const TuneRow& c, float* us) -> int {
        float best = 0;
        for (int rep = 0; rep < 2; ++rep) {
            MarchStats ms;
            const int rc = run_march<float>(st, jobs, adj, steps, c.nt_block, c.rows, tz, &ms);
            if (rc) return rc;
            const float v = 1000.f * ms.span_ms / (float)steps;
            if (rep == 0 || v < best) best = v;
        }
        *us = best;
        return kOk;
    };
    int rc;
    {
        float dummy;
        if ((rc = measure(jf, false, cand[0], &dummy)) || (rc = measure(ja, true, cand[0], &dummy))) return rc;
    }
    for (TuneRow& c : cand) {
        if ((rc = measure(jf, false, c, &c.fwd_us)) || (rc = measure(ja, true, c, &c.adj_us))) return rc;
        table.push_back(c);
    }
    const TuneRow* bf = &table[0];
    const TuneRow* ba = &table[0];
    for (const TuneRow& r : table) {
        if (r.fwd_us > 0 && r.fwd_us < bf->fwd_us) bf = &r;
        if (r.adj_us > 0 && r.adj_us < ba->adj_us) ba = &r;
    }
    for (Engine* e : es) {
        e->nt_block[0] = bf->nt_block;
        e->tb_rows[0] = bf->nt_block > 1 ? bf->rows : 0;
        e->nt_block[1] = ba->nt_block;
        e->tb_rows[1] = ba->nt_block > 1 ? ba->rows : 0;
    }
    // a single engine with its own snapshot allocation (not one carved from a
    // server's arena) also picks where its snapshots live
    if (es.size() == 1 && !es[0]->snaps.arena && !current_arena())
        if ((rc = choose_snapshot_placement(es[0], tz))) return rc;
    return kOk;
}

}  // namespace swd
