// The batched, device-resident SWD engine: one model, all shots of an
// inversion iteration resident in HBM (state, gathers, strain snapshots,
// spectra, adjoint sources, images).  Also the unit of work of the
// multi-model batch runner and of the serving loop.
#pragma once

#include <vector>

#include "swd_abi.h"
#include "swd_disp_plan.h"
#include "swd_host.h"

namespace swd {

// Radon geometry (frequency bins, velocities) of a problem.
int problem_geometry(const swd_problem& p, DispGeom& g);

// Batched dispersion plan of a problem's (shot, side) columns: the RTr/RTl
// windows of each shot index and the branch guards of SWD.m:175,181.
int build_disp_plan(const swd_problem& p, const std::vector<int>& is_index, hipStream_t st, DispPlan& disp);

int launch_assemble_gradient(hipStream_t st, const double* imgs, int nshots, int nz, int nx, const double* vs,
                             double* dk, double* gcl, double* gcm, double* gil);

struct StageTimes {
    float fwd_ms = 0, disp_ms = 0, adj_ms = 0, asm_ms = 0;
    float fwd_kernel_us = 0, adj_kernel_us = 0;  // mean duration of one march launch (sampled)
    int fwd_spl = 1, adj_spl = 1;                // time steps advanced by one sampled launch
    int fwd_rows = 0, adj_rows = 0;              // band rows used
};

class Engine {
public:
    swd_problem p{};
    std::vector<int> sx, sz, is_index;  // copies of the problem arrays
    int pt_m = 0, nzbc = 0, nxbc = 0, pitch = 0, nrec = 0, npair = 0, np = 0;
    size_t plane = 0, splane = 0;
    hipStream_t st = nullptr;
    bool own_stream = false;
    int fd_path = 0;  // 0 fused batched, 1 plain one-shot-per-launch
    int tz = 16;        // band rows of the single-step march
    int nt_block[2] = {4, 4};  // time steps per launch, forward / adjoint (1: single-step kernels only)
    int tb_rows[2] = {0, 0};   // band rows of the time-blocked march, forward / adjoint (0: from occupancy)
    bool have_model = false, have_wavelet = false;
    double vpmax = 0;
    double vp_ratio = 0;   // > 0: device model updates set vp = vp_ratio * vs (field scripts); 0: vp fixed (SWD.m)
    bool b1_same = false;  // the b1 plane equals b (jobs then read b for b1)
    bool use_graph = false;    // FD phases replay captured HIP graphs
    MarchGraph graphs[2][2][2];   // [adjoint][snapshots][multi-engine batch]
    hipStream_t batch_fd[2] = {nullptr, nullptr};  // forward / adjoint launch streams of batches this engine leads

    DevBuf coefs;     // 6 planes float: temp, b, b1, ca, cl, cm1
    DevBuf vs_d;      // nz x nx double (MATLAB layout) for the assembly
    DevBuf wavelet;   // nt float
    DevBuf state;     // nshots x 10 planes float
    DevBuf strain;    // 4 planes (plain path scratch)
    DevBuf traces;    // [shot][t][g] float
    DevBuf snaps;     // [shot][nrec][nz][nx][3] float (fux, bwz, fuz + bwx of each cell)
    DevBuf adjsrc;    // [shot][t][g] float
    DevBuf imgs;      // [shot][3][nz][nx] double
    DevBuf imgp;      // [shot][nz][nx][3] float, partial image sums of the fused kernels
    DevBuf outs;      // dk_vs, g_cl, g_cm, g_il (nz x nx double each)
    DevBuf djobs, dbmap;
    DevBuf vp_d;      // nz x nx double, the current vp (column-major)
    DevBuf geo;       // nzbc x nxbc double, damping geometry factors (column-major)
    DevBuf vs_cur;    // nz x nx double, model at the start of an iteration
    DevBuf vs_tr;     // nz x nx double, trial model of the line search
    DevBuf dk_sm;     // nz x nx double, smoothed gradient
    DevBuf stats;     // 4 doubles: min, max, sum of squares (model_stats)
    DispPlan disp;
    StageTimes times;
    // snapshot placements the autotuner timed (full forward + adjoint march,
    // ms each) and the index of the one kept (0: the first allocation)
    float snap_placement_ms[3] = {0, 0, 0};
    int snap_placements = 0, snap_choice = 0;
    int timing_samples = 8;

    ~Engine();
    int create(const swd_problem* prob, hipStream_t stream);
    int set_model(const double* vp, const double* vs);
    int set_wavelet(const double* s);
    int set_observed_picks(const double* cr0_r, const double* cr0_l);
    // rebuild the coefficient planes for a device-resident vs (vp unchanged,
    // or vp = vp_ratio * vs when vp_ratio > 0)
    int set_vs_device(const double* vs_dev);
    // one SWD iteration (SWD.m:171-335): gradient, smoothing, line search,
    // model update, all on the device
    int iterate(swd_iteration* it);
    double mean_right(const double* res_r) const;
    size_t footprint_bytes() const;

    // building blocks (all enqueue on st; nothing synchronises unless stated)
    void make_jobs(std::vector<FdJob<float>>& jobs, bool adjoint, bool snapshots);
    int forward(bool snapshots);
    int dispersion();
    int adjoint_source();
    int adjoint();
    int assemble();

    // complete operations (synchronous)
    int gradient(double* dk_vs, double* res_r, double* res_l, double* cr_r, double* cr_l);
    int forward_picks(double* res_r, double* res_l, double* cr_r, double* cr_l);
    int observe(double* cr0_r, double* cr0_l);
    int read_picks(double* res_r, double* res_l, double* cr_r, double* cr_l);

private:
    int run_fd(bool adjoint, bool snapshots);
};

// Gradient of several engines (models of different grid and nt) with their
// FD phases overlapped; dk_vs[i] (nz x nx, may be null) receives engine i's
// gradient, kernel_us (may be null) each phase's event-timed span divided by
// its launch count.  Needs time-blocked launch shapes (nt_block >= 2) in
// both phases.
int batch_gradient_overlapped(const std::vector<Engine*>& engines, double** dk_vs, double* kernel_us);

// Run the FD phase of several engines in shared launches (one launch covers
// every shot of every engine; engines may differ in grid size and nt) with
// the first engine's time-block settings.  adjoint selects the reverse-time
// phase.
int run_fd_batch(const std::vector<Engine*>& engines, bool adjoint, bool snapshots, hipStream_t st, int tz,
                 MarchStats* stats);

// One measured candidate of the FD autotuner.
struct TuneRow {
    int nt_block = 1, rows = 0;
    float fwd_us = 0, adj_us = 0;  // event-timed microseconds per time step over the batch
};

// Measure the FD march of a batch of engines for every time block 1..6 and a
// set of band heights, on the engines' own buffers over `steps` steps of each
// phase, and install the fastest forward and adjoint settings on every
// engine.  Overwrites the FD state, gathers, snapshots, adjoint sources and
// images (the next gradient recomputes all of them).
int autotune_fd_batch(const std::vector<Engine*>& engines, int steps, std::vector<TuneRow>& table);

}  // namespace swd
