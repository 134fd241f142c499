/*
 * swd_abi.h - C ABI of libswd, the MI355X (gfx950) engine for the hot path of
 * Skeletonized Wave-Equation Dispersion Spectrum Inversion (zhangchang271/SWD).
 *
 * Plain C types only (pointers and sizes, no torch/HIP types).  Unless stated
 * otherwise, array arguments are host memory in MATLAB layout: column-major,
 * first index fastest (depth z for model grids, time t for gathers).
 * Every entry point returns SWD_OK (0) or a negative SWD_ERR_* code; the
 * message of the last failure on the calling thread is swd_last_error().
 * Entry points are thread-safe; each call uses its own HIP stream.
 */
#ifndef SWD_ABI_H
#define SWD_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

#define SWD_OK 0
#define SWD_ERR_ARG (-2)      /* bad descriptor, or input the reference MEX would turn into UB */
#define SWD_ERR_CFL (-3)      /* time step violates the staggered-grid stability bound */
#define SWD_ERR_NOMEM (-4)    /* device / host allocation failed, or admission refused */
#define SWD_ERR_HIP (-5)      /* HIP runtime error */
#define SWD_ERR_PICK (-6)     /* picking produced fewer than two kept columns (interp1 would error) */
#define SWD_ERR_NODEV (-7)    /* no usable gfx950 device */
#define SWD_ERR_BUSY (-8)     /* handle in use / cancelled */

/* Library identity and diagnostics. */
const char *swd_version(void);
const char *swd_last_error(void);
int swd_device_info(int *n_devices, char *arch, int arch_len);

/* ------------------------------------------------------------------------
 * FD kernels with the MEX argument lists.
 * ------------------------------------------------------------------------ */

/*
 * Forward modelling.  Replaces the MEX eigen_staggerfd_single
 * (core/eigen_staggerfd_single.cpp:68, MexFunction::operator()) and, with
 * double arrays, the binary-only eigen_staggerfd_double, as called by
 * core/staggerfd_eigen.m:139 / core/staggerfd_eigend.m:139:
 *   [seismo_u, seismo_w, S] = eigen_staggerfd_*(iv, temp, ca, cl, cm, cm1, b, b1, s)
 * iv[16] = [nt nzbc nxbc dtx ng sz sx gz gx dg source_type fd_order in_wf nz nx format]
 *          (sz, sx, gz, gx 1-based on the padded grid, cpp:98-118).
 * temp..b1: nzbc x nxbc; s: nt.  cm is accepted and unused, as in the MEX.
 * Outputs (caller-allocated): seismo_u, seismo_w nt x ng; fux, fuz, bwx, bwz
 * nz x nx x (nt/in_wf) strain snapshots (the struct S of cpp:375-379; the
 * format-2 shape nz x (nx*nt/in_wf) is the same memory).  seismo_u and the
 * four snapshot pointers may be NULL to skip them (forward-only use).
 */
int swd_staggerfd_f32(const float *iv, const float *temp, const float *ca, const float *cl, const float *cm,
                      const float *cm1, const float *b, const float *b1, const float *s, float *seismo_u,
                      float *seismo_w, float *fux, float *fuz, float *bwx, float *bwz);
int swd_staggerfd_f64(const double *iv, const double *temp, const double *ca, const double *cl, const double *cm,
                      const double *cm1, const double *b, const double *b1, const double *s, double *seismo_u,
                      double *seismo_w, double *fux, double *fuz, double *bwx, double *bwz);

/*
 * Adjoint propagation + imaging.  Replaces the MEX eigen_e2drtm_single
 * (core/eigen_e2drtm_single.cpp:67) and the binary-only eigen_e2drtm_double,
 * as called by core/e2drtm_eigen.m:122 / core/e2drtm_eigend.m:122:
 *   [cl_img, cm_img, illum_div] = eigen_e2drtm_*(iv, temp, ca, cl, cm, cm1, b, b1, adj, fux, fuz, bwx, bwz)
 * iv[14] = [nt nzbc nxbc dtx ng nbc gz gx dg fd_order in_wf nz nx dt] (cpp:69-84).
 * adj: nt x ng adjoint source; snapshots nz x nx x (nt/in_wf).
 * Outputs nz x nx.  Images accumulate in double on the device.
 */
int swd_e2drtm_f32(const float *iv, const float *temp, const float *ca, const float *cl, const float *cm,
                   const float *cm1, const float *b, const float *b1, const float *adj, const float *fux,
                   const float *fuz, const float *bwx, const float *bwz, float *cl_img, float *cm_img,
                   float *illum_div);
int swd_e2drtm_f64(const double *iv, const double *temp, const double *ca, const double *cl, const double *cm,
                   const double *cm1, const double *b, const double *b1, const double *adj, const double *fux,
                   const double *fuz, const double *bwx, const double *bwz, double *cl_img, double *cm_img,
                   double *illum_div);

/* Select the FD code path of the two entry points above:
 * 0 = automatic (fused kernels when fd_order 22 and sources inside the
 * update block, plain kernels otherwise), 1 = always plain.  Returns the
 * previous value.  Used by tests and the unfused comparator. */
int swd_set_fd_path(int mode);

/* Time blocking of the fused FD kernels: nt_block steps per launch (1..6;
 * 1 = one launch per time step) and band rows of the time-blocked kernels
 * (0 = chosen from the kernels' occupancy and the grid, the default).
 * Results do not depend on either setting (bitwise).  No reference
 * counterpart: the MEX time loops have no launch structure. */
int swd_set_fd_time_block(int nt_block, int rows);

/*
 * Model preparation of the wrappers: core/staggerfd_eigen.m:37-107 (with
 * core/calparam.m:3-5, core/pad.m:3-9, core/damp_circle.m:5-67) for
 * parameter_type 0 and a free surface (isfs 1; the wrapper leaves b, b1 and
 * cm1 undefined otherwise).  vp, vs: nz x nx.  Outputs nzbc x nxbc with
 * nzbc = nz + pad_top + nbc, nxbc = nx + 2 nbc, pad_top = (fd_order-20)/2+1.
 * e2drtm_eigen.m:44-101 prepares the same arrays.
 */
int swd_prepare_model(const double *vp, const double *vs, int nz, int nx, int nbc, double dx, double dt,
                      int fd_order, double *temp, double *ca, double *cl, double *cm, double *cm1, double *b,
                      double *b1);

/* Ricker wavelet, core/ricker.m:18-35 with nargin == 3 (w has nt samples). */
int swd_ricker(double f, double dt, int nt, double *w);

/* ------------------------------------------------------------------------
 * Dispersion path (FP32 on the GPU: DFT-at-bins and slant stack on MFMA).
 * ------------------------------------------------------------------------ */

/*
 * Linear Radon transform of one side of one shot gather.  Replaces
 * core/RTr.m:2 (side 0) / core/RTl.m:2 (side 1) and their AD twins
 * core/RTrAD.m:2 / core/RTlAD.m:2:
 *   [ml, dataLen(, saveForBackward)] = RTx(seismo, is, df, dt, np, vmin, vmax, fmin, fmax, a, b, ds, offset, m, M)
 * seismo: nt x ng (a = nt, b = ng).  Outputs: ml np x npair (column-max
 * normalised |mm|, npair = nf-lf+1); optional mlr np x npair (|mm|, the AD
 * field mlr/mll) and mm np x npair complex interleaved (re, im) = columns
 * lf..nf of the AD field mm.  info[4] = {lf, nf, ccn, first trace (1-based)
 * of window position 0}.  swd_radon takes double gathers (SWD.m),
 * swd_radon_f32 single ones (SWD_single.m).
 */
int swd_radon(int side, const double *seismo, int nt, int ng, int is, double df, double dt, int np, double vmin,
              double vmax, double fmin, double fmax, double ds, int offset, int m, int M, float *ml, int *dataLen,
              float *mlr, float *mm, int *info);
int swd_radon_f32(int side, const float *seismo, int nt, int ng, int is, double df, double dt, int np, double vmin,
                  double vmax, double fmin, double fmax, double ds, int offset, int m, int M, float *ml,
                  int *dataLen, float *mlr, float *mm, int *info);

/*
 * Dispersion pick and misfit.  Replaces core/LHDispPick.m:2
 *   [res, cr_pre] = LHDispPick(ml, vmin, cr_0, method, ini)
 * (method 1 = pickFDC1(ml, ini, 2, 4), core/pickFDC1.m:1; method 2 = column
 * argmax).  ml: np x npair; cr_0, cr_pre: npair.  Returns SWD_ERR_PICK where
 * MATLAB's interp1 would error (fewer than two kept columns).
 */
int swd_lhdisppick(const float *ml, int np, int npair, double vmin, const double *cr_0, int method, int ini,
                   double *res, double *cr_pre);
/* The same pick on FP64 spectra (SWD.m passes double ml to LHDispPick; the
 * single entry above is SWD_single.m's). */
int swd_lhdisppick_f64(const double *ml, int np, int npair, double vmin, const double *cr_0, int method, int ini,
                       double *res, double *cr_pre);

/*
 * swd_lhdisppick without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results (bitwise) and errors as swd_lhdisppick; every
 * device buffer comes from a fixed 4 MiB array per GPU (held until the
 * process ends), so np * npair * 4 + 2 * npair * 8 bytes plus 1 KiB must fit
 * in 4 MiB, else SWD_ERR_ARG (701 x 1024 fits; 2048 x 1024 does not).  Calls
 * run on the per-thread default stream; calls on one GPU are serialised.
 */
int swd_lhdisppick_noalloc(const float *ml, int np, int npair, double vmin, const double *cr_0, int method, int ini,
                           double *res, double *cr_pre);

/*
 * SoftMax adjoint source.  Replaces core/ADWDgrad_1.m:1
 *   a_data_res = ADWDgrad_1(nt, ng, ns, npair, is, w, m, M, SoftArgNorm, gor, gol, space_M, sfbR, sfbL)
 * The saveForBackward structs are passed as their fields: mm (np x npair
 * complex interleaved), mlr/mll (np x npair), the window as its position-0
 * trace pos (1-based: uxtposr(1), resp. uxtposl(end)) and length L, and the
 * scalars lf, ccn; ll0 = 1i*2*pi*df*(1./(vmin:vmin+np-1))'*((m+(0:L-1))*ds)
 * is rebuilt from np, vmin, df, ds, m.  space_M is 1..np in every column
 * (SWD.m:133-136).  Branch guards as ADWDgrad_1.m:13/47.  a_data_res: nt x ng.
 */
int swd_adwdgrad1(int nt, int ng, int ns, int npair, int is, double w, int m, int M, double SoftArgNorm,
                  const double *grad_outputr, const double *grad_outputl, const float *mmr, const float *mlr,
                  int posr, int Lr, const float *mml, const float *mll, int posl, int Ll, int np, double vmin,
                  double df, double ds, int lf, int ccn, float *a_data_res);

/* ------------------------------------------------------------------------
 * Batched, device-resident engine: all shots of one SWD iteration of one
 * model in HBM, one FD launch per time step for the whole batch.
 * ------------------------------------------------------------------------ */

typedef struct swd_engine swd_engine;

/* One SWD problem (the constants of SWD.m:78-116 for one model).
 * Positions are 1-based on the model grid, as SWD.m passes them to the
 * wrappers: shot s at (sz[s], sx[s]); receiver g at (gz, gx0 + g*dg). */
typedef struct swd_problem {
    int nz, nx, nbc, fd_order, nt, in_wf;
    double dx, dt;
    int nshots;
    const int *sx, *sz;
    int ng, gx0, dg, gz;
    double df, fmin, fmax, vmin, vmax, ds;  /* Radon axes; ds = receiver spacing in metres */
    int offset, m, M, ini;                  /* windows (RTr.m:5-12), pick start column */
    double w, beta;                         /* branch-guard w (SWD.m:112), SoftArgNorm */
    const int *is_index;                    /* shot number used by windows and guards; NULL: s+1 */
    int ns;                                 /* ns of the guards; 0: nshots */
} swd_problem;

/* The parfor body SWD.m:171-194 for every shot plus SWD.m:199-201:
 * swd_engine_gradient runs forward FD (+ snapshots), RTrAD/RTlAD,
 * LHDispPick, ADWDgrad_1, adjoint FD + imaging, the shot reduction and
 * dk_vs = -4 vs g_cl + 2 vs g_cm (dk_vs nz x nx, before smooth2a).
 * res_r/res_l: nshots; cr_r/cr_l: npair x nshots picks (any may be NULL).
 * swd_engine_observe computes the observed picks cr_0 / cr_0l of
 * SWD.m:150-162 from the current model and keeps them for the gradient;
 * swd_engine_set_observed_picks installs given ones.  swd_engine_forward_picks
 * is the forward-only line-search pass of SWD.m:215-228. */
int swd_engine_create(const swd_problem *p, swd_engine **out);
int swd_engine_destroy(swd_engine *e);
int swd_engine_set_model(swd_engine *e, const double *vp, const double *vs);
int swd_engine_set_wavelet(swd_engine *e, const double *s);
int swd_engine_set_observed_picks(swd_engine *e, const double *cr0_r, const double *cr0_l);
int swd_engine_observe(swd_engine *e, double *cr0_r, double *cr0_l);
int swd_engine_gradient(swd_engine *e, double *dk_vs, double *res_r, double *res_l, double *cr_r, double *cr_l);
int swd_engine_forward_picks(swd_engine *e, double *res_r, double *res_l, double *cr_r, double *cr_l);
/* 0 = fused batched kernels (default); 1 = the plain per-statement kernels,
 * one shot per launch (the unfused comparator).  Returns the previous mode. */
int swd_engine_set_fd_path(swd_engine *e, int mode);
/* on != 0: each FD phase of the fused path replays a captured HIP graph of
 * its launch sequence (captured on first use, again after a change of
 * launch shape or batch); results are bitwise those of the eager launches.
 * The setting of the first engine of a swd_batch_gradient batch holds for
 * the batch.  swd_engine_timings then gives the dominant launch's duration
 * as the event-timed span of its graph over the launch count.  Returns the
 * previous setting. */
int swd_engine_set_graph(swd_engine *e, int on);
/* Variants of the field-data scripts (fieldexamples/outcrop/outcrop.m and
 * fieldexamples/QadimahFault/QademahSWD3.m with their core/ copies); all
 * zero is SWD.m.
 *   norm_trace     divide every gather trace by its peak |x| before the Radon
 *                  of each dispersion pass (norm_trace on the predicted data,
 *                  outcrop.m:164, QademahSWD3.m:164; core/norm_trace.m:8-16).
 *                  No derivative is taken through it, as in the scripts; a
 *                  zero trace gives NaN spectra (0/0 in norm_trace).
 *   auto_start     pick from the frequency column holding the largest |mm|
 *                  instead of the problem's ini ([~,ini] = max(max(mlr)),
 *                  outcrop/core/LHDispPick.m:3-9, QadimahFault/core/LHDispPick.m:4-10),
 *                  starting at that column's peak.
 *   res_first, res_last  the pick misfit averages cr(res_first:end-res_last)
 *                  (1-based; 15, 15 in outcrop/core/LHDispPick.m:21; 50, 30 in
 *                  QadimahFault/core/LHDispPick.m:25); 0, 0 = every column.
 *   adj_peak_norm  divide each shot's adjoint source by its peak |a|
 *                  (a_data_res./max(abs(a_data_res(:))), ADWDgrad_1.m:79 of
 *                  both field copies).
 *   vp_ratio       > 0: every model change of swd_engine_iterate sets
 *                  vp = vp_ratio * vs1 (vp = vs1*1.732, outcrop.m:217,245,271,
 *                  301,331); 0 keeps vp fixed (SWD.m).  The dispersion batch
 *                  ignores it.
 *   fk_filter      the QadimahFault RTr/RTl branch FK == 1 (RTr.m:14-16,
 *                  RTl.m:15-17; core/fk_filter.m:48-53,283-296): each window,
 *                  flipped for the left side, passes the down-dip f-k filter
 *                  before its spectrum (no derivative through it, as there).
 *                  Windows of at most 2048 traces.
 * The variant applies to every later dispersion pass of the handle (observed
 * picks included). */
typedef struct swd_variant {
    int norm_trace, auto_start, res_first, res_last, adj_peak_norm;
    double vp_ratio;
    int fk_filter;
} swd_variant;
int swd_engine_set_variant(swd_engine *e, const swd_variant *v);
/* One stage of swd_engine_gradient on the resident buffers: 1 forward FD +
 * snapshots, 2 Radon + picks, 3 adjoint sources, 4 adjoint FD + imaging,
 * 5 reduction + dk_vs.  swd_engine_set_picks overwrites the picks between
 * stages 2 and 3 (npair x nshots each; NULL keeps a side). */
int swd_engine_run(swd_engine *e, int stage);
int swd_engine_set_picks(swd_engine *e, const double *cr_r, const double *cr_l);
int swd_engine_set_tile_rows(swd_engine *e, int tz);
/* Time block of both FD phases (rows 0 = automatic), or of one phase
 * (phase 0 forward, 1 adjoint, -1 both). */
int swd_engine_set_time_block(swd_engine *e, int nt_block, int rows);
int swd_engine_set_phase_time_block(swd_engine *e, int phase, int nt_block, int rows);
/* FD autotuner (no reference counterpart): times every time block 1..6 with
 * band heights that split the tallest grid evenly (for a batch of several
 * engines: with band rows 0, chosen from occupancy for each launch table),
 * over `steps` (<= 0: 60) forward and adjoint steps of the batch on its own
 * buffers, and installs the fastest forward and adjoint settings on every
 * engine.  choice[4] =
 * {forward steps/launch, forward band rows, adjoint steps/launch, adjoint
 * band rows}; table[table_rows][4] = {steps/launch, rows, forward us/step,
 * adjoint us/step} per candidate; *n_rows = candidates timed.  Overwrites FD
 * state, gathers, snapshots, adjoint sources and images. */
int swd_batch_autotune(swd_engine **engines, int n, int steps, int *choice, double *table, int table_rows,
                       int *n_rows);
int swd_engine_autotune(swd_engine *e, int steps, int *choice, double *table, int table_rows, int *n_rows);
/* vals[0..9] = nzbc, nxbc, pitch, nt/in_wf, npair, np, device bytes, ccn, lf, tile rows */
int swd_engine_info(swd_engine *e, long long *vals, int n);
/* ms[0..9] = forward, dispersion, adjoint, assembly (ms) of the last gradient,
 * mean forward / adjoint march-kernel launch duration (us, event-timed
 * sample), forward / adjoint time steps per launch, forward / adjoint band
 * rows; ms[10..14] = snapshot placements swd_engine_autotune timed, the index
 * of the one kept, and the full forward + adjoint march time (ms) of each.
 * The adjoint's launch time depends on where the snapshot buffer lands in
 * physical memory, so the autotuner of a single engine times one full march
 * on its buffer and on up to two more, allocated while the earlier ones are
 * held (each only while 8 GiB of device memory stays free beside it), and
 * keeps the fastest. */
int swd_engine_timings(swd_engine *e, double *ms, int n);
/* Copy a device buffer to host: 0 gathers [shot][t][g] f32, 1 adjoint sources
 * [shot][t][g] f32, 2 {dk_vs, sum cl_img, sum cm_img, sum illum} nz x nx f64,
 * 3 spectra [col][npair][np] f32, 4 per-shot images [shot][3][nz][nx] f64,
 * 5 the smoothed dk_vs of the last swd_engine_iterate, nz x nx f64. */
int swd_engine_read(swd_engine *e, int what, void *dst, long long bytes);
/*
 * One SWD inversion iteration on the device (SWD.m:171-335 for one k):
 * gradient of the engine's model, dk_vs = smooth2a(dk_vs, smooth_z,
 * smooth_x), alpha = ||vs|| / ||dk_vs|| * pur, the line search of
 * forward-only batches (halve f1 while the right-side misfit grows, else try
 * 2 f1), the parabola step gama, and vs = clip(vs + alpha gama dk_vs,
 * vsmin, vsmax) with the reference's fallback to the better trial.  The
 * model stays on the device (swd_engine_read_model reads it); vp is the one
 * given to swd_engine_set_model.  f1, f2, res2 carry the script's state
 * between iterations (f1 = 0.5 at the first; f2 and res2 are used only when
 * the halving loop does not run).  The offset schedule of SWD.m:326-333
 * (rebuilding windows and observed picks) is the caller's.
 */
typedef struct swd_iteration {
    double pur, vsmin, vsmax;  /* SWD.m:76,99 */
    int smooth_z, smooth_x;    /* SWD.m:116 */
    double f1, f2, res2;       /* carried between iterations */
    double residual;           /* out: residual(k), mean right-side misfit of the gradient pass */
    double res0;               /* out: misfit of the accepted model */
    double alpha, gama;        /* out: step scale and accepted step factor */
    double res1, res3;         /* out: misfits of the last halving / parabola trials */
    int forwards;              /* out: forward-only batches run */
} swd_iteration;
int swd_engine_iterate(swd_engine *e, swd_iteration *it);
int swd_engine_read_model(swd_engine *e, double *vs);
/* smooth2a (core/smooth2a.m:52-76) on the device: (2 nr + 1) x (2 nc + 1) box
 * mean with truncated edges, NaNs ignored; nz x nx column-major. */
int swd_smooth2a(const double *in, int nz, int nx, int nr, int nc, double *out);

/* Gradients of n engines (models of different sizes and nt) with shared FD
 * launches; dk_vs[i] is nz_i x nx_i.  kernel_us[2] = sampled forward/adjoint
 * march-kernel duration. */
int swd_batch_gradient(swd_engine **engines, int n, double **dk_vs, double *kernel_us);

/* HBM counter calibration (no reference counterpart): copies `bytes` between
 * two device buffers `reps` times with a plain vector copy kernel; *ms = mean
 * event-timed duration of one copy.  Its rocprofv3 FETCH_SIZE / WRITE_SIZE
 * readings scale the FD kernels' readings to bytes.  Both copy functions
 * check the destination against the source after the timed copies and
 * return SWD_ERR_HIP if any 16-byte element differs. */
int swd_calib_copy(long long bytes, int reps, double *ms);

/* Practical HBM ceiling (no reference counterpart): the fastest device copy
 * this library has measured on MI355X (non-temporal 16-byte loads and
 * stores, four in flight per lane); copies `bytes` `reps` times, *ms = mean
 * event-timed duration of one copy.  Read + write bytes over *ms is the rate
 * a streaming kernel can reach, beside the 8 TB/s specification. */
int swd_copy_peak(long long bytes, int reps, double *ms);

/*
 * The dispersion path alone, batched over every (shot, side) column of
 * nshots gathers (the per-shot RTr/RTl + LHDispPick of SWD.m:173-185 and
 * RTrAD/RTlAD + ADWDgrad_1 of SWD.m:190, with the branch guards of
 * SWD.m:175,181), on gathers the caller provides.  Of the problem struct it
 * uses the Radon, pick and window fields: nt, dt, ng, nshots, is_index, ns,
 * df, fmin, fmax, vmin, vmax, ds, offset, m, M, ini, w, beta; the FD fields
 * are ignored.
 * swd_dispersion_info: vals[0..7] = npair, np, lf, nf, ccn, longest window,
 * total window length and number of enabled (shot, side) columns.
 * swd_dispersion_set_gathers: [shot][t][g] float (nshots x nt x ng, g fastest).
 * swd_dispersion_pick: picks cr_r / cr_l (npair x nshots column-major, in
 * m/s, cr_pre = idx + vmin), misfits res_r / res_l (nshots); *ms = device
 * time of Radon + picks.  SWD_ERR_PICK when an enabled column's pick keeps
 * fewer than two columns.
 * swd_dispersion_adjoint: adjoint sources [shot][t][g] of the last picks
 * against the observed picks (beta <= 0: the problem's beta); *ms = device
 * time.
 */
typedef struct swd_dispersion swd_dispersion;
int swd_dispersion_create(const swd_problem *p, swd_dispersion **out);
int swd_dispersion_destroy(swd_dispersion *d);
int swd_dispersion_info(swd_dispersion *d, long long *vals, int n);
int swd_dispersion_set_gathers(swd_dispersion *d, const float *gathers);
int swd_dispersion_set_observed_picks(swd_dispersion *d, const double *cr0_r, const double *cr0_l);
/* The field-script variants of swd_engine_set_variant (vp_ratio ignored). */
int swd_dispersion_set_variant(swd_dispersion *d, const swd_variant *v);
int swd_dispersion_pick(swd_dispersion *d, double *cr_r, double *cr_l, double *res_r, double *res_l, double *ms);
int swd_dispersion_adjoint(swd_dispersion *d, double beta, float *adj, double *ms);

/* Arithmetic of the Radon DFT and inverse-DFT GEMMs (engine, dispersion
 * batch and RT* entry points): 0 = FP32 MFMA (default); 1 = BF16 split
 * precision, each FP32 operand split into bf16 hi + lo and three bf16 MFMA
 * products accumulated in FP32 (about 1e-5 relative).  Process-wide; returns
 * the previous mode.  No reference counterpart. */
int swd_set_radon_precision(int mode);

/* Number of Radon frequency bins (rows of the pick arrays) of a problem. */
int swd_problem_npair(const swd_problem *p, int *npair);

/*
 * Continuous-batching gradient server (no reference counterpart: the serving
 * form of the per-iteration gradient of SWD.m:171-201 for callers that solve
 * many models).  A scheduler thread admits queued requests in arrival order
 * while a slot is free and the request's buffers fit in the device budget;
 * admitted requests' shots advance together in shared FD launches, joining
 * and leaving the launch sequence independently, and each request's result
 * is bitwise the result of swd_engine_gradient on the same inputs.  Each
 * launch advances the requests with the earliest deadlines until their
 * shots fill the GPU once; the others wait with their wavefields resident.
 *
 * swd_server_create: hbm_bytes = device budget for request buffers (0: free
 * memory less 2 GiB); max_active = requests on the GPU at once (1: one
 * request at a time); nt_block_fwd / nt_block_adj = FD steps per launch
 * (2..6).  swd_server_destroy fails requests still queued, waits for the
 * admitted ones and frees everything.
 * swd_server_submit: copies the problem (and its arrays), the current model
 * vp / vs (nz x nx column-major), the wavelet (nt) and the observed picks
 * cr0_r / cr0_l (npair x nshots column-major); returns a ticket.
 * swd_server_wait: waits up to timeout_ms (< 0: no limit) for the ticket
 * (SWD_ERR_BUSY when it is still running), copies dk_vs (nz x nx column-major)
 * and the pick misfits res_r / res_l (nshots) (NULL skips), and times_ms[4] =
 * {queued, on the GPU, total, submit time since server start}, then forgets
 * the ticket; returns the request's own status (SWD_ERR_PICK etc.).
 * swd_server_stats: vals[0..9] = queued, active, completed, failed, device
 * bytes in use, device budget, FD launches, peak active, peak device bytes,
 * launch-table rebuilds.
 */
typedef struct swd_server swd_server;
int swd_server_create(long long hbm_bytes, int max_active, int nt_block_fwd, int nt_block_adj, swd_server **out);
int swd_server_destroy(swd_server *s);
int swd_server_submit(swd_server *s, const swd_problem *p, const double *vp, const double *vs,
                      const double *wavelet, const double *cr0_r, const double *cr0_l, long long *ticket);
int swd_server_wait(swd_server *s, long long ticket, double timeout_ms, double *dk_vs, double *res_r, double *res_l,
                    double *times_ms);
int swd_server_stats(swd_server *s, long long *vals, int n);
/* Launch selection: requests are advanced earliest deadline first until their
 * shots fill this many resident waves (0: the device's own count, the
 * default); the others wait with their wavefields resident.  A request's
 * deadline is its arrival time plus slack x its work (cell-steps) at a nominal
 * 250 Gcell-steps/s, so new small requests come before large ones, and a large
 * request overtakes the ones that arrive after it (slack > 0, default 2). */
int swd_server_set_wave_slots(swd_server *s, long long waves);
int swd_server_set_deadline_slack(swd_server *s, double slack);
/* Cancel a request: a queued one ends at once, an admitted one leaves the
 * launch sequence at the next launch boundary and its memory returns to the
 * arenas once the launches already queued have drained; the other requests'
 * results do not change.  swd_server_wait on it then returns SWD_ERR_BUSY.
 * SWD_ERR_ARG for an unknown or finished ticket. */
int swd_server_cancel(swd_server *s, long long ticket);

#ifdef __cplusplus
}
#endif

#endif /* SWD_ABI_H */
