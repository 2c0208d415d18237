/*
 * hgpu.h -- C ABI of the MI355X GRMHD substep library (libhgpu.so).
 *
 * Drop-in boundary for coolharm3d's per-step advance: the reference driver
 * calls step_ch() (src/globals.h:898, src/main.c:81), which reads and writes the
 * primitive state p[][][][] and the failure bookkeeping through process
 * globals.  This library keeps the state resident in HBM and exposes the same
 * step as hgpu_step(); a C shim for the reference tree maps
 * step_ch() onto these calls.  All entry points are plain C: pointers, sizes,
 * int status (0 = ok, < 0 = error, message in hgpu_last_error()).  Numerical
 * failures are never returned as errors: they are recorded in the pflag/nfail
 * maps exactly as fail() does (src/diag.c:159-492).
 *
 * Array layout at the boundary is the reference's own: all zones including
 * NG = 3 ghost layers, ordering [i][j][k][var] with var fastest (8 doubles per
 * zone), i.e. what the pointer trees p[i][j][k][l] hold (src/decomp.c:148-175).
 */
#ifndef HGPU_H
#define HGPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* Run configuration: the compile-time switches of src/options.h the advance
 * honours, as run-time values, and the grid parameters of init_base(). */
typedef struct hgpu_config {
  int n1, n2, n3;          /* interior zones N1, N2, N3 (src/harm_mpi.c:1470-1472)          */
  int metric;              /* 0: METRIC_MINK_CARTESIAN, 2: METRIC_KS_SPHERICAL (src/options.h:73-83),
                              3: the time-dependent binary-black-hole metric of hgpu_pn_geometry()
                              (METRIC_GENERAL_DYNAMIC with a PN binary, src/metric.h:104-171)     */
  int coords;              /* 0: COORD_IDENTITY, 1: COORD_DIAGONAL, 4: COORD_DIAGONAL3
                              (src/options.h:136-147; the last two with h_slope, th_beg,
                              th_length, and diag3_exponent/diag3_factor below)                  */
  int recon;               /* 2: RECON_MC, 5: RECON_PPM_MC (src/options.h:151-157)                   */
  int flux;                /* 0: FLUX_LF, 1: FLUX_HLLE (src/options.h:169-172)                        */
  int ct_para;             /* USE_FLUX_CT_PARA (src/options.h:165)                                    */
  int cooling;             /* USE_COOLING_FUNCTION == 1 (src/options.h:194)                           */
  int rescale_r;           /* RESCALE_R (src/options.h:267)                                           */
  int bc;                  /* 0: BC_CARTESIAN_OUTFLOW, 10: BC_SPHERICAL_OUTFLOW3 (src/options.h:108-132) */
  int track_nfail;         /* keep the per-zone nfail[15][] counters (src/globals.h:310)              */
  double startx[4];        /* startx[] incl. the -NG*dx ghost offset (src/init.torus.c:100-110)       */
  double dx[4];            /* dx[1..3]; dx[0] is the initial time step                                */
  double gam;              /* adiabatic index                                                          */
  double cour;             /* Courant factor                                                           */
  double t_final;          /* GridLength[0] + startx[0] (timestep clamp, src/step_ch.c:615)           */
  double a, R0, h_slope, th_beg, th_length;  /* spin and coordinate parameters (src/init.torus.c:174-200) */
  double r_isco;           /* ISCO radius used by the cooling function (src/phys.c:815-845)           */
  double h_o_r;            /* cooling target H/R (src/phys.c:655)                                     */
  double rhomin, uumin, rhopower, uupower;  /* floor model (src/options.h:232-241)                 */
  int precision;           /* 0: all FP64 (the reference's arithmetic);
                              1: mixed -- reconstruction, face fluxes and EMF averaging in FP32
                              (face states formed from FP32 differences around the FP64 zone
                              values), FP64 conserved update with compensated flux-divergence
                              sums, FP64 con-to-prim, metric, connection and cooling.
                              No counterpart in the reference.                                     */
  /* metric = 3 only: hole masses, separation and orbital phase at t = 0, and whether the
   * orbit shrinks by radiation reaction (1) or stays circular (0)                      */
  double bbh_m1, bbh_m2, bbh_sep, bbh_phi0;
  int bbh_inspiral;
  /* Domain decomposition (zero for a single block).  bc_nbr: bit 2*(d-1) + side is set when
   * face side (0 lower, 1 upper) of direction x_d borders another block, i.e. bc_pid[d][side] !=
   * BC_PHYS (src/harm_mpi.c:1560-1600): that face gets no physical boundary condition, no
   * fix_flux() (src/bounds.c:2259-2270) and no inflow_check(), and a flagged zone near it raises
   * boundary_mpi_pflag instead of boundary_phys_pflag (check_boundary_pflag(),
   * src/harm_mpi.c:1388-1430); its ghosts are filled by the caller (hgpu_upload_box()).
   * gpos: the block's globalpos[1..3] (src/harm_mpi.c:1445-1480), the global index of its first
   * interior zone, so zone coordinates are startx + (gpos + i - NG + 1/2) dx as coord() forms them
   * from the global index (src/coord.c:182-228); startx[] stays the global grid's.             */
  int bc_nbr;
  int gpos[3];
  /* 1: the conserved update's inversion stops at convergence, without the two confirming
   * Newton iterations Utoprim_2d_fast() runs after it (EXTRA_NEWT_ITER, src/u2p_defs.h:79);
   * zones the update hands to the full recovery tree keep them.  0 (default): the reference's
   * iteration count.  No counterpart in the reference (an algorithm change, in either precision). */
  int no_confirm_iters;
  /* coords = 4 (COORD_DIAGONAL3, src/coord.c:270-276): theta = pi/2 (1 + h_slope f +
   * diag3_factor f^diag3_exponent), f = 2 seesaw(X2) - 1; diag3_exponent an odd integer
   * >= 3 (the reference's globals of the same names, set by set_special_coord(),
   * src/init.nz_pn_bbh_disk.c:357-360: 9 and 1 - 2 th_cutout/pi - h_slope).           */
  int diag3_exponent;
  double diag3_factor;
} hgpu_config;

typedef struct hgpu_ctx hgpu_ctx;

/* Returns the last error message of the calling thread ("" if none). */
const char *hgpu_last_error(void);

/* Allocates the device state for one grid and computes its static geometry
 * (calc_all_geom(), src/metric.c:4111) and the x1 tables of cooling, radial
 * rescaling and floors.  Replaces set_arrays()+init_base() for the device. */
int hgpu_create(const hgpu_config *cfg, hgpu_ctx **out);
int hgpu_destroy(hgpu_ctx *ctx);

/* Grid sizes: out[0..6] = N1, N2, N3, N1TOT, N2TOT, N3TOT, NG */
int hgpu_dims(const hgpu_ctx *ctx, int *out);

/* Primitive state transfer, which = 0: p, 1: ph.  Host layout [i][j][k][8]. */
int hgpu_upload_prims(hgpu_ctx *ctx, int which, const double *host);
int hgpu_download_prims(hgpu_ctx *ctx, int which, double *host);

/* Replaces the computed geometry by the host's geom_arr / conn records
 * (struct of_geom layout, src/globals.h:90-98; conn[N_CONN][4][4][4]).
 * n_geom / n_conn must equal the counts hgpu_geometry_counts() reports.
 * With metric = 3 (no upload): the download gives the records of the state's
 * time level (zone centres) and, for conn, the 28 metric derivatives of the
 * last stage's level (t + dt/2 of the last step) stored [28][n_conn]
 * (d g_ab / dx^nu for the non-zero components ab = tt, t1, t2, t3, 11, 22,
 * 33, index nu*7 + component). */
int hgpu_geometry_counts(const hgpu_ctx *ctx, long long *n_geom, long long *n_conn);
int hgpu_upload_geometry(hgpu_ctx *ctx, const double *geom, const double *conn);
int hgpu_download_geometry(hgpu_ctx *ctx, double *geom, double *conn);

/* Integrator clock: t and dt (= dx[0]), and dt_old. */
int hgpu_set_time(hgpu_ctx *ctx, double t, double dt);
int hgpu_get_time(hgpu_ctx *ctx, double *t, double *dt, double *dt_old);

/* fixup(p) + bounds(p,0) of the initial data (src/fixup.c:8, src/bounds.c:31). */
int hgpu_fixup_bounds(hgpu_ctx *ctx);

/* step_ch() (src/step_ch.c:50-234) repeated nsteps times: predictor and
 * corrector advance() with boundaries and fixups, then t += dt and the new
 * dt = timestep().  Enqueued on the library stream; returns without waiting. */
int hgpu_step(hgpu_ctx *ctx, int nsteps);

/* on != 0: hgpu_step() captures one step_ch() into a HIP graph on first use and
 * replays it (one launch per step instead of ~60 kernel launches); per-kernel
 * profile scopes are then not recorded, "step" still is.  No counterpart in
 * the reference (launch mechanics only; results are identical). */
int hgpu_set_graph(hgpu_ctx *ctx, int on);

/* Launch shapes of the stencil kernel families (0-2: face fluxes in x1, x2,
 * x3; 3: EMFs; 4: conserved update; 5: inversion): workgroup size bs (64, 128,
 * 256), tile k-extent kc (0 = bs consecutive (j,k) points, else kc x bs/kc
 * tiles; kc divides bs), flags bit 0: XCD-contiguous tile order, bit 1: x1
 * fastest (a (j,k) tile through consecutive x1-planes), bit 2: the
 * register-capped kernel variant (families 0-2 and 4), bit 3 (family 4 only):
 * the conserved update and the inversion in one kernel, bits 4-5 (families
 * 0-2, PPM reconstruction): 0, 1, 2 for 1, 2, 4 consecutive faces per thread
 * sharing each zone's parabola, bit 6 (family 2, FP64, 2 or 4 faces per
 * thread): the reconstruction reads LDS-staged x3 pencils (each wavefront's
 * zones loaded once, coalesced, into LDS) instead of per-lane stencil loads.
 * hgpu_autotune() times every candidate of families 0-4 on the current state
 * (reps launches each; family 4 as the whole predictor update, split or fused)
 * and keeps the fastest; report (NULL or [5][4]) gets
 * bs, kc, flags, ms per launch.  The state (p, ph), p_gamma, the flags and
 * the counters are unchanged; the face fluxes hgpu_download_fluxes() returns
 * are then those of the tuning pass.  Results do not depend on the shapes.
 * No counterpart in the reference. */
int hgpu_autotune(hgpu_ctx *ctx, int reps, double *report);
int hgpu_set_launch_shape(hgpu_ctx *ctx, int family, int bs, int kc, int flags);
int hgpu_get_launch_shape(hgpu_ctx *ctx, int family, int *bs, int *kc, int *flags);

/* Binary metric (metric = 3): in_kernels = 1 evaluates every geometry record
 * inside the kernels that read it, from the holes' orbit at each level
 * (computed once per step), and forms each stage's source terms in a kernel
 * of their own from the metric derivatives; 0 fills and reads stored levels
 * (records and metric derivatives).  Both give the same results bit for bit.
 * Default 1 where possible: the GPU build with PPM reconstruction and no
 * degenerate direction (an error is returned otherwise).  The stored levels
 * (three positions-by-zone record levels and, in mixed precision, two FP32
 * face copies) are allocated when first needed: on the switch to 0, or by
 * hgpu_download_geometry.  The switch to 0 fills the running step's levels,
 * so it may come between the predictor and the corrector.  No counterpart in the reference (whose binary
 * metric bodies are absent). */
int hgpu_set_metric_eval(hgpu_ctx *ctx, int in_kernels);
int hgpu_get_metric_eval(hgpu_ctx *ctx, int *in_kernels);

/* M_tot after the last advance() stage: the sum of the updated conserved
 * density U[RHO] (sqrt(-g) rho u^t) over the block's zones, as advance()
 * accumulates it (src/step_ch.c:309, :389) and the driver prints it
 * (src/main.c:71).  The device adds the zones' terms in a fixed order of its
 * own (deterministic; the reference's serial loop order differs, so the last
 * bits may too). */
int hgpu_get_m_tot(hgpu_ctx *ctx, double *m_tot);

/* Device memory the context holds: the bytes its allocations requested
 * (state, face fluxes, queues, geometry and metric buffers).  No counterpart
 * in the reference. */
int hgpu_device_bytes(hgpu_ctx *ctx, long long *bytes);

/* One advance() stage (0: predictor into ph, 1: corrector into p) followed by
 * bounds and fixups, as step_ch() sequences them (src/step_ch.c:63-121,
 * :159-180), the flag map cleared first.  hgpu_substep(ctx, 0), hgpu_substep(ctx,
 * 1), hgpu_end_step(ctx, NULL) is one hgpu_step(ctx, 1). */
int hgpu_substep(hgpu_ctx *ctx, int stage);

/* The same stage in the two parts a decomposed run separates with its halo
 * exchange (bounds(prim,0) then bounds(prim,1) around the fixups,
 * src/step_ch.c:103-121):
 *   hgpu_advance():     failures cleared (src/step_ch.c:65-69, :117-121),
 *                       advance() and the physical boundary conditions of
 *                       bounds(prim,0) (faces not in bc_nbr; src/bounds.c:31-579);
 *   -- the caller exchanges the ghosts of the bc_nbr faces (hgpu_download_box(),
 *      bounds_mpi(), bounds_pflag_mpi(), hgpu_upload_box(), src/harm_mpi.c:1631-1744)
 *      and reduces the boundary flags (mpi_global_imax2(), src/bounds.c:2164-2166) --
 *   hgpu_fixup_stage(): with the boundary flags set to phys / mpi (the
 *                       globally reduced values): fixup_interp_v(),
 *                       fixup_interp_prim() (src/fixup.c:499-761), then the
 *                       physical part of bounds(prim,1) if phys (the flags
 *                       are cleared when the next stage starts).  The caller
 *                       exchanges ghosts again if mpi.
 * hgpu_get_boundary_flags() reads failure_exists, boundary_phys_pflag and
 * boundary_mpi_pflag of the stage (waits for the stream). */
int hgpu_advance(hgpu_ctx *ctx, int stage);
int hgpu_get_boundary_flags(hgpu_ctx *ctx, int *failure_exists, int *phys, int *mpi);
int hgpu_fixup_stage(hgpu_ctx *ctx, int stage, int phys, int mpi);

/* End of step_ch() after the corrector (src/step_ch.c:213-223): t_old = t,
 * t += dt, dt_old = dt, dt = timestep() from max_char_speed[1..3] (vmax4 !=
 * NULL: these values, e.g. after mpi_global_vmax(), src/step_ch.c:561; NULL:
 * the block's own).  hgpu_set_dt() replaces dt (after mpi_global_min(),
 * src/step_ch.c:618). */
int hgpu_end_step(hgpu_ctx *ctx, const double *vmax4);
int hgpu_set_dt(hgpu_ctx *ctx, double dt);

/* Copies of a box of zones lo[d] <= i_d <= hi[d] (indices including ghosts,
 * 0 <= lo <= hi < N_dTOT) of p (which = 0) or ph (1): primitives [i][j][k][8]
 * and pflag [i][j][k] over the box, i slowest (either pointer may be NULL).
 * The halo slabs bounds_mpi() sends and receives (src/harm_mpi.c:470-560:
 * NG layers along the face, every zone across it) are such boxes. */
int hgpu_download_box(hgpu_ctx *ctx, int which, const int *lo, const int *hi, double *prims, int *pflag);
int hgpu_upload_box(hgpu_ctx *ctx, int which, const int *lo, const int *hi, const double *prims,
                    const int *pflag);

/* Failure bookkeeping: pflag[NT] (src/macros.h:256-259), p_gamma[NT],
 * nfail[15][NT] (src/macros.h:236-251), nfailtot[15]. */
int hgpu_download_pflag(hgpu_ctx *ctx, int *host);
int hgpu_download_gamma(hgpu_ctx *ctx, double *host);
int hgpu_download_nfail(hgpu_ctx *ctx, int *host);
int hgpu_get_nfailtot(hgpu_ctx *ctx, long long *out15);
int hgpu_reset_nfail(hgpu_ctx *ctx);

/* failure_exists and max_char_speed[1..3] of the last numerical_flux(). */
int hgpu_get_status(hgpu_ctx *ctx, int *failure_exists, double *vmax4);

/* check_for_crash() (src/step_ch.c:1244-1339) evaluated on the device at the
 * end of every step: the step's new Utoprim_1d_ee failures (ntot1) and
 * interpolation-fixup failures (ntot2) from nfailtot, n_bad = max of the two,
 * against half of 2 x N1 N2 N3 (N0 = 2 time levels, src/macros.h:36).  The
 * reference ends the run at the first such step; the library keeps stepping
 * and latches it: crash_step = that step (counted from 1 since hgpu_create),
 * 0 if none; n_bad its count.  A decomposed run's reference compares with the
 * global cell count; the library uses the block's. */
int hgpu_get_crash(hgpu_ctx *ctx, long long *crash_step, long long *n_bad);

/* Face fluxes of the last advance(): F[3][8][NT] (the conserved fluxes; the
 * entropy face fluxes are evaluated only for zones that take the entropy fix). */
int hgpu_download_fluxes(hgpu_ctx *ctx, double *host);

/* Maximum |div B| at zone corners with the constrained-transport stencil
 * (divb_calc(), src/diag.c:503-602) over the interior corners, from p. */
int hgpu_divb_max(hgpu_ctx *ctx, double *out);

/* Waits for all work enqueued by this context. */
int hgpu_sync(hgpu_ctx *ctx);

/* Kernel timing with HIP events on the library stream.  Names:
 * "flux", "emf", "update", "bounds", "fixup", "step". */
int hgpu_profile_enable(hgpu_ctx *ctx, int enable);
int hgpu_profile_read(hgpu_ctx *ctx, const char *name, double *total_ms, long long *count);

/* Batched con-to-prim of independent zones (Utoprim_2d_fast(),
 * src/utoprim_2d_fast.c:135, called by recover_primitives(), src/step_ch.c:832)
 * on device buffers laid out structure-of-arrays: U[8][n], geometry records
 * geom[n] (struct of_geom), prim[8][n] (guess in, solution out), gamma[n],
 * code[n] (return code), iters[n] (Newton iterations).  Pointers are device
 * addresses; the call is enqueued on the library stream. */
int hgpu_con2prim_batch(hgpu_ctx *ctx, long long n, const double *U, const double *geom,
                        double *prim, double *gamma, int *code, int *iters);

/* ------------------------------------------------------------------------
 * Batched con-to-prim of independent zones on Kerr-Schild records, with a
 * learned initial guess.  No context: every call takes the stream to enqueue
 * on (a hipStream_t, NULL = the null stream) and works only in memory the
 * caller passes (device addresses; host addresses in the host build).  Arrays
 * are structure-of-arrays: U[8][n], rth[2][n] (r, theta), prim[8][n].
 * Spin a and adiabatic index gam are arguments.  n < 2^31.
 * No counterpart in the reference: the reference inverts zone by zone inside
 * recover_primitives() (src/step_ch.c:832) with Utoprim_2d_fast()
 * (src/utoprim_2d_fast.c:135), which is the per-zone algorithm used here.
 * ------------------------------------------------------------------------ */

/* Device plumbing for callers without a GPU array library: device count,
 * device buffers, copies (kind 0: host to device, 1: device to host -- waits,
 * 2: device to device) on a stream, and events for timing work on a stream.
 * No counterpart in the reference. */
int hgpu_dev_count(int *out);
int hgpu_dev_alloc(long long bytes, void **out);
int hgpu_dev_free(void *p);
int hgpu_dev_copy(void *dst, const void *src, long long bytes, int kind, void *stream);
/* device-to-device copy by a plain streaming kernel (bytes % 16 == 0, 16-byte
 * aligned): traffic of known size for calibrating HBM byte counters */
int hgpu_dev_copy_kernel(void *dst, const void *src, long long bytes, void *stream);
int hgpu_stream_sync(void *stream);
int hgpu_event_create(void **ev);
int hgpu_event_destroy(void *ev);
int hgpu_event_record(void *ev, void *stream);
int hgpu_event_elapsed(void *ev0, void *ev1, float *ms);

/* Samples of the inversion workload from seed: r log-uniform in [1.3, 100],
 * theta uniform in [0.05, pi - 0.05], Lorentz factor log-uniform in [1, 50],
 * b^2/rho log-uniform in [1e-4, 1e2], u/rho log-uniform in [1e-4, 10], rho = 1,
 * velocity and field directions isotropic in the normal observer's frame.
 * Writes U, rth, the true primitives prim, and (if guess != NULL) the
 * perturbed-truth guess guess[8][n] (rho, u, u~ times 1 +- 10 %). */
int hgpu_c2p_samples(void *stream, long long n, unsigned long long seed, double a, double gam,
                     double *U, double *rth, double *prim, double *guess);

/* Workspace hgpu_c2p_solve() needs (bytes; 0 when compact == 0). */
long long hgpu_c2p_workspace_bytes(long long n, int compact);

/* Utoprim_2d_fast() on every zone.  Initial guess: mlp == NULL -> the plain
 * guess; else the network's guess, restarting once from the plain guess if that
 * attempt fails.  plain_guess: 0 = the incoming prim (the reference's guess),
 * 1 = cold (rho = D, u = 1e-3 D, u~ = 0).  compact = 1 runs a persistent first
 * pass of budget_first Newton iterations per zone, then passes of budget_next
 * over a compacted queue of the zones still iterating (0 = defaults); compact =
 * 0 solves each zone to the end in one pass.  Both give identical results.
 * Out: prim (B from U; rho, u, u~ the solution, or the plain guess where no
 * attempt succeeded), code (the return code of Utoprim_2d_fast()), iters
 * (Newton iterations over both attempts). */
int hgpu_c2p_solve(void *stream, long long n, double a, double gam, const double *U, const double *rth,
                   const float *mlp, int plain_guess, int compact, int budget_first, int budget_next,
                   void *work, long long work_bytes, double *prim, int *code, int *iters);

/* hgpu_c2p_solve() with compact = 1, allocating nothing: its compaction
 * queues live in a static device array of the library (80 MiB on each GPU it
 * runs on, held until the process ends), so n is at most 262144 zones per
 * call; larger n fails with status -3 and a message.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * caller passes inputs and outputs as for hgpu_c2p_solve(); results are
 * identical to it.  Every call on one GPU shares that GPU's array, so calls
 * that overlap on one GPU (concurrent streams or host threads) race. */
int hgpu_c2p_solve_static(void *stream, long long n, double a, double gam, const double *U, const double *rth,
                          const float *mlp, int plain_guess, int budget_first, int budget_next, double *prim,
                          int *code, int *iters);

/* The network (FP32): 8 inputs -> 64 GELU -> 64 GELU -> 2 outputs (log(W/D),
 * log gamma).  out[0..5] = parameter count, trainable count, offset of the
 * trainable part, inputs, hidden width, outputs. */
int hgpu_mlp_sizes(int *out);

/* Features feat[8][n] of each zone and, if prim != NULL (solved primitives),
 * the training targets target[2][n]. */
int hgpu_mlp_features(void *stream, long long n, double a, double gam, const double *U, const double *rth,
                      const double *prim, float *feat, float *target);

/* Network outputs y[2][n]. */
int hgpu_mlp_forward(void *stream, long long n, const float *w, const float *feat, float *y);

/* Workspace of hgpu_mlp_grad() for n samples / hgpu_mlp_train() for a batch. */
long long hgpu_mlp_workspace_bytes(long long batch);

/* Mean loss 0.5 |o - t_n|^2 over the n samples and its gradient with respect
 * to the trainable parameters, grad[trainable count]. */
int hgpu_mlp_grad(void *stream, long long n, const float *w, const float *feat, const float *target,
                  void *work, long long work_bytes, float *grad, float *loss);

/* Input/output normalisation from the data and seeded initial weights. */
int hgpu_mlp_init(void *stream, unsigned long long seed, long long n, const float *feat, const float *target,
                  float *w);

/* steps Adam steps (cosine-decayed learning rate lr) on batches of `batch`
 * samples drawn with replacement; loss_hist[steps] gets each batch's loss. */
int hgpu_mlp_train(void *stream, long long n, const float *feat, const float *target, unsigned long long seed,
                   int steps, int batch, float lr, void *work, long long work_bytes, float *w,
                   float *loss_hist);

/* ------------------------------------------------------------------------
 * Time-dependent metric of an inspiralling black-hole binary (the circumbinary
 * disk configuration).  The reference declares these metrics
 * (dyn_35pn_nz_gcov_func, dyn_fullpn_nz_gcov_func, src/metric.h:104-171) without
 * bodies; this one is the 1PN near-zone metric of two point masses in harmonic
 * coordinates (Blanchet, Faye & Ponsot 1998) with a quasi-circular orbit
 * shrinking by leading-order radiation reaction (Peters 1964), on spherical
 * numerical coordinates X1 = log(r - R0), theta(X2) as the modified
 * Kerr-Schild map, X3 = phi, about the centre of mass.
 * ------------------------------------------------------------------------ */
typedef struct hgpu_pn_config {
  double m1, m2;             /* hole masses                                          */
  double b0, phi0;           /* separation and orbital phase at t = 0               */
  int inspiral;              /* 1: radiation-reaction inspiral, 0: fixed circular orbit */
  double R0, th_beg, th_length, h_slope;  /* numerical-coordinate map              */
  int n1, n2, n3;            /* zones of the block (ghosts included if wanted)     */
  double startx[4], dx[4];   /* X of the block's first zone corner, zone widths    */
  int coords;                /* theta map: 0 or 1 COORD_DIAGONAL, 4 COORD_DIAGONAL3 */
  int diag3_exponent;        /* COORD_DIAGONAL3 parameters (as in hgpu_config)     */
  double diag3_factor;
} hgpu_pn_config;

/* Geometry records (struct of_geom, 42 doubles) of every zone of the block at
 * position pos (0 FACE1, 1 FACE2, 2 FACE3, 3 CENT, 4 CORN as coord(),
 * src/coord.c:182-228) and time t: geom[(i*n2 + j)*n3 + k][42]. */
int hgpu_pn_geometry(void *stream, const hgpu_pn_config *cfg, double t, int pos, double *geom);

/* Connection Gamma^d_{ab} (time derivatives included) at zone centres and
 * time t: conn[(i*n2 + j)*n3 + k][d][a][b]. */
int hgpu_pn_connection(void *stream, const hgpu_pn_config *cfg, double t, double *conn);

#ifdef __cplusplus
}
#endif
#endif /* HGPU_H */
