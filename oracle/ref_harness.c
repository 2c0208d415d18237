/*
 * ref_harness.c -- test infrastructure only.
 *
 * Drives the unmodified coolharm3d C sources (compiled in place from the
 * reference tree by oracle/Makefile, with per-configuration options.h edits)
 * as a shared library that the tests load through ctypes.  Nothing in the
 * product links or loads this file.
 *
 * The reference's own driver file main.c is compiled into this translation
 * unit with its main() renamed, so its static set-up routines
 * (set_initial_globals, set_arrays, global_init; src/main.c:548, :139, :283)
 * run exactly as in the reference executable.  The time loop of
 * src/main.c:64-115 is replaced by ref_step(), which calls step_ch()
 * (src/step_ch.c:50) the requested number of times, so state can be read back
 * between steps.
 */
#define main coolharm3d_reference_main
#include "main.c"
#undef main

extern void fixup(double ****pv);                       /* src/fixup.c:8   */
extern void bounds(double ****prim_arr, int skip_mpi);  /* src/bounds.c:31 */
extern void set_levi_civita(void);

/* Parameters handed to the problem set-up (ref_init_*.c reads them). */
double ref_params[64];

static int ref_is_setup = 0;

/* ref_setup(): mirrors src/main.c:17-62 up to (not including) init(). */
int ref_setup(int n1, int n2, int n3)
{
  char a0[] = "harm3d", a1[] = "1", a2[] = "1", a3[] = "1";
  char s1[32], s2[32], s3[32];
  char *argv[8];
  snprintf(s1, sizeof(s1), "%d", n1);
  snprintf(s2, sizeof(s2), "%d", n2);
  snprintf(s3, sizeof(s3), "%d", n3);
  argv[0] = a0; argv[1] = a1; argv[2] = a2; argv[3] = a3;
  argv[4] = s1; argv[5] = s2; argv[6] = s3; argv[7] = NULL;
  if (ref_is_setup) return -1;   /* the reference allocates once per process */
  set_initial_globals();
  treat_floor_as_failure = 0;
  setup_mpi(7, argv);
  set_arrays();
  ref_is_setup = 1;
  return 0;
}

/* Runs the linked problem's init() (src/main.c:37) then global_init(). */
void ref_init(void)
{
  init();
  t = t_old = startx[0];
  global_init();
}

/* Re-runs global_init() after parameters were changed by the caller. */
void ref_global_init(void) { global_init(); }

void ref_dims(int *out)
{
  out[0] = N1; out[1] = N2; out[2] = N3;
  out[3] = N1TOT; out[4] = N2TOT; out[5] = N3TOT; out[6] = NG;
}

/* Scalars: t, dt (=dx[0]), cour, gam, dx1..3, startx1..3, a, R0, h_slope,
   th_length, th_beg, r_isco, M */
void ref_get_scalars(double *out)
{
  out[0] = t;        out[1] = dx[0];     out[2] = cour;     out[3] = gam;
  out[4] = dx[1];    out[5] = dx[2];     out[6] = dx[3];
  out[7] = startx[1]; out[8] = startx[2]; out[9] = startx[3];
  out[10] = a;       out[11] = R0;       out[12] = h_slope;
  out[13] = th_length; out[14] = th_beg; out[15] = r_isco;  out[16] = M;
  out[17] = GridLength[0]; out[18] = startx[0]; out[19] = dt_old;
}

void ref_set_time(double tt, double dt, double cour_in)
{
  t = t_old = tt;
  dx[0] = dt;
  cour = cour_in;
}

void ref_set_end_time(double tend) { GridLength[0] = tend - startx[0]; }

/* All cells, ordering [i][j][k][l] with l fastest (the reference's own). */
void ref_get_prims(double *out)
{
  int i, j, k, l;
  size_t n = 0;
  ALL_LOOP PLOOP { out[n++] = p[i][j][k][l]; }
}

/* half-step state ph (the predictor's result after bounds and fixups) */
void ref_get_ph(double *out)
{
  int i, j, k, l;
  size_t n = 0;
  ALL_LOOP PLOOP { out[n++] = ph[i][j][k][l]; }
}

void ref_set_prims(const double *in)
{
  int i, j, k, l;
  size_t n = 0;
  ALL_LOOP PLOOP { p[i][j][k][l] = in[n++]; }
}

void ref_get_gamma(double *out)
{
  int i, j, k;
  size_t n = 0;
  ALL_LOOP { out[n++] = p_gamma[i][j][k]; }
}

void ref_get_pflag(int *out)
{
  int i, j, k;
  size_t n = 0;
  ALL_LOOP { out[n++] = pflag[i][j][k]; }
}

/* nfail counters: [N_NFAIL][i][j][k] */
void ref_get_nfail(int *out)
{
  int i, j, k, l;
  size_t n = 0;
  for (l = 0; l < N_NFAIL; l++) ALL_LOOP { out[n++] = nfail[l][i][j][k]; }
}

/* zero nfail[][] and nfailtot[] (the dump routines do this after writing them,
   src/harm_mpi.c:917-927) */
void ref_reset_nfail(void)
{
  int i, j, k, l;
  for (l = 0; l < N_NFAIL; l++) { nfailtot[l] = 0; ALL_LOOP { nfail[l][i][j][k] = 0; } }
}

void ref_get_nfailtot(long *out)
{
  int l;
  for (l = 0; l < N_NFAIL; l++) out[l] = (long) nfailtot[l];
}

/* Conserved variables of the newest substep (U_gf[n_U]) and of substep 0. */
void ref_get_U(int level, double *out)
{
  int i, j, k, l;
  size_t n = 0;
  ALL_LOOP PLOOP { out[n++] = U_gf[level][i][j][k][l]; }
}

/* Face fluxes F[i][j][k][d][l] of the last advance() */
void ref_get_F(double *out)
{
  int i, j, k, d, l;
  size_t n = 0;
  ALL_LOOP FACE_LOOP PLOOP { out[n++] = F[i][j][k][d][l]; }
}

/* Geometry at (i,j) for position pos (KS: 2D cache; Minkowski: one entry).
   Layout per entry: gcov[16], gcon[16], g, g_inv, alpha, beta[3], ncon[4] = 42 */
int ref_get_geom(int i, int j, int k, int pos, double *out)
{
  struct of_geom *geom;
  int a1, b1, n = 0;
  get_geometry(i, j, k, pos, ncurr, geom);
  for (a1 = 0; a1 < NDIM; a1++) for (b1 = 0; b1 < NDIM; b1++) out[n++] = geom->gcov[a1][b1];
  for (a1 = 0; a1 < NDIM; a1++) for (b1 = 0; b1 < NDIM; b1++) out[n++] = geom->gcon[a1][b1];
  out[n++] = geom->g; out[n++] = geom->g_inv; out[n++] = geom->alpha;
  out[n++] = geom->beta[0]; out[n++] = geom->beta[1]; out[n++] = geom->beta[2];
  out[n++] = geom->ncon[0]; out[n++] = geom->ncon[1]; out[n++] = geom->ncon[2]; out[n++] = geom->ncon[3];
  return n;
}

/* Connection at interior cell (i,j,k): conn[a][b][c] = Gamma^a_{bc} */
int ref_get_conn(int i, int j, int k, double *out)
{
  int l = CONN_ID(i, j, k), a1, b1, c1, n = 0;
  for (a1 = 0; a1 < NDIM; a1++) for (b1 = 0; b1 < NDIM; b1++) for (c1 = 0; c1 < NDIM; c1++)
    out[n++] = conn[l][a1][b1][c1];
  return n;
}

/* Coordinate record at (i,j,k,pos): x[4], xp[4], r, rhoflr, uuflr */
int ref_get_coord(int i, int j, int k, int pos, double *out)
{
  struct of_coord *coords;
  int n = 0, l;
  get_coord(i, j, k, pos, ncurr, coords);
  for (l = 0; l < NDIM; l++) out[n++] = coords->x[l];
  for (l = 0; l < NDIM; l++) out[n++] = coords->xp[l];
  out[n++] = coords->r; out[n++] = coords->rhoflr; out[n++] = coords->uuflr;
  return n;
}

/* init-time fixups and boundaries as the problem set-ups call them
   (src/init.torus.c:157-163) */
void ref_fixup_bounds(void)
{
  fixup(p);
  bounds(p, 0);
}

void ref_bounds(void) { bounds(p, 0); }

/* Advance n full steps with the reference integrator. */
void ref_step(int nsteps)
{
  int s;
  for (s = 0; s < nsteps; s++) {
    step_ch();
    nstep++;
  }
}

/* Diagnostics carried by the reference globals. */
void ref_get_diag(double *out)
{
  out[0] = M_tot;
  out[1] = (double) failure_exists;
  out[2] = (double) nstep;
}

/* Corner divergence of B, src/diag.c:503 (divb_calc); returns max |div B|. */
extern double divb_calc(int i, int j, int k);

/* Batched inversion of independent Kerr-Schild zones with the reference's own
   Utoprim_2d_fast() (src/utoprim_2d_fast.c:135) on records built by its metric
   functions ks_gcov_func(), ks_gcon_func() (src/metric.c:1168-1252) and
   bl_gdet_func() (:523-539) in Kerr-Schild coordinates, with the lapse and
   shift of calc_all_geom() (:4221-4232).  Arrays are structure-of-arrays:
   U[8][n], rth[2][n], prim[8][n] (guess in, result out).  Needs no ref_setup(). */
extern int Utoprim_2d_fast(double *U, struct of_geom *geom, double *prim, double *gamma, double *bsq);
extern void bl_gdet_func(double *x, double *gdet);
void ref_c2p_batch(long n, double a_spin, double gam_in, const double *U, const double *rth, double *prim,
                   int *code)
{
  long z;
  int l;
  M = 1.;
  a = a_spin;
  asq = a * a;
  gam = gam_in;
  gam_m1_o_gam = (gam - 1.) / gam;
  for (z = 0; z < n; z++) {
    struct of_geom geom;
    double x[NDIM], Uz[NP], pz[NP], gamma = 1., bsq = 0.;
    x[0] = 0.; x[1] = rth[z]; x[2] = rth[n + z]; x[3] = 0.;
    ks_gcov_func(x, geom.gcov);
    ks_gcon_func(x, geom.gcon);
    bl_gdet_func(x, &(geom.g));
    geom.g_inv = 1. / geom.g;
    geom.ncon[0] = sqrt(-geom.gcon[0][0]);
    geom.alpha = 1. / geom.ncon[0];
    geom.ncon[1] = -geom.alpha * geom.gcon[0][1];
    geom.ncon[2] = -geom.alpha * geom.gcon[0][2];
    geom.ncon[3] = -geom.alpha * geom.gcon[0][3];
    geom.beta[0] = -geom.gcon[0][1] / geom.gcon[0][0];
    geom.beta[1] = -geom.gcon[0][2] / geom.gcon[0][0];
    geom.beta[2] = -geom.gcon[0][3] / geom.gcon[0][0];
    for (l = 0; l < NP; l++) { Uz[l] = U[l * n + z]; pz[l] = prim[l * n + z]; }
    code[z] = Utoprim_2d_fast(Uz, &geom, pz, &gamma, &bsq);
    for (l = 0; l < NP; l++) prim[l * n + z] = pz[l];
  }
}

/* ------------------------------------------------------------------------
 * Branch coverage of the recovery tree.  The library is linked with
 * -Wl,--wrap for the solvers recover_primitives() and fixup_entropy_eq() call
 * (src/step_ch.c:832, :842; src/fixup.c:401, :409) and for the two
 * interpolating fixups (src/step_ch.c:106-107, :172-173), so every call from
 * the reference's own objects passes through a counter here on its way to the
 * unmodified function.  ref_branch[s][0] counts calls of solver s,
 * ref_branch[s][1] the calls that returned 0; for s = 0..3 the solvers are
 * Utoprim_2d_fast, Utoprim_1d, Utoprim_1d_ee, Utoprim_1d_ee2; s = 4, 5 are
 * fixup_interp_v and fixup_interp_prim (calls only).  The return codes of each
 * solver are tallied in ref_codes[s][0..REF_NCODES-1] as (code, count) pairs.
 * ------------------------------------------------------------------------ */
#define REF_NBRANCH 6
#define REF_NCODES 32
long ref_branch[REF_NBRANCH][2];
long ref_codes[REF_NBRANCH][REF_NCODES][2];

static void ref_tally(int s, int rc)
{
  int c;
  ref_branch[s][0]++;
  if (rc == 0) ref_branch[s][1]++;
  for (c = 0; c < REF_NCODES; c++) {
    if (ref_codes[s][c][1] == 0) { ref_codes[s][c][0] = rc; ref_codes[s][c][1] = 1; return; }
    if (ref_codes[s][c][0] == rc) { ref_codes[s][c][1]++; return; }
  }
}

void ref_reset_branch(void)
{
  memset(ref_branch, 0, sizeof(ref_branch));
  memset(ref_codes, 0, sizeof(ref_codes));
}

int __real_Utoprim_2d_fast(double *U, struct of_geom *geom, double *prim, double *gamma, double *bsq);
int __wrap_Utoprim_2d_fast(double *U, struct of_geom *geom, double *prim, double *gamma, double *bsq)
{
  int rc = __real_Utoprim_2d_fast(U, geom, prim, gamma, bsq);
  ref_tally(0, rc);
  return rc;
}

int __real_Utoprim_1d(double U[NP], double gcov[NDIM][NDIM], double gcon[NDIM][NDIM], double gdet,
                      double prim[NP], double *gamma, double *bsq);
int __wrap_Utoprim_1d(double U[NP], double gcov[NDIM][NDIM], double gcon[NDIM][NDIM], double gdet,
                      double prim[NP], double *gamma, double *bsq)
{
  int rc = __real_Utoprim_1d(U, gcov, gcon, gdet, prim, gamma, bsq);
  ref_tally(1, rc);
  return rc;
}

int __real_Utoprim_1d_ee(double U[NP], double S, struct of_geom *geom, double prim[NP], double *gamma);
int __wrap_Utoprim_1d_ee(double U[NP], double S, struct of_geom *geom, double prim[NP], double *gamma)
{
  int rc = __real_Utoprim_1d_ee(U, S, geom, prim, gamma);
  ref_tally(2, rc);
  return rc;
}

int __real_Utoprim_1d_ee2(double U[NP], double S, struct of_geom *geom, double prim[NP], double *gamma);
int __wrap_Utoprim_1d_ee2(double U[NP], double S, struct of_geom *geom, double prim[NP], double *gamma)
{
  int rc = __real_Utoprim_1d_ee2(U, S, geom, prim, gamma);
  ref_tally(3, rc);
  return rc;
}

void __real_fixup_interp_v(double ****pv);
void __wrap_fixup_interp_v(double ****pv) { ref_tally(4, 0); __real_fixup_interp_v(pv); }
void __real_fixup_interp_prim(double ****pv);
void __wrap_fixup_interp_prim(double ****pv) { ref_tally(5, 0); __real_fixup_interp_prim(pv); }

/* Checkpoints through the reference's own ASCII restart files (rdump0/1.<pid>
   in the working directory; src/restart.c:22-75 write, :80-220 read): */
extern void restart_write(void);
extern int restart_init(void);
void ref_restart_write(void) { restart_write(); }
/* restart_init() instead of init() (src/main.c:49-59); -1 when no file exists */
int ref_restart_init(void)
{
  if (!restart_init()) return -1;
  global_init();
  return 0;
}
void ref_set_n_restart(int n) { n_restart = n; }
int ref_get_nstep(void) { return nstep; }

/* The driver's per-step crash test (src/main.c:100 -> check_for_crash(),
   src/step_ch.c:1244-1339): ends the process through fail(FAIL_BASIC) once a
   step's new failures exceed its threshold. */
extern void check_for_crash(void);
void ref_check_for_crash(void) { check_for_crash(); }
