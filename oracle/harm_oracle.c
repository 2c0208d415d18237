/* harm_oracle.c -- CPU restatement of coolharm3d's step_ch() (TEST INFRASTRUCTURE).
 *
 * The checker of the GPU library: only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg load it.  Plain C, single-threaded, written from
 * the reference's algorithm with the reference's operation order (built with
 * -ffp-contract=off) so that, on the configurations it supports, its results
 * equal the reference build's bit for bit except where noted below.
 *
 * Structure follows the reference, not the GPU kernels:
 *   state p, ph, p_old          [i][j][k][8]          (src/globals.h:282-287)
 *   p_L, p_R, F                 [i][j][k][3][8]       (src/globals.h:284-286)
 *   c_gf, ucon_L, ucon_R        [i][j][k][3][2]       (src/step_ch.c:743-750)
 *   U_gf[2]                     [i][j][k][8]          (src/globals.h:288)
 *   emf[3]                      [i][j][k]             (src/globals.h:293)
 *   geometry                    [pos][i][j] records of struct of_geom (metric_dim 2)
 * and the passes are the reference's: reconstruct_fast(), numerical_flux(),
 * fix_flux(), flux_ct() / flux_ct_para_fast2(), the per-zone loop of advance(),
 * bounds(), fixup_interp_v()/fixup_interp_prim(), timestep().
 *
 * Differences from the reference build, each pinned in tests/test_oracle.py:
 *  - the Kerr-Schild connection is computed by forward-mode automatic
 *    differentiation of the metric in numerical coordinates (Christoffel
 *    formula), not by the reference's expanded expression
 *    (src/metric.c:1259-1697): equal to round-off (relative 1e-15);
 *  - the first cooled zone's fail() counts are recorded at that zone; the
 *    reference records them at (N1TOT-1, N2S, N3S) because the one-time
 *    cooling-table set-up overwrites the current-zone globals
 *    (src/phys.c:690-695 via get_coord()).
 */
#include "harm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NG 3
#define NP 8
#define NPOS 5
#define GEOMD 42
enum { RHO = 0, UU = 1, U1 = 2, U2 = 3, U3 = 4, B1 = 5, B2 = 6, B3 = 7 };
enum { FACE1 = 0, FACE2 = 1, FACE3 = 2, CENT = 3, CORN = 4 };
enum { PF_NONE = 0, PF_INTERP_V = 1, PF_INTERP_PRIM = 2 };
/* NFAIL_* indices, src/macros.h:236-251 */
enum { NF_FLOOR = 0, NF_TMAX = 1, NF_GAMMA_MAX = 2, NF_GAMMA_CALC = 3, NF_USE_FULL_INV = 4,
       NF_USE_EE = 5, NF_UTOPRIM = 6, NF_UTOPRIM_EE = 7, NF_USE_INTERP_PRIM = 8,
       NF_USE_INTERP_V = 9, NF_INTERP_PRIM = 10, NF_INTERP_V = 11, N_NF = 15 };

/* constants of src/macros.h, src/options.h, src/u2p_defs.h */
static const double SMALL = 1.e-15;                 /* src/macros.h:441 */
static const double SAFE = 1.3;                     /* src/macros.h:442 */
static const double SINGSMALL = 1.e-16;             /* src/options.h:271 */
static const double GAMMAMAX = 50.;                 /* src/options.h:264 */
static const double GAMMAMAX2 = 1.e4;               /* src/options.h:265 */
static const double BETA_MIN = 1.e-2;               /* src/options.h:212 */
static const double PPM_FACTOR = 0.16666666666666666667;  /* src/recon.h:10-11 */
#define MAX_NEWT_ITER 30                            /* src/u2p_defs.h:76-91 */
static const double NEWT_TOL = 1.0e-10;
static const double MIN_NEWT_TOL = 1.0e-10;
#define EXTRA_NEWT_ITER 2
static const double NEWT_TOL2 = 1.0e-15;
static const double W_TOO_BIG = 1.e20;
static const double UTSQ_TOO_BIG = 1.e20;
static const double FAIL_VAL = 1.e30;

typedef struct {           /* struct of_geom, src/globals.h:90-98 */
  double gcon[4][4], gcov[4][4];
  double g, g_inv, alpha, beta[3], ncon[4];
} geom_t;

typedef struct {           /* struct of_state, src/globals.h:100-109 */
  double ucon[4], ucov[4], bcon[4], bcov[4];
  double bsq, p, eta, ptot;
} state_t;

struct orc {
  orc_config c;
  int n[3], nt3[3];
  long nt;
  int do_recon[3], degenerate, nsd, metric_dim, buf;
  double dx[4], invdx[4], gam, gmog;
  double t, dt, t_old, dt_old;
  double *p, *ph, *p_old, *pL, *pR, *F, *cgf, *uconL, *uconR, *Ugf, *emf, *pgamma;
  double *rhoflr, *uuflr;
  int *pflag, *nfail;
  long long nfailtot[N_NF];
  int failure_exists, failure_type[3], boundary_phys_pflag;
  double vmax[4];
  geom_t* geom;
  double* conn;
  double *omega, *tr, *rscale, *irscale;
  double *pvect, *dqvect, *pLvect, *pRvect;
  int n_substep, n_U;
  int ic, jc, kc, pc;          /* icurr, jcurr, kcurr, pcurr of the reference */
};

/* ------------------------------------------------------------------------ */
/* indexing                                                                  */
static long IX(const orc_t* o, int i, int j, int k) { return ((long)i * o->nt3[1] + j) * o->nt3[2] + k; }
static double* P8(double* a, const orc_t* o, int i, int j, int k) { return a + IX(o, i, j, k) * NP; }
static double* F8(double* a, const orc_t* o, int i, int j, int k, int d) { return a + (IX(o, i, j, k) * 3 + d) * NP; }
static double* C2(double* a, const orc_t* o, int i, int j, int k, int d) { return a + (IX(o, i, j, k) * 3 + d) * 2; }
static double* EMF(orc_t* o, int d, int i, int j, int k) { return o->emf + d * o->nt + IX(o, i, j, k); }

/* get_geometry(), src/macros.h:590-605: [pos][i][j] for axisymmetric metrics */
static const geom_t* GEOM(const orc_t* o, int i, int j, int pos) {
  if (o->metric_dim == 0) return o->geom;
  return o->geom + ((long)pos * o->nt3[0] + i) * o->nt3[1] + j;
}

static double dmax(double a, double b) { return (a > b) ? a : b; }   /* MAX(), src/macros.h:474 */

/* ------------------------------------------------------------------------ */
/* coordinates and metric: src/coord.c, src/metric.c                          */

/* seesaw(x, 0, 1), src/coord.c:4316-4324 */
static double seesaw01(double x) {
  const double x0 = 0., xn = 1., xn2 = 2. * xn - x0;
  int shift = (x < x0) ? (1 + (int)(fabs(x - x0) / (xn2 - x0))) : (-(int)(fabs(x - x0) / (xn2 - x0)));
  double x1 = x + (xn2 - x0) * shift;
  return (x1 < xn) ? (x1 - x0) : (xn2 - x1);
}

/* coord(), src/coord.c:182-228 (single domain: global index = local index) */
static void coord_xp(const orc_t* o, int i, int j, int k, int pos, double* xp) {
  const double* sx = o->c.startx;
  const double* dx = o->dx;
  double fi = (pos == FACE1 || pos == CORN) ? i : i + 0.5;
  double fj = (pos == FACE2 || pos == CORN) ? j : j + 0.5;
  double fk = (pos == FACE3 || pos == CORN) ? k : k + 0.5;
  xp[0] = 0.;
  xp[1] = sx[1] + fi * dx[1];
  xp[2] = sx[2] + fj * dx[2];
  xp[3] = sx[3] + fk * dx[3];
}

/* x_of_xp(), src/coord.c:240-276, and dx_dxp_calc(), src/coord.c:435-452 */
static void x_of_xp(const orc_t* o, const double* xp, double* x, double J[4][4]) {
  int a, b;
  for (a = 0; a < 4; a++) for (b = 0; b < 4; b++) J[a][b] = (a == b) ? 1. : 0.;
  if (o->c.coords == 1) {
    double xp2 = seesaw01(xp[2]);
    x[0] = xp[0];
    x[1] = o->c.R0 + exp(xp[1]);
    x[2] = o->c.th_length * xp2 + o->c.th_beg + o->c.h_slope * sin(2 * M_PI * xp2);
    x[3] = xp[3];
    J[1][1] = x[1] - o->c.R0;
    J[2][2] = o->c.th_length + 2 * M_PI * o->c.h_slope * cos(2 * M_PI * xp[2]);
  } else if (o->c.coords == 4) {
    /* COORD_DIAGONAL3, src/coord.c:270-276 and :468-472 */
    double f = 2 * seesaw01(xp[2]) - 1.;
    x[0] = xp[0];
    x[1] = o->c.R0 + exp(xp[1]);
    x[2] = 0.5 * M_PI * (1. + o->c.h_slope * f + o->c.diag3_factor * pow(f, o->c.diag3_exponent));
    x[3] = xp[3];
    J[1][1] = x[1] - o->c.R0;
    J[2][2] = M_PI * (o->c.h_slope + o->c.diag3_exponent * o->c.diag3_factor * pow(f, o->c.diag3_exponent - 1));
  } else {
    for (a = 0; a < 4; a++) x[a] = xp[a];
  }
}

/* trigvalues(), src/metric.c:29-40 */
static void trigvalues(double th, double* cth, double* sth) {
  *sth = sin(th);
  *cth = cos(th);
  if (fabs(*sth) < SINGSMALL) *sth = ((*sth) >= 0) ? SINGSMALL : -SINGSMALL;
}

/* ks_gcov_func(), ks_gcon_func(), src/metric.c:1168-1252; bl_gdet_func(), :523-539 */
static void ks_gcov(const orc_t* o, double r, double th, double g[4][4]) {
  const double M = 1., a = o->c.a, asq = a * a;
  double c, s;
  int p, q;
  trigvalues(th, &c, &s);
  for (p = 0; p < 4; p++) for (q = 0; q < 4; q++) g[p][q] = 0.;
  double rr = r * r, sig = rr + asq * (c * c), m = M * r / sig, z = 2.0 * m, s2 = s * s;
  g[0][0] = z - 1.0;
  g[0][1] = g[1][0] = z;
  g[0][3] = g[3][0] = -2.0 * m * a * s2;
  g[1][1] = 1.0 + z;
  g[1][3] = g[3][1] = -a * g[1][1] * s2;
  g[2][2] = sig;
  g[3][3] = s2 * (rr + asq * (1.0 + 2.0 * m * s2));
}
static void ks_gcon(const orc_t* o, double r, double th, double g[4][4]) {
  const double M = 1., a = o->c.a, asq = a * a;
  double c, s;
  int p, q;
  trigvalues(th, &c, &s);
  for (p = 0; p < 4; p++) for (q = 0; q < 4; q++) g[p][q] = 0.;
  double mr = M * r, rr = r * r, isig = 1 / (rr + asq * (c * c)), z = 2.0 * mr * isig;
  g[0][0] = -1.0 - z;
  g[0][1] = g[1][0] = z;
  g[1][1] = (rr - 2.0 * mr + asq) * isig;
  g[1][3] = g[3][1] = a * isig;
  g[2][2] = isig;
  g[3][3] = isig / (s * s);
}
static double ks_gdet(const orc_t* o, double r, double th) {
  double c, s;
  trigvalues(th, &c, &s);
  double cos_2th = (c - s) * (c + s), a2 = o->c.a * o->c.a, r2 = r * r;
  return r2 * fabs(s) * (1. + 0.5 * (a2 / r2) * (1. + cos_2th));
}

/* calc_all_geom(), src/metric.c:4111-4233, for one position: metric in x
   coordinates, transform_all() (src/coord.c:2340-2353, full rank-2 sums of
   transform_rank2cov2()/transform_rank2con2(), :1852-1872, :1975-1995, with
   dxp_dx from dxp_dx_calc2_default(), :1629-1654), then lapse and shift. */
static void make_geom(const orc_t* o, int i, int j, int k, int pos, geom_t* G) {
  double xp[4], x[4], J[4][4], K[4][4], gc[4][4], gn[4][4], g;
  int a, b, p, q;
  coord_xp(o, i, j, k, pos, xp);
  x_of_xp(o, xp, x, J);
  if (o->c.metric == 2) {
    ks_gcov(o, x[1], x[2], gc);
    ks_gcon(o, x[1], x[2], gn);
    g = ks_gdet(o, x[1], x[2]);
  } else {
    for (a = 0; a < 4; a++) for (b = 0; b < 4; b++) gc[a][b] = gn[a][b] = (a == b) ? ((a == 0) ? -1. : 1.) : 0.;
    g = 1.;
  }
  {
    double a11 = J[1][1], a12 = J[1][2], a21 = J[2][1], a22 = J[2][2];
    double inv_det = 1. / (a11 * a22 - a12 * a21);
    for (a = 0; a < 4; a++) for (b = 0; b < 4; b++) K[a][b] = (a == b) ? 1. : 0.;
    K[1][1] = a22 * inv_det;
    K[1][2] = -a12 * inv_det;
    K[2][1] = -a21 * inv_det;
    K[2][2] = a11 * inv_det;
  }
  for (p = 0; p < 4; p++) for (q = 0; q < 4; q++) {
    double s1 = 0., s2 = 0.;
    for (a = 0; a < 4; a++) for (b = 0; b < 4; b++) {
      s1 += gc[a][b] * J[a][p] * J[b][q];
      s2 += gn[a][b] * K[p][a] * K[q][b];
    }
    G->gcov[p][q] = s1;
    G->gcon[p][q] = s2;
  }
  /* det_dx_dxp_calc2_default(), src/coord.c:837-841 */
  G->g = g * (J[0][0] * J[3][3] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]));
  G->g_inv = 1. / G->g;
  G->ncon[0] = sqrt(-G->gcon[0][0]);
  G->alpha = 1. / G->ncon[0];
  G->ncon[1] = -G->alpha * G->gcon[0][1];
  G->ncon[2] = -G->alpha * G->gcon[0][2];
  G->ncon[3] = -G->alpha * G->gcon[0][3];
  G->beta[0] = -G->gcon[0][1] / G->gcon[0][0];
  G->beta[1] = -G->gcon[0][2] / G->gcon[0][0];
  G->beta[2] = -G->gcon[0][3] / G->gcon[0][0];
}

/* Forward-mode derivative numbers in (x1, x2): value and d/dx1, d/dx2. */
typedef struct { double v, d1, d2; } dn;
static dn dn_c(double v) { dn r = {v, 0., 0.}; return r; }
static dn dn_add(dn a, dn b) { dn r = {a.v + b.v, a.d1 + b.d1, a.d2 + b.d2}; return r; }
static dn dn_sub(dn a, dn b) { dn r = {a.v - b.v, a.d1 - b.d1, a.d2 - b.d2}; return r; }
static dn dn_mul(dn a, dn b) { dn r = {a.v * b.v, a.d1 * b.v + a.v * b.d1, a.d2 * b.v + a.v * b.d2}; return r; }
static dn dn_div(dn a, dn b) {
  double iv = 1. / b.v;
  dn r = {a.v * iv, (a.d1 * b.v - a.v * b.d1) * iv * iv, (a.d2 * b.v - a.v * b.d2) * iv * iv};
  return r;
}
static dn dn_scale(dn a, double s) { dn r = {a.v * s, a.d1 * s, a.d2 * s}; return r; }
static dn dn_sin(dn a) { double c = cos(a.v); dn r = {sin(a.v), c * a.d1, c * a.d2}; return r; }
static dn dn_cos(dn a) { double s = sin(a.v); dn r = {cos(a.v), -s * a.d1, -s * a.d2}; return r; }
static dn dn_exp(dn a) { double e = exp(a.v); dn r = {e, e * a.d1, e * a.d2}; return r; }
/* a^n for an integer n >= 1, derivative n a^(n-1) (defined at a = 0) */
static dn dn_powi(dn a, int n) {
  double f = pow(a.v, n), df = n * pow(a.v, n - 1);
  dn r = {f, df * a.d1, df * a.d2};
  return r;
}

/* Kerr-Schild metric in numerical coordinates, g'_ab = g_ab J_a J_b (the
   modified Kerr-Schild Jacobian is diagonal), as derivative numbers. */
static void ks_gcov_num_dn(const orc_t* o, double x1, double x2, dn g[4][4]) {
  const double a = o->c.a, asq = a * a;
  dn X1 = {x1, 1., 0.}, X2 = {x2, 0., 1.};
  dn r, th, J1, J2;
  int p, q;
  if (o->c.coords == 1) {
    dn e = dn_exp(X1);
    r = dn_add(dn_c(o->c.R0), e);
    J1 = e;
    dn w = dn_scale(X2, 2 * M_PI);
    th = dn_add(dn_add(dn_scale(X2, o->c.th_length), dn_c(o->c.th_beg)), dn_scale(dn_sin(w), o->c.h_slope));
    J2 = dn_add(dn_c(o->c.th_length), dn_scale(dn_cos(w), 2 * M_PI * o->c.h_slope));
  } else if (o->c.coords == 4) {
    /* COORD_DIAGONAL3 at interior zone centres (0 < x2 < 1: no reflection) */
    const int n = o->c.diag3_exponent;
    const double fc = o->c.diag3_factor;
    dn e = dn_exp(X1);
    dn f = dn_sub(dn_scale(X2, 2.), dn_c(1.));
    r = dn_add(dn_c(o->c.R0), e);
    J1 = e;
    th = dn_scale(dn_add(dn_add(dn_c(1.), dn_scale(f, o->c.h_slope)), dn_scale(dn_powi(f, n), fc)), 0.5 * M_PI);
    J2 = dn_scale(dn_add(dn_c(o->c.h_slope), dn_scale(dn_powi(f, n - 1), n * fc)), M_PI);
  } else {
    r = X1; th = X2; J1 = dn_c(1.); J2 = dn_c(1.);
  }
  dn c = dn_cos(th), s = dn_sin(th), s2 = dn_mul(s, s);
  dn rr = dn_mul(r, r), sig = dn_add(rr, dn_scale(dn_mul(c, c), asq));
  dn z = dn_div(dn_scale(r, 2.), sig);                 /* 2 M r / Sigma, M = 1 */
  dn gg[4][4];
  for (p = 0; p < 4; p++) for (q = 0; q < 4; q++) gg[p][q] = dn_c(0.);
  gg[0][0] = dn_sub(z, dn_c(1.));
  gg[0][1] = gg[1][0] = z;
  gg[0][3] = gg[3][0] = dn_scale(dn_mul(z, s2), -a);
  gg[1][1] = dn_add(dn_c(1.), z);
  gg[1][3] = gg[3][1] = dn_scale(dn_mul(gg[1][1], s2), -a);
  gg[2][2] = sig;
  gg[3][3] = dn_mul(s2, dn_add(rr, dn_scale(dn_add(dn_c(1.), dn_mul(z, s2)), asq)));
  dn Jd[4] = {dn_c(1.), J1, J2, dn_c(1.)};
  for (p = 0; p < 4; p++) for (q = 0; q < 4; q++) g[p][q] = dn_mul(dn_mul(gg[p][q], Jd[p]), Jd[q]);
}

/* Connection Gamma^d_{ab} = g^{dc} (d_a g_cb + d_b g_ca - d_c g_ab) / 2 at
   the centre of zone (i,j), conn[(d*4+a)*4+b] (the reference's conn[l][d][a][b],
   src/macros.h:576-587).  Replaces ks_conn_func_1(), src/metric.c:1259-1697. */
static void make_conn(const orc_t* o, int i, int j, double* conn) {
  double xp[4];
  dn g[4][4];
  double dg[4][4][4];   /* dg[c][a][b] = d_c g_ab */
  int a, b, c, d;
  coord_xp(o, i, j, NG, CENT, xp);
  ks_gcov_num_dn(o, xp[1], xp[2], g);
  for (a = 0; a < 4; a++) for (b = 0; b < 4; b++) {
    dg[0][a][b] = 0.;
    dg[1][a][b] = g[a][b].d1;
    dg[2][a][b] = g[a][b].d2;
    dg[3][a][b] = 0.;
  }
  const geom_t* G = GEOM(o, i, j, CENT);
  for (d = 0; d < 4; d++) for (a = 0; a < 4; a++) for (b = 0; b < 4; b++) {
    double sum = 0.;
    for (c = 0; c < 4; c++) sum += G->gcon[d][c] * 0.5 * (dg[a][c][b] + dg[b][c][a] - dg[c][a][b]);
    conn[(d * 4 + a) * 4 + b] = sum;
  }
}

/* ------------------------------------------------------------------------ */
/* per-zone physics: src/phys.c, src/metric.c:338-383                         */

static void lower(const double* v, const geom_t* g, double* w) {
  int m;
  for (m = 0; m < 4; m++) w[m] = g->gcov[m][0] * v[0] + g->gcov[m][1] * v[1] + g->gcov[m][2] * v[2] + g->gcov[m][3] * v[3];
}
static void raise(const double* v, const geom_t* g, double* w) {
  int m;
  for (m = 0; m < 4; m++) w[m] = g->gcon[m][0] * v[0] + g->gcon[m][1] * v[1] + g->gcon[m][2] * v[2] + g->gcon[m][3] * v[3];
}
static double DOT(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]; }

/* sum of g_ij v^i v^j in the order of gamma_calc(), src/phys.c:248-254 */
static double qsq_of(const geom_t* g, const double* pr) {
  return g->gcov[1][1] * pr[U1] * pr[U1] + g->gcov[2][2] * pr[U2] * pr[U2] + g->gcov[3][3] * pr[U3] * pr[U3]
       + 2 * (g->gcov[1][2] * pr[U1] * pr[U2] + g->gcov[1][3] * pr[U1] * pr[U3] + g->gcov[2][3] * pr[U2] * pr[U3]);
}

/* gamma_calc(), src/phys.c:244-268 */
static int gamma_calc(const double* pr, const geom_t* g, double* gamma) {
  double qsq = qsq_of(g, pr);
  if (qsq < 0.) {
    if (fabs(qsq) > 1.E-10) { *gamma = 1.; return 1; }
    qsq = 1.E-10;
  }
  *gamma = sqrt(1. + qsq);
  return 0;
}

static void fail(orc_t* o, int nf, int state);

/* ucon_calc(), src/phys.c:216-241 (records FAIL_GAMMA_CALC when o is given) */
static void ucon_calc(orc_t* o, const double* pr, const geom_t* g, double* ucon) {
  double gamma;
  if (gamma_calc(pr, g, &gamma) && o) fail(o, NF_GAMMA_CALC, 0);
  ucon[0] = gamma * g->ncon[0];
  ucon[1] = pr[U1] - ucon[0] * g->beta[0];
  ucon[2] = pr[U2] - ucon[0] * g->beta[1];
  ucon[3] = pr[U3] - ucon[0] * g->beta[2];
}

/* get_state(), src/phys.c:199-213, with bcon_calc(), :75-90 */
static void get_state(orc_t* o, const double* pr, const geom_t* g, double gam, state_t* q) {
  ucon_calc(o, pr, g, q->ucon);
  lower(q->ucon, g, q->ucov);
  double inv_ut = 1. / q->ucon[0];
  q->bcon[0] = pr[B1] * q->ucov[1] + pr[B2] * q->ucov[2] + pr[B3] * q->ucov[3];
  q->bcon[1] = (pr[B1] + q->bcon[0] * q->ucon[1]) * inv_ut;
  q->bcon[2] = (pr[B2] + q->bcon[0] * q->ucon[2]) * inv_ut;
  q->bcon[3] = (pr[B3] + q->bcon[0] * q->ucon[3]) * inv_ut;
  lower(q->bcon, g, q->bcov);
  q->bsq = DOT(q->bcon, q->bcov);
  q->p = (gam - 1.) * pr[UU];
  q->eta = q->p + pr[RHO] + pr[UU] + q->bsq;
  q->ptot = q->p + 0.5 * q->bsq;
}

/* mhd_calc(), src/phys.c:93-116 */
static void mhd_calc(const state_t* q, int dir, double* mhd) {
  double eu = q->eta * q->ucon[dir], bd = q->bcon[dir];
  int m;
  for (m = 0; m < 4; m++) mhd[m] = eu * q->ucov[m] - bd * q->bcov[m];
  mhd[dir] += q->ptot;
}

/* primtoflux(), src/phys.c:28-63 */
static void primtoflux(const double* pr, const state_t* q, int dir, const geom_t* g, double* f) {
  double mhd[4];
  int l;
  mhd_calc(q, dir, mhd);
  f[RHO] = pr[RHO] * q->ucon[dir];
  f[UU] = mhd[0] + f[RHO];
  f[U1] = mhd[1];
  f[U2] = mhd[2];
  f[U3] = mhd[3];
  f[B1] = q->bcon[1] * q->ucon[dir] - q->bcon[dir] * q->ucon[1];
  f[B2] = q->bcon[2] * q->ucon[dir] - q->bcon[dir] * q->ucon[2];
  f[B3] = q->bcon[3] * q->ucon[dir] - q->bcon[dir] * q->ucon[3];
  for (l = 0; l < NP; l++) f[l] *= g->g;
}

/* vchar_fast(), src/phys.c:432-501 */
static void vchar_fast(const state_t* q, const geom_t* g, int js, double gam, double* vmax, double* vmin) {
  double cms2 = (gam * q->p + q->bsq) / q->eta;
  if (cms2 < 0.) cms2 = SMALL;
  else if (cms2 > 1.) cms2 = 1.;
  double Au2 = q->ucon[js] * q->ucon[js], Bu2 = q->ucon[0] * q->ucon[0], AuBu = q->ucon[js] * q->ucon[0];
  double A = Bu2 - (g->gcon[0][0] + Bu2) * cms2;
  double B = AuBu - (g->gcon[0][js] + AuBu) * cms2;
  double C = Au2 - (g->gcon[js][js] + Au2) * cms2;
  double discr = B * B - A * C;
  if (discr < 0.) discr = 0.;
  discr = sqrt(discr);
  double denom = 1. / A;
  double vp = -(-B + discr) * denom, vm = -(-B - discr) * denom;
  if (vp > vm) { *vmax = vp; *vmin = vm; }
  else { *vmax = vm; *vmin = vp; }
}

/* check_gamma(), src/fixup.c:133-141 */
static int check_gamma(double gamma) {
  if (gamma <= GAMMAMAX) return 0;
  return (gamma > GAMMAMAX2) ? 2 : 1;
}

/* ------------------------------------------------------------------------ */
/* con-to-prim solvers.  The reference keeps each solver's state in
   file-scope globals (src/utoprim_2d_fast.c:81, src/utoprim_1d.c:55,
   src/utoprim_1d_ee.c:91-95, src/utoprim_1d_ee2.c:88-90); here one struct each. */

typedef struct { double Bsq, QdotBsq, Qtsq, Qdotn, D, half_Bsq, gmog; } s2d_t;

/* func_vsq() with eos_info(), src/utoprim_2d_fast.c:613-665, :728-740 */
static void s2d_func(const s2d_t* s, const double* x, double* dx, double* f, double* df) {
  double W = x[0], vsq = x[1];
  double ftmp = 1. - vsq, gtmp = sqrt(ftmp);
  double dPdW = s->gmog * ftmp;
  double dPdvsq = s->gmog * (0.5 * s->D / gtmp - W);
  double p_tmp = s->gmog * (W * ftmp - s->D * gtmp);
  double t2 = -s->half_Bsq + dPdvsq;
  double t3 = s->Bsq + W;
  double t4 = t3 * t3;
  double t23 = 1 / W;
  double t16 = s->QdotBsq * t23 * t23;
  double t11 = s->Qtsq - vsq * t4 + t16 * (s->Bsq + W + W);
  double t18 = -s->Qdotn - s->half_Bsq * (1.0 + vsq) + 0.5 * t16 - W + p_tmp;
  double t24 = t16 * t23;
  double t25 = -1.0 + dPdW - t24;
  double t35 = t25 * t3 + (s->Bsq - 2.0 * dPdvsq) * (t16 + vsq * W) * t23;
  double t21 = 1 / (t3 * t35);
  double t40;
  dx[0] = -(t2 * t11 + t4 * t18) * t21;
  t40 = -2 * (vsq + t24) * t3;
  dx[1] = -(-t25 * t11 + t40 * t18) * t21;
  *df = -t11 * t11 - t18 * t18;
  *f = -0.5 * (*df);
}

/* general_newton_raphson(), src/utoprim_2d_fast.c:489-592 */
static int s2d_newton(const s2d_t* s, double* x, int* iters) {
  double f = 1., df = 1., dx[2], x_old[2], errx = 1.;
  int n_iter = 0, i_extra = 0, doing_extra = 0, keep = 1;
  const double dv = (1. - 1.e-15);
  while (keep) {
    s2d_func(s, x, dx, &f, &df);
    x_old[0] = x[0]; x_old[1] = x[1];
    x[0] += dx[0];
    x[1] += dx[1];
    errx = (x[0] == 0.) ? fabs(dx[0]) : fabs(dx[0] / x[0]);
    if (x[0] < 0.) x[0] = fabs(x[0]);
    else if (x[0] > W_TOO_BIG) x[0] = x_old[0];
    if (x[1] < 0.) x[1] = 0.;
    else if (x[1] > 1.) x[1] = dv;
    if ((fabs(errx) <= NEWT_TOL) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) doing_extra = 1;
    if (doing_extra == 1) i_extra++;
    if (((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
        (n_iter >= (MAX_NEWT_ITER - 1))) keep = 0;
    n_iter++;
  }
  if (iters) *iters = n_iter;
  if (!isfinite(f) || !isfinite(df)) return 2;
  if (fabs(errx) <= NEWT_TOL) return 0;
  if ((fabs(errx) <= MIN_NEWT_TOL) && (fabs(errx) > NEWT_TOL)) return 0;
  return 1;
}

/* Utoprim_2d_fast() and Utoprim_new_body(), src/utoprim_2d_fast.c:135-408 */
static int utoprim_2d_fast(const double* U, const geom_t* geom, double gam, double* prim,
                           double* gamma_out, double* bsq, int* iters) {
  double Ut[NP], pt[NP];
  int l;
  if (iters) *iters = 0;
  if (U[0] <= 0.) return -100;
  double inv_gdet = geom->g_inv, alpha = geom->alpha, gf = alpha * inv_gdet;
  prim[B1] = U[B1] * inv_gdet;
  prim[B2] = U[B2] * inv_gdet;
  prim[B3] = U[B3] * inv_gdet;
  Ut[RHO] = gf * U[RHO];
  Ut[UU] = gf * (U[UU] - U[RHO]);
  for (l = U1; l <= B3; l++) Ut[l] = gf * U[l];
  for (l = 0; l < B1; l++) pt[l] = prim[l];

  s2d_t s;
  s.gmog = (gam - 1.) / gam;
  double Bcon[4] = {0., Ut[B1], Ut[B2], Ut[B3]}, Qcov[4] = {Ut[UU], Ut[U1], Ut[U2], Ut[U3]}, Qcon[4];
  s.Bsq = geom->gcov[1][1] * Bcon[1] * Bcon[1] + geom->gcov[2][2] * Bcon[2] * Bcon[2] + geom->gcov[3][3] * Bcon[3] * Bcon[3]
        + 2 * (geom->gcov[1][2] * Bcon[1] * Bcon[2] + geom->gcov[1][3] * Bcon[1] * Bcon[3] + geom->gcov[2][3] * Bcon[2] * Bcon[3]);
  raise(Qcov, geom, Qcon);
  double QdotB = Qcov[1] * Bcon[1] + Qcov[2] * Bcon[2] + Qcov[3] * Bcon[3];
  s.QdotBsq = QdotB * QdotB;
  s.Qdotn = -alpha * Qcon[0];
  s.Qtsq = DOT(Qcov, Qcon) + s.Qdotn * s.Qdotn;
  s.D = Ut[RHO];
  s.half_Bsq = 0.5 * s.Bsq;

  double utsq = qsq_of(geom, pt);
  if ((utsq < 0.) && (fabs(utsq) < 1.0e-13)) utsq = fabs(utsq);
  if (utsq < 0. || utsq > UTSQ_TOO_BIG) return 2;
  double gammasq = 1. + utsq, gamma = sqrt(gammasq);
  double rho0 = s.D / gamma, u = pt[UU], p = (gam - 1.) * u, w = rho0 + u + p;
  double W_last = w * gammasq;
  int i_increase = 0;
  while (((W_last * W_last * W_last * (W_last + 2. * s.Bsq) - s.QdotBsq * (2. * W_last + s.Bsq))
          <= W_last * W_last * (s.Qtsq - s.Bsq * s.Bsq)) && (i_increase < 10)) {
    W_last *= 10.;
    i_increase++;
  }
  double x[2];
  x[0] = fabs(W_last);
  {   /* x1_of_x0() with vsq_calc(), src/utoprim_2d_fast.c:418-452 */
    double Wsq = W_last * W_last, BW = s.Bsq + W_last, Xsq = BW * BW;
    double vsq = fabs((Wsq * s.Qtsq + s.QdotBsq * (BW + W_last)) / (Wsq * Xsq));
    x[1] = (vsq > 1.) ? (1.0 - 1.e-15) : vsq;
  }
  int ret = s2d_newton(&s, x, iters);
  double W = x[0], vsq = x[1];
  if ((ret != 0) || (W == FAIL_VAL)) return ret * 100 + 1;
  if (W <= 0. || W > W_TOO_BIG) return 3;
  if (vsq >= 1.) return 4;
  double gtmp = sqrt(1. - vsq);
  gamma = 1. / gtmp;
  rho0 = s.D * gtmp;
  *gamma_out = gamma;
  w = W * (1. - vsq);
  p = (gam - 1.) * (w - rho0) / gam;     /* pressure_rho0_w(), src/u2p_util.c:387-390 */
  u = w - (rho0 + p);
  double g_o_WBsq = gamma / (W + s.Bsq), QdB_o_W = QdotB / W;
  *bsq = s.Bsq * (1. - vsq) + QdB_o_W * QdB_o_W;
  prim[RHO] = rho0;
  prim[UU] = u;
  for (l = 1; l < 4; l++) prim[U1 + l - 1] = g_o_WBsq * (Qcon[l] + geom->ncon[l] * s.Qdotn + QdB_o_W * Bcon[l]);
  return 0;
}

typedef struct { double Bsq, QdotBsq, Qtsq, Qdotn, D, gam; } s1d_t;

/* vsq_calc(), dvsq_dW(), src/utoprim_1d.c:354-381 */
static double s1d_vsq(const s1d_t* s, double W) {
  double Wsq = W * W, Xsq = (s->Bsq + W) * (s->Bsq + W);
  return (Wsq * s->Qtsq + s->QdotBsq * (s->Bsq + 2. * W)) / (Wsq * Xsq);
}
static double s1d_dvsq(const s1d_t* s, double W) {
  double X = s->Bsq + W, W3 = W * W * W, X3 = X * X * X;
  return -2. * (s->Qtsq / X3 + s->QdotBsq * (3 * W * X + s->Bsq * s->Bsq) / (W3 * X3));
}
/* func_1d_orig() with the EOS functions, src/utoprim_1d.c:536-625 */
static void s1d_func(const s1d_t* s, const double* x, double* dx, double* f, double* df) {
  const double dv = 1.0e-15, G = s->gam;
  double W = x[0], Wsq = W * W;
  double vsq = s1d_vsq(s, W);
  vsq = (vsq < -dv) ? 0. : fabs(vsq);
  vsq = (vsq > 1.) ? (1. - dv) : vsq;
  double gtmp = 1. - vsq;
  double p_tmp = (G - 1.) * (W * gtmp - s->D * sqrt(gtmp)) / G;
  double dvsq = s1d_dvsq(s, W);
  double dp1 = (G - 1.) * (1. - vsq) / G;
  double dp2 = (G - 1.) * (0.5 * s->D / sqrt(1. - vsq) - W) / G;
  double dpdW = dp1 + dp2 * dvsq;
  double r0 = +W + 0.5 * s->Bsq * (1. + vsq) - 0.5 * s->QdotBsq / Wsq + s->Qdotn - p_tmp;
  double drdW = 1. - dpdW + s->QdotBsq / (Wsq * W) + 0.5 * s->Bsq * dvsq;
  dx[0] = -r0 / drdW;
  *f = 0.5 * r0 * r0;
  *df = -2. * (*f);
}
/* general_newton_raphson(), src/utoprim_1d.c:395-520 (with its W back-off) */
static int s1d_newton(const s1d_t* s, double* x) {
  double f = 1., df = 1., dx[1], errx = 1.;
  int n_iter = 0, i_extra = 0, doing_extra = 0, keep = 1;
  while (keep) {
    s1d_func(s, x, dx, &f, &df);
    x[0] += dx[0];
    int i_increase = 0;
    while (((x[0] * x[0] * x[0] * (x[0] + 2. * s->Bsq) - s->QdotBsq * (2. * x[0] + s->Bsq))
            <= x[0] * x[0] * (s->Qtsq - s->Bsq * s->Bsq)) && (i_increase < 10)) {
      x[0] -= (1. * i_increase) * dx[0] / 10.;
      i_increase++;
    }
    errx = (x[0] == 0.) ? fabs(dx[0]) : fabs(dx[0] / x[0]);
    x[0] = fabs(x[0]);
    if ((fabs(errx) <= NEWT_TOL) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) doing_extra = 1;
    if (doing_extra == 1) i_extra++;
    if (((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
        (n_iter >= (MAX_NEWT_ITER - 1))) keep = 0;
    n_iter++;
  }
  if (!isfinite(f) || !isfinite(df)) return 2;
  if (fabs(errx) > MIN_NEWT_TOL) return 1;
  return 0;
}

/* Utoprim_1d() and its Utoprim_new_body(), src/utoprim_1d.c:118-343 */
static int utoprim_1d(const double* U, const geom_t* geom, double gam, double* prim, double* gamma_out, double* bsq) {
  const double gdet = geom->g;
  double Ut[NP], pt[NP], Bcon[4], Bcov[4], Qcov[4], Qcon[4], ncov[4], ncon[4];
  int i, j;
  for (i = B1; i <= B3; i++) prim[i] = U[i] / gdet;
  double alpha = 1.0 / sqrt(-geom->gcon[0][0]);
  Ut[RHO] = alpha * U[RHO] / gdet;
  Ut[UU] = alpha * (U[UU] - U[RHO]) / gdet;
  for (i = U1; i <= U3; i++) Ut[i] = alpha * U[i] / gdet;
  for (i = B1; i <= B3; i++) Ut[i] = alpha * U[i] / gdet;
  for (i = 0; i < B1; i++) pt[i] = prim[i];

  s1d_t s;
  s.gam = gam;
  Bcon[0] = 0.;
  for (i = 1; i < 4; i++) Bcon[i] = Ut[B1 + i - 1];
  lower(Bcon, geom, Bcov);                    /* lower_g(), src/u2p_util.c:238-263 */
  for (i = 0; i < 4; i++) Qcov[i] = Ut[UU + i];
  raise(Qcov, geom, Qcon);                    /* raise_g(), src/u2p_util.c:195-218 */
  s.Bsq = 0.;
  for (i = 1; i < 4; i++) s.Bsq += Bcon[i] * Bcov[i];
  double QdotB = 0.;
  for (i = 0; i < 4; i++) QdotB += Qcov[i] * Bcon[i];
  s.QdotBsq = QdotB * QdotB;
  ncov[0] = -sqrt(-1. / geom->gcon[0][0]);     /* ncov_calc(), src/u2p_util.c:272-284 */
  ncov[1] = ncov[2] = ncov[3] = 0.;
  raise(ncov, geom, ncon);
  s.Qdotn = Qcon[0] * ncov[0];
  double Qsq = 0.;
  for (i = 0; i < 4; i++) Qsq += Qcov[i] * Qcon[i];
  s.Qtsq = Qsq + s.Qdotn * s.Qdotn;
  s.D = Ut[RHO];

  double utsq = 0.;
  for (i = 1; i < 4; i++) for (j = 1; j < 4; j++) utsq += geom->gcov[i][j] * pt[U1 + i - 1] * pt[U1 + j - 1];
  if ((utsq < 0.) && (fabs(utsq) < 1.0e-13)) utsq = fabs(utsq);
  if (utsq < 0. || utsq > UTSQ_TOO_BIG) return 2;
  double gammasq = 1. + utsq, gamma = sqrt(gammasq);
  double rho0 = s.D / gamma, u = pt[UU], p = (gam - 1.) * u, w = rho0 + u + p;
  double W_last = w * gammasq;
  int i_increase = 0;
  while (((W_last * W_last * W_last * (W_last + 2. * s.Bsq) - s.QdotBsq * (2. * W_last + s.Bsq))
          <= W_last * W_last * (s.Qtsq - s.Bsq * s.Bsq)) && (i_increase < 10)) {
    W_last *= 10.;
    i_increase++;
  }
  double x[1] = {W_last};
  int ret = s1d_newton(&s, x);
  double W = x[0];
  if ((ret != 0) || (W == FAIL_VAL)) return ret * 100 + 1;
  if (W <= 0. || W > W_TOO_BIG) return 3;
  double vsq = s1d_vsq(&s, W);
  if (vsq >= 1.) return 4;
  double gtmp = sqrt(1. - vsq);
  gamma = 1. / gtmp;
  rho0 = s.D * gtmp;
  *gamma_out = gamma;
  w = W * (1. - vsq);
  p = (gam - 1.) * (w - rho0) / gam;
  u = w - (rho0 + p);
  prim[RHO] = rho0;
  prim[UU] = u;
  *bsq = s.Bsq * (1. - vsq) + s.QdotBsq / (W * W);
  double Qtcon[4];
  for (i = 1; i < 4; i++) Qtcon[i] = Qcon[i] + ncon[i] * s.Qdotn;
  for (i = 1; i < 4; i++) prim[U1 + i - 1] = gamma / (W + s.Bsq) * (Qtcon[i] + QdotB * Bcon[i] / W);
  return 0;
}

/* Entropy-equation solvers.  State of src/utoprim_1d_ee.c:91-95. */
typedef struct {
  double Bsq, QdotBsq, Qtsq, Qdotn, D, Sc, gam, gm1, gm1_inv, g_o_gm1;
  double W_g, rho_g, W_g_old, rho_g_old;
  double t2, t4, t7, t24, t300, t400, s100, s200, two_Bsq;
} see_t;

/* func_rho(), src/utoprim_1d_ee.c:813-836 (f, df are left untouched) */
static void see_func_rho(see_t* s, const double* x, double* dx) {
  double rho = x[0], W = s->W_g;
  double t40 = pow(rho, s->gm1);
  double resid = (rho * W + (-t40 * s->s100 - s->D) * s->D);
  double t14 = t40 / rho;
  double jac = -t14 * s->s200 + W;
  dx[0] = -resid / jac;
}
/* gnr2(), src/utoprim_1d_ee.c:640-720: its "extra iterations" switch never turns on */
static int see_gnr2(see_t* s, double* x) {
  double f = 1., df = 1., dx[1], errx = 1.;
  int n_iter = 0, i_extra = 0, doing_extra = 0, keep = 1;
  while (keep) {
    see_func_rho(s, x, dx);
    x[0] += dx[0];
    errx = (x[0] == 0.) ? fabs(dx[0]) : fabs(dx[0] / x[0]);
    x[0] = fabs(x[0]);
    if ((fabs(errx) <= NEWT_TOL2) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) doing_extra = 0;
    if (doing_extra == 1) i_extra++;
    if (((fabs(errx) <= NEWT_TOL2) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
        (n_iter >= (MAX_NEWT_ITER - 1))) keep = 0;
    n_iter++;
  }
  if (!isfinite(f) || !isfinite(df) || !isfinite(x[0])) return 2;
  if (fabs(errx) <= NEWT_TOL) return 0;
  if ((fabs(errx) <= MIN_NEWT_TOL) && (fabs(errx) > NEWT_TOL)) return 0;
  return 1;
}
/* rho(W) by gnr2 with up to 10 restarts from 10x larger guesses (src/utoprim_1d_ee.c:762-767) */
static int see_solve_rho(see_t* s, double* x_rho) {
  double rho_g = x_rho[0];
  int ret, ntries = 0;
  while ((ret = see_gnr2(s, x_rho)) && (ntries++ < 10)) {
    rho_g *= 10.;
    x_rho[0] = rho_g;
  }
  return ret;
}
/* func_W(), src/utoprim_1d_ee.c:746-795 */
static void see_func_W(see_t* s, const double* x, double* dx) {
  double W = x[0], x_rho[1];
  s->W_g_old = s->W_g;
  s->W_g = W;
  x_rho[0] = s->rho_g;
  see_solve_rho(s, x_rho);
  s->rho_g_old = s->rho_g;
  double rho = s->rho_g = x_rho[0];
  double rho_gm1 = pow(rho, s->gm1);
  double drho_dW = -rho * rho / (-rho_gm1 * s->s200 + W * rho);
  double t15 = -(s->D - rho) * (s->D + rho);
  double t200 = W + s->two_Bsq;
  double t1000 = rho * drho_dW;
  double resid = (s->t300 + (s->t4 + s->t4 + (s->t400 + t15 * (s->t7 + (t200) * W)) * W) * W) * s->t24;
  double jac = 2 * (s->t4 + (s->t400 + t15 * s->t7 + (3.0 * t15 * s->Bsq + s->t7 * t1000 + (t15 + t15 + t1000 * (t200)) * W) * W) * W) * s->t24;
  dx[0] = -resid / jac;
}
/* general_newton_raphson(), src/utoprim_1d_ee.c:560-630 */
static int see_newton_W(see_t* s, double* x) {
  double f = 1., df = 1., dx[1], errx = 1.;
  int n_iter = 0, i_extra = 0, doing_extra = 0, keep = 1;
  while (keep) {
    see_func_W(s, x, dx);
    x[0] += dx[0];
    errx = (x[0] == 0.) ? fabs(dx[0]) : fabs(dx[0] / x[0]);
    x[0] = fabs(x[0]);
    if ((fabs(errx) <= NEWT_TOL) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) doing_extra = 1;
    if (doing_extra == 1) i_extra++;
    if (((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
        (n_iter >= (MAX_NEWT_ITER - 1))) keep = 0;
    n_iter++;
  }
  if (!isfinite(f) || !isfinite(df) || !isfinite(x[0])) return 2;
  if (fabs(errx) <= NEWT_TOL) return 0;
  if ((fabs(errx) <= MIN_NEWT_TOL) && (fabs(errx) > NEWT_TOL)) return 0;
  return 1;
}

/* Utoprim_1d_ee() and its Utoprim_new_body(), src/utoprim_1d_ee.c:156-470 */
static int utoprim_1d_ee(const double* U, double S, const geom_t* geom, double gam, double* prim, double* gamma_out) {
  double Ut[NP], pt[NP];
  int l;
  if (U[0] <= 0.) return -100;
  see_t s;
  s.gam = gam;
  s.gm1 = (gam - 1.);
  s.gm1_inv = 1. / s.gm1;
  s.g_o_gm1 = gam * s.gm1_inv;
  double inv_gdet = geom->g_inv, alpha = geom->alpha, gf = alpha * inv_gdet;
  prim[B1] = U[B1] * inv_gdet;
  prim[B2] = U[B2] * inv_gdet;
  prim[B3] = U[B3] * inv_gdet;
  s.Sc = gf * S;
  Ut[RHO] = gf * U[RHO];
  Ut[UU] = gf * (U[UU] - U[RHO]);
  for (l = U1; l <= B3; l++) Ut[l] = gf * U[l];
  for (l = 0; l < B1; l++) pt[l] = prim[l];

  double Bcon[4] = {0., Ut[B1], Ut[B2], Ut[B3]}, Qcov[4] = {Ut[UU], Ut[U1], Ut[U2], Ut[U3]}, Qcon[4];
  s.Bsq = geom->gcov[1][1] * Bcon[1] * Bcon[1] + geom->gcov[2][2] * Bcon[2] * Bcon[2] + geom->gcov[3][3] * Bcon[3] * Bcon[3]
        + 2 * (geom->gcov[1][2] * Bcon[1] * Bcon[2] + geom->gcov[1][3] * Bcon[1] * Bcon[3] + geom->gcov[2][3] * Bcon[2] * Bcon[3]);
  raise(Qcov, geom, Qcon);
  double QdotB = Qcov[1] * Bcon[1] + Qcov[2] * Bcon[2] + Qcov[3] * Bcon[3];
  s.QdotBsq = QdotB * QdotB;
  s.Qdotn = -alpha * Qcon[0];
  s.Qtsq = DOT(Qcov, Qcon) + s.Qdotn * s.Qdotn;
  s.D = Ut[RHO];
  s.t2 = s.D * s.D;
  s.t4 = s.QdotBsq * s.t2;
  s.t7 = s.Bsq * s.Bsq;
  s.t24 = 1 / s.t2;
  s.two_Bsq = s.Bsq + s.Bsq;
  s.t300 = s.QdotBsq * s.Bsq * s.t2;
  s.t400 = s.Qtsq * s.t2;
  s.s200 = s.D * gam * s.Sc;
  s.s100 = s.g_o_gm1 * s.Sc;

  double utsq = qsq_of(geom, pt);
  if ((utsq < 0.) && (fabs(utsq) < 1.0e-13)) utsq = fabs(utsq);
  if (utsq < 0. || utsq > UTSQ_TOO_BIG) return 2;
  double gammasq = 1. + utsq, gamma = sqrt(gammasq);
  double rho0 = s.D / gamma;
  double rho_gm1 = pow(rho0, s.gm1);
  double p = s.Sc * rho_gm1 / gamma, u = p * s.gm1_inv, w = rho0 + u + p;
  double W_last = w * gammasq;
  int i_increase = 0;
  while (((W_last * W_last * W_last * (W_last + 2. * s.Bsq) - s.QdotBsq * (2. * W_last + s.Bsq))
          <= W_last * W_last * (s.Qtsq - s.Bsq * s.Bsq)) && (i_increase < 10)) {
    W_last *= 10.;
    i_increase++;
  }
  s.W_g = s.W_g_old = W_last;
  s.rho_g = s.rho_g_old = rho0;
  double x[1] = {W_last};
  int ret = see_newton_W(&s, x);
  double W = x[0];
  if ((ret != 0) || (W == FAIL_VAL)) return ret * 100 + 1;
  if (W <= 0. || W > W_TOO_BIG) return 3;
  double x_rho[1] = {s.rho_g};
  ret = see_solve_rho(&s, x_rho);
  s.rho_g = x_rho[0];
  if (ret != 0) return 10;
  rho0 = s.rho_g;
  rho_gm1 = pow(rho0, s.gm1);
  utsq = (s.D - rho0) * (s.D + rho0) / (rho0 * rho0);
  double gamma_sq = 1. + utsq;
  gamma = sqrt(gamma_sq);
  *gamma_out = gamma;
  if (utsq < 0.) return 4;
  p = s.Sc * rho_gm1 / gamma;
  u = p * s.gm1_inv;
  prim[RHO] = rho0;
  prim[UU] = u;
  double g_o_WBsq = gamma / (W + s.Bsq), QdB_o_W = QdotB / W;
  for (l = 1; l < 4; l++) prim[U1 + l - 1] = g_o_WBsq * (Qcon[l] + geom->ncon[l] * s.Qdotn + QdB_o_W * Bcon[l]);
  return 0;
}

/* Utoprim_1d_ee2() and its Utoprim_new_body(), src/utoprim_1d_ee2.c:152-423,
   with func_rho(), :545-586 and general_newton_raphson(), :440-530 */
typedef struct { double Bsq, QdotBsq, Qtsq, D, Sc, gam, g_o_gm1; } see2_t;
static void see2_func(const see2_t* s, const double* x, double* dx, double* f, double* df) {
  const double G = s->gam;
  double rho = x[0], rhosq = rho * rho;
  double t200 = 1. / (s->D * s->D);
  double t100 = s->g_o_gm1 * s->Sc * pow(rho, (G - 1.));
  double W = s->D * (s->D + t100) / rho;
  double dWdrho = s->D * (-s->D + t100 * (G - 2.)) / rhosq;
  double t1 = W * W, t2 = s->Bsq + W, t3 = t2 * t2;
  double vsq = (s->D - rho) * (s->D + rho) * t200;
  double dvsqdrho = -2 * rho * t200;
  double resid = t1 * (s->Qtsq - vsq * t3) + s->QdotBsq * (t2 + W);
  double t12 = s->Bsq * s->Bsq, t17 = dWdrho * vsq;
  double jac = 2 * s->QdotBsq * dWdrho
    + ((s->Qtsq - vsq * t12) * 2 * dWdrho + (-6 * t17 * s->Bsq - dvsqdrho * t12
        + (-2 * dvsqdrho * s->Bsq - 4 * t17 - dvsqdrho * W) * W) * W) * W;
  dx[0] = -resid / jac;
  *f = 0.5 * resid * resid;
  *df = -2. * (*f);
}
static int see2_newton(const see2_t* s, double* x) {
  double f = 1., df = 1., dx[1], errx = 1.;
  int n_iter = 0, i_extra = 0, doing_extra = 0, keep = 1;
  while (keep) {
    see2_func(s, x, dx, &f, &df);
    x[0] += dx[0];
    errx = (x[0] == 0.) ? fabs(dx[0]) : fabs(dx[0] / x[0]);
    x[0] = fabs(x[0]);
    if ((fabs(errx) <= NEWT_TOL) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) doing_extra = 1;
    if (doing_extra == 1) i_extra++;
    if (((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
        (n_iter >= (MAX_NEWT_ITER - 1))) keep = 0;
    n_iter++;
  }
  if (!isfinite(f) || !isfinite(df) || !isfinite(x[0])) return 2;
  if (fabs(errx) > MIN_NEWT_TOL) return 1;
  return 0;
}
static int utoprim_1d_ee2(const double* U, double S, const geom_t* geom, double gam, double* prim, double* gamma_out) {
  double Ut[NP], pt[NP], Bcon[4], Bcov[4], Qcov[4], Qcon[4], ncov[4], ncon[4];
  int l;
  if (U[0] <= 0.) return -100;
  see2_t s;
  s.gam = gam;
  s.g_o_gm1 = gam / (gam - 1.);
  double inv_gdet = 1. / geom->g;
  prim[B1] = U[B1] * inv_gdet;
  prim[B2] = U[B2] * inv_gdet;
  prim[B3] = U[B3] * inv_gdet;
  double alpha = geom->alpha, gf = alpha * inv_gdet;
  s.Sc = gf * S;
  Ut[RHO] = gf * U[RHO];
  Ut[UU] = gf * (U[UU] - U[RHO]);
  for (l = U1; l <= B3; l++) Ut[l] = gf * U[l];
  for (l = 0; l < B1; l++) pt[l] = prim[l];

  Bcon[0] = 0.; Bcon[1] = Ut[B1]; Bcon[2] = Ut[B2]; Bcon[3] = Ut[B3];
  lower(Bcon, geom, Bcov);
  Qcov[0] = Ut[UU]; Qcov[1] = Ut[U1]; Qcov[2] = Ut[U2]; Qcov[3] = Ut[U3];
  raise(Qcov, geom, Qcon);
  s.Bsq = Bcon[1] * Bcov[1] + Bcon[2] * Bcov[2] + Bcon[3] * Bcov[3];
  double QdotB = Qcov[1] * Bcon[1] + Qcov[2] * Bcon[2] + Qcov[3] * Bcon[3];
  s.QdotBsq = QdotB * QdotB;
  ncov[0] = -geom->alpha;
  ncov[1] = ncov[2] = ncov[3] = 0.;
  for (l = 0; l < 4; l++) ncon[l] = -geom->alpha * geom->gcon[0][l];
  double Qdotn = Qcon[0] * ncov[0];
  s.Qtsq = DOT(Qcov, Qcon) + Qdotn * Qdotn;
  s.D = Ut[RHO];

  double utsq = qsq_of(geom, pt);
  if ((utsq < 0.) && (fabs(utsq) < 1.0e-13)) utsq = fabs(utsq);
  if (utsq < 0. || utsq > UTSQ_TOO_BIG) return 2;
  double gammasq = 1. + utsq, gamma = sqrt(gammasq);
  double rho0 = s.D / gamma;
  double x[1] = {rho0};
  int ret = see2_newton(&s, x);
  rho0 = x[0];
  if ((ret != 0) || (rho0 == FAIL_VAL)) return ret * 100 + 1;
  if (rho0 > W_TOO_BIG) return 3;
  utsq = (s.D - rho0) * (s.D + rho0) / (rho0 * rho0);
  double gamma_sq = 1. + utsq;
  gamma = sqrt(gamma_sq);
  *gamma_out = gamma;
  if (utsq < 0.) return 4;
  double rho_gm1 = pow(rho0, (gam - 1.));
  double p = s.Sc * rho_gm1 / gamma, u = p / (gam - 1.), w = rho0 + u + p;
  double W = w * gamma_sq;
  prim[RHO] = rho0;
  prim[UU] = u;
  double g_o_WBsq = gamma / (W + s.Bsq), QdB_o_W = QdotB / W;
  for (l = 1; l < 4; l++) prim[U1 + l - 1] = g_o_WBsq * (Qcon[l] + ncon[l] * Qdotn + QdB_o_W * Bcon[l]);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* fail() bookkeeping, src/diag.c:159-492, check_boundary_pflag(),
   src/harm_mpi.c:1388-1430 (one domain: every boundary is physical)         */
static void fail(orc_t* o, int nf, int state) {
  static const int pflag_response[3] = {0, PF_INTERP_PRIM, PF_INTERP_V};
  int i = o->ic, j = o->jc, k = o->kc;
  if (state && (nf == NF_FLOOR || nf == NF_TMAX || nf == NF_GAMMA_MAX || nf == NF_GAMMA_CALC ||
                nf == NF_UTOPRIM || nf == NF_UTOPRIM_EE)) {
    int fi = i, fj = j, fk = k;
    if (nf == NF_GAMMA_CALC) {
      if (o->pc == FACE1) fi--;
      else if (o->pc == FACE2) fj--;
      else if (o->pc == FACE3) fk--;
    }
    o->failure_exists = 1;
    o->failure_type[pflag_response[state]] = 1;
    if (fi < NG + NG || fi > o->n[0] + NG - 1 - NG || fj < NG + NG || fj > o->n[1] + NG - 1 - NG ||
        fk < NG + NG || fk > o->n[2] + NG - 1 - NG)
      o->boundary_phys_pflag = 1;
    o->pflag[IX(o, fi, fj, fk)] = pflag_response[state];
  }
  o->nfail[(long)nf * o->nt + IX(o, i, j, k)]++;
  o->nfailtot[nf]++;
}
static void set_cur(orc_t* o, int i, int j, int k, int pos) { o->ic = i; o->jc = j; o->kc = k; o->pc = pos; }

/* ------------------------------------------------------------------------ */
/* reconstruction, src/recon.c                                                */

/* linear_recon(), MC limiter, src/recon.c:88-132 */
static double linear_recon(double y1, double y2, double y3) {
  double Dqm = 2. * (y2 - y1), Dqp = 2. * (y3 - y2), Dqc = 0.5 * (y3 - y1), s = Dqm * Dqp;
  if (s <= 0.) return 0.;
  double aDqm = fabs(Dqm), aDqp = fabs(Dqp), aDqc = fabs(Dqc);
  if ((aDqm <= aDqp) && (aDqm <= aDqc)) return Dqm;
  else if (aDqp <= aDqc) return Dqp;
  return Dqc;
}

/* linear_recon_vect(), MC form, src/recon.c:237-264: note the first
   difference is always y[1]-y[0] (callers start at ibeg = 1) */
static void linear_recon_vect(const double* y, double* Dq, int ibeg, int iend) {
  double a, b = y[1] - y[0];
  int i;
  for (i = ibeg; i <= iend; i++) {
    a = b;
    b = y[i + 1] - y[i];
    double s = a * b, Dqi;
    if (s <= 0.) Dqi = 0.;
    else {
      double a_abs = fabs(a), b_abs = fabs(b), a3 = a_abs + a_abs + a_abs, b3 = b_abs + b_abs + b_abs;
      if (a3 <= b_abs) Dqi = a + a;
      else if (b3 <= a_abs) Dqi = b + b;
      else Dqi = 0.5 * (a + b);
    }
    Dq[i] = Dqi;
  }
}

/* para_vect(), src/recon.c:611-641: yL[i], yR[i] are the left and right edge
   values of zone i (callers start at ibeg = 2) */
static void para_vect(const double* y, const double* dq, double* yL, double* yR, int ibeg, int iend) {
  double r = 0.5 * (y[2] + y[1]) - PPM_FACTOR * (dq[2] - dq[1]), yip = y[2], l, yi;
  int i;
  for (i = ibeg; i <= iend; i++) {
    yi = yip;
    yip = y[i + 1];
    l = r;
    r = 0.5 * (yip + yi) - PPM_FACTOR * (dq[i + 1] - dq[i]);
    double qa = (r - yi) * (yi - l), lnew, rnew;
    if (qa <= 0.) { lnew = rnew = yi; }
    else {
      double r2 = r + r, l2 = l + l, y3 = yi + yi + yi;
      lnew = l; rnew = r;
      if (r > l) {
        if (y3 > (l + r2)) lnew = y3 - r2;
        if (y3 < (l2 + r)) rnew = y3 - l2;
      } else {
        if (y3 < (l + r2)) lnew = y3 - r2;
        if (y3 > (l2 + r)) rnew = y3 - l2;
      }
    }
    yL[i] = lnew;
    yR[i] = rnew;
  }
}

/* rescale_prims_r(), src/recon.c:1120-1237 (tables built at creation) */
static void rescale_prims_r(orc_t* o, double* ph, int type) {
  const int n1t = o->nt3[0];
  int i, j, k, l;
  for (i = 0; i < n1t; i++) {
    const double* rs = type ? o->irscale : o->rscale;
    double f0 = rs[0 * n1t + i], f1 = rs[1 * n1t + i], f2 = rs[2 * n1t + i];
    for (j = 0; j < o->nt3[1]; j++) for (k = 0; k < o->nt3[2]; k++) {
      double* q = P8(ph, o, i, j, k);
      q[RHO] *= f0;
      q[UU] *= f2;
      for (l = U1; l <= B3; l++) q[l] *= f1;
    }
    if (type) {
      double f3 = rs[3 * n1t + i], f4 = rs[4 * n1t + i], f5 = rs[5 * n1t + i];
      int side;
      for (side = 0; side < 2; side++)
        for (j = 0; j < o->nt3[1]; j++) for (k = 0; k < o->nt3[2]; k++) {
          double* q = F8(side ? o->pR : o->pL, o, i, j, k, 0);
          q[RHO] *= f3;
          q[UU] *= f5;
          for (l = U1; l <= B3; l++) q[l] *= f4;
        }
    }
  }
}

/* reconstruct_fast(), src/recon.c:894-1103 */
static void reconstruct_fast(orc_t* o, double* ph) {
  const int BUF = o->buf;
  int RS[3], RE[3], d, i, j, k, l;
  for (d = 0; d < 3; d++) {
    if (o->n[d] == 1) { RS[d] = RE[d] = NG; }
    else { RS[d] = NG - BUF; RE[d] = NG + o->n[d] - 1 + BUF; }
  }
  if (o->c.recon != 5) {
    for (d = 0; d < 3; d++) {
      if (!o->do_recon[d]) continue;
      int im = (d == 0), jm = (d == 1), km = (d == 2);
      if (o->c.rescale_r && d == 0) rescale_prims_r(o, ph, 0);
      for (i = RS[0] + (BUF - 1) * im; i <= RE[0] - (BUF - 1) * im; i++)
        for (j = RS[1] + (BUF - 1) * jm; j <= RE[1] - (BUF - 1) * jm; j++)
          for (k = RS[2] + (BUF - 1) * km; k <= RE[2] - (BUF - 1) * km; k++)
            for (l = 0; l < NP; l++) {
              double dq = 0.5 * linear_recon(P8(ph, o, i - im, j - jm, k - km)[l], P8(ph, o, i, j, k)[l],
                                             P8(ph, o, i + im, j + jm, k + km)[l]);
              F8(o->pL, o, i + im, j + jm, k + km, d)[l] = P8(ph, o, i, j, k)[l] + dq;
              F8(o->pR, o, i, j, k, d)[l] = P8(ph, o, i, j, k)[l] - dq;
            }
      if (o->c.rescale_r && d == 0) rescale_prims_r(o, ph, 1);
    }
  } else {
    if (o->do_recon[0]) {
      const int N1S = NG, N1E = NG + o->n[0] - 1;
      if (o->c.rescale_r) rescale_prims_r(o, ph, 0);
      for (j = RS[1]; j <= RE[1]; j++) for (k = RS[2]; k <= RE[2]; k++) for (l = 0; l < NP; l++) {
        for (i = 0; i < o->nt3[0]; i++) o->pvect[i] = P8(ph, o, i, j, k)[l];
        linear_recon_vect(o->pvect, o->dqvect, N1S - 2, N1E + 2);
        para_vect(o->pvect, o->dqvect, o->pLvect, o->pRvect, N1S - 1, N1E + 1);
        for (i = RS[0] + BUF - 1; i <= RE[0] - BUF + 1; i++) {
          F8(o->pL, o, i + 1, j, k, 0)[l] = o->pRvect[i];
          F8(o->pR, o, i, j, k, 0)[l] = o->pLvect[i];
        }
      }
      if (o->c.rescale_r) rescale_prims_r(o, ph, 1);
    }
    if (o->do_recon[1]) {
      const int N2S = NG, N2E = NG + o->n[1] - 1;
      for (i = RS[0]; i <= RE[0]; i++) for (k = RS[2]; k <= RE[2]; k++) for (l = 0; l < NP; l++) {
        for (j = 0; j < o->nt3[1]; j++) o->pvect[j] = P8(ph, o, i, j, k)[l];
        linear_recon_vect(o->pvect, o->dqvect, N2S - 2, N2E + 2);
        para_vect(o->pvect, o->dqvect, o->pLvect, o->pRvect, N2S - 1, N2E + 1);
        for (j = RS[1] + BUF - 1; j <= RE[1] - BUF + 1; j++) {
          F8(o->pL, o, i, j + 1, k, 1)[l] = o->pRvect[j];
          F8(o->pR, o, i, j, k, 1)[l] = o->pLvect[j];
        }
      }
    }
    if (o->do_recon[2]) {
      const int N3S = NG, N3E = NG + o->n[2] - 1;
      for (i = RS[0]; i <= RE[0]; i++) for (j = RS[1]; j <= RE[1]; j++) for (l = 0; l < NP; l++) {
        for (k = 0; k < o->nt3[2]; k++) o->pvect[k] = P8(ph, o, i, j, k)[l];
        linear_recon_vect(o->pvect, o->dqvect, N3S - 2, N3E + 2);
        para_vect(o->pvect, o->dqvect, o->pLvect, o->pRvect, N3S - 1, N3E + 1);
        for (k = RS[2] + BUF - 1; k <= RE[2] - BUF + 1; k++) {
          F8(o->pL, o, i, j, k + 1, 2)[l] = o->pRvect[k];
          F8(o->pR, o, i, j, k, 2)[l] = o->pLvect[k];
        }
      }
    }
  }
  /* degenerate directions: copy L/R states into the ghosts, src/recon.c:1050-1075 */
  if (o->degenerate) {
    int g, dd;
    if (!o->do_recon[0])
      for (j = 0; j < o->nt3[1]; j++) for (k = 0; k < o->nt3[2]; k++) {
        for (l = 0; l < NP; l++) F8(o->pL, o, NG, j, k, 0)[l] = F8(o->pR, o, NG, j, k, 0)[l] = P8(ph, o, NG, j, k)[l];
        for (g = 1; g <= BUF; g++) for (dd = 0; dd < 3; dd++) for (l = 0; l < NP; l++)
          F8(o->pL, o, NG - g, j, k, dd)[l] = F8(o->pL, o, NG + g, j, k, dd)[l] = F8(o->pL, o, NG, j, k, dd)[l];
        for (g = 1; g <= BUF; g++) for (dd = 0; dd < 3; dd++) for (l = 0; l < NP; l++)
          F8(o->pR, o, NG - g, j, k, dd)[l] = F8(o->pR, o, NG + g, j, k, dd)[l] = F8(o->pR, o, NG, j, k, dd)[l];
      }
    if (!o->do_recon[1])
      for (i = 0; i < o->nt3[0]; i++) for (k = 0; k < o->nt3[2]; k++) {
        for (l = 0; l < NP; l++) F8(o->pL, o, i, NG, k, 1)[l] = F8(o->pR, o, i, NG, k, 1)[l] = P8(ph, o, i, NG, k)[l];
        for (g = 1; g <= BUF; g++) for (dd = 0; dd < 3; dd++) for (l = 0; l < NP; l++)
          F8(o->pL, o, i, NG - g, k, dd)[l] = F8(o->pL, o, i, NG + g, k, dd)[l] = F8(o->pL, o, i, NG, k, dd)[l];
        for (g = 1; g <= BUF; g++) for (dd = 0; dd < 3; dd++) for (l = 0; l < NP; l++)
          F8(o->pR, o, i, NG - g, k, dd)[l] = F8(o->pR, o, i, NG + g, k, dd)[l] = F8(o->pR, o, i, NG, k, dd)[l];
      }
    if (!o->do_recon[2])
      for (i = 0; i < o->nt3[0]; i++) for (j = 0; j < o->nt3[1]; j++) {
        for (l = 0; l < NP; l++) F8(o->pL, o, i, j, NG, 2)[l] = F8(o->pR, o, i, j, NG, 2)[l] = P8(ph, o, i, j, NG)[l];
        for (g = 1; g <= BUF; g++) for (dd = 0; dd < 3; dd++) for (l = 0; l < NP; l++)
          F8(o->pL, o, i, j, NG - g, dd)[l] = F8(o->pL, o, i, j, NG + g, dd)[l] = F8(o->pL, o, i, j, NG, dd)[l];
        for (g = 1; g <= BUF; g++) for (dd = 0; dd < 3; dd++) for (l = 0; l < NP; l++)
          F8(o->pR, o, i, j, NG - g, dd)[l] = F8(o->pR, o, i, j, NG + g, dd)[l] = F8(o->pR, o, i, j, NG, dd)[l];
      }
  }
}

/* ------------------------------------------------------------------------ */
/* fluxes, src/step_ch.c:636-790; fix_flux(), src/bounds.c:2231-2290          */

static void hll_or_lf(int flux, double cmax, double cmin, double ctop, const double* Fl, const double* Fr,
                      const double* Ul, const double* Ur, double* F) {
  int l;
  if (flux == 1) for (l = 0; l < NP; l++) F[l] = (cmax * Fl[l] + cmin * Fr[l] - cmax * cmin * (Ur[l] - Ul[l])) / (cmax + cmin + SMALL);
  else for (l = 0; l < NP; l++) F[l] = 0.5 * (Fl[l] + Fr[l] - ctop * (Ur[l] - Ul[l]));
}

static void numerical_flux(orc_t* o) {
  int d, i, j, k;
  const int N1S = NG, N2S = NG, N3S = NG, N1E = NG + o->n[0] - 1, N2E = NG + o->n[1] - 1, N3E = NG + o->n[2] - 1;
  for (d = 0; d < 4; d++) o->vmax[d] = 0.;
  for (d = 0; d < 3; d++) {
    int id = (d == 0), jd = (d == 1), kd = (d == 2), is, js, ks, ie, je, ke;
    if (o->c.ct_para) {
      is = N1S + NG * (id - 1); js = N2S + NG * (jd - 1); ks = N3S + NG * (kd - 1);
      ie = N1E + 1 + (NG - 1) * (1 - id); je = N2E + 1 + (NG - 1) * (1 - jd); ke = N3E + 1 + (NG - 1) * (1 - kd);
    } else {
      is = N1S + id - 1; js = N2S + jd - 1; ks = N3S + kd - 1;
      ie = N1E + 1; je = N2E + 1; ke = N3E + 1;
    }
    const int dim = d + 1;
    for (i = is; i <= ie; i++) for (j = js; j <= je; j++) for (k = ks; k <= ke; k++) {
      const geom_t* g = GEOM(o, i, j, d);
      set_cur(o, i, j, k, d);
      const double* pl = F8(o->pL, o, i, j, k, d);
      const double* pr = F8(o->pR, o, i, j, k, d);
      state_t ql, qr;
      double Ul[NP], Ur[NP], Fl[NP], Fr[NP], cmaxl, cminl, cmaxr, cminr;
      get_state(o, pl, g, o->gam, &ql);
      get_state(o, pr, g, o->gam, &qr);
      primtoflux(pl, &ql, 0, g, Ul);
      primtoflux(pr, &qr, 0, g, Ur);
      primtoflux(pl, &ql, dim, g, Fl);
      primtoflux(pr, &qr, dim, g, Fr);
      vchar_fast(&ql, g, dim, o->gam, &cmaxl, &cminl);
      vchar_fast(&qr, g, dim, o->gam, &cmaxr, &cminr);
      double cmax = fabs(dmax(dmax(0., cmaxl), cmaxr));
      double cmin = fabs(dmax(dmax(0., -cminl), -cminr));
      double ctop = dmax(cmin, cmax);
      if (ctop > o->vmax[dim] && i >= N1S && i <= N1E && j >= N2S && j <= N2E && k >= N3S && k <= N3E)
        o->vmax[dim] = ctop;
      double* c = C2(o->cgf, o, i, j, k, d);
      c[0] = cmin; c[1] = cmax;
      c = C2(o->uconL, o, i, j, k, d);
      c[0] = g->g * ql.ucon[0]; c[1] = g->g * ql.ucon[dim];
      c = C2(o->uconR, o, i, j, k, d);
      c[0] = g->g * qr.ucon[0]; c[1] = g->g * qr.ucon[dim];
      hll_or_lf(o->c.flux, cmax, cmin, ctop, Fl, Fr, Ul, Ur, F8(o->F, o, i, j, k, d));
    }
  }
}

static void fix_flux(orc_t* o) {
  int j, k;
  if (o->c.bc != 10) return;
  for (j = 0; j < o->nt3[1]; j++) for (k = 0; k < o->nt3[2]; k++) {
    double* f = F8(o->F, o, NG, j, k, 0);
    if (f[RHO] > 0) f[RHO] = 0.;
    f = F8(o->F, o, NG + o->n[0], j, k, 0);
    if (f[RHO] < 0) f[RHO] = 0.;
  }
}

/* rebuild the B fluxes from the corner EMFs with weight cf (0.5 Toth,
   0.125 PPM EMFs), src/flux_ct.c:88-122 and :853-933 */
static void flux_from_emf(orc_t* o, double cf) {
  const int N1S = NG, N2S = NG, N3S = NG, N1E = NG + o->n[0] - 1, N2E = NG + o->n[1] - 1, N3E = NG + o->n[2] - 1;
  int i, j, k, d;
  for (i = N1S; i <= N1E + 1; i++) for (j = N2S; j <= N2E + 1; j++) for (k = N3S; k <= N3E + 1; k++)
    for (d = 0; d < 3; d++) { double* f = F8(o->F, o, i, j, k, d); f[B1] = f[B2] = f[B3] = 0.; }
  for (i = N1S; i <= N1E + 1; i++) for (j = N2S; j <= N2E; j++) for (k = N3S; k <= N3E; k++) {
    F8(o->F, o, i, j, k, 0)[B2] += cf * (*EMF(o, 2, i, j, k) + *EMF(o, 2, i, j + 1, k));
    F8(o->F, o, i, j, k, 0)[B3] += -cf * (*EMF(o, 1, i, j, k) + *EMF(o, 1, i, j, k + 1));
  }
  for (i = N1S; i <= N1E; i++) for (j = N2S; j <= N2E + 1; j++) for (k = N3S; k <= N3E; k++) {
    F8(o->F, o, i, j, k, 1)[B1] += -cf * (*EMF(o, 2, i, j, k) + *EMF(o, 2, i + 1, j, k));
    F8(o->F, o, i, j, k, 1)[B3] += cf * (*EMF(o, 0, i, j, k) + *EMF(o, 0, i, j, k + 1));
  }
  for (i = N1S; i <= N1E; i++) for (j = N2S; j <= N2E; j++) for (k = N3S; k <= N3E + 1; k++) {
    F8(o->F, o, i, j, k, 2)[B1] += cf * (*EMF(o, 1, i, j, k) + *EMF(o, 1, i + 1, j, k));
    F8(o->F, o, i, j, k, 2)[B2] += -cf * (*EMF(o, 0, i, j, k) + *EMF(o, 0, i, j + 1, k));
  }
}

/* flux_ct(), src/flux_ct.c:26-128 (Toth average) */
static void flux_ct(orc_t* o) {
  const int N1S = NG, N2S = NG, N3S = NG, N1E = NG + o->n[0] - 1, N2E = NG + o->n[1] - 1, N3E = NG + o->n[2] - 1;
  int i, j, k;
  for (i = N1S; i <= N1E + 1; i++) for (j = N2S; j <= N2E + 1; j++) for (k = N3S; k <= N3E + 1; k++) {
    *EMF(o, 0, i, j, k) = 0.25 * (F8(o->F, o, i, j, k, 1)[B3] + F8(o->F, o, i, j, k - 1, 1)[B3]
                                - F8(o->F, o, i, j, k, 2)[B2] - F8(o->F, o, i, j - 1, k, 2)[B2]);
    *EMF(o, 1, i, j, k) = 0.25 * (F8(o->F, o, i, j, k, 2)[B1] + F8(o->F, o, i - 1, j, k, 2)[B1]
                                - F8(o->F, o, i, j, k, 0)[B3] - F8(o->F, o, i, j, k - 1, 0)[B3]);
    *EMF(o, 2, i, j, k) = 0.25 * (F8(o->F, o, i, j, k, 0)[B2] + F8(o->F, o, i, j - 1, k, 0)[B2]
                                - F8(o->F, o, i, j, k, 1)[B1] - F8(o->F, o, i - 1, j, k, 1)[B1]);
  }
  flux_from_emf(o, 0.5);
}

/* one PPM pencil of a face flux component, accumulated with sign sg into the
   EMF e at both edges of every zone along dim (flux_ct_para_fast2(), src/flux_ct.c:780-846) */
static void emf_pencil(orc_t* o, int e, int fd, int var, int dim, int i, int j, int k, double sg) {
  int m, nn = o->nt3[dim], s = NG, en = NG + o->n[dim] - 1;
  for (m = 0; m < nn; m++) {
    int ii = (dim == 0) ? m : i, jj = (dim == 1) ? m : j, kk = (dim == 2) ? m : k;
    o->pvect[m] = F8(o->F, o, ii, jj, kk, fd)[var];
  }
  linear_recon_vect(o->pvect, o->dqvect, s - 2, en + 2);
  para_vect(o->pvect, o->dqvect, o->pLvect, o->pRvect, s - 1, en + 1);
  for (m = s - 1; m <= en + 1; m++) {
    int ii = (dim == 0) ? m : i, jj = (dim == 1) ? m : j, kk = (dim == 2) ? m : k;
    *EMF(o, e, ii, jj, kk) += sg * o->pLvect[m];
    ii += (dim == 0); jj += (dim == 1); kk += (dim == 2);
    *EMF(o, e, ii, jj, kk) += sg * o->pRvect[m];
  }
}

static void flux_ct_para_fast2(orc_t* o) {
  const int N1S = NG, N2S = NG, N3S = NG, N1E = NG + o->n[0] - 1, N2E = NG + o->n[1] - 1, N3E = NG + o->n[2] - 1;
  int i, j, k;
  memset(o->emf, 0, sizeof(double) * 3 * o->nt);
  /* emf0: + F1[B3] along x3, - F2[B2] along x2 */
  for (i = N1S; i <= N1E + 1; i++) for (j = N2S; j <= N2E + 1; j++) emf_pencil(o, 0, 1, B3, 2, i, j, 0, 1.);
  for (i = N1S; i <= N1E + 1; i++) for (k = N3S; k <= N3E + 1; k++) emf_pencil(o, 0, 2, B2, 1, i, 0, k, -1.);
  /* emf1: + F2[B1] along x1, - F0[B3] along x3 */
  for (j = N2S; j <= N2E + 1; j++) for (k = N3S; k <= N3E + 1; k++) emf_pencil(o, 1, 2, B1, 0, 0, j, k, 1.);
  for (i = N1S; i <= N1E + 1; i++) for (j = N2S; j <= N2E + 1; j++) emf_pencil(o, 1, 0, B3, 2, i, j, 0, -1.);
  /* emf2: + F0[B2] along x2, - F1[B1] along x1 */
  for (i = N1S; i <= N1E + 1; i++) for (k = N3S; k <= N3E + 1; k++) emf_pencil(o, 2, 0, B2, 1, i, 0, k, 1.);
  for (j = N2S; j <= N2E + 1; j++) for (k = N3S; k <= N3E + 1; k++) emf_pencil(o, 2, 1, B1, 0, 0, j, k, -1.);
  flux_from_emf(o, 0.125);
}

/* ------------------------------------------------------------------------ */
/* source terms, src/phys.c:121-187, :541-582, :649-784                       */
static void source(orc_t* o, const double* ph, const geom_t* g, int ii, int jj, int kk, double* dU) {
  int l, a, b;
  state_t q;
  for (l = 0; l < NP; l++) dU[l] = 0.;
  if (o->metric_dim > 0) {
    const double* cn = o->conn + ((long)(ii - NG) * o->n[1] + (jj - NG)) * 64;
    double mhd[4][4];
    get_state(o, ph, g, o->gam, &q);
    for (a = 0; a < 4; a++) mhd_calc(&q, a, mhd[a]);
    for (a = 0; a < 4; a++) for (b = 0; b < 4; b++) {
      dU[U1] += mhd[a][b] * cn[(b * 4 + 1) * 4 + a];
      dU[U2] += mhd[a][b] * cn[(b * 4 + 2) * 4 + a];
    }
  }
  if (o->c.cooling) {
    if (o->metric_dim == 0) get_state(o, ph, g, o->gam, &q);
    int bound = ((q.ucov[0] * (q.p + ph[UU] + ph[RHO] + q.bsq)) > (-ph[RHO])) ? 1 : 0;
    double f_cool = 0.;
    if (bound) {   /* cooling_func_hr_disk(): s_norm = 1, q_exp = 0.5 */
      double E_ratio = (o->gam - 1.) * ph[UU] / (ph[RHO] * o->tr[ii]);
      double dE = E_ratio - 1.;
      f_cool = 1. * ph[UU] * o->omega[ii] * pow((dE + fabs(dE)), 0.5);
    }
    dU[UU] -= f_cool * q.ucov[0];
    dU[U1] -= f_cool * q.ucov[1];
    dU[U2] -= f_cool * q.ucov[2];
    dU[U3] -= f_cool * q.ucov[3];
  }
  dU[UU] *= g->g;
  dU[U1] *= g->g;
  dU[U2] *= g->g;
  dU[U3] *= g->g;
}

/* ------------------------------------------------------------------------ */
/* fixups, src/fixup.c                                                        */

static void fixup_gamma(orc_t* o, double* pv, double* gamma) {    /* :148-158 */
  double f = sqrt((GAMMAMAX - 1.) * (GAMMAMAX + 1.) / ((*gamma - 1.) * (*gamma + 1.)));
  pv[U1] *= f;
  pv[U2] *= f;
  pv[U3] *= f;
  *gamma = GAMMAMAX;
  fail(o, NF_GAMMA_MAX, 0);
}

static void fixup_floor(orc_t* o, double* pv, int i, int j, int k) {   /* :110-121 */
  long c = IX(o, i, j, k);
  int ret = 0;
  if (pv[RHO] < o->rhoflr[c]) { pv[RHO] = o->rhoflr[c]; ret = 1; }
  if (pv[UU] < o->uuflr[c]) { pv[UU] = o->uuflr[c]; ret = 1; }
  if (ret) fail(o, NF_FLOOR, 0);
}

/* entropy flux of face (i,j,k,d) from the stored face data, src/fixup.c:363-398 */
static double entropy_face_flux(orc_t* o, int i, int j, int k, int d) {
  const double* c = C2(o->cgf, o, i, j, k, d);
  double cmin = c[0], cmax = c[1], ctop = dmax(cmin, cmax);
  const double* pl = F8(o->pL, o, i, j, k, d);
  const double* pr = F8(o->pR, o, i, j, k, d);
  const double* ul = C2(o->uconL, o, i, j, k, d);
  const double* ur = C2(o->uconR, o, i, j, k, d);
  double s = (o->gam - 1.) * pl[UU] * pow(pl[RHO], 1. - o->gam);
  double Ul = s * ul[0], Fl = s * ul[1];
  s = (o->gam - 1.) * pr[UU] * pow(pr[RHO], 1. - o->gam);
  double Ur = s * ur[0], Fr = s * ur[1];
  if (o->c.flux == 1) return (cmax * Fl + cmin * Fr - cmax * cmin * (Ur - Ul)) / (cmax + cmin + SMALL);
  return 0.5 * (Fl + Fr - ctop * (Ur - Ul));
}

/* fixup_entropy_eq(), src/fixup.c:335-454 */
static int fixup_entropy_eq(orc_t* o, int i, int j, int k, const double* prim_old, double* prim,
                            const geom_t* g, double* gamma_out) {
  double backup[NP], gamma, Fm[3], Fp[3];
  int l, d, ret;
  for (l = 0; l < NP; l++) backup[l] = prim[l];
  double Dt = (o->n_substep == 1) ? o->dt : 0.5 * o->dt;
  gamma_calc(prim_old, g, &gamma);
  double S_old = g->g * (o->gam - 1.) * prim_old[UU] * gamma / (g->alpha * pow(prim_old[RHO], (o->gam - 1.)));
  for (d = 0; d < 3; d++) {
    Fm[d] = entropy_face_flux(o, i, j, k, d);
    Fp[d] = entropy_face_flux(o, i + (d == 0), j + (d == 1), k + (d == 2), d);
  }
  double S = S_old + Dt * (-(Fp[0] - Fm[0]) * o->invdx[1] - (Fp[1] - Fm[1]) * o->invdx[2] - (Fp[2] - Fm[2]) * o->invdx[3]);
  const double* U = P8(o->Ugf + (long)o->n_U * o->nt * NP, o, i, j, k);
  ret = utoprim_1d_ee(U, S, g, o->gam, prim, &gamma);
  if ((!ret) && (check_gamma(gamma) == 2)) {
    for (l = 0; l < NP; l++) prim[l] = backup[l];
    ret = 1;
    gamma = 1.;
  }
  if (ret) {
    ret = utoprim_1d_ee2(U, S, g, o->gam, prim, &gamma);
    if ((!ret) && (check_gamma(gamma) == 2)) {
      for (l = 0; l < NP; l++) prim[l] = backup[l];
      ret = 1;
    }
    if (ret) gamma = 1.;
  }
  *gamma_out = gamma;
  if (!ret) fail(o, NF_USE_EE, 0);
  return ret;
}

/* recover_primitives(), src/step_ch.c:801-1024 (FIXUP_TREE 3, entropy fix on) */
static void recover_primitives(orc_t* o, int i, int j, int k, double* pf, const double* U, const geom_t* g) {
  double pg[NP], pf2[NP], gamma = 1., bsq = 0.;
  int l, ret, use_ee = 0;
  double* pgam = &o->pgamma[IX(o, i, j, k)];
  for (l = 0; l < NP; l++) pg[l] = pf[l];
  ret = utoprim_2d_fast(U, g, o->gam, pf, &gamma, &bsq, NULL);
  if ((!ret) && (check_gamma(gamma) == 2)) {
    for (l = 0; l < B1; l++) pf[l] = pg[l];
    ret = 1;
    gamma = 1.;
  }
  if (ret) {
    ret = utoprim_1d(U, g, o->gam, pf, &gamma, &bsq);
    if ((!ret) && (check_gamma(gamma) == 2)) {
      for (l = 0; l < B1; l++) pf[l] = pg[l];
      ret = 1;
    }
    if (ret) gamma = 1.;
  }
  *pgam = gamma;
  if (ret || (pf[UU] < BETA_MIN * bsq) || (pf[RHO] < BETA_MIN * bsq)) use_ee = 1;
  if (!use_ee) {
    if (!ret) fail(o, NF_USE_FULL_INV, 0);
    else fail(o, NF_UTOPRIM, 1);
  } else {
    if (!ret) { for (l = 0; l < NP; l++) pf2[l] = pf[l]; }
    else fail(o, NF_UTOPRIM, 0);
    for (l = 0; l < B1; l++) pf[l] = pg[l];
    int ret_ee = fixup_entropy_eq(o, i, j, k, P8(o->p_old, o, i, j, k), pf, g, &gamma);
    if (ret_ee) {
      if (ret) {
        fail(o, NF_UTOPRIM_EE, 1);
        return;
      }
      fail(o, NF_UTOPRIM_EE, 0);
      for (l = 0; l < NP; l++) pf[l] = pf2[l];
      gamma = *pgam;
      fail(o, NF_USE_FULL_INV, 0);
    }
  }
  *pgam = gamma;
  fixup_floor(o, pf, i, j, k);
  if (check_gamma(gamma)) {
    fail(o, NF_GAMMA_MAX, 2);
    return;
  }
  *pgam = gamma;
}

/* fixup_interp_prim() (which = PF_INTERP_PRIM, src/fixup.c:499-610) and
   fixup_interp_v() (which = PF_INTERP_V, :659-761) */
static void fixup_interp(orc_t* o, double* pv, int which) {
  static const int ngood_min[] = {13, 9, 6, 4};
  const int o13 = 13;
  int i, j, k, ii, jj, kk, l;
  if (!o->failure_exists) return;
  for (i = NG; i < NG + o->n[0]; i++) for (j = NG; j < NG + o->n[1]; j++) for (k = NG; k < NG + o->n[2]; k++) {
    if (o->pflag[IX(o, i, j, k)] != which) continue;
    set_cur(o, i, j, k, CENT);
    int pf[27], ngood = 0, ngood2, i_ret;
    double pold[NP], gamma, f, gamma_old;
    double* p = P8(pv, o, i, j, k);
    double* pgam = &o->pgamma[IX(o, i, j, k)];
    const int lbeg = (which == PF_INTERP_PRIM) ? RHO : U1;
    for (l = 0; l < NP; l++) pold[l] = p[l];
    for (ii = -1; ii <= 1; ii++) for (jj = -1; jj <= 1; jj++) for (kk = -1; kk <= 1; kk++)
      pf[o13 + ii + 3 * (jj + 3 * kk)] = !o->pflag[IX(o, i + ii, j + jj, k + kk)];
    for (ii = 1; ii <= 13; ii++) {
      if (pf[o13 + ii] && pf[o13 - ii]) ngood++;
      else pf[o13 + ii] = pf[o13 - ii] = 0;
    }
    const geom_t* g = GEOM(o, i, j, CENT);
    if (ngood < ngood_min[o->nsd]) {
      if (which == PF_INTERP_PRIM) {
        gamma_calc(p, g, pgam);
        fail(o, NF_INTERP_PRIM, 0);
      } else {
        fixup_gamma(o, p, pgam);
        fail(o, NF_INTERP_V, 0);
      }
      continue;
    }
    for (l = lbeg; l <= U3; l++) p[l] = 0.;
    for (ii = -1; ii <= 1; ii++) for (jj = -1; jj <= 1; jj++) for (kk = -1; kk <= 1; kk++)
      if (pf[o13 + ii + 3 * (jj + 3 * kk)]) {
        const double* q = P8(pv, o, i + ii, j + jj, k + kk);
        for (l = lbeg; l <= U3; l++) p[l] += q[l];
      }
    ngood *= 2;
    for (l = lbeg; l <= U3; l++) p[l] /= ngood;
    i_ret = gamma_calc(p, g, &gamma);
    if (check_gamma(gamma)) {
      gamma_old = gamma;
      gamma = 0.;
      ngood2 = 0;
      for (ii = -1; ii <= 1; ii++) for (jj = -1; jj <= 1; jj++) for (kk = -1; kk <= 1; kk++)
        if (pf[o13 + ii + 3 * (jj + 3 * kk)]) {
          i_ret = gamma_calc(P8(pv, o, i + ii, j + jj, k + kk), GEOM(o, i + ii, j + jj, CENT), &f);
          if (!i_ret) { gamma += f; ngood2++; }
        }
      gamma /= ngood2;
      f = sqrt((gamma - 1.) * (gamma + 1.) / ((gamma_old - 1.) * (gamma_old + 1.)));
      p[U1] *= f;
      p[U2] *= f;
      p[U3] *= f;
      i_ret = gamma_calc(p, g, &gamma);
    }
    if (which == PF_INTERP_PRIM) {
      if (i_ret) {
        for (l = U1; l <= U3; l++) p[l] = pold[l];
        gamma_calc(p, g, &gamma);
        fail(o, NF_INTERP_PRIM, 0);
      }
      fail(o, NF_USE_INTERP_PRIM, 0);
      if ((!i_ret) && check_gamma(gamma)) fixup_gamma(o, p, &gamma);
      *pgam = gamma;
    } else {
      if (i_ret || check_gamma(gamma)) {
        for (l = U1; l <= U3; l++) p[l] = pold[l];
        fixup_gamma(o, p, pgam);
        fail(o, NF_INTERP_V, 0);
      } else {
        fail(o, NF_USE_INTERP_V, 0);
        *pgam = gamma;
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* boundaries, src/bounds.c:31-579, inflow_check(), :2304-2380                */
static void inflow_check(orc_t* o, double* pr, int ii, int jj, int kk, int type) {
  const int dir = 1;
  const geom_t* g = GEOM(o, ii, jj, CENT);
  double ucon[4], gamma, vsq;
  int j, k;
  set_cur(o, ii, jj, kk, CENT);
  ucon_calc(o, pr, g, ucon);
  if (((ucon[dir] > 0.) && (type == 0)) || ((ucon[dir] < 0.) && (type == 1))) {
    gamma = g->alpha * ucon[0];
    pr[U1] /= gamma;
    pr[U2] /= gamma;
    pr[U3] /= gamma;
    pr[U1 + dir - 1] = g->beta[dir - 1] * g->ncon[0];
    vsq = 0.;
    for (j = 1; j < 4; j++) for (k = 1; k < 4; k++) vsq += g->gcov[j][k] * pr[U1 + j - 1] * pr[U1 + k - 1];
    if (fabs(vsq) < 1.e-13) vsq = fabs(vsq);
    if (vsq < 0.) vsq = 0.;
    if (vsq >= 1.) vsq = 1. - 1. / (GAMMAMAX * GAMMAMAX);
    gamma = 1. / sqrt(1. - vsq);
    pr[U1] *= gamma;
    pr[U2] *= gamma;
    pr[U3] *= gamma;
  }
}

static void copy8(double* dst, const double* src) { memcpy(dst, src, sizeof(double) * NP); }

static void bounds(orc_t* o, double* pa, int skip_mpi) {
  const int N1S = NG, N2S = NG, N3S = NG, N1E = NG + o->n[0] - 1, N2E = NG + o->n[1] - 1, N3E = NG + o->n[2] - 1;
  const int fe = o->failure_exists;
  int i, j, k, g;
  int* pf = o->pflag;
  if (skip_mpi && !o->boundary_phys_pflag) return;
  if (o->c.bc == 0) {          /* BC_CARTESIAN_OUTFLOW, :58-133 */
    for (g = 0; g < NG; g++) for (j = N2S; j <= N2E; j++) for (k = N3S; k <= N3E; k++) {
      copy8(P8(pa, o, N1S - 1 - g, j, k), P8(pa, o, N1S, j, k));
      if (fe) pf[IX(o, N1S - 1 - g, j, k)] = pf[IX(o, N1S, j, k)];
    }
    for (g = 0; g < NG; g++) for (j = N2S; j <= N2E; j++) for (k = N3S; k <= N3E; k++) {
      copy8(P8(pa, o, N1E + 1 + g, j, k), P8(pa, o, N1E, j, k));
      if (fe) pf[IX(o, N1E + 1 + g, j, k)] = pf[IX(o, N1E, j, k)];
    }
    for (i = 0; i < o->nt3[0]; i++) for (g = 0; g < NG; g++) for (k = N3S; k <= N3E; k++) {
      copy8(P8(pa, o, i, N2S - 1 - g, k), P8(pa, o, i, N2S, k));
      if (fe) pf[IX(o, i, N2S - 1 - g, k)] = pf[IX(o, i, N2S, k)];
    }
    for (i = 0; i < o->nt3[0]; i++) for (g = 0; g < NG; g++) for (k = N3S; k <= N3E; k++) {
      copy8(P8(pa, o, i, N2E + 1 + g, k), P8(pa, o, i, N2E, k));
      if (fe) pf[IX(o, i, N2E + 1 + g, k)] = pf[IX(o, i, N2E, k)];
    }
    for (i = 0; i < o->nt3[0]; i++) for (j = 0; j < o->nt3[1]; j++) for (g = 0; g < NG; g++) {
      copy8(P8(pa, o, i, j, N3S - 1 - g), P8(pa, o, i, j, N3S));
      if (fe) pf[IX(o, i, j, N3S - 1 - g)] = pf[IX(o, i, j, N3S)];
    }
    for (i = 0; i < o->nt3[0]; i++) for (j = 0; j < o->nt3[1]; j++) for (g = 0; g < NG; g++) {
      copy8(P8(pa, o, i, j, N3E + 1 + g), P8(pa, o, i, j, N3E));
      if (fe) pf[IX(o, i, j, N3E + 1 + g)] = pf[IX(o, i, j, N3E)];
    }
  } else {                     /* BC_SPHERICAL_OUTFLOW3, :492-579 */
    for (g = 0; g < NG; g++) for (j = N2S; j <= N2E; j++) for (k = N3S; k <= N3E; k++)
      copy8(P8(pa, o, N1S - 1 - g, j, k), P8(pa, o, N1S, j, k));
    for (g = 0; g < NG; g++) for (j = N2S; j <= N2E; j++) for (k = N3S; k <= N3E; k++)
      inflow_check(o, P8(pa, o, N1S - 1 - g, j, k), N1S - 1 - g, j, k, 0);
    if (fe) for (g = 0; g < NG; g++) for (j = N2S; j <= N2E; j++) for (k = N3S; k <= N3E; k++)
      pf[IX(o, N1S - 1 - g, j, k)] = pf[IX(o, N1S, j, k)];
    for (g = 0; g < NG; g++) for (j = N2S; j <= N2E; j++) for (k = N3S; k <= N3E; k++)
      copy8(P8(pa, o, N1E + 1 + g, j, k), P8(pa, o, N1E, j, k));
    for (g = 0; g < NG; g++) for (j = N2S; j <= N2E; j++) for (k = N3S; k <= N3E; k++)
      inflow_check(o, P8(pa, o, N1E + 1 + g, j, k), N1E + 1 + g, j, k, 1);
    if (fe) for (g = 0; g < NG; g++) for (j = N2S; j <= N2E; j++) for (k = N3S; k <= N3E; k++)
      pf[IX(o, N1E + 1 + g, j, k)] = pf[IX(o, N1E, j, k)];
    for (i = 0; i < o->nt3[0]; i++) for (g = 0; g < NG; g++) for (k = N3S; k <= N3E; k++) {
      double* d = P8(pa, o, i, N2S - 1 - g, k);
      copy8(d, P8(pa, o, i, N2S + g, k));
      d[U2] *= -1;
      d[B2] *= -1;
      if (fe) pf[IX(o, i, N2S - 1 - g, k)] = pf[IX(o, i, N2S + g, k)];
    }
    for (i = 0; i < o->nt3[0]; i++) for (g = 0; g < NG; g++) for (k = N3S; k <= N3E; k++) {
      double* d = P8(pa, o, i, N2E + 1 + g, k);
      copy8(d, P8(pa, o, i, N2E - g, k));
      d[U2] *= -1;
      d[B2] *= -1;
      if (fe) pf[IX(o, i, N2E + 1 + g, k)] = pf[IX(o, i, N2E - g, k)];
    }
    for (i = 0; i < o->nt3[0]; i++) for (j = 0; j < o->nt3[1]; j++) for (g = 0; g < NG; g++) {
      copy8(P8(pa, o, i, j, N3S - 1 - g), P8(pa, o, i, j, N3E - g));
      if (fe) pf[IX(o, i, j, N3S - 1 - g)] = pf[IX(o, i, j, N3E - g)];
    }
    for (i = 0; i < o->nt3[0]; i++) for (j = 0; j < o->nt3[1]; j++) for (g = 0; g < NG; g++) {
      copy8(P8(pa, o, i, j, N3E + 1 + g), P8(pa, o, i, j, N3S + g));
      if (fe) pf[IX(o, i, j, N3E + 1 + g)] = pf[IX(o, i, j, N3S + g)];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* advance(), src/step_ch.c:248-451                                           */
static void advance(orc_t* o, double* p_i, double* p_h, double Dt, double* p_f) {
  int i, j, k, l;
  reconstruct_fast(o, p_h);
  numerical_flux(o);
  fix_flux(o);
  if (o->c.ct_para) flux_ct_para_fast2(o);
  else flux_ct(o);
  double* U0 = o->Ugf;
  double* Un = o->Ugf + (long)o->n_U * o->nt * NP;
  for (i = NG; i < NG + o->n[0]; i++) for (j = NG; j < NG + o->n[1]; j++) for (k = NG; k < NG + o->n[2]; k++) {
    const geom_t* g = GEOM(o, i, j, CENT);
    double U[NP], dU[NP];
    set_cur(o, i, j, k, CENT);
    if (o->n_substep == 0) {
      state_t q;
      get_state(o, P8(p_i, o, i, j, k), g, o->gam, &q);
      primtoflux(P8(p_i, o, i, j, k), &q, 0, g, U);
      copy8(P8(U0, o, i, j, k), U);
    } else {
      copy8(U, P8(U0, o, i, j, k));
    }
    source(o, P8(p_h, o, i, j, k), g, i, j, k, dU);
    for (l = 0; l < NP; l++)
      U[l] += Dt * (-(F8(o->F, o, i + 1, j, k, 0)[l] - F8(o->F, o, i, j, k, 0)[l]) * o->invdx[1]
                    - (F8(o->F, o, i, j + 1, k, 1)[l] - F8(o->F, o, i, j, k, 1)[l]) * o->invdx[2]
                    - (F8(o->F, o, i, j, k + 1, 2)[l] - F8(o->F, o, i, j, k, 2)[l]) * o->invdx[3]
                    + dU[l]);
    copy8(P8(Un, o, i, j, k), U);
    recover_primitives(o, i, j, k, P8(p_f, o, i, j, k), U, g);
  }
}

static void clear_failures(orc_t* o) {
  if (o->failure_exists) {
    o->failure_exists = 0;
    o->failure_type[0] = o->failure_type[1] = o->failure_type[2] = 0;
    memset(o->pflag, 0, sizeof(int) * o->nt);
  }
}

static void bounds_and_fixups(orc_t* o, double* pa) {
  bounds(o, pa, 0);
  if (o->failure_exists) {
    if (o->failure_type[PF_INTERP_V]) fixup_interp(o, pa, PF_INTERP_V);
    if (o->failure_type[PF_INTERP_PRIM]) fixup_interp(o, pa, PF_INTERP_PRIM);
  }
  if (o->boundary_phys_pflag) {
    bounds(o, pa, 1);
    o->boundary_phys_pflag = 0;
  }
}

/* timestep(), src/step_ch.c:546-623 (TIMESTEP_ORIGINAL) */
static double timestep(orc_t* o) {
  double c1 = o->vmax[1], c2 = o->vmax[2], c3 = o->vmax[3];
  double dtsqij = sqrt((1. / 3.) / ((c1 * c1) / (o->dx[1] * o->dx[1]) + (c2 * c2) / (o->dx[2] * o->dx[2]) +
                                    (c3 * c3) / (o->dx[3] * o->dx[3])));
  double dtt = o->c.cour * dtsqij;
  if (dtt > SAFE * o->dt) dtt = SAFE * o->dt;
  if (o->t + dtt > o->c.t_final) dtt = o->c.t_final - o->t;
  return dtt;
}

/* step_ch(), src/step_ch.c:50-234 */
static void step_ch(orc_t* o) {
  clear_failures(o);
  o->n_substep = 0;
  o->n_U = 1;
  memcpy(o->p_old, o->p, sizeof(double) * NP * o->nt);
  memcpy(o->ph, o->p, sizeof(double) * NP * o->nt);
  advance(o, o->p, o->p, 0.5 * o->dt, o->ph);
  bounds_and_fixups(o, o->ph);
  clear_failures(o);
  o->n_substep = 1;
  o->n_U = 0;
  advance(o, o->p, o->ph, o->dt, o->p);
  bounds_and_fixups(o, o->p);
  o->t_old = o->t;
  o->t += o->dt;
  o->dt_old = o->dt;
  o->dt = timestep(o);
}

/* ------------------------------------------------------------------------ */
/* set-up                                                                     */

/* omega_circular_equatorial(), target_temperature(), src/phys.c:805-907,
   vel_circular_equatorial(), :950-972; ISCO references at coord_of_r(r_isco),
   src/coord.c:3866-3905 (xp_at_eq(), :390-410), with the Kerr-Schild metric
   in x coordinates (get_special_geometry(), src/metric.c:4277-4310). */
static void ks_special(const orc_t* o, double r, geom_t* G) {
  double xp[4] = {0., 0., 0.5, 0.}, x[4], J[4][4];
  xp[1] = (o->c.coords != 0) ? log(r - o->c.R0) : r;
  x_of_xp(o, xp, x, J);
  ks_gcov(o, x[1], x[2], G->gcov);
  ks_gcon(o, x[1], x[2], G->gcon);
}
static void vel_circ(const orc_t* o, double r, double* v) {
  double a = o->c.a, r3h = pow(r, 1.5), uphi = 2 * a * r3h + r * r * (r - 3.);
  if (uphi <= 0.) {
    r = o->c.r_isco;
    r3h = pow(r, 1.5);
    uphi = 2 * a * r3h + r * r * (r - 3.);
  }
  uphi = 1. / sqrt(uphi);
  v[1] = v[2] = 0.;
  v[3] = uphi;
  v[0] = uphi * (r3h + a);
}
static void cooling_tables(orc_t* o) {
  const double a = o->c.a, M = 1., r_isco = o->c.r_isco, hor = o->c.h_o_r;
  double vcov_isco[4] = {0, 0, 0, 0}, ut_sq_isco, uphi_sq_isco;
  int i;
  if (fabs(r_isco) > 0.) {
    geom_t G;
    double v[4];
    ks_special(o, r_isco, &G);
    vel_circ(o, r_isco, v);
    lower(v, &G, vcov_isco);
  }
  {
    double r = r_isco, rm1 = 1. / r, x = sqrt(r), xm3 = 1. / (x * x * x), t1 = a * a * rm1;
    double C_f = 1. - 3 * rm1 + 2 * a * xm3, F_f = 1. - 2 * a * xm3 + t1 * rm1, G_f = 1. - 2 * rm1 + a * xm3;
    ut_sq_isco = G_f * G_f / C_f;
    uphi_sq_isco = F_f * F_f * r / C_f;
  }
  for (i = 0; i < o->nt3[0]; i++) {
    double xp[4], x[4], J[4][4];
    coord_xp(o, i, NG, NG, CENT, xp);
    x_of_xp(o, xp, x, J);
    double r = x[1];
    /* orbital frequency */
    if (r >= r_isco) o->omega[i] = sqrt(M) / (pow(r, 1.5) + a);
    else {
      geom_t G;
      double v[4];
      ks_special(o, r, &G);
      raise(vcov_isco, &G, v);
      o->omega[i] = v[3] / v[0];
    }
    /* target temperature T/(gam-1) */
    double rm1 = 1. / r, xs = sqrt(r), xm3 = 1. / (xs * xs * xs), t1 = a * a * rm1, R_z;
    if (r > r_isco) {
      double C_f = 1. - 3 * rm1 + 2 * a * xm3, F_f = 1. - 2 * a * xm3 + t1 * rm1, G_f = 1. - 2 * rm1 + a * xm3;
      double ut_sq = G_f * G_f / C_f, uphi_sq = F_f * F_f * r / C_f;
      R_z = uphi_sq * rm1 - t1 * (ut_sq - 1.);
    } else {
      R_z = uphi_sq_isco * rm1 - t1 * (ut_sq_isco - 1.);
    }
    o->tr[i] = 0.5 * M_PI * R_z * hor * hor * rm1;
  }
}

orc_t* orc_create(const orc_config* cfg) {
  if (!cfg || cfg->n1 < 1 || cfg->n2 < 1 || cfg->n3 < 1) return NULL;
  if (cfg->metric != 0 && cfg->metric != 2) return NULL;
  if (cfg->coords != 0 && cfg->coords != 1 && cfg->coords != 4) return NULL;
  if (cfg->coords == 4 && (cfg->diag3_exponent < 3 || cfg->diag3_exponent % 2 == 0)) return NULL;
  if (cfg->recon != 2 && cfg->recon != 5) return NULL;
  if (cfg->bc != 0 && cfg->bc != 10) return NULL;
  if (cfg->cooling && cfg->metric != 2) return NULL;
  orc_t* o = (orc_t*)calloc(1, sizeof(orc_t));
  int d, i, j, k;
  o->c = *cfg;
  o->n[0] = cfg->n1; o->n[1] = cfg->n2; o->n[2] = cfg->n3;
  for (d = 0; d < 3; d++) { o->nt3[d] = o->n[d] + 2 * NG; o->do_recon[d] = (o->n[d] != 1); }
  o->degenerate = !(o->do_recon[0] && o->do_recon[1] && o->do_recon[2]);
  o->nsd = o->do_recon[0] + o->do_recon[1] + o->do_recon[2];
  o->nt = (long)o->nt3[0] * o->nt3[1] * o->nt3[2];
  o->metric_dim = (cfg->metric == 2) ? 2 : 0;
  o->buf = cfg->ct_para ? NG : 1;
  if (!o->do_recon[0]) o->c.rescale_r = 0;
  for (d = 0; d < 4; d++) { o->dx[d] = cfg->dx[d]; o->invdx[d] = 1. / cfg->dx[d]; }
  o->gam = cfg->gam;
  o->gmog = (cfg->gam - 1.) / cfg->gam;
  long nt = o->nt;
  o->p = calloc(nt * NP, sizeof(double));
  o->ph = calloc(nt * NP, sizeof(double));
  o->p_old = calloc(nt * NP, sizeof(double));
  o->pL = calloc(nt * 3 * NP, sizeof(double));
  o->pR = calloc(nt * 3 * NP, sizeof(double));
  o->F = calloc(nt * 3 * NP, sizeof(double));
  o->cgf = calloc(nt * 3 * 2, sizeof(double));
  o->uconL = calloc(nt * 3 * 2, sizeof(double));
  o->uconR = calloc(nt * 3 * 2, sizeof(double));
  o->Ugf = calloc(2 * nt * NP, sizeof(double));
  o->emf = calloc(3 * nt, sizeof(double));
  o->pgamma = calloc(nt, sizeof(double));
  o->rhoflr = calloc(nt, sizeof(double));
  o->uuflr = calloc(nt, sizeof(double));
  o->pflag = calloc(nt, sizeof(int));
  o->nfail = calloc(nt * N_NF, sizeof(int));
  int nmax = o->nt3[0];
  if (o->nt3[1] > nmax) nmax = o->nt3[1];
  if (o->nt3[2] > nmax) nmax = o->nt3[2];
  o->pvect = calloc(nmax, sizeof(double));
  o->dqvect = calloc(nmax, sizeof(double));
  o->pLvect = calloc(nmax, sizeof(double));
  o->pRvect = calloc(nmax, sizeof(double));
  o->omega = calloc(o->nt3[0], sizeof(double));
  o->tr = calloc(o->nt3[0], sizeof(double));
  o->rscale = calloc(6 * o->nt3[0], sizeof(double));
  o->irscale = calloc(6 * o->nt3[0], sizeof(double));
  /* geometry */
  if (o->metric_dim == 2) {
    int pos;
    o->geom = calloc((long)NPOS * o->nt3[0] * o->nt3[1], sizeof(geom_t));
    for (pos = 0; pos < NPOS; pos++) for (i = 0; i < o->nt3[0]; i++) for (j = 0; j < o->nt3[1]; j++)
      make_geom(o, i, j, NG, pos, (geom_t*)GEOM(o, i, j, pos));
    o->conn = calloc((long)o->n[0] * o->n[1] * 64, sizeof(double));
    for (i = NG; i < NG + o->n[0]; i++) for (j = NG; j < NG + o->n[1]; j++)
      make_conn(o, i, j, o->conn + ((long)(i - NG) * o->n[1] + (j - NG)) * 64);
  } else {
    o->geom = calloc(1, sizeof(geom_t));
    make_geom(o, NG, NG, NG, CENT, o->geom);
    o->conn = calloc(64, sizeof(double));
  }
  /* floors, coord_of_xp(), src/coord.c:3777-3797: r = x^1 for a spherical
     topology, |x| for a Cartesian one */
  for (i = 0; i < o->nt3[0]; i++) for (j = 0; j < o->nt3[1]; j++) for (k = 0; k < o->nt3[2]; k++) {
    double xp[4], x[4], J[4][4], r;
    coord_xp(o, i, j, k, CENT, xp);
    x_of_xp(o, xp, x, J);
    r = (cfg->metric == 2) ? x[1] : sqrt(x[1] * x[1] + x[2] * x[2] + x[3] * x[3]);
    o->rhoflr[IX(o, i, j, k)] = cfg->rhomin * pow(r, cfg->rhopower);
    o->uuflr[IX(o, i, j, k)] = cfg->uumin * pow(r, cfg->uupower);
  }
  /* radial rescaling tables, rescale_prims_r() first call, src/recon.c:1141-1167 */
  for (i = 0; i < o->nt3[0]; i++) {
    int pz;
    for (pz = 0; pz < 2; pz++) {
      double xp[4], x[4], J[4][4];
      coord_xp(o, i, NG, NG, pz ? FACE1 : CENT, xp);
      x_of_xp(o, xp, x, J);
      double rl = x[1], rf2 = rl * rl, rf0 = sqrt(rf2 * rl), rf1 = rf2 * rl;
      o->rscale[(3 * pz + 0) * o->nt3[0] + i] = rf0;
      o->rscale[(3 * pz + 1) * o->nt3[0] + i] = rf1;
      o->rscale[(3 * pz + 2) * o->nt3[0] + i] = rl;
    }
    for (d = 0; d < 6; d++) o->irscale[d * o->nt3[0] + i] = 1 / o->rscale[d * o->nt3[0] + i];
  }
  if (cfg->cooling) cooling_tables(o);
  return o;
}

void orc_destroy(orc_t* o) {
  if (!o) return;
  free(o->p); free(o->ph); free(o->p_old); free(o->pL); free(o->pR); free(o->F);
  free(o->cgf); free(o->uconL); free(o->uconR); free(o->Ugf); free(o->emf); free(o->pgamma);
  free(o->rhoflr); free(o->uuflr); free(o->pflag); free(o->nfail);
  free(o->pvect); free(o->dqvect); free(o->pLvect); free(o->pRvect);
  free(o->omega); free(o->tr); free(o->rscale); free(o->irscale); free(o->geom); free(o->conn);
  free(o);
}

void orc_dims(const orc_t* o, int* out) {
  out[0] = o->n[0]; out[1] = o->n[1]; out[2] = o->n[2];
  out[3] = o->nt3[0]; out[4] = o->nt3[1]; out[5] = o->nt3[2];
}
void orc_set_prims(orc_t* o, const double* p) { memcpy(o->p, p, sizeof(double) * NP * o->nt); }
void orc_get_prims(const orc_t* o, double* p) { memcpy(p, o->p, sizeof(double) * NP * o->nt); }
void orc_set_time(orc_t* o, double t, double dt) { o->t = t; o->dt = dt; }
void orc_get_time(const orc_t* o, double* out) { out[0] = o->t; out[1] = o->dt; out[2] = o->t_old; out[3] = o->dt_old; }

/* fixup(), src/fixup.c:8-66, then bounds(p, 0) */
void orc_fixup_bounds(orc_t* o) {
  int i, j, k;
  for (i = NG; i < NG + o->n[0]; i++) for (j = NG; j < NG + o->n[1]; j++) for (k = NG; k < NG + o->n[2]; k++) {
    double* pv = P8(o->p, o, i, j, k), gamma;
    set_cur(o, i, j, k, CENT);
    if (gamma_calc(pv, GEOM(o, i, j, CENT), &gamma)) { fail(o, NF_GAMMA_CALC, 0); gamma = 1.; }
    fixup_floor(o, pv, i, j, k);
    if (check_gamma(gamma)) fixup_gamma(o, pv, &gamma);
  }
  bounds(o, o->p, 0);
}

void orc_step(orc_t* o, int n) { int s; for (s = 0; s < n; s++) step_ch(o); }
void orc_get_pflag(const orc_t* o, int* out) { memcpy(out, o->pflag, sizeof(int) * o->nt); }
void orc_get_gamma(const orc_t* o, double* out) { memcpy(out, o->pgamma, sizeof(double) * o->nt); }
void orc_get_nfail(const orc_t* o, int* out) { memcpy(out, o->nfail, sizeof(int) * o->nt * N_NF); }
void orc_get_nfailtot(const orc_t* o, long long* out) { int l; for (l = 0; l < N_NF; l++) out[l] = o->nfailtot[l]; }
void orc_reset_nfail(orc_t* o) { memset(o->nfail, 0, sizeof(int) * o->nt * N_NF); memset(o->nfailtot, 0, sizeof(o->nfailtot)); }
void orc_get_vmax(const orc_t* o, double* out) { int l; for (l = 0; l < 4; l++) out[l] = o->vmax[l]; }

void orc_get_geom(const orc_t* o, int i, int j, int pos, double* out) { memcpy(out, GEOM(o, i, j, pos), sizeof(geom_t)); }
void orc_get_conn(const orc_t* o, int i, int j, double* out) {
  if (o->metric_dim == 2) memcpy(out, o->conn + ((long)(i - NG) * o->n[1] + (j - NG)) * 64, sizeof(double) * 64);
  else memset(out, 0, sizeof(double) * 64);
}

/* max |div B| over corners whose 8 zones are interior, divb_calc(), src/diag.c:503-602 */
double orc_divb_max(const orc_t* o) {
  int i, j, k, di, dj, dk;
  double m = 0.;
  for (i = NG + (o->n[0] > 1); i < NG + o->n[0]; i++)
    for (j = NG + (o->n[1] > 1); j < NG + o->n[1]; j++)
      for (k = NG + (o->n[2] > 1); k < NG + o->n[2]; k++) {
        double s1 = 0., s2 = 0., s3 = 0.;
        for (di = -1; di <= 0; di++) for (dj = -1; dj <= 0; dj++) for (dk = -1; dk <= 0; dk++) {
          const double* q = o->p + IX(o, i + di, j + dj, k + dk) * NP;
          double g = GEOM(o, i + di, j + dj, CENT)->g;
          s1 += ((di == 0) ? q[B1] : -q[B1]) * g;
          s2 += ((dj == 0) ? q[B2] : -q[B2]) * g;
          s3 += ((dk == 0) ? q[B3] : -q[B3]) * g;
        }
        double dv = fabs(0.25 * (s1 * o->invdx[1] + s2 * o->invdx[2] + s3 * o->invdx[3]));
        if (dv > m) m = dv;
      }
  return m;
}

/* ------------------------------------------------------------------------ */
/* per-zone entry points for the unit tests                                   */
void orc_prim_to_flux(const double* g42, double gam, const double* prim, int dir, double* out) {
  state_t q;
  get_state(NULL, prim, (const geom_t*)g42, gam, &q);
  primtoflux(prim, &q, dir, (const geom_t*)g42, out);
}
void orc_vchar(const double* g42, double gam, const double* prim, int dir, double* cmax, double* cmin) {
  state_t q;
  get_state(NULL, prim, (const geom_t*)g42, gam, &q);
  vchar_fast(&q, (const geom_t*)g42, dir, gam, cmax, cmin);
}
int orc_utoprim_2d(const double* g42, double gam, const double* U, double* prim, double* gamma, double* bsq, int* iters) {
  return utoprim_2d_fast(U, (const geom_t*)g42, gam, prim, gamma, bsq, iters);
}
/* Kerr-Schild record at (r, th) in Kerr-Schild coordinates (identity
   coordinate map): ks_gcov_func(), ks_gcon_func(), bl_gdet_func() and the lapse
   and shift of calc_all_geom(), src/metric.c:1168-1252, :523-539, :4235-4250. */
void orc_ks_geom_rth(double a, double r, double th, double* g42) {
  orc_t tmp;
  geom_t* G = (geom_t*)g42;
  memset(&tmp, 0, sizeof(tmp));
  tmp.c.a = a;
  ks_gcov(&tmp, r, th, G->gcov);
  ks_gcon(&tmp, r, th, G->gcon);
  G->g = ks_gdet(&tmp, r, th);
  G->g_inv = 1. / G->g;
  G->ncon[0] = sqrt(-G->gcon[0][0]);
  G->alpha = 1. / G->ncon[0];
  G->ncon[1] = -G->alpha * G->gcon[0][1];
  G->ncon[2] = -G->alpha * G->gcon[0][2];
  G->ncon[3] = -G->alpha * G->gcon[0][3];
  G->beta[0] = -G->gcon[0][1] / G->gcon[0][0];
  G->beta[1] = -G->gcon[0][2] / G->gcon[0][0];
  G->beta[2] = -G->gcon[0][3] / G->gcon[0][0];
}
int orc_utoprim_1d(const double* g42, double gam, const double* U, double* prim, double* gamma, double* bsq) {
  return utoprim_1d(U, (const geom_t*)g42, gam, prim, gamma, bsq);
}
void orc_face_flux(const double* g42, double gam, int flux, int dir, const double* pl, const double* pr,
                   double* out, double* ctop_out) {
  const geom_t* g = (const geom_t*)g42;
  state_t ql, qr;
  double Ul[NP], Ur[NP], Fl[NP], Fr[NP], cmaxl, cminl, cmaxr, cminr;
  get_state(NULL, pl, g, gam, &ql);
  get_state(NULL, pr, g, gam, &qr);
  primtoflux(pl, &ql, 0, g, Ul);
  primtoflux(pr, &qr, 0, g, Ur);
  primtoflux(pl, &ql, dir, g, Fl);
  primtoflux(pr, &qr, dir, g, Fr);
  vchar_fast(&ql, g, dir, gam, &cmaxl, &cminl);
  vchar_fast(&qr, g, dir, gam, &cmaxr, &cminr);
  double cmax = fabs(dmax(dmax(0., cmaxl), cmaxr)), cmin = fabs(dmax(dmax(0., -cminl), -cminr));
  double ctop = dmax(cmin, cmax);
  hll_or_lf(flux, cmax, cmin, ctop, Fl, Fr, Ul, Ur, out);
  if (ctop_out) *ctop_out = ctop;
}
void orc_recon_line(int recon, int n, const double* y, double* left, double* right) {
  int i;
  if (recon == 5) {
    double* dq = calloc(n, sizeof(double));
    double* yl = calloc(n, sizeof(double));
    double* yr = calloc(n, sizeof(double));
    linear_recon_vect(y, dq, 1, n - 2);
    para_vect(y, dq, yl, yr, 2, n - 3);
    for (i = 3; i <= n - 3; i++) { left[i] = yr[i - 1]; right[i] = yl[i]; }
    free(dq); free(yl); free(yr);
  } else {
    for (i = 3; i <= n - 3; i++) {
      left[i] = y[i - 1] + 0.5 * linear_recon(y[i - 2], y[i - 1], y[i]);
      right[i] = y[i] - 0.5 * linear_recon(y[i - 1], y[i], y[i + 1]);
    }
  }
}
