// hgpu_geom.h -- static geometry of the grid: metric records at the five
// zone positions, the connection at zone centres, and the per-x1 tables used
// by cooling, radial rescaling and floors.  Computed once at set-up.
//
// Metric functions restate the reference's Kerr-Schild/Minkowski formulas
// (src/metric.c:1168-1252, :523-539) and its numerical-coordinate maps
// (src/coord.c:182-500: COORD_IDENTITY and COORD_DIAGONAL "modified Kerr-Schild").
// The connection is not the reference's expanded Maple expression
// (src/metric.c:1259-1697); it is derived here from the Christoffel formula
// Gamma^d_ab = g^dc (d_a g_cb + d_b g_ca - d_c g_ab)/2 with closed-form r- and
// theta-derivatives of the Kerr-Schild metric and the diagonal coordinate
// Jacobian, which gives the same coefficients to round-off.
#pragma once
#include "hgpu_types.h"
#include <math.h>

namespace hgpu {

struct CoordParams {
  int metric;        // 0 Minkowski Cartesian, 2 Kerr-Schild spherical
  int coords;        // 0 identity, 1 modified Kerr-Schild (COORD_DIAGONAL), 4 COORD_DIAGONAL3
  double startx[4], dx[4];
  double a, asq, M, R0, h_slope, th_beg, th_length;
  int diag3_exponent;
  double diag3_factor;
  int gpos[3];       // global index of the block's zone 0 minus the local one (globalpos[1..3])
};

// seesaw(x, 0, 1), src/coord.c:4316-4324 with PERIODIC(), src/macros.h:489-490
inline double seesaw01(double x) {
  const double x0 = 0., xn = 1.;
  double xn2 = 2*xn - x0;
  int shift = (x < x0) ? (1 + (int)(fabs(x - x0)/(xn2 - x0))) : (-(int)(fabs(x - x0)/(xn2 - x0)));
  double x1 = x + (xn2 - x0)*shift;
  return (x1 < xn) ? (x1 - x0) : (xn2 - x1);
}

// coord(), src/coord.c:182-228, from the global index as get_global_ijk()
// gives it (src/coord.c:188): the block's offset gpos plus the local index
inline void coord_xp(const CoordParams& C, int i, int j, int k, int pos, double* xp) {
  xp[0] = 0.;
  i += C.gpos[0]; j += C.gpos[1]; k += C.gpos[2];
  switch (pos) {
    case FACE1: xp[1] = C.startx[1] + i*C.dx[1];         xp[2] = C.startx[2] + (j + 0.5)*C.dx[2]; xp[3] = C.startx[3] + (k + 0.5)*C.dx[3]; break;
    case FACE2: xp[1] = C.startx[1] + (i + 0.5)*C.dx[1]; xp[2] = C.startx[2] + j*C.dx[2];         xp[3] = C.startx[3] + (k + 0.5)*C.dx[3]; break;
    case FACE3: xp[1] = C.startx[1] + (i + 0.5)*C.dx[1]; xp[2] = C.startx[2] + (j + 0.5)*C.dx[2]; xp[3] = C.startx[3] + k*C.dx[3]; break;
    case CENT:  xp[1] = C.startx[1] + (i + 0.5)*C.dx[1]; xp[2] = C.startx[2] + (j + 0.5)*C.dx[2]; xp[3] = C.startx[3] + (k + 0.5)*C.dx[3]; break;
    default:    xp[1] = C.startx[1] + i*C.dx[1];         xp[2] = C.startx[2] + j*C.dx[2];         xp[3] = C.startx[3] + k*C.dx[3]; break;
  }
}

// x_of_xp(), dx_dxp_calc(), src/coord.c:240-500
inline void x_of_xp(const CoordParams& C, const double* xp, double* x, double J[4]) {
  for (int l = 0; l < 4; l++) J[l] = 1.;
  if (C.coords == 1) {
    double xp2 = seesaw01(xp[2]);
    x[0] = xp[0];
    x[1] = C.R0 + exp(xp[1]);
    x[2] = C.th_length * xp2 + C.th_beg + C.h_slope * sin(2 * M_PI * xp2);
    x[3] = xp[3];
    J[1] = x[1] - C.R0;
    J[2] = C.th_length + 2 * M_PI * C.h_slope * cos(2 * M_PI * xp[2]);
  } else if (C.coords == 4) {
    // COORD_DIAGONAL3: theta a polynomial in f = 2 seesaw(x2) - 1 (src/coord.c:270-276, :468-472)
    const double f = 2 * seesaw01(xp[2]) - 1.;
    x[0] = xp[0];
    x[1] = C.R0 + exp(xp[1]);
    x[2] = 0.5*M_PI*(1. + C.h_slope*f + C.diag3_factor*pow(f, (double)C.diag3_exponent));
    x[3] = xp[3];
    J[1] = x[1] - C.R0;
    J[2] = M_PI*(C.h_slope + C.diag3_exponent*C.diag3_factor*pow(f, (double)(C.diag3_exponent - 1)));
  } else {
    for (int l = 0; l < 4; l++) x[l] = xp[l];
  }
}

// trigvalues(), src/metric.c:29-40 (COORDSINGFIX)
HD void trigvalues(double th, double* cth, double* sth) {
  *sth = sin(th);
  *cth = cos(th);
  if (fabs(*sth) < 1.E-16) {
    if ((*sth) >= 0) *sth = 1.E-16;
    if ((*sth) < 0) *sth = -1.E-16;
  }
}

// Kerr-Schild metric in (t, r, th, phi): ks_gcov_func(), ks_gcon_func(),
// bl_gdet_func() of src/metric.c:1168-1252, :523-539
HD void ks_metric_rth(double a, double M, double r, double th, double gcov[4][4], double gcon[4][4],
                      double* gdet) {
  const double asq = a*a;
  double t4, t12;
  trigvalues(th, &t4, &t12);
  {
    double t2 = r*r, t3 = asq, t5 = t4*t4, t7 = t2 + t3*t5;
    double t9 = M*r/t7, t10 = 2.0*t9, t13 = t12*t12;
    for (int p = 0; p < 4; p++) for (int q = 0; q < 4; q++) gcov[p][q] = 0.;
    gcov[0][0] = t10 - 1.0;
    gcov[0][1] = t10;
    gcov[0][3] = -2.0*t9*a*t13;
    gcov[1][0] = t10;
    gcov[1][1] = 1.0 + t10;
    gcov[1][3] = -a*gcov[1][1]*t13;
    gcov[2][2] = t7;
    gcov[3][0] = gcov[0][3];
    gcov[3][1] = gcov[1][3];
    gcov[3][3] = t13*(t2 + t3*(1.0 + 2.0*t9*t13));
  }
  {
    double t1 = M*r, t2 = r*r, t3 = asq, t5 = t4*t4;
    double t8 = 1/(t2 + t3*t5), t10 = 2.0*t1*t8, t14 = t12*t12;
    for (int p = 0; p < 4; p++) for (int q = 0; q < 4; q++) gcon[p][q] = 0.;
    gcon[0][0] = -1.0 - t10;
    gcon[0][1] = t10;
    gcon[1][0] = t10;
    gcon[1][1] = (t2 - 2.0*t1 + t3)*t8;
    gcon[1][3] = a*t8;
    gcon[2][2] = t8;
    gcon[3][1] = gcon[1][3];
    gcon[3][3] = t8/t14;
  }
  {
    double cth, sth;
    trigvalues(th, &cth, &sth);
    double cos_2th = (cth - sth)*(cth + sth);
    double a2 = asq, r2 = r*r;
    *gdet = r2*fabs(sth)*(1. + 0.5*(a2/r2)*(1. + cos_2th));
  }
}

inline void ks_metric(const CoordParams& C, double r, double th, double gcov[4][4], double gcon[4][4], double* gdet) {
  ks_metric_rth(C.a, C.M, r, th, gcov, gcon, gdet);
}

// Lapse, shift and normal observer of a record whose gcov, gcon and g are set
// (calc_all_geom(), src/metric.c:4235-4250).
HD void geom_finish(Geom* G) {
  G->g_inv = 1./G->g;
  G->ncon[0] = sqrt(-G->gcon[0][0]);
  G->alpha = 1./G->ncon[0];
  G->ncon[1] = -G->alpha * G->gcon[0][1];
  G->ncon[2] = -G->alpha * G->gcon[0][2];
  G->ncon[3] = -G->alpha * G->gcon[0][3];
  G->beta[0] = -G->gcon[0][1]/G->gcon[0][0];
  G->beta[1] = -G->gcon[0][2]/G->gcon[0][0];
  G->beta[2] = -G->gcon[0][3]/G->gcon[0][0];
}

// Geometry record at (i,j,k,pos): calc_all_geom(), src/metric.c:4111-4233,
// with transform_all(), src/coord.c:2340-2353 (diagonal Jacobian).
inline void geom_record(const CoordParams& C, int i, int j, int k, int pos, Geom* G) {
  double xp[4], x[4], J[4];
  coord_xp(C, i, j, k, pos, xp);
  x_of_xp(C, xp, x, J);
  double gcov[4][4], gcon[4][4], gdet;
  if (C.metric == 2) {
    ks_metric(C, x[1], x[2], gcov, gcon, &gdet);
  } else {
    for (int a = 0; a < 4; a++) for (int b = 0; b < 4; b++) gcov[a][b] = gcon[a][b] = (a == b) ? ((a == 0) ? -1. : 1.) : 0.;
    gdet = 1.;
  }
  // dxp_dx_calc2_default(), src/coord.c:1629-1654
  double inv_det = 1. / (J[1]*J[2] - 0.*0.);
  double K[4] = {1., J[2]*inv_det, J[1]*inv_det, 1.};
  for (int a = 0; a < 4; a++) for (int b = 0; b < 4; b++) {
    G->gcov[a][b] = (gcov[a][b]*J[a])*J[b];
    G->gcon[a][b] = (gcon[a][b]*K[a])*K[b];
  }
  G->g = gdet * (1.*1.*(J[1]*J[2] - 0.*0.));
  G->g_inv = 1./G->g;
  G->ncon[0] = sqrt(-G->gcon[0][0]);
  G->alpha = 1./G->ncon[0];
  G->ncon[1] = -G->alpha * G->gcon[0][1];
  G->ncon[2] = -G->alpha * G->gcon[0][2];
  G->ncon[3] = -G->alpha * G->gcon[0][3];
  G->beta[0] = -G->gcon[0][1]/G->gcon[0][0];
  G->beta[1] = -G->gcon[0][2]/G->gcon[0][0];
  G->beta[2] = -G->gcon[0][3]/G->gcon[0][0];
}

// Connection Gamma^d_{ab} at the centre of zone (i,j): Kerr-Schild in
// modified Kerr-Schild coordinates (R0 = 0 makes dr/dx1 = r - R0 = d2r/dx1^2).
// conn[(d*4 + a)*4 + b], the reference's conn[l][d][a][b] layout.
inline void ks_connection(const CoordParams& C, int i, int j, double* conn) {
  double xp[4], x[4], J[4];
  coord_xp(C, i, j, NG, CENT, xp);
  x_of_xp(C, xp, x, J);
  const double M = C.M, a = C.a, a2 = C.asq;
  const double r = x[1], th = x[2];
  double c, s;
  trigvalues(th, &c, &s);
  const double S = r*r + a2*c*c;
  const double z = 2.*M*r/S;
  const double dz_r = 2.*M*(S - 2.*r*r)/(S*S);
  const double dz_t = 4.*M*r*a2*c*s/(S*S);
  const double s2 = s*s;
  // g_{ab}(r,th) and its derivatives
  double g[4][4] = {{0}}, gr[4][4] = {{0}}, gt[4][4] = {{0}};
  g[0][0] = z - 1.;              gr[0][0] = dz_r;                   gt[0][0] = dz_t;
  g[0][1] = z;                   gr[0][1] = dz_r;                   gt[0][1] = dz_t;
  g[0][3] = -z*a*s2;             gr[0][3] = -a*s2*dz_r;             gt[0][3] = -a*(dz_t*s2 + z*2.*s*c);
  g[1][1] = 1. + z;              gr[1][1] = dz_r;                   gt[1][1] = dz_t;
  g[1][3] = -a*(1. + z)*s2;      gr[1][3] = -a*s2*dz_r;             gt[1][3] = -a*(dz_t*s2 + (1. + z)*2.*s*c);
  g[2][2] = S;                   gr[2][2] = 2.*r;                   gt[2][2] = -2.*a2*c*s;
  g[3][3] = s2*(r*r + a2*(1. + z*s2));
  gr[3][3] = s2*(2.*r + a2*s2*dz_r);
  gt[3][3] = 2.*s*c*(r*r + a2*(1. + z*s2)) + s2*a2*(dz_t*s2 + z*2.*s*c);
  for (int p = 0; p < 4; p++) for (int q = 0; q < p; q++) { g[p][q] = g[q][p]; gr[p][q] = gr[q][p]; gt[p][q] = gt[q][p]; }
  // numerical-coordinate Jacobian and its derivatives (diagonal, J_a depends on x^a only)
  double Jd[4] = {0., 0., 0., 0.};            // dJ_a/dx^a
  if (C.coords == 1) {
    Jd[1] = r - C.R0;                          // d^2 r/dx1^2 = exp(x1)
    Jd[2] = -4.*M_PI*M_PI*C.h_slope*sin(2.*M_PI*xp[2]);
  } else if (C.coords == 4) {
    // dx_dxp_dxp_calc(), src/coord.c:942-947
    const double f = 2 * seesaw01(xp[2]) - 1.;
    const int n = C.diag3_exponent;
    Jd[1] = r - C.R0;
    Jd[2] = 2*M_PI*n*(n - 1)*C.diag3_factor*pow(f, (double)(n - 2));
  }
  // dg'_{ab}/dx^c in numerical coordinates, c = 1 (x1) and 2 (x2)
  double dg[4][4][4] = {{{0}}};   // dg[c][a][b]
  for (int p = 0; p < 4; p++) for (int q = 0; q < 4; q++) {
    double gab = g[p][q] * J[p] * J[q];
    (void)gab;
    // c = 1
    double v1 = gr[p][q] * J[1] * J[p] * J[q];
    if (p == 1) v1 += g[p][q] * Jd[1] * J[q];
    if (q == 1) v1 += g[p][q] * J[p] * Jd[1];
    double v2 = gt[p][q] * J[2] * J[p] * J[q];
    if (p == 2) v2 += g[p][q] * Jd[2] * J[q];
    if (q == 2) v2 += g[p][q] * J[p] * Jd[2];
    dg[1][p][q] = v1;
    dg[2][p][q] = v2;
  }
  // inverse metric in numerical coordinates
  Geom G;
  geom_record(C, i, j, NG, CENT, &G);
  double low[4][4][4];   // Gamma_{c,ab}
  for (int cc = 0; cc < 4; cc++) for (int p = 0; p < 4; p++) for (int q = 0; q < 4; q++)
    low[cc][p][q] = 0.5*(dg[p][cc][q] + dg[q][cc][p] - dg[cc][p][q]);
  for (int d = 0; d < 4; d++) for (int p = 0; p < 4; p++) for (int q = 0; q < 4; q++) {
    double sum = 0.;
    for (int cc = 0; cc < 4; cc++) sum += G.gcon[d][cc]*low[cc][p][q];
    conn[(d*4 + p)*4 + q] = sum;
  }
}

// omega_circular_equatorial() and target_temperature(), src/phys.c:805-907,
// with the ISCO reference values computed as the reference does
// (coord_of_r() at the equator, src/coord.c:3866-3905; Kerr-Schild metric
// without the numerical-coordinate transform, src/metric.c:4277-4310).
struct CoolTables {
  double r_isco, h_o_r;
  double vcov_isco[4];
  double ut_sq_isco, uphi_sq_isco;
};

inline void ks_at_equator(const CoordParams& C, double r, double* xr, double* xth) {
  double xp[4] = {0., 0., 0.5, 0.};
  xp[1] = (C.coords != 0) ? log(r - C.R0) : r;
  double x[4], J[4];
  x_of_xp(C, xp, x, J);
  *xr = x[1];
  *xth = x[2];
}

inline void vel_circular_equatorial(const CoordParams& C, double r_isco, double r, double* vcon) {
  double r3halves = pow(r, 1.5);
  double uphi = 2*C.a*r3halves + r*r*(r - 3.);
  if (uphi <= 0.) {
    r = r_isco;
    r3halves = pow(r, 1.5);
    uphi = 2*C.a*r3halves + r*r*(r - 3.);
  }
  uphi = 1. / sqrt(uphi);
  vcon[1] = vcon[2] = 0.;
  vcon[3] = uphi;
  vcon[0] = uphi * (r3halves + C.a);
}

inline void cool_setup(const CoordParams& C, double r_isco, double h_o_r, CoolTables* T) {
  T->r_isco = r_isco;
  T->h_o_r = h_o_r;
  double xr, xth, gcov[4][4], gcon[4][4], gdet, vcon[4];
  ks_at_equator(C, r_isco, &xr, &xth);
  ks_metric(C, xr, xth, gcov, gcon, &gdet);
  vel_circular_equatorial(C, r_isco, r_isco, vcon);
  for (int a1 = 0; a1 < 4; a1++)
    T->vcov_isco[a1] = gcov[a1][0]*vcon[0] + gcov[a1][1]*vcon[1] + gcov[a1][2]*vcon[2] + gcov[a1][3]*vcon[3];
  double r = r_isco;
  double rm1 = 1./r, x = sqrt(r), xm3 = 1./(x*x*x), term1 = C.a*C.a*rm1;
  double C_f = 1. - 3*rm1 + 2*C.a*xm3;
  double F_f = 1. - 2*C.a*xm3 + term1*rm1;
  double G_f = 1. - 2*rm1 + C.a*xm3;
  T->ut_sq_isco = G_f * G_f / C_f;
  T->uphi_sq_isco = F_f * F_f * r / C_f;
}

inline double omega_circular_equatorial(const CoordParams& C, const CoolTables& T, double r_loc) {
  double r = r_loc;
  if (r >= T.r_isco) return sqrt(C.M)/(pow(r, 1.5) + C.a);
  double xr, xth, gcov[4][4], gcon[4][4], gdet, vcon[4];
  ks_at_equator(C, r, &xr, &xth);
  ks_metric(C, xr, xth, gcov, gcon, &gdet);
  const double* v = T.vcov_isco;
  for (int a1 = 0; a1 < 4; a1++)
    vcon[a1] = gcon[a1][0]*v[0] + gcon[a1][1]*v[1] + gcon[a1][2]*v[2] + gcon[a1][3]*v[3];
  return vcon[3]/vcon[0];
}

inline double target_temperature(const CoordParams& C, const CoolTables& T, double r_loc) {
  double r = r_loc;
  double rm1 = 1./r, x = sqrt(r), xm3 = 1./(x*x*x), term1 = C.a*C.a*rm1, R_z;
  if (r > T.r_isco) {
    double C_f = 1. - 3*rm1 + 2*C.a*xm3;
    double F_f = 1. - 2*C.a*xm3 + term1*rm1;
    double G_f = 1. - 2*rm1 + C.a*xm3;
    double ut_sq = G_f * G_f / C_f;
    double uphi_sq = F_f * F_f * r / C_f;
    R_z = uphi_sq*rm1 - term1*(ut_sq - 1.);
    term1 = 0.5*M_PI*R_z*T.h_o_r*T.h_o_r*rm1;
  } else {
    R_z = T.uphi_sq_isco*rm1 - term1*(T.ut_sq_isco - 1.);
    term1 = 0.5*M_PI*R_z*T.h_o_r*T.h_o_r*rm1;
  }
  return term1;
}

// risco_calc(1), src/metric.c:458-472 (with asq = a^2)
inline double risco_prograde(double a) {
  double term1 = pow(1. + a, 1./3.), term2 = pow(1. - a, 1./3.);
  double Z1 = 1. + term1*term2*(term1 + term2);
  double Z2 = sqrt(3.*a*a + Z1*Z1);
  return (3. + Z2 - sqrt((3. - Z1)*(3. + Z1 + 2.*Z2)));
}

}  // namespace hgpu
