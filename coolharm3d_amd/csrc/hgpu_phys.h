// hgpu_phys.h -- per-zone GRMHD physics as __host__ __device__ functions.
//
// Every routine restates the reference's arithmetic in the reference's
// operation order (the FP64 step is checked against the reference build to
// 1e-12 and the con-to-prim flag map bit for bit), with the reference's
// file-scope solver globals (src/utoprim_2d_fast.c:81, src/utoprim_1d_ee.c:84)
// turned into a per-thread struct so the solvers are re-entrant.
#pragma once
#include "hgpu_types.h"
#include <math.h>

namespace hgpu {

constexpr double SMALL = 1.e-15;          // src/macros.h:441
constexpr double GAMMAMAX = 50.;          // src/options.h:264
constexpr double GAMMAMAX2 = 1.e4;        // src/options.h:265
constexpr double BETA_MIN = 1.e-2;        // src/options.h:212
constexpr double PPM_FACTOR = 0.16666666666666666667;  // src/recon.h:10-11
constexpr int MAX_NEWT_ITER = 30;         // src/u2p_defs.h:76-79
constexpr double NEWT_TOL = 1.0e-10;
constexpr double MIN_NEWT_TOL = 1.0e-10;
constexpr int EXTRA_NEWT_ITER = 2;
constexpr double NEWT_TOL2 = 1.0e-15;
constexpr double W_TOO_BIG = 1.e20;
constexpr double UTSQ_TOO_BIG = 1.e20;
constexpr double FAIL_VAL = 1.e30;

// ---------------------------------------------------------------------------
// atomics: device builds use hardware atomics; the single-threaded host
// check build of these bodies updates memory directly.
HD void atomic_add_i32(int* a, int v) {
#ifdef __HIP_DEVICE_COMPILE__
  atomicAdd(a, v);
#else
  *a += v;
#endif
}
HD void atomic_add_u64(unsigned long long* a, unsigned long long v) {
#ifdef __HIP_DEVICE_COMPILE__
  atomicAdd(a, v);
#else
  *a += v;
#endif
}
HD void atomic_max_u64(unsigned long long* a, unsigned long long v) {
#ifdef __HIP_DEVICE_COMPILE__
  atomicMax(a, v);
#else
  if (v > *a) *a = v;
#endif
}
HD void atomic_or_u64(unsigned long long* a, unsigned long long v) {
#ifdef __HIP_DEVICE_COMPILE__
  atomicOr(a, v);
#else
  *a |= v;
#endif
}

HD bool is_finite(double x) { return isfinite(x); }

// ---------------------------------------------------------------------------
// The functions of the per-zone path that read a geometry record take its
// pointer as a template parameter GP: a plain `const Geom*`, or on the device a
// pointer into the constant address space when every lane of the wavefront
// reads the same record (its fields then load through the scalar cache).

// Metric contractions, src/metric.c:338-383
template <class GP>
HD void lower(const double* ucon, GP g, double* ucov) {
  ucov[0] = g->gcov[0][0]*ucon[0] + g->gcov[0][1]*ucon[1] + g->gcov[0][2]*ucon[2] + g->gcov[0][3]*ucon[3];
  ucov[1] = g->gcov[1][0]*ucon[0] + g->gcov[1][1]*ucon[1] + g->gcov[1][2]*ucon[2] + g->gcov[1][3]*ucon[3];
  ucov[2] = g->gcov[2][0]*ucon[0] + g->gcov[2][1]*ucon[1] + g->gcov[2][2]*ucon[2] + g->gcov[2][3]*ucon[3];
  ucov[3] = g->gcov[3][0]*ucon[0] + g->gcov[3][1]*ucon[1] + g->gcov[3][2]*ucon[2] + g->gcov[3][3]*ucon[3];
}
template <class GP>
HD void raise_v(const double* ucov, GP g, double* ucon) {
  ucon[0] = g->gcon[0][0]*ucov[0] + g->gcon[0][1]*ucov[1] + g->gcon[0][2]*ucov[2] + g->gcon[0][3]*ucov[3];
  ucon[1] = g->gcon[1][0]*ucov[0] + g->gcon[1][1]*ucov[1] + g->gcon[1][2]*ucov[2] + g->gcon[1][3]*ucov[3];
  ucon[2] = g->gcon[2][0]*ucov[0] + g->gcon[2][1]*ucov[1] + g->gcon[2][2]*ucov[2] + g->gcon[2][3]*ucov[3];
  ucon[3] = g->gcon[3][0]*ucov[0] + g->gcon[3][1]*ucov[1] + g->gcon[3][2]*ucov[2] + g->gcon[3][3]*ucov[3];
}
HD double dot4(const double* a, const double* b) {
  return a[0]*b[0] + a[1]*b[1] + a[2]*b[2] + a[3]*b[3];
}

// gamma_calc(), src/phys.c:244-268.  Returns 1 on failure (gamma = 1).
template <class GP>
HD int gamma_calc(const double* pr, GP g, double* gamma) {
  double qsq = g->gcov[1][1]*pr[U1]*pr[U1]
             + g->gcov[2][2]*pr[U2]*pr[U2]
             + g->gcov[3][3]*pr[U3]*pr[U3]
             + 2*(g->gcov[1][2]*pr[U1]*pr[U2]
                + g->gcov[1][3]*pr[U1]*pr[U3]
                + g->gcov[2][3]*pr[U2]*pr[U3]);
  if (qsq < 0.) {
    if (fabs(qsq) > 1.E-10) { *gamma = 1.; return 1; }
    else qsq = 1.E-10;
  }
  *gamma = sqrt(1. + qsq);
  return 0;
}

// get_state(), src/phys.c:199-241 (ucon_calc, lower, bcon_calc, lower).
// Returns the gamma_calc failure flag (the caller records FAIL_GAMMA_CALC).
template <class GP>
HD int get_state(const double* pr, GP g, double gam, State* q) {
  double gamma;
  int fl = gamma_calc(pr, g, &gamma);
  q->ucon[0] = gamma*g->ncon[0];
  q->ucon[1] = pr[U1] - q->ucon[0]*g->beta[0];
  q->ucon[2] = pr[U2] - q->ucon[0]*g->beta[1];
  q->ucon[3] = pr[U3] - q->ucon[0]*g->beta[2];
  lower(q->ucon, g, q->ucov);
  double inv_ut = 1./q->ucon[0];
  q->bcon[0] = pr[B1]*q->ucov[1] + pr[B2]*q->ucov[2] + pr[B3]*q->ucov[3];
  q->bcon[1] = (pr[B1] + q->bcon[0]*q->ucon[1])*inv_ut;
  q->bcon[2] = (pr[B2] + q->bcon[0]*q->ucon[2])*inv_ut;
  q->bcon[3] = (pr[B3] + q->bcon[0]*q->ucon[3])*inv_ut;
  lower(q->bcon, g, q->bcov);
  q->bsq = dot4(q->bcon, q->bcov);
  q->p = (gam - 1.)*pr[UU];
  q->eta = q->p + pr[RHO] + pr[UU] + q->bsq;
  q->ptot = q->p + 0.5*q->bsq;
  return fl;
}

// mhd_calc(), src/phys.c:93-116: T^dir_mu (first index up, second down)
HD void mhd_calc(const State* q, int dir, double* mhd) {
  double eta_ucon = q->eta * q->ucon[dir];
  double bcon_dir = q->bcon[dir];
  mhd[0] = eta_ucon * q->ucov[0] - bcon_dir * q->bcov[0];
  mhd[1] = eta_ucon * q->ucov[1] - bcon_dir * q->bcov[1];
  mhd[2] = eta_ucon * q->ucov[2] - bcon_dir * q->bcov[2];
  mhd[3] = eta_ucon * q->ucov[3] - bcon_dir * q->bcov[3];
  mhd[dir] += q->ptot;
}

// primtoflux(), src/phys.c:28-63
template <class GP>
HD void primtoflux(const double* pr, const State* q, int dir, GP g, double* flux) {
  double mhd[4];
  mhd_calc(q, dir, mhd);
  flux[RHO] = pr[RHO]*q->ucon[dir];
  flux[UU] = mhd[0] + flux[RHO];
  flux[U1] = mhd[1];
  flux[U2] = mhd[2];
  flux[U3] = mhd[3];
  flux[B1] = q->bcon[1]*q->ucon[dir] - q->bcon[dir]*q->ucon[1];
  flux[B2] = q->bcon[2]*q->ucon[dir] - q->bcon[dir]*q->ucon[2];
  flux[B3] = q->bcon[3]*q->ucon[dir] - q->bcon[dir]*q->ucon[3];
  for (int l = 0; l < NP; l++) flux[l] *= g->g;
}

// vchar_fast(), src/phys.c:432-501
HD void vchar_fast(const State* q, const Geom* g, int js, double gam, double* vmax, double* vmin) {
  double cms2 = (gam * q->p + q->bsq) / q->eta;
  if (cms2 < 0.) cms2 = SMALL;
  else if (cms2 > 1.) cms2 = 1.;
  double Au2 = q->ucon[js]*q->ucon[js];
  double Bu2 = q->ucon[0]*q->ucon[0];
  double AuBu = q->ucon[js]*q->ucon[0];
  double A = Bu2 - (g->gcon[0][0] + Bu2)*cms2;
  double B = AuBu - (g->gcon[0][js] + AuBu)*cms2;
  double C = Au2 - (g->gcon[js][js] + Au2)*cms2;
  double discr = B*B - A*C;
  if (discr < 0.) discr = 0.;   // both branches of src/phys.c:473-484 set 0
  discr = sqrt(discr);
  double denom = 1. / A;
  double vp = -(-B + discr)*denom;
  double vm = -(-B - discr)*denom;
  if (vp > vm) { *vmax = vp; *vmin = vm; }
  else         { *vmax = vm; *vmin = vp; }
}

// ---------------------------------------------------------------------------
// Reconstruction (src/recon.c)

// linear_recon(), MC limiter, src/recon.c:88-132
HD double mc_slope(double y1, double y2, double y3) {
  double Dqm = 2.*(y2 - y1);
  double Dqp = 2.*(y3 - y2);
  double Dqc = 0.5*(y3 - y1);
  double s = Dqm*Dqp;
  if (s <= 0.) return 0.;
  double aDqm = fabs(Dqm), aDqp = fabs(Dqp), aDqc = fabs(Dqc);
  if ((aDqm <= aDqp) && (aDqm <= aDqc)) return Dqm;
  else if (aDqp <= aDqc) return Dqp;
  else return Dqc;
}

// linear_recon_vect() MC form, src/recon.c:237-264 (slope of the middle point)
HD double mc_slope_vect(double ym, double y0, double yp) {
  double a = y0 - ym, b = yp - y0, s = a*b;
  if (s <= 0.) return 0.;
  double a_abs = fabs(a), b_abs = fabs(b);
  double a3 = a_abs + a_abs + a_abs, b3 = b_abs + b_abs + b_abs;
  if (a3 <= b_abs) return a + a;
  else if (b3 <= a_abs) return b + b;
  else return 0.5*(a + b);
}

// para_vect() limiter for one zone, src/recon.c:611-641
HD void ppm_limit(double l, double r, double yi, double* lnew, double* rnew) {
  double qa = (r - yi)*(yi - l);
  if (qa <= 0.) { *lnew = *rnew = yi; return; }
  double r2 = r + r, l2 = l + l, y3 = yi + yi + yi;
  *lnew = l; *rnew = r;
  if (r > l) {
    if (y3 > (l + r2)) *lnew = y3 - r2;
    if (y3 < (l2 + r)) *rnew = y3 - l2;
  } else {
    if (y3 < (l + r2)) *lnew = y3 - r2;
    if (y3 > (l2 + r)) *rnew = y3 - l2;
  }
}

// PPM states at the interface between window points 2 and 3 of y[0..5]:
// *yl = right edge of zone 2 (the reference's pR_vect[2]) and
// *yr = left edge of zone 3 (pL_vect[3]), src/recon.c:237-264, :611-641.
HD void ppm_interface(const double* y, double* yl, double* yr) {
  double dq1 = mc_slope_vect(y[0], y[1], y[2]);
  double dq2 = mc_slope_vect(y[1], y[2], y[3]);
  double dq3 = mc_slope_vect(y[2], y[3], y[4]);
  double dq4 = mc_slope_vect(y[3], y[4], y[5]);
  double f12 = 0.5*(y[2] + y[1]) - PPM_FACTOR*(dq2 - dq1);
  double f23 = 0.5*(y[3] + y[2]) - PPM_FACTOR*(dq3 - dq2);
  double f34 = 0.5*(y[4] + y[3]) - PPM_FACTOR*(dq4 - dq3);
  double l2, r2, l3, r3;
  ppm_limit(f12, f23, y[2], &l2, &r2);
  ppm_limit(f23, f34, y[3], &l3, &r3);
  *yl = r2;
  *yr = l3;
}

// ---------------------------------------------------------------------------
// Con-to-prim: 2D (W, v^2) Newton, src/utoprim_2d_fast.c:135-740

struct U2d {   // file-scope globals of src/utoprim_2d_fast.c:81
  double Bsq, QdotBsq, Qtsq, Qdotn, D, half_Bsq, gmog;
};

HD double u2d_vsq_calc(const U2d& s, double W) {
  double Wsq = W*W;
  double Bsq_W = (s.Bsq + W);
  double Xsq = Bsq_W * Bsq_W;
  return (Wsq * s.Qtsq + s.QdotBsq * (Bsq_W + W)) / (Wsq*Xsq);
}

HD double u2d_x1_of_x0(const U2d& s, double x0) {
  double dv = 1.e-15;
  double vsq = fabs(u2d_vsq_calc(s, x0));
  return (vsq > 1.) ? (1.0 - dv) : vsq;
}

HD void u2d_func_vsq(const U2d& s, const double* x, double* dx, double* f, double* df) {
  double W = x[0], vsq = x[1];
  double ftmp = 1. - vsq;
  double gtmp = sqrt(ftmp);
  double dPdW = s.gmog * ftmp;
  double dPdvsq = s.gmog * (0.5 * s.D/gtmp - W);
  double p_tmp = s.gmog * (W * ftmp - s.D * gtmp);
  double t2 = -s.half_Bsq + dPdvsq;
  double t3 = s.Bsq + W;
  double t4 = t3*t3;
  double t23 = 1/W;
  double t16 = s.QdotBsq*t23*t23;
  double t11 = s.Qtsq - vsq*t4 + t16*(s.Bsq + W + W);
  double t18 = -s.Qdotn - s.half_Bsq*(1.0 + vsq) + 0.5*t16 - W + p_tmp;
  double t24 = t16*t23;
  double t25 = -1.0 + dPdW - t24;
  double t35 = t25*t3 + (s.Bsq - 2.0*dPdvsq)*(t16 + vsq*W)*t23;
  double t21 = 1/(t3*t35);
  dx[0] = -(t2*t11 + t4*t18)*t21;
  double t40 = -2*(vsq + t24)*t3;
  dx[1] = -(-t25*t11 + t40*t18)*t21;
  double r0 = t11, r1 = t18;
  *df = -r0*r0 - r1*r1;
  *f = -0.5 * (*df);
}

// general_newton_raphson(), src/utoprim_2d_fast.c:489-592, as a resumable
// iteration: u2d_newton_step() performs one pass of the reference's loop body
// and returns whether to keep iterating, so a persistent kernel can interleave
// cells; u2d_newton() runs it to completion.
struct NewtonIter {
  double x[2];
  double f, df, errx;
  int n_iter, i_extra, doing_extra;
  int n_extra;   // confirming iterations after convergence (EXTRA_NEWT_ITER)
};

HD void u2d_newton_init(NewtonIter& it, const double* x, int n_extra = EXTRA_NEWT_ITER) {
  it.x[0] = x[0]; it.x[1] = x[1];
  it.f = 1.; it.df = 1.; it.errx = 1.;
  it.n_iter = 0; it.i_extra = 0; it.doing_extra = 0;
  it.n_extra = n_extra;
}

HD bool u2d_newton_step(const U2d& s, NewtonIter& it) {
  const double dv = (1. - 1.e-15);
  double dx[2], x_old[2];
  u2d_func_vsq(s, it.x, dx, &it.f, &it.df);
  it.errx = 0.;
  x_old[0] = it.x[0]; x_old[1] = it.x[1];
  it.x[0] += dx[0];
  it.x[1] += dx[1];
  it.errx = (it.x[0] == 0.) ? fabs(dx[0]) : fabs(dx[0]/it.x[0]);
  if (it.x[0] < 0.) { it.x[0] = fabs(it.x[0]); }
  else { if (it.x[0] > W_TOO_BIG) { it.x[0] = x_old[0]; } }
  if (it.x[1] < 0.) { it.x[1] = 0.; }
  else { if (it.x[1] > 1.) { it.x[1] = dv; } }
  if ((fabs(it.errx) <= NEWT_TOL) && (it.doing_extra == 0) && (it.n_extra > 0)) it.doing_extra = 1;
  if (it.doing_extra == 1) it.i_extra++;
  bool keep = !(((fabs(it.errx) <= NEWT_TOL) && (it.doing_extra == 0)) || (it.i_extra > it.n_extra) ||
                (it.n_iter >= (MAX_NEWT_ITER - 1)));
  it.n_iter++;
  return keep;
}

HD int u2d_newton_result(const NewtonIter& it) {
  if ((!is_finite(it.f)) || (!is_finite(it.df))) return 2;
  if (fabs(it.errx) <= NEWT_TOL) return 0;
  else if ((fabs(it.errx) <= MIN_NEWT_TOL) && (fabs(it.errx) > NEWT_TOL)) return 0;
  return 1;
}

HD int u2d_newton(const U2d& s, double* x, int* iters, int n_extra = EXTRA_NEWT_ITER) {
  NewtonIter it;
  u2d_newton_init(it, x, n_extra);
  while (u2d_newton_step(s, it)) {}
  x[0] = it.x[0]; x[1] = it.x[1];
  if (iters) *iters = it.n_iter;
  return u2d_newton_result(it);
}

// Quantities of Utoprim_new_body() that the final step needs besides the
// Newton solution (src/utoprim_2d_fast.c:395-406).
struct U2dAux {
  double QdotB, Qcon[4], Bcon[4], ncon[4];
};

// Utoprim_2d_fast() + Utoprim_new_body() up to the initial guess
// (src/utoprim_2d_fast.c:155-300): the invariants of U the Newton solve works
// with.  Sets prim's B components.  Returns -100 when U[RHO] <= 0.
template <class GP>
HD int u2d_prepare(const double* U, GP geom, double gmog, double* prim, U2d& s, U2dAux& aux) {
  if (U[0] <= 0.) return -100;
  double inv_gdet = geom->g_inv;
  double alpha = geom->alpha;
  double geomfactor = alpha * inv_gdet;
  prim[B1] = U[B1] * inv_gdet;
  prim[B2] = U[B2] * inv_gdet;
  prim[B3] = U[B3] * inv_gdet;
  double Ut[NP];
  Ut[RHO] = geomfactor * U[RHO];
  Ut[UU] = geomfactor * (U[UU] - U[RHO]);
  Ut[U1] = geomfactor * U[U1];
  Ut[U2] = geomfactor * U[U2];
  Ut[U3] = geomfactor * U[U3];
  Ut[B1] = geomfactor * U[B1];
  Ut[B2] = geomfactor * U[B2];
  Ut[B3] = geomfactor * U[B3];

  s.gmog = gmog;
  double* Bcon = aux.Bcon;
  Bcon[0] = 0.; Bcon[1] = Ut[B1]; Bcon[2] = Ut[B2]; Bcon[3] = Ut[B3];
  s.Bsq = geom->gcov[1][1]*Bcon[1]*Bcon[1]
        + geom->gcov[2][2]*Bcon[2]*Bcon[2]
        + geom->gcov[3][3]*Bcon[3]*Bcon[3]
        + 2*(geom->gcov[1][2]*Bcon[1]*Bcon[2]
           + geom->gcov[1][3]*Bcon[1]*Bcon[3]
           + geom->gcov[2][3]*Bcon[2]*Bcon[3]);
  double Qcov[4] = {Ut[UU], Ut[U1], Ut[U2], Ut[U3]};
  raise_v(Qcov, geom, aux.Qcon);
  aux.QdotB = Qcov[1]*Bcon[1] + Qcov[2]*Bcon[2] + Qcov[3]*Bcon[3];
  s.QdotBsq = aux.QdotB*aux.QdotB;
  s.Qdotn = -alpha*aux.Qcon[0];
  double Qsq = dot4(Qcov, aux.Qcon);
  s.Qtsq = Qsq + s.Qdotn*s.Qdotn;
  s.D = Ut[RHO];
  s.half_Bsq = 0.5*s.Bsq;
  for (int l = 0; l < 4; l++) aux.ncon[l] = geom->ncon[l];
  return 0;
}

// Initial (W, v^2) from an estimate W_last (src/utoprim_2d_fast.c:313-330):
// W is raised by `factor` (the reference: tenfold, at most 10 times) until it
// lies above the physical root's lower bound.  vsq_guess >= 0 replaces v^2(W)
// when no increase was needed.
HD void u2d_guess_w(const U2d& s, double W_last, double* x2d, double vsq_guess = -1., double factor = 10.,
                    int max_increase = 10) {
  int i_increase = 0;
  while (((W_last*W_last*W_last * (W_last + 2.*s.Bsq)
           - s.QdotBsq*(2.*W_last + s.Bsq)) <= W_last*W_last*(s.Qtsq - s.Bsq*s.Bsq))
         && (i_increase < max_increase)) {
    W_last *= factor;
    i_increase++;
  }
  x2d[0] = fabs(W_last);
  x2d[1] = (vsq_guess >= 0. && i_increase == 0) ? vsq_guess : u2d_x1_of_x0(s, W_last);
}

// The reference's initial guess from the incoming primitives' u and u~
// (src/utoprim_2d_fast.c:302-330).  Returns 2 when their u~^2 is unusable.
template <class GP>
HD int u2d_guess(const U2d& s, GP geom, double gam, const double* prim, double* x2d) {
  double utsq = geom->gcov[1][1]*prim[U1]*prim[U1]
              + geom->gcov[2][2]*prim[U2]*prim[U2]
              + geom->gcov[3][3]*prim[U3]*prim[U3]
              + 2*(geom->gcov[1][2]*prim[U1]*prim[U2]
                 + geom->gcov[1][3]*prim[U1]*prim[U3]
                 + geom->gcov[2][3]*prim[U2]*prim[U3]);
  if ((utsq < 0.) && (fabs(utsq) < 1.0e-13)) utsq = fabs(utsq);
  if (utsq < 0. || utsq > UTSQ_TOO_BIG) return 2;
  double gammasq = 1. + utsq;
  double gamma = sqrt(gammasq);
  double rho0 = s.D / gamma;
  double u = prim[UU];
  double p = (gam - 1.)*u;
  double w = rho0 + u + p;
  u2d_guess_w(s, w*gammasq, x2d);
  return 0;
}

template <class GP>
HD int u2d_setup(const double* U, GP geom, double gam, double gmog, double* prim, U2d& s,
                 U2dAux& aux, double* x2d) {
  int rc = u2d_prepare(U, geom, gmog, prim, s, aux);
  if (rc) return rc;
  return u2d_guess(s, geom, gam, prim, x2d);
}

// Utoprim_new_body() after the Newton solve (src/utoprim_2d_fast.c:345-408).
HD int u2d_finish(const U2d& s, const U2dAux& aux, double gam, const double* x2d, int retval, double* prim,
                  double* gamma_out, double* bsq_out) {
  double W = x2d[0];
  double vsq = x2d[1];
  if ((retval != 0) || (W == FAIL_VAL)) return retval*100 + 1;
  if (W <= 0. || W > W_TOO_BIG) return 3;
  if (vsq >= 1.) return 4;
  double gtmp = sqrt(1. - vsq);
  double gamma = 1./gtmp;
  double rho0 = s.D * gtmp;
  *gamma_out = gamma;
  double w = W * (1. - vsq);
  double p = (gam - 1.)*(w - rho0)/gam;        // pressure_rho0_w, src/u2p_util.c:387-390
  double u = w - (rho0 + p);
  prim[RHO] = rho0;
  prim[UU] = u;
  double g_o_WBsq = gamma/(W + s.Bsq);
  double QdB_o_W = aux.QdotB / W;
  *bsq_out = s.Bsq * (1. - vsq) + QdB_o_W*QdB_o_W;
  prim[U1] = g_o_WBsq * (aux.Qcon[1] + aux.ncon[1] * s.Qdotn + QdB_o_W*aux.Bcon[1]);
  prim[U2] = g_o_WBsq * (aux.Qcon[2] + aux.ncon[2] * s.Qdotn + QdB_o_W*aux.Bcon[2]);
  prim[U3] = g_o_WBsq * (aux.Qcon[3] + aux.ncon[3] * s.Qdotn + QdB_o_W*aux.Bcon[3]);
  return 0;
}

// Utoprim_2d_fast() + Utoprim_new_body(), src/utoprim_2d_fast.c:135-408.
// prim: guess on input; rho,u,u~ replaced only on success, B always set from U
// (as the reference does before the solve).  gam: adiabatic index,
// gmog = (gam-1)/gam (gam_m1_o_gam, src/main.c:299).
// n_extra: confirming iterations after convergence (the reference's
// EXTRA_NEWT_ITER unless the caller asks for fewer).
template <class GP>
HD int utoprim_2d_fast(const double* U, GP geom, double gam, double gmog, double* prim,
                       double* gamma_out, double* bsq_out, int* iters, int n_extra = EXTRA_NEWT_ITER) {
  if (iters) *iters = 0;
  U2d s;
  U2dAux aux;
  double x2d[2];
  int rc = u2d_setup(U, geom, gam, gmog, prim, s, aux, x2d);
  if (rc) return rc;
  int retval = u2d_newton(s, x2d, iters, n_extra);
  return u2d_finish(s, aux, gam, x2d, retval, prim, gamma_out, bsq_out);
}

// ---------------------------------------------------------------------------
// 1D-W Newton fallback, src/utoprim_1d.c:118-629
struct U1d { double Bsq, QdotBsq, Qtsq, Qdotn, D, gam; };

HD double u1d_vsq_calc(const U1d& s, double W) {
  double Wsq = W*W;
  double Xsq = (s.Bsq + W) * (s.Bsq + W);
  return (Wsq * s.Qtsq + s.QdotBsq * (s.Bsq + 2.*W)) / (Wsq*Xsq);
}
HD double u1d_dvsq_dW(const U1d& s, double W) {
  double X = s.Bsq + W;
  double W3 = W*W*W;
  double X3 = X*X*X;
  return -2.*(s.Qtsq/X3 + s.QdotBsq * (3*W*X + s.Bsq*s.Bsq) / (W3 * X3));
}
HD void u1d_func(const U1d& s, const double* x, double* dx, double* f, double* df) {
  const double dv = 1.0e-15;
  double W = x[0];
  double Wsq = W*W;
  double vsq = u1d_vsq_calc(s, W);
  vsq = (vsq < -dv) ? 0. : fabs(vsq);
  vsq = (vsq > 1.) ? (1. - dv) : vsq;
  double gtmp = 1. - vsq;
  double p_tmp = (s.gam - 1.) * (W * gtmp - s.D * sqrt(gtmp)) / s.gam;
  double dvsq = u1d_dvsq_dW(s, W);
  double dp1 = (s.gam - 1.) * (1. - vsq) / s.gam;
  double dp2 = (s.gam - 1.) * (0.5 * s.D / sqrt(1. - vsq) - W) / s.gam;
  double dpdW = dp1 + dp2*dvsq;
  double r0 = + W + 0.5 * s.Bsq * (1. + vsq) - 0.5*s.QdotBsq/Wsq + s.Qdotn - p_tmp;
  double drdW = 1. - dpdW + s.QdotBsq/(Wsq*W) + 0.5*s.Bsq*dvsq;
  dx[0] = -r0/drdW;
  *f = 0.5*r0*r0;
  *df = -2. * (*f);
}
HD int u1d_newton(const U1d& s, double* x) {
  double f = 1., df = 1., dx[1], errx = 1.;
  int n_iter = 0, i_extra = 0, doing_extra = 0, keep_iterating = 1;
  while (keep_iterating) {
    u1d_func(s, x, dx, &f, &df);
    errx = 0.;
    x[0] += dx[0];
    int i_increase = 0;
    while (((x[0]*x[0]*x[0] * (x[0] + 2.*s.Bsq) - s.QdotBsq*(2.*x[0] + s.Bsq))
            <= x[0]*x[0]*(s.Qtsq - s.Bsq*s.Bsq)) && (i_increase < 10)) {
      x[0] -= (1.*i_increase) * dx[0] / 10.;
      i_increase++;
    }
    errx = (x[0] == 0.) ? fabs(dx[0]) : fabs(dx[0]/x[0]);
    x[0] = fabs(x[0]);
    if ((fabs(errx) <= NEWT_TOL) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) doing_extra = 1;
    if (doing_extra == 1) i_extra++;
    if (((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
        (n_iter >= (MAX_NEWT_ITER - 1))) keep_iterating = 0;
    n_iter++;
  }
  if ((!is_finite(f)) || (!is_finite(df))) return 2;
  if (fabs(errx) > MIN_NEWT_TOL) return 1;
  return 0;
}

// Utoprim_1d(), src/utoprim_1d.c:118-343
HD int utoprim_1d(const double* U, const Geom* geom, double gam, double* prim,
                  double* gamma_out, double* bsq_out) {
  const double gdet = geom->g;
  for (int i = B1; i <= B3; i++) prim[i] = U[i] / gdet;
  double alpha = 1.0/sqrt(-geom->gcon[0][0]);
  double Ut[NP];
  Ut[RHO] = alpha * U[RHO] / gdet;
  Ut[UU] = alpha * (U[UU] - U[RHO]) / gdet;
  for (int i = U1; i <= U3; i++) Ut[i] = alpha * U[i] / gdet;
  for (int i = B1; i <= B3; i++) Ut[i] = alpha * U[i] / gdet;
  double pt[5];
  for (int i = 0; i < B1; i++) pt[i] = prim[i];

  U1d s;
  s.gam = gam;
  double Bcon[4] = {0., Ut[B1], Ut[B2], Ut[B3]}, Bcov[4];
  // lower_g(), src/u2p_util.c:238-263 (same sums as lower())
  for (int i = 0; i < 4; i++)
    Bcov[i] = geom->gcov[i][0]*Bcon[0] + geom->gcov[i][1]*Bcon[1] + geom->gcov[i][2]*Bcon[2] + geom->gcov[i][3]*Bcon[3];
  double Qcov[4] = {Ut[UU], Ut[U1], Ut[U2], Ut[U3]}, Qcon[4];
  raise_v(Qcov, geom, Qcon);
  s.Bsq = 0.;
  for (int i = 1; i < 4; i++) s.Bsq += Bcon[i]*Bcov[i];
  double QdotB = 0.;
  for (int i = 0; i < 4; i++) QdotB += Qcov[i]*Bcon[i];
  s.QdotBsq = QdotB*QdotB;
  // ncov_calc(), raise_g(): n_mu = (-lapse,0,0,0), lapse = sqrt(-1/gcon00)
  double lapse = sqrt(-1./geom->gcon[0][0]);
  double ncov[4] = {-lapse, 0., 0., 0.}, ncon[4];
  raise_v(ncov, geom, ncon);
  s.Qdotn = Qcon[0]*ncov[0];
  double Qsq = 0.;
  for (int i = 0; i < 4; i++) Qsq += Qcov[i]*Qcon[i];
  s.Qtsq = Qsq + s.Qdotn*s.Qdotn;
  s.D = Ut[RHO];
  double utsq = 0.;
  for (int i = 1; i < 4; i++)
    for (int j = 1; j < 4; j++) utsq += geom->gcov[i][j]*pt[U1+i-1]*pt[U1+j-1];
  if ((utsq < 0.) && (fabs(utsq) < 1.0e-13)) utsq = fabs(utsq);
  if (utsq < 0. || utsq > UTSQ_TOO_BIG) return 2;
  double gammasq = 1. + utsq;
  double gamma = sqrt(gammasq);
  double rho0 = s.D / gamma;
  double u = pt[UU];
  double p = (gam - 1.)*u;
  double w = rho0 + u + p;
  double W_last = w*gammasq;
  int i_increase = 0;
  while (((W_last*W_last*W_last * (W_last + 2.*s.Bsq)
           - s.QdotBsq*(2.*W_last + s.Bsq)) <= W_last*W_last*(s.Qtsq - s.Bsq*s.Bsq))
         && (i_increase < 10)) {
    W_last *= 10.;
    i_increase++;
  }
  double x1d[1] = {W_last};
  int retval = u1d_newton(s, x1d);
  double W = x1d[0];
  if ((retval != 0) || (W == FAIL_VAL)) return retval*100 + 1;
  if (W <= 0. || W > W_TOO_BIG) return 3;
  double vsq = u1d_vsq_calc(s, W);
  if (vsq >= 1.) return 4;
  double gtmp = sqrt(1. - vsq);
  gamma = 1./gtmp;
  rho0 = s.D * gtmp;
  *gamma_out = gamma;
  w = W * (1. - vsq);
  p = (gam - 1.)*(w - rho0)/gam;
  u = w - (rho0 + p);
  prim[RHO] = rho0;
  prim[UU] = u;
  *bsq_out = s.Bsq * (1. - vsq) + s.QdotBsq / (W*W);
  double Qtcon[4];
  for (int i = 1; i < 4; i++) Qtcon[i] = Qcon[i] + ncon[i] * s.Qdotn;
  for (int i = 1; i < 4; i++) prim[U1+i-1] = gamma/(W + s.Bsq) * (Qtcon[i] + QdotB*Bcon[i]/W);
  return 0;
}

// ---------------------------------------------------------------------------
// Entropy-equation inversions, src/utoprim_1d_ee.c:156-847, src/utoprim_1d_ee2.c:152-590
struct Uee {
  double Bsq, QdotBsq, Qtsq, Qdotn, D, Sc, rho_gm1, two_Bsq;
  double W_for_gnr2, rho_for_gnr2, W_for_gnr2_old, rho_for_gnr2_old, drho_dW;
  double g_o_gm1, gm1, gm1_inv, gam;
  double t2, t4, t7, t24, t300, t400, s100, s200;
};

// func_rho() of src/utoprim_1d_ee.c:818-838; like the reference it leaves the
// residual norms f, df untouched (they stay 1, so only a non-finite x stops
// the solver as divergent)
HD void uee_func_rho(Uee& s, const double* x, double* dx, double*, double*) {
  double rho = x[0];
  double W = s.W_for_gnr2;
  double t40 = pow(rho, s.gm1);
  s.rho_gm1 = t40;
  double r0 = (rho*W + (-t40*s.s100 - s.D)*s.D);
  double t14 = t40/rho;
  double j00 = -t14*s.s200 + W;
  dx[0] = -r0/j00;
}

// gnr2(), src/utoprim_1d_ee.c:628-700 (never enters the extra-iteration mode)
HD int uee_gnr2(Uee& s, double* x) {
  double f = 1., df = 1., dx[1], errx = 1.;
  int n_iter = 0, i_extra = 0, doing_extra = 0, keep_iterating = 1;
  while (keep_iterating) {
    uee_func_rho(s, x, dx, &f, &df);
    x[0] += dx[0];
    errx = (x[0] == 0.) ? fabs(dx[0]) : fabs(dx[0]/x[0]);
    x[0] = fabs(x[0]);
    if (doing_extra == 1) i_extra++;
    if (((fabs(errx) <= NEWT_TOL2) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
        (n_iter >= (MAX_NEWT_ITER - 1))) keep_iterating = 0;
    n_iter++;
  }
  if ((!is_finite(f)) || (!is_finite(df)) || (!is_finite(x[0]))) return 2;
  if (fabs(errx) <= NEWT_TOL) return 0;
  else if ((fabs(errx) <= MIN_NEWT_TOL) && (fabs(errx) > NEWT_TOL)) return 0;
  return 1;
}

// func_W(), src/utoprim_1d_ee.c:742-795 (f, df untouched, as func_rho())
HD void uee_func_W(Uee& s, const double* x, double* dx, double*, double*) {
  double W = x[0];
  s.W_for_gnr2_old = s.W_for_gnr2;
  s.W_for_gnr2 = W;
  double x_rho[1];
  double rho_g = x_rho[0] = s.rho_for_gnr2;
  int ntries = 0;
  while ((uee_gnr2(s, x_rho)) && (ntries++ < 10)) {
    rho_g *= 10.;
    x_rho[0] = rho_g;
  }
  s.rho_for_gnr2_old = s.rho_for_gnr2;
  double rho = s.rho_for_gnr2 = x_rho[0];
  s.rho_gm1 = pow(rho, s.gm1);
  s.drho_dW = -rho*rho/(-s.rho_gm1*s.s200 + W*rho);
  double t15 = -(s.D - rho)*(s.D + rho);
  double t200 = W + s.two_Bsq;
  double t1000 = rho*s.drho_dW;
  double r0 = (s.t300 + (s.t4 + s.t4 + (s.t400 + t15*(s.t7 + (t200)*W))*W)*W)*s.t24;
  double j00 = 2*(s.t4 + (s.t400 + t15*s.t7 + (3.0*t15*s.Bsq + s.t7*t1000 + (t15 + t15 + t1000*(t200))*W)*W)*W)*s.t24;
  dx[0] = -r0/j00;
}

// general_newton_raphson(), src/utoprim_1d_ee.c:534-618
HD int uee_newton(Uee& s, double* x) {
  double f = 1., df = 1., dx[1], errx = 1.;
  int n_iter = 0, i_extra = 0, doing_extra = 0, keep_iterating = 1;
  while (keep_iterating) {
    uee_func_W(s, x, dx, &f, &df);
    errx = 0.;
    x[0] += dx[0];
    errx = (x[0] == 0.) ? fabs(dx[0]) : fabs(dx[0]/x[0]);
    x[0] = fabs(x[0]);
    if ((fabs(errx) <= NEWT_TOL) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) doing_extra = 1;
    if (doing_extra == 1) i_extra++;
    if (((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
        (n_iter >= (MAX_NEWT_ITER - 1))) keep_iterating = 0;
    n_iter++;
  }
  if ((!is_finite(f)) || (!is_finite(df)) || (!is_finite(x[0]))) return 2;
  if (fabs(errx) <= NEWT_TOL) return 0;
  else if ((fabs(errx) <= MIN_NEWT_TOL) && (fabs(errx) > NEWT_TOL)) return 0;
  return 1;
}

// Utoprim_1d_ee(), src/utoprim_1d_ee.c:156-478
HD int utoprim_1d_ee(const double* U, double S, const Geom* geom, double gam, double* prim, double* gamma_out) {
  if (U[0] <= 0.) return -100;
  Uee s;
  s.gam = gam;
  s.gm1 = (gam - 1.);
  s.gm1_inv = 1./s.gm1;
  s.g_o_gm1 = gam*s.gm1_inv;
  double inv_gdet = geom->g_inv;
  double alpha = geom->alpha;
  double geomfactor = alpha * inv_gdet;
  prim[B1] = U[B1] * inv_gdet;
  prim[B2] = U[B2] * inv_gdet;
  prim[B3] = U[B3] * inv_gdet;
  s.Sc = geomfactor * S;
  double Ut[NP];
  Ut[RHO] = geomfactor * U[RHO];
  Ut[UU] = geomfactor * (U[UU] - U[RHO]);
  Ut[U1] = geomfactor * U[U1];
  Ut[U2] = geomfactor * U[U2];
  Ut[U3] = geomfactor * U[U3];
  Ut[B1] = geomfactor * U[B1];
  Ut[B2] = geomfactor * U[B2];
  Ut[B3] = geomfactor * U[B3];

  double Bcon[4] = {0., Ut[B1], Ut[B2], Ut[B3]};
  s.Bsq = geom->gcov[1][1]*Bcon[1]*Bcon[1]
        + geom->gcov[2][2]*Bcon[2]*Bcon[2]
        + geom->gcov[3][3]*Bcon[3]*Bcon[3]
        + 2*(geom->gcov[1][2]*Bcon[1]*Bcon[2]
           + geom->gcov[1][3]*Bcon[1]*Bcon[3]
           + geom->gcov[2][3]*Bcon[2]*Bcon[3]);
  double Qcov[4] = {Ut[UU], Ut[U1], Ut[U2], Ut[U3]}, Qcon[4];
  raise_v(Qcov, geom, Qcon);
  double QdotB = Qcov[1]*Bcon[1] + Qcov[2]*Bcon[2] + Qcov[3]*Bcon[3];
  s.QdotBsq = QdotB*QdotB;
  s.Qdotn = -alpha * Qcon[0];
  double Qsq = dot4(Qcov, Qcon);
  s.Qtsq = Qsq + s.Qdotn*s.Qdotn;
  s.D = Ut[RHO];
  s.t2 = s.D*s.D;
  s.t4 = s.QdotBsq*s.t2;
  s.t7 = s.Bsq*s.Bsq;
  s.t24 = 1/s.t2;
  s.two_Bsq = s.Bsq + s.Bsq;
  s.t300 = s.QdotBsq*s.Bsq*s.t2;
  s.t400 = s.Qtsq*s.t2;
  s.s200 = s.D*gam*s.Sc;
  s.s100 = s.g_o_gm1*s.Sc;

  double utsq = geom->gcov[1][1]*prim[U1]*prim[U1]
              + geom->gcov[2][2]*prim[U2]*prim[U2]
              + geom->gcov[3][3]*prim[U3]*prim[U3]
              + 2*(geom->gcov[1][2]*prim[U1]*prim[U2]
                 + geom->gcov[1][3]*prim[U1]*prim[U3]
                 + geom->gcov[2][3]*prim[U2]*prim[U3]);
  if ((utsq < 0.) && (fabs(utsq) < 1.0e-13)) utsq = fabs(utsq);
  if (utsq < 0. || utsq > UTSQ_TOO_BIG) return 2;
  double gammasq = 1. + utsq;
  double gamma = sqrt(gammasq);
  double rho0 = s.D / gamma;
  s.rho_gm1 = pow(rho0, s.gm1);
  double p = s.Sc * s.rho_gm1 / gamma;
  double u = p * s.gm1_inv;
  double w = rho0 + u + p;
  double W_last = w*gammasq;
  int i_increase = 0;
  while (((W_last*W_last*W_last * (W_last + 2.*s.Bsq)
           - s.QdotBsq*(2.*W_last + s.Bsq)) <= W_last*W_last*(s.Qtsq - s.Bsq*s.Bsq))
         && (i_increase < 10)) {
    W_last *= 10.;
    i_increase++;
  }
  s.W_for_gnr2 = s.W_for_gnr2_old = W_last;
  s.rho_for_gnr2 = s.rho_for_gnr2_old = rho0;
  double x1d[1] = {W_last};
  int retval = uee_newton(s, x1d);
  double W = x1d[0];
  if ((retval != 0) || (W == FAIL_VAL)) return retval*100 + 1;
  if (W <= 0. || W > W_TOO_BIG) return 3;
  double x_rho[1];
  double rho_g = x_rho[0] = s.rho_for_gnr2;
  int ntries = 0;
  while ((retval = uee_gnr2(s, x_rho)) && (ntries++ < 10)) {
    rho_g *= 10.;
    x_rho[0] = rho_g;
  }
  s.rho_for_gnr2 = x_rho[0];
  if (retval != 0) return 10;
  rho0 = s.rho_for_gnr2;
  s.rho_gm1 = pow(rho0, s.gm1);
  utsq = (s.D - rho0)*(s.D + rho0)/(rho0*rho0);
  double gamma_sq = 1. + utsq;
  gamma = sqrt(gamma_sq);
  *gamma_out = gamma;
  if (utsq < 0.) return 4;
  w = W / gamma_sq;
  p = s.Sc * s.rho_gm1 / gamma;
  u = p * s.gm1_inv;
  prim[RHO] = rho0;
  prim[UU] = u;
  double g_o_WBsq = gamma/(W + s.Bsq);
  double QdB_o_W = QdotB / W;
  prim[U1] = g_o_WBsq * (Qcon[1] + geom->ncon[1] * s.Qdotn + QdB_o_W*Bcon[1]);
  prim[U2] = g_o_WBsq * (Qcon[2] + geom->ncon[2] * s.Qdotn + QdB_o_W*Bcon[2]);
  prim[U3] = g_o_WBsq * (Qcon[3] + geom->ncon[3] * s.Qdotn + QdB_o_W*Bcon[3]);
  return 0;
}

// Utoprim_1d_ee2(), src/utoprim_1d_ee2.c:152-590 (1D Newton in rho)
struct Uee2 { double Bsq, QdotBsq, Qtsq, Qdotn, D, Sc, g_o_gm1, gam; };

HD void uee2_func_rho(const Uee2& s, const double* x, double* dx, double* f, double* df) {
  double rho = x[0];
  double rhosq = rho*rho;
  double t200 = 1./(s.D*s.D);
  double t100 = s.g_o_gm1*s.Sc*pow(rho, (s.gam - 1.));
  double W = s.D*(s.D + t100) / rho;
  double dWdrho = s.D * (-s.D + t100*(s.gam - 2.)) / rhosq;
  double t1 = W*W;
  double t2 = s.Bsq + W;
  double t3 = t2*t2;
  double vsq = (s.D - rho)*(s.D + rho)*t200;
  double dvsqdrho = -2*rho*t200;
  double r0 = t1*(s.Qtsq - vsq*t3) + s.QdotBsq*(t2 + W);
  double t12 = s.Bsq*s.Bsq;
  double t17 = dWdrho*vsq;
  double j00 = 2*s.QdotBsq*dWdrho
    + ((s.Qtsq - vsq*t12)*2*dWdrho + (-6*t17*s.Bsq - dvsqdrho*t12
       + (-2*dvsqdrho*s.Bsq - 4*t17 - dvsqdrho*W)*W)*W)*W;
  dx[0] = -r0/j00;
  *f = 0.5*r0*r0;
  *df = -2. * (*f);
}

HD int uee2_newton(const Uee2& s, double* x) {
  double f = 1., df = 1., dx[1], errx = 1.;
  int n_iter = 0, i_extra = 0, doing_extra = 0, keep_iterating = 1;
  while (keep_iterating) {
    uee2_func_rho(s, x, dx, &f, &df);
    errx = 0.;
    x[0] += dx[0];
    errx = (x[0] == 0.) ? fabs(dx[0]) : fabs(dx[0]/x[0]);
    x[0] = fabs(x[0]);
    if ((fabs(errx) <= NEWT_TOL) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) doing_extra = 1;
    if (doing_extra == 1) i_extra++;
    if (((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
        (n_iter >= (MAX_NEWT_ITER - 1))) keep_iterating = 0;
    n_iter++;
  }
  if ((is_finite(f) == 0) || (is_finite(df) == 0) || (is_finite(x[0]) == 0)) return 2;
  if (fabs(errx) > MIN_NEWT_TOL) return 1;
  return 0;
}

HD int utoprim_1d_ee2(const double* U, double S, const Geom* geom, double gam, double* prim, double* gamma_out) {
  if (U[0] <= 0.) return -100;
  Uee2 s;
  s.gam = gam;
  s.g_o_gm1 = gam/(gam - 1.);
  double inv_gdet = 1./geom->g;
  prim[B1] = U[B1] * inv_gdet;
  prim[B2] = U[B2] * inv_gdet;
  prim[B3] = U[B3] * inv_gdet;
  double alpha = geom->alpha;
  double geomfactor = alpha*inv_gdet;
  s.Sc = geomfactor * S;
  double Ut[NP];
  Ut[RHO] = geomfactor * U[RHO];
  Ut[UU] = geomfactor * (U[UU] - U[RHO]);
  Ut[U1] = geomfactor * U[U1];
  Ut[U2] = geomfactor * U[U2];
  Ut[U3] = geomfactor * U[U3];
  Ut[B1] = geomfactor * U[B1];
  Ut[B2] = geomfactor * U[B2];
  Ut[B3] = geomfactor * U[B3];
  double Bcon[4] = {0., Ut[B1], Ut[B2], Ut[B3]}, Bcov[4];
  lower(Bcon, geom, Bcov);
  double Qcov[4] = {Ut[UU], Ut[U1], Ut[U2], Ut[U3]}, Qcon[4];
  raise_v(Qcov, geom, Qcon);
  s.Bsq = Bcon[1]*Bcov[1] + Bcon[2]*Bcov[2] + Bcon[3]*Bcov[3];
  double QdotB = Qcov[1]*Bcon[1] + Qcov[2]*Bcon[2] + Qcov[3]*Bcon[3];
  s.QdotBsq = QdotB*QdotB;
  double ncov0 = -geom->alpha;
  double ncon[4];
  ncon[0] = -geom->alpha * geom->gcon[0][0];
  ncon[1] = -geom->alpha * geom->gcon[0][1];
  ncon[2] = -geom->alpha * geom->gcon[0][2];
  ncon[3] = -geom->alpha * geom->gcon[0][3];
  s.Qdotn = Qcon[0]*ncov0;
  double Qsq = dot4(Qcov, Qcon);
  s.Qtsq = Qsq + s.Qdotn*s.Qdotn;
  s.D = Ut[RHO];
  double utsq = geom->gcov[1][1]*prim[U1]*prim[U1]
              + geom->gcov[2][2]*prim[U2]*prim[U2]
              + geom->gcov[3][3]*prim[U3]*prim[U3]
              + 2*(geom->gcov[1][2]*prim[U1]*prim[U2]
                 + geom->gcov[1][3]*prim[U1]*prim[U3]
                 + geom->gcov[2][3]*prim[U2]*prim[U3]);
  if ((utsq < 0.) && (fabs(utsq) < 1.0e-13)) utsq = fabs(utsq);
  if (utsq < 0. || utsq > UTSQ_TOO_BIG) return 2;
  double gammasq = 1. + utsq;
  double gamma = sqrt(gammasq);
  double rho0 = s.D / gamma;
  double x1d[1] = {rho0};
  int retval = uee2_newton(s, x1d);
  rho0 = x1d[0];
  if ((retval != 0) || (rho0 == FAIL_VAL)) return retval*100 + 1;
  if (rho0 > W_TOO_BIG) return 3;
  utsq = (s.D - rho0)*(s.D + rho0)/(rho0*rho0);
  double gamma_sq = 1. + utsq;
  gamma = sqrt(gamma_sq);
  *gamma_out = gamma;
  if (utsq < 0.) return 4;
  double rho_gm1 = pow(rho0, (gam - 1.));
  double p = s.Sc * rho_gm1 / gamma;
  double u = p / (gam - 1.);
  double w = rho0 + u + p;
  double W = w * gamma_sq;
  prim[RHO] = rho0;
  prim[UU] = u;
  double g_o_WBsq = gamma/(W + s.Bsq);
  double QdB_o_W = QdotB / W;
  prim[U1] = g_o_WBsq * (Qcon[1] + ncon[1] * s.Qdotn + QdB_o_W*Bcon[1]);
  prim[U2] = g_o_WBsq * (Qcon[2] + ncon[2] * s.Qdotn + QdB_o_W*Bcon[2]);
  prim[U3] = g_o_WBsq * (Qcon[3] + ncon[3] * s.Qdotn + QdB_o_W*Bcon[3]);
  return 0;
}

// check_gamma(), src/fixup.c:133-141
HD int check_gamma(double gamma) {
  if (gamma <= GAMMAMAX) return 0;
  if (gamma > GAMMAMAX2) return 2;
  return 1;
}

// --------------------------------------------------------------------------
// small helpers

HD double dmax(double a, double b) { return (a > b) ? a : b; }   // MAX(), src/macros.h:474
HD int imin(int a, int b) { return (a < b) ? a : b; }

HD unsigned long long dbl_bits(double x) {
  union { double d; unsigned long long u; } c;
  c.d = x;
  return c.u;
}
HD double bits_dbl(unsigned long long u) {
  union { double d; unsigned long long u; } c;
  c.u = u;
  return c.d;
}

// Index of the x1 rescaling factor per variable: rho -> r^1.5, u -> r, rest -> r^3
// (rscale_prims_r(), src/recon.c:1173-1189).
HD int rs_slot(int v) { return (v == RHO) ? 0 : ((v == UU) ? 2 : 1); }

// Value of a primitive after the in-place x1 rescale/unscale round trip that
// reconstruct_fast() applies to every zone (src/recon.c:996, :1008, :1173-1210).
// Without RESCALE_R the tables hold exact ones, so the products return x
// unchanged; applying them unconditionally keeps the kernels free of a
// branch per loaded value, which would make each load wait for the last.
HD double rt_val(const Params& P, const Ptrs& X, int v, int i, double x) {
  int s = rs_slot(v);
  return (x * X.rscale[s*P.n1t + i]) * X.irscale[s*P.n1t + i];
}

// Tally of fail() calls of one zone; written out by the kernel wrapper.
struct FailRec {
  int nf[N_NFAIL];
  int pflag;        // last pflag response (0: none)
  int boundary;     // check_boundary_pflag(): bit 0 a physical boundary, bit 1 one shared with another block
  double vmax;      // fastest characteristic speed of an interior face (0: none)
  HD void clear() {
    for (int l = 0; l < N_NFAIL; l++) nf[l] = 0;
    pflag = 0;
    boundary = 0;
    vmax = 0.;
  }
};

// fail(type, state) bookkeeping, src/diag.c:159-300: state 1 -> INTERP_PRIM,
// state 2 -> INTERP_V; check_boundary_pflag(), src/harm_mpi.c:1388-1430.
HD void fail_set(const Params& P, FailRec& r, int nfail_id, int state, int i, int j, int k) {
  r.nf[nfail_id]++;
  if (state) {
    r.pflag = (state == 1) ? PFLAG_INTERP_PRIM : PFLAG_INTERP_V;
    // within NG zones of a face: boundary_phys_pflag, or boundary_mpi_pflag
    // when the face borders another block
    const int idx[3] = {i, j, k}, n[3] = {P.n1, P.n2, P.n3};
    for (int d = 0; d < 3; d++) {
      int side = -1;
      if (idx[d] < NG + NG) side = 0;
      else if (idx[d] > n[d] - 1) side = 1;   // > N_dE - NG with N_dE = n + NG - 1
      if (side >= 0) r.boundary |= ((P.bc_nbr >> (2*d + side)) & 1) ? 2 : 1;
    }
  }
}

HD const Geom* geom_at(const Params& P, const Ptrs& X, int i, int j, int k, int pos) {
  return X.geom + geom_id(P, i, j, k, pos);
}

}  // namespace hgpu
