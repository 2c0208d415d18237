// hgpu_pn.h -- time-dependent metric of an inspiralling black-hole binary for
// the circumbinary-disk configuration (host/device bodies).
//
// The reference declares its post-Newtonian binary metrics
// (dyn_35pn_nz_gcov_func, dyn_fullpn_nz_gcov_func, src/metric.h:104-171) but
// ships no body for them, so this is written from the published method:
//
// * Metric: the 1PN near-zone metric of two point masses in harmonic
//   coordinates (Blanchet, Faye & Ponsot 1998, Phys. Rev. D 58, 124002), G = c = 1:
//     g00 = -1 + 2V - 2U^2,  g0i = -4 V_i,  gij = delta_ij (1 + 2U),
//   with U = sum_a m_a/r_a the Newtonian potential (U^2 and V^2 differ beyond
//   1PN order) and V the 1PN retarded potential of the two bodies,
//     V = sum_a m_a/r_a [1 + 2 v_a^2 - (n_a.v_a)^2/2] - m_a m_b/(r_a r_ab)
//         + m_a m_b (r_b^2 - r_a^2 - r_ab^2)/(4 r_a r_ab^3)     (b != a),
//   and V_i = sum_a m_a v_a^i / r_a  (y_b at distance r_b from the field point).
// * Orbit: quasi-circular, centre of mass at rest at the origin, Newtonian
//   Kepler frequency omega = sqrt(M/b^3), separation shrinking by the leading
//   (2.5PN) radiation reaction, db/dt = -(64/5) m1 m2 M / b^3 (Peters 1964):
//     b(t)^4 = b0^4 - (256/5) m1 m2 M t,
//     phi(t) = phi0 + (8 sqrt(M) / (5 k)) (b0^(5/2) - b(t)^(5/2)),  k = (256/5) m1 m2 M.
// * Coordinates: spherical polar about the centre of mass, numerical
//   coordinates X1 = log(r - R0), X2 in [0,1] with
//   theta = th_length X2 + th_beg + h_slope sin(2 pi X2) (the modified
//   Kerr-Schild map, src/coord.c:240-382), X3 = phi.  g in numerical
//   coordinates is J^T g_harmonic J with J = d(t, x, y, z)/d(t, X1, X2, X3),
//   which for this conformally flat spatial metric and orthogonal map has a
//   closed form (pn_gcov_num()).
// * Everything is templated on the scalar: double for values, Dual4 for values
//   and the four partial derivatives d/d(t, X1, X2, X3) (forward-mode
//   differentiation), from which the connection Gamma^a_bc, including the
//   time derivatives of the moving metric, is formed exactly.
#pragma once
#include "hgpu_types.h"
#include <math.h>

namespace hgpu {

// ---------------------------------------------------------------------------
// forward-mode numbers in (t, X1, X2, X3); their functions are found by
// argument-dependent lookup, so double arguments keep the libm functions
namespace ad {
struct Dual4 {
  double v, d[4];
};
HD Dual4 dc(double v) { Dual4 r; r.v = v; r.d[0] = r.d[1] = r.d[2] = r.d[3] = 0.; return r; }
HD Dual4 dvar(double v, int k) { Dual4 r = dc(v); r.d[k] = 1.; return r; }
HD Dual4 operator+(const Dual4& a, const Dual4& b) {
  Dual4 r; r.v = a.v + b.v; for (int k = 0; k < 4; k++) r.d[k] = a.d[k] + b.d[k]; return r;
}
HD Dual4 operator-(const Dual4& a, const Dual4& b) {
  Dual4 r; r.v = a.v - b.v; for (int k = 0; k < 4; k++) r.d[k] = a.d[k] - b.d[k]; return r;
}
HD Dual4 operator-(const Dual4& a) {
  Dual4 r; r.v = -a.v; for (int k = 0; k < 4; k++) r.d[k] = -a.d[k]; return r;
}
HD Dual4 operator*(const Dual4& a, const Dual4& b) {
  Dual4 r; r.v = a.v * b.v; for (int k = 0; k < 4; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k]; return r;
}
HD Dual4 operator/(const Dual4& a, const Dual4& b) {
  const double ib = 1. / b.v;
  Dual4 r; r.v = a.v * ib;
  for (int k = 0; k < 4; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
  return r;
}
HD Dual4 operator+(const Dual4& a, double s) { Dual4 r = a; r.v += s; return r; }
HD Dual4 operator+(double s, const Dual4& a) { return a + s; }
HD Dual4 operator-(const Dual4& a, double s) { Dual4 r = a; r.v -= s; return r; }
HD Dual4 operator-(double s, const Dual4& a) { return (-a) + s; }
HD Dual4 operator*(const Dual4& a, double s) {
  Dual4 r; r.v = a.v * s; for (int k = 0; k < 4; k++) r.d[k] = a.d[k] * s; return r;
}
HD Dual4 operator*(double s, const Dual4& a) { return a * s; }
HD Dual4 operator/(const Dual4& a, double s) { return a * (1. / s); }
HD Dual4 operator/(double s, const Dual4& a) { return dc(s) / a; }
HD Dual4 chain(const Dual4& a, double f, double df) {
  Dual4 r; r.v = f; for (int k = 0; k < 4; k++) r.d[k] = df * a.d[k]; return r;
}
HD Dual4 sqrt(const Dual4& a) { const double s = ::sqrt(a.v); return chain(a, s, 0.5 / s); }
HD Dual4 sin(const Dual4& a) { return chain(a, ::sin(a.v), ::cos(a.v)); }
HD Dual4 cos(const Dual4& a) { return chain(a, ::cos(a.v), -::sin(a.v)); }
HD Dual4 exp(const Dual4& a) { const double e = ::exp(a.v); return chain(a, e, e); }
HD Dual4 pow(const Dual4& a, double p) { const double f = ::pow(a.v, p); return chain(a, f, p * f / a.v); }
}  // namespace ad
using ad::Dual4;
using ad::dc;
using ad::dvar;
HD double value(double x) { return x; }
HD double value(const Dual4& x) { return x.v; }
HD void set_value(double& x, double v) { x = v; }
HD void set_value(Dual4& x, double v) { x.v = v; }

// seesaw(x, 0, 1) of the theta coordinate, src/coord.c:4316-4324: ghost zones
// beyond a pole reflect back into [0, 1]
HD double seesaw_unit(double x) {
  const double x0 = 0., xn = 1., xn2 = 2.*xn - x0;
  const int shift = (x < x0) ? (1 + (int)(fabs(x - x0)/(xn2 - x0))) : (-(int)(fabs(x - x0)/(xn2 - x0)));
  const double x1 = x + (xn2 - x0)*shift;
  return (x1 < xn) ? (x1 - x0) : (xn2 - x1);
}

// ---------------------------------------------------------------------------

// Positions and velocities of the two holes at time t.
template <class T>
HD void pn_orbit(const PnBinary& B, const T& t, T y1[3], T y2[3], T v1[3], T v2[3]) {
  const double M = B.m1 + B.m2;
  const double k = (256. / 5.) * B.m1 * B.m2 * M;
  T b, bdot, phi, omega;
  if (B.inspiral && k > 0.) {
    const T b4 = B.b0 * B.b0 * B.b0 * B.b0 - k * t;
    b = pow(b4, 0.25);
    bdot = -(k / 4.) / (b * b * b);
    phi = B.phi0 + (8. * ::sqrt(M) / (5. * k)) * (::pow(B.b0, 2.5) - pow(b4, 0.625));
    omega = ::sqrt(M) / pow(b4, 0.375);
  } else {
    b = t * 0. + B.b0;
    bdot = t * 0.;
    omega = t * 0. + ::sqrt(M / (B.b0 * B.b0 * B.b0));
    phi = B.phi0 + omega * t;
  }
  const T c = cos(phi), s = sin(phi);
  // relative separation vector y1 - y2 = b (c, s, 0); centre of mass at rest
  const double f1 = B.m2 / M, f2 = -B.m1 / M;
  const T rx = b * c, ry = b * s;
  const T vx = bdot * c - b * omega * s, vy = bdot * s + b * omega * c;
  y1[0] = f1 * rx; y1[1] = f1 * ry; y1[2] = t * 0.;
  y2[0] = f2 * rx; y2[1] = f2 * ry; y2[2] = t * 0.;
  v1[0] = f1 * vx; v1[1] = f1 * vy; v1[2] = t * 0.;
  v2[0] = f2 * vx; v2[1] = f2 * vy; v2[2] = t * 0.;
}

// Positions and velocities of the holes at one time level (pn_orbit()) and
// their time derivatives, kept so that metric evaluations at that level need
// not recompute the orbit.  The first 12 values are the positions and
// velocities.
struct PnOrbit {
  double y[2][3], v[2][3];
  double ydot[2][3], vdot[2][3];
};

HD void pn_orbit_at(const PnBinary& B, double t, PnOrbit* o) {
  pn_orbit(B, t, o->y[0], o->y[1], o->v[0], o->v[1]);
  Dual4 yd[2][3], vd[2][3];
  pn_orbit(B, dvar(t, 0), yd[0], yd[1], vd[0], vd[1]);
  for (int a = 0; a < 2; a++)
    for (int i = 0; i < 3; i++) { o->ydot[a][i] = yd[a][i].d[0]; o->vdot[a][i] = vd[a][i].d[0]; }
}

// Potentials of the 1PN near-zone metric at Cartesian point x for holes at y
// moving with v: the Newtonian U, the 1PN V and the vector potential V_i, so
// that g00 = -1 + 2V - 2U^2, g0i = -4 V_i, gij = (1 + 2U) delta_ij.
template <class T>
HD void pn_potentials_yv(const PnBinary& B, const T y[2][3], const T v[2][3], const T x[3], T& U, T& V, T Vi[3]) {
  const double m[2] = {B.m1, B.m2};
  T r[2], nv[2], vsq[2];
  for (int a = 0; a < 2; a++) {
    const T dx = x[0] - y[a][0], dy = x[1] - y[a][1], dz = x[2] - y[a][2];
    r[a] = sqrt(dx * dx + dy * dy + dz * dz);
    nv[a] = (dx * v[a][0] + dy * v[a][1] + dz * v[a][2]) / r[a];
    vsq[a] = v[a][0] * v[a][0] + v[a][1] * v[a][1] + v[a][2] * v[a][2];
  }
  const T ex = y[0][0] - y[1][0], ey = y[0][1] - y[1][1], ez = y[0][2] - y[1][2];
  const T r12 = sqrt(ex * ex + ey * ey + ez * ez);
  const T r12_3 = r12 * r12 * r12;
  U = r12 * 0.;
  V = r12 * 0.;
  for (int a = 0; a < 2; a++) {
    const int b = 1 - a;
    const T mr = m[a] / r[a];
    U = U + mr;
    V = V + mr * (1. + 2. * vsq[a] - 0.5 * nv[a] * nv[a]) - m[a] * m[b] / (r[a] * r12)
          + (m[a] * m[b] / 4.) * (r[b] * r[b] - r[a] * r[a] - r12 * r12) / (r[a] * r12_3);
  }
  for (int i = 0; i < 3; i++) {
    Vi[i] = r12 * 0.;
    for (int a = 0; a < 2; a++) Vi[i] = Vi[i] + m[a] * v[a][i] / r[a];
  }
}

// The potentials at Cartesian point x, time t.
template <class T>
HD void pn_potentials(const PnBinary& B, const T& t, const T x[3], T& U, T& V, T Vi[3]) {
  T y[2][3], v[2][3];
  pn_orbit(B, t, y[0], y[1], v[0], v[1]);
  pn_potentials_yv<T>(B, y, v, x, U, V, Vi);
}

// 1PN near-zone metric in harmonic coordinates at Cartesian point x, time t.
template <class T>
HD void pn_gcov_harmonic(const PnBinary& B, const T& t, const T x[3], T g[4][4]) {
  T U, V, Vi[3];
  pn_potentials(B, t, x, U, V, Vi);
  g[0][0] = -1. + 2. * V - 2. * U * U;
  for (int i = 0; i < 3; i++) g[0][i + 1] = g[i + 1][0] = -4. * Vi[i];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) g[i + 1][j + 1] = (i == j) ? 1. + 2. * U : t * 0.;
}

// Metric in the numerical coordinates (t, X1, X2, X3) of the spherical grid.
// The spatial metric is conformally flat, (1 + 2U) delta_ij, and the map to
// (r, theta, phi) is orthogonal, so J^T g J has the closed form
//   g_tt = g00,  g_tn = -4 V_i J^i_n,
//   g_11 = A r'^2,  g_22 = A (r theta')^2,  g_33 = A (r sin theta)^2,  A = 1 + 2U,
// with no other spatial components (r' = dr/dX1, theta' = dtheta/dX2).
// pot(x, U, V, Vi) evaluates the potentials at the Cartesian point x.
// The coordinate functions of the spherical map, one numerical coordinate
// each: r and dr/dX1 of X1; sin, cos of theta and dtheta/dX2 of X2; sin, cos
// of phi = X3.
template <class T>
struct PnCoords {
  T r, dr, st, ct, dth, sp, cp;
};
template <class T>
HD void pn_coord1(const PnBinary& B, const T& X1, T* r, T* dr) {
  const T er = exp(X1);
  *r = er + B.R0;
  *dr = er;
}
// x^p for an integer p >= 1 as pow() forms it, with the derivative
// p x^(p-1) (defined at x = 0, where the chain rule through pow's log is not)
HD double pow_int(double x, int p) { return pow(x, (double)p); }
HD Dual4 pow_int(const Dual4& x, int p) { return ad::chain(x, ::pow(x.v, (double)p), p * ::pow(x.v, (double)(p - 1))); }

template <class T>
HD void pn_coord2(const PnBinary& B, const T& X2, T* st, T* ct, T* dth) {
  // theta from the reflected X2, as x_of_xp() and dx_dxp_calc() form it
  // (src/coord.c:240-382): COORD_DIAGONAL's derivative from X2 itself,
  // COORD_DIAGONAL3's from f = 2 seesaw(X2) - 1 (src/coord.c:270-276, :468-472)
  const T X2s = X2 + (seesaw_unit(value(X2)) - value(X2));
  T th;
  if (B.coords == 4) {
    const T f = 2. * X2s - 1.;
    th = 0.5 * M_PI * (1. + B.h_slope * f + B.diag3_factor * pow_int(f, B.diag3_exponent));
    *dth = M_PI * (B.h_slope + (B.diag3_exponent * B.diag3_factor) * pow_int(f, B.diag3_exponent - 1));
  } else {
    th = B.th_length * X2s + B.th_beg + B.h_slope * sin(2. * M_PI * X2s);
    *dth = B.th_length + 2. * M_PI * B.h_slope * cos(2. * M_PI * X2);
  }
  *st = sin(th);
  *ct = cos(th);
  if (fabs(value(*st)) < 1.E-16) set_value(*st, value(*st) >= 0. ? 1.E-16 : -1.E-16);   // trigvalues(), src/metric.c:29-40
}
template <class T>
HD void pn_coord3(const T& X3, T* sp, T* cp) {
  *sp = sin(X3);
  *cp = cos(X3);
}

// The metric from the coordinate functions; zero is a zero of type T.
template <class T, class Pot>
HD void pn_gcov_from_coords(const PnCoords<T>& C, const T& zero, T g[4][4], const Pot& pot) {
  const T r = C.r, dr = C.dr, st = C.st, ct = C.ct, dth = C.dth, sp = C.sp, cp = C.cp;
  const T rs = r * st;
  T x[3] = {rs * cp, rs * sp, r * ct};
  T U, V, Vi[3];
  pot(x, U, V, Vi);
  // columns n = 1..3 of J = d(x, y, z)/d(X1, X2, X3)
  const T J1[3] = {st * cp * dr, st * sp * dr, ct * dr};
  const T rdth = r * dth;
  const T J2[3] = {ct * cp * rdth, ct * sp * rdth, -st * rdth};
  const T J3[3] = {-rs * sp, rs * cp, zero};
  const T A = 1. + 2. * U;
  g[0][0] = -1. + 2. * V - 2. * U * U;
  g[0][1] = g[1][0] = -4. * (Vi[0] * J1[0] + Vi[1] * J1[1] + Vi[2] * J1[2]);
  g[0][2] = g[2][0] = -4. * (Vi[0] * J2[0] + Vi[1] * J2[1] + Vi[2] * J2[2]);
  g[0][3] = g[3][0] = -4. * (Vi[0] * J3[0] + Vi[1] * J3[1]);
  g[1][1] = A * dr * dr;
  g[2][2] = A * rdth * rdth;
  g[3][3] = A * rs * rs;
  g[1][2] = g[2][1] = g[1][3] = g[3][1] = g[2][3] = g[3][2] = zero;
}

template <class T, class Pot>
HD void pn_gcov_num_pot(const PnBinary& B, const T& X1, const T& X2, const T& X3, T g[4][4], const Pot& pot) {
  PnCoords<T> C;
  pn_coord1(B, X1, &C.r, &C.dr);
  pn_coord2(B, X2, &C.st, &C.ct, &C.dth);
  pn_coord3(X3, &C.sp, &C.cp);
  pn_gcov_from_coords(C, X1 * 0., g, pot);
}

template <class T>
HD void pn_gcov_num(const PnBinary& B, const T& t, const T& X1, const T& X2, const T& X3, T g[4][4]) {
  pn_gcov_num_pot(B, X1, X2, X3, g, [&](const T x[3], T& U, T& V, T Vi[3]) { pn_potentials(B, t, x, U, V, Vi); });
}

// The same at a time level whose orbit is known.
HD void pn_gcov_num_orb(const PnBinary& B, const PnOrbit& o, double X1, double X2, double X3, double g[4][4]) {
  pn_gcov_num_pot(B, X1, X2, X3, g, [&](const double x[3], double& U, double& V, double Vi[3]) {
    pn_potentials_yv<double>(B, o.y, o.v, x, U, V, Vi);
  });
}

// Components of the numerical metric that are not identically zero (the
// spatial block is diagonal), in the order the derivative records use.
constexpr int PN_NG = 7;
HD void pn_gcomp(int m, int* a, int* b) {
  const int A[PN_NG] = {0, 0, 0, 0, 1, 2, 3}, Bv[PN_NG] = {0, 1, 2, 3, 1, 2, 3};
  *a = A[m]; *b = Bv[m];
}
constexpr int PN_NDG = 4 * PN_NG;   // dg[nu * PN_NG + m] = d g_{ab(m)} / dx^nu, x = (t, X1, X2, X3)

// First derivatives of the numerical metric at (X1, X2, X3) and the level of
// orbit o, by forward-mode differentiation with the orbit's time derivatives.
HD void pn_metric_dg(const PnBinary& B, const PnOrbit& o, double X1, double X2, double X3, double* dg) {
  Dual4 yd[2][3], vd[2][3];
  for (int a = 0; a < 2; a++)
    for (int i = 0; i < 3; i++) {
      yd[a][i] = dc(o.y[a][i]); yd[a][i].d[0] = o.ydot[a][i];
      vd[a][i] = dc(o.v[a][i]); vd[a][i].d[0] = o.vdot[a][i];
    }
  Dual4 g[4][4];
  pn_gcov_num_pot(B, dvar(X1, 1), dvar(X2, 2), dvar(X3, 3), g,
                  [&](const Dual4 x[3], Dual4& U, Dual4& V, Dual4 Vi[3]) {
                    pn_potentials_yv<Dual4>(B, yd, vd, x, U, V, Vi);
                  });
  for (int nu = 0; nu < 4; nu++)
    for (int m = 0; m < PN_NG; m++) {
      int a, b;
      pn_gcomp(m, &a, &b);
      dg[nu * PN_NG + m] = g[a][b].d[nu];
    }
}

// Inverse and determinant of a 4x4 matrix by the Laplace expansion in 2x2
// minors (closed form: no pivoting, no indexed scratch).
HD double invert4(const double a[4][4], double b[4][4]) {
  const double s0 = a[0][0]*a[1][1] - a[1][0]*a[0][1], s1 = a[0][0]*a[1][2] - a[1][0]*a[0][2];
  const double s2 = a[0][0]*a[1][3] - a[1][0]*a[0][3], s3 = a[0][1]*a[1][2] - a[1][1]*a[0][2];
  const double s4 = a[0][1]*a[1][3] - a[1][1]*a[0][3], s5 = a[0][2]*a[1][3] - a[1][2]*a[0][3];
  const double c5 = a[2][2]*a[3][3] - a[3][2]*a[2][3], c4 = a[2][1]*a[3][3] - a[3][1]*a[2][3];
  const double c3 = a[2][1]*a[3][2] - a[3][1]*a[2][2], c2 = a[2][0]*a[3][3] - a[3][0]*a[2][3];
  const double c1 = a[2][0]*a[3][2] - a[3][0]*a[2][2], c0 = a[2][0]*a[3][1] - a[3][0]*a[2][1];
  const double det = s0*c5 - s1*c4 + s2*c3 + s3*c2 - s4*c1 + s5*c0;
  const double id = 1. / det;
  b[0][0] = ( a[1][1]*c5 - a[1][2]*c4 + a[1][3]*c3) * id;
  b[0][1] = (-a[0][1]*c5 + a[0][2]*c4 - a[0][3]*c3) * id;
  b[0][2] = ( a[3][1]*s5 - a[3][2]*s4 + a[3][3]*s3) * id;
  b[0][3] = (-a[2][1]*s5 + a[2][2]*s4 - a[2][3]*s3) * id;
  b[1][0] = (-a[1][0]*c5 + a[1][2]*c2 - a[1][3]*c1) * id;
  b[1][1] = ( a[0][0]*c5 - a[0][2]*c2 + a[0][3]*c1) * id;
  b[1][2] = (-a[3][0]*s5 + a[3][2]*s2 - a[3][3]*s1) * id;
  b[1][3] = ( a[2][0]*s5 - a[2][2]*s2 + a[2][3]*s1) * id;
  b[2][0] = ( a[1][0]*c4 - a[1][1]*c2 + a[1][3]*c0) * id;
  b[2][1] = (-a[0][0]*c4 + a[0][1]*c2 - a[0][3]*c0) * id;
  b[2][2] = ( a[3][0]*s4 - a[3][1]*s2 + a[3][3]*s0) * id;
  b[2][3] = (-a[2][0]*s4 + a[2][1]*s2 - a[2][3]*s0) * id;
  b[3][0] = (-a[1][0]*c3 + a[1][1]*c1 - a[1][2]*c0) * id;
  b[3][1] = ( a[0][0]*c3 - a[0][1]*c1 + a[0][2]*c0) * id;
  b[3][2] = (-a[3][0]*s3 + a[3][1]*s1 - a[3][2]*s0) * id;
  b[3][3] = ( a[2][0]*s3 - a[2][1]*s1 + a[2][2]*s0) * id;
  return det;
}

// Geometry record (struct of_geom layout) at (t, X): metric, inverse, sqrt(-g),
// lapse, shift and normal observer, as calc_all_geom() forms them
// (src/metric.c:4221-4232).
// The numerical metric of this binary has a diagonal spatial block D and a
// shift row b, so its inverse is closed-form through the Schur complement
// S = g_tt - sum b_i^2 / D_i (= -alpha^2): g^tt = 1/S, g^ti = -q_i/S,
// g^ij = delta_ij / D_i + q_i q_j / S with q_i = b_i / D_i, det g = S D1 D2 D3.
HD void pn_geom_from_gcov(const double g[4][4], Geom* G) {
  for (int a = 0; a < 4; a++) for (int b = 0; b < 4; b++) G->gcov[a][b] = g[a][b];
  const double iD[3] = {1. / g[1][1], 1. / g[2][2], 1. / g[3][3]};
  const double q[3] = {g[0][1] * iD[0], g[0][2] * iD[1], g[0][3] * iD[2]};
  const double S = g[0][0] - (g[0][1] * q[0] + g[0][2] * q[1] + g[0][3] * q[2]);
  const double iS = 1. / S;
  G->gcon[0][0] = iS;
  for (int a = 0; a < 3; a++) G->gcon[0][a + 1] = G->gcon[a + 1][0] = -q[a] * iS;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) G->gcon[a + 1][b + 1] = ((a == b) ? iD[a] : 0.) + q[a] * q[b] * iS;
  const double det = S * g[1][1] * g[2][2] * g[3][3];
  G->g = ::sqrt(-det);
  G->g_inv = 1. / G->g;
  G->ncon[0] = ::sqrt(-G->gcon[0][0]);
  G->alpha = 1. / G->ncon[0];
  G->ncon[1] = -G->alpha * G->gcon[0][1];
  G->ncon[2] = -G->alpha * G->gcon[0][2];
  G->ncon[3] = -G->alpha * G->gcon[0][3];
  G->beta[0] = -G->gcon[0][1] / G->gcon[0][0];
  G->beta[1] = -G->gcon[0][2] / G->gcon[0][0];
  G->beta[2] = -G->gcon[0][3] / G->gcon[0][0];
}
HD void pn_geom(const PnBinary& B, double t, double X1, double X2, double X3, Geom* G) {
  double g[4][4];
  pn_gcov_num(B, t, X1, X2, X3, g);
  pn_geom_from_gcov(g, G);
}
// The record at a time level whose orbit is known: the same values as
// pn_geom() at that level's time.
HD void pn_geom_orb(const PnBinary& B, const PnOrbit& o, double X1, double X2, double X3, Geom* G) {
  double g[4][4];
  pn_gcov_num_orb(B, o, X1, X2, X3, g);
  pn_geom_from_gcov(g, G);
}

// Binary and grid of a context whose metric is the binary's: the numerical
// coordinates of zone (i,j,k) at position pos (coord(), src/coord.c:182-228).
HD long long pn_tab_size(const Params& P) { return 2LL * (2 * P.n1t + 3 * P.n2t + 2 * P.n3t); }
HD void zone_X(const Params& P, const PnGrid& G, int i, int j, int k, int pos, double* X) {
  const double h1 = (pos == FACE1 || pos == CORN) ? 0. : 0.5;
  const double h2 = (pos == FACE2 || pos == CORN) ? 0. : 0.5;
  const double h3 = (pos == FACE3 || pos == CORN) ? 0. : 0.5;
  X[1] = G.x0[1] + (i + G.gpos[0] + h1) * P.dx[1];
  X[2] = G.x0[2] + (j + G.gpos[1] + h2) * P.dx[2];
  X[3] = G.x0[3] + (k + G.gpos[2] + h3) * P.dx[3];
}
// Coordinate functions of zone (i,j,k) at position pos from the tables (the
// values pn_gcov_num() forms from the coordinates).
HD void pn_zone_coords(const Params& P, const PnGrid& G, int i, int j, int k, int pos, PnCoords<double>& C) {
  const int h1 = (pos == FACE1 || pos == CORN) ? 0 : 1;
  const int h2 = (pos == FACE2 || pos == CORN) ? 0 : 1;
  const int h3 = (pos == FACE3 || pos == CORN) ? 0 : 1;
  const double* t1 = G.tab + 2 * (h1 * P.n1t + i);
  const double* t2 = G.tab + 4 * P.n1t + 3 * (h2 * P.n2t + j);
  const double* t3 = G.tab + 4 * P.n1t + 6 * P.n2t + 2 * (h3 * P.n3t + k);
  C.r = t1[0]; C.dr = t1[1];
  C.st = t2[0]; C.ct = t2[1]; C.dth = t2[2];
  C.sp = t3[0]; C.cp = t3[1];
}
// Fill of the tables (host or device; one call per entry index e < pn_tab_size / 2)
HD void pn_tab_entry(const Params& P, const PnGrid& G, double* tab, long long e) {
  if (e < 2LL * P.n1t) {
    const int h = (int)(e / P.n1t), i = (int)(e % P.n1t);
    pn_coord1(G.B, G.x0[1] + (i + G.gpos[0] + 0.5 * h) * P.dx[1], &tab[2 * e], &tab[2 * e + 1]);
    return;
  }
  e -= 2LL * P.n1t;
  if (e < 2LL * P.n2t) {
    const int h = (int)(e / P.n2t), j = (int)(e % P.n2t);
    double* t = tab + 4 * P.n1t + 3 * e;
    pn_coord2(G.B, G.x0[2] + (j + G.gpos[1] + 0.5 * h) * P.dx[2], &t[0], &t[1], &t[2]);
    return;
  }
  e -= 2LL * P.n2t;
  if (e < 2LL * P.n3t) {
    const int h = (int)(e / P.n3t), k = (int)(e % P.n3t);
    double* t = tab + 4 * P.n1t + 6 * P.n2t + 2 * e;
    pn_coord3(G.x0[3] + (k + G.gpos[2] + 0.5 * h) * P.dx[3], &t[0], &t[1]);
  }
}
HD long long pn_tab_entries(const Params& P) { return 2LL * (P.n1t + P.n2t + P.n3t); }

// FP32 geometry record of zone (i,j,k) at position pos and the level of orbit
// o, evaluated in FP32 from the tables (for the FP32 face solve).  1 - alpha
// is formed without cancellation: 1 - alpha = (1 + S) / (1 + alpha) with
// S = g_tt - sum b_i q_i = -alpha^2 and 1 + S = 2V - 2U^2 - sum b_i q_i.
HD void pn_zone_geomf(const Params& P, const PnGrid& G, const PnOrbit& o, int i, int j, int k, int pos,
                      GeomF* out) {
  PnCoords<double> C;
  pn_zone_coords(P, G, i, j, k, pos, C);
  const float r = (float)C.r, dr = (float)C.dr, st = (float)C.st, ct = (float)C.ct, dth = (float)C.dth;
  const float sp = (float)C.sp, cp = (float)C.cp;
  const float rs = r * st;
  const float x[3] = {rs * cp, rs * sp, r * ct};
  const float m[2] = {(float)G.B.m1, (float)G.B.m2};
  float y[2][3], v[2][3], ra[2], nv[2], vsq[2];
  for (int a = 0; a < 2; a++)
    for (int q = 0; q < 3; q++) { y[a][q] = (float)o.y[a][q]; v[a][q] = (float)o.v[a][q]; }
  for (int a = 0; a < 2; a++) {
    const float dx = x[0] - y[a][0], dy = x[1] - y[a][1], dz = x[2] - y[a][2];
    ra[a] = sqrtf(dx * dx + dy * dy + dz * dz);
    nv[a] = (dx * v[a][0] + dy * v[a][1] + dz * v[a][2]) / ra[a];
    vsq[a] = v[a][0] * v[a][0] + v[a][1] * v[a][1] + v[a][2] * v[a][2];
  }
  const float ex = y[0][0] - y[1][0], ey = y[0][1] - y[1][1], ez = y[0][2] - y[1][2];
  const float r12 = sqrtf(ex * ex + ey * ey + ez * ez);
  const float r12_3 = r12 * r12 * r12;
  float U = 0.f, V = 0.f, Vi[3] = {0.f, 0.f, 0.f};
  for (int a = 0; a < 2; a++) {
    const int b = 1 - a;
    const float mr = m[a] / ra[a];
    U = U + mr;
    V = V + mr * (1.f + 2.f * vsq[a] - 0.5f * nv[a] * nv[a]) - m[a] * m[b] / (ra[a] * r12)
          + (m[a] * m[b] / 4.f) * (ra[b] * ra[b] - ra[a] * ra[a] - r12 * r12) / (ra[a] * r12_3);
  }
  for (int q = 0; q < 3; q++)
    for (int a = 0; a < 2; a++) Vi[q] = Vi[q] + m[a] * v[a][q] / ra[a];
  const float J1[3] = {st * cp * dr, st * sp * dr, ct * dr};
  const float rdth = r * dth;
  const float J2[3] = {ct * cp * rdth, ct * sp * rdth, -st * rdth};
  const float J3[2] = {-rs * sp, rs * cp};
  const float A = 1.f + 2.f * U;
  const float b[3] = {-4.f * (Vi[0] * J1[0] + Vi[1] * J1[1] + Vi[2] * J1[2]),
                      -4.f * (Vi[0] * J2[0] + Vi[1] * J2[1] + Vi[2] * J2[2]),
                      -4.f * (Vi[0] * J3[0] + Vi[1] * J3[1])};
  const float D[3] = {A * dr * dr, A * rdth * rdth, A * rs * rs};
  const float opg00 = 2.f * V - 2.f * U * U;          // 1 + g_tt
  for (int a = 0; a < 4; a++)
    for (int c = 0; c < 4; c++) out->gcov[a][c] = 0.f;
  out->gcov[0][0] = opg00 - 1.f;
  for (int a = 0; a < 3; a++) {
    out->gcov[0][a + 1] = out->gcov[a + 1][0] = b[a];
    out->gcov[a + 1][a + 1] = D[a];
  }
  const float iD[3] = {1.f / D[0], 1.f / D[1], 1.f / D[2]};
  const float q[3] = {b[0] * iD[0], b[1] * iD[1], b[2] * iD[2]};
  const float bq = b[0] * q[0] + b[1] * q[1] + b[2] * q[2];
  const float S = out->gcov[0][0] - bq;
  const float iS = 1.f / S;
  out->gcon[0][0] = iS;
  for (int a = 0; a < 3; a++) out->gcon[0][a + 1] = out->gcon[a + 1][0] = -q[a] * iS;
  for (int a = 0; a < 3; a++)
    for (int c = 0; c < 3; c++) out->gcon[a + 1][c + 1] = ((a == c) ? iD[a] : 0.f) + q[a] * q[c] * iS;
  out->g = sqrtf(-(S * D[0] * D[1] * D[2]));
  out->g_inv = 1.f / out->g;
  const float alpha = sqrtf(-S);
  out->alpha = alpha;
  out->ncon[0] = 1.f / alpha;
  for (int a = 0; a < 3; a++) {
    out->ncon[a + 1] = -alpha * out->gcon[0][a + 1];
    out->beta[a] = -out->gcon[0][a + 1] / out->gcon[0][0];
  }
  out->om_alpha = (opg00 - bq) / (1.f + alpha);
}

// Geometry record of zone (i,j,k) at position pos and the level of orbit o
HD void pn_zone_geom(const Params& P, const PnGrid& G, const PnOrbit& o, int i, int j, int k, int pos, Geom* out) {
  if (!G.tab) {
    double X[4];
    zone_X(P, G, i, j, k, pos, X);
    pn_geom_orb(G.B, o, X[1], X[2], X[3], out);
    return;
  }
  PnCoords<double> C;
  pn_zone_coords(P, G, i, j, k, pos, C);
  double g[4][4];
  pn_gcov_from_coords(C, 0., g, [&](const double x[3], double& U, double& V, double Vi[3]) {
    pn_potentials_yv<double>(G.B, o.y, o.v, x, U, V, Vi);
  });
  pn_geom_from_gcov(g, out);
}

// Connection Gamma^d_{ab} = g^{dc} (d_a g_cb + d_b g_ca - d_c g_ab) / 2 at
// (t, X), time derivatives included; conn[(d*4 + a)*4 + b] (conn[l][d][a][b]).
HD void pn_conn(const PnBinary& B, double t, double X1, double X2, double X3, double* conn) {
  // the connection from the same metric derivatives the step's source term
  // uses (pn_metric_dg()), at the orbit of time t
  PnOrbit o;
  pn_orbit_at(B, t, &o);
  double gv[4][4], gi[4][4], dg[PN_NDG];
  pn_gcov_num_orb(B, o, X1, X2, X3, gv);
  pn_metric_dg(B, o, X1, X2, X3, dg);
  invert4(gv, gi);
  double d[4][4][4];   // d[c][a][b] = d g_ab / dx^c
  for (int c = 0; c < 4; c++)
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) d[c][a][b] = 0.;
  for (int c = 0; c < 4; c++)
    for (int m = 0; m < PN_NG; m++) {
      int a, b;
      pn_gcomp(m, &a, &b);
      d[c][a][b] = d[c][b][a] = dg[c * PN_NG + m];
    }
  double low[4][4][4];   // Gamma_{c a b}
  for (int c = 0; c < 4; c++)
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) low[c][a][b] = 0.5 * (d[a][c][b] + d[b][c][a] - d[c][a][b]);
  for (int e = 0; e < 4; e++)
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) {
        double s = 0.;
        for (int c = 0; c < 4; c++) s += gi[e][c] * low[c][a][b];
        conn[(e * 4 + a) * 4 + b] = s;
      }
}

}  // namespace hgpu
