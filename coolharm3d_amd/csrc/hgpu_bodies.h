// hgpu_bodies.h -- per-thread bodies of the substep kernels.
//
// Each body handles one face, edge or zone and is called by a thin __global__
// wrapper (hgpu_kernels.hip).  The same bodies are compiled for the host by the
// check build (hgpu_hostcheck.cpp), which runs them in loops so the kernels'
// arithmetic can be compared with the reference on a machine without a GPU.
#pragma once
#include "hgpu_phys.h"
#include "hgpu_mixed.h"
#include "hgpu_pn.h"

namespace hgpu {

// Binary metric evaluated in a kernel: the grid and the orbit of the level.
struct FaceOtf {
  PnGrid G;
  const PnOrbit* o;
};

// Zone-centre record at the level X.geom describes: read, or (binary metric
// evaluated in the kernels, P.otf) evaluated into buf from the orbit of level
// X.cent_lvl.
HD const Geom* cent_rec(const Params& P, const Ptrs& X, int i, int j, int k, Geom& buf) {
  if (P.otf) {
    pn_zone_geom(P, X.pn, X.orbit[X.cent_lvl], i, j, k, CENT, &buf);
    return &buf;
  }
  return geom_at(P, X, i, j, k, CENT);
}

// --------------------------------------------------------------------------
// numerical_flux() for one face, src/step_ch.c:636-790 (with the face states of
// reconstruct_fast(), src/recon.c:894-1103).  src is the state being
// reconstructed (p in the predictor, ph in the corrector).  The direction is a
// template parameter so every 4-vector component the face selects (u^dir,
// b^dir, g^{t dir}) is a compile-time index and stays in registers.
template <int d>
HD void flux_recon(const Params& P, const Ptrs& X, const double* __restrict__ src,
                   int i, int j, int k, double* pl, double* pr) {
  const long long NT = P.nt;
  constexpr int di = (d == 0), dj = (d == 1), dk = (d == 2);

  if (P.degen[d]) {
    // degenerate direction: p_L = p_R = p_h of the zone at the direction's
    // first index; degenerate directions after d are clamped by the later
    // copy passes, earlier ones are not (src/recon.c:1050-1075)
    int ii = i, jj = j, kk = k;
    if (d == 0) { ii = NG; if (P.degen[1]) jj = NG; if (P.degen[2]) kk = NG; }
    if (d == 1) { jj = NG; if (P.degen[2]) kk = NG; }
    if (d == 2) { kk = NG; }
    long long c = cidx(P, ii, jj, kk);
    for (int v = 0; v < NP; v++) pl[v] = pr[v] = rt_val(P, X, v, ii, src[v*NT + c]);
  } else {
    int ii = i, jj = j, kk = k;
    if (P.degen[0]) ii = NG;
    if (P.degen[1]) jj = NG;
    if (P.degen[2]) kk = NG;
    if (P.recon == RECON_PPM_MC) {
      long long c0 = cidx(P, ii - 3*di, jj - 3*dj, kk - 3*dk);
      long long st = cidx(P, di, dj, dk) - cidx(P, 0, 0, 0);
      for (int v = 0; v < NP; v++) {
        double y[6];
        for (int m = 0; m < 6; m++) {
          double x = src[v*NT + c0 + m*st];
          if (d == 0) {
            x *= X.rscale[rs_slot(v)*P.n1t + (ii - 3 + m)];
          } else {
            x = rt_val(P, X, v, ii, x);
          }
          y[m] = x;
        }
        ppm_interface(y, &pl[v], &pr[v]);
        if (d == 0) {
          double f = X.irscale[(3 + rs_slot(v))*P.n1t + i];
          pl[v] *= f;
          pr[v] *= f;
        }
      }
    } else {
      // MC slopes, src/recon.c:960-986
      long long c0 = cidx(P, ii - 2*di, jj - 2*dj, kk - 2*dk);
      long long st = cidx(P, di, dj, dk) - cidx(P, 0, 0, 0);
      for (int v = 0; v < NP; v++) {
        double y[4];
        for (int m = 0; m < 4; m++) {
          double x = src[v*NT + c0 + m*st];
          if (d == 0) {
            x *= X.rscale[rs_slot(v)*P.n1t + (ii - 2 + m)];
          } else {
            x = rt_val(P, X, v, ii, x);
          }
          y[m] = x;
        }
        double dql = 0.5*mc_slope(y[0], y[1], y[2]);
        double dqr = 0.5*mc_slope(y[1], y[2], y[3]);
        pl[v] = y[1] + dql;
        pr[v] = y[2] - dqr;
        if (d == 0) {
          double f = X.irscale[(3 + rs_slot(v))*P.n1t + i];
          pl[v] *= f;
          pr[v] *= f;
        }
      }
    }
  }
}

// PPM+MC face states of F consecutive faces along d starting at face (i,j,k),
// each zone's limited parabola computed once (para_vect(), src/recon.c:994-1043,
// evaluated per zone as the reference does; flux_recon() evaluates both
// zones of one face).  Direction d must not be degenerate.  Zones past the
// array end, which only faces beyond a partial group read, are clamped.
template <int d, int F>
HD void ppm_faces(const Params& P, const Ptrs& X, const double* __restrict__ src,
                  int i, int j, int k, double (*pl)[NP], double (*pr)[NP]) {
  const long long NT = P.nt;
  constexpr int di = (d == 0), dj = (d == 1), dk = (d == 2);
  constexpr int M = F + 5;
  int ii = i, jj = j, kk = k;
  if (d != 0 && P.degen[0]) ii = NG;
  if (d != 1 && P.degen[1]) jj = NG;
  if (d != 2 && P.degen[2]) kk = NG;
  const int pos = (d == 0) ? ii : ((d == 1) ? jj : kk);
  const int nd = (d == 0) ? P.n1t : ((d == 1) ? P.n2t : P.n3t);
  const long long st = cidx(P, di, dj, dk) - cidx(P, 0, 0, 0);
  const long long c0 = cidx(P, ii - 3*di, jj - 3*dj, kk - 3*dk);
  int zm[M];
#pragma unroll
  for (int m = 0; m < M; m++) zm[m] = imin(m, nd - 1 - (pos - 3));
  // The next primitive's stencil is loaded before this one's limiters run
  // (their branches keep the compiler from moving loads across them); this
  // primitive's rescale factors are loaded before that, so waiting for them
  // leaves the stencil loads in flight.  rt_val()'s round trip is spelled out
  // with the factors loaded ahead.
  double raw[2][M];
#pragma unroll
  for (int m = 0; m < M; m++) raw[0][m] = src[c0 + zm[m]*st];
#pragma unroll
  for (int v = 0; v < NP; v++) {
    const int s = rs_slot(v);
    double fz[M], fo[F], ra = 1., ri = 1.;
    if (d == 0) {
#pragma unroll
      for (int m = 0; m < M; m++) fz[m] = X.rscale[s*P.n1t + (ii - 3 + zm[m])];
#pragma unroll
      for (int f = 0; f < F; f++) fo[f] = X.irscale[(3 + s)*P.n1t + imin(i + f, P.n1t - 1)];
    } else {
      ra = X.rscale[s*P.n1t + ii];
      ri = X.irscale[s*P.n1t + ii];
    }
    if (v + 1 < NP) {
#pragma unroll
      for (int m = 0; m < M; m++) raw[(v + 1) & 1][m] = src[(v + 1)*NT + c0 + zm[m]*st];
    }
    double y[M];
#pragma unroll
    for (int m = 0; m < M; m++) y[m] = (d == 0) ? raw[v & 1][m]*fz[m] : (raw[v & 1][m]*ra)*ri;
    double dq[M - 1], fi[M - 2], le[M - 2], re[M - 2];
#pragma unroll
    for (int m = 1; m < M - 1; m++) dq[m] = mc_slope_vect(y[m - 1], y[m], y[m + 1]);
#pragma unroll
    for (int m = 1; m < M - 2; m++) fi[m] = 0.5*(y[m + 1] + y[m]) - PPM_FACTOR*(dq[m + 1] - dq[m]);
#pragma unroll
    for (int m = 2; m < M - 2; m++) ppm_limit(fi[m - 1], fi[m], y[m], &le[m], &re[m]);
#pragma unroll
    for (int f = 0; f < F; f++) {
      double a = re[2 + f], b = le[3 + f];
      if (d == 0) {
        a *= fo[f];
        b *= fo[f];
      }
      pl[f][v] = a;
      pr[f][v] = b;
    }
  }
}

// Signal speeds of a face from its two states (src/step_ch.c:706-714)
HD void face_speeds(const State& ql, const State& qr, const Geom* g, int dim, double gam,
                    double* cmax, double* cmin, double* ctop) {
  double cmaxl, cminl, cmaxr, cminr;
  vchar_fast(&ql, g, dim, gam, &cmaxl, &cminl);
  vchar_fast(&qr, g, dim, gam, &cmaxr, &cminr);
  *cmax = fabs(dmax(dmax(0., cmaxl), cmaxr));
  *cmin = fabs(dmax(dmax(0., -cminl), -cminr));
  *ctop = dmax(*cmin, *cmax);
}

// Riemann solve of face (i,j,k) from its two states: fluxes and the fastest
// interior speed (the rest of numerical_flux(), src/step_ch.c:681-764).  The
// fluxes go to X.F, or to out[8] when the caller stores them itself.
// g: the face's geometry record (flux_riemann() reads the stored one).
template <int d>
HD void flux_riemann_g(const Params& P, const Ptrs& X, const double* pl, const double* pr,
                       int i, int j, int k, FailRec& fr, double* out, const Geom* g) {
  const long long NT = P.nt;
  constexpr int dim = d + 1;
  State ql, qr;
  if (get_state(pl, g, P.gam, &ql)) fr.nf[NFAIL_GAMMA_CALC]++;
  if (get_state(pr, g, P.gam, &qr)) fr.nf[NFAIL_GAMMA_CALC]++;
  double Ul[NP], Ur[NP], Fl[NP], Fr[NP];
  primtoflux(pl, &ql, 0, g, Ul);
  primtoflux(pr, &qr, 0, g, Ur);
  primtoflux(pl, &ql, dim, g, Fl);
  primtoflux(pr, &qr, dim, g, Fr);
  double cmax, cmin, ctop;
  face_speeds(ql, qr, g, dim, P.gam, &cmax, &cmin, &ctop);

  // fastest speed over interior faces (timestep(), src/step_ch.c:716-724),
  // reduced over the launch by the caller
  if (i >= NG && i < NG + P.n1 && j >= NG && j < NG + P.n2 && k >= NG && k < NG + P.n3 && ctop > 0.)
    fr.vmax = dmax(fr.vmax, ctop);

  double Fv[NP];
  if (P.flux == FLUX_HLLE) {
    for (int l = 0; l < NP; l++)
      Fv[l] = (cmax*Fl[l] + cmin*Fr[l] - cmax*cmin*(Ur[l] - Ul[l])) / (cmax + cmin + SMALL);
  } else {
    for (int l = 0; l < NP; l++)
      Fv[l] = 0.5*(Fl[l] + Fr[l] - ctop*(Ur[l] - Ul[l]));
  }
  if (out) {
    for (int l = 0; l < NP; l++) out[l] = Fv[l];
  } else {
    const long long c = cidx(P, i, j, k);
    // the normal field's flux is not stored: flux_ct() zeroes it and nothing
    // reads it (src/flux_ct.c:853-863)
    for (int l = 0; l < NP; l++)
      if (l != B1 + d) X.F[(d*NFLUX + l)*NT + c] = Fv[l];
  }
}

template <int d>
HD void flux_riemann(const Params& P, const Ptrs& X, const double* pl, const double* pr,
                     int i, int j, int k, FailRec& fr, double* out = nullptr) {
  flux_riemann_g<d>(P, X, pl, pr, i, j, k, fr, out, geom_at(P, X, i, j, k, d));
}

// Entropy flux of face (i,j,k) in direction d for fixup_entropy_eq()
// (src/fixup.c:363-398), from the face data the reference keeps in
// c_gf/ucon_L/ucon_R (src/step_ch.c:743-750).  Only the rare zones that take
// the entropy fix read it, so it is computed here, from the stage's
// reconstruction source and face geometry, with the arithmetic of the face
// solve, instead of being stored by every face.
template <int d>
HD double entropy_face_d(const Params& P, const Ptrs& X, const double* src, int i, int j, int k,
                         const FaceOtf* fo) {
  Ptrs Xs = X;
  if (X.geom_s) Xs.geom = X.geom_s;
  double pl[NP], pr[NP];
  flux_recon<d>(P, Xs, src, i, j, k, pl, pr);
  Geom gb;
  if (fo) pn_zone_geom(P, fo->G, *fo->o, i, j, k, d, &gb);
  const Geom* g = fo ? &gb : geom_at(P, Xs, i, j, k, d);
  constexpr int dim = d + 1;
  State ql, qr;
  get_state(pl, g, P.gam, &ql);   // failures were counted by the face solve
  get_state(pr, g, P.gam, &qr);
  double cmax, cmin, ctop;
  face_speeds(ql, qr, g, dim, P.gam, &cmax, &cmin, &ctop);
  double uL0 = g->g * ql.ucon[0], uL1 = g->g * ql.ucon[dim];
  double uR0 = g->g * qr.ucon[0], uR1 = g->g * qr.ucon[dim];
  double s = (P.gam - 1.) * pl[UU] * pow(pl[RHO], 1. - P.gam);
  double Sl = s * uL0, Gl = s * uL1;
  s = (P.gam - 1.) * pr[UU] * pow(pr[RHO], 1. - P.gam);
  double Sr = s * uR0, Gr = s * uR1;
  if (P.flux == FLUX_HLLE) return (cmax*Gl + cmin*Gr - cmax*cmin*(Sr - Sl)) / (cmax + cmin + SMALL);
  return 0.5*(Gl + Gr - ctop*(Sr - Sl));
}

template <int d>
HD void flux_face(const Params& P, const Ptrs& X, const double* __restrict__ src,
                  int i, int j, int k, FailRec& fr) {
  double pl[NP], pr[NP];
  flux_recon<d>(P, X, src, i, j, k, pl, pr);
  flux_riemann<d>(P, X, pl, pr, i, j, k, fr);
}

// --------------------------------------------------------------------------
// Corner EMFs, flux_ct_para_fast2() (src/flux_ct.c:759-846) or flux_ct()
// (src/flux_ct.c:26-41) for the edge (i,j,k).
HD void emf_edge(const Params& P, const Ptrs& X, int i, int j, int k) {
  const long long NT = P.nt;
  const long long c = cidx(P, i, j, k);
  const long long si = (long long)P.n2t * P.n3t, sj = P.n3t, sk = 1;
  const double* F = X.F;
  const double* F0B2 = F + (0*NFLUX + B2)*NT;
  const double* F0B3 = F + (0*NFLUX + B3)*NT;
  const double* F1B1 = F + (1*NFLUX + B1)*NT;
  const double* F1B3 = F + (1*NFLUX + B3)*NT;
  const double* F2B1 = F + (2*NFLUX + B1)*NT;
  const double* F2B2 = F + (2*NFLUX + B2)*NT;
  double e0, e1, e2;
  if (P.ct_para) {
    // emf0: + PPM of F1[B3] along x3, - PPM of F2[B2] along x2;
    // emf1: + PPM of F2[B1] along x1, - PPM of F0[B3] along x3;
    // emf2: + PPM of F0[B2] along x2, - PPM of F1[B1] along x1.
    // All six stencils are loaded before any limiter runs, so their loads
    // are in flight together.
    const double* line[6] = {F1B3, F2B2, F2B1, F0B3, F0B2, F1B1};
    const long long str[6] = {sk, sj, si, sk, sj, si};
    double y[6][6], yl[6], yr[6];
#pragma unroll
    for (int l = 0; l < 6; l++)
#pragma unroll
      for (int m = 0; m < 6; m++) y[l][m] = line[l][c + (m - 3)*str[l]];
#pragma unroll
    for (int l = 0; l < 6; l++) ppm_interface(y[l], &yl[l], &yr[l]);
    e0 = (((0.0 + yl[0]) + yr[0]) - yl[1]) - yr[1];
    e1 = (((0.0 + yl[2]) + yr[2]) - yl[3]) - yr[3];
    e2 = (((0.0 + yl[4]) + yr[4]) - yl[5]) - yr[5];
  } else {
    e0 = 0.25*(F1B3[c] + F1B3[c - sk] - F2B2[c] - F2B2[c - sj]);
    e1 = 0.25*(F2B1[c] + F2B1[c - si] - F0B3[c] - F0B3[c - sk]);
    e2 = 0.25*(F0B2[c] + F0B2[c - sj] - F1B1[c] - F1B1[c - si]);
  }
  X.emf[0*NT + c] = e0;
  X.emf[1*NT + c] = e1;
  X.emf[2*NT + c] = e2;
}

// Face flux of variable l through face (i,j,k) in direction d as the update
// reads it: B components rebuilt from the corner EMFs (src/flux_ct.c:853-933),
// the x1 mass flux limited by fix_flux() (src/bounds.c:2255-2270).
HD double face_flux(const Params& P, const Ptrs& X, int d, int l, int i, int j, int k) {
  const long long NT = P.nt;
  const long long c = cidx(P, i, j, k);
  if (l >= B1) {
    const double cf = P.ct_para ? 0.125 : 0.5;
    const long long si = (long long)P.n2t * P.n3t, sj = P.n3t, sk = 1;
    const double* E0 = X.emf;
    const double* E1 = X.emf + NT;
    const double* E2 = X.emf + 2*NT;
    if (d == 0) {
      if (l == B1) return 0.;
      if (l == B2) return 0. + cf*(E2[c] + E2[c + sj]);
      return 0. + (-cf)*(E1[c] + E1[c + sk]);
    } else if (d == 1) {
      if (l == B1) return 0. + (-cf)*(E2[c] + E2[c + si]);
      if (l == B2) return 0.;
      return 0. + cf*(E0[c] + E0[c + sk]);
    } else {
      if (l == B1) return 0. + cf*(E1[c] + E1[c + si]);
      if (l == B2) return 0. + (-cf)*(E0[c] + E0[c + sj]);
      return 0.;
    }
  }
  double f = X.F[(d*NFLUX + l)*NT + c];
  if (l == RHO && d == 0 && P.bc == BC_SPHERICAL_OUTFLOW3) {
    if (i == NG && f > 0 && !(P.bc_nbr & 1)) f = 0.;
    if (i == NG + P.n1 && f < 0 && !(P.bc_nbr & 2)) f = 0.;
  }
  return f;
}

// --------------------------------------------------------------------------
// source(), src/phys.c:121-187, with cooling_func_hr_disk(), src/phys.c:649-784.
// The binary metric (P.dyn) contracts the stress tensor with the metric's
// first derivatives, dU_nu = T^{mu lambda} d_nu g_{mu lambda} / 2, which equals
// T^mu_kappa Gamma^kappa_{nu mu} for the metric connection; its derivative
// records are the stage's stored level X.conn ([PN_NDG][zones]).
// MET: -1 decides between the static and the binary metric at run time;
// 0 / 1 compile only the static / binary branch (kernels built per metric, so
// the branch not taken costs no registers).
// Binary-metric source terms dU_nu = 1/2 T^{mu lambda} d_nu g_{mu lambda}
// (equal to T^mu_kappa Gamma^kappa_{nu mu} of source(), src/phys.c:143-153,
// for the metric connection) plus the cooling term, in two parts so a caller
// can form the state's part before evaluating the metric derivatives: the
// weights of the 7 metric components' derivatives (T^tt, T^t1 + T^1t,
// T^t2 + T^2t, T^t3 + T^3t, T^11, T^22, T^33 with T^{mu lambda} =
// T^mu_kappa g^{kappa lambda}) and the cooling rate (src/phys.c:541-582),
// then the contraction, the cooling term and sqrt(-g).
struct SrcDyn {
  double w[PN_NG];
  double f_cool, ucov[4];
  int gfail;
};
HD void source_dyn_state(const Params& P, const Ptrs& X, const double* ph, const Geom* g, int i, SrcDyn& s) {
  State q;
  s.gfail = get_state(ph, g, P.gam, &q);
  double mhd[4][4];
  mhd_calc(&q, 0, mhd[0]); mhd_calc(&q, 1, mhd[1]);
  mhd_calc(&q, 2, mhd[2]); mhd_calc(&q, 3, mhd[3]);
  auto Tu = [&](int a, int b) {
    return mhd[a][0]*g->gcon[0][b] + mhd[a][1]*g->gcon[1][b] + mhd[a][2]*g->gcon[2][b]
         + mhd[a][3]*g->gcon[3][b];
  };
  s.w[0] = Tu(0, 0); s.w[1] = Tu(0, 1) + Tu(1, 0); s.w[2] = Tu(0, 2) + Tu(2, 0); s.w[3] = Tu(0, 3) + Tu(3, 0);
  s.w[4] = Tu(1, 1); s.w[5] = Tu(2, 2); s.w[6] = Tu(3, 3);
  for (int a = 0; a < 4; a++) s.ucov[a] = q.ucov[a];
  s.f_cool = 0.;
  if (P.cooling) {
    int bound = ((q.ucov[0]*(q.p + ph[UU] + ph[RHO] + q.bsq)) > (-ph[RHO])) ? 1 : 0;
    if (bound) {
      double T_r = X.tr[i];
      double Omega_r = X.omega[i];
      double E_ratio = (P.gam - 1.)*ph[UU] / (ph[RHO] * T_r);
      double dE = E_ratio - 1.;
      double base = dE + fabs(dE);
      double pw = (P.q_exp == 0.5) ? sqrt(base) : pow(base, P.q_exp);
      s.f_cool = P.s_norm * ph[UU] * Omega_r * pw;
    }
  }
}
// dg: the derivatives d_nu g_m at dg[(nu*7 + m)*cs]; gdet: sqrt(-g).
HD void source_dyn_finish(const Params& P, const SrcDyn& s, const double* dg, long long cs, double gdet,
                          double* dU) {
  for (int l = 0; l < NP; l++) dU[l] = 0.;
  for (int nu = 0; nu < 4; nu++) {
    double sum = 0.;
    for (int m = 0; m < PN_NG; m++) sum += s.w[m]*dg[(nu*PN_NG + m)*cs];
    dU[UU + nu] += 0.5*sum;
  }
  if (P.cooling) {
    dU[UU] -= s.f_cool * s.ucov[0];
    dU[U1] -= s.f_cool * s.ucov[1];
    dU[U2] -= s.f_cool * s.ucov[2];
    dU[U3] -= s.f_cool * s.ucov[3];
  }
  dU[UU] *= gdet;
  dU[U1] *= gdet;
  dU[U2] *= gdet;
  dU[U3] *= gdet;
}

// MET: -1 from P; 0 static metric; 1 binary metric, derivatives from the
// stored level X.conn; 2 binary metric, the terms themselves precomputed in
// X.conn ([4][nconn], UU..U3, by k_source_otf, which counts their failures);
// 3 binary metric, derivatives passed in dgl[28].
// UNI: every lane of the calling wavefront is in the zone's (i, j) column
// (row-aligned tiles of a 2D static metric), so the connection record is read
// through the scalar cache.
template <int MET = -1, bool UNI = false, class GP = const Geom*>
HD int source(const Params& P, const Ptrs& X, const double* ph, GP g, int i, int j, int k,
              double* dU, const double* dgl = nullptr) {
  for (int l = 0; l < NP; l++) dU[l] = 0.;
  if constexpr (MET == 2) {
    const long long cid = conn_id(P, i, j, k);
    for (int m = 0; m < 4; m++) dU[UU + m] = X.conn[m*P.nconn + cid];
    return 0;
  }
  if constexpr (MET != 0) {
    const bool dyn = (MET < 0) ? (P.dyn != 0) : (MET >= 1);
    if (dyn) {   // the binary metric (3D, not flat)
      SrcDyn s;
      source_dyn_state(P, X, ph, g, i, s);
      if constexpr (MET == 3) {
        source_dyn_finish(P, s, dgl, 1, g->g, dU);
      } else {
        source_dyn_finish(P, s, X.conn + conn_id(P, i, j, k), P.nconn, g->g, dU);
      }
      return s.gfail;
    }
  }
  int gfail = 0;
  State q;
  bool have_q = false;
  if (P.metric_dim > 0) {
    gfail += get_state(ph, g, P.gam, &q);
    have_q = true;
    if (!P.flat) {
      double mhd[4][4];
      mhd_calc(&q, 0, mhd[0]); mhd_calc(&q, 1, mhd[1]);
      mhd_calc(&q, 2, mhd[2]); mhd_calc(&q, 3, mhd[3]);
      const long long cid = conn_id(P, i, j, k);
      {
        // static metric: records [zone][64] of Gamma^b_{nu a}; each sum runs
        // over (a, b) in the same order, the 3D test outside the loop so the
        // record's loads issue together
        const double* cn = X.conn + cid * 64;
#ifdef __HIP_DEVICE_COMPILE__
        if constexpr (UNI) {
          typedef const __attribute__((address_space(4))) double* CDp;
          const CDp cs = (CDp)(X.conn + (long long)__builtin_amdgcn_readfirstlane((int)cid) * 64);
          for (int a = 0; a < 4; a++) for (int b = 0; b < 4; b++) {
            dU[U1] += mhd[a][b]*cs[((b*4 + 1)*4 + a)];
            dU[U2] += mhd[a][b]*cs[((b*4 + 2)*4 + a)];
          }
        } else
#endif
        for (int a = 0; a < 4; a++) for (int b = 0; b < 4; b++) {
          dU[U1] += mhd[a][b]*cn[((b*4 + 1)*4 + a)];
          dU[U2] += mhd[a][b]*cn[((b*4 + 2)*4 + a)];
        }
        if (P.metric_dim > 2) {
          for (int a = 0; a < 4; a++) for (int b = 0; b < 4; b++) {
            dU[UU] += mhd[a][b]*cn[((b*4 + 0)*4 + a)];
            dU[U3] += mhd[a][b]*cn[((b*4 + 3)*4 + a)];
          }
        }
      }
    }
  }
  if (P.cooling) {
    if (!have_q) gfail += get_state(ph, g, P.gam, &q);
    int bound = ((q.ucov[0]*(q.p + ph[UU] + ph[RHO] + q.bsq)) > (-ph[RHO])) ? 1 : 0;
    double f_cool = 0.;
    if (bound) {
      double T_r = X.tr[i];
      double Omega_r = X.omega[i];
      double E_ratio = (P.gam - 1.)*ph[UU] / (ph[RHO] * T_r);
      double dE = E_ratio - 1.;
      double base = dE + fabs(dE);
      double pw = (P.q_exp == 0.5) ? sqrt(base) : pow(base, P.q_exp);
      f_cool = P.s_norm * ph[UU] * Omega_r * pw;
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
  return gfail;
}

// --------------------------------------------------------------------------
// Entropy flux of face (i,j,k) in direction d, FP64 or FP32 face solve
// fo: the face records are evaluated (binary metric) rather than stored.
template <bool MIXED>
HD double entropy_flux(const Params& P, const Ptrs& X, const double* src, int d, int i, int j, int k,
                       const FaceOtf* fo) {
  if constexpr (MIXED) {
    GeomF gf;
    const GeomF* gp = nullptr;
    if (fo) {
      pn_zone_geomf(P, fo->G, *fo->o, i, j, k, d, &gf);
      gp = &gf;
    }
    if (d == 0) return entropy_face_mixed<0>(P, X, src, i, j, k, gp);
    if (d == 1) return entropy_face_mixed<1>(P, X, src, i, j, k, gp);
    return entropy_face_mixed<2>(P, X, src, i, j, k, gp);
  } else {
    if (d == 0) return entropy_face_d<0>(P, X, src, i, j, k, fo);
    if (d == 1) return entropy_face_d<1>(P, X, src, i, j, k, fo);
    return entropy_face_d<2>(P, X, src, i, j, k, fo);
  }
}

// fixup_entropy_eq(), src/fixup.c:335-454
template <bool MIXED>
HD int fixup_entropy_eq(const Params& P, const Ptrs& X, int i, int j, int k, double Dt,
                        const double* prim_old, double* prim, const Geom* g, const double* U,
                        double* gamma_out, const double* src, const FaceOtf* fo, FailRec& fr) {
  double prim_backup[NP];
  for (int l = 0; l < NP; l++) prim_backup[l] = prim[l];
  double gamma;
  gamma_calc(prim_old, g, &gamma);
  double S_old = g->g * (P.gam - 1.)*prim_old[UU]*gamma / (g->alpha * pow(prim_old[RHO], (P.gam - 1.)));
  const double Fm0 = entropy_flux<MIXED>(P, X, src, 0, i, j, k, fo), Fp0 = entropy_flux<MIXED>(P, X, src, 0, i + 1, j, k, fo);
  const double Fm1 = entropy_flux<MIXED>(P, X, src, 1, i, j, k, fo), Fp1 = entropy_flux<MIXED>(P, X, src, 1, i, j + 1, k, fo);
  const double Fm2 = entropy_flux<MIXED>(P, X, src, 2, i, j, k, fo), Fp2 = entropy_flux<MIXED>(P, X, src, 2, i, j, k + 1, fo);
  double S = S_old + Dt*(
      - (Fp0 - Fm0) * P.invdx[1]
      - (Fp1 - Fm1) * P.invdx[2]
      - (Fp2 - Fm2) * P.invdx[3]);
  int ret = utoprim_1d_ee(U, S, g, P.gam, prim, &gamma);
  if ((!ret) && (check_gamma(gamma) == 2)) {
    for (int l = 0; l < NP; l++) prim[l] = prim_backup[l];
    ret = 1;
    gamma = 1.;
  }
  if (ret) {
    ret = utoprim_1d_ee2(U, S, g, P.gam, prim, &gamma);
    if ((!ret) && (check_gamma(gamma) == 2)) {
      for (int l = 0; l < NP; l++) prim[l] = prim_backup[l];
      ret = 1;
    }
    if (ret) gamma = 1.;
  }
  *gamma_out = gamma;
  if (!ret) fail_set(P, fr, NFAIL_USE_EE, 0, i, j, k);
  return ret;
}

// --------------------------------------------------------------------------
// recover_primitives(), src/step_ch.c:801-1024 (FIXUP_TREE 3, entropy fix on)
// src: the stage's reconstruction source, for the entropy face fluxes.
template <bool MIXED>
HD void recover_primitives(const Params& P, const Ptrs& X, int i, int j, int k, double Dt,
                           double* pf, const double* U, const double* p_old, const Geom* g,
                           double rhoflr, double uuflr, double* p_gamma, const double* src, const FaceOtf* fo,
                           FailRec& fr) {
  double pg[NP], pf2[NP];
  for (int l = 0; l < NP; l++) pg[l] = pf[l];
  double gamma = 1., bsq = 0.;
  int ret = utoprim_2d_fast(U, g, P.gam, P.gam_m1_o_gam, pf, &gamma, &bsq, nullptr);
  if ((!ret) && (check_gamma(gamma) == 2)) {
    for (int l = 0; l < B1; l++) pf[l] = pg[l];
    ret = 1;
    gamma = 1.;
  }
  if (ret) {
    ret = utoprim_1d(U, g, P.gam, pf, &gamma, &bsq);
    if ((!ret) && (check_gamma(gamma) == 2)) {
      for (int l = 0; l < B1; l++) pf[l] = pg[l];
      ret = 1;
    }
    if (ret) gamma = 1.;
  }
  *p_gamma = gamma;

  int use_ee = 0;
  if (ret || (pf[UU] < BETA_MIN * bsq) || (pf[RHO] < BETA_MIN * bsq)) use_ee = 1;

  if (!use_ee) {
    if (!ret) fail_set(P, fr, NFAIL_USE_FULL_INV, 0, i, j, k);
    else fail_set(P, fr, NFAIL_UTOPRIM, 1, i, j, k);
  } else {
    if (!ret) { for (int l = 0; l < NP; l++) pf2[l] = pf[l]; }
    else fail_set(P, fr, NFAIL_UTOPRIM, 0, i, j, k);
    for (int l = 0; l < B1; l++) pf[l] = pg[l];
    int ret_ee = fixup_entropy_eq<MIXED>(P, X, i, j, k, Dt, p_old, pf, g, U, &gamma, src, fo, fr);
    if (ret_ee) {
      if (ret) {
        fail_set(P, fr, NFAIL_UTOPRIM_EE, 1, i, j, k);
        return;
      } else {
        fail_set(P, fr, NFAIL_UTOPRIM_EE, 0, i, j, k);
        for (int l = 0; l < NP; l++) pf[l] = pf2[l];
        gamma = *p_gamma;
        fail_set(P, fr, NFAIL_USE_FULL_INV, 0, i, j, k);
      }
    }
  }
  *p_gamma = gamma;

  // fixup_floor(), src/fixup.c:110-121
  int fl = 0;
  if (pf[RHO] < rhoflr) { pf[RHO] = rhoflr; fl = 1; }
  if (pf[UU] < uuflr) { pf[UU] = uuflr; fl = 1; }
  if (fl) fail_set(P, fr, NFAIL_FLOOR, 0, i, j, k);

  if (check_gamma(gamma)) {
    fail_set(P, fr, NFAIL_GAMMA_MAX, 2, i, j, k);
    return;
  }
  *p_gamma = gamma;
}

HD double floor_rho(const Params& P, const Ptrs& X, int i, long long c) {
  return (P.floor_dim == 1) ? X.rhoflr[i] : X.rhoflr[c];
}
HD double floor_uu(const Params& P, const Ptrs& X, int i, long long c) {
  return (P.floor_dim == 1) ? X.uuflr[i] : X.uuflr[c];
}

// --------------------------------------------------------------------------
// Per-zone part of advance(), src/step_ch.c:312-443: conserved update and
// recovery of the primitives for interior zone (i,j,k).
//   stage 0 (predictor): p_i = p_h = p, result into ph, guess = p before the
//                        rescale round trip (the reference's ph copy)
//   stage 1 (corrector): p_h = ph, U^n recomputed from p, guess = p, result into p
// MIXED reads the FP32 face fluxes and EMFs and sums the divergence with
// compensation; otherwise the FP64 arrays with the reference's sum.
//
// The work is split so the device can run it in two passes: zone_conserved()
// + recover_fast() finish every zone whose 2D Newton inversion succeeds
// without the entropy fix; the others keep U^{n+1} and go through the full
// recover_primitives() tree in a second, compacted pass (recover_zone()).

// U^{n+1} of the zone (src/step_ch.c:330-375)
// gn: geometry of U^n (time t); g: geometry of the stage (source terms).  With a
// static metric both are the zone's one record.
template <bool MIXED, int MET = -1, bool UNI = false, class GP = const Geom*>
HD void zone_conserved(const Params& P, const Ptrs& X, int stage, double Dt, int i, int j, int k,
                       GP gn, GP g, FailRec& fr, double* U, double* pn_out = nullptr, bool want_pn = false) {
  const long long NT = P.nt;
  const long long c = cidx(P, i, j, k);
  // U^n = primtoflux(p_i) with geom_i (src/step_ch.c:319, :339-341): kept for
  // the corrector as KEEP_CONSERVED_VARS keeps U_gf (src/step_ch.c:321-333), or,
  // in the mixed-precision step on a static metric, formed again by the
  // corrector from p^n (which it reads for the Newton guess anyway): the same
  // operations on the same values, so the same bits, for 128 B less traffic
  // per zone and step.  Its gamma_calc failure is counted once, as the
  // reference counts it, in the predictor.
  const bool keep_un = !MIXED || ((MET < 0) ? P.dyn != 0 : MET != 0);
  double pi[NP];
  if (stage == 0 || !keep_un) {
#pragma unroll
    for (int v = 0; v < NP; v++) pi[v] = rt_val(P, X, v, i, X.p[v*NT + c]);
    State q;
    const int gfail = get_state(pi, gn, P.gam, &q);
    if (gfail && stage == 0) fr.nf[NFAIL_GAMMA_CALC]++;
    primtoflux(pi, &q, 0, gn, U);
    if (stage == 0 && keep_un) {
#pragma unroll
      for (int v = 0; v < NP; v++) X.Un[v*NT + c] = U[v];
    }
    if (want_pn) {
#pragma unroll
      for (int v = 0; v < NP; v++) pn_out[v] = pi[v];
    }
  } else {
#pragma unroll
    for (int v = 0; v < NP; v++) U[v] = X.Un[v*NT + c];
  }
  double dU[NP];
  if constexpr (MET == 2) {
    source<2>(P, X, nullptr, g, i, j, k, dU);
  } else {
    double phz[NP];
    if (MIXED && stage == 0) {
      // the predictor's source terms are evaluated at p^n: the values just
      // read (the mixed update is 4 % faster for it; the FP64 update, whose
      // registers are tighter, 6 % slower, and reads p^n again)
#pragma unroll
      for (int v = 0; v < NP; v++) phz[v] = pi[v];
    } else {
      const double* src = (stage == 0) ? X.p : X.ph;
#pragma unroll
      for (int v = 0; v < NP; v++) phz[v] = rt_val(P, X, v, i, src[v*NT + c]);
    }
    fr.nf[NFAIL_GAMMA_CALC] += source<MET, UNI>(P, X, phz, g, i, j, k, dU);
  }
  if (MIXED) {
#pragma unroll
    for (int l = 0; l < NP; l++) {
      CompSum acc;
      acc.add(-(face_flux_mixed(P, X, 0, l, i + 1, j, k) - face_flux_mixed(P, X, 0, l, i, j, k)) * P.invdx[1]);
      acc.add(-(face_flux_mixed(P, X, 1, l, i, j + 1, k) - face_flux_mixed(P, X, 1, l, i, j, k)) * P.invdx[2]);
      acc.add(-(face_flux_mixed(P, X, 2, l, i, j, k + 1) - face_flux_mixed(P, X, 2, l, i, j, k)) * P.invdx[3]);
      acc.add(dU[l]);
      U[l] += Dt*acc.value();
    }
  } else {
#pragma unroll
    for (int l = 0; l < NP; l++) {
      double f0m = face_flux(P, X, 0, l, i, j, k), f0p = face_flux(P, X, 0, l, i + 1, j, k);
      double f1m = face_flux(P, X, 1, l, i, j, k), f1p = face_flux(P, X, 1, l, i, j + 1, k);
      double f2m = face_flux(P, X, 2, l, i, j, k), f2p = face_flux(P, X, 2, l, i, j, k + 1);
      U[l] += Dt*(
          - (f0p - f0m) * P.invdx[1]
          - (f1p - f1m) * P.invdx[2]
          - (f2p - f2m) * P.invdx[3]
          + dU[l]);
    }
  }
}

// Newton guess of recover_primitives(): the predictor's result array holds p
// before the rescale round trip, the corrector's holds p after it.
HD void zone_guess(const Params& P, const Ptrs& X, int stage, int i, long long c, double* pf) {
  const long long NT = P.nt;
#pragma unroll
  for (int v = 0; v < NP; v++) {
    const double x = X.p[v*NT + c];
    pf[v] = (stage == 0) ? x : rt_val(P, X, v, i, x);
  }
}

// The common branch of recover_primitives() (src/step_ch.c:832-1020): the 2D
// inversion succeeds with an acceptable gamma and the entropy fix is not
// needed.  Returns false, with pf, fr and the zone untouched in memory, when
// the zone needs the full tree.
template <class GP>
HD bool recover_fast(const Params& P, const Ptrs& X, int i, int j, int k, double* pf, const double* U,
                     GP g, double rhoflr, double uuflr, double* p_gamma, FailRec& fr) {
  double pt[NP];
#pragma unroll
  for (int l = 0; l < NP; l++) pt[l] = pf[l];
  double gamma = 1., bsq = 0.;
  int ret = utoprim_2d_fast(U, g, P.gam, P.gam_m1_o_gam, pt, &gamma, &bsq, nullptr, P.n_extra_newt);
  if (ret || check_gamma(gamma) == 2) return false;
  if ((pt[UU] < BETA_MIN * bsq) || (pt[RHO] < BETA_MIN * bsq)) return false;
#pragma unroll
  for (int l = 0; l < NP; l++) pf[l] = pt[l];
  fail_set(P, fr, NFAIL_USE_FULL_INV, 0, i, j, k);
  int fl = 0;
  if (pf[RHO] < rhoflr) { pf[RHO] = rhoflr; fl = 1; }
  if (pf[UU] < uuflr) { pf[UU] = uuflr; fl = 1; }
  if (fl) fail_set(P, fr, NFAIL_FLOOR, 0, i, j, k);
  *p_gamma = gamma;
  if (check_gamma(gamma)) fail_set(P, fr, NFAIL_GAMMA_MAX, 2, i, j, k);
  return true;
}

HD void store_zone(const Params& P, const Ptrs& X, int stage, long long c, const double* pf, double pgam,
                   const FailRec& fr) {
  const long long NT = P.nt;
  double* dst = (stage == 0) ? X.ph : X.p;
#pragma unroll
  for (int v = 0; v < NP; v++) dst[v*NT + c] = pf[v];
  X.gamma[c] = pgam;
  if (fr.pflag) X.pflag[c] = fr.pflag;
}

// Full recover_primitives() tree for a zone whose U^{n+1} is known.
// fo: the stage's face records are evaluated (binary metric) for the
// entropy fluxes.
template <bool MIXED>
HD void recover_zone(const Params& P, const Ptrs& X, int stage, double dt, int i, int j, int k, const double* U,
                     FailRec& fr, const FaceOtf* fo = nullptr) {
  const long long NT = P.nt;
  const long long c = cidx(P, i, j, k);
  const double Dt = (stage == 0) ? 0.5*dt : dt;
  Geom gb;
  const Geom* g = cent_rec(P, X, i, j, k, gb);
  double pf[NP], p_raw[NP];
  zone_guess(P, X, stage, i, c, pf);
#pragma unroll
  for (int v = 0; v < NP; v++) p_raw[v] = X.p[v*NT + c];
  double pgam = 1.;
  recover_primitives<MIXED>(P, X, i, j, k, Dt, pf, U, p_raw, g, floor_rho(P, X, i, c), floor_uu(P, X, i, c),
                            &pgam, (stage == 0) ? X.p : X.ph, fo, fr);
  store_zone(P, X, stage, c, pf, pgam, fr);
}

// Whole zone in one pass (the host build's sequence; the device runs the two
// passes).  Returns false and leaves U filled when the full tree is needed.
// gn, g, gnew: the zone's records at U^n's level, the stage's and the result's.
template <bool MIXED, int MET = -1, bool UNI = false, class GP = const Geom*>
HD bool update_zone_fast_g(const Params& P, const Ptrs& X, int stage, double dt, int i, int j, int k,
                           FailRec& fr, double* U, GP gn, GP g, GP gnew) {
  const long long c = cidx(P, i, j, k);
  const double Dt = (stage == 0) ? 0.5*dt : dt;
  double pf[NP];
  // the mixed corrector forms U^n from p^n, which is also its Newton guess:
  // the values already read
  const bool reuse = MIXED && stage == 1 && !((MET < 0) ? P.dyn != 0 : MET != 0);
  zone_conserved<MIXED, MET, UNI>(P, X, stage, Dt, i, j, k, gn, g, fr, U, pf, reuse);
  if (!reuse) zone_guess(P, X, stage, i, c, pf);
  double pgam = 1.;
  if (!recover_fast(P, X, i, j, k, pf, U, gnew, floor_rho(P, X, i, c), floor_uu(P, X, i, c), &pgam, fr))
    return false;
  store_zone(P, X, stage, c, pf, pgam, fr);
  return true;
}

template <bool MIXED, int MET = -1, bool UNI = false>
HD bool update_zone_fast(const Params& P, const Ptrs& X, int stage, double dt, int i, int j, int k,
                         FailRec& fr, double* U) {
  const int gid = geom_id(P, i, j, k, CENT);
#ifdef __HIP_DEVICE_COMPILE__
  if constexpr (UNI) {
    // the wavefront's zones share one record: read through the scalar cache
    typedef const __attribute__((address_space(4))) Geom* CGp;
    const int gu = __builtin_amdgcn_readfirstlane(gid);
    return update_zone_fast_g<MIXED, MET, UNI>(P, X, stage, dt, i, j, k, fr, U, (CGp)(X.geom_n + gu),
                                               (CGp)(X.geom + gu), (CGp)(X.geom_new + gu));
  }
#endif
  return update_zone_fast_g<MIXED, MET, UNI>(P, X, stage, dt, i, j, k, fr, U, X.geom_n + gid, X.geom + gid,
                                             X.geom_new + gid);
}

template <bool MIXED>
// Returns U^{n+1}[RHO], the zone's term of M_tot (src/step_ch.c:389).
HD double update_zone(const Params& P, const Ptrs& X, int stage, double dt, int i, int j, int k,
                      FailRec& fr) {
  double U[NP];
  if (!update_zone_fast<MIXED>(P, X, stage, dt, i, j, k, fr, U)) {
    Ptrs Xn = X;
    Xn.geom = X.geom_new;   // the full tree works at the result's time level
    recover_zone<MIXED>(P, Xn, stage, dt, i, j, k, U, fr);
  }
  return U[RHO];
}

// --------------------------------------------------------------------------
// Boundary conditions, src/bounds.c:31-579 (BC_CARTESIAN_OUTFLOW :58-133,
// BC_SPHERICAL_OUTFLOW3 :492-579) and inflow_check(), src/bounds.c:2304-2380.

HD void inflow_check(const Params& P, const Ptrs& X, double* pr, int ii, int jj, int kk, int type,
                     FailRec& fr) {
  const int dir = 1;
  Geom gb;
  const Geom* g = cent_rec(P, X, ii, jj, kk, gb);
  double gamma;
  if (gamma_calc(pr, g, &gamma)) fr.nf[NFAIL_GAMMA_CALC]++;
  double ucon[4];
  ucon[0] = gamma*g->ncon[0];
  ucon[1] = pr[U1] - ucon[0]*g->beta[0];
  ucon[2] = pr[U2] - ucon[0]*g->beta[1];
  ucon[3] = pr[U3] - ucon[0]*g->beta[2];
  if (((ucon[dir] > 0.) && (type == 0)) || ((ucon[dir] < 0.) && (type == 1))) {
    gamma = g->alpha * ucon[0];
    pr[U1] /= gamma;
    pr[U2] /= gamma;
    pr[U3] /= gamma;
    pr[U1 + dir - 1] = g->beta[dir - 1]*g->ncon[0];
    double vsq = 0.;
    for (int a = 1; a < 4; a++) for (int b = 1; b < 4; b++)
      vsq += g->gcov[a][b]*pr[U1 + a - 1]*pr[U1 + b - 1];
    if (fabs(vsq) < 1.e-13) vsq = fabs(vsq);
    if (vsq < 0.) vsq = 0.;
    if (vsq >= 1.) vsq = 1. - 1./(GAMMAMAX*GAMMAMAX);
    gamma = 1./sqrt(1. - vsq);
    pr[U1] *= gamma;
    pr[U2] *= gamma;
    pr[U3] *= gamma;
  }
}

// One ghost zone of direction dir.  t enumerates (g, transverse...) for the
// lower (side 0) and upper (side 1) faces.
HD void bound_zone(const Params& P, const Ptrs& X, double* S, int dir, int side, int g, int a, int b,
                   int do_pflag, FailRec& fr) {
  // a face shared with another block gets its ghosts from that block
  if ((P.bc_nbr >> (2*dir + side)) & 1) return;
  const long long NT = P.nt;
  const int n1e = P.n1 + NG - 1, n2e = P.n2 + NG - 1, n3e = P.n3 + NG - 1;
  int di, dj, dk, si, sj, sk;
  bool reflect = false;
  if (dir == 0) {
    // a = j in [N2S,N2E], b = k in [N3S,N3E]
    dj = sj = a; dk = sk = b;
    di = side ? (n1e + 1 + g) : (NG - 1 - g);
    si = side ? n1e : NG;
  } else if (dir == 1) {
    // a = i (all), b = k in [N3S,N3E]
    di = si = a; dk = sk = b;
    dj = side ? (n2e + 1 + g) : (NG - 1 - g);
    if (P.bc == BC_SPHERICAL_OUTFLOW3) { sj = side ? (n2e - g) : (NG + g); reflect = true; }
    else sj = side ? n2e : NG;
  } else {
    // a = i (all), b = j (all)
    di = si = a; dj = sj = b;
    dk = side ? (n3e + 1 + g) : (NG - 1 - g);
    if (P.bc == BC_SPHERICAL_OUTFLOW3) sk = side ? (NG + g) : (n3e - g);
    else sk = side ? n3e : NG;
  }
  const long long cd = cidx(P, di, dj, dk), cs = cidx(P, si, sj, sk);
  double pr[NP];
  for (int v = 0; v < NP; v++) pr[v] = S[v*NT + cs];
  if (reflect) { pr[U2] *= -1; pr[B2] *= -1; }
  if (dir == 0 && P.bc == BC_SPHERICAL_OUTFLOW3) inflow_check(P, X, pr, di, dj, dk, side, fr);
  for (int v = 0; v < NP; v++) S[v*NT + cd] = pr[v];
  if (do_pflag) X.pflag[cd] = X.pflag[cs];
}

// --------------------------------------------------------------------------
// fixup_gamma(), src/fixup.c:148-158
HD void fixup_gamma(double* pv, double* gamma, FailRec& fr) {
  double f = sqrt((GAMMAMAX - 1.)*(GAMMAMAX + 1.) / ((*gamma - 1.)*(*gamma + 1.)));
  pv[U1] *= f;
  pv[U2] *= f;
  pv[U3] *= f;
  *gamma = GAMMAMAX;
  fr.nf[NFAIL_GAMMA_MAX]++;
}

// fixup_interp_v() (which == PFLAG_INTERP_V, src/fixup.c:659-761) and
// fixup_interp_prim() (which == PFLAG_INTERP_PRIM, src/fixup.c:499-610).
HD void fixup_interp_zone(const Params& P, const Ptrs& X, double* pv, int which, int i, int j, int k,
                          FailRec& fr) {
  const long long NT = P.nt;
  const long long c = cidx(P, i, j, k);
  if (X.pflag[c] != which) return;
  const int o = 13;
  const int ngood_min[4] = {13, 9, 6, 4};
  int pfl[27];
  for (int ii = -1; ii <= 1; ii++) for (int jj = -1; jj <= 1; jj++) for (int kk = -1; kk <= 1; kk++) {
    int ind1 = ii + 3*(jj + 3*kk);
    pfl[o + ind1] = !X.pflag[cidx(P, i + ii, j + jj, k + kk)];
  }
  int ngood = 0;
  for (int ii = 1; ii <= 13; ii++) {
    if (pfl[o + ii] && pfl[o - ii]) ngood++;
    else pfl[o + ii] = pfl[o - ii] = 0;
  }
  Geom gb;
  const Geom* g = cent_rec(P, X, i, j, k, gb);
  const int lbeg = (which == PFLAG_INTERP_V) ? U1 : RHO;
  double p[NP], pold[NP];
  for (int v = 0; v < NP; v++) p[v] = pv[v*NT + c];
  for (int v = 0; v < NP; v++) pold[v] = p[v];
  double gamma;
  if (ngood < ngood_min[P.n_spatial_dims]) {
    if (which == PFLAG_INTERP_V) {
      double gm = X.gamma[c];
      fixup_gamma(p, &gm, fr);
      X.gamma[c] = gm;
      fr.nf[NFAIL_INTERP_V]++;
    } else {
      double gm;
      gamma_calc(p, g, &gm);
      X.gamma[c] = gm;
      fr.nf[NFAIL_INTERP_PRIM]++;
    }
  } else {
    for (int l = lbeg; l <= U3; l++) p[l] = 0.;
    for (int ii = -1; ii <= 1; ii++) for (int jj = -1; jj <= 1; jj++) for (int kk = -1; kk <= 1; kk++) {
      if (pfl[o + ii + 3*(jj + 3*kk)]) {
        long long cn = cidx(P, i + ii, j + jj, k + kk);
        for (int l = lbeg; l <= U3; l++) p[l] += pv[l*NT + cn];
      }
    }
    ngood *= 2;
    for (int l = lbeg; l <= U3; l++) p[l] /= ngood;
    int i_ret = gamma_calc(p, g, &gamma);
    if (check_gamma(gamma)) {
      double gamma_old = gamma;
      gamma = 0.;
      int ngood2 = 0;
      for (int ii = -1; ii <= 1; ii++) for (int jj = -1; jj <= 1; jj++) for (int kk = -1; kk <= 1; kk++) {
        if (pfl[o + ii + 3*(jj + 3*kk)]) {
          long long cn = cidx(P, i + ii, j + jj, k + kk);
          double pn[NP];
          for (int v = 0; v < NP; v++) pn[v] = pv[v*NT + cn];
          double f;
          Geom gn;
          i_ret = gamma_calc(pn, cent_rec(P, X, i + ii, j + jj, k + kk, gn), &f);
          if (!i_ret) { gamma += f; ngood2++; }
        }
      }
      gamma /= ngood2;
      double f = sqrt((gamma - 1.)*(gamma + 1.) / ((gamma_old - 1.)*(gamma_old + 1.)));
      p[U1] *= f;
      p[U2] *= f;
      p[U3] *= f;
      i_ret = gamma_calc(p, g, &gamma);
    }
    if (which == PFLAG_INTERP_V) {
      if (i_ret || check_gamma(gamma)) {
        for (int l = U1; l <= U3; l++) p[l] = pold[l];
        double gm = X.gamma[c];
        fixup_gamma(p, &gm, fr);
        X.gamma[c] = gm;
        fr.nf[NFAIL_INTERP_V]++;
      } else {
        fr.nf[NFAIL_USE_INTERP_V]++;
        X.gamma[c] = gamma;
      }
    } else {
      if (i_ret) {
        for (int l = U1; l <= U3; l++) p[l] = pold[l];
        gamma_calc(p, g, &gamma);
        fr.nf[NFAIL_INTERP_PRIM]++;
      }
      fr.nf[NFAIL_USE_INTERP_PRIM]++;
      if ((!i_ret) && check_gamma(gamma)) fixup_gamma(p, &gamma, fr);
      X.gamma[c] = gamma;
    }
  }
  for (int v = 0; v < NP; v++) pv[v*NT + c] = p[v];
}

// --------------------------------------------------------------------------
// fixup(), src/fixup.c:8-24 with fixup1zone(), :35-66 (initial data)
HD void fixup_init_zone(const Params& P, const Ptrs& X, int i, int j, int k, FailRec& fr) {
  const long long NT = P.nt;
  const long long c = cidx(P, i, j, k);
  Geom gb;
  const Geom* g = cent_rec(P, X, i, j, k, gb);
  double p[NP];
  for (int v = 0; v < NP; v++) p[v] = X.p[v*NT + c];
  double gamma;
  if (gamma_calc(p, g, &gamma)) { fr.nf[NFAIL_GAMMA_CALC]++; gamma = 1.; }
  int fl = 0;
  double rf = floor_rho(P, X, i, c), uf = floor_uu(P, X, i, c);
  if (p[RHO] < rf) { p[RHO] = rf; fl = 1; }
  if (p[UU] < uf) { p[UU] = uf; fl = 1; }
  if (fl) fr.nf[NFAIL_FLOOR]++;
  if (check_gamma(gamma)) fixup_gamma(p, &gamma, fr);
  for (int v = 0; v < NP; v++) X.p[v*NT + c] = p[v];
}

// --------------------------------------------------------------------------
// timestep(), src/step_ch.c:546-623 (TIMESTEP_ORIGINAL); also advances t as
// step_ch() does at src/step_ch.c:218-223.  One thread.
// check_for_crash(), src/step_ch.c:1244-1339, once per step on the device: the
// step's new entropy-solve failures (ntot1) and interpolation-fixup failures
// (ntot2) from the running nfailtot words (a reset shows as a decrease: no new
// failures), n_bad = max of the two, against nfail_fraction = 0.5 of N0 = 2
// time levels of the block's cells (src/macros.h:36).  The reference ends the
// run there (fail(FAIL_BASIC)); the device latches the step and n_bad, which
// the host reads with hgpu_get_crash().
HD void crash_check(const Params& P, const Ptrs& X) {
  unsigned long long* st = X.status;
  const unsigned long long n1 = st[ST_NFAILTOT + NFAIL_UTOPRIM_EE];
  const unsigned long long n2 = st[ST_NFAILTOT + NFAIL_INTERP_PRIM] + st[ST_NFAILTOT + NFAIL_INTERP_V];
  const unsigned long long d1 = (n1 < st[ST_CRASH_N1]) ? 0ull : n1 - st[ST_CRASH_N1];
  const unsigned long long d2 = (n2 < st[ST_CRASH_N2]) ? 0ull : n2 - st[ST_CRASH_N2];
  st[ST_CRASH_N1] = n1;
  st[ST_CRASH_N2] = n2;
  st[ST_NSTEP] += 1ull;
  const unsigned long long n_bad = (d1 > d2) ? d1 : d2;
  const double ncells = 2. * (double)P.n1 * (double)P.n2 * (double)P.n3;
  if ((double)n_bad / ncells > 0.5 && st[ST_CRASH] == 0ull) {
    st[ST_CRASH] = st[ST_NSTEP];
    st[ST_CRASH_NBAD] = n_bad;
  }
}

HD void timestep_update(const Params& P, const Ptrs& X) {
  double* sc = X.scal;
  sc[SC_T_OLD] = sc[SC_T];
  sc[SC_T] += sc[SC_DT];
  sc[SC_DT_OLD] = sc[SC_DT];
  double c1 = bits_dbl(X.status[ST_VMAX + 1]);
  double c2 = bits_dbl(X.status[ST_VMAX + 2]);
  double c3 = bits_dbl(X.status[ST_VMAX + 3]);
  double dtsqij = sqrt((1./3.)/(
      (c1*c1)/(P.dx[1]*P.dx[1]) +
      (c2*c2)/(P.dx[2]*P.dx[2]) +
      (c3*c3)/(P.dx[3]*P.dx[3])));
  double dtt = P.cour*dtsqij;
  if (dtt > 1.3*sc[SC_DT]) dtt = 1.3*sc[SC_DT];
  if (sc[SC_T] + dtt > P.t_final) dtt = P.t_final - sc[SC_T];
  sc[SC_DT] = dtt;
  crash_check(P, X);
}

// --------------------------------------------------------------------------
// divb_calc(), src/diag.c:503-602: |div B| at the corner of zone (i,j,k)
HD double divb_corner(const Params& P, const Ptrs& X, const double* S, int i, int j, int k) {
  const long long NT = P.nt;
  double det_g[8], b1[8], b2[8], b3[8];
  int id = 0;
  for (int di = -1; di <= 0; di++) for (int dj = -1; dj <= 0; dj++) for (int dk = -1; dk <= 0; dk++) {
    long long c = cidx(P, i + di, j + dj, k + dk);
    Geom gb;
    det_g[id] = cent_rec(P, X, i + di, j + dj, k + dk, gb)->g;
    double B1v = S[B1*NT + c], B2v = S[B2*NT + c], B3v = S[B3*NT + c];
    b1[id] = (di == 0) ? B1v : -B1v;
    b2[id] = (dj == 0) ? B2v : -B2v;
    b3[id] = (dk == 0) ? B3v : -B3v;
    id++;
  }
  double divb = 0., ftmp = 0.;
  for (id = 0; id < 8; id++) ftmp += b1[id]*det_g[id];
  divb += ftmp*P.invdx[1];
  ftmp = 0.;
  for (id = 0; id < 8; id++) ftmp += b2[id]*det_g[id];
  divb += ftmp*P.invdx[2];
  ftmp = 0.;
  for (id = 0; id < 8; id++) ftmp += b3[id]*det_g[id];
  divb += ftmp*P.invdx[3];
  divb *= 0.25;
  return fabs(divb);
}

// --------------------------------------------------------------------------
// Index boxes of the kernels (inclusive bounds)
struct Box {
  int lo[3], hi[3];
  HD long long count() const {
    return (long long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
  }
  HD void at(long long t, int& i, int& j, int& k) const {
    const int nk = hi[2] - lo[2] + 1, nj = hi[1] - lo[1] + 1;
    k = lo[2] + (int)(t % nk);
    t /= nk;
    j = lo[1] + (int)(t % nj);
    i = lo[0] + (int)(t / nj);
  }
};

HD void dim_n(const Params& P, int* n) { n[0] = P.n1; n[1] = P.n2; n[2] = P.n3; }

// numerical_flux() face range, src/step_ch.c:656-664
HD Box flux_box(const Params& P, int d) {
  int n[3]; dim_n(P, n);
  Box b;
  for (int e = 0; e < 3; e++) {
    if (e == d) { b.lo[e] = NG; b.hi[e] = NG + n[e]; }
    else if (P.ct_para) { b.lo[e] = 0; b.hi[e] = n[e] + 2*NG - 1; }
    else { b.lo[e] = NG - 1; b.hi[e] = NG + n[e]; }
  }
  return b;
}
// corner EMFs, src/flux_ct.c:31 and :780-846
HD Box edge_box(const Params& P) {
  Box b;
  b.lo[0] = b.lo[1] = b.lo[2] = NG;
  b.hi[0] = NG + P.n1; b.hi[1] = NG + P.n2; b.hi[2] = NG + P.n3;
  return b;
}
HD Box interior_box(const Params& P) {
  Box b;
  b.lo[0] = b.lo[1] = b.lo[2] = NG;
  b.hi[0] = NG + P.n1 - 1; b.hi[1] = NG + P.n2 - 1; b.hi[2] = NG + P.n3 - 1;
  return b;
}
// ghost layers of direction dir, both sides: (side, g, a, b) flattened as
// lo = {0, 0, a0, b0}: index box over (side*NG + g, a, b)
HD Box ghost_box(const Params& P, int dir) {
  Box b;
  b.lo[0] = 0; b.hi[0] = 2*NG - 1;
  if (dir == 0) { b.lo[1] = NG; b.hi[1] = NG + P.n2 - 1; b.lo[2] = NG; b.hi[2] = NG + P.n3 - 1; }
  else if (dir == 1) { b.lo[1] = 0; b.hi[1] = P.n1t - 1; b.lo[2] = NG; b.hi[2] = NG + P.n3 - 1; }
  else { b.lo[1] = 0; b.hi[1] = P.n1t - 1; b.lo[2] = 0; b.hi[2] = P.n2t - 1; }
  return b;
}
// interior corners for div B: (N1S+1..N1E, ...) so that all 8 zones are interior
HD Box divb_box(const Params& P) {
  Box b;
  b.lo[0] = NG + (P.n1 > 1); b.hi[0] = NG + P.n1 - 1;
  b.lo[1] = NG + (P.n2 > 1); b.hi[1] = NG + P.n2 - 1;
  b.lo[2] = NG + (P.n3 > 1); b.hi[2] = NG + P.n3 - 1;
  return b;
}

}  // namespace hgpu
