// hgpu_mixed.h -- mixed-precision variant of the face fluxes and corner EMFs.
//
// Reconstruction, the face Riemann solve (HLL or Lax-Friedrichs) and the EMF
// averaging run in FP32; everything that integrates (conserved update,
// con-to-prim, metric, connection, cooling) stays FP64.  To keep the FP32
// rounding relative to the small quantities:
//   * a face state is the FP64 zone value plus an FP32 offset computed from
//     FP32 differences of neighbouring FP64 zone values (PPM/MC on
//     differences), so smooth large-valued profiles lose no digits;
//   * the energy flux and energy density T^mu_t + rho u^mu are formed
//     without the rest-mass cancellation: with u_t = -alpha gamma + beta_i u~^i,
//     1 + u_t = gamma (1 - alpha) - (gamma - 1) + beta_i u~^i, where
//     gamma - 1 = |u~|^2 / (gamma + 1) and 1 - alpha comes from FP64;
//   * fluxes are stored in FP32 and the flux divergence is summed in FP64
//     with Neumaier compensation (update_zone<true>).
// The algorithms are the FP64 ones of hgpu_phys.h / hgpu_bodies.h (the
// reference's recon.c, phys.c, step_ch.c, flux_ct.c) evaluated in float.
#pragma once
#include "hgpu_phys.h"

namespace hgpu {

// The device build of this file's arithmetic fuses a multiply and an add
// written in one expression into one FMA (one rounding instead of two, and
// fewer instructions; within an expression only, so every inlined copy of a
// body fuses alike); the FP64 code elsewhere keeps -ffp-contract=off so it
// stays bit-identical to the reference build.  The compensated FP64 sum at
// the end of the file is excluded (Neumaier's correction needs its
// separately rounded steps).
#ifdef __HIP_DEVICE_COMPILE__
#pragma clang fp contract(on)
#endif

// ---------------------------------------------------------------------------
// FP32 reciprocal, square root and power with the hardware's single-instruction
// forms on the device (v_rcp_f32, v_sqrt_f32: about 1 ulp; v_log_f32 and
// v_exp_f32 for x^a = 2^(a log2 x), relative error about
// |a log2 x| 2^-23 ln 2).  The FP32 face solve's own rounding is of the same
// order; the host build uses the libm forms.
HD float rcp_f(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.f/x;
#endif
}
HD float sqrt_f(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}
HD float pow_f(float x, float a) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_exp2f(a*__builtin_amdgcn_logf(x));
#else
  return powf(x, a);
#endif
}

// ---------------------------------------------------------------------------
// FP32 restatements of get_state()/primtoflux()/vchar_fast() (src/phys.c)
struct StateF {
  float ucon[4], ucov[4], bcon[4], bcov[4];
  float bsq, p, eta, ptot;
  float opu_t;     // 1 + u_t without cancellation
};

HD void lower_f(const float* v, const GeomF* g, float* w) {
#pragma unroll
  for (int m = 0; m < 4; m++) w[m] = g->gcov[m][0]*v[0] + g->gcov[m][1]*v[1] + g->gcov[m][2]*v[2] + g->gcov[m][3]*v[3];
}

HD int gamma_calc_f(const float* pr, const GeomF* g, float* gamma, float* qsq_out) {
  float qsq = g->gcov[1][1]*pr[U1]*pr[U1] + g->gcov[2][2]*pr[U2]*pr[U2] + g->gcov[3][3]*pr[U3]*pr[U3]
            + 2.f*(g->gcov[1][2]*pr[U1]*pr[U2] + g->gcov[1][3]*pr[U1]*pr[U3] + g->gcov[2][3]*pr[U2]*pr[U3]);
  if (qsq < 0.f) {
    if (fabsf(qsq) > 1.E-10f) { *gamma = 1.f; *qsq_out = 0.f; return 1; }
    qsq = 1.E-10f;
  }
  *gamma = sqrt_f(1.f + qsq);
  *qsq_out = qsq;
  return 0;
}

HD int get_state_f(const float* pr, const GeomF* g, float gam, StateF* q) {
  float gamma, qsq;
  int fl = gamma_calc_f(pr, g, &gamma, &qsq);
  q->opu_t = gamma*g->om_alpha - qsq*rcp_f(gamma + 1.f)
           + (g->gcov[0][1]*pr[U1] + g->gcov[0][2]*pr[U2] + g->gcov[0][3]*pr[U3]);
  q->ucon[0] = gamma*g->ncon[0];
  q->ucon[1] = pr[U1] - q->ucon[0]*g->beta[0];
  q->ucon[2] = pr[U2] - q->ucon[0]*g->beta[1];
  q->ucon[3] = pr[U3] - q->ucon[0]*g->beta[2];
  lower_f(q->ucon, g, q->ucov);
  float inv_ut = rcp_f(q->ucon[0]);
  q->bcon[0] = pr[B1]*q->ucov[1] + pr[B2]*q->ucov[2] + pr[B3]*q->ucov[3];
  q->bcon[1] = (pr[B1] + q->bcon[0]*q->ucon[1])*inv_ut;
  q->bcon[2] = (pr[B2] + q->bcon[0]*q->ucon[2])*inv_ut;
  q->bcon[3] = (pr[B3] + q->bcon[0]*q->ucon[3])*inv_ut;
  lower_f(q->bcon, g, q->bcov);
  q->bsq = q->bcon[0]*q->bcov[0] + q->bcon[1]*q->bcov[1] + q->bcon[2]*q->bcov[2] + q->bcon[3]*q->bcov[3];
  q->p = (gam - 1.f)*pr[UU];
  q->eta = q->p + pr[RHO] + pr[UU] + q->bsq;
  q->ptot = q->p + 0.5f*q->bsq;
  return fl;
}

template <int dir>
HD void primtoflux_f(const float* pr, const StateF* q, const GeomF* g, float* flux) {
  const float eu = q->eta*q->ucon[dir], bd = q->bcon[dir];
  float mhd[4];
#pragma unroll
  for (int m = 0; m < 4; m++) mhd[m] = eu*q->ucov[m] - bd*q->bcov[m];
  mhd[dir] += q->ptot;
  flux[RHO] = pr[RHO]*q->ucon[dir];
  // T^dir_t + rho u^dir = rho u^dir (1 + u_t) + (u + p + b^2) u^dir u_t + ptot delta^dir_t - b^dir b_t
  flux[UU] = flux[RHO]*q->opu_t + (q->eta - pr[RHO])*q->ucon[dir]*q->ucov[0] - bd*q->bcov[0]
           + ((dir == 0) ? q->ptot : 0.f);
  flux[U1] = mhd[1];
  flux[U2] = mhd[2];
  flux[U3] = mhd[3];
  flux[B1] = q->bcon[1]*q->ucon[dir] - q->bcon[dir]*q->ucon[1];
  flux[B2] = q->bcon[2]*q->ucon[dir] - q->bcon[dir]*q->ucon[2];
  flux[B3] = q->bcon[3]*q->ucon[dir] - q->bcon[dir]*q->ucon[3];
#pragma unroll
  for (int l = 0; l < NP; l++) flux[l] *= g->g;
}

template <int js>
HD void vchar_f(const StateF* q, const GeomF* g, float gam, float* vmax, float* vmin) {
  float cms2 = (gam*q->p + q->bsq)*rcp_f(q->eta);
  if (cms2 < 0.f) cms2 = (float)SMALL;
  else if (cms2 > 1.f) cms2 = 1.f;
  const float Au2 = q->ucon[js]*q->ucon[js], Bu2 = q->ucon[0]*q->ucon[0], AuBu = q->ucon[js]*q->ucon[0];
  const float A = Bu2 - (g->gcon[0][0] + Bu2)*cms2;
  const float B = AuBu - (g->gcon[0][js] + AuBu)*cms2;
  const float C = Au2 - (g->gcon[js][js] + Au2)*cms2;
  float discr = B*B - A*C;
  if (discr < 0.f) discr = 0.f;
  discr = sqrt_f(discr);
  const float denom = rcp_f(A);
  const float vp = -(-B + discr)*denom, vm = -(-B - discr)*denom;
  if (vp > vm) { *vmax = vp; *vmin = vm; }
  else { *vmax = vm; *vmin = vp; }
}

// ---------------------------------------------------------------------------
// Reconstruction on differences.  D[m] = y[m+1] - y[m].

// MC slope of the middle point from its two one-sided differences
// (linear_recon_vect(), src/recon.c:237-264)
HD float mc_slope_diff(float a, float b) {
  if (a*b <= 0.f) return 0.f;
  const float aa = fabsf(a), ab = fabsf(b);
  if (aa + aa + aa <= ab) return a + a;
  if (ab + ab + ab <= aa) return b + b;
  return 0.5f*(a + b);
}

// para_vect() limiter (src/recon.c:611-641) on edge values l, r taken
// relative to the zone value
HD void ppm_limit_rel(float l, float r, float* lnew, float* rnew) {
  if (-r*l <= 0.f) { *lnew = *rnew = 0.f; return; }
  const float r2 = r + r, l2 = l + l;
  *lnew = l; *rnew = r;
  if (r > l) {
    if (0.f > (l + r2)) *lnew = -r2;
    if (0.f < (l2 + r)) *rnew = -l2;
  } else {
    if (0.f < (l + r2)) *lnew = -r2;
    if (0.f > (l2 + r)) *rnew = -l2;
  }
}

// PPM+MC states at the interface between points 2 and 3 of a 6-point window,
// as offsets: *o2 = (right edge of zone 2) - y2, *o3 = (left edge of zone 3) - y3
HD void ppm_offsets(const float* D, float* o2, float* o3) {
  const float K = (float)PPM_FACTOR;
  const float dq1 = mc_slope_diff(D[0], D[1]), dq2 = mc_slope_diff(D[1], D[2]);
  const float dq3 = mc_slope_diff(D[2], D[3]), dq4 = mc_slope_diff(D[3], D[4]);
  float l2n, r2n, l3n, r3n;
  ppm_limit_rel(-0.5f*D[1] - K*(dq2 - dq1), 0.5f*D[2] - K*(dq3 - dq2), &l2n, &r2n);
  ppm_limit_rel(-0.5f*D[2] - K*(dq3 - dq2), 0.5f*D[3] - K*(dq4 - dq3), &l3n, &r3n);
  *o2 = r2n;
  *o3 = l3n;
}

// linear_recon() MC limiter (src/recon.c:88-132) on the differences a = y2-y1, b = y3-y2
HD float mc_linear_diff(float a, float b) {
  const float Dqm = 2.f*a, Dqp = 2.f*b, Dqc = 0.5f*(a + b);
  if (Dqm*Dqp <= 0.f) return 0.f;
  const float am = fabsf(Dqm), ap = fabsf(Dqp), ac = fabsf(Dqc);
  if ((am <= ap) && (am <= ac)) return Dqm;
  if (ap <= ac) return Dqp;
  return Dqc;
}

// ---------------------------------------------------------------------------
// Face flux in mixed precision (the FP64 flux_face<d> in FP32)
template <int d>
HD void flux_recon_mixed(const Params& P, const Ptrs& X, const double* __restrict__ src,
                         int i, int j, int k, float* pl, float* pr) {
  const long long NT = P.nt;
  constexpr int di = (d == 0), dj = (d == 1), dk = (d == 2);
  if (P.degen[d]) {
    int ii = i, jj = j, kk = k;
    if (d == 0) { ii = NG; if (P.degen[1]) jj = NG; if (P.degen[2]) kk = NG; }
    if (d == 1) { jj = NG; if (P.degen[2]) kk = NG; }
    if (d == 2) { kk = NG; }
    const long long c = cidx(P, ii, jj, kk);
#pragma unroll
    for (int v = 0; v < NP; v++) pl[v] = pr[v] = (float)src[v*NT + c];
  } else {
    int ii = i, jj = j, kk = k;
    if (P.degen[0]) ii = NG;
    if (P.degen[1]) jj = NG;
    if (P.degen[2]) kk = NG;
    const long long st = cidx(P, di, dj, dk) - cidx(P, 0, 0, 0);
    constexpr bool rs = (d == 0);   // ones in the tables without RESCALE_R
#pragma unroll
    for (int v = 0; v < NP; v++) {
      double y[6];
      if (P.recon == RECON_PPM_MC) {
        const long long c0 = cidx(P, ii - 3*di, jj - 3*dj, kk - 3*dk);
#pragma unroll
        for (int m = 0; m < 6; m++) {
          double x = src[v*NT + c0 + m*st];
          if (rs) x *= X.rscale[rs_slot(v)*P.n1t + (ii - 3 + m)];
          y[m] = x;
        }
        float D[5];
#pragma unroll
        for (int m = 0; m < 5; m++) D[m] = (float)(y[m + 1] - y[m]);
        float o2, o3;
        ppm_offsets(D, &o2, &o3);
        double l = y[2] + (double)o2, r = y[3] + (double)o3;
        if (rs) { const double f = X.irscale[(3 + rs_slot(v))*P.n1t + i]; l *= f; r *= f; }
        pl[v] = (float)l;
        pr[v] = (float)r;
      } else {
        const long long c0 = cidx(P, ii - 2*di, jj - 2*dj, kk - 2*dk);
#pragma unroll
        for (int m = 0; m < 4; m++) {
          double x = src[v*NT + c0 + m*st];
          if (rs) x *= X.rscale[rs_slot(v)*P.n1t + (ii - 2 + m)];
          y[m] = x;
        }
        const float D0 = (float)(y[1] - y[0]), D1 = (float)(y[2] - y[1]), D2 = (float)(y[3] - y[2]);
        double l = y[1] + (double)(0.5f*mc_linear_diff(D0, D1));
        double r = y[2] - (double)(0.5f*mc_linear_diff(D1, D2));
        if (rs) { const double f = X.irscale[(3 + rs_slot(v))*P.n1t + i]; l *= f; r *= f; }
        pl[v] = (float)l;
        pr[v] = (float)r;
      }
    }
  }
}

// PPM+MC states of F consecutive faces along d from face (i,j,k), each zone's
// limited parabola computed once (ppm_offsets() split per zone).  Direction d
// must not be degenerate; zones past the array end are clamped (they feed
// only faces beyond a partial group).
template <int d, int F>
HD void ppm_faces_mixed(const Params& P, const Ptrs& X, const double* __restrict__ src,
                        int i, int j, int k, float (*pl)[NP], float (*pr)[NP]) {
  const long long NT = P.nt;
  constexpr int di = (d == 0), dj = (d == 1), dk = (d == 2);
  constexpr int M = F + 5;
  const float K = (float)PPM_FACTOR;
  int ii = i, jj = j, kk = k;
  if (d != 0 && P.degen[0]) ii = NG;
  if (d != 1 && P.degen[1]) jj = NG;
  if (d != 2 && P.degen[2]) kk = NG;
  const int pos = (d == 0) ? ii : ((d == 1) ? jj : kk);
  const int nd = (d == 0) ? P.n1t : ((d == 1) ? P.n2t : P.n3t);
  const long long st = cidx(P, di, dj, dk) - cidx(P, 0, 0, 0);
  const long long c0 = cidx(P, ii - 3*di, jj - 3*dj, kk - 3*dk);
  constexpr bool rs = (d == 0);   // ones in the tables without RESCALE_R
  int zm[M];
#pragma unroll
  for (int m = 0; m < M; m++) zm[m] = imin(m, nd - 1 - (pos - 3));
  // the next primitive's stencil is loaded before this one's limiters run,
  // this one's x1 rescale factors before that (waiting for them leaves the
  // stencil loads in flight)
  double raw[2][M];
#pragma unroll
  for (int m = 0; m < M; m++) raw[0][m] = src[c0 + zm[m]*st];
#pragma unroll
  for (int v = 0; v < NP; v++) {
    double fz[M], fo[F];
    if (rs) {
#pragma unroll
      for (int m = 0; m < M; m++) fz[m] = X.rscale[rs_slot(v)*P.n1t + (ii - 3 + zm[m])];
#pragma unroll
      for (int f = 0; f < F; f++) fo[f] = X.irscale[(3 + rs_slot(v))*P.n1t + imin(i + f, P.n1t - 1)];
    }
    if (v + 1 < NP) {
#pragma unroll
      for (int m = 0; m < M; m++) raw[(v + 1) & 1][m] = src[(v + 1)*NT + c0 + zm[m]*st];
    }
    double y[M];
#pragma unroll
    for (int m = 0; m < M; m++) y[m] = rs ? raw[v & 1][m]*fz[m] : raw[v & 1][m];
    float D[M - 1], dq[M - 1], ln[M - 2], rn[M - 2];
#pragma unroll
    for (int m = 0; m < M - 1; m++) D[m] = (float)(y[m + 1] - y[m]);
#pragma unroll
    for (int m = 1; m < M - 1; m++) dq[m] = mc_slope_diff(D[m - 1], D[m]);
#pragma unroll
    for (int m = 2; m < M - 2; m++)
      ppm_limit_rel(-0.5f*D[m - 1] - K*(dq[m] - dq[m - 1]), 0.5f*D[m] - K*(dq[m + 1] - dq[m]), &ln[m], &rn[m]);
#pragma unroll
    for (int f = 0; f < F; f++) {
      double l = y[2 + f] + (double)rn[2 + f], r = y[3 + f] + (double)ln[3 + f];
      if (rs) {
        l *= fo[f];
        r *= fo[f];
      }
      pl[f][v] = (float)l;
      pr[f][v] = (float)r;
    }
  }
}

// FP32 Riemann solve of face (i,j,k) from its two states; fluxes to X.Ff or
// to out[8]
template <int d>
HD void flux_riemann_mixed_g(const Params& P, const Ptrs& X, const float* pl, const float* pr,
                             int i, int j, int k, FailRec& fr, float* out, const GeomF* g) {
  const long long NT = P.nt;
  constexpr int dim = d + 1;
  const float gam = (float)P.gam;
  StateF ql, qr;
  if (get_state_f(pl, g, gam, &ql)) fr.nf[NFAIL_GAMMA_CALC]++;
  if (get_state_f(pr, g, gam, &qr)) fr.nf[NFAIL_GAMMA_CALC]++;
  float cmaxl, cminl, cmaxr, cminr;
  vchar_f<dim>(&ql, g, gam, &cmaxl, &cminl);
  vchar_f<dim>(&qr, g, gam, &cmaxr, &cminr);
  const float cmax = fabsf(fmaxf(fmaxf(0.f, cmaxl), cmaxr));
  const float cmin = fabsf(fmaxf(fmaxf(0.f, -cminl), -cminr));
  const float ctop = fmaxf(cmin, cmax);
  if (i >= NG && i < NG + P.n1 && j >= NG && j < NG + P.n2 && k >= NG && k < NG + P.n3 && ctop > 0.f)
    fr.vmax = dmax(fr.vmax, (double)ctop);

  float Ul[NP], Ur[NP], Fl[NP], Fr[NP];
  primtoflux_f<0>(pl, &ql, g, Ul);
  primtoflux_f<0>(pr, &qr, g, Ur);
  primtoflux_f<dim>(pl, &ql, g, Fl);
  primtoflux_f<dim>(pr, &qr, g, Fr);
  float Fv[NP];
  if (P.flux == FLUX_HLLE) {
    const float den = rcp_f(cmax + cmin + (float)SMALL);
#pragma unroll
    for (int l = 0; l < NP; l++) Fv[l] = (cmax*Fl[l] + cmin*Fr[l] - cmax*cmin*(Ur[l] - Ul[l]))*den;
  } else {
#pragma unroll
    for (int l = 0; l < NP; l++) Fv[l] = 0.5f*(Fl[l] + Fr[l] - ctop*(Ur[l] - Ul[l]));
  }
  if (out) {
#pragma unroll
    for (int l = 0; l < NP; l++) out[l] = Fv[l];
  } else {
    const long long c = cidx(P, i, j, k);
#pragma unroll
    for (int l = 0; l < NP; l++)
      if (l != B1 + d) X.Ff[(d*NFLUX + l)*NT + c] = Fv[l];   // the normal field's flux: never read
  }
}

template <int d>
HD void flux_riemann_mixed(const Params& P, const Ptrs& X, const float* pl, const float* pr,
                           int i, int j, int k, FailRec& fr, float* out = nullptr) {
  flux_riemann_mixed_g<d>(P, X, pl, pr, i, j, k, fr, out, X.geomf + geom_id(P, i, j, k, d));
}

// Entropy flux of face (i,j,k) for fixup_entropy_eq() (src/fixup.c:363-398)
// from the FP32 face solve's states and speeds, computed for the zones that
// take the entropy fix (the stage's face geometry is X.geomf throughout).
template <int d>
HD double entropy_face_mixed(const Params& P, const Ptrs& X, const double* __restrict__ src, int i, int j, int k,
                             const GeomF* gin) {
  float pl[NP], pr[NP];
  flux_recon_mixed<d>(P, X, src, i, j, k, pl, pr);
  const GeomF* g = gin ? gin : X.geomf + geom_id(P, i, j, k, d);
  constexpr int dim = d + 1;
  const float gam = (float)P.gam;
  StateF ql, qr;
  get_state_f(pl, g, gam, &ql);   // failures were counted by the face solve
  get_state_f(pr, g, gam, &qr);
  float cmaxl, cminl, cmaxr, cminr;
  vchar_f<dim>(&ql, g, gam, &cmaxl, &cminl);
  vchar_f<dim>(&qr, g, gam, &cmaxr, &cminr);
  const float cmax = fabsf(fmaxf(fmaxf(0.f, cmaxl), cmaxr));
  const float cmin = fabsf(fmaxf(fmaxf(0.f, -cminl), -cminr));
  const float ctop = fmaxf(cmin, cmax);
  const float gm1 = gam - 1.f;
  float s = gm1*pl[UU]*pow_f(pl[RHO], -gm1);
  const float Sl = s*(g->g*ql.ucon[0]), Gl = s*(g->g*ql.ucon[dim]);
  s = gm1*pr[UU]*pow_f(pr[RHO], -gm1);
  const float Sr = s*(g->g*qr.ucon[0]), Gr = s*(g->g*qr.ucon[dim]);
  float Fs;
  if (P.flux == FLUX_HLLE) Fs = (cmax*Gl + cmin*Gr - cmax*cmin*(Sr - Sl))*rcp_f(cmax + cmin + (float)SMALL);
  else Fs = 0.5f*(Gl + Gr - ctop*(Sr - Sl));
  return (double)Fs;
}

template <int d>
HD void flux_face_mixed(const Params& P, const Ptrs& X, const double* __restrict__ src,
                        int i, int j, int k, FailRec& fr) {
  float pl[NP], pr[NP];
  flux_recon_mixed<d>(P, X, src, i, j, k, pl, pr);
  flux_riemann_mixed<d>(P, X, pl, pr, i, j, k, fr);
}

// Corner EMFs from the FP32 face fluxes (flux_ct_para_fast2(), src/flux_ct.c:780-846;
// flux_ct(), :31-41)
HD float ppm_pair_f(const float* y) {
  float D[5];
#pragma unroll
  for (int m = 0; m < 5; m++) D[m] = y[m + 1] - y[m];
  float o2, o3;
  ppm_offsets(D, &o2, &o3);
  return (0.f + (y[2] + o2)) + (y[3] + o3);
}

HD void emf_edge_mixed(const Params& P, const Ptrs& X, int i, int j, int k) {
  const long long NT = P.nt;
  const long long c = cidx(P, i, j, k);
  const long long si = (long long)P.n2t*P.n3t, sj = P.n3t, sk = 1;
  const float* F = X.Ff;
  const float* F0B2 = F + (0*NFLUX + B2)*NT;
  const float* F0B3 = F + (0*NFLUX + B3)*NT;
  const float* F1B1 = F + (1*NFLUX + B1)*NT;
  const float* F1B3 = F + (1*NFLUX + B3)*NT;
  const float* F2B1 = F + (2*NFLUX + B1)*NT;
  const float* F2B2 = F + (2*NFLUX + B2)*NT;
  float e0, e1, e2;
  if (P.ct_para) {
    // the six stencils loaded before any limiter runs (loads in flight together)
    const float* line[6] = {F1B3, F2B2, F2B1, F0B3, F0B2, F1B1};
    const long long str[6] = {sk, sj, si, sk, sj, si};
    float y[6][6];
#pragma unroll
    for (int l = 0; l < 6; l++)
#pragma unroll
      for (int m = 0; m < 6; m++) y[l][m] = line[l][c + (m - 3)*str[l]];
    e0 = ppm_pair_f(y[0]) - ppm_pair_f(y[1]);
    e1 = ppm_pair_f(y[2]) - ppm_pair_f(y[3]);
    e2 = ppm_pair_f(y[4]) - ppm_pair_f(y[5]);
  } else {
    e0 = 0.25f*(F1B3[c] + F1B3[c - sk] - F2B2[c] - F2B2[c - sj]);
    e1 = 0.25f*(F2B1[c] + F2B1[c - si] - F0B3[c] - F0B3[c - sk]);
    e2 = 0.25f*(F0B2[c] + F0B2[c - sj] - F1B1[c] - F1B1[c - si]);
  }
  X.emff[0*NT + c] = e0;
  X.emff[1*NT + c] = e1;
  X.emff[2*NT + c] = e2;
}

// Face flux of variable l as the mixed update reads it, widened to FP64
// (B components from the corner EMFs, src/flux_ct.c:911-933; fix_flux(),
// src/bounds.c:2255-2270).
HD double face_flux_mixed(const Params& P, const Ptrs& X, int d, int l, int i, int j, int k) {
  const long long NT = P.nt;
  const long long c = cidx(P, i, j, k);
  if (l >= B1) {
    const double cf = P.ct_para ? 0.125 : 0.5;
    const long long si = (long long)P.n2t*P.n3t, sj = P.n3t, sk = 1;
    const float* E0 = X.emff;
    const float* E1 = X.emff + NT;
    const float* E2 = X.emff + 2*NT;
    if (d == 0) {
      if (l == B1) return 0.;
      if (l == B2) return cf*((double)E2[c] + (double)E2[c + sj]);
      return -cf*((double)E1[c] + (double)E1[c + sk]);
    } else if (d == 1) {
      if (l == B1) return -cf*((double)E2[c] + (double)E2[c + si]);
      if (l == B2) return 0.;
      return cf*((double)E0[c] + (double)E0[c + sk]);
    } else {
      if (l == B1) return cf*((double)E1[c] + (double)E1[c + si]);
      if (l == B2) return -cf*((double)E0[c] + (double)E0[c + sj]);
      return 0.;
    }
  }
  double f = (double)X.Ff[(d*NFLUX + l)*NT + c];
  if (l == RHO && d == 0 && P.bc == BC_SPHERICAL_OUTFLOW3) {
    if (i == NG && f > 0 && !(P.bc_nbr & 1)) f = 0.;
    if (i == NG + P.n1 && f < 0 && !(P.bc_nbr & 2)) f = 0.;
  }
  return f;
}

#ifdef __HIP_DEVICE_COMPILE__
#pragma clang fp contract(off)
#endif

// Neumaier-compensated accumulation
struct CompSum {
  double s = 0., c = 0.;
  HD void add(double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  HD double value() const { return s + c; }
};

}  // namespace hgpu
