// hgpu_c2p.h -- batched con-to-prim of independent zones and its learned
// initial guess (host/device bodies shared by the HIP kernels and the host build).
//
// * Counter-based sampling: Philox4x32-10 (Salmon, Moraes, Dror & Shaw,
//   "Parallel random numbers: as easy as 1, 2, 3", SC11) keyed by the seed and
//   counted by the sample index, so any device or host draws identical samples.
// * Samples: Kerr-Schild records at (r, th) (the metric of
//   src/metric.c:1168-1252, :523-539 in Kerr-Schild coordinates, lapse and shift
//   as calc_all_geom(), src/metric.c:4235-4250), primitives with log-uniform
//   Lorentz factor, magnetisation and u/rho, velocity and field directions
//   isotropic in the normal observer's orthonormal frame, and their conserved
//   variables (primtoflux(dir 0), src/phys.c:28-63).  rho = 1: the inversion is
//   scale-free in rho (U, W, b^2 all scale with it).
// * The solve is Utoprim_2d_fast() (src/utoprim_2d_fast.c:135-408) split into
//   resumable pieces (C2pCell) so a persistent kernel can park cells that are
//   still iterating and compact them into a queue.
// * The learned guess: an MLP 8 -> 64 -> 64 -> 2 (GELU) in FP32 maps
//   scale-free invariants of U (those the 2D Newton solves from) to log(W/D) and
//   log(gamma); the Newton solve starts from it and restarts from the plain guess
//   if that attempt fails.
#pragma once
#include "hgpu_phys.h"
#include "hgpu_geom.h"

namespace hgpu {

// ---------------------------------------------------------------------------
// Philox4x32-10
struct Philox4 { unsigned v[4]; };

HD void philox_round(unsigned* c, const unsigned* k) {
  const unsigned M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const unsigned long long p0 = (unsigned long long)M0 * c[0];
  const unsigned long long p1 = (unsigned long long)M1 * c[2];
  const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
  const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
  const unsigned n0 = hi1 ^ c[1] ^ k[0];
  const unsigned n1 = lo1;
  const unsigned n2 = hi0 ^ c[3] ^ k[1];
  const unsigned n3 = lo0;
  c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
}

HD Philox4 philox4x32_10(const unsigned* ctr, const unsigned* key) {
  const unsigned W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  unsigned c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  unsigned k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; r++) {
    philox_round(c, k);
    k[0] += W0;
    k[1] += W1;
  }
  Philox4 out;
  for (int l = 0; l < 4; l++) out.v[l] = c[l];
  return out;
}

// Uniform double in (0,1) from two 32-bit words (53 random bits, never 0 or 1).
HD double u01(unsigned hi, unsigned lo) {
  const unsigned long long m = ((unsigned long long)hi << 21) ^ (unsigned long long)(lo >> 11);
  return ((double)(m & ((1ull << 53) - 1ull)) + 0.5) * (1.0 / 9007199254740992.0);
}

// `count` uniforms of draw n in stream `stream`: Philox blocks with counter
// (n_lo, n_hi, stream, block) under key (seed_lo, seed_hi), two uniforms per block.
HD void draw_uniforms(unsigned long long seed, unsigned stream, unsigned long long n, double* u, int count) {
  const unsigned key[2] = {(unsigned)seed, (unsigned)(seed >> 32)};
  int m = 0;
  for (unsigned blk = 0; m < count; blk++) {
    const unsigned ctr[4] = {(unsigned)n, (unsigned)(n >> 32), stream, blk};
    const Philox4 r = philox4x32_10(ctr, key);
    u[m++] = u01(r.v[0], r.v[1]);
    if (m < count) u[m++] = u01(r.v[2], r.v[3]);
  }
}

// ---------------------------------------------------------------------------
// Kerr-Schild geometry record at (r, th) in Kerr-Schild coordinates (M = 1).
HD void ks_record_rth(double a, double r, double th, Geom* G) {
  double gdet;
  ks_metric_rth(a, 1., r, th, G->gcov, G->gcon, &gdet);
  G->g = gdet;
  geom_finish(G);
}

// Sampling ranges of the batched inversion workload.
constexpr double C2P_R_MIN = 1.3, C2P_R_MAX = 100., C2P_TH_PAD = 0.05;
constexpr double C2P_GAMMA_MAX = 50.;
constexpr double C2P_SIG_MIN = 1e-4, C2P_SIG_MAX = 1e2;   // b^2 / rho
constexpr double C2P_UOR_MIN = 1e-4, C2P_UOR_MAX = 1e1;   // u / rho
constexpr double C2P_GUESS_JITTER = 0.1;                   // perturbed-truth guess: x (1 +- 10 %)
constexpr int C2P_NUNIF = 14;

// Coordinate components x of the vector whose components in the orthonormal
// frame of gamma_ij = gcov[1..3][1..3] are e: x = L^{-T} e with gamma = L L^T.
HD void frame_vector(const Geom* G, const double* e, double* x) {
  double L[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
  for (int p = 0; p < 3; p++) {
    for (int q = 0; q <= p; q++) {
      double sum = G->gcov[p + 1][q + 1];
      for (int m = 0; m < q; m++) sum -= L[p][m]*L[q][m];
      L[p][q] = (p == q) ? sqrt(sum) : sum / L[q][q];
    }
  }
  for (int p = 2; p >= 0; p--) {
    double sum = e[p];
    for (int m = p + 1; m < 3; m++) sum -= L[m][p]*x[m];
    x[p] = sum / L[p][p];
  }
}

// Sample n: position rth[2], primitives prim[8] (rho = 1), conserved U[8] and,
// when guess is given, the perturbed-truth guess (rho, u, u~ each times
// 1 +- 10 %, B copied).
HD void c2p_sample(unsigned long long seed, unsigned long long n, double a, double gam, double* U, double* rth,
                   double* prim, double* guess) {
  double u[C2P_NUNIF];
  draw_uniforms(seed, 0u, n, u, C2P_NUNIF);
  const double r = exp(log(C2P_R_MIN) + u[0]*(log(C2P_R_MAX) - log(C2P_R_MIN)));
  const double th = C2P_TH_PAD + u[1]*(M_PI - 2.*C2P_TH_PAD);
  const double gamma = exp(u[2]*log(C2P_GAMMA_MAX));
  const double sig = exp(log(C2P_SIG_MIN) + u[3]*(log(C2P_SIG_MAX) - log(C2P_SIG_MIN)));
  const double uor = exp(log(C2P_UOR_MIN) + u[4]*(log(C2P_UOR_MAX) - log(C2P_UOR_MIN)));
  const double cv = 2.*u[5] - 1., pv = 2.*M_PI*u[6];
  const double cb = 2.*u[7] - 1., pb = 2.*M_PI*u[8];
  Geom G;
  ks_record_rth(a, r, th, &G);
  const double sv = sqrt(fmax(0., 1. - cv*cv)), sb = sqrt(fmax(0., 1. - cb*cb));
  const double ev[3] = {sv*cos(pv), sv*sin(pv), cv};
  const double eb[3] = {sb*cos(pb), sb*sin(pb), cb};
  double xv[3], xb[3];
  frame_vector(&G, ev, xv);
  frame_vector(&G, eb, xb);
  const double ut = sqrt((gamma - 1.)*(gamma + 1.));
  prim[RHO] = 1.;
  prim[UU] = uor;
  for (int l = 0; l < 3; l++) prim[U1 + l] = ut*xv[l];
  for (int l = 0; l < 3; l++) prim[B1 + l] = xb[l];
  // b^2 is quadratic in B: scale B to b^2 = sig * rho
  State q;
  get_state(prim, &G, gam, &q);
  const double scale = sqrt(sig*prim[RHO]/q.bsq);
  for (int l = 0; l < 3; l++) prim[B1 + l] *= scale;
  get_state(prim, &G, gam, &q);
  primtoflux(prim, &q, 0, &G, U);
  rth[0] = r;
  rth[1] = th;
  if (guess) {
    for (int l = 0; l < 5; l++) guess[l] = prim[l]*(1. + C2P_GUESS_JITTER*(2.*u[9 + l] - 1.));
    for (int l = B1; l < NP; l++) guess[l] = prim[l];
  }
}

// ---------------------------------------------------------------------------
// MLP of the learned guess: 8 -> 64 -> 64 -> 2, GELU, FP32.
constexpr int MLP_IN = 8, MLP_H = 64, MLP_OUT = 2;
// Parameter blob (floats): input mean[8] and inverse scale[8], W1[8][64],
// b1[64], W2[64][64], b2[64], W3[64][2], b3[2], output mean[2] and scale[2].
// The network sees x = (f - mean) * inv_scale and outputs o; y = o*scale + mean.
constexpr int MLP_OFF_MU = 0, MLP_OFF_ISD = 8, MLP_OFF_W1 = 16, MLP_OFF_B1 = MLP_OFF_W1 + MLP_IN*MLP_H,
              MLP_OFF_W2 = MLP_OFF_B1 + MLP_H, MLP_OFF_B2 = MLP_OFF_W2 + MLP_H*MLP_H,
              MLP_OFF_W3 = MLP_OFF_B2 + MLP_H, MLP_OFF_B3 = MLP_OFF_W3 + MLP_H*MLP_OUT,
              MLP_OFF_YMU = MLP_OFF_B3 + MLP_OUT, MLP_OFF_YSD = MLP_OFF_YMU + MLP_OUT,
              MLP_NPARAM = MLP_OFF_YSD + MLP_OUT;
// trainable part: W1 .. b3, contiguous
constexpr int MLP_TRAIN_OFF = MLP_OFF_W1, MLP_NTRAIN = MLP_OFF_YMU - MLP_OFF_W1;

// GELU in its tanh form (Hendrycks & Gimpel 2016), written with the logistic
// function: gelu(x) = x * sigma(k (x + c x^3)), k = 2 sqrt(2/pi), c = 0.044715.
constexpr float GELU_K = 1.5957691216057308f, GELU_C = 0.044715f;
HD float gelu_sigma(float x) { return 1.f/(1.f + expf(-GELU_K*(x + GELU_C*x*x*x))); }
HD float gelu(float x) { return x*gelu_sigma(x); }
HD float gelu_grad(float x) {
  const float s = gelu_sigma(x);
  return s + x*s*(1.f - s)*GELU_K*(1.f + 3.f*GELU_C*x*x);
}

// Features: scale-free invariants of the inversion (U2d quantities over D).
HD void mlp_features(const U2d& s, float* f) {
  const double D = s.D, E = -s.Qdotn;   // energy density incl. rest mass, normal frame
  const double tiny = 1e-30;
  f[0] = (float)log(s.Bsq/D + tiny);
  f[1] = (float)log(s.Qtsq/(D*D) + tiny);
  f[2] = (float)log(s.QdotBsq/(D*D*D) + tiny);
  f[3] = (float)log(E/D);
  f[4] = (float)log(fabs(E - D - s.half_Bsq)/D + 1e-12);
  f[5] = (float)(sqrt(s.Qtsq)/E);
  f[6] = (float)(s.Bsq/E);
  f[7] = (float)(s.QdotBsq/(s.Bsq*s.Qtsq + tiny));
}

// Targets of a solved zone: log(W/D) = log(h gamma), log(gamma).
HD void mlp_targets(const double* prim, double gamma, double gam, float* y) {
  const double h = (prim[RHO] + gam*prim[UU])/prim[RHO];
  y[0] = (float)(log(h) + log(gamma));
  y[1] = (float)log(gamma);
}

// Forward pass of one sample.  w is read with uniform indices only, so on the
// device every weight is a scalar (SGPR) operand shared by the wave.
HD void mlp_forward(const float* __restrict__ w, const float* f, float* y) {
  float x[MLP_IN], h1[MLP_H];
  for (int i = 0; i < MLP_IN; i++) x[i] = (f[i] - w[MLP_OFF_MU + i]) * w[MLP_OFF_ISD + i];
  for (int j = 0; j < MLP_H; j++) {
    float z = w[MLP_OFF_B1 + j];
    for (int i = 0; i < MLP_IN; i++) z = fmaf(x[i], w[MLP_OFF_W1 + i*MLP_H + j], z);
    h1[j] = gelu(z);
  }
  float o0 = w[MLP_OFF_B3 + 0], o1 = w[MLP_OFF_B3 + 1];
  for (int j = 0; j < MLP_H; j++) {
    float z = w[MLP_OFF_B2 + j];
    for (int i = 0; i < MLP_H; i++) z = fmaf(h1[i], w[MLP_OFF_W2 + i*MLP_H + j], z);
    const float h2 = gelu(z);
    o0 = fmaf(h2, w[MLP_OFF_W3 + j*MLP_OUT + 0], o0);
    o1 = fmaf(h2, w[MLP_OFF_W3 + j*MLP_OUT + 1], o1);
  }
  y[0] = o0 * w[MLP_OFF_YSD + 0] + w[MLP_OFF_YMU + 0];
  y[1] = o1 * w[MLP_OFF_YSD + 1] + w[MLP_OFF_YMU + 1];
}

// Backward pass of one sample for the loss 0.5 |o - t_n|^2 (o: raw network
// output, t_n = (t - mean)/scale); the gradient of the trainable part is added
// to g (blob layout from MLP_TRAIN_OFF).  Returns the sample's loss.
HD float mlp_backward_add(const float* __restrict__ w, const float* f, const float* t, float* g) {
  float x[MLP_IN], z1[MLP_H], h1[MLP_H], z2[MLP_H], h2[MLP_H];
  for (int i = 0; i < MLP_IN; i++) x[i] = (f[i] - w[MLP_OFF_MU + i]) * w[MLP_OFF_ISD + i];
  for (int j = 0; j < MLP_H; j++) {
    float z = w[MLP_OFF_B1 + j];
    for (int i = 0; i < MLP_IN; i++) z = fmaf(x[i], w[MLP_OFF_W1 + i*MLP_H + j], z);
    z1[j] = z;
    h1[j] = gelu(z);
  }
  float o[MLP_OUT] = {w[MLP_OFF_B3 + 0], w[MLP_OFF_B3 + 1]};
  for (int j = 0; j < MLP_H; j++) {
    float z = w[MLP_OFF_B2 + j];
    for (int i = 0; i < MLP_H; i++) z = fmaf(h1[i], w[MLP_OFF_W2 + i*MLP_H + j], z);
    z2[j] = z;
    h2[j] = gelu(z);
    for (int k = 0; k < MLP_OUT; k++) o[k] = fmaf(h2[j], w[MLP_OFF_W3 + j*MLP_OUT + k], o[k]);
  }
  float dy[MLP_OUT], loss = 0.f;
  for (int k = 0; k < MLP_OUT; k++) {
    const float tn = (t[k] - w[MLP_OFF_YMU + k]) / w[MLP_OFF_YSD + k];
    dy[k] = o[k] - tn;
    loss += 0.5f*dy[k]*dy[k];
  }
  float* gW1 = g + (MLP_OFF_W1 - MLP_TRAIN_OFF);
  float* gb1 = g + (MLP_OFF_B1 - MLP_TRAIN_OFF);
  float* gW2 = g + (MLP_OFF_W2 - MLP_TRAIN_OFF);
  float* gb2 = g + (MLP_OFF_B2 - MLP_TRAIN_OFF);
  float* gW3 = g + (MLP_OFF_W3 - MLP_TRAIN_OFF);
  float* gb3 = g + (MLP_OFF_B3 - MLP_TRAIN_OFF);
  for (int k = 0; k < MLP_OUT; k++) gb3[k] += dy[k];
  float dz2[MLP_H];
  for (int j = 0; j < MLP_H; j++) {
    float dh = 0.f;
    for (int k = 0; k < MLP_OUT; k++) {
      gW3[j*MLP_OUT + k] += h2[j]*dy[k];
      dh = fmaf(w[MLP_OFF_W3 + j*MLP_OUT + k], dy[k], dh);
    }
    dz2[j] = dh*gelu_grad(z2[j]);
    gb2[j] += dz2[j];
  }
  for (int i = 0; i < MLP_H; i++) {
    float dh = 0.f;
    for (int j = 0; j < MLP_H; j++) {
      gW2[i*MLP_H + j] += h1[i]*dz2[j];
      dh = fmaf(w[MLP_OFF_W2 + i*MLP_H + j], dz2[j], dh);
    }
    const float dz1 = dh*gelu_grad(z1[i]);
    gb1[i] += dz1;
    for (int k = 0; k < MLP_IN; k++) gW1[k*MLP_H + i] += x[k]*dz1;
  }
  return loss;
}

// ---------------------------------------------------------------------------
// Resumable per-zone solve.
//
// Attempt 0 starts from the learned guess when the network is given, else from
// the plain guess; a failed learned attempt restarts once from the plain guess.
// The plain guess is the incoming primitives (the reference's own guess) or the
// cold guess rho = D, u = 1e-3 D, u~ = 0.
struct C2pCell {
  U2d s;
  U2dAux aux;
  NewtonIter it;
  double x_plain[2];
  int attempt;       // 0 first guess, 1 fallback from the plain guess
  int has_fallback;  // the plain guess is usable and attempt 0 used the network
  int total;         // Newton iterations over both attempts
};

constexpr int C2P_RUNNING = 0x7fffffff;

HD void cold_guess(const U2d& s, double* prim) {
  prim[RHO] = s.D;
  prim[UU] = 1e-3*s.D;
  prim[U1] = prim[U2] = prim[U3] = 0.;
}

HD void c2p_start_learned(C2pCell& c, int rc_plain, const float* y);

// Geometry-dependent set-up of one zone.  prim: incoming primitives (read when
// plain_guess == 0), written with B from U and the plain guess's rho, u, u~ --
// what the zone holds if no attempt succeeds.  Returns an early return code of
// Utoprim_2d_fast() (-100, 2) or C2P_RUNNING.
HD int c2p_begin(const double* U, const Geom* geom, double gam, double gmog, const float* mlp, int plain_guess,
                 double* prim, C2pCell& c) {
  c.total = 0;
  c.attempt = 0;
  c.has_fallback = 0;
  int rc = u2d_prepare(U, geom, gmog, prim, c.s, c.aux);
  if (rc) return rc;
  if (plain_guess == 1) cold_guess(c.s, prim);
  const int rc_plain = u2d_guess(c.s, geom, gam, prim, c.x_plain);
  if (mlp) {
    float f[MLP_IN], y[MLP_OUT];
    mlp_features(c.s, f);
    mlp_forward(mlp, f, y);
    c2p_start_learned(c, rc_plain, y);
  } else {
    if (rc_plain) return rc_plain;
    u2d_newton_init(c.it, c.x_plain);
  }
  return C2P_RUNNING;
}

// The two halves of c2p_begin() with the learned guess for callers that
// evaluate the network for many zones at once (the device's wavefront-wide
// matrix-core forward pass): c2p_prepare_learned() sets the zone up and
// writes its features (returns an early code of Utoprim_2d_fast() or
// C2P_RUNNING, with *rc_plain the plain guess's code; f may be null when the
// outputs are already known); c2p_start_learned()
// starts the Newton iteration from the network's outputs y.
HD int c2p_prepare_learned(const double* U, const Geom* geom, double gam, double gmog, int plain_guess,
                           double* prim, C2pCell& c, int* rc_plain, float* f) {
  c.total = 0;
  c.attempt = 0;
  c.has_fallback = 0;
  int rc = u2d_prepare(U, geom, gmog, prim, c.s, c.aux);
  if (rc) return rc;
  if (plain_guess == 1) cold_guess(c.s, prim);
  *rc_plain = u2d_guess(c.s, geom, gam, prim, c.x_plain);
  if (f) mlp_features(c.s, f);
  return C2P_RUNNING;
}

HD void c2p_start_learned(C2pCell& c, int rc_plain, const float* y) {
  double x0[2];
  const double lg = fmax((double)y[1], 0.);
  const double vsq = -expm1(-2.*lg);     // 1 - 1/gamma^2
  // a learned W close to the bracket's lower bound is nudged up by 2 % steps
  // rather than tenfold, so a guess within a few per cent stays that close
  u2d_guess_w(c.s, c.s.D*exp((double)y[0]), x0, vsq, 1.02, 200);
  c.has_fallback = (rc_plain == 0);
  u2d_newton_init(c.it, x0);
}

// Up to `budget` Newton iterations; on convergence or failure the zone is
// finished (prim and gamma written on success) or restarted from the plain
// guess.  Returns the final code or C2P_RUNNING.
HD int c2p_advance(C2pCell& c, double gam, int budget, double* prim, double* gamma) {
  for (int b = 0; b < budget; b++) {
    const bool keep = u2d_newton_step(c.s, c.it);
    c.total++;
    if (keep) continue;
    const int retval = u2d_newton_result(c.it);
    double bsq = 0.;
    const int ret = u2d_finish(c.s, c.aux, gam, c.it.x, retval, prim, gamma, &bsq);
    if (ret == 0 || !c.has_fallback || c.attempt == 1) return ret;
    c.attempt = 1;
    u2d_newton_init(c.it, c.x_plain);
  }
  return C2P_RUNNING;
}

// Parked-zone record of the compaction queue: 19 doubles + 2 ints, stored
// structure-of-arrays (field f of slot q at d[f*cap + q]).
constexpr int C2P_QD = 19;
struct C2pQueue {
  double* d;        // [C2P_QD][cap]
  int* cell;        // [cap]
  int* meta;        // [cap]
  long long cap;
};

HD void c2p_park(const C2pQueue& Q, long long q, int cell, const C2pCell& c) {
  const long long K = Q.cap;
  double* d = Q.d;
  d[0*K + q] = c.s.D; d[1*K + q] = c.s.Bsq; d[2*K + q] = c.s.QdotBsq; d[3*K + q] = c.s.Qtsq;
  d[4*K + q] = c.s.Qdotn; d[5*K + q] = c.aux.QdotB;
  for (int l = 0; l < 3; l++) {
    d[(6 + l)*K + q] = c.aux.Qcon[1 + l];
    d[(9 + l)*K + q] = c.aux.Bcon[1 + l];
    d[(12 + l)*K + q] = c.aux.ncon[1 + l];
  }
  d[15*K + q] = c.it.x[0]; d[16*K + q] = c.it.x[1];
  d[17*K + q] = c.x_plain[0]; d[18*K + q] = c.x_plain[1];
  Q.cell[q] = cell;
  Q.meta[q] = c.it.n_iter | (c.it.i_extra << 6) | (c.it.doing_extra << 9) | (c.attempt << 10) |
              (c.has_fallback << 11) | (c.total << 12);
}

HD int c2p_unpark(const C2pQueue& Q, long long q, double gmog, C2pCell& c) {
  const long long K = Q.cap;
  const double* d = Q.d;
  c.s.D = d[0*K + q]; c.s.Bsq = d[1*K + q]; c.s.QdotBsq = d[2*K + q]; c.s.Qtsq = d[3*K + q];
  c.s.Qdotn = d[4*K + q]; c.aux.QdotB = d[5*K + q];
  c.s.half_Bsq = 0.5*c.s.Bsq;
  c.s.gmog = gmog;
  c.aux.Qcon[0] = c.aux.Bcon[0] = c.aux.ncon[0] = 0.;   // not read after the solve
  for (int l = 0; l < 3; l++) {
    c.aux.Qcon[1 + l] = d[(6 + l)*K + q];
    c.aux.Bcon[1 + l] = d[(9 + l)*K + q];
    c.aux.ncon[1 + l] = d[(12 + l)*K + q];
  }
  c.it.x[0] = d[15*K + q]; c.it.x[1] = d[16*K + q];
  c.x_plain[0] = d[17*K + q]; c.x_plain[1] = d[18*K + q];
  const int m = Q.meta[q];
  c.it.n_iter = m & 63;
  c.it.i_extra = (m >> 6) & 7;
  c.it.doing_extra = (m >> 9) & 1;
  c.it.n_extra = EXTRA_NEWT_ITER;
  c.attempt = (m >> 10) & 1;
  c.has_fallback = (m >> 11) & 1;
  c.total = m >> 12;
  c.it.f = c.it.df = c.it.errx = 1.;   // recomputed by the next iteration before use
  return Q.cell[q];
}

// Newton iterations one zone can take: general_newton_raphson() stops at
// n_iter = MAX_NEWT_ITER (the extra iterations fall inside that bound), twice
// with the fallback.
constexpr int C2P_MAX_ITERS = 2*MAX_NEWT_ITER;

struct C2pArgs {
  long long n;
  double a, gam, gmog;
  const double* U;     // [8][n]
  const double* rth;   // [2][n]
  int plain_guess;
  double* prim;        // [8][n]
  int* code;
  int* iters;
};

// First pass over zone t: set-up, the learned guess (mlp != nullptr) and up
// to `budget` iterations.  prim[.][t] gets B and the plain guess (or the
// solution).  Returns true when the zone is still iterating; c then holds it.
HD bool c2p_first_zone(const C2pArgs& A, const float* mlp, long long t, int budget, C2pCell& c) {
  const long long n = A.n;
  double u[NP], pr[NP];
  for (int l = 0; l < NP; l++) u[l] = A.U[l*n + t];
  for (int l = 0; l < NP; l++) pr[l] = (A.plain_guess == 0) ? A.prim[l*n + t] : 0.;
  Geom G;
  ks_record_rth(A.a, A.rth[t], A.rth[n + t], &G);
  double gm = 1.;
  int rc = c2p_begin(u, &G, A.gam, A.gmog, mlp, A.plain_guess, pr, c);
  if (rc == C2P_RUNNING) rc = c2p_advance(c, A.gam, budget, pr, &gm);
  for (int l = 0; l < NP; l++) A.prim[l*n + t] = pr[l];
  if (rc == C2P_RUNNING) return true;
  A.code[t] = rc;
  A.iters[t] = c.total;
  return false;
}

// The rest of c2p_first_zone() once the network's outputs are known (the
// zone's set-up returned rc, C2P_RUNNING to iterate): iterations, output.
HD bool c2p_first_zone_finish(const C2pArgs& A, long long t, int budget, int rc, int rc_plain, const float* y,
                              double* pr, C2pCell& c) {
  const long long n = A.n;
  double gm = 1.;
  if (rc == C2P_RUNNING) {
    c2p_start_learned(c, rc_plain, y);
    rc = c2p_advance(c, A.gam, budget, pr, &gm);
  }
  for (int l = 0; l < NP; l++) A.prim[l*n + t] = pr[l];
  if (rc == C2P_RUNNING) return true;
  A.code[t] = rc;
  A.iters[t] = c.total;
  return false;
}

// Later pass over queue slot q of Q: up to `budget` more iterations.  Returns
// true when the zone is still iterating (c holds it, *cell its index).
HD bool c2p_next_zone(const C2pArgs& A, const C2pQueue& Q, long long q, int budget, C2pCell& c, int* cell) {
  const long long n = A.n;
  const int t = c2p_unpark(Q, q, A.gmog, c);
  *cell = t;
  double pr[NP], gm = 1.;
  const int rc = c2p_advance(c, A.gam, budget, pr, &gm);
  if (rc == C2P_RUNNING) return true;
  if (rc == 0)
    for (int l = 0; l < B1; l++) A.prim[l*n + t] = pr[l];
  A.code[t] = rc;
  A.iters[t] = c.total;
  return false;
}

}  // namespace hgpu
