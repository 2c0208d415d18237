// hgpu_c2p_host.cpp -- host build of the batched con-to-prim and network entry
// points (same bodies as hgpu_c2p.hip, executed zone by zone on the CPU with
// host pointers; the stream argument is ignored).  Linked only into the host
// check library, which the CPU tests use to exercise the kernels' arithmetic
// and the compaction queue without a GPU.
#include "hgpu_c2p_api.h"
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace hgpu;

namespace {

void grad_host(long long n, const float* w, const float* feat, const float* target, long long batch, int random,
               unsigned long long seed, int step, float* grad /* [MLP_PART] */) {
  // per-workgroup partials summed in the GPU's order (workgroups of MLP_GT)
  const long long nblk = (batch + MLP_GT - 1) / MLP_GT;
  std::vector<float> acc(MLP_PART, 0.f), part(MLP_PART);
  for (long long b = 0; b < nblk; b++) {
    std::fill(part.begin(), part.end(), 0.f);
    float loss = 0.f;
    for (int lane = 0; lane < MLP_GT; lane++) {
      const long long s = b * MLP_GT + lane;
      if (s >= batch) break;
      const long long k = random ? mlp_batch_index(seed, step, s, batch, n) : s;
      float f[MLP_IN], t[MLP_OUT];
      for (int i = 0; i < MLP_IN; i++) f[i] = feat[i*n + k];
      for (int q = 0; q < MLP_OUT; q++) t[q] = target[q*n + k];
      loss += mlp_backward_add(w, f, t, part.data());
    }
    part[MLP_NTRAIN] = loss;
    for (int p = 0; p < MLP_PART; p++) acc[p] += part[p];
  }
  for (int p = 0; p < MLP_PART; p++) grad[p] = acc[p] / (float)batch;
}

}  // namespace

extern "C" {

// device plumbing of the host build: "device" memory is host memory
int hgpu_dev_count(int* out) {
  if (!out) return set_error(-1, "null argument");
  *out = 0;
  return 0;
}
int hgpu_dev_alloc(long long bytes, void** out) {
  if (!out || bytes < 0) return set_error(-1, "bad argument");
  *out = std::malloc((size_t)(bytes > 0 ? bytes : 8));
  return *out ? 0 : set_error(-4, "out of host memory");
}
int hgpu_dev_free(void* p) { std::free(p); return 0; }
int hgpu_dev_copy(void* dst, const void* src, long long bytes, int kind, void*) {
  if (bytes < 0 || (bytes > 0 && (!dst || !src))) return set_error(-1, "bad argument");
  if (kind < 0 || kind > 2) return set_error(-2, "kind must be 0 (to device), 1 (to host) or 2 (device to device)");
  if (bytes) std::memmove(dst, src, (size_t)bytes);
  return 0;
}
int hgpu_dev_copy_kernel(void* dst, const void* src, long long bytes, void*) {
  if (bytes < 0 || bytes % 16 || (bytes > 0 && (!dst || !src))) return set_error(-1, "bytes must be a multiple of 16");
  if (bytes) std::memmove(dst, src, (size_t)bytes);
  return 0;
}
int hgpu_stream_sync(void*) { return 0; }
int hgpu_event_create(void** ev) {
  if (!ev) return set_error(-1, "null argument");
  *ev = new double(0.);
  return 0;
}
int hgpu_event_destroy(void* ev) { delete static_cast<double*>(ev); return 0; }
int hgpu_event_record(void* ev, void*) {
  *static_cast<double*>(ev) = std::chrono::duration<double, std::milli>(
      std::chrono::steady_clock::now().time_since_epoch()).count();
  return 0;
}
int hgpu_event_elapsed(void* ev0, void* ev1, float* ms) {
  if (!ms) return set_error(-1, "null argument");
  *ms = (float)(*static_cast<double*>(ev1) - *static_cast<double*>(ev0));
  return 0;
}

int hgpu_c2p_samples(void*, long long n, unsigned long long seed, double a, double gam, double* U, double* rth,
                     double* prim, double* guess) {
  if (int rc = c2p_check_common(n, a, gam)) return rc;
  if (n == 0) return 0;
  if (!U || !rth || !prim) return set_error(-1, "null array argument");
  for (long long t = 0; t < n; t++) {
    double u[NP], r[2], p[NP], g[NP];
    c2p_sample(seed, (unsigned long long)t, a, gam, u, r, p, guess ? g : nullptr);
    for (int l = 0; l < NP; l++) { U[l*n + t] = u[l]; prim[l*n + t] = p[l]; }
    rth[t] = r[0];
    rth[n + t] = r[1];
    if (guess)
      for (int l = 0; l < NP; l++) guess[l*n + t] = g[l];
  }
  return 0;
}

long long hgpu_c2p_workspace_bytes(long long n, int compact) { return c2p_work_bytes(n, compact); }

int hgpu_c2p_solve(void*, long long n, double a, double gam, const double* U, const double* rth, const float* mlp,
                   int plain_guess, int compact, int budget_first, int budget_next, void* work, long long work_bytes,
                   double* prim, int* code, int* iters) {
  if (int rc = c2p_check_solve(n, a, gam, U, rth, plain_guess, compact, budget_first, budget_next, work,
                               work_bytes, prim, code, iters))
    return rc;
  if (n == 0) return 0;
  C2pArgs A{n, a, gam, (gam - 1.)/gam, U, rth, plain_guess, prim, code, iters};
  if (!compact) {
    for (long long t = 0; t < n; t++) {
      C2pCell c;
      if (c2p_first_zone(A, mlp, t, C2P_MAX_ITERS, c)) return set_error(-6, "zone still iterating after the full budget");
    }
    return 0;
  }
  C2pWork W = c2p_work_layout(work, n);
  const int passes = c2p_npasses(budget_first, budget_next);
  for (int p = 0; p < 64; p++) W.count[p] = 0u;
  for (long long t = 0; t < n; t++) {
    C2pCell c;
    if (c2p_first_zone(A, mlp, t, budget_first, c)) c2p_park(W.q[0], W.count[0]++, (int)t, c);
  }
  for (int p = 1; p <= passes; p++) {
    const int budget = (p == passes) ? C2P_MAX_ITERS : budget_next;
    const C2pQueue& Qin = W.q[(p - 1) & 1];
    const C2pQueue& Qout = W.q[p & 1];
    for (long long q = 0; q < (long long)W.count[p - 1]; q++) {
      C2pCell c;
      int cell = 0;
      if (c2p_next_zone(A, Qin, q, budget, c, &cell)) c2p_park(Qout, W.count[p]++, cell, c);
    }
  }
  if (W.count[passes]) return set_error(-6, "zones still iterating after the last pass");
  return 0;
}

int hgpu_mlp_sizes(int* out) {
  if (!out) return set_error(-1, "null argument");
  out[0] = MLP_NPARAM; out[1] = MLP_NTRAIN; out[2] = MLP_TRAIN_OFF;
  out[3] = MLP_IN; out[4] = MLP_H; out[5] = MLP_OUT;
  return 0;
}

int hgpu_mlp_features(void*, long long n, double a, double gam, const double* U, const double* rth,
                      const double* prim, float* feat, float* target) {
  if (int rc = c2p_check_common(n, a, gam)) return rc;
  if (n == 0) return 0;
  if (!U || !rth || !feat) return set_error(-1, "null array argument");
  if (prim && !target) return set_error(-1, "targets requested without an output array");
  for (long long t = 0; t < n; t++) {
    double u[NP], pr[NP] = {0., 0., 0., 0., 0., 0., 0., 0.};
    for (int l = 0; l < NP; l++) u[l] = U[l*n + t];
    Geom G;
    ks_record_rth(a, rth[t], rth[n + t], &G);
    U2d s;
    U2dAux aux;
    float f[MLP_IN];
    if (u2d_prepare(u, &G, (gam - 1.)/gam, pr, s, aux) == 0) mlp_features(s, f);
    else for (int i = 0; i < MLP_IN; i++) f[i] = 0.f;
    for (int i = 0; i < MLP_IN; i++) feat[i*n + t] = f[i];
    if (prim) {
      for (int l = 0; l < NP; l++) pr[l] = prim[l*n + t];
      double gamma = 1.;
      gamma_calc(pr, &G, &gamma);
      float y[MLP_OUT];
      mlp_targets(pr, gamma, gam, y);
      for (int k = 0; k < MLP_OUT; k++) target[k*n + t] = y[k];
    }
  }
  return 0;
}

int hgpu_mlp_forward(void*, long long n, const float* w, const float* feat, float* y) {
  if (n < 0 || n > INT_MAX) return set_error(-2, "n must be in [0, 2^31)");
  if (n == 0) return 0;
  if (!w || !feat || !y) return set_error(-1, "null array argument");
  for (long long t = 0; t < n; t++) {
    float f[MLP_IN], o[MLP_OUT];
    for (int i = 0; i < MLP_IN; i++) f[i] = feat[i*n + t];
    mlp_forward(w, f, o);
    for (int k = 0; k < MLP_OUT; k++) y[k*n + t] = o[k];
  }
  return 0;
}

long long hgpu_mlp_workspace_bytes(long long batch) { return mlp_work_bytes(batch); }

int hgpu_mlp_grad(void*, long long n, const float* w, const float* feat, const float* target, void* work,
                  long long work_bytes, float* grad, float* loss) {
  if (n < 1 || n > INT_MAX) return set_error(-2, "n must be in [1, 2^31)");
  if (!w || !feat || !target || !work || !grad || !loss) return set_error(-1, "null argument");
  if (work_bytes < mlp_work_bytes(n)) return set_error(-3, "workspace smaller than hgpu_mlp_workspace_bytes()");
  MlpWork W = mlp_work_layout(work, n);
  grad_host(n, w, feat, target, n, 0, 0ull, 0, W.grad);
  for (int p = 0; p < MLP_NTRAIN; p++) grad[p] = W.grad[p];
  *loss = W.grad[MLP_NTRAIN];
  return 0;
}

int hgpu_mlp_init(void*, unsigned long long seed, long long n, const float* feat, const float* target, float* w) {
  if (n < 1 || n > INT_MAX) return set_error(-2, "n must be in [1, 2^31)");
  if (!feat || !target || !w) return set_error(-1, "null argument");
  // the GPU's reduction order: 256 strided partial sums, then summed in order
  const int CBH = 256, M = MLP_IN + MLP_OUT;
  std::vector<double> ps(CBH * M, 0.), pq(CBH * M, 0.);
  for (int t = 0; t < CBH; t++)
    for (long long k = t; k < n; k += CBH)
      for (int i = 0; i < M; i++) {
        const double v = (i < MLP_IN) ? (double)feat[i*n + k] : (double)target[(i - MLP_IN)*n + k];
        ps[t*M + i] += v;
        pq[t*M + i] += v*v;
      }
  double sum[M], sumsq[M];
  for (int i = 0; i < M; i++) {
    sum[i] = sumsq[i] = 0.;
    for (int t = 0; t < CBH; t++) { sum[i] += ps[t*M + i]; sumsq[i] += pq[t*M + i]; }
  }
  mlp_initial_blob(w, seed, sum, sumsq, n);
  return 0;
}

int hgpu_mlp_train(void*, long long n, const float* feat, const float* target, unsigned long long seed, int steps,
                   int batch, float lr, void* work, long long work_bytes, float* w, float* loss_hist) {
  if (n < 1 || n > INT_MAX) return set_error(-2, "n must be in [1, 2^31)");
  if (steps < 0 || batch < 1 || !(lr > 0.f)) return set_error(-2, "steps >= 0, batch >= 1 and lr > 0 required");
  if (!feat || !target || !w || !work) return set_error(-1, "null argument");
  if (work_bytes < mlp_work_bytes(batch)) return set_error(-3, "workspace smaller than hgpu_mlp_workspace_bytes()");
  MlpWork W = mlp_work_layout(work, batch);
  for (int p = 0; p < MLP_NTRAIN; p++) W.m[p] = W.v[p] = 0.f;
  float c1 = 1.f, c2 = 1.f;
  for (int t = 1; t <= steps; t++) {
    grad_host(n, w, feat, target, batch, 1, seed, t, W.grad);
    if (loss_hist) loss_hist[t - 1] = W.grad[MLP_NTRAIN];
    c1 *= ADAM_B1;
    c2 *= ADAM_B2;
    const float lr_t = cosine_lr(lr, t - 1, steps);
    for (int p = 0; p < MLP_NTRAIN; p++)
      adam_update(&w[MLP_TRAIN_OFF + p], W.grad[p], &W.m[p], &W.v[p], lr_t, 1.f - c1, 1.f - c2);
  }
  return 0;
}

}  // extern "C"
