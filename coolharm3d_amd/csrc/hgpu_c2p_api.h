// hgpu_c2p_api.h -- host-side argument checks and workspace layouts of the
// batched con-to-prim and network entry points, shared by the HIP library and
// the host build.
#pragma once
#include "hgpu_core.h"
#include "hgpu_c2p.h"
#include <climits>

namespace hgpu {

constexpr int C2P_BUDGET_FIRST = 16, C2P_BUDGET_NEXT = 8;   // measured best on the 2^24-zone workload
constexpr int MLP_GT = 64;                  // samples per gradient workgroup
constexpr int MLP_PART = MLP_NTRAIN + 1;    // per-workgroup partial: gradient + loss

inline int c2p_npasses(int budget_first, int budget_next) {
  // passes after the first; the last one runs to completion
  int rest = C2P_MAX_ITERS - budget_first;
  return rest <= 0 ? 0 : (rest + budget_next - 1) / budget_next;
}

struct C2pWork {
  unsigned* count;     // [C2P_MAX_ITERS + 2] queue sizes, one per pass
  C2pQueue q[2];
};

inline long long c2p_queue_bytes(long long n) {
  return (long long)(C2P_QD * sizeof(double) + 2 * sizeof(int)) * n;
}

inline long long c2p_work_bytes(long long n, int compact) {
  if (!compact) return 0;
  return 256 + 2 * c2p_queue_bytes(n);
}

inline C2pWork c2p_work_layout(void* work, long long n) {
  C2pWork W;
  char* p = static_cast<char*>(work);
  W.count = reinterpret_cast<unsigned*>(p);
  p += 256;
  for (int b = 0; b < 2; b++) {
    W.q[b].cap = n;
    W.q[b].d = reinterpret_cast<double*>(p);
    p += sizeof(double) * C2P_QD * n;
    W.q[b].cell = reinterpret_cast<int*>(p);
    p += sizeof(int) * n;
    W.q[b].meta = reinterpret_cast<int*>(p);
    p += sizeof(int) * n;
  }
  return W;
}

inline int c2p_check_common(long long n, double a, double gam) {
  if (n < 0 || n > INT_MAX) return set_error(-2, "n must be in [0, 2^31)");
  if (!(fabs(a) < 1.)) return set_error(-2, "spin a must satisfy |a| < 1");
  if (!(gam > 1.)) return set_error(-2, "gam must exceed 1");
  return 0;
}

inline int c2p_check_solve(long long n, double a, double gam, const double* U, const double* rth, int plain_guess,
                           int compact, int& budget_first, int& budget_next, void* work, long long work_bytes,
                           double* prim, int* code, int* iters) {
  if (int rc = c2p_check_common(n, a, gam)) return rc;
  if (n > 0 && (!U || !rth || !prim || !code || !iters)) return set_error(-1, "null array argument");
  if (plain_guess != 0 && plain_guess != 1) return set_error(-2, "plain_guess must be 0 (incoming prim) or 1 (cold)");
  if (compact != 0 && compact != 1) return set_error(-2, "compact must be 0 or 1");
  if (budget_first == 0) budget_first = C2P_BUDGET_FIRST;
  if (budget_next == 0) budget_next = C2P_BUDGET_NEXT;
  if (budget_first < 1 || budget_first > C2P_MAX_ITERS || budget_next < 1 || budget_next > C2P_MAX_ITERS)
    return set_error(-2, "Newton budgets must lie in [1, 60]");
  if (compact && n > 0) {
    if (!work) return set_error(-1, "compact = 1 needs the workspace");
    if (work_bytes < c2p_work_bytes(n, 1)) return set_error(-3, "workspace smaller than hgpu_c2p_workspace_bytes()");
  }
  return 0;
}

inline long long mlp_work_bytes(long long batch) {
  const long long nblk = (batch + MLP_GT - 1) / MLP_GT;
  return (long long)sizeof(float) * (nblk * MLP_PART + MLP_PART + 2 * MLP_NTRAIN) + 256;
}

struct MlpWork {
  float* part;   // [nblk][MLP_PART]
  float* grad;   // [MLP_PART]
  float* m;      // [MLP_NTRAIN] Adam moments
  float* v;      // [MLP_NTRAIN]
};

inline MlpWork mlp_work_layout(void* work, long long batch) {
  const long long nblk = (batch + MLP_GT - 1) / MLP_GT;
  MlpWork W;
  float* p = static_cast<float*>(work);
  W.part = p; p += nblk * MLP_PART;
  W.grad = p; p += MLP_PART;
  W.m = p; p += MLP_NTRAIN;
  W.v = p;
  return W;
}

// Adam (Kingma & Ba 2015) with a cosine-decayed step size.
constexpr float ADAM_B1 = 0.9f, ADAM_B2 = 0.999f, ADAM_EPS = 1e-8f;
inline float cosine_lr(float lr, int step, int steps) {
  return lr * 0.5f * (1.f + cosf(3.14159265358979f * (float)step / (float)steps));
}

HD void adam_update(float* w, float g, float* m, float* v, float lr_t, float c1, float c2) {
  *m = ADAM_B1 * (*m) + (1.f - ADAM_B1) * g;
  *v = ADAM_B2 * (*v) + (1.f - ADAM_B2) * g * g;
  const float mh = (*m) / c1, vh = (*v) / c2;
  *w -= lr_t * mh / (sqrtf(vh) + ADAM_EPS);
}

// Seeded initial weights: uniform in +-sqrt(6/(fan_in + fan_out)) (Glorot &
// Bengio 2010), zero biases.
HD float mlp_init_value(unsigned long long seed, int p) {
  int fan_in = 0, fan_out = 0;
  if (p >= MLP_OFF_W1 && p < MLP_OFF_B1) { fan_in = MLP_IN; fan_out = MLP_H; }
  else if (p >= MLP_OFF_W2 && p < MLP_OFF_B2) { fan_in = MLP_H; fan_out = MLP_H; }
  else if (p >= MLP_OFF_W3 && p < MLP_OFF_B3) { fan_in = MLP_H; fan_out = MLP_OUT; }
  else return 0.f;
  double u;
  draw_uniforms(seed, 2u, (unsigned long long)p, &u, 1);
  return (float)((2.*u - 1.) * sqrt(6. / (fan_in + fan_out)));
}

// Index of sample s of training step `step`, drawn with replacement.
HD long long mlp_batch_index(unsigned long long seed, int step, long long s, long long batch, long long n) {
  double u;
  draw_uniforms(seed, 3u, (unsigned long long)step * (unsigned long long)batch + (unsigned long long)s, &u, 1);
  long long k = (long long)(u * (double)n);
  return k < n ? k : n - 1;
}

// Initial parameter blob: normalisation of inputs and outputs from their
// moments (sum[i], sumsq[i] over n samples; i < 8 features, then 2 targets)
// and seeded weights.
inline void mlp_initial_blob(float* blob, unsigned long long seed, const double* sum, const double* sumsq,
                             long long n) {
  for (int p = 0; p < MLP_NPARAM; p++) blob[p] = mlp_init_value(seed, p);
  for (int i = 0; i < MLP_IN + MLP_OUT; i++) {
    const double mu = sum[i] / (double)n;
    const double var = sumsq[i] / (double)n - mu * mu;
    const double sd = var > 1e-12 ? sqrt(var) : 1.;
    if (i < MLP_IN) {
      blob[MLP_OFF_MU + i] = (float)mu;
      blob[MLP_OFF_ISD + i] = (float)(1. / sd);
    } else {
      blob[MLP_OFF_YMU + (i - MLP_IN)] = (float)mu;
      blob[MLP_OFF_YSD + (i - MLP_IN)] = (float)sd;
    }
  }
}

}  // namespace hgpu
