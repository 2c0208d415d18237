// hgpu_c2p.hip -- MI355X kernels of the batched con-to-prim and of its learned
// initial guess (forward, backward and Adam), behind the context-free entry
// points of include/hgpu.h.
//
// Solve: a persistent first pass (a machine-filling grid striding over the
// zones) sets every zone up and runs a fixed Newton budget; zones still
// iterating are parked in a queue (wave-aggregated slot allocation, SoA
// records).  Each later pass strides over the previous queue only, so every
// wavefront works on live zones; the last pass runs to completion.  The
// per-zone decision sequence is the one-pass solve's exactly.
//
// Network: inference is one thread per zone with the weights as wave-uniform
// (scalar) operands.  The gradient is one wavefront per 64 samples: the
// forward pass is per sample, the weight gradients are sums over the
// wavefront's samples formed from activations staged in LDS (rows padded to 65
// words, conflict-free), and the per-workgroup partials are summed in fixed
// order, so gradients are deterministic.
#include "hgpu_c2p_api.h"
#include <hip/hip_runtime.h>
#include <algorithm>
#include <mutex>
#include <string>

namespace hgpu {
namespace {

constexpr int CB = 256;   // workgroup size of the zone kernels

int hipc(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  return set_error(-4, std::string(what) + ": " + hipGetErrorString(e));
}

int device_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 256;
  }
  return cus;
}

// workgroups of a persistent pass: 8 per CU (2 per SIMD) or fewer for small n
unsigned persistent_blocks(long long n) {
  const long long need = (n + CB - 1) / CB;
  return (unsigned)std::max<long long>(1, std::min<long long>(need, 8LL * device_cus()));
}

__global__ __launch_bounds__(CB) void k_c2p_samples(long long n, unsigned long long seed, double a, double gam,
                                                   double* U, double* rth, double* prim, double* guess) {
  const long long t = (long long)blockIdx.x * CB + threadIdx.x;
  if (t >= n) return;
  double u[NP], r[2], p[NP], g[NP];
  c2p_sample(seed, (unsigned long long)t, a, gam, u, r, p, guess ? g : nullptr);
  for (int l = 0; l < NP; l++) { U[l*n + t] = u[l]; prim[l*n + t] = p[l]; }
  rth[t] = r[0];
  rth[n + t] = r[1];
  if (guess)
    for (int l = 0; l < NP; l++) guess[l*n + t] = g[l];
}

// Plain streaming copy (16 B per lane per iteration): the reference kernel of
// known traffic that calibrates the profiler's HBM byte counters.
__global__ __launch_bounds__(CB) void k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, long long n16) {
  const long long stride = (long long)gridDim.x * CB;
  for (long long t = (long long)blockIdx.x * CB + threadIdx.x; t < n16; t += stride) dst[t] = src[t];
}

__global__ void k_fill_u32(unsigned* p, int m) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) p[t] = 0u;
}

// Slot in *count for each lane with park set: one atomic per wavefront.
__device__ __forceinline__ long long wave_slot(bool park, unsigned* count) {
  const unsigned long long ballot = __ballot(park);
  if (!ballot) return -1;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)ballot) - 1;
  unsigned base = 0;
  if (lane == leader) base = atomicAdd(count, (unsigned)__popcll(ballot));
  base = __shfl(base, leader, 64);
  return park ? (long long)base + __popcll(ballot & ((1ull << lane) - 1ull)) : -1;
}

template <bool MLP>
__global__ __launch_bounds__(CB) void k_c2p_first(C2pArgs A, const float* __restrict__ mlp, int budget,
                                                 C2pQueue Q, unsigned* count) {
  const long long stride = (long long)gridDim.x * CB;
  for (long long base = (long long)blockIdx.x * CB; base < A.n; base += stride) {
    const long long t = base + threadIdx.x;
    C2pCell c;
    bool park = false;
    if (t < A.n) park = c2p_first_zone(A, MLP ? mlp : nullptr, t, budget, c);
    if (Q.cap == 0) continue;   // one-pass solve: the budget covers every attempt
    const long long q = wave_slot(park, count);
    if (park) c2p_park(Q, q, (int)t, c);
  }
}

__global__ __launch_bounds__(CB) void k_c2p_next(C2pArgs A, C2pQueue Qin, const unsigned* count_in, int budget,
                                                C2pQueue Qout, unsigned* count_out) {
  const long long m = *count_in;
  const long long stride = (long long)gridDim.x * CB;
  for (long long base = (long long)blockIdx.x * CB; base < m; base += stride) {
    const long long q = base + threadIdx.x;
    C2pCell c;
    int cell = 0;
    bool park = false;
    if (q < m) park = c2p_next_zone(A, Qin, q, budget, c, &cell);
    const long long s = wave_slot(park, count_out);
    if (park) c2p_park(Qout, s, cell, c);
  }
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(CB) void k_mlp_features(long long n, double a, double gam, const double* U,
                                                    const double* rth, const double* prim, float* feat,
                                                    float* target) {
  const long long t = (long long)blockIdx.x * CB + threadIdx.x;
  if (t >= n) return;
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
  if (prim && target) {
    for (int l = 0; l < NP; l++) pr[l] = prim[l*n + t];
    double gamma = 1.;
    gamma_calc(pr, &G, &gamma);
    float y[MLP_OUT];
    mlp_targets(pr, gamma, gam, y);
    for (int k = 0; k < MLP_OUT; k++) target[k*n + t] = y[k];
  }
}

__global__ __launch_bounds__(CB) void k_mlp_forward(long long n, const float* __restrict__ w, const float* feat,
                                                   float* y) {
  const long long t = (long long)blockIdx.x * CB + threadIdx.x;
  if (t >= n) return;
  float f[MLP_IN], o[MLP_OUT];
  for (int i = 0; i < MLP_IN; i++) f[i] = feat[i*n + t];
  mlp_forward(w, f, o);
  for (int k = 0; k < MLP_OUT; k++) y[k*n + t] = o[k];
}

// Gradient partial of one workgroup of MLP_GT samples.  Samples are
// s = blockIdx.x*MLP_GT + lane (< batch): index s itself, or with random set
// the draw of training step `step`.
constexpr int LP = MLP_H + 1;   // padded LDS row
__global__ __launch_bounds__(MLP_GT) void k_mlp_grad(long long n, const float* __restrict__ w, const float* feat,
                                                    const float* target, long long batch, int random,
                                                    unsigned long long seed, int step, float* part) {
  __shared__ float sX[MLP_GT][MLP_IN];
  __shared__ float sH1[MLP_GT][LP];
  __shared__ float sH2[MLP_GT][LP];    // h2, later dz1
  __shared__ float sDZ2[MLP_GT][LP];
  __shared__ float sDY[MLP_GT][MLP_OUT];
  __shared__ float sLoss[MLP_GT];
  const int lane = threadIdx.x;
  const long long s = (long long)blockIdx.x * MLP_GT + lane;
  const bool valid = s < batch;
  long long k = 0;
  if (valid) k = random ? mlp_batch_index(seed, step, s, batch, n) : s;
  // forward, one sample per lane
  float x[MLP_IN], z1[MLP_H];
  for (int i = 0; i < MLP_IN; i++)
    x[i] = valid ? (feat[i*n + k] - w[MLP_OFF_MU + i]) * w[MLP_OFF_ISD + i] : 0.f;
  for (int i = 0; i < MLP_IN; i++) sX[lane][i] = x[i];
  for (int j = 0; j < MLP_H; j++) {
    float z = w[MLP_OFF_B1 + j];
    for (int i = 0; i < MLP_IN; i++) z = fmaf(x[i], w[MLP_OFF_W1 + i*MLP_H + j], z);
    z1[j] = z;
    sH1[lane][j] = gelu(z);
  }
  float o[MLP_OUT] = {w[MLP_OFF_B3 + 0], w[MLP_OFF_B3 + 1]};
  float z2l[MLP_H];
  for (int j = 0; j < MLP_H; j++) {
    float z = w[MLP_OFF_B2 + j];
    for (int i = 0; i < MLP_H; i++) z = fmaf(sH1[lane][i], w[MLP_OFF_W2 + i*MLP_H + j], z);
    z2l[j] = z;
    const float h2 = gelu(z);
    sH2[lane][j] = h2;
    for (int q = 0; q < MLP_OUT; q++) o[q] = fmaf(h2, w[MLP_OFF_W3 + j*MLP_OUT + q], o[q]);
  }
  float dy[MLP_OUT], loss = 0.f;
  for (int q = 0; q < MLP_OUT; q++) {
    const float tn = valid ? (target[q*n + k] - w[MLP_OFF_YMU + q]) / w[MLP_OFF_YSD + q] : 0.f;
    dy[q] = valid ? o[q] - tn : 0.f;
    loss += 0.5f*dy[q]*dy[q];
    sDY[lane][q] = dy[q];
  }
  sLoss[lane] = loss;
  for (int j = 0; j < MLP_H; j++) {
    float dh = 0.f;
    for (int q = 0; q < MLP_OUT; q++) dh = fmaf(w[MLP_OFF_W3 + j*MLP_OUT + q], dy[q], dh);
    sDZ2[lane][j] = dh*gelu_grad(z2l[j]);
  }
  __syncthreads();
  float* P = part + (long long)blockIdx.x * MLP_PART;   // gradient slot of blob offset o: P[o - MLP_TRAIN_OFF]
  constexpr int GW1 = MLP_OFF_W1 - MLP_TRAIN_OFF, GB1 = MLP_OFF_B1 - MLP_TRAIN_OFF, GW2 = MLP_OFF_W2 - MLP_TRAIN_OFF,
                GB2 = MLP_OFF_B2 - MLP_TRAIN_OFF, GW3 = MLP_OFF_W3 - MLP_TRAIN_OFF, GB3 = MLP_OFF_B3 - MLP_TRAIN_OFF;
  // sums over the samples: lane = hidden unit j (or i)
  {
    const int j = lane;
    float g3[MLP_OUT] = {0.f, 0.f}, gb2 = 0.f;
    for (int r = 0; r < MLP_GT; r++) {
      const float h2 = sH2[r][j];
      for (int q = 0; q < MLP_OUT; q++) g3[q] = fmaf(h2, sDY[r][q], g3[q]);
      gb2 += sDZ2[r][j];
    }
    for (int q = 0; q < MLP_OUT; q++) P[GW3 + j*MLP_OUT + q] = g3[q];
    P[GB2 + j] = gb2;
    if (lane < MLP_OUT) {
      float gb3 = 0.f;
      for (int r = 0; r < MLP_GT; r++) gb3 += sDY[r][lane];
      P[GB3 + lane] = gb3;
    }
    if (lane == 0) {
      float l = 0.f;
      for (int r = 0; r < MLP_GT; r++) l += sLoss[r];
      part[(long long)blockIdx.x * MLP_PART + MLP_NTRAIN] = l;
    }
  }
  {
    const int i = lane;
    float acc[MLP_H];
    for (int jj = 0; jj < MLP_H; jj++) acc[jj] = 0.f;
    for (int r = 0; r < MLP_GT; r++) {
      const float h1 = sH1[r][i];
      for (int jj = 0; jj < MLP_H; jj++) acc[jj] = fmaf(h1, sDZ2[r][jj], acc[jj]);
    }
    for (int jj = 0; jj < MLP_H; jj++) P[GW2 + i*MLP_H + jj] = acc[jj];
  }
  __syncthreads();
  // dz1 of this lane's sample into sH2 (h2 is no longer needed)
  for (int i = 0; i < MLP_H; i++) {
    float dh = 0.f;
    for (int j = 0; j < MLP_H; j++) dh = fmaf(w[MLP_OFF_W2 + i*MLP_H + j], sDZ2[lane][j], dh);
    sH2[lane][i] = dh*gelu_grad(z1[i]);
  }
  __syncthreads();
  {
    const int i = lane;
    float g1[MLP_IN], gb1 = 0.f;
    for (int q = 0; q < MLP_IN; q++) g1[q] = 0.f;
    for (int r = 0; r < MLP_GT; r++) {
      const float dz1 = sH2[r][i];
      gb1 += dz1;
      for (int q = 0; q < MLP_IN; q++) g1[q] = fmaf(sX[r][q], dz1, g1[q]);
    }
    for (int q = 0; q < MLP_IN; q++) P[GW1 + q*MLP_H + i] = g1[q];
    P[GB1 + i] = gb1;
  }
}

// grad[p] = sum over workgroups of part[b][p] / batch (fixed order); the last
// slot is the mean loss, also copied to *loss_out when given.
__global__ __launch_bounds__(CB) void k_mlp_reduce(const float* part, long long nblk, long long batch,
                                                  float* grad, float* loss_out) {
  const int p = blockIdx.x * CB + threadIdx.x;
  if (p >= MLP_PART) return;
  float acc = 0.f;
  for (long long b = 0; b < nblk; b++) acc += part[b * MLP_PART + p];
  acc /= (float)batch;
  grad[p] = acc;
  if (p == MLP_NTRAIN && loss_out) *loss_out = acc;
}

__global__ __launch_bounds__(CB) void k_mlp_adam(float* w, const float* grad, float* m, float* v, float lr_t,
                                                float c1, float c2) {
  const int p = blockIdx.x * CB + threadIdx.x;
  if (p >= MLP_NTRAIN) return;
  adam_update(&w[MLP_TRAIN_OFF + p], grad[p], &m[p], &v[p], lr_t, c1, c2);
}

__global__ __launch_bounds__(CB) void k_zero_f32(float* p, long long m) {
  const long long t = (long long)blockIdx.x * CB + threadIdx.x;
  if (t < m) p[t] = 0.f;
}

// Moments of features and targets: one workgroup, fixed reduction order.
__global__ __launch_bounds__(CB) void k_mlp_moments(long long n, const float* feat, const float* target,
                                                   double* out /* [2][10] */) {
  __shared__ double sS[CB][MLP_IN + MLP_OUT + 1];
  __shared__ double sQ[CB][MLP_IN + MLP_OUT + 1];
  const int t = threadIdx.x;
  double s[MLP_IN + MLP_OUT], q[MLP_IN + MLP_OUT];
  for (int i = 0; i < MLP_IN + MLP_OUT; i++) s[i] = q[i] = 0.;
  for (long long k = t; k < n; k += CB) {
    for (int i = 0; i < MLP_IN + MLP_OUT; i++) {
      const double v = (i < MLP_IN) ? (double)feat[i*n + k] : (double)target[(i - MLP_IN)*n + k];
      s[i] += v;
      q[i] += v*v;
    }
  }
  for (int i = 0; i < MLP_IN + MLP_OUT; i++) { sS[t][i] = s[i]; sQ[t][i] = q[i]; }
  __syncthreads();
  if (t < MLP_IN + MLP_OUT) {
    double a = 0., b = 0.;
    for (int r = 0; r < CB; r++) { a += sS[r][t]; b += sQ[r][t]; }
    out[t] = a;
    out[MLP_IN + MLP_OUT + t] = b;
  }
}

unsigned nb(long long n) { return (unsigned)((n + CB - 1) / CB); }
hipStream_t st(void* s) { return static_cast<hipStream_t>(s); }

int grad_launch(hipStream_t stream, long long n, const float* w, const float* feat, const float* target,
                long long batch, int random, unsigned long long seed, int step, const MlpWork& W, float* loss_out) {
  const long long nblk = (batch + MLP_GT - 1) / MLP_GT;
  hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)nblk), dim3(MLP_GT), 0, stream, n, w, feat, target, batch, random,
                     seed, step, W.part);
  hipLaunchKernelGGL(k_mlp_reduce, dim3(nb(MLP_PART)), dim3(CB), 0, stream, W.part, nblk, batch, W.grad, loss_out);
  return hipc(hipGetLastError(), "k_mlp_grad");
}

// Launches of a compacted solve: the persistent first pass of budget_first
// iterations, then passes of budget_next over the queue of zones still
// iterating (the last pass runs to completion).  The queue sizes are cleared
// by a kernel (no memset node, so the sequence can be captured in a graph).
int c2p_compacted(hipStream_t s, const C2pArgs& A, const float* mlp, int budget_first, int budget_next,
                  const C2pWork& W, unsigned G) {
  const int passes = c2p_npasses(budget_first, budget_next);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, W.count, 64);
  if (mlp) hipLaunchKernelGGL(k_c2p_first<true>, dim3(G), dim3(CB), 0, s, A, mlp, budget_first, W.q[0], W.count);
  else hipLaunchKernelGGL(k_c2p_first<false>, dim3(G), dim3(CB), 0, s, A, mlp, budget_first, W.q[0], W.count);
  for (int p = 1; p <= passes; p++) {
    const int budget = (p == passes) ? C2P_MAX_ITERS : budget_next;
    hipLaunchKernelGGL(k_c2p_next, dim3(G), dim3(CB), 0, s, A, W.q[(p - 1) & 1], W.count + (p - 1), budget,
                       W.q[p & 1], W.count + p);
  }
  return hipc(hipGetLastError(), "k_c2p_next");
}

// Static scratch of hgpu_c2p_solve_static(): the queue sizes and both
// compaction queues for C2P_STATIC_MAX zones (256 + 2 x 160 B per zone,
// 80 MiB).  Every GPU the library runs on holds its own copy until the process
// ends.  Host code never uses the array's name as an address: each device's
// address comes from hipGetSymbolAddress, once per device, with that device
// current, and is kept in scratch_ptr.
constexpr long long C2P_STATIC_MAX = 1LL << 18;
constexpr int MAX_DEVICES = 64;
static __device__ __align__(256) char scratch[256 + 2 * (C2P_QD * sizeof(double) + 2 * sizeof(int)) * C2P_STATIC_MAX];
static_assert(sizeof(scratch) == (size_t)(256 + 2 * (C2P_QD * sizeof(double) + 2 * sizeof(int)) * C2P_STATIC_MAX),
              "scratch must hold the compaction layout of C2P_STATIC_MAX zones");
char* scratch_ptr[MAX_DEVICES];
std::once_flag scratch_once[MAX_DEVICES];
int scratch_rc[MAX_DEVICES];

int static_scratch(char** out) {
  int dev = 0;
  if (int rc = hipc(hipGetDevice(&dev), "hipGetDevice")) return rc;
  if (dev < 0 || dev >= MAX_DEVICES) return set_error(-4, "device index beyond the static scratch table");
  std::call_once(scratch_once[dev], [dev] {
    void* p = nullptr;
    scratch_rc[dev] = hipc(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    scratch_ptr[dev] = static_cast<char*>(p);
  });
  if (scratch_rc[dev]) return scratch_rc[dev];
  if (!scratch_ptr[dev]) return set_error(-4, "hipGetSymbolAddress(scratch) returned no address");
  *out = scratch_ptr[dev];
  return 0;
}

}  // namespace
}  // namespace hgpu

using namespace hgpu;

extern "C" {

int hgpu_dev_count(int* out) {
  if (!out) return set_error(-1, "null argument");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *out = n;
  return 0;
}

int hgpu_dev_alloc(long long bytes, void** out) {
  if (!out || bytes < 0) return set_error(-1, "bad argument");
  *out = nullptr;
  return hipc(hipMalloc(out, (size_t)(bytes > 0 ? bytes : 8)), "hipMalloc");
}

int hgpu_dev_free(void* p) { return p ? hipc(hipFree(p), "hipFree") : 0; }

int hgpu_dev_copy(void* dst, const void* src, long long bytes, int kind, void* stream) {
  if (bytes < 0 || (bytes > 0 && (!dst || !src))) return set_error(-1, "bad argument");
  if (bytes == 0) return 0;
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                        : kind == 2 ? hipMemcpyDeviceToDevice : hipMemcpyHostToHost;
  if (kind < 0 || kind > 2) return set_error(-2, "kind must be 0 (to device), 1 (to host) or 2 (device to device)");
  if (int rc = hipc(hipMemcpyAsync(dst, src, (size_t)bytes, k, st(stream)), "hipMemcpyAsync")) return rc;
  return kind == 1 ? hipc(hipStreamSynchronize(st(stream)), "hipStreamSynchronize") : 0;
}

int hgpu_dev_copy_kernel(void* dst, const void* src, long long bytes, void* stream) {
  if (bytes < 0 || bytes % 16 || (bytes > 0 && (!dst || !src))) return set_error(-1, "bytes must be a multiple of 16");
  if (reinterpret_cast<uintptr_t>(dst) % 16 || reinterpret_cast<uintptr_t>(src) % 16)
    return set_error(-2, "buffers must be 16-byte aligned");
  if (bytes == 0) return 0;
  const long long n16 = bytes / 16;
  hipLaunchKernelGGL(k_copy16, dim3(persistent_blocks(n16)), dim3(CB), 0, st(stream), static_cast<uint4*>(dst),
                     static_cast<const uint4*>(src), n16);
  return hipc(hipGetLastError(), "k_copy16");
}

int hgpu_stream_sync(void* stream) { return hipc(hipStreamSynchronize(st(stream)), "hipStreamSynchronize"); }

int hgpu_event_create(void** ev) {
  if (!ev) return set_error(-1, "null argument");
  return hipc(hipEventCreate(reinterpret_cast<hipEvent_t*>(ev)), "hipEventCreate");
}
int hgpu_event_destroy(void* ev) { return ev ? hipc(hipEventDestroy(static_cast<hipEvent_t>(ev)), "hipEventDestroy") : 0; }
int hgpu_event_record(void* ev, void* stream) {
  return hipc(hipEventRecord(static_cast<hipEvent_t>(ev), st(stream)), "hipEventRecord");
}
int hgpu_event_elapsed(void* ev0, void* ev1, float* ms) {
  if (!ms) return set_error(-1, "null argument");
  if (int rc = hipc(hipEventSynchronize(static_cast<hipEvent_t>(ev1)), "hipEventSynchronize")) return rc;
  return hipc(hipEventElapsedTime(ms, static_cast<hipEvent_t>(ev0), static_cast<hipEvent_t>(ev1)), "hipEventElapsedTime");
}

int hgpu_c2p_samples(void* stream, long long n, unsigned long long seed, double a, double gam, double* U,
                     double* rth, double* prim, double* guess) {
  if (int rc = c2p_check_common(n, a, gam)) return rc;
  if (n == 0) return 0;
  if (!U || !rth || !prim) return set_error(-1, "null array argument");
  hipLaunchKernelGGL(k_c2p_samples, dim3(nb(n)), dim3(CB), 0, st(stream), n, seed, a, gam, U, rth, prim, guess);
  return hipc(hipGetLastError(), "k_c2p_samples");
}

long long hgpu_c2p_workspace_bytes(long long n, int compact) { return c2p_work_bytes(n, compact); }

int hgpu_c2p_solve(void* stream, long long n, double a, double gam, const double* U, const double* rth,
                   const float* mlp, int plain_guess, int compact, int budget_first, int budget_next, void* work,
                   long long work_bytes, double* prim, int* code, int* iters) {
  if (int rc = c2p_check_solve(n, a, gam, U, rth, plain_guess, compact, budget_first, budget_next, work,
                               work_bytes, prim, code, iters))
    return rc;
  if (n == 0) return 0;
  hipStream_t s = st(stream);
  C2pArgs A{n, a, gam, (gam - 1.)/gam, U, rth, plain_guess, prim, code, iters};
  const unsigned G = persistent_blocks(n);
  if (!compact) {
    C2pQueue none{nullptr, nullptr, nullptr, 0};
    if (mlp) hipLaunchKernelGGL(k_c2p_first<true>, dim3(G), dim3(CB), 0, s, A, mlp, C2P_MAX_ITERS, none, nullptr);
    else hipLaunchKernelGGL(k_c2p_first<false>, dim3(G), dim3(CB), 0, s, A, mlp, C2P_MAX_ITERS, none, nullptr);
    return hipc(hipGetLastError(), "k_c2p_first");
  }
  return c2p_compacted(s, A, mlp, budget_first, budget_next, c2p_work_layout(work, n), G);
}

/* Compacted solve with its queues in the library's static scratch array
 * (no allocation; at most C2P_STATIC_MAX zones per call). */
int hgpu_c2p_solve_static(void* stream, long long n, double a, double gam, const double* U, const double* rth,
                          const float* mlp, int plain_guess, int budget_first, int budget_next, double* prim,
                          int* code, int* iters) {
  if (n > C2P_STATIC_MAX)
    return set_error(-3, "n exceeds the static scratch capacity (" + std::to_string(C2P_STATIC_MAX) +
                             " zones per call); use hgpu_c2p_solve() with a workspace");
  char* ws = nullptr;
  if (n > 0)
    if (int rc = static_scratch(&ws)) return rc;
  if (int rc = c2p_check_solve(n, a, gam, U, rth, plain_guess, 1, budget_first, budget_next, ws,
                               (long long)sizeof(scratch), prim, code, iters))
    return rc;
  if (n == 0) return 0;
  C2pArgs A{n, a, gam, (gam - 1.)/gam, U, rth, plain_guess, prim, code, iters};
  return c2p_compacted(st(stream), A, mlp, budget_first, budget_next, c2p_work_layout(ws, n), persistent_blocks(n));
}

int hgpu_mlp_sizes(int* out) {
  if (!out) return set_error(-1, "null argument");
  out[0] = MLP_NPARAM; out[1] = MLP_NTRAIN; out[2] = MLP_TRAIN_OFF;
  out[3] = MLP_IN; out[4] = MLP_H; out[5] = MLP_OUT;
  return 0;
}

int hgpu_mlp_features(void* stream, long long n, double a, double gam, const double* U, const double* rth,
                      const double* prim, float* feat, float* target) {
  if (int rc = c2p_check_common(n, a, gam)) return rc;
  if (n == 0) return 0;
  if (!U || !rth || !feat) return set_error(-1, "null array argument");
  if (prim && !target) return set_error(-1, "targets requested without an output array");
  hipLaunchKernelGGL(k_mlp_features, dim3(nb(n)), dim3(CB), 0, st(stream), n, a, gam, U, rth, prim, feat, target);
  return hipc(hipGetLastError(), "k_mlp_features");
}

int hgpu_mlp_forward(void* stream, long long n, const float* w, const float* feat, float* y) {
  if (n < 0 || n > INT_MAX) return set_error(-2, "n must be in [0, 2^31)");
  if (n == 0) return 0;
  if (!w || !feat || !y) return set_error(-1, "null array argument");
  hipLaunchKernelGGL(k_mlp_forward, dim3(nb(n)), dim3(CB), 0, st(stream), n, w, feat, y);
  return hipc(hipGetLastError(), "k_mlp_forward");
}

long long hgpu_mlp_workspace_bytes(long long batch) { return mlp_work_bytes(batch); }

int hgpu_mlp_grad(void* stream, long long n, const float* w, const float* feat, const float* target, void* work,
                  long long work_bytes, float* grad, float* loss) {
  if (n < 1 || n > INT_MAX) return set_error(-2, "n must be in [1, 2^31)");
  if (!w || !feat || !target || !work || !grad || !loss) return set_error(-1, "null argument");
  if (work_bytes < mlp_work_bytes(n)) return set_error(-3, "workspace smaller than hgpu_mlp_workspace_bytes()");
  hipStream_t s = st(stream);
  MlpWork W = mlp_work_layout(work, n);
  if (int rc = grad_launch(s, n, w, feat, target, n, 0, 0ull, 0, W, loss)) return rc;
  return hipc(hipMemcpyAsync(grad, W.grad, sizeof(float) * MLP_NTRAIN, hipMemcpyDeviceToDevice, s),
              "hipMemcpyAsync");
}

int hgpu_mlp_init(void* stream, unsigned long long seed, long long n, const float* feat, const float* target,
                  float* w) {
  if (n < 1 || n > INT_MAX) return set_error(-2, "n must be in [1, 2^31)");
  if (!feat || !target || !w) return set_error(-1, "null argument");
  if (reinterpret_cast<uintptr_t>(w) % 8) return set_error(-2, "w must be 8-byte aligned");
  hipStream_t s = st(stream);
  double mom[2 * (MLP_IN + MLP_OUT)];
  // the moments are staged in the first 160 bytes of w, which is rewritten below
  hipLaunchKernelGGL(k_mlp_moments, dim3(1), dim3(CB), 0, s, n, feat, target, reinterpret_cast<double*>(w));
  if (int rc = hipc(hipGetLastError(), "k_mlp_moments")) return rc;
  if (int rc = hipc(hipMemcpyAsync(mom, w, sizeof(mom), hipMemcpyDeviceToHost, s), "hipMemcpyAsync")) return rc;
  if (int rc = hipc(hipStreamSynchronize(s), "hipStreamSynchronize")) return rc;
  float blob[MLP_NPARAM];
  mlp_initial_blob(blob, seed, mom, mom + MLP_IN + MLP_OUT, n);
  if (int rc = hipc(hipMemcpyAsync(w, blob, sizeof(blob), hipMemcpyHostToDevice, s), "hipMemcpyAsync")) return rc;
  return hipc(hipStreamSynchronize(s), "hipStreamSynchronize");
}

int hgpu_mlp_train(void* stream, long long n, const float* feat, const float* target, unsigned long long seed,
                   int steps, int batch, float lr, void* work, long long work_bytes, float* w, float* loss_hist) {
  if (n < 1 || n > INT_MAX) return set_error(-2, "n must be in [1, 2^31)");
  if (steps < 0 || batch < 1 || !(lr > 0.f)) return set_error(-2, "steps >= 0, batch >= 1 and lr > 0 required");
  if (!feat || !target || !w || !work) return set_error(-1, "null argument");
  if (work_bytes < mlp_work_bytes(batch)) return set_error(-3, "workspace smaller than hgpu_mlp_workspace_bytes()");
  hipStream_t s = st(stream);
  MlpWork W = mlp_work_layout(work, batch);
  hipLaunchKernelGGL(k_zero_f32, dim3(nb(2 * MLP_NTRAIN)), dim3(CB), 0, s, W.m, (long long)(2 * MLP_NTRAIN));
  float c1 = 1.f, c2 = 1.f;
  for (int t = 1; t <= steps; t++) {
    if (int rc = grad_launch(s, n, w, feat, target, batch, 1, seed, t, W, loss_hist ? loss_hist + (t - 1) : nullptr))
      return rc;
    c1 *= ADAM_B1;
    c2 *= ADAM_B2;
    hipLaunchKernelGGL(k_mlp_adam, dim3(nb(MLP_NTRAIN)), dim3(CB), 0, s, w, W.grad, W.m, W.v,
                       cosine_lr(lr, t - 1, steps), 1.f - c1, 1.f - c2);
  }
  return hipc(hipGetLastError(), "k_mlp_adam");
}

}  // extern "C"
