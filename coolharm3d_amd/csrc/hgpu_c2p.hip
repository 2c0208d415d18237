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

// ---------------------------------------------------------------------------
// The guess network on the matrix cores: one wavefront evaluates its 64 zones
// with FP32 MFMA (v_mfma_f32_16x16x4_f32: A 16x4, B 4x16, C 16x16; lane l
// supplies A[l % 16][l / 16] and B[l / 16][l % 16] and holds C[4 (l / 16) + r]
// [l % 16] in element r).  Activations are kept transposed, units x zones:
//   H1^T (64 x 64) = W1^T (64 x 8) X^T (8 x 64)         32 MFMA
//   H2^T (64 x 64) = W2^T (64 x 64) gelu(H1^T + b1)     256 MFMA
// so element r of an H1^T tile (unit 16 mt + 4 g + r, zone 16 nt + c, g = l / 16,
// c = l % 16) is already lane l's B operand of the layer-2 step that sums the
// units {16 mt + 4 g' + r : g' = 0..3}: the k-steps run over (mt, r) and the
// A operand (W2, stored transposed in LDS) takes the same unit order.  The
// output layer sums each lane's 16 units per zone and then across the four
// lane groups.  FP32 throughout; the summation order and the activation's
// exp2/reciprocal (gelu_hw) differ from the per-zone form (mlp_forward()), so
// the outputs agree to FP32 rounding.
typedef float mfma_f4 __attribute__((ext_vector_type(4)));

// LDS layout of the network's parameters for the wavefront-wide pass
constexpr int NL_W2T = 0;                       // W2 transposed: [out][in]
constexpr int NL_W1 = NL_W2T + MLP_H * MLP_H;   // W1 [in][unit]
constexpr int NL_B1 = NL_W1 + MLP_IN * MLP_H;
constexpr int NL_B2 = NL_B1 + MLP_H;
constexpr int NL_W3 = NL_B2 + MLP_H;            // W3 [unit][out]
constexpr int NL_MU = NL_W3 + MLP_H * MLP_OUT;
constexpr int NL_ISD = NL_MU + MLP_IN;
constexpr int NL_TAIL = NL_ISD + MLP_IN;        // b3, ymu, ysd
constexpr int NL_SIZE = NL_TAIL + 3 * MLP_OUT;
constexpr int NL_FEAT = MLP_IN + 1;             // padded feature row of one zone

// gelu() with the hardware's single-instruction exp2 and reciprocal (about 1
// ulp each), for the guess network's inference: x sigma(k (x + c x^3)).
__device__ __forceinline__ float gelu_hw(float x) {
  const float z = GELU_K * (x + GELU_C * x * x * x);
  return x * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}

// Every thread of the workgroup: the parameters into LDS.
__device__ void mlp_stage(const float* __restrict__ w, float* nl) {
  for (int e = threadIdx.x; e < MLP_H * MLP_H; e += blockDim.x) {
    const int o = e / MLP_H, i = e - o * MLP_H;
    nl[NL_W2T + e] = w[MLP_OFF_W2 + i * MLP_H + o];
  }
  for (int e = threadIdx.x; e < MLP_IN * MLP_H; e += blockDim.x) nl[NL_W1 + e] = w[MLP_OFF_W1 + e];
  for (int e = threadIdx.x; e < MLP_H; e += blockDim.x) {
    nl[NL_B1 + e] = w[MLP_OFF_B1 + e];
    nl[NL_B2 + e] = w[MLP_OFF_B2 + e];
  }
  for (int e = threadIdx.x; e < MLP_H * MLP_OUT; e += blockDim.x) nl[NL_W3 + e] = w[MLP_OFF_W3 + e];
  if (threadIdx.x < MLP_IN) {
    nl[NL_MU + threadIdx.x] = w[MLP_OFF_MU + threadIdx.x];
    nl[NL_ISD + threadIdx.x] = w[MLP_OFF_ISD + threadIdx.x];
  }
  if (threadIdx.x < MLP_OUT) {
    nl[NL_TAIL + threadIdx.x] = w[MLP_OFF_B3 + threadIdx.x];
    nl[NL_TAIL + MLP_OUT + threadIdx.x] = w[MLP_OFF_YMU + threadIdx.x];
    nl[NL_TAIL + 2 * MLP_OUT + threadIdx.x] = w[MLP_OFF_YSD + threadIdx.x];
  }
}

// Network outputs y of this lane's zone from its features f; every lane of
// the wavefront calls it (lanes without a zone pass finite features).
// feat: the wavefront's 64 x NL_FEAT LDS rows.
__device__ void mlp_forward_wave(const float* nl, const float* f, float* y, float* feat) {
  const int l = threadIdx.x & 63, g = l >> 4, c = l & 15;
#pragma unroll
  for (int i = 0; i < MLP_IN; i++) feat[l * NL_FEAT + i] = (f[i] - nl[NL_MU + i]) * nl[NL_ISD + i];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  mfma_f4 h1[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      mfma_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < MLP_IN / 4; s++) {
        const float a = nl[NL_W1 + (4 * s + g) * MLP_H + 16 * mt + c];
        const float b = feat[(16 * nt + c) * NL_FEAT + 4 * s + g];
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) acc[r] = gelu_hw(acc[r] + nl[NL_B1 + 16 * mt + 4 * g + r]);
      h1[mt][nt] = acc;
    }
  float o0[4], o1[4];
#pragma unroll
  for (int nt = 0; nt < 4; nt++) { o0[nt] = 0.f; o1[nt] = 0.f; }
#pragma unroll
  for (int ot = 0; ot < 4; ot++) {
    mfma_f4 acc[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) acc[nt] = (mfma_f4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      // W2^T[out 16 ot + c][in 16 mt + 4 g + r], r = 0..3: four consecutive floats
      const mfma_f4 a4 = *reinterpret_cast<const mfma_f4*>(&nl[NL_W2T + (16 * ot + c) * MLP_H + 16 * mt + 4 * g]);
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
          acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[r], h1[mt][nt][r], acc[nt], 0, 0, 0);
    }
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int u = 16 * ot + 4 * g + r;
        const float h = gelu_hw(acc[nt][r] + nl[NL_B2 + u]);
        o0[nt] = fmaf(h, nl[NL_W3 + u * MLP_OUT + 0], o0[nt]);
        o1[nt] = fmaf(h, nl[NL_W3 + u * MLP_OUT + 1], o1[nt]);
      }
  }
  float m0 = 0.f, m1 = 0.f;
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    float a = o0[nt], b = o1[nt];
    a += __shfl_xor(a, 16, 64); a += __shfl_xor(a, 32, 64);
    b += __shfl_xor(b, 16, 64); b += __shfl_xor(b, 32, 64);
    if (nt == g) { m0 = a; m1 = b; }   // zone 16 g + c is this lane's
  }
  y[0] = (m0 + nl[NL_TAIL + 0]) * nl[NL_TAIL + 2 * MLP_OUT + 0] + nl[NL_TAIL + MLP_OUT + 0];
  y[1] = (m1 + nl[NL_TAIL + 1]) * nl[NL_TAIL + 2 * MLP_OUT + 1] + nl[NL_TAIL + MLP_OUT + 1];
}

// The network's guess of every zone, one wavefront per 64 zones
// (mlp_forward_wave()), handed to the first pass through the code and
// iteration arrays, which the solve overwrites: y[0], y[1] as FP32 bits.  A
// kernel of its own keeps the matrix-core pass's registers (its activations)
// out of the Newton iteration's occupancy.
__global__ __launch_bounds__(CB) void k_c2p_guess(C2pArgs A, const float* __restrict__ mlp) {
  __shared__ __align__(16) float nl[NL_SIZE];
  __shared__ float featbuf[(CB / 64) * 64 * NL_FEAT];
  mlp_stage(mlp, nl);
  __syncthreads();
  const long long stride = (long long)gridDim.x * CB;
  const long long n = A.n;
  for (long long base = (long long)blockIdx.x * CB; base < n; base += stride) {
    const long long t = base + threadIdx.x;
    float f[MLP_IN], y[MLP_OUT];
    for (int i = 0; i < MLP_IN; i++) f[i] = 0.f;
    if (t < n) {
      double u[NP], pr[NP];
      for (int l = 0; l < NP; l++) u[l] = A.U[l*n + t];
      for (int l = 0; l < NP; l++) pr[l] = (A.plain_guess == 0) ? A.prim[l*n + t] : 0.;
      Geom G;
      ks_record_rth(A.a, A.rth[t], A.rth[n + t], &G);
      C2pCell c;
      int rc_plain = 0;
      if (c2p_prepare_learned(u, &G, A.gam, A.gmog, A.plain_guess, pr, c, &rc_plain, f) != C2P_RUNNING)
        for (int i = 0; i < MLP_IN; i++) f[i] = 0.f;
    }
    mlp_forward_wave(nl, f, y, featbuf + (threadIdx.x >> 6) * 64 * NL_FEAT);
    if (t < n) {
      A.code[t] = __float_as_int(y[0]);
      A.iters[t] = __float_as_int(y[1]);
    }
  }
}

// First pass; with MLP the network's outputs come from k_c2p_guess.
template <bool MLP>
__global__ __launch_bounds__(CB) void k_c2p_first(C2pArgs A, int budget, C2pQueue Q, unsigned* count) {
  const long long stride = (long long)gridDim.x * CB;
  for (long long base = (long long)blockIdx.x * CB; base < A.n; base += stride) {
    const long long t = base + threadIdx.x;
    C2pCell c;
    bool park = false;
    if (t < A.n) {
      if constexpr (MLP) {
        const long long n = A.n;
        double u[NP], pr[NP];
        const float y[MLP_OUT] = {__int_as_float(A.code[t]), __int_as_float(A.iters[t])};
        for (int l = 0; l < NP; l++) u[l] = A.U[l*n + t];
        for (int l = 0; l < NP; l++) pr[l] = (A.plain_guess == 0) ? A.prim[l*n + t] : 0.;
        Geom G;
        ks_record_rth(A.a, A.rth[t], A.rth[n + t], &G);
        int rc_plain = 0;
        const int rc = c2p_prepare_learned(u, &G, A.gam, A.gmog, A.plain_guess, pr, c, &rc_plain, nullptr);
        park = c2p_first_zone_finish(A, t, budget, rc, rc_plain, y, pr, c);
      } else {
        park = c2p_first_zone(A, nullptr, t, budget, c);
      }
    }
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
  if (mlp) {
    hipLaunchKernelGGL(k_c2p_guess, dim3(G), dim3(CB), 0, s, A, mlp);
    hipLaunchKernelGGL(k_c2p_first<true>, dim3(G), dim3(CB), 0, s, A, budget_first, W.q[0], W.count);
  } else {
    hipLaunchKernelGGL(k_c2p_first<false>, dim3(G), dim3(CB), 0, s, A, budget_first, W.q[0], W.count);
  }
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
    if (mlp) {
      hipLaunchKernelGGL(k_c2p_guess, dim3(G), dim3(CB), 0, s, A, mlp);
      hipLaunchKernelGGL(k_c2p_first<true>, dim3(G), dim3(CB), 0, s, A, C2P_MAX_ITERS, none, nullptr);
    } else {
      hipLaunchKernelGGL(k_c2p_first<false>, dim3(G), dim3(CB), 0, s, A, C2P_MAX_ITERS, none, nullptr);
    }
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
