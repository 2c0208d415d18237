// hgpu_pn.hip -- MI355X kernels of the binary-black-hole metric: geometry
// records and connection recomputed per zone at any time (one thread per zone;
// the metric and its derivatives are evaluated in registers, nothing is cached).
#include "hgpu_pn_api.h"
#include <hip/hip_runtime.h>

namespace hgpu {
namespace {

constexpr int PB = 256;

int hipc_pn(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  return set_error(-4, std::string(what) + ": " + hipGetErrorString(e));
}

__global__ __launch_bounds__(PB) void k_pn_geometry(PnBinary B, PnBlock K, double t, int pos, double* geom) {
  const long long n = (long long)K.n1 * K.n2 * K.n3;
  const long long z = (long long)blockIdx.x * PB + threadIdx.x;
  if (z >= n) return;
  double X[4];
  pn_zone_x(K, z, pos, X);
  Geom G;
  pn_geom(B, t, X[1], X[2], X[3], &G);
  const double* src = reinterpret_cast<const double*>(&G);
  double* dst = geom + z * GEOM_DOUBLES;
  for (int l = 0; l < GEOM_DOUBLES; l++) dst[l] = src[l];
}

__global__ __launch_bounds__(PB) void k_pn_connection(PnBinary B, PnBlock K, double t, double* conn) {
  const long long n = (long long)K.n1 * K.n2 * K.n3;
  const long long z = (long long)blockIdx.x * PB + threadIdx.x;
  if (z >= n) return;
  double X[4];
  pn_zone_x(K, z, CENT, X);
  double c[64];
  pn_conn(B, t, X[1], X[2], X[3], c);
  for (int l = 0; l < 64; l++) conn[z * 64 + l] = c[l];
}

}  // namespace
}  // namespace hgpu

using namespace hgpu;

extern "C" {

int hgpu_pn_geometry(void* stream, const hgpu_pn_config* cfg, double t, int pos, double* geom) {
  if (int rc = pn_check(cfg)) return rc;
  if (pos < 0 || pos >= NPOS) return set_error(-2, "pos must be 0..4");
  if (!geom) return set_error(-1, "null output");
  const long long n = (long long)cfg->n1 * cfg->n2 * cfg->n3;
  hipLaunchKernelGGL(k_pn_geometry, dim3((unsigned)((n + PB - 1) / PB)), dim3(PB), 0,
                     static_cast<hipStream_t>(stream), pn_binary(cfg), pn_block(cfg), t, pos, geom);
  return hipc_pn(hipGetLastError(), "k_pn_geometry");
}

int hgpu_pn_connection(void* stream, const hgpu_pn_config* cfg, double t, double* conn) {
  if (int rc = pn_check(cfg)) return rc;
  if (!conn) return set_error(-1, "null output");
  const long long n = (long long)cfg->n1 * cfg->n2 * cfg->n3;
  hipLaunchKernelGGL(k_pn_connection, dim3((unsigned)((n + PB - 1) / PB)), dim3(PB), 0,
                     static_cast<hipStream_t>(stream), pn_binary(cfg), pn_block(cfg), t, conn);
  return hipc_pn(hipGetLastError(), "k_pn_connection");
}

}  // extern "C"
