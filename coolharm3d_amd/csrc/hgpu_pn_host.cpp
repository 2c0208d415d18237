// hgpu_pn_host.cpp -- host build of the binary-metric entry points (same
// bodies as hgpu_pn.hip, zone by zone; the stream argument is ignored).
#include "hgpu_pn_api.h"

using namespace hgpu;

extern "C" {

int hgpu_pn_geometry(void*, const hgpu_pn_config* cfg, double t, int pos, double* geom) {
  if (int rc = pn_check(cfg)) return rc;
  if (pos < 0 || pos >= NPOS) return set_error(-2, "pos must be 0..4");
  if (!geom) return set_error(-1, "null output");
  const PnBinary B = pn_binary(cfg);
  const PnBlock K = pn_block(cfg);
  const long long n = (long long)cfg->n1 * cfg->n2 * cfg->n3;
  for (long long z = 0; z < n; z++) {
    double X[4];
    pn_zone_x(K, z, pos, X);
    pn_geom(B, t, X[1], X[2], X[3], reinterpret_cast<Geom*>(geom + z * GEOM_DOUBLES));
  }
  return 0;
}

int hgpu_pn_connection(void*, const hgpu_pn_config* cfg, double t, double* conn) {
  if (int rc = pn_check(cfg)) return rc;
  if (!conn) return set_error(-1, "null output");
  const PnBinary B = pn_binary(cfg);
  const PnBlock K = pn_block(cfg);
  const long long n = (long long)cfg->n1 * cfg->n2 * cfg->n3;
  for (long long z = 0; z < n; z++) {
    double X[4];
    pn_zone_x(K, z, CENT, X);
    pn_conn(B, t, X[1], X[2], X[3], conn + z * 64);
  }
  return 0;
}

}  // extern "C"
