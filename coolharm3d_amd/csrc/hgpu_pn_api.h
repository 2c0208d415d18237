// hgpu_pn_api.h -- argument checks and per-zone bodies of the binary-metric
// entry points, shared by the HIP library and the host build.
#pragma once
#include "hgpu_core.h"
#include "hgpu_pn.h"

namespace hgpu {

inline int pn_check(const hgpu_pn_config* c) {
  if (!c) return set_error(-1, "null configuration");
  if (!(c->m1 >= 0. && c->m2 >= 0. && c->m1 + c->m2 > 0.)) return set_error(-2, "masses must be >= 0 with a positive sum");
  if (!(c->b0 > 0.)) return set_error(-2, "separation b0 must be > 0");
  if (c->n1 < 1 || c->n2 < 1 || c->n3 < 1) return set_error(-2, "block sizes must be >= 1");
  if ((long long)c->n1 * c->n2 * c->n3 > (1LL << 31) - 1) return set_error(-2, "block larger than 2^31 zones");
  if (c->coords != 0 && c->coords != 1 && c->coords != 4)
    return set_error(-2, "coords must be 0 or 1 (COORD_DIAGONAL) or 4 (COORD_DIAGONAL3)");
  if (c->coords == 4 && (c->diag3_exponent < 3 || c->diag3_exponent % 2 == 0))
    return set_error(-2, "COORD_DIAGONAL3 needs an odd diag3_exponent >= 3");
  return 0;
}

inline PnBinary pn_binary(const hgpu_pn_config* c) {
  PnBinary B;
  B.m1 = c->m1; B.m2 = c->m2; B.b0 = c->b0; B.phi0 = c->phi0; B.inspiral = c->inspiral;
  B.R0 = c->R0; B.th_beg = c->th_beg; B.th_length = c->th_length; B.h_slope = c->h_slope;
  B.coords = (c->coords == 4) ? 4 : 1;
  B.diag3_exponent = c->diag3_exponent; B.diag3_factor = c->diag3_factor;
  return B;
}

struct PnBlock {
  int n1, n2, n3;
  double x0[4], dx[4];
};

inline PnBlock pn_block(const hgpu_pn_config* c) {
  PnBlock K;
  K.n1 = c->n1; K.n2 = c->n2; K.n3 = c->n3;
  for (int l = 0; l < 4; l++) { K.x0[l] = c->startx[l]; K.dx[l] = c->dx[l]; }
  return K;
}

// Numerical coordinates of zone z at position pos (coord(), src/coord.c:182-228).
HD void pn_zone_x(const PnBlock& K, long long z, int pos, double* X) {
  const int k = (int)(z % K.n3);
  const int j = (int)((z / K.n3) % K.n2);
  const int i = (int)(z / ((long long)K.n3 * K.n2));
  const double h1 = (pos == FACE1 || pos == CORN) ? 0. : 0.5;
  const double h2 = (pos == FACE2 || pos == CORN) ? 0. : 0.5;
  const double h3 = (pos == FACE3 || pos == CORN) ? 0. : 0.5;
  X[1] = K.x0[1] + (i + h1) * K.dx[1];
  X[2] = K.x0[2] + (j + h2) * K.dx[2];
  X[3] = K.x0[3] + (k + h3) * K.dx[3];
}

}  // namespace hgpu
