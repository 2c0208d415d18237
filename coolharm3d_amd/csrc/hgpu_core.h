// hgpu_core.h -- context, set-up and step sequencing shared by the GPU
// library (hgpu_kernels.hip) and the host check build of the kernel bodies
// (hgpu_hostcheck.cpp).  BE is the backend: it owns memory and launches.
#pragma once
#include "../../include/hgpu.h"
#include "hgpu_bodies.h"
#include "hgpu_geom.h"
#include "hgpu_pn.h"
#include <string>
#include <vector>
#include <map>
#include <cstring>
#include <cstdio>
#include <cstdlib>

namespace hgpu {

inline thread_local std::string g_last_error;
inline int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

}  // namespace hgpu

namespace hgpu {
// Time of level tsel of the step: 0 t, 1 t + dt/2, 2 t + dt.
HD double level_time(const Ptrs& X, int tsel) {
  const double t = X.scal[SC_T], dt = X.scal[SC_DT];
  return (tsel == 0) ? t : ((tsel == 1) ? t + 0.5*dt : t + dt);
}
}  // namespace hgpu

struct hgpu_ctx {
  hgpu_config cfg;
  hgpu::Params P;
  hgpu::Ptrs X;
  // time-dependent metric: geometry levels at t, t + dt/2, t + dt, their FP32
  // copies (t, t + dt/2) and connections (t, t + dt/2); level 0 is X.geom's buffer.
  // The binary metric evaluated in the kernels reads no geometry level: its
  // levels are allocated when first needed (lev_alloc).
  hgpu::Geom* lev[3] = {nullptr, nullptr, nullptr};
  hgpu::GeomF* levf[2] = {nullptr, nullptr};
  bool lev_alloc = false;
  long long dev_bytes = 0;         // device memory the context's allocations requested
  long long mpart_cap = 0;         // wavefronts an update launch may have (M_tot parts)
  double* connlev[2] = {nullptr, nullptr};
  hgpu::PnGrid pn;
  hgpu::PnOrbit* orbit = nullptr;  // device: the orbit at levels t, t + dt/2, t + dt
  double* pn_tab = nullptr;        // device: coordinate-function tables of the binary metric
  int otf_stage = 0;               // level of the running stage (binary metric evaluated in kernels)
  long long n_geom = 0, n_conn = 0;
  int n1_conn = 0;
  std::vector<double> stage;       // host staging buffer
  void* be = nullptr;              // backend-private data
  unsigned long long host_status[hgpu::ST_WORDS];
  double host_scal[hgpu::SC_WORDS];
};

namespace hgpu {

// Fills Params from the configuration (set_global_domain_variables(),
// src/decomp.c:20-130, and global_init(), src/main.c:283-300).
inline int make_params(const hgpu_config& c, Params& P) {
  if (c.n1 < 1 || c.n2 < 1 || c.n3 < 1) return set_error(-2, "grid sizes must be >= 1");
  if (c.metric != 0 && c.metric != 2 && c.metric != 3)
    return set_error(-2, "metric must be 0 (Minkowski Cartesian), 2 (Kerr-Schild) or 3 (binary black hole)");
  if (c.metric == 3 && c.coords != 1 && c.coords != 4)
    return set_error(-2, "the binary-black-hole metric needs spherical coordinates (coords = 1 or 4)");
  if (c.metric == 3 && !(c.bbh_m1 >= 0. && c.bbh_m2 >= 0. && c.bbh_m1 + c.bbh_m2 > 0. && c.bbh_sep > 0.))
    return set_error(-2, "binary: masses >= 0 with a positive sum and a positive separation required");
  if (c.coords != 0 && c.coords != 1 && c.coords != 4)
    return set_error(-2, "coords must be 0 (identity), 1 (COORD_DIAGONAL) or 4 (COORD_DIAGONAL3)");
  // set_special_coord()'s check of the exponent (src/init.nz_pn_bbh_disk.c:360)
  if (c.coords == 4 && (c.diag3_exponent < 3 || c.diag3_exponent % 2 == 0))
    return set_error(-2, "COORD_DIAGONAL3 needs an odd diag3_exponent >= 3");
  if (c.recon != RECON_MC && c.recon != RECON_PPM_MC) return set_error(-2, "recon must be 2 (MC) or 5 (PPM+MC)");
  if (c.flux != FLUX_LF && c.flux != FLUX_HLLE) return set_error(-2, "flux must be 0 (LF) or 1 (HLLE)");
  if (c.bc != BC_CARTESIAN_OUTFLOW && c.bc != BC_SPHERICAL_OUTFLOW3) return set_error(-2, "bc must be 0 or 10");
  if (c.ct_para && c.recon != RECON_PPM_MC)
    return set_error(-2, "PPM EMFs (ct_para) use the PPM limiter of the reconstruction: recon must be 5");
  if (c.cooling && c.metric == 0) return set_error(-2, "the h/r cooling function needs a black-hole metric");
  if (c.bc == BC_SPHERICAL_OUTFLOW3 && (c.n2 < NG || c.n3 < NG))
    return set_error(-2, "reflecting/periodic boundaries need at least NG interior zones in x2 and x3");
  if (c.bc_nbr < 0 || c.bc_nbr > 63) return set_error(-2, "bc_nbr holds 6 face bits");
  {
    const int n[3] = {c.n1, c.n2, c.n3};
    for (int d = 0; d < 3; d++)
      if (((c.bc_nbr >> (2*d)) & 3) && n[d] < NG)
        return set_error(-2, "a face shared with another block needs at least NG interior zones across it");
  }
  std::memset(&P, 0, sizeof(P));
  P.n1 = c.n1; P.n2 = c.n2; P.n3 = c.n3;
  P.n1t = c.n1 + 2*NG; P.n2t = c.n2 + 2*NG; P.n3t = c.n3 + 2*NG;
  P.nt = (long long)P.n1t * P.n2t * P.n3t;
  P.metric_dim = (c.metric == 3) ? 3 : ((c.metric == 2) ? 2 : 0);
  P.flat = (c.metric == 0);
  P.dyn = (c.metric == 3);
  P.nconn = (P.metric_dim == 3) ? (long long)c.n1 * c.n2 * c.n3 : ((P.metric_dim == 2) ? (long long)c.n1 * c.n2 : 1);
  P.recon = c.recon;
  P.flux = c.flux;
  P.ct_para = c.ct_para;
  P.cooling = c.cooling;
  P.degen[0] = (c.n1 == 1); P.degen[1] = (c.n2 == 1); P.degen[2] = (c.n3 == 1);
  P.rescale_r = c.rescale_r && !P.degen[0];
  P.bc = c.bc;
  P.bc_nbr = c.bc_nbr;
  P.n_extra_newt = c.no_confirm_iters ? 0 : EXTRA_NEWT_ITER;
  P.floor_dim = (c.coords != 0) ? 1 : 3;   // spherical maps: floors depend on r, i.e. on x1 only
  P.n_spatial_dims = (c.n1 > 1) + (c.n2 > 1) + (c.n3 > 1);
  P.track_nfail = c.track_nfail;
  if (c.precision != 0 && c.precision != 1) return set_error(-2, "precision must be 0 (FP64) or 1 (mixed)");
  P.mixed = c.precision;
  for (int l = 0; l < 4; l++) { P.dx[l] = c.dx[l]; P.invdx[l] = 1./c.dx[l]; }
  P.gam = c.gam;
  P.gam_m1_o_gam = (c.gam - 1.)/c.gam;
  P.cour = c.cour;
  P.t_final = c.t_final;
  P.q_exp = 0.5;
  P.s_norm = 1.;
  return 0;
}

// Coordinate parameters of the host tables.  The binary metric uses the
// spherical map with a non-spinning total mass for its per-x1 tables (cooling
// targets, rescaling, floors); its geometry itself is filled on the device.
inline CoordParams coord_params(const hgpu_config& c) {
  CoordParams C;
  C.metric = (c.metric == 3) ? 2 : c.metric;
  C.coords = c.coords;
  for (int l = 0; l < 4; l++) { C.startx[l] = c.startx[l]; C.dx[l] = c.dx[l]; }
  C.a = (c.metric == 3) ? 0. : c.a;
  C.asq = C.a*C.a; C.M = 1.; C.R0 = c.R0;
  C.h_slope = c.h_slope; C.th_beg = c.th_beg; C.th_length = c.th_length;
  C.diag3_exponent = c.diag3_exponent; C.diag3_factor = c.diag3_factor;
  for (int d = 0; d < 3; d++) C.gpos[d] = c.gpos[d];
  return C;
}

// Host tables (computed with the host libm as the reference computes them).
struct HostTables {
  std::vector<double> geom, conn, omega, tr, rscale, irscale, rhoflr, uuflr;
};

inline void build_tables(const hgpu_config& c, const Params& P, long long n_geom, long long n_conn, HostTables& T) {
  CoordParams C = coord_params(c);
  T.geom.assign(P.dyn ? 0 : n_geom * GEOM_DOUBLES, 0.);
  if (P.dyn) {
    // filled on the device at every step
  } else if (P.metric_dim == 2) {
    for (int pos = 0; pos < NPOS; pos++)
      for (int i = 0; i < P.n1t; i++)
        for (int j = 0; j < P.n2t; j++)
          geom_record(C, i, j, NG, pos, reinterpret_cast<Geom*>(&T.geom[(long long)geom_id(P, i, j, NG, pos) * GEOM_DOUBLES]));
  } else {
    geom_record(C, NG, NG, NG, CENT, reinterpret_cast<Geom*>(&T.geom[0]));
  }
  T.conn.assign(P.dyn ? 0 : n_conn * 64, 0.);
  if (P.metric_dim == 2 && !P.flat) {
    for (int i = NG; i < NG + P.n1; i++)
      for (int j = NG; j < NG + P.n2; j++)
        ks_connection(C, i, j, &T.conn[(long long)conn_id(P, i, j, NG) * 64]);
  }
  // per-x1 tables
  T.omega.assign(P.n1t, 0.); T.tr.assign(P.n1t, 0.);
  // ones without RESCALE_R: the kernels multiply by these tables unconditionally
  T.rscale.assign(6 * P.n1t, 1.); T.irscale.assign(6 * P.n1t, 1.);
  std::vector<double> rc(P.n1t), rf(P.n1t);
  for (int i = 0; i < P.n1t; i++) {
    double xp[4], x[4], J[4];
    coord_xp(C, i, NG, NG, CENT, xp); x_of_xp(C, xp, x, J); rc[i] = x[1];
    coord_xp(C, i, NG, NG, FACE1, xp); x_of_xp(C, xp, x, J); rf[i] = x[1];
  }
  if (c.cooling) {
    CoolTables CT;
    cool_setup(C, c.r_isco, c.h_o_r, &CT);
    for (int i = 0; i < P.n1t; i++) {
      T.omega[i] = omega_circular_equatorial(C, CT, rc[i]);
      T.tr[i] = target_temperature(C, CT, rc[i]);
    }
  }
  if (P.rescale_r) {
    // rescale_prims_r() first call, src/recon.c:1141-1167
    for (int i = 0; i < P.n1t; i++) {
      for (int pz = 0; pz < 2; pz++) {
        double rloc = pz ? rf[i] : rc[i];
        double rf2 = rloc*rloc, rf0 = sqrt(rf2 * rloc), rf1 = rf2*rloc;
        T.rscale[(3*pz + 0)*P.n1t + i] = rf0;
        T.rscale[(3*pz + 1)*P.n1t + i] = rf1;
        T.rscale[(3*pz + 2)*P.n1t + i] = rloc;
      }
      for (int s = 0; s < 6; s++) T.irscale[s*P.n1t + i] = 1/T.rscale[s*P.n1t + i];
    }
  }
  // floors: coord_of_xp(), src/coord.c:3790-3797 (r = x^1 for spherical
  // topology, |x| for Cartesian)
  if (P.floor_dim == 1) {
    T.rhoflr.resize(P.n1t); T.uuflr.resize(P.n1t);
    for (int i = 0; i < P.n1t; i++) {
      T.rhoflr[i] = c.rhomin*pow(rc[i], c.rhopower);
      T.uuflr[i] = c.uumin*pow(rc[i], c.uupower);
    }
  } else {
    T.rhoflr.resize(P.nt); T.uuflr.resize(P.nt);
    for (int i = 0; i < P.n1t; i++) for (int j = 0; j < P.n2t; j++) for (int k = 0; k < P.n3t; k++) {
      double xp[4], x[4], J[4];
      coord_xp(C, i, j, k, CENT, xp); x_of_xp(C, xp, x, J);
      double r = sqrt(x[1]*x[1] + x[2]*x[2] + x[3]*x[3]);
      long long cc = cidx(P, i, j, k);
      T.rhoflr[cc] = c.rhomin*pow(r, c.rhopower);
      T.uuflr[cc] = c.uumin*pow(r, c.uupower);
    }
  }
}

// FP32 copies of geometry records (mixed precision)
inline void geom_to_f32(const double* g42, long long n, std::vector<GeomF>& out) {
  out.resize(n);
  for (long long r = 0; r < n; r++) {
    const Geom* g = reinterpret_cast<const Geom*>(g42 + r * GEOM_DOUBLES);
    GeomF& f = out[r];
    for (int a = 0; a < 4; a++) for (int b = 0; b < 4; b++) {
      f.gcon[a][b] = (float)g->gcon[a][b];
      f.gcov[a][b] = (float)g->gcov[a][b];
    }
    f.g = (float)g->g; f.g_inv = (float)g->g_inv; f.alpha = (float)g->alpha;
    for (int a = 0; a < 3; a++) f.beta[a] = (float)g->beta[a];
    for (int a = 0; a < 4; a++) f.ncon[a] = (float)g->ncon[a];
    f.om_alpha = (float)(1. - g->alpha);
  }
}

// AoS [zone][8] <-> SoA [8][zone]
inline void aos_to_soa(const double* aos, double* soa, long long nt, int nv) {
  for (long long c = 0; c < nt; c++) for (int v = 0; v < nv; v++) soa[v*nt + c] = aos[c*nv + v];
}
inline void soa_to_aos(const double* soa, double* aos, long long nt, int nv) {
  for (long long c = 0; c < nt; c++) for (int v = 0; v < nv; v++) aos[c*nv + v] = soa[v*nt + c];
}

template <class BE>
struct Api {
  static int dalloc(hgpu_ctx* c, void** p, size_t bytes) {
    const int rc = BE::alloc(c, p, bytes);
    if (!rc) c->dev_bytes += (long long)bytes;
    return rc;
  }
  // Geometry records in memory: the static metric's cache (level 0, FP32 copy
  // in mixed precision), or the binary metric's three levels and two FP32 face
  // copies.  A binary-metric context evaluating the metric in its kernels
  // allocates them here when it first needs them (the stored mode, a geometry
  // download); the level pointers of the current state are set again.
  static int alloc_levels(hgpu_ctx* c) {
    if (c->lev_alloc) return 0;
    const Params& P = c->P;
    const int nl = P.dyn ? 3 : 1, nlf = P.dyn ? 2 : 1;
    int rc = 0;
    for (int l = 0; l < nl; l++) rc |= dalloc(c, (void**)&c->lev[l], sizeof(double) * GEOM_DOUBLES * c->n_geom);
    for (int l = 0; l < nlf; l++) rc |= dalloc(c, (void**)&c->levf[l], sizeof(GeomF) * (P.mixed ? c->n_geom : 1));
    if (rc) return set_error(-3, "device allocation of the geometry records failed: " + g_last_error);
    c->lev_alloc = true;
    Ptrs& X = c->X;
    const int l = P.dyn ? X.cent_lvl : 0;
    X.geom = X.geom_n = X.geom_new = c->lev[l];
    X.geomf = c->levf[l < nlf ? l : nlf - 1];
    return 0;
  }
  static int create(const hgpu_config* cfg, hgpu_ctx** out) {
    if (!cfg || !out) return set_error(-1, "null argument");
    hgpu_ctx* c = new hgpu_ctx();
    c->cfg = *cfg;
    int rc = make_params(*cfg, c->P);
    if (rc) { delete c; return rc; }
    const Params& P = c->P;
    c->n_geom = (P.metric_dim == 3) ? (long long)NPOS * P.nt : ((P.metric_dim == 2) ? (long long)NPOS * P.n1t * P.n2t : 1);
    c->n_conn = (P.metric_dim == 3) ? (long long)P.n1 * P.n2 * P.n3 : ((P.metric_dim == 2) ? (long long)P.n1 * P.n2 : 1);
    if (P.dyn) {
      PnBinary& B = c->pn.B;
      B.m1 = cfg->bbh_m1; B.m2 = cfg->bbh_m2; B.b0 = cfg->bbh_sep; B.phi0 = cfg->bbh_phi0;
      B.inspiral = cfg->bbh_inspiral;
      B.R0 = cfg->R0; B.th_beg = cfg->th_beg; B.th_length = cfg->th_length; B.h_slope = cfg->h_slope;
      B.coords = cfg->coords; B.diag3_exponent = cfg->diag3_exponent; B.diag3_factor = cfg->diag3_factor;
      for (int l = 0; l < 4; l++) c->pn.x0[l] = cfg->startx[l];
      for (int d = 0; d < 3; d++) c->pn.gpos[d] = cfg->gpos[d];
    }
    rc = BE::init(c);
    if (rc) { delete c; return rc; }
    c->P.otf = metric_otf_possible(c) ? 1 : 0;
    HostTables T;
    build_tables(*cfg, P, c->n_geom, c->n_conn, T);
    Ptrs& X = c->X;
    const long long NT = P.nt;
    rc = 0;
    rc |= dalloc(c, (void**)&X.p, sizeof(double) * NP * NT);
    rc |= dalloc(c, (void**)&X.ph, sizeof(double) * NP * NT);
    const long long nF = P.mixed ? 1 : 3 * NFLUX * NT, nE = P.mixed ? 1 : 3 * NT;
    rc |= dalloc(c, (void**)&X.F, sizeof(double) * nF);
    rc |= dalloc(c, (void**)&X.emf, sizeof(double) * nE);
    rc |= dalloc(c, (void**)&X.Ff, sizeof(float) * (P.mixed ? 3 * NFLUX * NT : 2));
    rc |= dalloc(c, (void**)&X.emff, sizeof(float) * (P.mixed ? 3 * NT : 2));
    // connection records [zone][64] of a static metric; derivative levels
    // [PN_NDG][zone] of the binary metric
    const long long nconn_rec = P.dyn ? PN_NDG : 64;
    rc |= dalloc(c, (void**)&c->connlev[0], sizeof(double) * nconn_rec * c->n_conn);
    X.conn = c->connlev[0];
    if (P.dyn) {
      rc |= dalloc(c, (void**)&c->connlev[1], sizeof(double) * nconn_rec * c->n_conn);
      rc |= dalloc(c, (void**)&c->orbit, sizeof(PnOrbit) * 3);
      X.orbit = c->orbit;
      rc |= dalloc(c, (void**)&c->pn_tab, sizeof(double) * pn_tab_size(P));
    }
    rc |= dalloc(c, (void**)&X.omega, sizeof(double) * P.n1t);
    rc |= dalloc(c, (void**)&X.tr, sizeof(double) * P.n1t);
    rc |= dalloc(c, (void**)&X.rscale, sizeof(double) * 6 * P.n1t);
    rc |= dalloc(c, (void**)&X.irscale, sizeof(double) * 6 * P.n1t);
    rc |= dalloc(c, (void**)&X.rhoflr, sizeof(double) * T.rhoflr.size());
    rc |= dalloc(c, (void**)&X.uuflr, sizeof(double) * T.uuflr.size());
    rc |= dalloc(c, (void**)&X.gamma, sizeof(double) * NT);
    const long long nint = (long long)P.n1 * P.n2 * P.n3;
    rc |= dalloc(c, (void**)&X.Uslow, sizeof(double) * NP * NT);
    // U^n for the corrector (the mixed step on a static metric forms it again)
    rc |= dalloc(c, (void**)&X.Un, sizeof(double) * ((P.mixed && !P.dyn) ? 1 : NP * NT));
    rc |= dalloc(c, (void**)&X.slow_idx, sizeof(unsigned) * nint);
    rc |= dalloc(c, (void**)&X.pflag, sizeof(int) * NT);
    rc |= dalloc(c, (void**)&X.nfail, sizeof(int) * (P.track_nfail ? N_NFAIL * NT : 1));
    rc |= dalloc(c, (void**)&X.status, sizeof(unsigned long long) * ST_WORDS);
    rc |= dalloc(c, (void**)&X.scal, sizeof(double) * SC_WORDS);
    // at most n1 (n2 n3 / 64 + n2 + n3 + 2) + 8 workgroups of at most 4
    // wavefronts over the interior, for any tile of 64 to 256 threads
    c->mpart_cap = 4 * ((long long)P.n1 * ((long long)P.n2 * P.n3 / 64 + P.n2 + P.n3 + 2) + 8);
    rc |= dalloc(c, (void**)&X.mpart, sizeof(double) * c->mpart_cap);
    rc |= dalloc(c, (void**)&X.mpart2, sizeof(double) * 1024);
    if (rc) { destroy(c); return set_error(-3, "device allocation failed: " + g_last_error); }
    if (!c->P.otf) rc |= alloc_levels(c);
    if (rc) { destroy(c); return rc; }
    if (!T.geom.empty()) rc |= BE::h2d(c, (void*)X.geom, T.geom.data(), sizeof(double) * T.geom.size());
    if (P.dyn) {
      // coordinate functions along each direction, evaluated once here and
      // read by every record evaluation of the binary metric
      std::vector<double> tab(pn_tab_size(P));
      PnGrid G0 = c->pn;
      for (long long e = 0; e < pn_tab_entries(P); e++) pn_tab_entry(P, G0, tab.data(), e);
      rc |= BE::h2d(c, c->pn_tab, tab.data(), sizeof(double) * tab.size());
      c->pn.tab = c->pn_tab;
      X.pn = c->pn;
    }
    rc |= BE::h2d(c, (void*)X.conn, T.conn.data(), sizeof(double) * T.conn.size());
    rc |= BE::h2d(c, (void*)X.omega, T.omega.data(), sizeof(double) * T.omega.size());
    rc |= BE::h2d(c, (void*)X.tr, T.tr.data(), sizeof(double) * T.tr.size());
    rc |= BE::h2d(c, (void*)X.rscale, T.rscale.data(), sizeof(double) * T.rscale.size());
    rc |= BE::h2d(c, (void*)X.irscale, T.irscale.data(), sizeof(double) * T.irscale.size());
    rc |= BE::h2d(c, (void*)X.rhoflr, T.rhoflr.data(), sizeof(double) * T.rhoflr.size());
    rc |= BE::h2d(c, (void*)X.uuflr, T.uuflr.data(), sizeof(double) * T.uuflr.size());
    rc |= BE::zero(c, X.p, sizeof(double) * NP * NT);
    rc |= BE::zero(c, X.ph, sizeof(double) * NP * NT);
    rc |= BE::zero(c, X.F, sizeof(double) * nF);
    rc |= BE::zero(c, X.emf, sizeof(double) * nE);
    if (P.mixed) {
      rc |= BE::zero(c, X.Ff, sizeof(float) * 3 * NFLUX * NT);
      rc |= BE::zero(c, X.emff, sizeof(float) * 3 * NT);
      if (!P.dyn) rc |= upload_geomf(c, T.geom.data());
    }
    rc |= BE::zero(c, X.gamma, sizeof(double) * NT);
    rc |= BE::zero(c, X.pflag, sizeof(int) * NT);
    rc |= BE::zero(c, X.nfail, sizeof(int) * (P.track_nfail ? N_NFAIL * NT : 1));
    rc |= BE::zero(c, X.status, sizeof(unsigned long long) * ST_WORDS);
    double sc[SC_WORDS] = {0};
    sc[SC_T] = cfg->startx[0];
    sc[SC_DT] = cfg->dx[0];
    rc |= BE::h2d(c, X.scal, sc, sizeof(sc));
    if (P.dyn) rc |= set_levels_at_state(c);
    rc |= BE::sync(c);
    if (rc) { destroy(c); return set_error(-3, "device initialisation failed: " + g_last_error); }
    *out = c;
    return 0;
  }

  static int destroy(hgpu_ctx* c) {
    if (!c) return 0;
    Ptrs& X = c->X;
    // X.geom, X.geomf, X.conn point into the level buffers (level 0 is the
    // buffer allocated for them), so the levels are what is released
    void* ptrs[] = {X.p, X.ph, X.F, X.emf, (void*)X.omega, (void*)X.tr,
                    (void*)X.rscale, (void*)X.irscale, (void*)X.rhoflr, (void*)X.uuflr, X.gamma,
                    X.pflag, X.nfail, X.status, X.scal, X.mpart, X.mpart2, X.Ff, X.emff, X.Uslow, X.Un,
                    X.slow_idx,
                    c->lev[0], c->lev[1], c->lev[2], c->levf[0], c->levf[1], c->connlev[0], c->connlev[1],
                    (void*)c->orbit, c->pn_tab};
    for (void* p : ptrs) if (p) BE::release(c, p);
    BE::fini(c);
    delete c;
    return 0;
  }

  // FP32 geometry copy for the mixed-precision face fluxes
  static int upload_geomf(hgpu_ctx* c, const double* geom_host) {
    if (!c->P.mixed) return 0;
    std::vector<GeomF> gf;
    geom_to_f32(geom_host, c->n_geom, gf);
    return BE::h2d(c, (void*)c->X.geomf, gf.data(), sizeof(GeomF) * gf.size());
  }

  static int upload_prims(hgpu_ctx* c, int which, const double* host) {
    if (!c || !host) return set_error(-1, "null argument");
    const long long NT = c->P.nt;
    c->stage.resize(NP * NT);
    aos_to_soa(host, c->stage.data(), NT, NP);
    double* dst = which ? c->X.ph : c->X.p;
    int rc = BE::h2d(c, dst, c->stage.data(), sizeof(double) * NP * NT);
    return rc ? rc : BE::sync(c);
  }

  static int download_prims(hgpu_ctx* c, int which, double* host) {
    if (!c || !host) return set_error(-1, "null argument");
    const long long NT = c->P.nt;
    c->stage.resize(NP * NT);
    const double* src = which ? c->X.ph : c->X.p;
    int rc = BE::d2h(c, c->stage.data(), src, sizeof(double) * NP * NT);
    if (rc) return rc;
    soa_to_aos(c->stage.data(), host, NT, NP);
    return 0;
  }

  static int set_time(hgpu_ctx* c, double t, double dt) {
    double sc[SC_WORDS] = {0};
    sc[SC_T] = t; sc[SC_DT] = dt; sc[SC_T_OLD] = t; sc[SC_DT_OLD] = dt;
    int rc = BE::h2d(c, c->X.scal, sc, sizeof(sc));
    return rc ? rc : BE::sync(c);
  }

  static int get_time(hgpu_ctx* c, double* t, double* dt, double* dt_old) {
    double sc[SC_WORDS];
    int rc = BE::d2h(c, sc, c->X.scal, sizeof(sc));
    if (rc) return rc;
    if (t) *t = sc[SC_T];
    if (dt) *dt = sc[SC_DT];
    if (dt_old) *dt_old = sc[SC_DT_OLD];
    return 0;
  }
  // M_tot of the last stage (src/step_ch.c:309, :389)
  static int get_m_tot(hgpu_ctx* c, double* m_tot) {
    double sc[SC_WORDS];
    const int rc = BE::d2h(c, sc, c->X.scal, sizeof(sc));
    if (!rc) *m_tot = sc[SC_MTOT];
    return rc;
  }

  // bounds(prim, skip_mpi): skip_mpi = 0 always applies the physical BCs;
  // skip_mpi = 1 only if boundary_phys_pflag (src/bounds.c:56)
  // after_update: the first bounds after an update (it completes M_tot)
  static int bounds(hgpu_ctx* c, double* S, int skip_mpi, bool after_update = false) {
    int rc = 0;
    for (int dir = 0; dir < 3; dir++) rc |= BE::bounds_dir(c, S, dir, skip_mpi, after_update && dir == 0);
    return rc;
  }

  // One advance() stage plus its boundaries and fixups (src/step_ch.c:92-121,
  // :159-180).
  // Time-dependent metric: the state p (time t of the coming step) is read with
  // level 0; fixup_bounds() and the diagnostics use it.
  static int set_levels_at_state(hgpu_ctx* c) {
    Ptrs& X = c->X;
    int rc = BE::orbit(c);
    if (!c->P.otf) rc |= BE::fill_geometry(c, c->lev[0], c->levf[0], 0, 1u << CENT, 0u);
    X.geom = X.geom_n = X.geom_new = c->lev[0];
    X.cent_lvl = 0;
    X.geomf = c->levf[0];
    X.conn = c->connlev[0];
    return rc;
  }

  // The stage itself; BE::begin_stage() has cleared its failure words, flag
  // map, boundary flags and fastest speeds and emptied the slow-zone queue.
  static int substep(hgpu_ctx* c, int stage) {
    Ptrs& X = c->X;
    double* src = (stage == 0) ? X.p : X.ph;
    double* dst = (stage == 0) ? X.ph : X.p;
    int rc = 0;
    if (c->P.dyn) {
      // fluxes and sources at the stage's time, U^n at t, the result at the
      // next level (ncurr = n_beg / n_mid / n_end, src/step_ch.c:74-162, :319-331)
      X.geom = c->lev[stage];
      X.geomf = c->levf[stage];
      X.conn = c->connlev[stage];
      X.geom_n = c->lev[0];
      X.geom_new = c->lev[stage + 1];
      X.geom_s = c->lev[stage];
      X.cent_lvl = stage;
      c->otf_stage = stage;
    }
    for (int d = 0; d < 3; d++) rc |= BE::flux(c, src, d);
    rc |= BE::emf(c);
    rc |= BE::update(c, stage, false);   // begin_stage() emptied the slow-zone queue
    if (c->P.dyn) { X.geom = X.geom_new; X.cent_lvl = stage + 1; }
    rc |= BE::recover_queued(c, stage);
    rc |= bounds(c, dst, 0, true);
    return rc;
  }

  // The rest of the stage after bounds(prim,0) (src/step_ch.c:105-114,
  // :171-180): the interpolating fixups, bounds(prim,1) where the boundary
  // flags ask for it, the flags cleared.
  static int stage_fixup(hgpu_ctx* c, int stage) {
    double* dst = (stage == 0) ? c->X.ph : c->X.p;
    int rc = BE::fixup_interp(c, dst);   // fixup_interp_v(), then fixup_interp_prim()
    rc |= bounds(c, dst, 1);   // the boundary flags are cleared when the next stage starts
    return rc;
  }

  // Geometry and connection of a step's three time levels for the
  // time-dependent metric (set_general_geometry()/set_general_conn(),
  // src/metric.c:3390-3647): levels t and t + dt/2 at all positions with their
  // connections, t + dt at zone centres.
  // Only what is read is filled: face records in FP32 for the mixed-precision
  // fluxes (FP64 otherwise), centre records in FP64; no corner records.
  // The binary metric can be evaluated inside the flux and update kernels:
  // the backend has those kernels, the reconstruction is PPM (the
  // several-faces-per-thread kernels) and no direction is degenerate.
  static bool metric_otf_possible(const hgpu_ctx* c) {
    const Params& P = c->P;
    return BE::HAS_METRIC_OTF && P.dyn && P.recon == RECON_PPM_MC && !P.degen[0] && !P.degen[1] && !P.degen[2];
  }
  static int set_metric_otf(hgpu_ctx* c, int on) {
    if (on && !metric_otf_possible(c))
      return set_error(-2, "the metric is evaluated in the kernels only for the binary metric with PPM "
                           "reconstruction, no degenerate direction, on the GPU build");
    int rc = BE::invalidate_graph(c);
    const bool was = c->P.otf != 0;
    if (!on) rc |= alloc_levels(c);
    if (rc) return rc;
    c->P.otf = on ? 1 : 0;
    // leaving the evaluated mode: the levels of the running step must exist in
    // memory (the switch may come between its predictor and corrector), the
    // state's centre level among them
    if (was && !on && c->P.dyn) rc |= fill_levels(c);
    return rc;
  }
  // Centre records of the state's level in memory (the evaluated mode keeps
  // none): for hgpu_download_geometry.
  static int materialize_state_geometry(hgpu_ctx* c) {
    if (!c->P.dyn || !c->P.otf) return 0;
    const int rc = alloc_levels(c);
    if (rc) return rc;
    return BE::fill_geometry(c, const_cast<Geom*>(c->X.geom), nullptr, c->X.cent_lvl, 1u << CENT, 0u);
  }
  // Metric derivatives of the last stage's level (the evaluated mode keeps the
  // source terms there instead): for hgpu_download_geometry.
  static int materialize_stage_derivatives(hgpu_ctx* c) {
    if (!c->P.dyn || !c->P.otf) return 0;
    const int lvl = (c->X.conn == c->connlev[1]) ? 1 : 0;
    return BE::fill_conn(c, const_cast<double*>(c->X.conn), lvl);
  }

  static int fill_levels(hgpu_ctx* c) {
    const unsigned faces = (1u << FACE1) | (1u << FACE2) | (1u << FACE3), cent = 1u << CENT;
    const unsigned m64 = c->P.mixed ? cent : (faces | cent), m32 = c->P.mixed ? faces : 0u;
    int rc = BE::orbit(c);   // the holes at t, t + dt/2, t + dt, read by every fill
    // every kernel evaluates the face and zone records it needs from the
    // orbit; each stage's update forms its source terms first (the stage's
    // connection buffer holds the 4 terms instead of 28 derivatives)
    if (c->P.otf) return rc;
    rc |= BE::fill_geometry(c, c->lev[0], c->levf[0], 0, m64, m32);
    rc |= BE::fill_conn(c, c->connlev[0], 0);
    rc |= BE::fill_geometry(c, c->lev[1], c->levf[1], 1, m64, m32);
    rc |= BE::fill_conn(c, c->connlev[1], 1);
    rc |= BE::fill_geometry(c, c->lev[2], nullptr, 2, cent, 0u);
    return rc;
  }

  // One advance() stage and bounds(prim,0)'s physical part, through the
  // public entry point: the step's geometry levels first (stage 0, binary
  // metric), the flag map and failure words cleared as step_ch() does before
  // each stage (src/step_ch.c:65-69, :117-121), with the stage's fastest speeds.
  static int advance_api(hgpu_ctx* c, int stage) {
    int rc = 0;
    if (c->P.dyn && stage == 0) rc |= fill_levels(c);
    rc |= BE::begin_stage(c);
    rc |= substep(c, stage);
    return rc;
  }
  // The stage's fixups and bounds(prim,1); set: the boundary flags take the
  // caller's (globally reduced) values first.
  static int fixup_api(hgpu_ctx* c, int stage, bool set, int phys, int mpi) {
    int rc = 0;
    if (set) rc |= BE::set_boundary_flags(c, phys, mpi);
    rc |= stage_fixup(c, stage);
    if (c->P.dyn && stage == 1) { c->X.geom = c->X.geom_n = c->X.geom_new = c->lev[2]; c->X.cent_lvl = 2; }
    return rc;
  }
  static int substep_api(hgpu_ctx* c, int stage) {
    int rc = advance_api(c, stage);
    rc |= fixup_api(c, stage, false, 0, 0);
    return rc;
  }
  // t += dt and the next dt (src/step_ch.c:213-223) from the given speeds
  // max_char_speed[0..3] or, without them, the block's own.
  static int end_step(hgpu_ctx* c, const double* vmax4) {
    int rc = 0;
    if (vmax4) rc |= BE::set_vmax(c, vmax4);
    rc |= BE::timestep(c);
    return rc;
  }
  static int set_dt(hgpu_ctx* c, double dt) {
    double sc[SC_WORDS];
    int rc = BE::d2h(c, sc, c->X.scal, sizeof(sc));
    if (rc) return rc;
    sc[SC_DT] = dt;
    rc = BE::h2d(c, c->X.scal, sc, sizeof(sc));
    return rc ? rc : BE::sync(c);
  }
  static int box(hgpu_ctx* c, int which, const int* lo, const int* hi, double* prims, int* pflag, int to_host) {
    if (!c || !lo || !hi) return set_error(-1, "null argument");
    if (which != 0 && which != 1) return set_error(-2, "which must be 0 (p) or 1 (ph)");
    const int nt[3] = {c->P.n1t, c->P.n2t, c->P.n3t};
    Box b;
    for (int d = 0; d < 3; d++) {
      if (lo[d] < 0 || hi[d] >= nt[d] || lo[d] > hi[d]) return set_error(-2, "box outside the grid or empty");
      b.lo[d] = lo[d];
      b.hi[d] = hi[d];
    }
    return BE::box_copy(c, which ? c->X.ph : c->X.p, b, prims, pflag, to_host);
  }

  // One step_ch(): predictor, corrector, new time step (src/step_ch.c:50-234).
  static int one_step(hgpu_ctx* c) {
    int rc = 0;
    if (c->P.dyn) rc |= fill_levels(c);
    rc |= BE::begin_stage(c);         // src/step_ch.c:65-69
    rc |= substep(c, 0);
    rc |= stage_fixup(c, 0);
    rc |= BE::begin_stage(c);         // src/step_ch.c:117-121
    rc |= substep(c, 1);
    rc |= stage_fixup(c, 1);
    rc |= BE::timestep(c);            // src/step_ch.c:218-223
    if (c->P.dyn) {
      // the state is now at t + dt: level 2 (centres) describes it
      c->X.geom = c->X.geom_n = c->X.geom_new = c->lev[2];
      c->X.cent_lvl = 2;
    }
    return rc;
  }

  static int step(hgpu_ctx* c, int nsteps) {
    if (!c) return set_error(-1, "null context");
    if (nsteps < 0) return set_error(-2, "nsteps must be >= 0");
    if (BE::graph_enabled(c)) return BE::graph_steps(c, nsteps, [c]() { return one_step(c); });
    int rc = 0;
    for (int s = 0; s < nsteps && !rc; s++) {
      rc |= BE::mark(c, "step", 0);
      rc |= one_step(c);
      rc |= BE::mark(c, "step", 1);
    }
    return rc;
  }

  static int fixup_bounds(hgpu_ctx* c) {
    int rc = 0;
    if (c->P.dyn) rc |= set_levels_at_state(c);
    rc |= BE::fixup_init(c);
    rc |= bounds(c, c->X.p, 0);
    rc |= BE::sync(c);
    return rc;
  }
};

}  // namespace hgpu
