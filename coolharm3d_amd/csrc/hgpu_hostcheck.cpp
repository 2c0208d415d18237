// hgpu_hostcheck.cpp -- host build of the kernel bodies (libhgpu_hostcheck.so).
//
// Runs exactly the per-thread bodies the GPU kernels run (hgpu_bodies.h), in
// sequential loops over the same index boxes, behind the same C ABI.  It exists
// so the kernels' arithmetic can be compared with the reference build on a
// machine without a GPU; it is never loaded by the product path (the Python
// wrapper loads libhgpu.so and fails if that is missing).
#include "hgpu_core.h"

namespace hgpu {

static void host_commit(const Params& P, const Ptrs& X, const FailRec& fr, long long c, bool count_zone) {
  const long long NT = P.nt;
  for (int l = 0; l < N_NFAIL; l++) {
    if (!fr.nf[l]) continue;
    if (count_zone && P.track_nfail) X.nfail[l*NT + c] += fr.nf[l];
    X.status[ST_NFAILTOT + l] += (unsigned long long)fr.nf[l];
  }
  if (fr.pflag) {
    X.status[ST_FAILURE_EXISTS] = 1;
    X.status[(fr.pflag == PFLAG_INTERP_PRIM) ? ST_FAILTYPE_PRIM : ST_FAILTYPE_V] = 1;
    if (fr.boundary & 1) X.status[ST_BOUNDARY_PHYS] = 1;
    if (fr.boundary & 2) X.status[ST_BOUNDARY_MPI] = 1;
  }
}

struct HostBE {
  static int init(hgpu_ctx*) { return 0; }
  static void fini(hgpu_ctx*) {}
  static int alloc(hgpu_ctx*, void** p, size_t bytes) {
    *p = calloc(1, bytes ? bytes : 8);
    return *p ? 0 : set_error(-3, "calloc failed");
  }
  static void release(hgpu_ctx*, void* p) { free(p); }
  static int h2d(hgpu_ctx*, void* d, const void* h, size_t bytes) { memcpy(d, h, bytes); return 0; }
  static int d2h(hgpu_ctx*, void* h, const void* d, size_t bytes) { memcpy(h, d, bytes); return 0; }
  static int zero(hgpu_ctx*, void* p, size_t bytes) { memset(p, 0, bytes); return 0; }
  static int sync(hgpu_ctx*) { return 0; }
  static int mark(hgpu_ctx*, const char*, int) { return 0; }
  static int set_graph(hgpu_ctx*, int) { return 0; }   // the host build runs steps directly
  static bool graph_enabled(hgpu_ctx*) { return false; }
  static int invalidate_graph(hgpu_ctx*) { return 0; }
  // launch shapes do not apply to the host build (zones run one by one)
  static int set_shape(hgpu_ctx*, int fam, int, int, int) {   // bs, kc, flags ignored
    return (fam < 0 || fam >= 6) ? set_error(-2, "kernel family out of range") : 0;
  }
  static int get_shape(hgpu_ctx*, int fam, int* bs, int* kc, int* remap) {
    if (fam < 0 || fam >= 6) return set_error(-2, "kernel family out of range");
    if (bs) *bs = 256;
    if (kc) *kc = 0;
    if (remap) *remap = 1;
    return 0;
  }
  static int autotune(hgpu_ctx*, int, double* report) {
    if (report) for (int f = 0; f < 5; f++) { report[f*4] = 256; report[f*4 + 1] = 0; report[f*4 + 2] = 1; report[f*4 + 3] = 0; }
    return 0;
  }
  template <class F>
  static int graph_steps(hgpu_ctx*, int, F&&) { return 0; }
  static int profile_enable(hgpu_ctx*, int) { return 0; }
  static int profile_read(hgpu_ctx*, const char*, double* ms, long long* n) {
    if (ms) *ms = 0.;
    if (n) *n = 0;
    return 0;
  }
  static int reset_nfailtot(hgpu_ctx* c) {
    for (int l = 0; l < N_NFAIL; l++) c->X.status[ST_NFAILTOT + l] = 0;
    return 0;
  }
  static int flux(hgpu_ctx* c, const double* src, int d) {
    Box b = flux_box(c->P, d);
    for (long long t = 0; t < b.count(); t++) {
      int i, j, k;
      b.at(t, i, j, k);
      FailRec fr; fr.clear();
      if (c->P.mixed) {
        if (d == 0) flux_face_mixed<0>(c->P, c->X, src, i, j, k, fr);
        else if (d == 1) flux_face_mixed<1>(c->P, c->X, src, i, j, k, fr);
        else flux_face_mixed<2>(c->P, c->X, src, i, j, k, fr);
      } else {
        if (d == 0) flux_face<0>(c->P, c->X, src, i, j, k, fr);
        else if (d == 1) flux_face<1>(c->P, c->X, src, i, j, k, fr);
        else flux_face<2>(c->P, c->X, src, i, j, k, fr);
      }
      host_commit(c->P, c->X, fr, cidx(c->P, i, j, k), true);
      if (fr.vmax > 0.) {
        unsigned long long& w = c->X.status[ST_VMAX + d + 1];
        if (dbl_bits(fr.vmax) > w) w = dbl_bits(fr.vmax);
      }
    }
    return 0;
  }
  static int emf(hgpu_ctx* c) {
    Box b = edge_box(c->P);
    for (long long t = 0; t < b.count(); t++) {
      int i, j, k;
      b.at(t, i, j, k);
      if (c->P.mixed) emf_edge_mixed(c->P, c->X, i, j, k);
      else emf_edge(c->P, c->X, i, j, k);
    }
    return 0;
  }
  static int recover_queued(hgpu_ctx*, int) { return 0; }   // update() runs the whole tree
  static int update(hgpu_ctx* c, int stage, bool = true) {
    Box b = interior_box(c->P);
    const double dt = c->X.scal[SC_DT];
    double mtot = 0.;   // in the reference's loop order, k fastest (src/step_ch.c:309, :389)
    for (long long t = 0; t < b.count(); t++) {
      int i, j, k;
      b.at(t, i, j, k);
      FailRec fr; fr.clear();
      if (c->P.mixed) mtot += update_zone<true>(c->P, c->X, stage, dt, i, j, k, fr);
      else mtot += update_zone<false>(c->P, c->X, stage, dt, i, j, k, fr);
      host_commit(c->P, c->X, fr, cidx(c->P, i, j, k), true);
    }
    c->X.scal[SC_MTOT] = mtot;
    return 0;
  }
  static int bounds_dir(hgpu_ctx* c, double* S, int dir, int skip_mpi, bool = false) {
    if (skip_mpi && !c->X.status[ST_BOUNDARY_PHYS]) return 0;
    const int do_pflag = c->X.status[ST_FAILURE_EXISTS] ? 1 : 0;
    Box b = ghost_box(c->P, dir);
    const int n1e = c->P.n1 + NG - 1;
    for (long long t = 0; t < b.count(); t++) {
      int sg, a, bb;
      b.at(t, sg, a, bb);
      const int side = sg / NG, g = sg % NG;
      FailRec fr; fr.clear();
      bound_zone(c->P, c->X, S, dir, side, g, a, bb, do_pflag, fr);
      if (dir == 0) host_commit(c->P, c->X, fr, cidx(c->P, side ? (n1e + 1 + g) : (NG - 1 - g), a, bb), true);
    }
    return 0;
  }
  // fixup_interp_v() over every zone, then fixup_interp_prim(), in the
  // reference's order
  static int fixup_interp(hgpu_ctx* c, double* S) {
    if (!c->X.status[ST_FAILURE_EXISTS]) return 0;
    for (int which : {PFLAG_INTERP_V, PFLAG_INTERP_PRIM}) {
      if (!c->X.status[(which == PFLAG_INTERP_PRIM) ? ST_FAILTYPE_PRIM : ST_FAILTYPE_V]) continue;
      Box b = interior_box(c->P);
      for (long long t = 0; t < b.count(); t++) {
        int i, j, k;
        b.at(t, i, j, k);
        FailRec fr; fr.clear();
        fixup_interp_zone(c->P, c->X, S, which, i, j, k, fr);
        host_commit(c->P, c->X, fr, cidx(c->P, i, j, k), true);
      }
    }
    return 0;
  }
  static int set_boundary_flags(hgpu_ctx* c, int phys, int mpi) {
    c->X.status[ST_BOUNDARY_PHYS] = phys ? 1 : 0;
    c->X.status[ST_BOUNDARY_MPI] = mpi ? 1 : 0;
    return 0;
  }
  static int set_vmax(hgpu_ctx* c, const double* vmax4) {
    for (int l = 0; l < 4; l++) c->X.status[ST_VMAX + l] = dbl_bits(vmax4[l]);
    return 0;
  }
  // a box of zones of S (and pflag) to (to_host) or from contiguous host
  // arrays [i][j][k][8] / [i][j][k]
  static int box_copy(hgpu_ctx* c, double* S, const Box& b, double* prims, int* pflag, int to_host) {
    const long long NT = c->P.nt;
    for (long long t = 0; t < b.count(); t++) {
      int i, j, k;
      b.at(t, i, j, k);
      const long long cc = cidx(c->P, i, j, k);
      if (prims) for (int v = 0; v < NP; v++) {
        if (to_host) prims[t*NP + v] = S[v*NT + cc];
        else S[v*NT + cc] = prims[t*NP + v];
      }
      if (pflag) {
        if (to_host) pflag[t] = c->X.pflag[cc];
        else c->X.pflag[cc] = pflag[t];
      }
    }
    return 0;
  }
  // before each stage: flag map, failure words and boundary flags cleared,
  // fastest speeds reset
  static int begin_stage(hgpu_ctx* c) {
    if (c->X.status[ST_FAILURE_EXISTS]) {
      for (long long t = 0; t < c->P.nt; t++) c->X.pflag[t] = 0;
    }
    c->X.status[ST_FAILURE_EXISTS] = 0;
    c->X.status[ST_FAILTYPE_V] = 0;
    c->X.status[ST_FAILTYPE_PRIM] = 0;
    for (int l = 0; l < 4; l++) c->X.status[ST_VMAX + l] = 0;
    c->X.status[ST_BOUNDARY_PHYS] = 0;
    c->X.status[ST_BOUNDARY_MPI] = 0;
    return 0;
  }
  static int timestep(hgpu_ctx* c) { timestep_update(c->P, c->X); return 0; }
  static int fixup_init(hgpu_ctx* c) {
    Box b = interior_box(c->P);
    for (long long t = 0; t < b.count(); t++) {
      int i, j, k;
      b.at(t, i, j, k);
      FailRec fr; fr.clear();
      fixup_init_zone(c->P, c->X, i, j, k, fr);
      host_commit(c->P, c->X, fr, cidx(c->P, i, j, k), true);
    }
    return 0;
  }
  static int divb_max(hgpu_ctx* c, double* out) {
    Box b = divb_box(c->P);
    double m = 0.;
    for (long long t = 0; t < b.count(); t++) {
      int i, j, k;
      b.at(t, i, j, k);
      double v = divb_corner(c->P, c->X, c->X.p, i, j, k);
      if (v > m) m = v;
    }
    *out = m;
    return 0;
  }
  static int fill_geometry(hgpu_ctx* c, Geom* dst, GeomF* dstf, int tsel, unsigned mask64, unsigned mask32) {
    const Params& P = c->P;
    for (int pos = 0; pos < NPOS; pos++) {
      const bool f64 = (mask64 >> pos) & 1u, f32 = ((mask32 >> pos) & 1u) && dstf;
      if (!f64 && !f32) continue;
      for (int i = 0; i < P.n1t; i++)
        for (int j = 0; j < P.n2t; j++)
          for (int k = 0; k < P.n3t; k++) {
            const int id = geom_id(P, i, j, k, pos);
            Geom g;
            pn_zone_geom(P, c->pn, c->orbit[tsel], i, j, k, pos, &g);
            if (f64) dst[id] = g;
            if (f32) pn_zone_geomf(P, c->pn, c->orbit[tsel], i, j, k, pos, &dstf[id]);
          }
    }
    return 0;
  }
  // The host build reads stored levels only (the core never sets P.otf here).
  static constexpr bool HAS_METRIC_OTF = false;
  static int orbit(hgpu_ctx* c) {
    for (int l = 0; l < 3; l++) pn_orbit_at(c->pn.B, level_time(c->X, l), &c->orbit[l]);
    return 0;
  }
  static int fill_conn(hgpu_ctx* c, double* dst, int tsel) {
    const Params& P = c->P;
    for (int i = NG; i < NG + P.n1; i++)
      for (int j = NG; j < NG + P.n2; j++)
        for (int k = NG; k < NG + P.n3; k++) {
          double Xc[4], dg[PN_NDG];
          zone_X(P, c->pn, i, j, k, CENT, Xc);
          pn_metric_dg(c->pn.B, c->orbit[tsel], Xc[1], Xc[2], Xc[3], dg);
          const long long cid = conn_id(P, i, j, k);
          for (int l = 0; l < PN_NDG; l++) dst[l * P.nconn + cid] = dg[l];
        }
    return 0;
  }
  static int con2prim_batch(hgpu_ctx* c, long long n, const double* U, const double* geom, double* prim,
                            double* gamma, int* code, int* iters) {
    const Geom* G = reinterpret_cast<const Geom*>(geom);
    for (long long t = 0; t < n; t++) {
      double u[NP], pr[NP];
      for (int l = 0; l < NP; l++) { u[l] = U[l*n + t]; pr[l] = prim[l*n + t]; }
      double gm = 1., bsq = 0.;
      int it = 0;
      int ret = utoprim_2d_fast(u, &G[t], c->P.gam, c->P.gam_m1_o_gam, pr, &gm, &bsq, &it);
      for (int l = 0; l < NP; l++) prim[l*n + t] = pr[l];
      gamma[t] = gm;
      code[t] = ret;
      if (iters) iters[t] = it;
    }
    return 0;
  }
};

}  // namespace hgpu

using BE = hgpu::HostBE;
#include "hgpu_capi.inc"
