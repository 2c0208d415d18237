// hgpu_types.h -- data layout of the per-substep GRMHD advance on one GPU.
//
// State in HBM is structure-of-arrays: variable v of cell (i,j,k) lives at
// base[v*NT + (i*N2T + j)*N3T + k] with k fastest, ghosts included (NG = 3
// layers per side, the reference's src/macros.h:41).  Variables follow the
// reference's mnemonics (src/macros.h:154-161): rho, u, u~1..3, B1..3.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HD __host__ __device__ __forceinline__

namespace hgpu {

constexpr int NG = 3;
constexpr int NP = 8;
constexpr int NDIM = 4;
constexpr int NFLUX = 8;      // conserved fluxes per face
constexpr int N_NFAIL = 15;   // src/macros.h:236-251
constexpr int GEOM_DOUBLES = 42;

enum { RHO = 0, UU = 1, U1 = 2, U2 = 3, U3 = 4, B1 = 5, B2 = 6, B3 = 7 };
enum { FACE1 = 0, FACE2 = 1, FACE3 = 2, CENT = 3, CORN = 4, NPOS = 5 };
enum { PFLAG_NONE = 0, PFLAG_INTERP_V = 1, PFLAG_INTERP_PRIM = 2 };
enum {
  NFAIL_FLOOR = 0, NFAIL_TMAX = 1, NFAIL_GAMMA_MAX = 2, NFAIL_GAMMA_CALC = 3,
  NFAIL_USE_FULL_INV = 4, NFAIL_USE_EE = 5, NFAIL_UTOPRIM = 6, NFAIL_UTOPRIM_EE = 7,
  NFAIL_USE_INTERP_PRIM = 8, NFAIL_USE_INTERP_V = 9, NFAIL_INTERP_PRIM = 10,
  NFAIL_INTERP_V = 11
};

// Reconstruction / flux / CT / BC choices (values mirror src/options.h).
enum { RECON_MC = 2, RECON_PPM_MC = 5 };
enum { FLUX_LF = 0, FLUX_HLLE = 1 };
enum { BC_CARTESIAN_OUTFLOW = 0, BC_SPHERICAL_OUTFLOW3 = 10 };

// Geometry record, same field order as struct of_geom (src/globals.h:90-98),
// so a host geom_arr can be uploaded unchanged.
struct Geom {
  double gcon[4][4];
  double gcov[4][4];
  double g;
  double g_inv;
  double alpha;
  double beta[3];
  double ncon[4];
};
static_assert(sizeof(Geom) == GEOM_DOUBLES * sizeof(double), "Geom layout");

// Single-precision copy of a geometry record, used by the mixed-precision
// reconstruction and face fluxes.
struct GeomF {
  float gcon[4][4];
  float gcov[4][4];
  float g;
  float g_inv;
  float alpha;
  float beta[3];
  float ncon[4];
  float om_alpha;          // 1 - alpha, formed in FP64 (energy flux without the rest-mass cancellation)
};

struct State {   // struct of_state, src/globals.h:100-109
  double ucon[4], ucov[4], bcon[4], bcov[4];
  double bsq, p, eta, ptot;
};

// Status words kept on the device (no host round trip inside a step).
enum {
  ST_FAILURE_EXISTS = 0,     // failure_exists
  ST_FAILTYPE_V = 1,         // failure_type[PFLAG_INTERP_V]
  ST_FAILTYPE_PRIM = 2,      // failure_type[PFLAG_INTERP_PRIM]
  ST_BOUNDARY_PHYS = 3,      // boundary_phys_pflag
  ST_VMAX = 4,               // 4..7: max_char_speed[0..3] as ordered bits
  ST_NFAILTOT = 8,           // 8..22: nfailtot[15]
  ST_SLOW = 23,              // zones queued for the full con-to-prim tree this stage
  ST_BOUNDARY_MPI = 24,      // boundary_mpi_pflag
  ST_CRASH_N1 = 25,          // check_for_crash(): ntot1, ntot2 of the previous step
  ST_CRASH_N2 = 26,
  ST_NSTEP = 27,             // steps completed since the context was created
  ST_CRASH = 28,             // step (1-based) that first crossed the crash threshold, 0 if none
  ST_CRASH_NBAD = 29,        // its n_bad
  ST_MPARTS = 30,            // M_tot parts of the last update (one per wavefront)
  ST_MPARTS2 = 31,           // their second-level parts (one per k_recover workgroup)
  ST_WORDS = 32
};

// Device scalars of the integrator (t, dt ... of src/globals.h:380-390).
// SC_MTOT: M_tot, the sum of U[RHO] over the last stage's zones (src/step_ch.c:309, :389).
enum { SC_T = 0, SC_DT = 1, SC_DT_OLD = 2, SC_T_OLD = 3, SC_MTOT = 4, SC_WORDS = 8 };

struct Params {
  int n1, n2, n3;          // interior cells
  int n1t, n2t, n3t;       // with ghosts
  long long nt;            // n1t*n2t*n3t
  int metric_dim;          // 0 (flat, one geometry record), 2 (axisymmetric 2D cache), 3
  int flat;                // no connection contribution in source()
  int recon;               // RECON_MC or RECON_PPM_MC
  int flux;                // FLUX_LF or FLUX_HLLE
  int ct_para;             // 1: flux_ct_para_fast2, 0: flux_ct (Toth)
  int cooling;             // USE_COOLING_FUNCTION == 1
  int rescale_r;           // RESCALE_R
  int bc;                  // BC_CARTESIAN_OUTFLOW or BC_SPHERICAL_OUTFLOW3
  int bc_nbr;              // faces bordering another block: bit 2*dir + side (dir 0..2)
  int n_extra_newt;        // confirming Newton iterations of the update's inversion (2, or 0 on request)
  int floor_dim;           // 1: floors per x1 index, 3: per cell
  int degen[3];            // N_d == 1
  int n_spatial_dims;
  int track_nfail;         // keep nfail[15][NT] per-cell counters
  int mixed;               // 1: FP32 reconstruction, face fluxes and EMFs
  int dyn;                 // 1: time-dependent metric, geometry levels refilled every step
  int otf;                 // with dyn: face records and the update's zone records evaluated in
                           // the kernels from X.orbit instead of read from stored levels
  long long nconn;         // connection records (interior zones); with dyn the levels are
                           // structure-of-arrays, component l of zone z at conn[l*nconn + z]
  double dx[4], invdx[4];
  double gam, gam_m1_o_gam;
  double cour;
  double t_final;          // GridLength[0] + startx[0]
  double q_exp, s_norm;    // cooling exponent / normalisation (src/phys.c:656-657)
};

// Device pointers of one context.
struct PnOrbit;

// The binary of the time-dependent metric (hgpu_pn.h evaluates it).
struct PnBinary {
  double m1, m2;          // masses (m1 + m2 = M)
  double b0, phi0;        // separation and phase at t = 0
  int inspiral;           // 1: radiation-reaction decay of b (when both masses are > 0);
                          // 0: circular orbit of fixed b0
  // numerical coordinates: theta map 1 (COORD_DIAGONAL) or 4 (COORD_DIAGONAL3)
  double R0, th_beg, th_length, h_slope;
  int coords, diag3_exponent;
  double diag3_factor;
};
// Binary and grid of a context whose metric is the binary's; x0: startx.
// tab: the coordinate functions at every zone's lower face (h = 0) and centre
// (h = 1) along each direction, [2][N1T][2] (r, dr), then [2][N2T][3]
// (sin theta, cos theta, dtheta/dX2), then [2][N3T][2] (sin phi, cos phi);
// null: evaluate them.
struct PnGrid {
  PnBinary B;
  double x0[4];
  int gpos[3];            // global index of the block's zone 0 (a decomposed grid's block offset)
  const double* tab;
};

struct Ptrs {
  double* p;               // primitive state p  (SoA, 8*NT)
  double* ph;              // half-step state ph (SoA, 8*NT)
  double* F;               // face fluxes [3][NFLUX][NT]
  double* emf;             // corner EMFs [3][NT]
  const Geom* geom;        // geometry records (the level of the stage being computed)
  const Geom* geom_n;      // level of U^n (time t of the step; = geom for a static metric)
  const Geom* geom_new;    // level of the stage's result (t + dt/2 or t + dt; = geom when static)
  const Geom* geom_s;      // stage's level for face records after geom moves to geom_new (null: geom)
  const PnOrbit* orbit;    // binary metric: the holes' orbit at the step's levels t, t + dt/2, t + dt
  PnGrid pn;               // binary metric: its grid (with P.otf, zone-centre records are evaluated
  int cent_lvl;            // at level cent_lvl of orbit instead of read from geom)
  const GeomF* geomf;      // their FP32 copy (mixed precision)
  float* Ff;               // FP32 face fluxes [3][NFLUX][NT] (mixed precision)
  float* emff;             // FP32 corner EMFs [3][NT] (mixed precision)
  const double* conn;      // connection [N_CONN][64]
  const double* omega;     // cooling: Omega(r) per x1 index
  const double* tr;        // cooling: T(r)/(gam-1) per x1 index
  const double* rscale;    // [6][N1T] r^1.5, r^3, r at CENT then FACE1
  const double* irscale;   // [6][N1T] reciprocals
  const double* rhoflr;    // floor tables (floor_dim 1: [N1T], 3: [NT])
  const double* uuflr;
  double* gamma;           // p_gamma [NT]
  double* Uslow;           // U^{n+1} of the stage [8][NT]
  double* Un;              // U^n of the step, kept from the predictor for the corrector [8][NT]
  unsigned* slow_idx;      // zones queued for the full con-to-prim tree [n_interior]
  int* pflag;              // [NT]
  int* nfail;              // [15][NT]
  unsigned long long* status;  // [ST_WORDS]
  double* scal;            // [SC_WORDS]
  double* mpart;           // M_tot parts of the update, one per wavefront
  double* mpart2;          // their sums, one per workgroup of k_recover (<= 1024)
};

// FP32 copy of a geometry record (mixed precision); 1 - alpha is formed in FP64.
HD void geom_f32(const Geom& g, GeomF& f) {
  for (int a = 0; a < 4; a++) for (int b = 0; b < 4; b++) {
    f.gcon[a][b] = (float)g.gcon[a][b];
    f.gcov[a][b] = (float)g.gcov[a][b];
  }
  f.g = (float)g.g; f.g_inv = (float)g.g_inv; f.alpha = (float)g.alpha;
  for (int a = 0; a < 3; a++) f.beta[a] = (float)g.beta[a];
  for (int a = 0; a < 4; a++) f.ncon[a] = (float)g.ncon[a];
  f.om_alpha = (float)(1. - g.alpha);
}

HD long long cidx(const Params& P, int i, int j, int k) {
  return ((long long)i * P.n2t + j) * P.n3t + k;
}

HD int geom_id(const Params& P, int i, int j, int k, int pos) {
  if (P.metric_dim == 3) return (int)(k + P.n3t * (j + P.n2t * (i + (long long)P.n1t * pos)));
  if (P.metric_dim == 2) return j + P.n2t * (i + P.n1t * pos);
  return 0;
}

HD int conn_id(const Params& P, int i, int j, int k) {
  if (P.metric_dim == 3) return (k - NG) + P.n3 * ((j - NG) + P.n2 * (i - NG));
  if (P.metric_dim == 2) return (j - NG) + P.n2 * (i - NG);
  return 0;
}

}  // namespace hgpu
