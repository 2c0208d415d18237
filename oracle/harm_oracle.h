/* harm_oracle.h -- CPU restatement of coolharm3d's step_ch() (TEST INFRASTRUCTURE).
 *
 * This library is the checker of the GPU library, never part of it: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it.
 * It follows the reference's own data layout and pass structure
 * (array-of-structures state [i][j][k][var]; reconstruct_fast() into full
 * p_L/p_R arrays; numerical_flux(); fix_flux(); flux_ct*(); the per-zone
 * update loop of advance(); bounds(); fixup_interp_*()) so that it checks the
 * GPU's fused kernels by a different route.
 *
 * Configuration fields carry the reference's compile-time options
 * (src/options.h) as run-time values.
 */
#ifndef HARM_ORACLE_H
#define HARM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int n1, n2, n3;          /* interior zones (N1, N2, N3) */
  int metric;              /* 0: METRIC_MINK_CARTESIAN, 2: METRIC_KS_SPHERICAL */
  int coords;              /* 0: COORD_IDENTITY, 1: COORD_DIAGONAL (modified Kerr-Schild), 4: COORD_DIAGONAL3 */
  int recon;               /* 2: RECON_MC, 5: RECON_PPM_MC */
  int flux;                /* 0: FLUX_LF, 1: FLUX_HLLE */
  int ct_para;             /* USE_FLUX_CT_PARA */
  int cooling;             /* USE_COOLING_FUNCTION == 1 */
  int rescale_r;           /* RESCALE_R */
  int bc;                  /* 0: BC_CARTESIAN_OUTFLOW, 10: BC_SPHERICAL_OUTFLOW3 */
  double startx[4], dx[4];
  double gam, cour, t_final;
  double a, R0, h_slope, th_beg, th_length, r_isco, h_o_r;
  double rhomin, uumin, rhopower, uupower;
  int diag3_exponent;      /* COORD_DIAGONAL3: theta = pi/2 (1 + h_slope f + diag3_factor f^n) */
  double diag3_factor;
} orc_config;

typedef struct orc orc_t;

orc_t* orc_create(const orc_config* cfg);     /* NULL on an invalid configuration */
void orc_destroy(orc_t* o);
void orc_dims(const orc_t* o, int* out6);       /* n1 n2 n3 n1t n2t n3t */

/* state, reference layout [i][j][k][8] over all zones including ghosts */
void orc_set_prims(orc_t* o, const double* p);
void orc_get_prims(const orc_t* o, double* p);
void orc_set_time(orc_t* o, double t, double dt);
void orc_get_time(const orc_t* o, double* out4);   /* t, dt, t_old, dt_old */

void orc_fixup_bounds(orc_t* o);                   /* fixup(p); bounds(p,0) of initial data */
void orc_step(orc_t* o, int nsteps);               /* step_ch() nsteps times */

void orc_get_pflag(const orc_t* o, int* out);      /* [nt] */
void orc_get_gamma(const orc_t* o, double* out);   /* [nt] */
void orc_get_nfail(const orc_t* o, int* out);      /* [15][nt] */
void orc_get_nfailtot(const orc_t* o, long long* out15);
void orc_reset_nfail(orc_t* o);
void orc_get_vmax(const orc_t* o, double* out4);   /* max_char_speed[] of the last flux pass */

/* geometry (tests of the metric/connection restatements) */
void orc_get_geom(const orc_t* o, int i, int j, int pos, double* out42);
void orc_get_conn(const orc_t* o, int i, int j, double* out64);
double orc_divb_max(const orc_t* o);

/* per-zone functions on a 42-double geometry record (struct of_geom layout) */
void orc_prim_to_flux(const double* geom42, double gam, const double* prim, int dir, double* out8);
void orc_vchar(const double* geom42, double gam, const double* prim, int dir, double* cmax, double* cmin);
int orc_utoprim_2d(const double* geom42, double gam, const double* U, double* prim,
                   double* gamma, double* bsq, int* iters);
/* Kerr-Schild record (struct of_geom layout) at (r, th), Kerr-Schild coordinates */
void orc_ks_geom_rth(double a, double r, double th, double* geom42);
int orc_utoprim_1d(const double* geom42, double gam, const double* U, double* prim,
                   double* gamma, double* bsq);
/* face flux (HLLE when flux == 1, LF when 0) of left/right primitive states */
void orc_face_flux(const double* geom42, double gam, int flux, int dir, const double* pl,
                   const double* pr, double* out8, double* ctop);
/* reconstruction of one line y[0..n-1]: face states at faces 3..n-3 (face f is
   between zones f-1 and f): left[f] and right[f]; recon 2 = MC, 5 = PPM+MC */
void orc_recon_line(int recon, int n, const double* y, double* left, double* right);

#ifdef __cplusplus
}
#endif
#endif
