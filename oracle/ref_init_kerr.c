/*
 * ref_init_kerr.c -- test infrastructure only (linked into oracle/_ref builds).
 *
 * Problem set-up for the Kerr-Schild / modified-Kerr-Schild torus runs of the
 * reference.  The reference's own torus set-up, src/init.torus.c, does not
 * compile as shipped (stray ';' after the init_prim prototype at
 * src/init.torus.c:355) and hard-codes its spin, radii and seed, so this file
 * provides init()/init_base()/init_data() in the same shape as
 * src/init.torus.c:19-158 and src/init.torus.c:174-248: it sets the grid and
 * coordinate globals from ref_params[] and calls the reference's
 * calc_all_geom() (src/metric.c:4111).  The primitive variables are supplied
 * afterwards by the caller (ref_set_prims + ref_fixup_bounds), so the GPU path
 * and the reference start from identical arrays.
 *
 * ref_params: [0]=a [1]=Rin [2]=Rout [3]=h_slope [4]=gam [5]=cour
 *             [6]=x3 extent [7]=t_end [8]=initial dt
 *             [9]=diag3_exponent (COORD_DIAGONAL3 builds; 0 = the reference's 9,
 *                 src/init.nz_pn_bbh_disk.c:358)
 */
#include "decs.h"

extern double ref_params[64];
extern double risco_calc(int do_prograde);
extern double rhorizon_calc(int pos_sign);

void init_base(void)
{
  int i;
  void calc_all_geom(void);

  gam  = ref_params[4];
  cour = ref_params[5];

  /* set_special_coord(), src/init.torus.c:174-248, with configurable values */
  a         = ref_params[0];
  /* risco_calc() reads asq, which calc_all_geom() sets only later
     (src/metric.c:4117); set it first so r_isco is the prograde ISCO of a. */
  asq       = a*a;
  r_isco    = risco_calc(1);
  r_horizon = rhorizon_calc(1);
  R0        = 0.;
  Rin       = ref_params[1];
  Rout      = ref_params[2];
  h_slope   = ref_params[3];
  X1_slope  = 1.;
  X1_0      = log(1.e6);
  th_cutout = SMALL * M_PI;
  th_beg    = th_cutout;
  th_end    = M_PI - th_cutout;
  th_length = th_end - th_beg;
  /* COORD_DIAGONAL3's parameters as set_special_coord() sets them,
     src/init.nz_pn_bbh_disk.c:357-359 (h_slope is its dtheta_min/(pi/2)) */
  diag3_exponent = (ref_params[9] > 0.) ? (int)ref_params[9] : 9;
  diag3_factor   = 1. - 2*th_cutout/M_PI - h_slope;

  t = 0.;
  GridLength[0] = ref_params[7];
  GridLength[1] = log((Rout-R0)/(Rin-R0));
  GridLength[2] = 1.;
  GridLength[3] = ref_params[6];
  SDLOOP1 dx[i] = GridLength[i] / totalsize[i];

  startx[0] = 0.;
  startx[1] = log(Rin-R0);
  startx[2] = 0.;
  startx[3] = 0.;
  SDLOOP1 startx[i] -= NG*dx[i];
  dV = dx[1]*dx[2]*dx[3];

  n_restart = 1000000;
  for (i = 0; i < N_OUT_TYPES; i++) DT_out[i] = 1.e30;

  dx[0] = ref_params[8];
  calc_all_geom();
}

void init_data(void) { }

void init(void)
{
  init_base();
  init_data();
}
