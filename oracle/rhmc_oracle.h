/*
 * rhmc_oracle.h -- FP64 CPU restatement of the reference RHMC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path
 * is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links
 * or calls it.
 *
 * Conventions (all follow the reference CPU code, coppolachan/RHMC-on-GPU):
 *   - full-lattice site index  i = parity*Vh + lex/2,
 *     lex = x + nx*(y + ny*(z + nz*t))          (lib/Main/Geometry/geometry.cc:5-11,82-86)
 *   - links U[mu*V + i], 3x3 complex row-major, staggered phase eta folded in
 *     (lib/Main/Conf/conf.cc:36,49)
 *   - spinors on one parity: psi[h][3] complex, h in [0,Vh)
 *   - su(3) fields (momenta P, ipdot) stored as full 3x3 complex matrices
 *   - complex numbers are interleaved (re,im) doubles, i.e. layout-compatible
 *     with std::complex<double>.
 *
 * Each function names the reference lines it restates.
 */
#ifndef RHMC_ORACLE_H
#define RHMC_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex orc_c;

typedef struct {
  int nx, ny, nz, nt;
  int antiperiodic_t; /* 1: antiperiodic temporal fermion b.c. (ferm_temp_bc==0) */
  long V, Vh;
  long *nnp; /* [V][4] */
  long *nnm; /* [V][4] */
  int *eta;  /* [4V]   */
} orc_lattice;

/* geometry.cc:26-153 */
void orc_set_threads(int n);
int orc_lattice_init(orc_lattice *lat, int nx, int ny, int nz, int nt, int antiperiodic_t);
void orc_lattice_free(orc_lattice *lat);

/* su3.cc:69-112 (sunitarize), conf.cc:235-251 (unitarize_with_eta) */
void orc_sunitarize(orc_c m[3][3]);
void orc_unitarize_with_eta(const orc_lattice *lat, orc_c (*U)[3][3]);

/* su3.cc:161-183 traceless anti-hermitian part */
void orc_ta(orc_c m[3][3]);
/* su3.cc:187-219 6th-order Taylor exponential + sunitarize */
void orc_su3_exp(orc_c m[3][3]);

/* fermionmatrix.cc:8-52  out_even = D_eo in_odd */
void orc_deo(const orc_lattice *lat, const orc_c (*U)[3][3], orc_c (*out)[3], const orc_c (*in)[3]);
/* fermionmatrix.cc:58-97 out_odd = D_oe in_even */
void orc_doe(const orc_lattice *lat, const orc_c (*U)[3][3], orc_c (*out)[3], const orc_c (*in)[3]);
/* inverter.cc:27-34 out = (m^2 - D_eo D_oe) in (even sites) */
void orc_mdagm(const orc_lattice *lat, const orc_c (*U)[3][3], double mass,
               orc_c (*out)[3], const orc_c (*in)[3]);

/* include/tools.h:25-57 pairwise in-place tree sum; returns vec[0] */
double orc_global_sum(double *vec, long n);

/* inverter.cc:142-339 Jegerlehner multi-shift CG for one right-hand side.
 * x[s] (n_shift half-lattice spinors) receives the solutions of
 * (M^dag M + shifts[s]) x = b.  Returns the iteration count. */
int orc_multishift(const orc_lattice *lat, const orc_c (*U)[3][3], double mass,
                   const orc_c (*b)[3], int n_shift, const double *shifts,
                   double res, int max_cg, orc_c (*x)[3] /* [n_shift*Vh] */);

/* inverter.cc:12-127 plain CG with a zero start (the reference starts from a
 * gaussian; the start is an argument here). */
int orc_cg(const orc_lattice *lat, const orc_c (*U)[3][3], double mass,
           const orc_c (*b)[3], double res, int max_cg, orc_c (*x)[3]);

/* fermionforce.cc:46-104: accumulate the fermion force of one pseudofermion
 * into F (4V matrices, not yet multiplied by U nor projected). */
void orc_fermion_force_accum(const orc_lattice *lat, const orc_c (*U)[3][3], int n_shift,
                             const double *num, const orc_c (*x)[3], orc_c (*F)[3][3]);

/* staples.cc:16-74 */
void orc_staples(const orc_lattice *lat, const orc_c (*U)[3][3], orc_c (*staple)[3][3]);

/* ipdot.cc:25-79  ipdot = TA(U * beta/3 * staple) */
void orc_ipdot_gauge(const orc_lattice *lat, const orc_c (*U)[3][3], double beta,
                     orc_c (*ipdot)[3][3]);
/* ipdot.cc:160-196  ipdot = TA(U * (beta/3 staple + F_fermion)); F supplied */
void orc_ipdot_total(const orc_lattice *lat, const orc_c (*U)[3][3], double beta,
                     const orc_c (*F)[3][3], int with_gauge, orc_c (*ipdot)[3][3]);

/* momenta.cc:42-78  U <- exp(p P) U, then unitarize_with_eta; p = i*step */
void orc_conf_left_exp_multiply(const orc_lattice *lat, orc_c (*U)[3][3],
                                const orc_c (*P)[3][3], double step);
/* momenta.cc:82-115  P <- P + (i*step) ipdot */
void orc_momenta_sum_multiply(const orc_lattice *lat, orc_c (*P)[3][3],
                              const orc_c (*ipdot)[3][3], double step);

/* action.cc:6-156 per-site action accumulation (init 0: set, -1: subtract) */
/* chiralmeas.cc:18-205, one Z2 noise vector, no imaginary chemical
 * potential; out[8] = (chiral, e_dens, b_dens, p_dens) as (re, im); returns
 * the CG iterations */
int orc_chiral_meas(const orc_lattice *lat, const orc_c (*U)[3][3], double mass, double n_flavours,
                    const orc_c (*rnd_e)[3], const orc_c (*rnd_o)[3], double res, int max_cg, double *out);
/* gaugemeas.cc:6-104: plaquettes and Polyakov loop of the stored links */
void orc_plaquette(const orc_lattice *lat, const orc_c (*U)[3][3], double *pls, double *plt);
void orc_polyakov(const orc_lattice *lat, const orc_c (*U)[3][3], double *re, double *im);
void orc_calc_action(const orc_lattice *lat, const orc_c (*U)[3][3], const orc_c (*P)[3][3],
                     double beta, int n_pf, const orc_c (*phi)[3], const orc_c (*chi)[3],
                     int with_fermions, double *value, int init);

/* findminmax.cc:12-137 power iterations from the given start vector */
int orc_findmaxeig(const orc_lattice *lat, const orc_c (*U)[3][3], double mass,
                   const orc_c (*start)[3], double tol, double *max);
int orc_findmineig(const orc_lattice *lat, const orc_c (*U)[3][3], double mass,
                   const orc_c (*start)[3], double tol, double max, double *min);

/* rationalapprox.cc:134-263 rescale normalised coefficients to [min,max]
 * (min,max already padded by the caller or not: pad=1 applies 0.95/1.05). */
void orc_rescale_approx(int order, double a0_in, const double *a_in, const double *b_in,
                        double min, double max, double exponent, int pad,
                        double *a0, double *a, double *b);

/* ---------------------------------------------------------------------- */
/* Trajectory (update.cc:58-152) with the MD integrators of
 * lib/Update/MD_integrators.  The state is supplied explicitly. */

enum { ORC_INT_OMELYAN = 0, ORC_INT_MULTISTEP_2MN = 1, ORC_INT_LEAPFROG = 2, ORC_INT_MULTISTEP_4MN = 3 };

typedef struct {
  double beta, mass;
  int n_flavours, n_pf;
  int no_md, gauge_scale, integrator;
  double res_md, res_metro;
  int max_cg;
  /* normalised (unrescaled) rational coefficients on [lambda_min,1] */
  int order_first, order_md, order_last;
  const double *first_a, *first_b; double first_a0;
  const double *md_a, *md_b; double md_a0;
  const double *last_a, *last_b; double last_a0;
  /* spectral bounds used for every rescaling (already measured) */
  double eig_min, eig_max;
  int pure_gauge;
} orc_traj_params;

/* Runs one trajectory from links U (modified in place to the MD end point),
 * momenta P (modified), pseudofermion phi [n_pf*Vh].  Writes the
 * per-component action differences and returns Delta H = S_new - S_old
 * (update.cc:110-114).  No Metropolis decision is taken here. */
double orc_trajectory(const orc_lattice *lat, const orc_traj_params *prm, orc_c (*U)[3][3],
                      orc_c (*P)[3][3], const orc_c (*phi)[3], orc_c (*chi_out)[3],
                      orc_c (*phi_new_out)[3]);

/* Single force evaluation as used inside the MD (ipdot.cc:83-196 +
 * fermionforce.cc): ipdot = TA(U (beta/3 staple [if with_gauge] + F)), where F
 * comes from a multi-shift solve at res on chi (n_pf pseudofermions). */
int orc_full_force(const orc_lattice *lat, const orc_c (*U)[3][3], double beta, double mass,
                   int n_pf, const orc_c (*chi)[3], int order, double a0, const double *a,
                   const double *b, double res, int max_cg, int with_gauge, orc_c (*ipdot)[3][3]);

#ifdef __cplusplus
}
#endif
#endif
