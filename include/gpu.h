/*
 * gpu.h -- C ABI of the MI355X RHMC hot path (gfx950, HIP).
 *
 * Replaces the device boundary of the reference (coppolachan/RHMC-on-GPU):
 *   include/gpu.h:6-8            gpu_init0 / gpu_init1 / gpu_end
 *   lib/Cuda/cuda_init.h:4-11    cuda_init0/1, cuda_end, cuda_get_conf, momenta get/put
 *   lib/Cuda/cuda_exponentiate.h:4, cuda_mom_sum_mul.h:4, cuda_deriv.h:4,
 *   cuda_fermion_force.h:4, cuda_sol_sum.h:4, cuda_find_min_max.h:6-7
 * The reference passes everything through global singletons and exits on
 * error; this ABI takes an explicit handle and plain pointers, and every call
 * returns 0 on success (nonzero: see gpu_last_error()).
 *
 * Host data layouts (identical to the reference's host arrays, so a
 * std::complex<double> array can be passed as double*):
 *   links / momenta / ipdot : double[4*V][3][3][2], entry mu*V + i with the
 *                             reference site index i = parity*V/2 + lex/2,
 *                             lex = x + nx*(y + ny*(z + nz*t)); links carry the
 *                             staggered phases (lib/Main/Conf/conf.cc:36).
 *   pseudofermions          : double[n_pf][V/2][3][2]   (even sites)
 *   shifted solutions       : double[n_pf][order][V/2][3][2]
 */
#ifndef RHMC_AMD_GPU_H
#define RHMC_AMD_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rhmc_gpu_ctx *rhmc_gpu_t;

typedef struct {
  int nx, ny, nz, nt;    /* even extents (lib/Tools/exception.cc:12-15)       */
  int antiperiodic_t;    /* 1 = antiperiodic temporal fermion b.c.            */
  double mass;           /* bare staggered mass (parameters.h `mass`)         */
  double beta;           /* Wilson coupling (parameters.h `beta`)             */
  int n_pf;              /* pseudofermions (parameters.h `no_ps`)             */
  int max_order;         /* largest number of poles used (<= 32)              */
  int device;            /* HIP device ordinal                                */
} rhmc_gpu_params;

/* Status codes.  RHMC_EMAXCG: max_cg exhausted before convergence (the
 * reference writes a warning and carries on, inverter.cc:290-296, or exits,
 * cu_inverter.cc:62-69); RHMC_ENONFINITE: a solver reduction produced NaN or
 * Inf (non-finite links or sources), the solve stops at once. */
enum { RHMC_OK = 0, RHMC_EINVAL = 1, RHMC_EMAXCG = 2, RHMC_ENONFINITE = 3, RHMC_EHIP = -1 };

/* Context lifetime (the reference's implicit global state; cuda_init.cu:5-59). */
int gpu_create(const rhmc_gpu_params *p, rhmc_gpu_t *out);
int gpu_destroy(rhmc_gpu_t h);
const char *gpu_last_error(rhmc_gpu_t h);

/* include/gpu.h:6 gpu_init0 -> cuda_init0 (cuda_init.cu:5-57): gauge field
 * to the device (FP64 18-real and the FP32 and FP64 12-real copies).  The
 * links must be unitary, with the staggered phase in row 2 as the reference
 * stores them: RHMC_EINVAL (with the largest deviation in gpu_last_error)
 * when row 2 differs from eta (row 0 x row 1)^* by more than 1e-10, since
 * the 12-real copies rebuild row 2.  The reference computes with such links
 * as given (conf.cc:86-92 does not re-project them); every configuration it
 * writes or evolves is unitary to rounding. */
int gpu_init0(rhmc_gpu_t h, const double *links);

/* include/gpu.h:7 gpu_init1 -> cuda_init1 (cuda_init.cu:61-108): pseudofermion
 * right-hand sides ("mf"), ipdot and momenta; NULL leaves a field unchanged. */
int gpu_init1(rhmc_gpu_t h, const double *phi, const double *ipdot, const double *momenta);

/* include/gpu.h:8 gpu_end -> cuda_end (cuda_init.cu:187-211): fields back to
 * the host; NULL skips a field.  Device state is kept (no free). */
int gpu_end(rhmc_gpu_t h, double *links, double *ipdot, double *momenta);

/* lib/Main/Conf/conf.cc:143-204 (Conf::write / write_last): the device links
 * to a text file in the reference's format (header "nx ny nz nt beta mass
 * no_flavours update_iteration", then every link as three rows of "(re,im)"
 * at precision 16); written to path.tmp and renamed. */
int gpu_write_conf(rhmc_gpu_t h, const char *path, double beta, double mass, double n_flavours,
                   int update_iteration);

/* lib/Main/Conf/conf.cc:55-96 (Conf(init=2)): links from such a file to the
 * device, as stored; the extents must match the handle's lattice, and the
 * links must be unitary as for gpu_init0.  Header values are returned
 * through the pointers (each may be NULL). */
int gpu_read_conf(rhmc_gpu_t h, const char *path, int *update_iteration, double *beta, double *mass,
                  double *n_flavours);

/* Pseudofermion buffers: which = 0 right-hand sides "mf" (reference
 * mf_device / chi_packed), 1 shifted solutions "smf" (psi_packed, n_pf*order
 * vectors of the last solve), 2 heatbath phi, 3 chi, 4 phi_new. */
int gpu_get_fermions(rhmc_gpu_t h, int which, double *out);
int gpu_put_fermions(rhmc_gpu_t h, int which, const double *in);

/* lib/Cuda/cuda_exponentiate.h:4 cuda_exp_su3(step) (momenta.cc:54):
 * U <- Pi_SU3[exp(i step P)] U, row 2 *= eta (FP64, 6th-order Taylor). */
int gpu_exp_su3(rhmc_gpu_t h, double step);

/* lib/Cuda/cuda_mom_sum_mul.h:4 cuda_momenta_sum_multiply(step) (momenta.cc:94):
 * P <- P + i*step*ipdot. */
int gpu_momenta_sum_multiply(rhmc_gpu_t h, double step);

/* lib/Cuda/cuda_deriv.h:4 cuda_gauge_deriv(ipdot_init) (ipdot.cc:37,162):
 * ipdot (=, or += when ipdot_init=1) TA(U beta/3 staple). */
int gpu_gauge_deriv(rhmc_gpu_t h, int ipdot_init);

/* lib/Cuda/cuda_fermion_force.h:4 cuda_fermion_force(num_shifts, numerators)
 * (fermionforce.cc:44): ipdot = TA(U F) from the shifted solutions of the last
 * cu_multips_shifted_invert; numerators are FP64 here (the reference rounds
 * them to float, fermionforce.cc:42-43).  accumulate=1 adds to ipdot. */
int gpu_fermion_force(rhmc_gpu_t h, int num_shifts, const double *numerators, int accumulate);

/* No counterpart in the reference: gpu_fermion_force for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and result; it
 * calls no hipMalloc/hipFree and creates nothing.  Its scratch (the
 * coefficients and the hop results D_oe s_j) is one fixed 64 MiB device
 * array per GPU, held until the process ends; the systems are processed in
 * groups that fit it, each group's force added to the previous groups'.
 * Limit: V/2 <= 1,398,058 sites (one system, 48 V/2 bytes, must fit);
 * larger lattices get RHMC_EINVAL.  The scratch is shared by every call on
 * that GPU and keeps the last call's contents: calls that overlap on one GPU
 * (several handles on different streams) race.  The Dslash launch shape is
 * the one already tuned for this handle, else the default (no autotuning). */
int gpu_fermion_force_noalloc(rhmc_gpu_t h, int num_shifts, const double *numerators, int accumulate);

/* lib/Cuda/cuda_sol_sum.h:4 cuda_sol_sum (rationalapprox.cc:288,340):
 * mf <- a0*mf + sum_k num[k]*smf[k] for every pseudofermion. */
int gpu_sol_sum(rhmc_gpu_t h, int num_shifts, double mult_const, const double *numerators);

/* lib/Action/fermionforce.cc:10-45 (the USE_GPU branch: the MD multi-shift
 * solve of the pseudofermion right-hand sides in "mf", then
 * cuda_fermion_force): ipdot <- fermion force of
 * S = sum_ps phi^+ (a0 + sum_i num_i (MdagM + shifts_i)^-1) phi, the adjoint
 * dS/dU.  pole_chunk < order solves the poles in groups of pole_chunk and
 * accumulates the force group by group, so only n_pf*pole_chunk solutions are
 * held at once (recompute instead of store; pole_chunk <= max_order);
 * pole_chunk >= order is a single solve.  iterations: summed CG iterations. */
int gpu_fermionforce(rhmc_gpu_t h, double res, int relative, int order, const double *shifts,
                     const double *numerators, int precision, int max_cg, int pole_chunk, int *iterations);

/* Pseudofermion action S = sum_ps Re <mf_ps, (a0 + sum_i num_i
 * (MdagM + shifts_i)^-1) mf_ps> with one multi-shift solve; the fermion part
 * of calc_action(-1) after last_inv_approx_calc (action.cc:87-101,
 * rationalapprox.cc:330-369).  FP64 reductions. */
int gpu_fermion_action(rhmc_gpu_t h, double res, int relative, int order, const double *shifts, double a0,
                       const double *numerators, int precision, int max_cg, double *action);

/* lib/Cuda/cuda_find_min_max.h:6-7 (findminmax.cc:140-183): power iterations
 * for the extreme eigenvalues of MdagM from a gaussian start drawn with seed;
 * tolerance on the relative change (reference: 1e-5). */
int gpu_find_max(rhmc_gpu_t h, double *max, double tol, unsigned long long seed);
int gpu_find_min(rhmc_gpu_t h, double *min, double max, double tol, unsigned long long seed);

/* Action bookkeeping on the device (action.cc:111-156, update.cc:110-114):
 * init=0 stores the per-site action of (U, P, phi); init=-1 subtracts the
 * action of (U, P, phi_new, chi).  gpu_action_sum returns the pairwise tree
 * sum (include/tools.h:25-57) of the per-site differences. */
int gpu_action(rhmc_gpu_t h, int init, int with_fermions);
int gpu_action_sum(rhmc_gpu_t h, double *sum);

/* Gaussian momenta / heatbath phi from the device counter-based generator
 * (not the reference's dSFMT stream). */
int gpu_gauss_momenta(rhmc_gpu_t h, unsigned long long seed, unsigned long long stream);
int gpu_gauss_phi(rhmc_gpu_t h, unsigned long long seed, unsigned long long stream);

/* Dslash on host data, for checking: out = D_oe in (par_out=1) or
 * D_eo in (par_out=0), nvec vectors; precision 0 = FP64 18-real links,
 * 1 = FP32 12-real links (vectors rounded to float). */
int gpu_dslash(rhmc_gpu_t h, int par_out, int precision, int nvec, const double *in, double *out);

/* lib/Measurements/gaugemeas.cc:6-104 (Conf::calc_plaq, Conf::calc_poly):
 * spatial and temporal plaquettes and the Polyakov loop of the device links,
 * with the reference's signs (staggered phases folded in), pairwise sums and
 * normalisation. */
int gpu_gauge_meas(rhmc_gpu_t h, double *plaq_spatial, double *plaq_temporal, double *poly_re, double *poly_im);

/* lib/Measurements/chiralmeas.cc:18-205 for one Z2 noise vector, without an
 * imaginary chemical potential: chi_e = (MdagM)^-1 (m rnd_e - D_eo rnd_o)
 * (FP64 CG to the absolute residual res), chi_o = (rnd_o - D_oe chi_e)/m;
 * out[8] = chiral condensate, energy, baryon and pressure densities as
 * (re, im) pairs with the reference's normalisation.  rnd_e/rnd_o: host
 * noise vectors (V/2 x 3 complex each), or both NULL to draw Z2 noise on the
 * device from seed. */
int gpu_chiral_meas(rhmc_gpu_t h, const double *rnd_e, const double *rnd_o, unsigned long long seed, double res,
                    int max_cg, double n_flavours, double *out, int *iterations);

/* Timing helper: run `reps` device-resident half-lattice Dslash launches on
 * nvec vectors and return the mean kernel time in ms, measured with HIP
 * events on the library's stream.  precision: 0 FP64 18-real links, 1 FP32
 * 12-real, 2 FP64 12-real. */
int gpu_time_dslash(rhmc_gpu_t h, int precision, int nvec, int reps, double *ms);

/* The same with every timed launch preceded by a read pass over flush_bytes
 * (<= 8e9) of other device memory (written once beforehand), so the operands
 * come from HBM rather than the L2s or the Infinity Cache and no dirty lines
 * are left for the timed launch to write back; mean of the per-launch event
 * times. */
int gpu_time_dslash_cold(rhmc_gpu_t h, int precision, int nvec, int reps, double flush_bytes, double *ms);

/* Live kernel timing inside real work: while enabled, the two launches of
 * each iteration of the MD solver (FP32 12-real links, the n_pf
 * pseudofermions) are timed on the device wall clock: [0] the half-lattice
 * hop D_oe, [1] the M^dag M launch after it (with the CG's residual
 * update).  Each interval runs from the end of the launch before to the end
 * of the timed launch's last workgroup, the per-dispatch interval a
 * kernel-trace profiler reports (dispatch, execution and end of kernel);
 * every workgroup stamps its own end.  The stamps sit in slots indexed by
 * the solver iteration, so replayed graphs are timed too.  Only iterations
 * that did work are counted, and not the first hop of each 8-iteration
 * chunk (it follows the chunk's bookkeeping launches instead).
 * mean_ms[2] and count[2] receive the mean durations and counts since the
 * last reset; enable = 1/0 then switches timing on or off and resets,
 * enable = -1 only reads.  No counterpart in the reference. */
int gpu_kernel_timing(rhmc_gpu_t h, int enable, double *mean_ms, long long *count);

/* Timing helper: a plain device copy of `bytes` (read + write), mean ms per
 * copy over `reps` launches (HIP events); the achievable-bandwidth reference
 * and the byte-counter calibration for the roofline. */
int gpu_time_copy(rhmc_gpu_t h, double bytes, int reps, double *ms);

/* Autotuner choices made so far (Dslash right-hand sides per thread) as a
 * JSON array; returns the length written (truncated to buflen-1). */
int gpu_tuning_table(rhmc_gpu_t h, char *buf, int buflen);

/* The context's device as reported by hipGetDeviceProperties, as a JSON
 * object (name, architecture, compute units, wavefront size, HBM bytes,
 * memory clock and bus width, L2 bytes, LDS bytes per CU and per
 * workgroup, shader clock), truncated to buflen-1.  No counterpart in the
 * reference; bench records it beside the roofline's assumed peaks. */
int gpu_device_info(rhmc_gpu_t h, char *buf, int buflen);

/* --------------------------------------------------------------------- */
/* One RHMC trajectory on the device: update.cc:58-152 with the integrators
 * of lib/Update/MD_integrators. */
enum { RHMC_INT_OMELYAN = 0, RHMC_INT_MULTISTEP_2MN = 1, RHMC_INT_LEAPFROG = 2, RHMC_INT_MULTISTEP_4MN = 3 };

typedef struct {
  int n_flavours;
  int no_md, gauge_scale, integrator;
  double res_md, res_metro;       /* multi-shift stopping residuals           */
  int residual_relative;          /* 1: residuals relative to |rhs|           */
  int precision;                  /* 0 FP64, 1 mixed (see fp32_threshold)     */
  double fp32_threshold;          /* mixed: solves with res >= this run FP32
                                     only (the reference's
                                     inv_single_double_prec dispatch,
                                     cu_inverter.cc:45-60); 0 = never       */
  int verify_fp64;                /* 1: recompute Delta H of the same
                                     endpoints with FP64 solves             */
  int max_cg;
  int order_first, order_md, order_last;
  const double *first_a, *first_b; double first_a0;   /* normalised on [eps,1] */
  const double *md_a, *md_b; double md_a0;
  const double *last_a, *last_b; double last_a0;
  double lambda_min;              /* eps of the normalised approximations     */
  int bounds_mode;                /* 0: fixed eig_min/eig_max, 1: power method */
  double eig_min, eig_max;
  double power_tol;
  int draw_fields;                /* 1: draw P and phi on the device (seed)   */
  unsigned long long seed;
  int metropolis;                 /* 1: accept/reject with u (below)          */
  double metropolis_u;
  int pure_gauge;
  int metro_precision;            /* mixed: solves with res < fp32_threshold
                                     (heatbath, Metropolis): 0 FP64 as the
                                     reference dispatches them
                                     (cu_inverter.cc:45-52), 1 FP32 iterations
                                     + FP64 defect correction for both, 2 for
                                     the Metropolis solve only; a refined
                                     Metropolis solve's pole terms are taken
                                     variationally (metro_correction)       */
  int per_pf_solves;              /* 1: one pseudofermion per solve with
                                     per-shift search directions, the
                                     reference's loop (cu_inverter.cc:42-75);
                                     0: all pseudofermions as one multi-RHS
                                     solve on the stored residual basis     */
  double verify_res;              /* verify_fp64: residual of the FP64
                                     recomputation (0: res_metro)           */
} rhmc_traj_params;

typedef struct {
  double delta_h;
  double delta_h_fp64;            /* FP64 recomputation (verify_fp64=1)       */
  int accepted;
  int cg_iterations;              /* summed over all solves                   */
  int cg_iterations_fp32;         /* of which in FP32 solves                  */
  int n_solves;
  double eig_min, eig_max;
  double seconds_total, seconds_solver, seconds_force, seconds_update;  /* without the verification */
  double seconds_verify;          /* the FP64 recomputation (verify_fp64=1)   */
  double metro_correction;        /* included in delta_h when the Metropolis
                                     solve is refined (metro_precision 1, 2):
                                     sum over pseudofermions and poles of
                                     a_s Re<x_s, r_s>, r_s = chi - (MdagM +
                                     b_s) x_s the FP64 true residual, so that
                                     each pole term is 2 Re<chi, x_s> -
                                     <x_s, (MdagM + b_s) x_s>, with an error
                                     quadratic in the solution's; 0 otherwise */
} rhmc_traj_result;

int gpu_trajectory(rhmc_gpu_t h, const rhmc_traj_params *p, rhmc_traj_result *r);

/* lib/Update/RevUpdate/revupdate.cc:19-97 (rev_update, the reference's
 * REVERSIBILITY_TEST): fields (draw_fields), bounds and heatbath as in
 * gpu_trajectory, then the MD of update_nometro, momenta reversed, the MD
 * again.  delta_conf = sqrt(sum |U - U_0|^2 / (V 4 8)) and delta_mom =
 * sqrt(sum |P + P_0|^2 / (V 4 8)) (revupdate.cc:86-92; zero for an exactly
 * reversible integration).  The links and momenta are left at the end of the
 * second MD; no Metropolis step. */
int gpu_rev_update(rhmc_gpu_t h, const rhmc_traj_params *p, double *delta_conf, double *delta_mom,
                   rhmc_traj_result *r);

/* Several independent lattices (distinct handles, any sizes), one trajectory
 * each, run concurrently on the device: independent Markov chains as in the
 * reference's update() per lattice.  status (may be NULL) receives each
 * lattice's return code; the return value is the first failure or RHMC_OK.
 * No counterpart in the reference, which runs one lattice per process. */
int gpu_multi_trajectory(rhmc_gpu_t *handles, int n, const rhmc_traj_params *params, rhmc_traj_result *results,
                         int *status);

#ifdef __cplusplus
}
#endif
#endif
