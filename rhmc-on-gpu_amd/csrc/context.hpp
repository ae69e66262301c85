// context.hpp -- device-resident state of one lattice and the internal
// launch API shared by the HIP translation units and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "lattice.hpp"

namespace rhmc {

constexpr int kMaxShifts = 32;   // poles per rational function
constexpr int kBlock = 256;      // threads per workgroup (4 wave64)

enum Prec { PREC_FP64 = 0, PREC_MIXED = 1, PREC_FP32 = 2 };

// per right-hand side state of the multi-shift CG (FP64 scalars on device)
struct MsState {
  double delta, omega, omega_prev, gammag, res;
  // res < 0 on entry: the solver's first launch sets res = max(tol_abs,
  // tol_rel |b|) from the |b|^2 it forms (bnorm2), so the host needs no
  // norm before the solve
  double tol_abs, tol_rel, bnorm2;
  int iter, max_iter, done, stop, nshift, nact;
  int nonfinite;  // a reduction or recurrence produced NaN/Inf (the solve stops)
  double shift[kMaxShifts], z1[kMaxShifts], z2[kMaxShifts], z3[kMaxShifts];
  double om[kMaxShifts], ga[kMaxShifts];
  int flag[kMaxShifts], flag_next[kMaxShifts], its[kMaxShifts];
};

// per system state of the batched single-shift CG used for FP64 refinement
struct CgState {
  double delta, alpha, beta, tol2, shift;
  int iter, max_iter, done, stop;
  int nonfinite;  // a reduction produced NaN/Inf (the system stops)
};

// status codes of the solver entry points (0 = converged)
constexpr int kStatusMaxCg = 2;      // max_cg exhausted (RHMC_EMAXCG)
constexpr int kStatusNonFinite = 3;  // NaN/Inf in a solver reduction (RHMC_ENONFINITE)

struct Workspace;

struct Context {
  LatticeDims L;
  long V = 0, Vh = 0;
  double mass = 0, beta = 0;
  int n_pf = 1, max_order = kMaxShifts;
  int device = 0;
  hipStream_t stream = nullptr;

  // gauge field (eta folded in)
  double2 *U64 = nullptr;       // FP64 18-real
  float4 *U32 = nullptr;        // FP32 12-real copy, refreshed with U64
  double2 *U64c = nullptr;      // FP64 12-real copy, refreshed with U64
  double2 *U64_save = nullptr;  // Metropolis backup
  // algebra fields (8 reals/link)
  double2 *P = nullptr;      // momenta (hermitian)
  double2 *ipdot = nullptr;  // force (anti-hermitian)
  // fermions (FP64, even sites): mf = right-hand sides, smf = shifted solutions
  double2 *mf = nullptr;     // [n_pf][3][Vh]
  double2 *smf = nullptr;    // [n_pf][max_order][3][Vh]
  double2 *phi0 = nullptr;   // heatbath pseudofermion
  double2 *chi = nullptr;    // chi = r_first(MdagM) phi0
  double2 *phinew = nullptr; // r_last(MdagM) chi at the end of the MD
  double *action = nullptr;  // [V] per-site action differences
  float2 *smf32 = nullptr;   // FP32 MD solutions [n_pf][max_order][3][Vh] (allocated on first use)
  int smf_order = 0;         // poles held in smf
  void *noalloc_scratch = nullptr;  // this device's fixed scratch (fields.hip), or null

  // Dslash traversal order: kBlock-site tiles listed so that each eighth of
  // the list (one XCD's share) is a compact sub-block of the lattice
  int *tiles = nullptr;
  int ntiles = 0;
  std::vector<int> tiles_host;
  int tile_split[3] = {1, 1, 1};  // sub-blocks in y, z, t

  // multi-shift algorithm: 0 stored residual basis (falls back to 1 when the
  // basis does not fit), 1 per-shift search directions (inverter.cc:142-339)
  int ms_algorithm = 0;
  // replay the stored-basis solver's iteration chunks as captured HIP graphs
  int use_graphs = 1;

  // live kernel timing of every FP32 hop with n_pf right-hand sides in the
  // stored-basis solver (the MD solver's Dslash), on the device wall clock;
  // graphs stay on
  int time_kernels = 0;
  unsigned long long *kt_clk = nullptr; // per solver iteration: end of the launch before the hop, end of the hop
  int kt_clk_cap = 0;
  unsigned long long *kt_wg = nullptr;  // per-workgroup end stamps of the timed launches
  size_t kt_wg_cap = 0;
  double kt_sum_ms = 0.0;               // hop durations on the device wall clock
  long kt_count = 0;
  double kt_sum2_ms = 0.0;              // durations of the M^dag M launch after it
  long kt_count2 = 0;
  double wall_khz = 0.0;

  Workspace *ws = nullptr;
  std::string last_error;

  // statistics of the last solve
  int last_iters = 0, last_refine_cycles = 0, last_iters_fp32 = 0, last_iters_refine = 0;
  double last_true_res_max = 0.0;  // largest |b - (MdagM+s) x| / target of the last checked solve
  int last_restarts = 0;           // continuations past a full stored basis in the last solve
  // With want_xr set, every FP64-checked solve appends Re<x_j, r_j> of each
  // shifted system j (x its solution, r = b - (MdagM + s_j) x its FP64 true
  // residual) to last_xr; the caller clears it.
  bool want_xr = false;
  std::vector<double> last_xr;
  // the fermion force's coefficients held by the device buffer force_coef_dev
  std::vector<double> force_coef_held;
  const double *force_coef_dev = nullptr;
  int graph_captures = 0;          // solver chunk graphs captured (instantiated) so far
};

// ------------------------------------------------------------------ launches
// all launches are enqueued on ctx.stream
void fill_dims(Context &c, int nx, int ny, int nz, int nt, int antiperiodic);

// Dslash family (solver.hip)
// out[j] = 0.5 * hop(in[j]) to parity par_out, FP64 18-real links
void launch_hop_d(Context &c, int par_out, const double2 *in, double2 *out, int nvec, long vstride_in,
                  long vstride_out);
// the same with the FP64 12-real links (row 2 rebuilt)
void launch_hop_d12(Context &c, int par_out, const double2 *in, double2 *out, int nvec, long vstride_in,
                    long vstride_out);
// r = b - ((m^2+shift_j) p_j - D_eo D_oe p_j), norms2[j]; b index = j / b_group
int residual_d(Context &c, const double2 *b, int b_group, const double2 *x, const double *shifts, int nvec,
               double2 *r, double *norms2);

// multi-shift CG.  b: [nrhs][3][Vh] FP64; x: [nrhs][nshift][3][Vh] FP64 out.
// prec PREC_FP64, PREC_MIXED (FP32 iterations + FP64 refinement) or PREC_FP32
// (FP32 iterations only; the target is floored at the FP32 accuracy).
// rel=0: res is an absolute residual (reference convention);
// rel=1: res is relative to |b| per right-hand side.
// x32 (PREC_FP32 only): FP32 solutions written there instead of x.
int multishift_solve(Context &c, const double2 *b, int nrhs, int nshift, const double *shifts, double res,
                     int rel, int prec, int max_iter, double2 *x, int *iters, float2 *x32 = nullptr);
// one multishift_solve per lattice, run in lockstep on one stream (every
// context's stream must be the same); rc and iters per lattice in the specs,
// the first failure (or 0) returned
struct MsSpec {
  const double2 *b;
  int nrhs, nshift;
  const double *shifts;
  double res;
  int rel, prec, max_iter;
  double2 *x;
  float2 *x32;
  int iters, rc;
};
int multishift_solve_group(std::vector<Context *> &cs, std::vector<MsSpec> &sp);
// the FP64 18-real hop with the tuned (or default) shape, never timing
// candidates: no event, file or allocation
void launch_hop_d_fixed(Context &c, int par_out, const double2 *in, double2 *out, int nvec, long vstride_in,
                        long vstride_out);
// out[j] = 0.5 * hop(in[j]) to parity par_out, FP32 12-real links
void launch_hop_s(Context &c, int par_out, const float2 *in, float2 *out, int nvec, long vstride_in,
                  long vstride_out);

// spectral bounds of MdagM by power iterations (findminmax.cc:12-137)
int power_max(Context &c, const double2 *start, double tol, double *out, int *loops);
int power_min(Context &c, const double2 *start, double tol, double mx, double *out, int *loops);

// fields.hip
int fermion_force(Context &c, int nsys_per_pf, const double *num, const double2 *x, int accumulate);
// the same from FP32 solutions: FP32 hop on 12-real links, FP32 loads, FP64
// accumulation and projection
int fermion_force32(Context &c, int nsys, const double *num, const float2 *x32, int accumulate);
// The allocation-free force: as fermion_force, with the per-system hop
// results and the coefficients in the library's fixed device scratch
// (kNoAllocScratchBytes per GPU), the systems taken in groups that fit it.
// Returns -2 when one system does not fit (c.last_error says why).
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
int fermion_force_noalloc(Context &c, int order, const double *num, int accumulate);
// once per device, with the device current: the device address of the
// fixed scratch (kept per device ordinal); sets c.noalloc_scratch
int noalloc_scratch_init(Context &c);
int gauge_force(Context &c, int accumulate);
int exp_update(Context &c, double step);
int momenta_update(Context &c, double step);
// P <- s P
int momenta_scale(Context &c, double s);
// sum of l2norm2(a - b) over the FP64 18-real links (kind 0) or of
// l2norm2(a + b) over hermitian algebra fields (kind 1), FP64
int pair_norm2(Context &c, const double2 *a, const double2 *b, int kind, double *out);
int sol_sum(Context &c, double2 *io, const double2 *x, int nshift, double a0, const double *num);
int action_local(Context &c, const double2 *phi, const double2 *chi, int with_fermions, int init);
int action_sum(Context &c, double *out);
// chiral condensate and energy / baryon / pressure densities for one Z2
// noise vector (host arrays, or drawn on the device from seed when NULL)
int chiral_meas(Context &c, const double *rnd_e, const double *rnd_o, unsigned long long seed, double res,
                int max_cg, double n_flavours, double *out, int *iters);
int upload_spinors(Context &c, const double *host, double2 *dst, int n);
// plaquettes (spatial, temporal) and Polyakov loop of the device links
int gauge_meas(Context &c, double *pls, double *plt, double *pre, double *pim);
int gauss_momenta(Context &c, unsigned long long seed, unsigned long long stream_id);
int gauss_spinor(Context &c, double2 *out, int n, unsigned long long seed, unsigned long long stream_id);
int refresh_links32(Context &c);
int vec_norms2(Context &c, const double2 *v, int nvec, long vstride, double *out);
// out[j] = Re<a_j, b_j> (fixed-order FP64 reduction)
int vec_dots(Context &c, const double2 *a, const double2 *b, int nvec, long vstride, double *out);
int vec_dot_diffs(Context &c, const double2 *a1, const double2 *b1, const double2 *a2, const double2 *b2, int nvec,
                  long vs, double *out);

// Dslash tile order for the lattice of c (solver.hip); allocates c.tiles
int build_tile_order(Context &c);
// fold the device-clock timing slots of the first n solver iterations into
// kt_sum_ms / kt_count (hop) and kt_sum2_ms / kt_count2 (M^dag M) and clear
// them; release frees the slots
void kernel_timing_harvest(Context &c, int n);
void kernel_timing_release(Context &c);
// allocate and clear the slots (timing switched on)
int kernel_timing_clock(Context &c);
// per-workgroup end stamps, at least n
int kernel_timing_stamps(Context &c, size_t n);

// persistent workspace buffers of the gauge-field kernels (grown on demand,
// freed with the context): the force's hop outputs and coefficients, the
// rational sums' coefficients
enum ScratchUse { SCR_FORCE_HOP = 0, SCR_FORCE_COEF, SCR_SOLSUM_COEF, SCR_COUNT };
void *scratch_get(Context &c, int use, size_t bytes);
// small host <-> device transfers through pinned staging (pageable copies
// cost a runtime staging pass with the GPU idle): h2d is enqueued on
// c.stream (the source may be reused at once), d2h_sync synchronises it
bool h2d(Context &c, void *dst, const void *src, size_t bytes);
bool d2h_sync(Context &c, void *dst, const void *src, size_t bytes);

// host helpers
bool hip_ok(Context &c, hipError_t e, const char *what);
void *dev_alloc(Context &c, size_t bytes);
void dev_free(void *p);
Workspace *workspace_create(Context &c);
void workspace_destroy(Workspace *w);

}  // namespace rhmc
