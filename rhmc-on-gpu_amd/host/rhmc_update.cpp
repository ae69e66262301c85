// rhmc_update.cpp -- RHMC trajectories driven on the device.
//
// Host-side mirror of the reference's algorithm layer:
//   lib/Update/update.cc:58-152              heatbath -> MD -> Metropolis
//   lib/Update/RevUpdate/revupdate.cc:19-97  reversibility check
//   lib/Tools/RationalApprox/rationalapprox.cc:134-369  rescaling, chi / phi_new
//   lib/Update/MD_integrators/minimum_norm2.cc:28-105    Omelyan 2MN
//   lib/Update/MD_integrators/multistep_2MN.cc:31-168    multistep 2MN
//   lib/Update/MD_integrators/multistep_4MN.cc:30-300    multistep 4MN
//   lib/Update/MD_integrators/leapfrog.cc:18-62          leapfrog, with the
//       momentum sign of the live integrators (P <- P - i eps ipdot)
//   lib/Update/ipdot.cc:83-196, lib/Update/momenta.cc:42-115
// Every field stays on the device; the host only sequences launches and
// reads back solver flags and the final Delta H.
//
// A trajectory advances one lattice, or several independent lattices (any
// extents, one Markov chain each) in lockstep: every multi-shift solve of
// the trajectory is one solve over all lattices (multishift_solve_group: one
// hop and one M^dag M launch per solver step covering every lattice's
// tiles, chunks of steps replayed as HIP graphs), the other kernels run
// lattice after lattice on the same stream.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "../csrc/context.hpp"
#include "gpu.h"

namespace rhmc {

namespace {

using clk = std::chrono::steady_clock;

double secs(clk::time_point t0) { return std::chrono::duration<double>(clk::now() - t0).count(); }

// Traj::solve's prec argument for the Metropolis solve's dispatch
constexpr int kPrecMetro = -2;

struct Coeffs {
  double a0 = 0.0;
  std::vector<double> a, b;
};

// rationalapprox.cc:134-263
Coeffs rescale(int order, double a0, const double *a, const double *b, double mn, double mx, double exponent) {
  (void)mn;
  const double mxp = 1.05 * mx;
  const double eps = std::pow(mxp, exponent);
  Coeffs c;
  c.a0 = a0 * eps;
  c.a.resize(order);
  c.b.resize(order);
  for (int i = 0; i < order; i++) {
    c.a[i] = a[i] * (mxp * eps);
    c.b[i] = b[i] * mxp;
  }
  return c;
}

// one lattice of a trajectory
struct Lat {
  Context &c;
  const rhmc_traj_params &p;
  rhmc_traj_result &r;
  Coeffs md;              // rescaled MD approximation
  double mn = 0, mx = 0;  // spectral bounds of this trajectory
  int status = 0;         // first solver/HIP failure (solver codes, -1 HIP)
  double metro_corr = 0;  // variational correction of the new fermion action

  Lat(Context &cc, const rhmc_traj_params &pp, rhmc_traj_result &rr) : c(cc), p(pp), r(rr) {}

  // mixed runs solves at or above fp32_threshold in FP32 only and the rest
  // in FP64, like the reference's inv_single_double_prec dispatch
  // (cu_inverter.cc:45-60); the rest run as FP32 iterations with FP64
  // defect correction instead with metro_precision=1 (heatbath and
  // Metropolis) or 2 (the Metropolis solve only, metro = true)
  int prec_for(double res, bool metro = false) const {
    if (p.precision != PREC_MIXED) return p.precision;
    if (p.fp32_threshold > 0.0 && res >= p.fp32_threshold) return PREC_FP32;
    return p.metro_precision == 1 || (p.metro_precision == 2 && metro) ? PREC_MIXED : PREC_FP64;
  }
  // MD solve in FP32 with FP32 solutions and the FP32-input force
  bool md_fp32() const { return prec_for(p.res_md) == PREC_FP32 && !p.per_pf_solves; }
  void fail(int rc) {
    if (rc && !status) status = rc;
  }
  // solver status -> ABI status
  int abi_status() const {
    if (status == 0) return RHMC_OK;
    if (status < 0) return RHMC_EHIP;
    return status == kStatusMaxCg ? RHMC_EMAXCG : status == kStatusNonFinite ? RHMC_ENONFINITE : RHMC_EHIP;
  }
};

struct Traj {
  std::vector<Lat> L;
  const rhmc_traj_params &p;  // integrator parameters (equal for every lattice)
  double t_solve = 0;
  // GPU time of the force and update sections: event pairs recorded around
  // them, read once at the end, so the MD has no host wait besides the solves
  std::vector<hipEvent_t> ev_force, ev_update;

  explicit Traj(const rhmc_traj_params &pp) : p(pp) {}
  ~Traj() {
    for (hipEvent_t e : ev_force)
      if (e) hipEventDestroy(e);
    for (hipEvent_t e : ev_update)
      if (e) hipEventDestroy(e);
  }
  void mark(std::vector<hipEvent_t> &v) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) e = nullptr;
    else hipEventRecord(e, stream());
    v.push_back(e);
  }
  // seconds between the events of each pair (all recorded work complete)
  static double gpu_seconds(std::vector<hipEvent_t> &v) {
    double t = 0.0;
    if (!v.empty() && v.back()) hipEventSynchronize(v.back());
    for (size_t i = 0; i + 1 < v.size(); i += 2) {
      float ms = 0.f;
      if (v[i] && v[i + 1] && hipEventElapsedTime(&ms, v[i], v[i + 1]) == hipSuccess) t += 1e-3 * ms;
    }
    for (hipEvent_t e : v)
      if (e) hipEventDestroy(e);
    v.clear();
    return t;
  }

  bool ok() const {
    for (const Lat &l : L)
      if (l.status) return false;
    return true;
  }
  int abi_status() const {
    for (const Lat &l : L)
      if (l.status) return l.abi_status();
    return RHMC_OK;
  }
  hipStream_t stream() const { return L[0].c.stream; }

  // One multi-shift solve per lattice of b[l] with the poles of co[l] to
  // residual res: prec < 0 takes each lattice's dispatch for res (kPrecMetro:
  // the dispatch of the Metropolis solve); f32 writes
  // FP32 solutions into smf32 (the MD path), else FP64 into x[l] (default
  // smf).  Lattices in lockstep unless one pseudofermion at a time is asked
  // (the reference's per-pseudofermion loop, per lattice).
  void solve(const std::vector<const double2 *> &b, const std::vector<const Coeffs *> &co, double res, int prec,
             bool f32, const std::vector<double2 *> *x = nullptr) {
    const auto t0 = clk::now();
    const int n = (int)L.size();
    std::vector<Context *> cs(n);
    std::vector<MsSpec> sp(n);
    for (int l = 0; l < n; l++) {
      Lat &la = L[l];
      Context &c = la.c;
      cs[l] = &c;
      const int order = (int)co[l]->b.size();
      const int pr = f32 ? PREC_FP32 : prec < 0 ? la.prec_for(res, prec == kPrecMetro) : prec;
      if (f32 && !c.smf32) {
        c.smf32 = (float2 *)dev_alloc(c, sizeof(float2) * 3 * c.Vh * (size_t)c.n_pf * c.max_order);
        if (!c.smf32) la.fail(-1);
      }
      sp[l] = MsSpec{b[l], c.n_pf, order, co[l]->b.data(), res, la.p.residual_relative, pr, la.p.max_cg,
                     x ? (*x)[l] : c.smf, f32 ? c.smf32 : nullptr, 0, 0};
    }
    if (!ok()) return;
    if (p.per_pf_solves) {
      // cu_inverter.cc:42-75: one pseudofermion at a time, per-shift search directions
      for (int l = 0; l < n; l++) {
        Context &c = *cs[l];
        const long v = 3 * c.Vh;
        const int saved = c.ms_algorithm;
        c.ms_algorithm = 1;
        int rc = 0;
        for (int ps = 0; ps < c.n_pf && !rc; ps++) {
          int it = 0;
          rc = multishift_solve(c, sp[l].b + ps * v, 1, sp[l].nshift, sp[l].shifts, res, sp[l].rel, sp[l].prec,
                                sp[l].max_iter, sp[l].x + (size_t)ps * sp[l].nshift * v, &it);
          sp[l].iters += it;
        }
        c.ms_algorithm = saved;
        sp[l].rc = rc;
      }
    } else if (n == 1) {
      sp[0].rc = multishift_solve(*cs[0], sp[0].b, sp[0].nrhs, sp[0].nshift, sp[0].shifts, res, sp[0].rel,
                                  sp[0].prec, sp[0].max_iter, sp[0].x, &sp[0].iters, sp[0].x32);
    } else if (multishift_solve_group(cs, sp) < 0) {
      for (auto &q : sp)
        if (!q.rc) q.rc = -1;
    }
    for (int l = 0; l < n; l++) {
      Lat &la = L[l];
      la.r.cg_iterations += sp[l].iters;
      if (sp[l].prec == PREC_FP32) la.r.cg_iterations_fp32 += sp[l].iters;
      la.r.n_solves++;
      la.c.smf_order = sp[l].nshift;
      la.fail(sp[l].rc);
    }
    t_solve += secs(t0);
  }

  // ipdot from the fermion force (+ gauge force when with_gauge)
  void force(bool fermion, bool gauge) {
    const int n = (int)L.size();
    if (fermion && !p.pure_gauge) {
      const bool f32 = L[0].md_fp32();
      std::vector<const double2 *> b(n);
      std::vector<const Coeffs *> co(n);
      for (int l = 0; l < n; l++) {
        b[l] = L[l].c.chi;
        co[l] = &L[l].md;
      }
      solve(b, co, p.res_md, -1, f32);
      if (!ok()) return;
      mark(ev_force);
      for (Lat &la : L) {
        Context &c = la.c;
        const int order = (int)la.md.a.size();
        std::vector<double> coef((size_t)c.n_pf * order);
        for (int ps = 0; ps < c.n_pf; ps++)
          for (int k = 0; k < order; k++) coef[(size_t)ps * order + k] = la.md.a[k];
        const int rcf = f32 ? fermion_force32(c, c.n_pf * order, coef.data(), c.smf32, 0)
                            : fermion_force(c, c.n_pf * order, coef.data(), c.smf, 0);
        if (rcf) la.fail(-1);
        if (gauge) gauge_force(c, 1);
      }
      mark(ev_force);
    } else {
      mark(ev_force);
      for (Lat &la : L) gauge_force(la.c, 0);
      mark(ev_force);
    }
  }
  void P(double step, bool fermion, bool gauge) {
    force(fermion, gauge);
    mark(ev_update);
    for (Lat &la : L) momenta_update(la.c, step);
    mark(ev_update);
  }
  void Q(double step) {
    mark(ev_update);
    for (Lat &la : L) exp_update(la.c, step);
    mark(ev_update);
  }

  static constexpr double kLambda = 0.1931833275037836;  // Omelyan et al.

  void omelyan() {  // minimum_norm2.cc:28-105
    const double eps = 1.0 / p.no_md, l = kLambda;
    const bool f = !p.pure_gauge;
    Q(eps * l);
    for (int md_ = 1; md_ < p.no_md; md_++) {
      P(-0.5 * eps, f, true);
      Q(eps * (1.0 - 2.0 * l));
      P(-0.5 * eps, f, true);
      Q(eps * 2.0 * l);
    }
    P(-0.5 * eps, f, true);
    Q(eps * (1.0 - 2.0 * l));
    P(-0.5 * eps, f, true);
    Q(eps * l);
  }
  void ms2mn_gauge(double scale) {  // multistep_2MN.cc:31-103
    const double eps = 1.0 / p.no_md, l = kLambda;
    P(-eps * scale * l, false, true);
    for (int md_ = 1; md_ < p.gauge_scale; md_++) {
      Q(0.5 * eps * scale);
      P(-eps * (1.0 - 2.0 * l) * scale, false, true);
      Q(0.5 * eps * scale);
      P(-eps * 2.0 * l * scale, false, true);
    }
    Q(0.5 * eps * scale);
    P(-eps * (1.0 - 2.0 * l) * scale, false, true);
    Q(0.5 * eps * scale);
    P(-eps * l * scale, false, true);
  }
  void ms2mn() {  // multistep_2MN.cc:104-168
    const double eps = 1.0 / p.no_md, l = kLambda, gs = 0.5 / (double)p.gauge_scale;
    P(-eps * l, true, false);
    for (int md_ = 1; md_ < p.no_md; md_++) {
      ms2mn_gauge(gs);
      P(-eps * (1.0 - 2.0 * l), true, false);
      ms2mn_gauge(gs);
      P(-eps * 2.0 * l, true, false);
    }
    ms2mn_gauge(gs);
    P(-eps * (1.0 - 2.0 * l), true, false);
    ms2mn_gauge(gs);
    P(-eps * l, true, false);
  }
  // multistep_4MN.cc:37-40 (Takaishi, de Forcrand)
  static constexpr double k4Theta = 0.08398315262876693, k4Rho = 0.2539785108410595,
                          k4Lambda = 0.6822365335719091, k4Mu = -0.03230286765269967;
  void ms4mn_gauge(double scale) {  // multistep_4MN.cc:30-165
    const double e = scale / p.no_md, th = k4Theta, rh = k4Rho, la = k4Lambda, mu = k4Mu;
    const double hf = 0.5 * (1.0 - 2.0 * (la + th)), mid = 1.0 - 2.0 * (mu + rh);
    P(-e * th, false, true);
    for (int md_ = 1; md_ < p.gauge_scale; md_++) {
      Q(e * rh); P(-e * la, false, true);
      Q(e * mu); P(-e * hf, false, true);
      Q(e * mid); P(-e * hf, false, true);
      Q(e * mu); P(-e * la, false, true);
      Q(e * rh); P(-e * 2.0 * th, false, true);
    }
    Q(e * rh); P(-e * la, false, true);
    Q(e * mu); P(-e * hf, false, true);
    Q(e * mid); P(-e * hf, false, true);
    Q(e * mu); P(-e * la, false, true);
    Q(e * rh); P(-e * th, false, true);
  }
  void ms4mn() {  // multistep_4MN.cc:167-300
    const double eps = 1.0 / p.no_md, gs = 1.0 / (double)p.gauge_scale;
    const double th = k4Theta, rh = k4Rho, la = k4Lambda, mu = k4Mu;
    const double hf = 0.5 * (1.0 - 2.0 * (la + th)), mid = 1.0 - 2.0 * (mu + rh);
    P(-eps * th, true, false);
    for (int md_ = 1; md_ < p.no_md; md_++) {
      ms4mn_gauge(rh * gs); P(-eps * la, true, false);
      ms4mn_gauge(mu * gs); P(-eps * hf, true, false);
      ms4mn_gauge(mid * gs); P(-eps * hf, true, false);
      ms4mn_gauge(mu * gs); P(-eps * la, true, false);
      ms4mn_gauge(rh * gs); P(-eps * 2.0 * th, true, false);
    }
    ms4mn_gauge(rh * gs); P(-eps * la, true, false);
    ms4mn_gauge(mu * gs); P(-eps * hf, true, false);
    ms4mn_gauge(mid * gs); P(-eps * hf, true, false);
    ms4mn_gauge(mu * gs); P(-eps * la, true, false);
    ms4mn_gauge(rh * gs); P(-eps * th, true, false);
  }
  void leapfrog() {  // leapfrog.cc:18-62 with the live momentum sign
    const double eps = 1.0 / p.no_md;
    const bool f = !p.pure_gauge;
    Q(0.5 * eps);
    for (int md_ = 1; md_ < p.no_md; md_++) {
      P(-eps, f, true);
      Q(eps);
    }
    P(-eps, f, true);
    Q(0.5 * eps);
  }
  // update_nometro's MD (update.cc:16-48): the configured integrator
  void run_md() {
    if (p.integrator == RHMC_INT_OMELYAN) omelyan();
    else if (p.integrator == RHMC_INT_MULTISTEP_2MN) ms2mn();
    else if (p.integrator == RHMC_INT_MULTISTEP_4MN) ms4mn();
    else leapfrog();
  }
};

int check_params(Context &c, const rhmc_traj_params &p) {
  if (p.no_md < 1 || p.integrator < 0 || p.integrator > RHMC_INT_MULTISTEP_4MN ||
      ((p.integrator == RHMC_INT_MULTISTEP_2MN || p.integrator == RHMC_INT_MULTISTEP_4MN) && p.gauge_scale < 1)) {
    c.last_error = "trajectory: bad MD parameters";
    return RHMC_EINVAL;
  }
  if (p.metro_precision < 0 || p.metro_precision > 2) {
    c.last_error = "trajectory: metro_precision must be 0, 1 or 2";
    return RHMC_EINVAL;
  }
  if (!p.pure_gauge && (p.order_first < 1 || p.order_md < 1 || p.order_last < 1 || p.order_first > c.max_order ||
                        p.order_md > c.max_order || p.order_last > c.max_order)) {
    c.last_error = "trajectory: bad rational orders";
    return RHMC_EINVAL;
  }
  return RHMC_OK;
}

double e_first(const rhmc_traj_params &p, const Context &c) { return (double)p.n_flavours / 8.0 / (double)c.n_pf; }
double e_md(const rhmc_traj_params &p, const Context &c) { return -(double)p.n_flavours / 4.0 / (double)c.n_pf; }

// Fields, spectral bounds and the heatbath of update.cc:64-79 /
// revupdate.cc:31-47 for every lattice: momenta and phi drawn on the device
// (draw_fields), the bounds (fixed or power method), chi = r_first(MdagM) phi
// (one solve over the lattices) and the rescaled MD approximation.
int prologue(Traj &T) {
  const int n = (int)T.L.size();
  for (Lat &la : T.L) {
    Context &c = la.c;
    const rhmc_traj_params &p = la.p;
    if (p.draw_fields) {
      gauss_momenta(c, p.seed, 0);
      gauss_spinor(c, c.phi0, c.n_pf, p.seed, 1);
    }
    la.mn = p.eig_min;
    la.mx = p.eig_max;
    if (!p.pure_gauge && p.bounds_mode == 1) {
      int loops;
      gauss_spinor(c, c.phinew, 1, p.seed, 2);
      if (power_max(c, c.phinew, p.power_tol, &la.mx, &loops)) return RHMC_EHIP;
      gauss_spinor(c, c.phinew, 1, p.seed, 3);
      if (power_min(c, c.phinew, p.power_tol, la.mx, &la.mn, &loops)) return RHMC_EHIP;
    }
    la.r.eig_min = la.mn;
    la.r.eig_max = la.mx;
  }
  if (T.p.pure_gauge) return RHMC_OK;
  // first_inv_approx_calc: chi = r_first(MdagM) phi
  std::vector<Coeffs> cf(n);
  std::vector<const double2 *> b(n);
  std::vector<const Coeffs *> co(n);
  for (int l = 0; l < n; l++) {
    Lat &la = T.L[l];
    cf[l] = rescale(la.p.order_first, la.p.first_a0, la.p.first_a, la.p.first_b, la.mn, la.mx, e_first(la.p, la.c));
    b[l] = la.c.phi0;
    co[l] = &cf[l];
  }
  T.solve(b, co, T.p.res_metro, -1, false);
  if (!T.ok()) return T.abi_status();
  for (int l = 0; l < n; l++) {
    Lat &la = T.L[l];
    Context &c = la.c;
    hipMemcpyAsync(c.chi, c.phi0, sizeof(double2) * 3 * c.Vh * c.n_pf, hipMemcpyDeviceToDevice, c.stream);
    if (sol_sum(c, c.chi, c.smf, la.p.order_first, cf[l].a0, cf[l].a.data())) return RHMC_EHIP;
    la.md = rescale(la.p.order_md, la.p.md_a0, la.p.md_a, la.p.md_b, la.mn, la.mx, e_md(la.p, c));
  }
  return RHMC_OK;
}

// Recompute the fermion action of the same endpoints with FP64 solves to
// verify_res: chi64 = r_first(MdagM[U_start]) phi, phi_new64 =
// r_last(MdagM[U_end]) chi64; gauge and momentum parts are identical and
// |phi|^2 needs no solve.  delta_h_fp64 = delta_h + the per-site summed
// difference of Re<phi_new, chi>.
int verify_delta_h(Lat &la) {
  Context &c = la.c;
  const rhmc_traj_params &p = la.p;
  rhmc_traj_result &r = la.r;
  const auto tv = clk::now();
  const long sp = 3 * c.Vh;
  const size_t bytes = sizeof(double2) * sp * c.n_pf;
  double2 *chi64 = nullptr, *phn64 = nullptr;
  if (!hip_ok(c, hipMalloc(&chi64, bytes), "verify")) return RHMC_EHIP;
  if (!hip_ok(c, hipMalloc(&phn64, bytes), "verify")) {
    hipFree(chi64);
    return RHMC_EHIP;
  }
  Coeffs cf = rescale(p.order_first, p.first_a0, p.first_a, p.first_b, la.mn, la.mx, e_first(p, c));
  Coeffs cl = rescale(p.order_last, p.last_a0, p.last_a, p.last_b, la.mn, la.mx, e_md(p, c));
  const double vres = p.verify_res > 0.0 ? p.verify_res : p.res_metro;
  // the start links for the chi solve (the compressed copies follow U64)
  std::swap(c.U64, c.U64_save);
  refresh_links32(c);
  int it = 0;
  int rc1 = multishift_solve(c, c.phi0, c.n_pf, p.order_first, cf.b.data(), vres, p.residual_relative, PREC_FP64,
                             p.max_cg, c.smf, &it);
  std::swap(c.U64, c.U64_save);
  refresh_links32(c);
  hipMemcpyAsync(chi64, c.phi0, bytes, hipMemcpyDeviceToDevice, c.stream);
  if (!rc1) sol_sum(c, chi64, c.smf, p.order_first, cf.a0, cf.a.data());
  int rc2 = rc1 ? rc1
                : multishift_solve(c, chi64, c.n_pf, p.order_last, cl.b.data(), vres, p.residual_relative, PREC_FP64,
                                   p.max_cg, c.smf, &it);
  hipMemcpyAsync(phn64, chi64, bytes, hipMemcpyDeviceToDevice, c.stream);
  if (!rc2) sol_sum(c, phn64, c.smf, p.order_last, cl.a0, cl.a.data());
  // sum over sites of Re<phi_new64, chi64> - Re<phi_new, chi>, differenced per site
  std::vector<double> d(c.n_pf);
  int rc3 = rc2 ? rc2 : vec_dot_diffs(c, phn64, chi64, c.phinew, c.chi, c.n_pf, sp, d.data());
  hipStreamSynchronize(c.stream);
  hipFree(chi64);
  hipFree(phn64);
  if (rc1 || rc2 || rc3) {
    la.fail(rc1 ? rc1 : rc2 ? rc2 : -1);
    return la.abi_status();
  }
  double diff = 0.0;
  for (int ps = 0; ps < c.n_pf; ps++) diff += d[ps];
  // the FP64 recomputation's new action is Re<phi_new64, chi64>, without the
  // variational correction of a refined Metropolis solve
  r.delta_h_fp64 = r.delta_h - r.metro_correction + diff;
  r.seconds_verify = secs(tv);
  return RHMC_OK;
}

}  // namespace

// Lattices that can share one lockstep trajectory: the same integrator, MD
// steps, precision dispatch and residuals (each lattice keeps its own
// extents, pseudofermions, rational functions, bounds and seeds).
bool trajectories_lockstep(const std::vector<const rhmc_traj_params *> &ps) {
  const rhmc_traj_params &a = *ps[0];
  for (const rhmc_traj_params *q : ps) {
    const rhmc_traj_params &b = *q;
    if (b.integrator != a.integrator || b.no_md != a.no_md || b.gauge_scale != a.gauge_scale ||
        b.pure_gauge != a.pure_gauge || b.precision != a.precision || b.fp32_threshold != a.fp32_threshold ||
        b.metro_precision != a.metro_precision || b.per_pf_solves || b.res_md != a.res_md ||
        b.res_metro != a.res_metro)
      return false;
  }
  return true;
}

// update.cc:58-152 for one lattice or several in lockstep (one stream: every
// context adopts the first one's stream for the trajectory); status[l]
// receives each lattice's code, the first failure is returned
int run_trajectory_group(std::vector<Context *> &cs, std::vector<const rhmc_traj_params *> &ps,
                         std::vector<rhmc_traj_result *> &rs, std::vector<int> &status) {
  const int n = (int)cs.size();
  status.assign(n, RHMC_OK);
  for (int l = 0; l < n; l++) {
    memset(rs[l], 0, sizeof(rhmc_traj_result));
    if (const int e = check_params(*cs[l], *ps[l])) status[l] = e;
  }
  for (int l = 0; l < n; l++)
    if (status[l]) return status[l];
  const auto t0 = clk::now();
  std::vector<hipStream_t> own(n);
  for (int l = 0; l < n; l++) {
    own[l] = cs[l]->stream;
    if (l > 0) {
      hipStreamSynchronize(own[l]);
      cs[l]->stream = cs[0]->stream;
    }
  }
  Traj T(*ps[0]);
  for (int l = 0; l < n; l++) T.L.emplace_back(*cs[l], *ps[l], *rs[l]);
  auto body = [&]() -> int {
    // Metropolis backup of the links (Conf::save / copy_saved)
    for (Lat &la : T.L)
      hipMemcpyAsync(la.c.U64_save, la.c.U64, sizeof(double2) * 72 * la.c.Vh, hipMemcpyDeviceToDevice, la.c.stream);
    if (const int e = prologue(T)) return e;
    for (Lat &la : T.L) action_local(la.c, la.c.phi0, la.c.chi, !la.p.pure_gauge, 0);
    T.run_md();
    if (!T.ok()) return T.abi_status();
    std::vector<Coeffs> cl(n);
    if (!T.p.pure_gauge) {
      // last_inv_approx_calc: phi_new = r_last(MdagM) chi
      std::vector<const double2 *> b(n);
      std::vector<const Coeffs *> co(n);
      for (int l = 0; l < n; l++) {
        Lat &la = T.L[l];
        cl[l] = rescale(la.p.order_last, la.p.last_a0, la.p.last_a, la.p.last_b, la.mn, la.mx, e_md(la.p, la.c));
        b[l] = la.c.chi;
        co[l] = &cl[l];
      }
      // A solve refined by FP64 defect correction is not Galerkin: Re<chi, x_s>
      // then carries an error linear in x_s's error.  The variational estimate
      // 2 Re<chi, x> - <x, (MdagM + s) x> = Re<chi, x> + Re<x, r> (r the true
      // residual) differs from <chi, (MdagM + s)^-1 chi> by -<e, (MdagM + s) e>
      // only, quadratic in the error e; its correction term sum_s a_s Re<x_s, r_s>
      // is added to the new action.  (For the basis solver from zero r is
      // orthogonal to the Krylov space that holds x, so the term vanishes.)
      for (Lat &la : T.L) {
        la.c.want_xr = la.prec_for(T.p.res_metro, true) == PREC_MIXED;
        la.c.last_xr.clear();
      }
      T.solve(b, co, T.p.res_metro, kPrecMetro, false);
      for (Lat &la : T.L) la.c.want_xr = false;
      if (!T.ok()) return T.abi_status();
      for (int l = 0; l < n; l++) {
        Lat &la = T.L[l];
        Context &c = la.c;
        hipMemcpyAsync(c.phinew, c.chi, sizeof(double2) * 3 * c.Vh * c.n_pf, hipMemcpyDeviceToDevice, c.stream);
        if (sol_sum(c, c.phinew, c.smf, la.p.order_last, cl[l].a0, cl[l].a.data())) return RHMC_EHIP;
        la.metro_corr = 0.0;
        const int order = la.p.order_last;
        if (la.prec_for(T.p.res_metro, true) == PREC_MIXED) {
          if (c.last_xr.size() != (size_t)c.n_pf * order) {
            c.last_error = "trajectory: Metropolis residual dots missing";
            return RHMC_EHIP;
          }
          for (int ps = 0; ps < c.n_pf; ps++)
            for (int k = 0; k < order; k++) la.metro_corr += cl[l].a[k] * c.last_xr[(size_t)ps * order + k];
        }
      }
    }
    const double t_force = Traj::gpu_seconds(T.ev_force), t_update = Traj::gpu_seconds(T.ev_update);
    for (Lat &la : T.L) {
      action_local(la.c, la.c.phinew, la.c.chi, !la.p.pure_gauge, -1);
      double s = 0.0;
      if (action_sum(la.c, &s)) return RHMC_EHIP;
      la.r.delta_h = -s + la.metro_corr;
      la.r.metro_correction = la.metro_corr;
      la.r.delta_h_fp64 = la.r.delta_h;
      la.r.seconds_total = secs(t0);
      la.r.seconds_solver = T.t_solve;
      la.r.seconds_force = t_force;
      la.r.seconds_update = t_update;
    }
    for (Lat &la : T.L) {
      if (la.p.verify_fp64 && !la.p.pure_gauge)
        if (const int e = verify_delta_h(la)) return e;
      la.r.accepted = 1;
      if (la.p.metropolis) {
        const double dh = la.r.delta_h;
        la.r.accepted = (dh < 0.0) || (la.p.metropolis_u < std::exp(-dh));
        if (!la.r.accepted) {
          hipMemcpyAsync(la.c.U64, la.c.U64_save, sizeof(double2) * 72 * la.c.Vh, hipMemcpyDeviceToDevice,
                         la.c.stream);
          refresh_links32(la.c);
        }
      }
    }
    return hip_ok(T.L[0].c, hipStreamSynchronize(T.stream()), "trajectory") ? RHMC_OK : RHMC_EHIP;
  };
  const int rc = body();
  hipStreamSynchronize(cs[0]->stream);
  for (int l = 0; l < n; l++) cs[l]->stream = own[l];
  for (int l = 0; l < n; l++) status[l] = rc ? (T.L[l].status ? T.L[l].abi_status() : rc) : RHMC_OK;
  return rc;
}

int run_trajectory(Context &c, const rhmc_traj_params &p, rhmc_traj_result &r) {
  std::vector<Context *> cs = {&c};
  std::vector<const rhmc_traj_params *> ps = {&p};
  std::vector<rhmc_traj_result *> rs = {&r};
  std::vector<int> st;
  return run_trajectory_group(cs, ps, rs, st);
}

// revupdate.cc:19-97 (rev_update): heatbath, MD, momenta reversed, MD again;
// distances per degree of freedom |U - U_0| and |P + P_0| with the
// reference's normalisation sqrt(sum) * sqrt(1/(V 4 8)).  The links and
// momenta are left at the end of the second MD (as the reference leaves them).
int run_rev_update(Context &c, const rhmc_traj_params &p, rhmc_traj_result &r, double *d_conf, double *d_mom) {
  memset(&r, 0, sizeof(r));
  if (const int e = check_params(c, p)) return e;
  const auto t0 = clk::now();
  Traj T(p);
  T.L.emplace_back(c, p, r);
  if (const int e = prologue(T)) return e;
  double2 *P0 = nullptr;
  if (!hip_ok(c, hipMalloc((void **)&P0, sizeof(double2) * 32 * c.Vh), "saved momenta")) return RHMC_EHIP;
  // save_momenta / the links at the start (Conf::save)
  hipMemcpyAsync(P0, c.P, sizeof(double2) * 32 * c.Vh, hipMemcpyDeviceToDevice, c.stream);
  hipMemcpyAsync(c.U64_save, c.U64, sizeof(double2) * 72 * c.Vh, hipMemcpyDeviceToDevice, c.stream);
  T.run_md();
  int rc = T.abi_status();
  if (rc == RHMC_OK) {
    momenta_scale(c, -1.0);  // reverse_momenta: P <- -P
    T.run_md();
    rc = T.abi_status();
  }
  double s1 = 0.0, s2 = 0.0;
  if (rc == RHMC_OK && (pair_norm2(c, c.U64, c.U64_save, 0, &s1) || pair_norm2(c, c.P, P0, 1, &s2)))
    rc = RHMC_EHIP;
  hipStreamSynchronize(c.stream);
  hipFree(P0);
  if (rc != RHMC_OK) return rc;
  const double norm = std::sqrt(1.0 / (double)c.V * 0.25 * 0.125);
  *d_conf = std::sqrt(s1) * norm;
  *d_mom = std::sqrt(s2) * norm;
  r.seconds_total = secs(t0);
  r.seconds_solver = T.t_solve;
  r.seconds_force = Traj::gpu_seconds(T.ev_force);
  r.seconds_update = Traj::gpu_seconds(T.ev_update);
  return RHMC_OK;
}

}  // namespace rhmc
