// rhmc_update.cpp -- one RHMC trajectory driven on the device.
//
// Host-side mirror of the reference's algorithm layer:
//   lib/Update/update.cc:58-152              heatbath -> MD -> Metropolis
//   lib/Tools/RationalApprox/rationalapprox.cc:134-369  rescaling, chi / phi_new
//   lib/Update/MD_integrators/minimum_norm2.cc:28-105    Omelyan 2MN
//   lib/Update/MD_integrators/multistep_2MN.cc:31-168    multistep 2MN
//   lib/Update/MD_integrators/multistep_4MN.cc:30-300    multistep 4MN
//   lib/Update/MD_integrators/leapfrog.cc:18-62          leapfrog, with the
//       momentum sign of the live integrators (P <- P - i eps ipdot)
//   lib/Update/ipdot.cc:83-196, lib/Update/momenta.cc:42-115
// Every field stays on the device; the host only sequences launches and
// reads back solver flags and the final Delta H.
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

struct Traj {
  Context &c;
  const rhmc_traj_params &p;
  rhmc_traj_result &r;
  Coeffs md;
  double t_solve = 0, t_force = 0, t_update = 0;
  int status = 0;

  Traj(Context &cc, const rhmc_traj_params &pp, rhmc_traj_result &rr) : c(cc), p(pp), r(rr) {}

  double secs(clk::time_point t0) { return std::chrono::duration<double>(clk::now() - t0).count(); }

  // mixed runs solves at or above fp32_threshold in FP32 only and the rest
  // in FP64, like the reference's inv_single_double_prec dispatch
  // (cu_inverter.cc:45-60), or with metro_precision=1 the rest as FP32
  // iterations with FP64 defect correction
  int prec_for(double res) const {
    if (p.precision != PREC_MIXED) return p.precision;
    if (p.fp32_threshold > 0.0 && res >= p.fp32_threshold) return PREC_FP32;
    return p.metro_precision == 1 ? PREC_MIXED : PREC_FP64;
  }

  int solve(const double2 *b, const Coeffs &co, double res, int prec = -1, double2 *x = nullptr) {
    auto t0 = clk::now();
    const int order = (int)co.b.size();
    const int pr = prec < 0 ? prec_for(res) : prec;
    double2 *xo = x ? x : c.smf;
    int rc = 0;
    if (p.per_pf_solves) {
      // cu_inverter.cc:42-75: one pseudofermion at a time
      const long sp = 3 * c.Vh;
      const int saved = c.ms_algorithm;
      c.ms_algorithm = 1;
      for (int ps = 0; ps < c.n_pf && !rc; ps++) {
        int it = 0;
        rc = multishift_solve(c, b + ps * sp, 1, order, co.b.data(), res, p.residual_relative, pr, p.max_cg,
                              xo + (size_t)ps * order * sp, &it);
        r.cg_iterations += it;
        if (pr == PREC_FP32) r.cg_iterations_fp32 += it;
      }
      c.ms_algorithm = saved;
    } else {
      int it = 0;
      rc = multishift_solve(c, b, c.n_pf, order, co.b.data(), res, p.residual_relative, pr, p.max_cg, xo, &it);
      r.cg_iterations += it;
      if (pr == PREC_FP32) r.cg_iterations_fp32 += it;
    }
    c.smf_order = order;
    r.n_solves++;
    t_solve += secs(t0);
    if (rc && !status) status = rc;
    return rc;
  }

  // MD solve in FP32 with FP32 solutions and the FP32-input force
  bool md_fp32() const { return prec_for(p.res_md) == PREC_FP32 && !p.per_pf_solves; }

  int solve32(const double2 *b, const Coeffs &co, double res) {
    auto t0 = clk::now();
    const int order = (int)co.b.size();
    if (!c.smf32) {
      c.smf32 = (float2 *)dev_alloc(c, sizeof(float2) * 3 * c.Vh * (size_t)c.n_pf * c.max_order);
      if (!c.smf32) { if (!status) status = -1; return -1; }
    }
    int it = 0;
    int rc = multishift_solve(c, b, c.n_pf, order, co.b.data(), res, p.residual_relative, PREC_FP32, p.max_cg,
                              c.smf, &it, c.smf32);
    r.cg_iterations += it;
    r.cg_iterations_fp32 += it;
    r.n_solves++;
    t_solve += secs(t0);
    if (rc && !status) status = rc;
    return rc;
  }

  // ipdot from the fermion force (+ gauge force when with_gauge)
  void force(bool fermion, bool gauge) {
    if (fermion && !p.pure_gauge) {
      const bool f32 = md_fp32();
      if (f32 ? solve32(c.chi, md, p.res_md) : solve(c.chi, md, p.res_md)) return;
      auto t0 = clk::now();
      const int order = (int)md.a.size();
      std::vector<double> coef((size_t)c.n_pf * order);
      for (int ps = 0; ps < c.n_pf; ps++)
        for (int k = 0; k < order; k++) coef[(size_t)ps * order + k] = md.a[k];
      const int rcf = f32 ? fermion_force32(c, c.n_pf * order, coef.data(), c.smf32, 0)
                          : fermion_force(c, c.n_pf * order, coef.data(), c.smf, 0);
      if (rcf && !status) status = -1;
      if (gauge) gauge_force(c, 1);
      hipStreamSynchronize(c.stream);
      t_force += secs(t0);
    } else {
      auto t0 = clk::now();
      gauge_force(c, 0);
      hipStreamSynchronize(c.stream);
      t_force += secs(t0);
    }
  }
  void P(double step, bool fermion, bool gauge) {
    force(fermion, gauge);
    auto t0 = clk::now();
    momenta_update(c, step);
    t_update += secs(t0);
  }
  void Q(double step) {
    auto t0 = clk::now();
    exp_update(c, step);
    t_update += secs(t0);
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
  // solver status -> ABI status
  int md_status() const {
    if (status == 0) return RHMC_OK;
    if (status < 0) return RHMC_EHIP;
    return status == 2 ? RHMC_EMAXCG : status;
  }
};

int check_params(Context &c, const rhmc_traj_params &p) {
  if (p.no_md < 1 || p.integrator < 0 || p.integrator > RHMC_INT_MULTISTEP_4MN ||
      ((p.integrator == RHMC_INT_MULTISTEP_2MN || p.integrator == RHMC_INT_MULTISTEP_4MN) && p.gauge_scale < 1)) {
    c.last_error = "trajectory: bad MD parameters";
    return RHMC_EINVAL;
  }
  if (!p.pure_gauge && (p.order_first < 1 || p.order_md < 1 || p.order_last < 1 || p.order_first > c.max_order ||
                        p.order_md > c.max_order || p.order_last > c.max_order)) {
    c.last_error = "trajectory: bad rational orders";
    return RHMC_EINVAL;
  }
  return RHMC_OK;
}

// Fields, spectral bounds and the heatbath of update.cc:64-79 /
// revupdate.cc:31-47: momenta and phi drawn on the device (draw_fields), the
// bounds (fixed or power method), chi = r_first(MdagM) phi and the rescaled
// MD approximation in T.md.
int prologue(Context &c, const rhmc_traj_params &p, Traj &T, double &mn, double &mx) {
  const long sp = 3 * c.Vh;
  if (p.draw_fields) {
    gauss_momenta(c, p.seed, 0);
    gauss_spinor(c, c.phi0, c.n_pf, p.seed, 1);
  }
  mn = p.eig_min;
  mx = p.eig_max;
  if (!p.pure_gauge && p.bounds_mode == 1) {
    int loops;
    gauss_spinor(c, c.phinew, 1, p.seed, 2);
    if (power_max(c, c.phinew, p.power_tol, &mx, &loops)) return RHMC_EHIP;
    gauss_spinor(c, c.phinew, 1, p.seed, 3);
    if (power_min(c, c.phinew, p.power_tol, mx, &mn, &loops)) return RHMC_EHIP;
  }
  T.r.eig_min = mn;
  T.r.eig_max = mx;
  if (!p.pure_gauge) {
    const double e_first = (double)p.n_flavours / 8.0 / (double)c.n_pf;
    const double e_md = -(double)p.n_flavours / 4.0 / (double)c.n_pf;
    // first_inv_approx_calc: chi = r_first(MdagM) phi
    Coeffs cf = rescale(p.order_first, p.first_a0, p.first_a, p.first_b, mn, mx, e_first);
    if (T.solve(c.phi0, cf, p.res_metro)) return T.md_status();
    hipMemcpyAsync(c.chi, c.phi0, sizeof(double2) * sp * c.n_pf, hipMemcpyDeviceToDevice, c.stream);
    if (sol_sum(c, c.chi, c.smf, p.order_first, cf.a0, cf.a.data())) return RHMC_EHIP;
    T.md = rescale(p.order_md, p.md_a0, p.md_a, p.md_b, mn, mx, e_md);
  }
  return RHMC_OK;
}

}  // namespace

// revupdate.cc:19-97 (rev_update): heatbath, MD, momenta reversed, MD again;
// distances per degree of freedom |U - U_0| and |P + P_0| with the
// reference's normalisation sqrt(sum) * sqrt(1/(V 4 8)).  The links and
// momenta are left at the end of the second MD (as the reference leaves them).
int run_rev_update(Context &c, const rhmc_traj_params &p, rhmc_traj_result &r, double *d_conf, double *d_mom) {
  memset(&r, 0, sizeof(r));
  if (const int e = check_params(c, p)) return e;
  const auto t0 = clk::now();
  Traj T(c, p, r);
  double mn, mx;
  if (const int e = prologue(c, p, T, mn, mx)) return e;
  double2 *P0 = nullptr;
  if (!hip_ok(c, hipMalloc((void **)&P0, sizeof(double2) * 32 * c.Vh), "saved momenta")) return RHMC_EHIP;
  // save_momenta / the links at the start (Conf::save)
  hipMemcpyAsync(P0, c.P, sizeof(double2) * 32 * c.Vh, hipMemcpyDeviceToDevice, c.stream);
  hipMemcpyAsync(c.U64_save, c.U64, sizeof(double2) * 72 * c.Vh, hipMemcpyDeviceToDevice, c.stream);
  T.run_md();
  int rc = T.md_status();
  if (rc == RHMC_OK) {
    momenta_scale(c, -1.0);  // reverse_momenta: P <- -P
    T.run_md();
    rc = T.md_status();
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
  r.seconds_total = std::chrono::duration<double>(clk::now() - t0).count();
  r.seconds_solver = T.t_solve;
  r.seconds_force = T.t_force;
  r.seconds_update = T.t_update;
  return RHMC_OK;
}

int run_trajectory(Context &c, const rhmc_traj_params &p, rhmc_traj_result &r) {
  memset(&r, 0, sizeof(r));
  if (const int e = check_params(c, p)) return e;
  const auto t0 = clk::now();
  Traj T(c, p, r);
  const long sp = 3 * c.Vh;
  // Metropolis backup of the links (Conf::save / copy_saved)
  hipMemcpyAsync(c.U64_save, c.U64, sizeof(double2) * 72 * c.Vh, hipMemcpyDeviceToDevice, c.stream);
  double mn, mx;
  if (const int e = prologue(c, p, T, mn, mx)) return e;
  const double e_first = (double)p.n_flavours / 8.0 / (double)c.n_pf;
  const double e_md = -(double)p.n_flavours / 4.0 / (double)c.n_pf;
  action_local(c, c.phi0, c.chi, !p.pure_gauge, 0);
  T.run_md();
  if (T.status) return T.md_status();
  if (!p.pure_gauge) {
    // last_inv_approx_calc: phi_new = r_last(MdagM) chi
    Coeffs cl = rescale(p.order_last, p.last_a0, p.last_a, p.last_b, mn, mx, e_md);
    if (T.solve(c.chi, cl, p.res_metro)) return T.md_status();
    hipMemcpyAsync(c.phinew, c.chi, sizeof(double2) * sp * c.n_pf, hipMemcpyDeviceToDevice, c.stream);
    if (sol_sum(c, c.phinew, c.smf, p.order_last, cl.a0, cl.a.data())) return RHMC_EHIP;
  }
  action_local(c, c.phinew, c.chi, !p.pure_gauge, -1);
  double s = 0.0;
  if (action_sum(c, &s)) return RHMC_EHIP;
  r.delta_h = -s;
  r.delta_h_fp64 = r.delta_h;
  r.seconds_total = std::chrono::duration<double>(clk::now() - t0).count();
  r.seconds_solver = T.t_solve;
  r.seconds_force = T.t_force;
  r.seconds_update = T.t_update;
  if (p.verify_fp64 && !p.pure_gauge) {
    const auto tv = clk::now();
    // Recompute the fermion action of the same endpoints with FP64 solves to
    // verify_res: chi64 = r_first(MdagM[U_start]) phi,
    // phi_new64 = r_last(MdagM[U_end]) chi64.
    // Gauge and momentum parts are identical, |phi|^2 needs no solve.
    const size_t bytes = sizeof(double2) * sp * c.n_pf;
    double2 *chi64 = nullptr, *phn64 = nullptr;
    if (!hip_ok(c, hipMalloc(&chi64, bytes), "verify") || !hip_ok(c, hipMalloc(&phn64, bytes), "verify"))
      return RHMC_EHIP;
    Coeffs cf = rescale(p.order_first, p.first_a0, p.first_a, p.first_b, mn, mx, e_first);
    Coeffs cl = rescale(p.order_last, p.last_a0, p.last_a, p.last_b, mn, mx, e_md);
    const double vres = p.verify_res > 0.0 ? p.verify_res : p.res_metro;
    // the start links for the chi solve (the compressed copies follow U64)
    std::swap(c.U64, c.U64_save);
    refresh_links32(c);
    int rc1 = T.solve(c.phi0, cf, vres, PREC_FP64);
    std::swap(c.U64, c.U64_save);
    refresh_links32(c);
    hipMemcpyAsync(chi64, c.phi0, bytes, hipMemcpyDeviceToDevice, c.stream);
    if (!rc1) sol_sum(c, chi64, c.smf, p.order_first, cf.a0, cf.a.data());
    int rc2 = rc1 ? rc1 : T.solve(chi64, cl, vres, PREC_FP64);
    hipMemcpyAsync(phn64, chi64, bytes, hipMemcpyDeviceToDevice, c.stream);
    if (!rc2) sol_sum(c, phn64, c.smf, p.order_last, cl.a0, cl.a.data());
    // sum over sites of Re<phi_new64, chi64> - Re<phi_new, chi>, differenced per site
    std::vector<double> d(c.n_pf);
    int rc3 = rc2 ? rc2 : vec_dot_diffs(c, phn64, chi64, c.phinew, c.chi, c.n_pf, sp, d.data());
    hipFree(chi64);
    hipFree(phn64);
    if (rc1 || rc2 || rc3) return RHMC_EHIP;
    double diff = 0.0;
    for (int ps = 0; ps < c.n_pf; ps++) diff += d[ps];
    r.delta_h_fp64 = r.delta_h + diff;
    r.seconds_verify = std::chrono::duration<double>(clk::now() - tv).count();
  }
  r.accepted = 1;
  if (p.metropolis) {
    const double dh = r.delta_h;
    r.accepted = (dh < 0.0) || (p.metropolis_u < std::exp(-dh));
    if (!r.accepted) {
      hipMemcpyAsync(c.U64, c.U64_save, sizeof(double2) * 72 * c.Vh, hipMemcpyDeviceToDevice, c.stream);
      refresh_links32(c);
    }
  }
  if (!hip_ok(c, hipStreamSynchronize(c.stream), "trajectory")) return RHMC_EHIP;
  return RHMC_OK;
}

}  // namespace rhmc
