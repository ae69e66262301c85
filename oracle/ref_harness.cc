// ref_harness.cc -- drives the reference's own FP64 CPU path (compiled from
// the unmodified reference sources under /root/reference) on inputs written
// by tests/golden/make_golden.py, and dumps its outputs.
//
// TEST INFRASTRUCTURE ONLY: it produces the golden vectors that pin the
// oracle restatement.  Built by oracle/build_ref.py into oracle/_ref/.
//
// The harness replaces two reference functions on purpose:
//   * rationalapprox_calc() (needs MPFR headers absent from this image) is not
//     linked: the normalised coefficients are read from approx.txt instead;
//   * findminmax() (random-start power method) is replaced by fixed spectral
//     bounds read from approx.txt, so the rescaled coefficients are exact
//     inputs rather than RNG-dependent.
// Everything else (Deo/Doe, multi-shift CG, fermion force, staples, ipdot,
// exponentiation, momenta update, action, integrators, update flow) is the
// reference code itself.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <string>
#include <unistd.h>

#define private public
#define DEFINE_GLOBALS
#include "include/global_var.h"
#include "include/geometry.h"
#include "include/tools.h"
#include "lib/Action/action.h"
#include "lib/Action/fermions.h"
#include "lib/Action/staples.h"
#include "lib/Tools/RationalApprox/rationalapprox.h"
#include "lib/Update/MD_integrators/md_integrators.h"
#include "lib/Update/ipdot.h"
#include "lib/Update/momenta.h"
#ifdef USE_GPU
#include "include/gpu.h"
#endif
#undef private

void multips_shifted_invert(ShiftMultiFermion *out, MultiFermion *in, REAL res, RationalApprox approx);
void chiral_meas(complex<REAL> &chiral, complex<REAL> &e_dens, complex<REAL> &b_dens, complex<REAL> &p_dens);
void update_nometro();
void fermionforce(void);

static double g_eig_min = 0.0, g_eig_max = 1.0;

void findminmax(REAL &min, REAL &max) {
  min = g_eig_min;
  max = g_eig_max;
}

static std::string g_dir;

static void read_raw(const char *name, void *dst, size_t bytes) {
  std::string p = g_dir + "/" + name;
  FILE *f = fopen(p.c_str(), "rb");
  if (!f) { fprintf(stderr, "missing %s\n", p.c_str()); exit(2); }
  if (fread(dst, 1, bytes, f) != bytes) { fprintf(stderr, "short %s\n", p.c_str()); exit(2); }
  fclose(f);
}
static void write_raw(const char *name, const void *src, size_t bytes) {
  std::string p = g_dir + "/" + name;
  FILE *f = fopen(p.c_str(), "wb");
  fwrite(src, 1, bytes, f);
  fclose(f);
}

struct Coeffs { int order; double a0; double a[64]; double b[64]; };

static void read_coeffs(std::ifstream &in, Coeffs &c) {
  in >> c.order >> c.a0;
  for (int i = 0; i < c.order; i++) in >> c.a[i];
  for (int i = 0; i < c.order; i++) in >> c.b[i];
}

int main(int argc, char **argv) {
  if (argc < 3) { fprintf(stderr, "usage: ref_harness <dir> <op> [arg]\n"); return 1; }
  g_dir = argv[1];
  std::string op = argv[2];
  double arg = argc > 3 ? atof(argv[3]) : 0.0;

  initrand(12345);
  eta = new int[no_links];
  init_geo();
  gauge_conf = new Conf(1);
  gauge_staples = new Staples();
  gauge_momenta = new Momenta();
  gauge_ipdot = new Ipdot();
  d_vector1 = new double[size];
  d_vector2 = new double[size];
  c_vector1 = new complex<double>[size];
  c_vector2 = new complex<double>[size];
  plaq_test = new double[size];
  fermion_phi = new MultiFermion;
  fermion_chi = new MultiFermion;
  loc_r = new Fermion; loc_h = new Fermion; loc_s = new Fermion; loc_p = new Fermion;
  p_shiftferm = new ShiftFermion;
  fermion_shiftmulti = new ShiftMultiFermion;
  update_iteration = 0;
#ifdef USE_GPU
  // host staging arrays of the reference's device boundary (init.cc:79-96)
  simple_fermion_packed = new float[6 * sizeh * 2];
  chi_packed = new float[6 * sizeh * no_ps * 2];
  psi_packed = new float[6 * sizeh * max_approx_order * no_ps * 2];
  gauge_field_packed = new float[12 * no_links * 2];
  ipdot_packed = new float[8 * no_links];
  momenta_packed = new float[8 * no_links];
  shift_table = new int[8 * size];
#endif

  // inputs
  read_raw("links.bin", gauge_conf->u_work, sizeof(Su3) * no_links);
  memcpy(gauge_conf->u_save, gauge_conf->u_work, sizeof(Su3) * no_links);

  Coeffs cf, cm, cl;
  {
    std::ifstream in((g_dir + "/approx.txt").c_str());
    in >> g_eig_min >> g_eig_max;
    read_coeffs(in, cf);
    read_coeffs(in, cm);
    read_coeffs(in, cl);
  }
  first_inv_approx_norm_coeff = new RationalApprox(lambda_min_metro, cf.order, cf.a0, cf.a, cf.b);
  md_inv_approx_norm_coeff = new RationalApprox(lambda_min_md, cm.order, cm.a0, cm.a, cm.b);
  last_inv_approx_norm_coeff = new RationalApprox(lambda_min_metro, cl.order, cl.a0, cl.a, cl.b);

  if (op == "dslash") {
    // in_e -> Doe -> out_o ; in_o -> Deo -> out_e
    Fermion *ie = new Fermion, *io = new Fermion, *oe = new Fermion, *oo = new Fermion;
    read_raw("in_e.bin", ie->fermion, sizeof(Vec3) * sizeh);
    read_raw("in_o.bin", io->fermion, sizeof(Vec3) * sizeh);
    Doe(oo, ie);
    Deo(oe, io);
    write_raw("out_doe.bin", oo->fermion, sizeof(Vec3) * sizeh);
    write_raw("out_deo.bin", oe->fermion, sizeof(Vec3) * sizeh);
  } else if (op == "multishift") {
    // shifts/numerators of the "md" entry used verbatim, residual = arg
    read_raw("chi.bin", fermion_chi->fermion, sizeof(Vec3) * sizeh * no_ps);
    RationalApprox ap(lambda_min_md, cm.order, cm.a0, cm.a, cm.b);
    multips_shifted_invert(fermion_shiftmulti, fermion_chi, arg, ap);
    for (int ps = 0; ps < no_ps; ps++)
      for (int k = 0; k < cm.order; k++) {
        char name[64];
        snprintf(name, sizeof name, "out_x_%d_%d.bin", ps, k);
        write_raw(name, fermion_shiftmulti->fermion[ps][k], sizeof(Vec3) * sizeh);
      }
  } else if (op == "force_fermion" || op == "force_full" || op == "force_gauge") {
    read_raw("chi.bin", fermion_chi->fermion, sizeof(Vec3) * sizeh * no_ps);
    if (op == "force_fermion") calc_ipdot_fermion();
    else if (op == "force_full") calc_ipdot();
    else calc_ipdot_gauge();
    write_raw("out_ipdot.bin", gauge_ipdot->ipdot, sizeof(Su3) * no_links);
  } else if (op == "exp") {
    read_raw("mom.bin", gauge_momenta->momenta, sizeof(Su3) * no_links);
    conf_left_exp_multiply(complex<REAL>(0.0, arg));
    write_raw("out_links.bin", gauge_conf->u_work, sizeof(Su3) * no_links);
  } else if (op == "mom") {
    read_raw("mom.bin", gauge_momenta->momenta, sizeof(Su3) * no_links);
    read_raw("ipdot.bin", gauge_ipdot->ipdot, sizeof(Su3) * no_links);
    momenta_sum_multiply(complex<REAL>(0.0, arg));
    write_raw("out_mom.bin", gauge_momenta->momenta, sizeof(Su3) * no_links);
  } else if (op == "action") {
    read_raw("mom.bin", gauge_momenta->momenta, sizeof(Su3) * no_links);
    read_raw("phi.bin", fermion_phi->fermion, sizeof(Vec3) * sizeh * no_ps);
    read_raw("chi.bin", fermion_chi->fermion, sizeof(Vec3) * sizeh * no_ps);
    calc_action(plaq_test, 0);
    write_raw("out_action0.bin", plaq_test, sizeof(double) * size);
    calc_action(plaq_test, -1);
    write_raw("out_action1.bin", plaq_test, sizeof(double) * size);
  } else if (op == "traj") {
    // update.cc:58-152 without create_momenta/create_phi (inputs) and
    // without the Metropolis draw.
    read_raw("mom.bin", gauge_momenta->momenta, sizeof(Su3) * no_links);
    read_raw("phi.bin", fermion_phi->fermion, sizeof(Vec3) * sizeh * no_ps);
#ifdef USE_GPU
    // the reference's device sequence: rhmc.cc:72 then update.cc:75-108
    gpu_init0();
    gpu_init1();
#endif
    first_inv_approx_calc(residue_metro);
    write_raw("out_chi.bin", fermion_chi->fermion, sizeof(Vec3) * sizeh * no_ps);
    calc_action(plaq_test, 0);
    update_nometro();
    last_inv_approx_calc(residue_metro);
#ifdef USE_GPU
    gpu_end();
#endif
    calc_action(plaq_test, -1);
    global_sum(plaq_test, size);
    double de = -plaq_test[0];
    write_raw("out_links.bin", gauge_conf->u_work, sizeof(Su3) * no_links);
    write_raw("out_mom.bin", gauge_momenta->momenta, sizeof(Su3) * no_links);
    write_raw("out_phinew.bin", fermion_phi->fermion, sizeof(Vec3) * sizeh * no_ps);
    write_raw("out_dh.bin", &de, sizeof(double));
  } else if (op == "meas_gauge") {
    // Conf::calc_plaq / calc_poly (gaugemeas.cc:6-104) on the stored links
    REAL ps, pt, pr, pi;
    gauge_conf->calc_plaq(ps, pt);
    gauge_conf->calc_poly(pr, pi);
    double v[4] = {ps, pt, pr, pi};
    write_raw("out_meas.bin", v, sizeof v);
  } else if (op == "meas_chiral") {
    // chiral_meas (chiralmeas.cc:18-205), rand_vect = 1: dumps the Z2 noise
    // it drew and its four complex results
    chi_e = new Fermion; chi_o = new Fermion; phi_e = new Fermion; phi_o = new Fermion;
    rnd_e = new Fermion; rnd_o = new Fermion;
    complex<REAL> ch, ed, bd, pd;
    chiral_meas(ch, ed, bd, pd);
    double v[8] = {ch.real(), ch.imag(), ed.real(), ed.imag(), bd.real(), bd.imag(), pd.real(), pd.imag()};
    write_raw("out_chiral.bin", v, sizeof v);
    write_raw("out_rnd_e.bin", rnd_e->fermion, sizeof(Vec3) * sizeh);
    write_raw("out_rnd_o.bin", rnd_o->fermion, sizeof(Vec3) * sizeh);
  } else if (op == "conf_write") {
    // Conf::write_last (conf.cc:179-204) of the input links, update_iteration = arg
    update_iteration = (int)arg;
    memcpy(gauge_conf->u_save, gauge_conf->u_work, sizeof(Su3) * no_links);
    if (chdir(g_dir.c_str()) != 0) return 2;
    gauge_conf->write_last();
  } else if (op == "conf_read") {
    // Conf(init=2) (conf.cc:55-96) on the file "config" in the directory
    if (chdir(g_dir.c_str()) != 0) return 2;
    Conf *c2 = new Conf(2);
    write_raw("out_links.bin", c2->u_work, sizeof(Su3) * no_links);
    double it = update_iteration;
    write_raw("out_iter.bin", &it, sizeof(double));
  } else {
    fprintf(stderr, "unknown op %s\n", op.c_str());
    return 1;
  }
  return 0;
}
