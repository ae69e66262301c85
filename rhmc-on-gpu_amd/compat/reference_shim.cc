// reference_shim.cc -- the reference's device entry points implemented on the
// C ABI of include/gpu.h and include/cu_inverter.h.
//
// coppolachan/RHMC-on-GPU's algorithm layer (lib/Update, lib/Action, compiled
// with -DUSE_GPU) calls the extern "C" functions of lib/Cuda/*.h and reads or
// writes the host staging arrays of include/global_var.h:90-98.  This file
// provides exactly those symbols, so the reference's own update(),
// integrators, fermionforce(), cu_multips_shifted_invert() and packers link
// unchanged against librhmc_amd.so.  It is compiled only together with the
// reference sources (see build_compat.py).
//
// Semantics kept from the reference device code:
//   cuda_init0/cuda_init1/cuda_end      lib/Cuda/cuda_init.cu:5-211
//   cuda_shifted_inverter[_d]           lib/Cuda/cuda_inverter[_d].cu (FP32 / FP64 paths)
//   cuda_sol_sum                        lib/Cuda/cuda_sol_sum.cu:168-221 (mf <- a0 mf + sum a_i x_i,
//                                       then chi_packed on the host)
//   cuda_fermion_force                  lib/Cuda/cuda_fermion_force.cu:346-418 (ipdot overwritten)
//   cuda_gauge_deriv(init)              lib/Cuda/cuda_deriv.cu:697-744
//   cuda_exp_su3 / cuda_momenta_sum_multiply   cuda_exponentiate.cu:783, cuda_mom_sum_mul.cu:45
//   cuda_find_max / cuda_find_min       lib/Cuda/cuda_find_min_max.cu:69-222 (tolerance 5e-4)
// The fields cross the boundary in FP64 from the reference's host objects;
// the hi/lo float staging arrays the reference's packers read after
// cuda_end / cuda_sol_sum / cuda_meas_end are filled with the reference's own
// packers, so the unmodified gpu_end() and *_inv_approx_calc() see the same
// data layout as with the CUDA backend.
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#define private public
#include "include/global_var.h"
#include "lib/Tools/Packer/packer.h"
#undef private

#include "cu_inverter.h"
#include "gpu.h"

namespace {

rhmc_gpu_t g_h = nullptr;

void check(int rc, const char *what) {
  if (rc != RHMC_OK) {
    fprintf(stderr, "rhmc_amd %s failed (%d): %s\n", what, rc, gpu_last_error(g_h));
    exit(-1);  // the reference's cudaSafe/cudaCheckError behaviour
  }
}

void ensure_handle() {
  if (g_h) return;
  rhmc_gpu_params p;
  p.nx = nx;
  p.ny = ny;
  p.nz = nz;
  p.nt = nt;
  p.antiperiodic_t = ferm_temp_bc == 0;
  p.mass = mass;
  p.beta = beta;
  p.n_pf = no_ps;
  p.max_order = max_approx_order;
  p.device = gpu_device_to_use;
  int rc = gpu_create(&p, &g_h);
  if (rc != RHMC_OK) {
    fprintf(stderr, "rhmc_amd gpu_create failed (%d)\n", rc);
    exit(-1);
  }
}

double *as_doubles(void *p) { return reinterpret_cast<double *>(p); }

// host copy of the pseudofermion right-hand sides -> chi_packed (hi/lo)
void mf_to_chi_packed() {
  MultiFermion *tmp = new MultiFermion;
  check(gpu_get_fermions(g_h, 0, as_doubles(tmp->fermion)), "gpu_get_fermions");
  smartpack_multifermion(chi_packed, tmp);
  delete tmp;
}

int solve_one(double res, const double *shifts, int n, int psferm, int *ncount, int precision) {
  ensure_handle();
  int it = 0;
  int rc = gpu_shifted_inverter(g_h, res, shifts, n, psferm, precision, &it);
  if (ncount) *ncount = rc == RHMC_EMAXCG ? max_cg : it;
  if (rc != RHMC_OK && rc != RHMC_EMAXCG) check(rc, "cuda_shifted_inverter");
  return rc;
}

}  // namespace

extern "C" {

void cuda_init0(void) {
  ensure_handle();
  check(gpu_init0(g_h, as_doubles(gauge_conf->u_work)), "gpu_init0");
}

void cuda_init1(void) {
  ensure_handle();
  check(gpu_init1(g_h, as_doubles(fermion_phi->fermion), as_doubles(gauge_ipdot->ipdot),
                  as_doubles(gauge_momenta->momenta)),
        "gpu_init1");
}

void cuda_end(void) {
  ensure_handle();
  check(gpu_end(g_h, as_doubles(gauge_conf->u_work), as_doubles(gauge_ipdot->ipdot),
                as_doubles(gauge_momenta->momenta)),
        "gpu_end");
  // the reference's gpu_end() unpacks these staging arrays next
  smartpack_gauge(gauge_field_packed, gauge_conf);
  smartpack_tamatrix(ipdot_packed, gauge_ipdot);
  smartpack_thmatrix(momenta_packed, gauge_momenta);
}

void cuda_get_conf(void) {
  ensure_handle();
  check(gpu_end(g_h, as_doubles(gauge_conf->u_work), nullptr, nullptr), "gpu_end");
  smartpack_gauge(gauge_field_packed, gauge_conf);
}

void cuda_get_momenta(void) {
  ensure_handle();
  check(gpu_end(g_h, nullptr, nullptr, as_doubles(gauge_momenta->momenta)), "gpu_end");
  smartpack_thmatrix(momenta_packed, gauge_momenta);
}

void cuda_put_momenta(void) {
  ensure_handle();
  smartunpack_thmatrix(gauge_momenta, momenta_packed);
  check(gpu_init1(g_h, nullptr, nullptr, as_doubles(gauge_momenta->momenta)), "gpu_init1");
}

// chiral measurements: one FP64 solve with the single right-hand side staged
// in simple_fermion_packed (cuda_init.cu:112-183)
void cuda_meas_init(void) {
  ensure_handle();
  Fermion *f = new Fermion;
  smartunpack_fermion_d(f, simple_fermion_packed);
  MultiFermion *m = new MultiFermion;
  check(gpu_get_fermions(g_h, 0, as_doubles(m->fermion)), "gpu_get_fermions");
  memcpy(m->fermion[0], f->fermion, sizeof(Vec3) * sizeh);
  check(gpu_put_fermions(g_h, 0, as_doubles(m->fermion)), "gpu_put_fermions");
  delete f;
  delete m;
}

void cuda_meas_end(void) {
  ensure_handle();
  std::vector<std::complex<double>> x((size_t)no_ps * sizeh * 3);
  check(gpu_get_fermions(g_h, 1, as_doubles(x.data())), "gpu_get_fermions");
  Fermion *f = new Fermion;
  memcpy(f->fermion, x.data(), sizeof(Vec3) * sizeh);  // pseudofermion 0, shift 0
  smartpack_fermion_d(simple_fermion_packed, f);
  delete f;
}

void cuda_shifted_inverter(const double residual, const double *shifts, const int num_shifts, const int psferm,
                           int *ncount) {
  solve_one(residual, shifts, num_shifts, psferm, ncount, RHMC_PREC_FP32);
}

void cuda_shifted_inverter_d(const double residual, const double *shifts, const int num_shifts, const int psferm,
                             int *ncount) {
  solve_one(residual, shifts, num_shifts, psferm, ncount, RHMC_PREC_FP64);
}

void cuda_sol_sum(const int num_shifts, const double mult_const, const double *numerators) {
  ensure_handle();
  check(gpu_sol_sum(g_h, num_shifts, mult_const, numerators), "gpu_sol_sum");
  mf_to_chi_packed();
}

void cuda_fermion_force(int num_shifts, float *numerators) {
  ensure_handle();
  std::vector<double> num(numerators, numerators + num_shifts);
  check(gpu_fermion_force(g_h, num_shifts, num.data(), 0), "gpu_fermion_force");
}

void cuda_gauge_deriv(int ipdot_init) {
  ensure_handle();
  check(gpu_gauge_deriv(g_h, ipdot_init), "gpu_gauge_deriv");
}

void cuda_exp_su3(float step) {
  ensure_handle();
  check(gpu_exp_su3(g_h, (double)step), "gpu_exp_su3");
}

void cuda_momenta_sum_multiply(float step) {
  ensure_handle();
  check(gpu_momenta_sum_multiply(g_h, (double)step), "gpu_momenta_sum_multiply");
}

void cuda_find_max(REAL *max) {
  ensure_handle();
  double v = 0.0;
  check(gpu_find_max(g_h, &v, 5e-4, 12345ull), "gpu_find_max");
  *max = v;
}

void cuda_find_min(REAL *min, const REAL max) {
  ensure_handle();
  double v = 0.0;
  check(gpu_find_min(g_h, &v, max, 5e-4, 54321ull), "gpu_find_min");
  *min = v;
}

void GPUDeviceInfo(const int gpu_device) { (void)gpu_device; }

}  // extern "C"
