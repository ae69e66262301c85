// capi.hip -- the extern "C" boundary (include/gpu.h, include/cu_inverter.h).
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <thread>
#include <cstdio>

#include <cmath>
#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "context.hpp"
#include "cu_inverter.h"
#include "gpu.h"

namespace rhmc {
int upload_links(Context &c, const double *host, double2 *dst, double tol, double *max_dev);
// links must be unitary to this (row 2 is rebuilt from rows 0 and 1 by the 12-real copies)
constexpr double kUnitarityTol = 1e-10;
int download_links(Context &c, const double2 *src, double *host);
int upload_alg(Context &c, const double *host, double2 *dst, int herm);
int download_alg(Context &c, const double2 *src, double *host, int herm);
int upload_spinors(Context &c, const double *host, double2 *dst, int n);
int download_spinors(Context &c, const double2 *src, double *host, int n);
int hop_s12(Context &c, int par_out, const float2 *in, float2 *out, int nvec);
int run_trajectory(Context &c, const rhmc_traj_params &p, rhmc_traj_result &r);
int run_rev_update(Context &c, const rhmc_traj_params &p, rhmc_traj_result &r, double *d_conf, double *d_mom);
bool trajectories_lockstep(const std::vector<const rhmc_traj_params *> &ps);
int run_trajectory_group(std::vector<Context *> &cs, std::vector<const rhmc_traj_params *> &ps,
                         std::vector<rhmc_traj_result *> &rs, std::vector<int> &status);
std::string tuning_table(Context &c);
int time_copy(Context &c, size_t bytes, int reps, double *ms);

bool hip_ok(Context &c, hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  c.last_error = std::string(what) + ": " + hipGetErrorString(e);
  return false;
}

// Zero-filled allocation, ordered on the context's stream.  The stream is
// non-blocking, so work on the legacy null stream (a plain hipMemset) would
// not be ordered before later launches on it.
void *dev_alloc(Context &c, size_t bytes) {
  void *p = nullptr;
  if (!hip_ok(c, hipMalloc(&p, bytes), "hipMalloc")) return nullptr;
  if (!hip_ok(c, hipMemsetAsync(p, 0, bytes, c.stream), "hipMemsetAsync")) {
    hipFree(p);
    return nullptr;
  }
  return p;
}

void dev_free(void *p) {
  if (p) hipFree(p);
}

void fill_dims(Context &c, int nx, int ny, int nz, int nt, int antiperiodic) {
  c.L.nx = nx; c.L.ny = ny; c.L.nz = nz; c.L.nt = nt;
  c.L.nxh = nx / 2;
  c.V = (long)nx * ny * nz * nt;
  c.Vh = c.V / 2;
  c.L.Vh = c.Vh;
  c.L.antiperiodic_t = antiperiodic;
  c.L.inv_nxh = 1.0 / c.L.nxh;
  c.L.inv_ny = 1.0 / ny;
  c.L.inv_nz = 1.0 / nz;
}

__global__ void k_f2d(double2 *o, const float2 *i, long n) {
  const long k = (long)blockIdx.x * 256 + threadIdx.x;
  if (k < n) o[k] = make_double2(i[k].x, i[k].y);
}
__global__ void k_d2f(float2 *o, const double2 *i, long n) {
  const long k = (long)blockIdx.x * 256 + threadIdx.x;
  if (k < n) o[k] = make_float2((float)i[k].x, (float)i[k].y);
}

}  // namespace rhmc

using namespace rhmc;

struct rhmc_gpu_ctx {
  Context c;
};

static int fail(rhmc_gpu_t h, const char *msg, int code = RHMC_EINVAL) {
  if (h) h->c.last_error = msg;
  return code;
}

extern "C" {

int gpu_create(const rhmc_gpu_params *p, rhmc_gpu_t *out) {
  if (!p || !out) return RHMC_EINVAL;
  *out = nullptr;
  if (p->nx <= 0 || p->ny <= 0 || p->nz <= 0 || p->nt <= 0 || (p->nx | p->ny | p->nz | p->nt) & 1) return RHMC_EINVAL;
  // 32-bit site arithmetic in the kernels: V/2 < 2^31
  if ((double)p->nx * p->ny * p->nz * p->nt / 2 >= 2147483647.0) return RHMC_EINVAL;
  if (p->n_pf < 1 || p->max_order < 1 || p->max_order > kMaxShifts || p->n_pf * p->max_order > 256)
    return RHMC_EINVAL;
  rhmc_gpu_t h = new (std::nothrow) rhmc_gpu_ctx();
  if (!h) return RHMC_EINVAL;
  Context &c = h->c;
  c.device = p->device;
  if (!hip_ok(c, hipSetDevice(p->device), "hipSetDevice")) { delete h; return RHMC_EHIP; }
  fill_dims(c, p->nx, p->ny, p->nz, p->nt, p->antiperiodic_t);
  c.mass = p->mass;
  c.beta = p->beta;
  c.n_pf = p->n_pf;
  c.max_order = p->max_order;
  if (!hip_ok(c, hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "stream")) { delete h; return RHMC_EHIP; }
  const long Vh = c.Vh;
  c.U64 = (double2 *)dev_alloc(c, sizeof(double2) * 72 * Vh);
  c.U64_save = (double2 *)dev_alloc(c, sizeof(double2) * 72 * Vh);
  c.U32 = (float4 *)dev_alloc(c, sizeof(float4) * 24 * Vh);
  c.U64c = (double2 *)dev_alloc(c, sizeof(double2) * 48 * Vh);
  c.P = (double2 *)dev_alloc(c, sizeof(double2) * 32 * Vh);
  c.ipdot = (double2 *)dev_alloc(c, sizeof(double2) * 32 * Vh);
  const size_t sp = sizeof(double2) * 3 * Vh;
  c.mf = (double2 *)dev_alloc(c, sp * c.n_pf);
  c.smf = (double2 *)dev_alloc(c, sp * c.n_pf * c.max_order);
  c.phi0 = (double2 *)dev_alloc(c, sp * c.n_pf);
  c.chi = (double2 *)dev_alloc(c, sp * c.n_pf);
  c.phinew = (double2 *)dev_alloc(c, sp * c.n_pf);
  c.action = (double *)dev_alloc(c, sizeof(double) * c.V);
  if (!c.U64 || !c.U64_save || !c.U32 || !c.U64c || !c.P || !c.ipdot || !c.mf || !c.smf || !c.phi0 || !c.chi || !c.phinew ||
      !c.action) {
    gpu_destroy(h);
    return RHMC_EHIP;
  }
  c.ws = workspace_create(c);
  if (!c.ws || build_tile_order(c)) { gpu_destroy(h); return RHMC_EHIP; }
  // the fixed scratch of the allocation-free entry points: its address on
  // this device, looked up once per device with the device current (on
  // failure those entry points report it; nothing else depends on it)
  noalloc_scratch_init(c);
  c.use_graphs = getenv("RHMC_NO_GRAPHS") == nullptr;
  *out = h;
  return RHMC_OK;
}

int gpu_destroy(rhmc_gpu_t h) {
  if (!h) return RHMC_OK;
  Context &c = h->c;
  hipSetDevice(c.device);
  if (c.stream) hipStreamSynchronize(c.stream);
  dev_free(c.U64); dev_free(c.U64_save); dev_free(c.U32); dev_free(c.U64c); dev_free(c.P); dev_free(c.ipdot);
  dev_free(c.mf); dev_free(c.smf); dev_free(c.phi0); dev_free(c.chi); dev_free(c.phinew); dev_free(c.action);
  dev_free(c.tiles);
  dev_free(c.smf32);
  kernel_timing_release(c);
  workspace_destroy(c.ws);
  if (c.stream) hipStreamDestroy(c.stream);
  delete h;
  return RHMC_OK;
}

const char *gpu_last_error(rhmc_gpu_t h) { return h ? h->c.last_error.c_str() : "null handle"; }

// Links to the device (FP64 18-real), checked for unitarity, then the
// 12-real copies.  Links with a row 2 that is not eta (row 0 x row 1)^* (to
// kUnitarityTol) are refused and the device keeps its links: the 12-real
// copies rebuild row 2 and would differ from it.  Every configuration the
// reference writes or evolves is unitary to rounding (its text format keeps
// 16 digits).
static int set_links(Context &c, const double *links, const char *who) {
  double dev = 0.0;
  const int up = upload_links(c, links, c.U64, kUnitarityTol, &dev);
  if (up < 0) return RHMC_EHIP;
  if (up > 0) {
    char msg[256];
    snprintf(msg, sizeof msg,
             "%s: links not unitary: row 2 differs from eta (row 0 x row 1)^* by up to %.3g (limit %.0e); "
             "the 12-real link copies would change the operator",
             who, dev, kUnitarityTol);
    c.last_error = msg;
    return RHMC_EINVAL;
  }
  if (refresh_links32(c)) return RHMC_EHIP;
  return hip_ok(c, hipStreamSynchronize(c.stream), who) ? RHMC_OK : RHMC_EHIP;
}

int gpu_init0(rhmc_gpu_t h, const double *links) {
  if (!h || !links) return RHMC_EINVAL;
  Context &c = h->c;
  hipSetDevice(c.device);
  return set_links(c, links, "gpu_init0");
}

int gpu_init1(rhmc_gpu_t h, const double *phi, const double *ipdot, const double *momenta) {
  if (!h) return RHMC_EINVAL;
  Context &c = h->c;
  hipSetDevice(c.device);
  if (phi && upload_spinors(c, phi, c.mf, c.n_pf)) return RHMC_EHIP;
  if (ipdot && upload_alg(c, ipdot, c.ipdot, 0)) return RHMC_EHIP;
  if (momenta && upload_alg(c, momenta, c.P, 1)) return RHMC_EHIP;
  return RHMC_OK;
}

int gpu_end(rhmc_gpu_t h, double *links, double *ipdot, double *momenta) {
  if (!h) return RHMC_EINVAL;
  Context &c = h->c;
  hipSetDevice(c.device);
  if (links && download_links(c, c.U64, links)) return RHMC_EHIP;
  if (ipdot && download_alg(c, c.ipdot, ipdot, 0)) return RHMC_EHIP;
  if (momenta && download_alg(c, c.P, momenta, 1)) return RHMC_EHIP;
  return RHMC_OK;
}

// ---- gauge configuration files in the reference's text format
// conf.cc:143-204 (Conf::write_last): header "nx ny nz nt beta mass
// no_flavours update_iteration", then every link (site order mu*V + i,
// phases folded in) as three rows of "(re,im)" at precision 16
// (su3.cc:225-232), a blank line after each matrix.
static void put_g16(std::string &o, double v) {
  char b[40];
  snprintf(b, sizeof b, "%.16g", v);
  o += b;
}

int gpu_write_conf(rhmc_gpu_t h, const char *path, double beta, double mass, double n_flavours, int update_iteration) {
  if (!h || !path) return RHMC_EINVAL;
  Context &c = h->c;
  hipSetDevice(c.device);
  const long nl = 4 * c.V;
  std::vector<double> u((size_t)nl * 18);
  if (download_links(c, c.U64, u.data())) return RHMC_EHIP;
  std::string o;
  o.reserve((size_t)nl * 9 * 48 + 128);
  o += std::to_string(c.L.nx) + " " + std::to_string(c.L.ny) + " " + std::to_string(c.L.nz) + " " +
       std::to_string(c.L.nt) + " ";
  put_g16(o, beta); o += " ";
  put_g16(o, mass); o += " ";
  put_g16(o, n_flavours); o += " ";
  o += std::to_string(update_iteration) + "\n";
  for (long r = 0; r < nl; r++)
    for (int i = 0; i < 3; i++) {
      for (int k = 0; k < 3; k++) {
        const double *z = &u[(size_t)r * 18 + (i * 3 + k) * 2];
        if (k) o += " ";
        o += "(";
        put_g16(o, z[0]);
        o += ",";
        put_g16(o, z[1]);
        o += ")";
      }
      o += i < 2 ? "\n" : "\n\n";
    }
  const std::string tmp = std::string(path) + ".tmp";
  FILE *f = fopen(tmp.c_str(), "wb");
  if (!f) return fail(h, "gpu_write_conf: cannot open file");
  const bool ok = fwrite(o.data(), 1, o.size(), f) == o.size();
  if (fclose(f) != 0 || !ok || rename(tmp.c_str(), path) != 0) return fail(h, "gpu_write_conf: write failed");
  return RHMC_OK;
}

// conf.cc:55-96 (Conf(init=2)): extents must match; links used as stored
int gpu_read_conf(rhmc_gpu_t h, const char *path, int *update_iteration, double *beta, double *mass,
                  double *n_flavours) {
  if (!h || !path) return RHMC_EINVAL;
  Context &c = h->c;
  hipSetDevice(c.device);
  FILE *f = fopen(path, "rb");
  if (!f) return fail(h, "gpu_read_conf: no stored configuration");
  std::string txt;
  char buf[1 << 16];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, f)) > 0) txt.append(buf, n);
  fclose(f);
  const char *p = txt.c_str();
  char *e = nullptr;
  long d[4];
  for (int k = 0; k < 4; k++) {
    d[k] = strtol(p, &e, 10);
    if (e == p) return fail(h, "gpu_read_conf: bad header");
    p = e;
  }
  if (d[0] != c.L.nx || d[1] != c.L.ny || d[2] != c.L.nz || d[3] != c.L.nt)
    return fail(h, "gpu_read_conf: stored configuration does not fit the lattice");
  double hv[3];
  for (int k = 0; k < 3; k++) {
    hv[k] = strtod(p, &e);
    if (e == p) return fail(h, "gpu_read_conf: bad header");
    p = e;
  }
  const long it = strtol(p, &e, 10);
  if (e == p) return fail(h, "gpu_read_conf: bad header");
  p = e;
  const long nl = 4 * c.V;
  std::vector<double> u((size_t)nl * 18);
  for (size_t q = 0; q < u.size(); q += 2) {
    while (*p && *p != '(') p++;
    if (!*p) return fail(h, "gpu_read_conf: truncated file");
    u[q] = strtod(p + 1, &e);
    if (*e != ',') return fail(h, "gpu_read_conf: bad complex number");
    u[q + 1] = strtod(e + 1, &e);
    if (*e != ')') return fail(h, "gpu_read_conf: bad complex number");
    p = e + 1;
  }
  if (const int rc = set_links(c, u.data(), "gpu_read_conf")) return rc;
  if (update_iteration) *update_iteration = (int)it;
  if (beta) *beta = hv[0];
  if (mass) *mass = hv[1];
  if (n_flavours) *n_flavours = hv[2];
  return RHMC_OK;
}

static double2 *fermion_buf(Context &c, int which, int *n) {
  switch (which) {
    case 0: *n = c.n_pf; return c.mf;
    case 1: *n = c.n_pf * (c.smf_order > 0 ? c.smf_order : c.max_order); return c.smf;
    case 2: *n = c.n_pf; return c.phi0;
    case 3: *n = c.n_pf; return c.chi;
    case 4: *n = c.n_pf; return c.phinew;
    default: *n = 0; return nullptr;
  }
}

int gpu_get_fermions(rhmc_gpu_t h, int which, double *out) {
  if (!h || !out) return RHMC_EINVAL;
  Context &c = h->c;
  hipSetDevice(c.device);
  int n;
  double2 *b = fermion_buf(c, which, &n);
  if (!b) return fail(h, "gpu_get_fermions: bad buffer id");
  return download_spinors(c, b, out, n) ? RHMC_EHIP : RHMC_OK;
}

int gpu_put_fermions(rhmc_gpu_t h, int which, const double *in) {
  if (!h || !in) return RHMC_EINVAL;
  Context &c = h->c;
  hipSetDevice(c.device);
  int n;
  double2 *b = fermion_buf(c, which, &n);
  if (!b) return fail(h, "gpu_put_fermions: bad buffer id");
  return upload_spinors(c, in, b, n) ? RHMC_EHIP : RHMC_OK;
}

int gpu_exp_su3(rhmc_gpu_t h, double step) {
  if (!h) return RHMC_EINVAL;
  hipSetDevice(h->c.device);
  if (exp_update(h->c, step)) return RHMC_EHIP;
  return hip_ok(h->c, hipStreamSynchronize(h->c.stream), "exp") ? RHMC_OK : RHMC_EHIP;
}

int gpu_momenta_sum_multiply(rhmc_gpu_t h, double step) {
  if (!h) return RHMC_EINVAL;
  hipSetDevice(h->c.device);
  if (momenta_update(h->c, step)) return RHMC_EHIP;
  return hip_ok(h->c, hipStreamSynchronize(h->c.stream), "mom") ? RHMC_OK : RHMC_EHIP;
}

int gpu_gauge_deriv(rhmc_gpu_t h, int ipdot_init) {
  if (!h) return RHMC_EINVAL;
  hipSetDevice(h->c.device);
  if (gauge_force(h->c, ipdot_init)) return RHMC_EHIP;
  return hip_ok(h->c, hipStreamSynchronize(h->c.stream), "gauge deriv") ? RHMC_OK : RHMC_EHIP;
}

int gpu_fermion_force(rhmc_gpu_t h, int num_shifts, const double *num, int accumulate) {
  if (!h || !num) return RHMC_EINVAL;
  Context &c = h->c;
  if (num_shifts != c.smf_order || num_shifts < 1) return fail(h, "gpu_fermion_force: order differs from the last solve");
  hipSetDevice(c.device);
  std::vector<double> coef((size_t)c.n_pf * num_shifts);
  for (int ps = 0; ps < c.n_pf; ps++)
    for (int k = 0; k < num_shifts; k++) coef[(size_t)ps * num_shifts + k] = num[k];
  return fermion_force(c, c.n_pf * num_shifts, coef.data(), c.smf, accumulate) ? RHMC_EHIP : RHMC_OK;
}

int gpu_fermion_force_noalloc(rhmc_gpu_t h, int num_shifts, const double *num, int accumulate) {
  if (!h || !num) return RHMC_EINVAL;
  Context &c = h->c;
  if ((size_t)sizeof(double2) * 3 * (size_t)c.Vh > kNoAllocScratchBytes - 256 * sizeof(double))
    return fail(h, "gpu_fermion_force_noalloc: one system (48 V/2 bytes) exceeds the 64 MiB fixed scratch");
  if (num_shifts != c.smf_order || num_shifts < 1)
    return fail(h, "gpu_fermion_force_noalloc: order differs from the last solve");
  hipSetDevice(c.device);
  const int rc = fermion_force_noalloc(c, num_shifts, num, accumulate);
  return rc == 0 ? RHMC_OK : rc == -2 ? RHMC_EINVAL : RHMC_EHIP;
}

int gpu_sol_sum(rhmc_gpu_t h, int num_shifts, double a0, const double *num) {
  if (!h || !num) return RHMC_EINVAL;
  Context &c = h->c;
  if (num_shifts != c.smf_order) return fail(h, "gpu_sol_sum: order differs from the last solve");
  hipSetDevice(c.device);
  return sol_sum(c, c.mf, c.smf, num_shifts, a0, num) ? RHMC_EHIP : RHMC_OK;
}

static int solve_status(int rc) {
  return rc == 0 ? RHMC_OK
         : rc == 1 ? RHMC_EINVAL
         : rc == kStatusMaxCg ? RHMC_EMAXCG
         : rc == kStatusNonFinite ? RHMC_ENONFINITE
                                 : RHMC_EHIP;
}

int gpu_fermionforce(rhmc_gpu_t h, double res, int relative, int order, const double *shifts, const double *num,
                     int precision, int max_cg, int pole_chunk, int *iterations) {
  if (!h || !shifts || !num || order < 1 || order > kMaxShifts) return RHMC_EINVAL;
  Context &c = h->c;
  if (pole_chunk < 1 || pole_chunk > order) pole_chunk = order;
  if (pole_chunk > c.max_order) return fail(h, "gpu_fermionforce: pole_chunk above max_order");
  hipSetDevice(c.device);
  int total = 0;
  for (int k0 = 0; k0 < order; k0 += pole_chunk) {
    const int nk = order - k0 < pole_chunk ? order - k0 : pole_chunk;
    int it = 0;
    const int rc = multishift_solve(c, c.mf, c.n_pf, nk, shifts + k0, res, relative, precision, max_cg, c.smf, &it);
    total += it;
    if (rc) return solve_status(rc);
    std::vector<double> coef((size_t)c.n_pf * nk);
    for (int ps = 0; ps < c.n_pf; ps++)
      for (int k = 0; k < nk; k++) coef[(size_t)ps * nk + k] = num[k0 + k];
    if (fermion_force(c, c.n_pf * nk, coef.data(), c.smf, k0 > 0)) return RHMC_EHIP;
    c.smf_order = nk;
  }
  if (iterations) *iterations = total;
  return hip_ok(c, hipStreamSynchronize(c.stream), "fermionforce") ? RHMC_OK : RHMC_EHIP;
}

int gpu_fermion_action(rhmc_gpu_t h, double res, int relative, int order, const double *shifts, double a0,
                       const double *num, int precision, int max_cg, double *action) {
  if (!h || !shifts || !num || !action || order < 1 || order > h->c.max_order) return RHMC_EINVAL;
  Context &c = h->c;
  hipSetDevice(c.device);
  int it = 0;
  const int rc = multishift_solve(c, c.mf, c.n_pf, order, shifts, res, relative, precision, max_cg, c.smf, &it);
  if (rc) return solve_status(rc);
  c.smf_order = order;
  const long sp = 3 * c.Vh;
  double2 *tmp = (double2 *)dev_alloc(c, sizeof(double2) * sp * c.n_pf);
  if (!tmp) return RHMC_EHIP;
  hipMemcpyAsync(tmp, c.mf, sizeof(double2) * sp * c.n_pf, hipMemcpyDeviceToDevice, c.stream);
  std::vector<double> d(c.n_pf);
  int err = sol_sum(c, tmp, c.smf, order, a0, num);
  if (!err) err = vec_dots(c, c.mf, tmp, c.n_pf, sp, d.data());
  dev_free(tmp);
  if (err) return RHMC_EHIP;
  double s = 0.0;
  for (int ps = 0; ps < c.n_pf; ps++) s += d[ps];
  *action = s;
  return RHMC_OK;
}

int gpu_find_max(rhmc_gpu_t h, double *mx, double tol, unsigned long long seed) {
  if (!h || !mx) return RHMC_EINVAL;
  Context &c = h->c;
  hipSetDevice(c.device);
  if (gauss_spinor(c, c.phinew, 1, seed, 7)) return RHMC_EHIP;
  int loops;
  return power_max(c, c.phinew, tol, mx, &loops) ? RHMC_EHIP : RHMC_OK;
}

int gpu_find_min(rhmc_gpu_t h, double *mn, double mx, double tol, unsigned long long seed) {
  if (!h || !mn) return RHMC_EINVAL;
  Context &c = h->c;
  hipSetDevice(c.device);
  if (gauss_spinor(c, c.phinew, 1, seed, 8)) return RHMC_EHIP;
  int loops;
  return power_min(c, c.phinew, tol, mx, mn, &loops) ? RHMC_EHIP : RHMC_OK;
}

int gpu_action(rhmc_gpu_t h, int init, int with_f) {
  if (!h || (init != 0 && init != -1)) return RHMC_EINVAL;
  Context &c = h->c;
  hipSetDevice(c.device);
  const double2 *phi = init == 0 ? c.phi0 : c.phinew;
  if (action_local(c, phi, c.chi, with_f, init)) return RHMC_EHIP;
  return hip_ok(c, hipStreamSynchronize(c.stream), "action") ? RHMC_OK : RHMC_EHIP;
}

int gpu_action_sum(rhmc_gpu_t h, double *sum) {
  if (!h || !sum) return RHMC_EINVAL;
  hipSetDevice(h->c.device);
  return action_sum(h->c, sum) ? RHMC_EHIP : RHMC_OK;
}

int gpu_gauss_momenta(rhmc_gpu_t h, unsigned long long seed, unsigned long long stream) {
  if (!h) return RHMC_EINVAL;
  hipSetDevice(h->c.device);
  if (gauss_momenta(h->c, seed, stream)) return RHMC_EHIP;
  return hip_ok(h->c, hipStreamSynchronize(h->c.stream), "gauss") ? RHMC_OK : RHMC_EHIP;
}

int gpu_gauss_phi(rhmc_gpu_t h, unsigned long long seed, unsigned long long stream) {
  if (!h) return RHMC_EINVAL;
  hipSetDevice(h->c.device);
  if (gauss_spinor(h->c, h->c.phi0, h->c.n_pf, seed, stream)) return RHMC_EHIP;
  return hip_ok(h->c, hipStreamSynchronize(h->c.stream), "gauss") ? RHMC_OK : RHMC_EHIP;
}

int gpu_dslash(rhmc_gpu_t h, int par_out, int precision, int nvec, const double *in, double *out) {
  if (!h || !in || !out || nvec < 1 || (par_out != 0 && par_out != 1)) return RHMC_EINVAL;
  Context &c = h->c;
  hipSetDevice(c.device);
  const long n = 3 * c.Vh * (long)nvec;
  double2 *a = nullptr, *b = nullptr;
  if (!hip_ok(c, hipMalloc(&a, sizeof(double2) * n), "malloc")) return RHMC_EHIP;
  if (!hip_ok(c, hipMalloc(&b, sizeof(double2) * n), "malloc")) { hipFree(a); return RHMC_EHIP; }
  int rc = RHMC_OK;
  if (upload_spinors(c, in, a, nvec)) rc = RHMC_EHIP;
  if (rc == RHMC_OK) {
    if (precision == 0) {
      launch_hop_d(c, par_out, a, b, nvec, 3 * c.Vh, 3 * c.Vh);
    } else {
      float2 *fa = nullptr, *fb = nullptr;
      hipMalloc(&fa, sizeof(float2) * n);
      hipMalloc(&fb, sizeof(float2) * n);
      const unsigned nb = (unsigned)((n + 255) / 256);
      hipLaunchKernelGGL(k_d2f, dim3(nb), dim3(256), 0, c.stream, fa, a, n);
      hop_s12(c, par_out, fa, fb, nvec);
      hipLaunchKernelGGL(k_f2d, dim3(nb), dim3(256), 0, c.stream, b, fb, n);
      hipStreamSynchronize(c.stream);
      hipFree(fa);
      hipFree(fb);
    }
    if (download_spinors(c, b, out, nvec)) rc = RHMC_EHIP;
  }
  hipFree(a);
  hipFree(b);
  return rc;
}

int gpu_time_dslash(rhmc_gpu_t h, int precision, int nvec, int reps, double *ms) {
  if (!h || !ms || nvec < 1 || reps < 1) return RHMC_EINVAL;
  Context &c = h->c;
  hipSetDevice(c.device);
  const long n = 3 * c.Vh * (long)nvec;
  if (precision < 0 || precision > 2) return RHMC_EINVAL;
  const size_t es = precision == 1 ? sizeof(float2) : sizeof(double2);
  void *a = nullptr, *b = nullptr;
  if (!hip_ok(c, hipMalloc(&a, es * n), "malloc")) return RHMC_EHIP;
  if (!hip_ok(c, hipMalloc(&b, es * n), "malloc")) { hipFree(a); return RHMC_EHIP; }
  hipMemsetAsync(a, 0, es * n, c.stream);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  auto launch = [&](int k) {
    const int par = k & 1;
    if (precision == 0)
      launch_hop_d(c, par, (const double2 *)((k & 1) ? b : a), (double2 *)((k & 1) ? a : b), nvec, 3 * c.Vh, 3 * c.Vh);
    else if (precision == 2)
      launch_hop_d12(c, par, (const double2 *)((k & 1) ? b : a), (double2 *)((k & 1) ? a : b), nvec, 3 * c.Vh,
                     3 * c.Vh);
    else
      hop_s12(c, par, (const float2 *)((k & 1) ? b : a), (float2 *)((k & 1) ? a : b), nvec);
  };
  for (int k = 0; k < 3; k++) launch(k);
  hipEventRecord(e0, c.stream);
  for (int k = 0; k < reps; k++) launch(k);
  hipEventRecord(e1, c.stream);
  hipEventSynchronize(e1);
  float t = 0.f;
  hipEventElapsedTime(&t, e0, e1);
  *ms = (double)t / reps;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(a);
  hipFree(b);
  return hip_ok(c, hipGetLastError(), "time dslash") ? RHMC_OK : RHMC_EHIP;
}

int gpu_gauge_meas(rhmc_gpu_t h, double *plaq_spatial, double *plaq_temporal, double *poly_re, double *poly_im) {
  if (!h || !plaq_spatial || !plaq_temporal || !poly_re || !poly_im) return RHMC_EINVAL;
  hipSetDevice(h->c.device);
  return gauge_meas(h->c, plaq_spatial, plaq_temporal, poly_re, poly_im) ? RHMC_EHIP : RHMC_OK;
}

int gpu_chiral_meas(rhmc_gpu_t h, const double *rnd_e, const double *rnd_o, unsigned long long seed, double res,
                    int max_cg, double n_flavours, double *out, int *iterations) {
  if (!h || !out || (!rnd_e) != (!rnd_o) || res <= 0.0 || h->c.mass == 0.0) return RHMC_EINVAL;
  hipSetDevice(h->c.device);
  return solve_status(chiral_meas(h->c, rnd_e, rnd_o, seed, res, max_cg, n_flavours, out, iterations));
}

int gpu_kernel_timing(rhmc_gpu_t h, int enable, double *mean_ms, long long *count) {
  if (!h) return RHMC_EINVAL;
  Context &c = h->c;
  hipSetDevice(c.device);
  if (mean_ms) {
    mean_ms[0] = c.kt_count ? c.kt_sum_ms / c.kt_count : 0.0;
    mean_ms[1] = c.kt_count2 ? c.kt_sum2_ms / c.kt_count2 : 0.0;
  }
  if (count) {
    count[0] = c.kt_count;
    count[1] = c.kt_count2;
  }
  if (enable >= 0) {
    if (enable && kernel_timing_clock(c)) return RHMC_EHIP;
    c.time_kernels = enable;
    c.kt_sum_ms = c.kt_sum2_ms = 0.0;
    c.kt_count = c.kt_count2 = 0;
  }
  return RHMC_OK;
}

__global__ void k_flush_fill(double4 *p, long n, double v) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = make_double4(v, v, v, v);
}

// read pass over n double4: every workgroup stores one sum (so the loads are
// kept), 32 KB of writes for a 4096-workgroup grid
__global__ __launch_bounds__(256) void k_flush_read(const double4 *p, long n, double *out) {
  const long stride = (long)gridDim.x * blockDim.x;
  double s = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double4 v = p[i];
    s += (v.x + v.y) + (v.z + v.w);
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  __shared__ double sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

int gpu_time_dslash_cold(rhmc_gpu_t h, int precision, int nvec, int reps, double flush_bytes, double *ms) {
  if (!h || !ms || nvec < 1 || reps < 1 || flush_bytes < 0 || flush_bytes > 8e9) return RHMC_EINVAL;
  Context &c = h->c;
  hipSetDevice(c.device);
  const long n = 3 * c.Vh * (long)nvec;
  if (precision < 0 || precision > 2) return RHMC_EINVAL;
  const size_t es = precision == 1 ? sizeof(float2) : sizeof(double2);
  const long nf = (long)(flush_bytes / sizeof(double4));
  constexpr int kFlushBlocks = 4096;
  void *a = nullptr, *b = nullptr, *f = nullptr;
  if (!hip_ok(c, hipMalloc(&a, es * n), "malloc")) return RHMC_EHIP;
  if (!hip_ok(c, hipMalloc(&b, es * n), "malloc")) { hipFree(a); return RHMC_EHIP; }
  if (nf > 0 && !hip_ok(c, hipMalloc(&f, nf * sizeof(double4) + sizeof(double) * kFlushBlocks), "malloc")) {
    hipFree(a);
    hipFree(b);
    return RHMC_EHIP;
  }
  hipMemsetAsync(a, 0, es * n, c.stream);
  // the eviction buffer is written once, before the timed launches
  if (nf > 0) hipLaunchKernelGGL(k_flush_fill, dim3(kFlushBlocks), dim3(256), 0, c.stream, (double4 *)f, nf, 1.0);
  std::vector<hipEvent_t> ev(2 * reps);
  for (auto &e : ev) hipEventCreate(&e);
  auto launch = [&](int k) {
    const int par = k & 1;
    if (precision == 0)
      launch_hop_d(c, par, (const double2 *)((k & 1) ? b : a), (double2 *)((k & 1) ? a : b), nvec, 3 * c.Vh, 3 * c.Vh);
    else if (precision == 2)
      launch_hop_d12(c, par, (const double2 *)((k & 1) ? b : a), (double2 *)((k & 1) ? a : b), nvec, 3 * c.Vh,
                     3 * c.Vh);
    else
      hop_s12(c, par, (const float2 *)((k & 1) ? b : a), (float2 *)((k & 1) ? a : b), nvec);
  };
  launch(0);
  // each timed launch follows a read pass over flush_bytes of other memory,
  // which evicts the operands from the L2s and the Infinity Cache without
  // leaving dirty lines for the timed launch to write back
  for (int k = 0; k < reps; k++) {
    if (nf > 0)
      hipLaunchKernelGGL(k_flush_read, dim3(kFlushBlocks), dim3(256), 0, c.stream, (const double4 *)f, nf,
                         (double *)((double4 *)f + nf));
    hipEventRecord(ev[2 * k], c.stream);
    launch(k);
    hipEventRecord(ev[2 * k + 1], c.stream);
  }
  hipStreamSynchronize(c.stream);
  double tot = 0.0;
  for (int k = 0; k < reps; k++) {
    float t = 0.f;
    hipEventElapsedTime(&t, ev[2 * k], ev[2 * k + 1]);
    tot += t;
  }
  *ms = tot / reps;
  for (auto &e : ev) hipEventDestroy(e);
  hipFree(a);
  hipFree(b);
  if (f) hipFree(f);
  return hip_ok(c, hipGetLastError(), "time dslash") ? RHMC_OK : RHMC_EHIP;
}

int cu_multips_shifted_invert(rhmc_gpu_t h, double res, int relative, int order, const double *shifts, int precision,
                              int max_cg, int *iterations) {
  if (!h || !shifts || order < 1 || order > h->c.max_order) return RHMC_EINVAL;
  Context &c = h->c;
  hipSetDevice(c.device);
  int it = 0;
  int rc = multishift_solve(c, c.mf, c.n_pf, order, shifts, res, relative, precision, max_cg, c.smf, &it);
  c.smf_order = order;
  if (iterations) *iterations = it;
  return solve_status(rc);
}

int gpu_shifted_inverter(rhmc_gpu_t h, double res, const double *shifts, int num_shifts, int psferm, int precision,
                          int *ncount) {
  if (!h || !shifts || num_shifts < 1 || num_shifts > h->c.max_order || psferm < 0 || psferm >= h->c.n_pf)
    return RHMC_EINVAL;
  Context &c = h->c;
  hipSetDevice(c.device);
  int it = 0;
  const long sp = 3 * c.Vh;
  int rc = multishift_solve(c, c.mf + psferm * sp, 1, num_shifts, shifts, res, 0, precision, 100000,
                            c.smf + (long)psferm * num_shifts * sp, &it);
  c.smf_order = num_shifts;
  if (ncount) *ncount = it;
  return solve_status(rc);
}

int cu_shifted_residuals(rhmc_gpu_t h, int order, const double *shifts, double *out) {
  if (!h || !shifts || !out || order != h->c.smf_order) return RHMC_EINVAL;
  Context &c = h->c;
  hipSetDevice(c.device);
  const int nsys = c.n_pf * order;
  std::vector<double> sh(nsys), n2(nsys);
  for (int j = 0; j < nsys; j++) sh[j] = shifts[j % order];
  double2 *r = nullptr;
  if (!hip_ok(c, hipMalloc(&r, sizeof(double2) * 3 * c.Vh * nsys), "malloc")) return RHMC_EHIP;
  int rc = residual_d(c, c.mf, order, c.smf, sh.data(), nsys, r, n2.data());
  hipFree(r);
  if (rc) return RHMC_EHIP;
  for (int j = 0; j < nsys; j++) out[j] = sqrt(n2[j]);
  return RHMC_OK;
}

int cu_last_solve_stats(rhmc_gpu_t h, int *i32, int *iref, int *cyc) {
  if (!h) return RHMC_EINVAL;
  if (i32) *i32 = h->c.last_iters_fp32;
  if (iref) *iref = h->c.last_iters_refine;
  if (cyc) *cyc = h->c.last_refine_cycles;
  return RHMC_OK;
}

int cu_solver_counters(rhmc_gpu_t h, int *restarts, int *graph_captures, double *true_res_max) {
  if (!h) return RHMC_EINVAL;
  if (restarts) *restarts = h->c.last_restarts;
  if (graph_captures) *graph_captures = h->c.graph_captures;
  if (true_res_max) *true_res_max = h->c.last_true_res_max;
  return RHMC_OK;
}

int gpu_time_copy(rhmc_gpu_t h, double bytes, int reps, double *ms) {
  if (!h || !ms || bytes < 1024 || reps < 1) return RHMC_EINVAL;
  hipSetDevice(h->c.device);
  return time_copy(h->c, (size_t)bytes, reps, ms) ? RHMC_EHIP : RHMC_OK;
}

int gpu_tuning_table(rhmc_gpu_t h, char *buf, int buflen) {
  if (!h || !buf || buflen < 1) return -1;
  const std::string t = tuning_table(h->c);
  const int n = (int)std::min<size_t>(t.size(), (size_t)buflen - 1);
  memcpy(buf, t.data(), n);
  buf[n] = 0;
  return n;
}

int gpu_device_info(rhmc_gpu_t h, char *buf, int buflen) {
  if (!h || !buf || buflen < 1) return RHMC_EINVAL;
  hipDeviceProp_t pr;
  if (hipGetDeviceProperties(&pr, h->c.device) != hipSuccess) return RHMC_EHIP;
  int lds_cu = 0;
  hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, h->c.device);
  char t[1024];
  const int n = snprintf(t, sizeof t,
                         "{\"name\": \"%s\", \"arch\": \"%s\", \"compute_units\": %d, \"wavefront\": %d, "
                         "\"hbm_bytes\": %zu, \"memory_clock_khz\": %d, \"memory_bus_bits\": %d, "
                         "\"l2_bytes\": %d, \"lds_bytes_per_cu\": %d, \"lds_bytes_per_workgroup\": %zu, "
                         "\"clock_khz\": %d}",
                         pr.name, pr.gcnArchName, pr.multiProcessorCount, pr.warpSize, pr.totalGlobalMem,
                         pr.memoryClockRate, pr.memoryBusWidth, pr.l2CacheSize, lds_cu, pr.sharedMemPerBlock,
                         pr.clockRate);
  if (n < 0 || n >= (int)sizeof t) return RHMC_EINVAL;
  const int m = std::min(n, buflen - 1);
  memcpy(buf, t, m);
  buf[m] = 0;
  return RHMC_OK;
}

int gpu_trajectory(rhmc_gpu_t h, const rhmc_traj_params *p, rhmc_traj_result *r) {
  if (!h || !p || !r) return RHMC_EINVAL;
  hipSetDevice(h->c.device);
  return run_trajectory(h->c, *p, *r);
}

int gpu_rev_update(rhmc_gpu_t h, const rhmc_traj_params *p, double *delta_conf, double *delta_mom,
                   rhmc_traj_result *r) {
  if (!h || !p || !r || !delta_conf || !delta_mom) return RHMC_EINVAL;
  hipSetDevice(h->c.device);
  return run_rev_update(h->c, *p, *r, delta_conf, delta_mom);
}

int gpu_multi_trajectory(rhmc_gpu_t *hs, int n, const rhmc_traj_params *ps, rhmc_traj_result *rs, int *status) {
  if (!hs || !ps || !rs || n < 1) return RHMC_EINVAL;
  for (int i = 0; i < n; i++) {
    if (!hs[i]) return RHMC_EINVAL;
    for (int k = 0; k < i; k++)
      if (hs[i] == hs[k]) return RHMC_EINVAL;
  }
  // Lattices with the same integrator, MD steps, precision dispatch and
  // residuals (each with its own extents, pseudofermions, rational
  // functions, bounds and seed) advance in lockstep on one device: every
  // solver step is one launch over all of their tiles.  Others: one host
  // thread per lattice, each sequencing its own stream.
  std::vector<const rhmc_traj_params *> pv(n);
  bool same_device = true;
  for (int i = 0; i < n; i++) {
    pv[i] = &ps[i];
    same_device = same_device && hs[i]->c.device == hs[0]->c.device;
  }
  if (same_device && trajectories_lockstep(pv)) {
    hipSetDevice(hs[0]->c.device);
    std::vector<Context *> cs(n);
    std::vector<rhmc_traj_result *> rv(n);
    for (int i = 0; i < n; i++) {
      cs[i] = &hs[i]->c;
      rv[i] = &rs[i];
    }
    std::vector<int> st;
    const int rc = run_trajectory_group(cs, pv, rv, st);
    if (status)
      for (int i = 0; i < n; i++) status[i] = st[i];
    return rc;
  }
  std::vector<int> rc(n, RHMC_OK);
  std::vector<std::thread> th;
  th.reserve(n);
  for (int i = 0; i < n; i++)
    th.emplace_back([&, i]() {
      hipSetDevice(hs[i]->c.device);
      rc[i] = run_trajectory(hs[i]->c, ps[i], rs[i]);
    });
  for (auto &t : th) t.join();
  int worst = RHMC_OK;
  for (int i = 0; i < n; i++) {
    if (status) status[i] = rc[i];
    if (rc[i] != RHMC_OK && worst == RHMC_OK) worst = rc[i];
  }
  return worst;
}

}  // extern "C"
