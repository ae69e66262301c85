// fields.hip -- gauge-field side of the RHMC molecular dynamics on gfx950:
// fermion force, gauge force, exp(i eps P) link update, momentum update,
// per-site action / Delta H, rational-function sums, RNG and the host-layout
// conversions of the C ABI.  All arithmetic here is FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <vector>

#include "context.hpp"

namespace rhmc {

static inline unsigned nblk(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

__device__ __forceinline__ Mat<double> link_at(const double2 *U, const LatticeDims &L, int x, int y, int z, int t,
                                               int mu) {
  const int par = (x + y + z + t) & 1;
  return load_link_d18(U, L.Vh, mu, par, half_index(L, x, y, z, t));
}

__device__ __forceinline__ void shift_coord(const LatticeDims &L, int &x, int &y, int &z, int &t, int mu, int d) {
  if (mu == 0) x = (x + d + L.nx) % L.nx;
  else if (mu == 1) y = (y + d + L.ny) % L.ny;
  else if (mu == 2) z = (z + d + L.nz) % L.nz;
  else t = (t + d + L.nt) % L.nt;
}

// --------------------------------------------------------------- fermion force
// fermionforce.cc:46-104 + ipdot.cc:83-155:
//   F_mu(e) =  sum_j a_j h_j(e+mu) (x) s_j(e)^*     (even sites)
//   F_mu(o) = -sum_j a_j s_j(o+mu) (x) h_j(o)^*     (odd sites)
//   ipdot_mu(x) = TA(U_mu(x) F_mu(x))   [+= existing ipdot if accumulate]
// s_j: shifted solutions (even), h_j = D_oe s_j (odd); one thread per site
// holds the four F_mu in registers (FP64) and reads each spinor of its site
// once.  C = double2 (FP64 solutions) or float2 (FP32 solutions of the FP32
// MD solves).
// One thread per site and system at a time: the 15 loads of a system
// (uniform base + 32-bit lane offset) are issued before its multiply-adds.
// Taking two systems per step keeps more loads in flight but needs ~280
// VGPRs (one wave per SIMD) and measured slower.  So did two lanes per site
// (two F_mu each, loads of system j+1 issued before system j's multiply-
// adds, 166 VGPRs): 522 against 415 us for 128 FP32 systems at 32^3x8 (both
// with the earlier parity-major grid).  The kernel is bound by load-
// instruction issue, and that split issues 18 loads per site and system
// instead of 15.
template <typename C, int NQ>
__device__ __forceinline__ void ff_systems(const C *own_v, const C *nb_v, long Vh, int j0, unsigned ho,
                                           const unsigned (&nbo)[4], const double *coef, double sgn,
                                           Mat<double> (&F)[4]) {
  C own[NQ][3], w[NQ][4][3];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const long base = (long)(j0 + q) * 3 * Vh;
#pragma unroll
    for (int k = 0; k < 3; k++) own[q][k] = ld_off(own_v + base + (long)k * Vh, ho);
#pragma unroll
    for (int mu = 0; mu < 4; mu++)
#pragma unroll
      for (int i = 0; i < 3; i++) w[q][mu][i] = ld_off(nb_v + base + (long)i * Vh, nbo[mu]);
  }
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const double a = sgn * coef[j0 + q];
    double2 o[3];
#pragma unroll
    for (int k = 0; k < 3; k++) o[k] = make_double2((double)own[q][k].x, -(double)own[q][k].y);  // conjugate
#pragma unroll
    for (int mu = 0; mu < 4; mu++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const double2 v = make_double2(a * (double)w[q][mu][i].x, a * (double)w[q][mu][i].y);
#pragma unroll
        for (int k = 0; k < 3; k++) cfma(F[mu].e[i][k], v, o[k]);
      }
  }
}

// Grid: both parities of each kBlock-site tile, consecutive on one XCD, the
// tiles in the XCD-compact order of the Dslash (tiles, or lattice order when
// null): the even sites read the odd spinors that the odd sites of the same
// tile read as their own, and the +mu neighbours stay in the XCD's share, so
// more of a spinor's five uses find it in that XCD's L2 (FP64, 128 systems
// at 32^3x8: 5.84 -> 4.36 GB fetched per call, 816 -> 648 us).
template <typename C>
__global__ __launch_bounds__(kBlock) void k_fermion_force(LatticeDims L, const int *tiles, const double2 *U,
                                                          const C *s, const C *hv, const double *coef, int nsys,
                                                          double2 *ipdot, int accumulate) {
  const long Vh = L.Vh;
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const int par = (int)(b & 1u);
  const int q = (int)(b >> 1);
  const int h = (tiles ? tiles[q] : q) * kBlock + (int)threadIdx.x;
  if (h >= (int)Vh) return;
  const SiteCoord c = site_coords_i(L, par, h);
  unsigned nbo[4];
#pragma unroll
  for (int mu = 0; mu < 4; mu++) nbo[mu] = (unsigned)neighbour_i(L, c, mu, +1) * (unsigned)sizeof(C);
  const unsigned ho = (unsigned)h * (unsigned)sizeof(C);
  Mat<double> F[4];
#pragma unroll
  for (int mu = 0; mu < 4; mu++)
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) F[mu].e[i][k] = make_double2(0.0, 0.0);
  const C *own_v = par == 0 ? s : hv;
  const C *nb_v = par == 0 ? hv : s;
  const double sgn = par == 0 ? 1.0 : -1.0;
  for (int j = 0; j < nsys; j++) ff_systems<C, 1>(own_v, nb_v, Vh, j, ho, nbo, coef, sgn, F);
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    const Mat<double> Um = load_link_d18(U, Vh, mu, par, h);
    Mat<double> T = ta(mmul(Um, F[mu]));
    if (accumulate) {
      const Mat<double> old = load_aherm(ipdot, Vh, mu, par, h);
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) T.e[i][k] = cadd(T.e[i][k], old.e[i][k]);
    }
    store_aherm(ipdot, Vh, mu, par, h, T);
  }
}

// The force's coefficients in device memory: uploaded only when they differ
// from the ones the buffer already holds (every MD force of a trajectory
// uses the same ones), so the force launches follow the solve without a copy.
static bool force_coef_upload(Context &c, double *dcoef, const double *num, int nsys) {
  std::vector<double> &held = c.force_coef_held;
  if (c.force_coef_dev == dcoef && held.size() == (size_t)nsys && std::equal(held.begin(), held.end(), num))
    return true;
  c.force_coef_dev = nullptr;
  if (!h2d(c, dcoef, num, sizeof(double) * nsys)) return false;
  held.assign(num, num + nsys);
  c.force_coef_dev = dcoef;
  return true;
}

int fermion_force(Context &c, int nsys, const double *num, const double2 *x, int accumulate) {
  // nsys = n_pf * order systems, x = [n_pf][order][3][Vh]; num has nsys entries
  const long Vh = c.Vh;
  size_t need = sizeof(double2) * 3 * Vh * (size_t)nsys;
  // h_j = D_oe s_j for all systems (one multi-RHS hop sharing link loads),
  // into workspace buffers kept for the context's lifetime
  double2 *hv = (double2 *)scratch_get(c, SCR_FORCE_HOP, need);
  double *coef = (double *)scratch_get(c, SCR_FORCE_COEF, sizeof(double) * 256);
  if (!hv || !coef || nsys > 256) return -1;
  if (!force_coef_upload(c, coef, num, nsys)) return -1;
  launch_hop_d(c, 1, x, hv, nsys, 3 * Vh, 3 * Vh);
  hipLaunchKernelGGL(k_fermion_force<double2>, dim3(2 * nblk(Vh)), dim3(kBlock), 0, c.stream, c.L, c.tiles, c.U64, x, hv, coef,
                     nsys, c.ipdot, accumulate);
  // stream-ordered for the caller (the coefficients are staged in pinned memory)
  return hip_ok(c, hipGetLastError(), "fermion force") ? 0 : -1;
}

int fermion_force32(Context &c, int nsys, const double *num, const float2 *x32, int accumulate) {
  const long Vh = c.Vh;
  float2 *hv = (float2 *)scratch_get(c, SCR_FORCE_HOP, sizeof(float2) * 3 * Vh * (size_t)nsys);
  double *coef = (double *)scratch_get(c, SCR_FORCE_COEF, sizeof(double) * 256);
  if (!hv || !coef || nsys > 256) return -1;
  if (!force_coef_upload(c, coef, num, nsys)) return -1;
  launch_hop_s(c, 1, x32, hv, nsys, 3 * Vh, 3 * Vh);
  hipLaunchKernelGGL(k_fermion_force<float2>, dim3(2 * nblk(Vh)), dim3(kBlock), 0, c.stream, c.L, c.tiles, c.U64, x32, hv, coef,
                     nsys, c.ipdot, accumulate);
  // stream-ordered for the caller (the coefficients are staged in pinned memory)
  return hip_ok(c, hipGetLastError(), "fermion force") ? 0 : -1;
}

// ------------------------------------------------ allocation-free fermion force
// For callers that cannot allocate device memory (graph capture, tight
// serving loops): the force's scratch -- the coefficients and the hop
// results h_j = D_oe s_j -- lives in one fixed file-scope device array.
// Every GPU holds its own copy of it until the process ends.  In host code
// the name `scratch` is only a stand-in for the symbol: kernels are given
// the per-device address that hipGetSymbolAddress returned.
constexpr size_t kNoAllocCoefBytes = 256 * sizeof(double);  // 2 KiB, keeps h 256-byte aligned
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDeviceOrdinal = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDeviceOrdinal];

int noalloc_scratch_init(Context &c) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  c.noalloc_scratch = nullptr;
  if (c.device < 0 || c.device >= kMaxDeviceOrdinal) return -1;
  if (!g_scratch_ptr[c.device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return -1;
    }
    g_scratch_ptr[c.device] = p;
  }
  c.noalloc_scratch = g_scratch_ptr[c.device];
  return 0;
}

// the coefficients travel by value in the kernel arguments, so nothing is
// staged through host memory
struct CoefTable {
  double v[256];
};
__global__ void k_coef_store(CoefTable t, int n, double *dst) {
  const int i = threadIdx.x;
  if (i < n) dst[i] = t.v[i];
}

int fermion_force_noalloc(Context &c, int order, const double *num, int accumulate) {
  const long Vh = c.Vh;
  const int nsys = c.n_pf * order;
  const size_t per_sys = sizeof(double2) * 3 * (size_t)Vh;
  const size_t room = kNoAllocScratchBytes - kNoAllocCoefBytes;
  if (per_sys > room) {
    c.last_error = "gpu_fermion_force_noalloc: one system (48 V/2 bytes) exceeds the 64 MiB fixed scratch";
    return -2;
  }
  if (!c.noalloc_scratch) {
    c.last_error = "gpu_fermion_force_noalloc: device address of the fixed scratch unavailable";
    return -1;
  }
  if (nsys < 1 || nsys > 256) return -2;
  CoefTable t;
  for (int j = 0; j < nsys; j++) t.v[j] = num[j % order];  // system j = ps * order + k
  double *coef = (double *)c.noalloc_scratch;
  double2 *hv = (double2 *)((unsigned char *)c.noalloc_scratch + kNoAllocCoefBytes);
  hipLaunchKernelGGL(k_coef_store, dim3(1), dim3(256), 0, c.stream, t, nsys, coef);
  // groups of systems whose hop results fit the scratch; the force of each
  // group is added to the previous groups' (TA and U F are linear)
  const int group = (int)std::min<size_t>((size_t)nsys, room / per_sys);
  for (int j0 = 0; j0 < nsys; j0 += group) {
    const int n = std::min(group, nsys - j0);
    const double2 *x = c.smf + (long)j0 * 3 * Vh;
    launch_hop_d_fixed(c, 1, x, hv, n, 3 * Vh, 3 * Vh);
    hipLaunchKernelGGL(k_fermion_force<double2>, dim3(2 * nblk(Vh)), dim3(kBlock), 0, c.stream, c.L, c.tiles, c.U64, x, hv,
                       coef + j0, n, c.ipdot, (j0 > 0 || accumulate) ? 1 : 0);
  }
  return hip_ok(c, hipGetLastError(), "fermion force (fixed scratch)") ? 0 : -1;
}

// ----------------------------------------------------------------- gauge force
// staples.cc:16-74, ipdot.cc:25-79:
//   ipdot_mu(x) = TA(U_mu(x) beta/3 sum_{nu!=mu} [U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
//                                              + U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)])
// one step along mu with the periodic wrap, no division
__device__ __forceinline__ void step_coord(const LatticeDims &L, int &x, int &y, int &z, int &t, int mu, int d) {
  if (mu == 0) x = d > 0 ? (x + 1 == L.nx ? 0 : x + 1) : (x == 0 ? L.nx - 1 : x - 1);
  else if (mu == 1) y = d > 0 ? (y + 1 == L.ny ? 0 : y + 1) : (y == 0 ? L.ny - 1 : y - 1);
  else if (mu == 2) z = d > 0 ? (z + 1 == L.nz ? 0 : z + 1) : (z == 0 ? L.nz - 1 : z - 1);
  else t = d > 0 ? (t + 1 == L.nt ? 0 : t + 1) : (t == 0 ? L.nt - 1 : t - 1);
}

// link (x, mu) from the FP64 12-real copy, row 2 rebuilt with the link's eta
__device__ __forceinline__ Mat<double> link12_at(const double2 *U12, const LatticeDims &L, int x, int y, int z, int t,
                                                 int mu) {
  return load_link_d12(U12, L.Vh, mu, (x + y + z + t) & 1, half_index_i(L, x, y, z, t),
                       eta_phase(L, x, y, z, t, mu));
}

// staples.cc:16-74 + ipdot.cc:25-79: ipdot_mu(x) = TA(U_mu(x) beta/3 sum_nu
// [U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ + U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)]).
// Links from the FP64 12-real copy (96 instead of 144 bytes per load; each
// thread loads 19), neighbours by compare-and-wrap in 32-bit arithmetic.
// Grid: the eight (parity, mu) workgroups of each kBlock-site tile
// consecutive on one XCD, the tiles in the XCD-compact order of the Dslash
// (tiles, or lattice order when null), so the links a tile's staples share
// are fetched into that XCD's L2 together.
__global__ __launch_bounds__(kBlock) void k_gauge_force(LatticeDims L, const int *tiles, const double2 *U12,
                                                        double b3, double2 *ipdot, int accumulate) {
  const long Vh = L.Vh;
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const int mu = (int)(b & 3u);
  const int par = (int)((b >> 2) & 1u);
  const int q = (int)(b >> 3);
  const int h = (tiles ? tiles[q] : q) * kBlock + (int)threadIdx.x;
  if (h >= (int)Vh) return;
  const SiteCoord c = site_coords_i(L, par, h);
  Mat<double> st;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) st.e[i][k] = make_double2(0.0, 0.0);
  int x1 = c.x, y1 = c.y, z1 = c.z, t1 = c.t;  // x + mu
  step_coord(L, x1, y1, z1, t1, mu, +1);
  for (int nu = 0; nu < 4; nu++) {
    if (nu == mu) continue;
    // upper: U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
    int x2 = c.x, y2 = c.y, z2 = c.z, t2 = c.t;
    step_coord(L, x2, y2, z2, t2, nu, +1);
    Mat<double> a = mmul_nd(link12_at(U12, L, x1, y1, z1, t1, nu), link12_at(U12, L, x2, y2, z2, t2, mu));
    a = mmul_nd(a, link12_at(U12, L, c.x, c.y, c.z, c.t, nu));
    // lower: U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)
    int x3 = c.x, y3 = c.y, z3 = c.z, t3 = c.t;
    step_coord(L, x3, y3, z3, t3, nu, -1);
    int x4 = x1, y4 = y1, z4 = z1, t4 = t1;
    step_coord(L, x4, y4, z4, t4, nu, -1);
    const Mat<double> l1 = link12_at(U12, L, x4, y4, z4, t4, nu);
    const Mat<double> l2 = link12_at(U12, L, x3, y3, z3, t3, mu);
    Mat<double> b = mmul_dn(l1, mdag(l2));
    b = mmul(b, link12_at(U12, L, x3, y3, z3, t3, nu));
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) st.e[i][k] = cadd(st.e[i][k], cadd(a.e[i][k], b.e[i][k]));
  }
  const Mat<double> Um = link12_at(U12, L, c.x, c.y, c.z, c.t, mu);
  Mat<double> T = mmul(Um, st);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) T.e[i][k] = cscale(T.e[i][k], b3);
  T = ta(T);
  if (accumulate) {
    const Mat<double> old = load_aherm(ipdot, Vh, mu, par, h);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) T.e[i][k] = cadd(T.e[i][k], old.e[i][k]);
  }
  store_aherm(ipdot, Vh, mu, par, h, T);
}

int gauge_force(Context &c, int accumulate) {
  hipLaunchKernelGGL(k_gauge_force, dim3(8 * nblk(c.Vh)), dim3(kBlock), 0, c.stream, c.L, c.tiles, c.U64c,
                     c.beta / 3.0, c.ipdot, accumulate);
  return hip_ok(c, hipGetLastError(), "gauge force") ? 0 : -1;
}

// ------------------------------------------------------------ link update
// momenta.cc:42-78 + su3.cc:187-219 + conf.cc:235-251:
//   U <- Pi_SU3[ exp6(i step P) ] U, then Gram-Schmidt and row 2 *= eta.
// Also refreshes the FP32 12-real copy used by the mixed-precision Dslash.
__global__ __launch_bounds__(kBlock) void k_exp_update(LatticeDims L, const double2 *P, double step, double2 *U,
                                                       float4 *U32, double2 *U64c) {
  const long Vh = L.Vh;
  const int bx = xcd_block(blockIdx.x, gridDim.x);
  const long h = (long)bx * kBlock + threadIdx.x;
  const int par = blockIdx.y, mu = blockIdx.z;
  if (h >= Vh) return;
  const Mat<double> Pm = load_herm(P, Vh, mu, par, h);
  Mat<double> X;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) X.e[i][k] = make_double2(-step * Pm.e[i][k].y, step * Pm.e[i][k].x);
  Mat<double> E, A = X;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) E.e[i][k] = make_double2((i == k) ? 1.0 : 0.0, 0.0);
  const double cf[6] = {1.0, 0.5, 1.0 / 3.0, 0.25, 0.2, 1.0 / 6.0};
#pragma unroll
  for (int n = 0; n < 6; n++) {
    if (n > 0) {
      A = mmul(A, X);
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) A.e[i][k] = cscale(A.e[i][k], cf[n]);
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) E.e[i][k] = cadd(E.e[i][k], A.e[i][k]);
  }
  sunitarize(E);
  Mat<double> Un = mmul(E, load_link_d18(U, Vh, mu, par, h));
  sunitarize(Un);
  const SiteCoord c = site_coords(L, par, h);
  const double eta = eta_phase(L, c.x, c.y, c.z, c.t, mu);
#pragma unroll
  for (int k = 0; k < 3; k++) Un.e[2][k] = cscale(Un.e[2][k], eta);
  store_link_d18(U, Vh, mu, par, h, Un);
  store_link_s12(U32, Vh, mu, par, h, Un);
  store_link_d12(U64c, Vh, mu, par, h, Un);
}

int exp_update(Context &c, double step) {
  hipLaunchKernelGGL(k_exp_update, dim3(nblk(c.Vh), 2, 4), dim3(kBlock), 0, c.stream, c.L, c.P, step, c.U64, c.U32,
                     c.U64c);
  return hip_ok(c, hipGetLastError(), "exp update") ? 0 : -1;
}

__global__ __launch_bounds__(kBlock) void k_refresh32(LatticeDims L, const double2 *U, float4 *U32, double2 *U64c) {
  const long Vh = L.Vh;
  const long h = (long)blockIdx.x * kBlock + threadIdx.x;
  const int par = blockIdx.y, mu = blockIdx.z;
  if (h >= Vh) return;
  const Mat<double> m = load_link_d18(U, Vh, mu, par, h);
  store_link_s12(U32, Vh, mu, par, h, m);
  store_link_d12(U64c, Vh, mu, par, h, m);
}

// the FP32 12-real and FP64 12-real copies of the FP64 links
int refresh_links32(Context &c) {
  hipLaunchKernelGGL(k_refresh32, dim3(nblk(c.Vh), 2, 4), dim3(kBlock), 0, c.stream, c.L, c.U64, c.U32, c.U64c);
  return hip_ok(c, hipGetLastError(), "refresh links32") ? 0 : -1;
}

// -------------------------------------------------------- momentum update
// momenta.cc:82-115: P <- P + (i step) ipdot on the packed 8-real fields
__global__ __launch_bounds__(kBlock) void k_mom_update(double2 *P, const double2 *A, double s, long Vh, long n) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int slot = (int)((i / Vh) & 3);
  double2 p = P[i];
  const double2 a = A[i];
  if (slot < 3) {
    p.x -= s * a.y;
    p.y += s * a.x;
  } else {
    p.x -= s * a.x;
    p.y -= s * a.y;
  }
  P[i] = p;
}

int momenta_update(Context &c, double step) {
  const long n = 32 * c.Vh;
  hipLaunchKernelGGL(k_mom_update, dim3(nblk(n)), dim3(kBlock), 0, c.stream, c.P, c.ipdot, step, c.Vh, n);
  return hip_ok(c, hipGetLastError(), "momenta update") ? 0 : -1;
}

// P <- s P (revupdate.cc:53, reverse_momenta with s = -1)
__global__ __launch_bounds__(kBlock) void k_scale_inplace(double2 *v, double s, long n) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double2 p = v[i];
  v[i] = make_double2(s * p.x, s * p.y);
}

int momenta_scale(Context &c, double s) {
  const long n = 32 * c.Vh;
  hipLaunchKernelGGL(k_scale_inplace, dim3(nblk(n)), dim3(kBlock), 0, c.stream, c.P, s, n);
  return hip_ok(c, hipGetLastError(), "momenta scale") ? 0 : -1;
}

// ------------------------------------------------------------ solution sums
// rationalapprox.cc:269-369 / cuda_sol_sum: io[ps] = a0 io[ps] + sum_k a_k x[ps][k]
__global__ __launch_bounds__(kBlock) void k_sol_sum(double2 *io, const double2 *x, int nshift, double a0,
                                                    const double *num, long n3) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  const int ps = blockIdx.y;
  if (i >= n3) return;
  double2 v = io[(long)ps * n3 + i];
  v.x *= a0;
  v.y *= a0;
  for (int k = 0; k < nshift; k++) {
    const double2 xv = x[((long)ps * nshift + k) * n3 + i];
    v.x += num[k] * xv.x;
    v.y += num[k] * xv.y;
  }
  io[(long)ps * n3 + i] = v;
}

int sol_sum(Context &c, double2 *io, const double2 *x, int nshift, double a0, const double *num) {
  double *d = (double *)scratch_get(c, SCR_SOLSUM_COEF, sizeof(double) * kMaxShifts);
  if (!d || nshift > kMaxShifts) return -1;
  std::vector<double> h(num, num + nshift);
  if (!h2d(c, d, h.data(), sizeof(double) * nshift)) return -1;
  const long n3 = 3 * c.Vh;
  hipLaunchKernelGGL(k_sol_sum, dim3(nblk(n3), c.n_pf), dim3(kBlock), 0, c.stream, io, x, nshift, a0, d, n3);
  return hip_ok(c, hipStreamSynchronize(c.stream), "sol_sum") ? 0 : -1;
}

// ------------------------------------------------------------------- action
// action.cc:6-156: per-site contributions in the reference's order
//   init 0 : v  = sum_plaq beta/3 Re tr(plaq) + sum_mu 1/2 Re tr P^2 + sum_ps |phi|^2
//   init -1: v -= the same with Re<phi_new, chi> for the fermions
__global__ __launch_bounds__(kBlock) void k_action(LatticeDims L, const double2 *U, const double2 *P, double b3,
                                                   const double2 *phi, const double2 *chi, int npf, int with_f,
                                                   double *value, int init) {
  const long Vh = L.Vh;
  const long h = (long)blockIdx.x * kBlock + threadIdx.x;
  const int par = blockIdx.y;
  if (h >= Vh) return;
  const long idx = (long)par * Vh + h;
  const SiteCoord c = site_coords(L, par, h);
  double v = init == 0 ? 0.0 : value[idx];
  const double sg = init == 0 ? 1.0 : -1.0;
  for (int mu = 0; mu < 3; mu++)
    for (int nu = mu + 1; nu < 4; nu++) {
      int x1 = c.x, y1 = c.y, z1 = c.z, t1 = c.t;
      shift_coord(L, x1, y1, z1, t1, mu, +1);
      int x2 = c.x, y2 = c.y, z2 = c.z, t2 = c.t;
      shift_coord(L, x2, y2, z2, t2, nu, +1);
      Mat<double> a = mmul(load_link_d18(U, Vh, mu, par, h), link_at(U, L, x1, y1, z1, t1, nu));
      a = mmul_nd(a, link_at(U, L, x2, y2, z2, t2, mu));
      a = mmul_nd(a, load_link_d18(U, Vh, nu, par, h));
      const double p = b3 * (a.e[0][0].x + a.e[1][1].x + a.e[2][2].x);
      v += sg * p;
    }
  for (int mu = 0; mu < 4; mu++) {
    const Mat<double> pm = load_herm(P, Vh, mu, par, h);
    const Mat<double> q = mmul(pm, pm);
    v += sg * 0.5 * (q.e[0][0].x + q.e[1][1].x + q.e[2][2].x);
  }
  if (with_f && par == 0) {
    for (int ps = 0; ps < npf; ps++) {
      double f = 0.0;
      for (int k = 0; k < 3; k++) {
        const double2 a = phi[((long)ps * 3 + k) * Vh + h];
        const double2 b = init == 0 ? a : chi[((long)ps * 3 + k) * Vh + h];
        f += a.x * b.x + a.y * b.y;
      }
      v += sg * f;
    }
  }
  value[idx] = v;
}

int action_local(Context &c, const double2 *phi, const double2 *chi, int with_f, int init) {
  hipLaunchKernelGGL(k_action, dim3(nblk(c.Vh), 2), dim3(kBlock), 0, c.stream, c.L, c.U64, c.P, c.beta / 3.0, phi,
                     chi, c.n_pf, with_f, c.action, init);
  return hip_ok(c, hipGetLastError(), "action") ? 0 : -1;
}

// include/tools.h:25-57 pairwise fold, same association order as the reference
__global__ __launch_bounds__(kBlock) void k_fold(double *v, long j, long off) {
  const long k = (long)blockIdx.x * kBlock + threadIdx.x;
  if (k < j) v[k] += v[k + off];
}

// the reference's pairwise tree sum (include/tools.h:25-57) of v[0..n), in place
static void pairwise_sum(Context &c, double *v, long n) {
  while (n > 1) {
    long j, off, next;
    if (n % 2 == 0) { j = n / 2; off = j; next = j; }
    else { j = n / 2; off = j + 1; next = j + 1; }
    hipLaunchKernelGGL(k_fold, dim3(nblk(j)), dim3(kBlock), 0, c.stream, v, j, off);
    n = next;
  }
}

int action_sum(Context &c, double *out) {
  pairwise_sum(c, c.action, c.V);
  return d2h_sync(c, out, c.action, sizeof(double)) ? 0 : -1;
}

// ------------------------------------------------------------ measurements
// gaugemeas.cc:6-66 (Conf::calc_plaq): per site minus Re tr of its three
// temporal (nu = 3) and three spatial plaquettes, at the reference site index
__global__ __launch_bounds__(kBlock) void k_plaq(LatticeDims L, const double2 *U, double *vt, double *vs) {
  const long Vh = L.Vh;
  const long h = (long)blockIdx.x * kBlock + threadIdx.x;
  const int par = blockIdx.y;
  if (h >= Vh) return;
  const SiteCoord c = site_coords(L, par, h);
  double st = 0.0, ss = 0.0;
  for (int mu = 0; mu < 3; mu++)
    for (int nu = mu + 1; nu < 4; nu++) {
      int x1 = c.x, y1 = c.y, z1 = c.z, t1 = c.t;
      shift_coord(L, x1, y1, z1, t1, mu, +1);
      int x2 = c.x, y2 = c.y, z2 = c.z, t2 = c.t;
      shift_coord(L, x2, y2, z2, t2, nu, +1);
      Mat<double> a = mmul(link_at(U, L, c.x, c.y, c.z, c.t, mu), link_at(U, L, x1, y1, z1, t1, nu));
      a = mmul_nd(a, link_at(U, L, x2, y2, z2, t2, mu));
      a = mmul_nd(a, link_at(U, L, c.x, c.y, c.z, c.t, nu));
      const double re = a.e[0][0].x + a.e[1][1].x + a.e[2][2].x;
      if (nu == 3) st -= re;
      else ss -= re;
    }
  const long idx = (long)par * Vh + h;
  vt[idx] = st;
  vs[idx] = ss;
}

// gaugemeas.cc:70-104 (Conf::calc_poly): one line per spatial point, from
// the sites r < vol3 (even parity, t = 0 or 1), minus the trace
__global__ __launch_bounds__(kBlock) void k_poly(LatticeDims L, const double2 *U, double *vre, double *vim,
                                                 long vol3) {
  const long r = (long)blockIdx.x * kBlock + threadIdx.x;
  if (r >= vol3) return;
  const SiteCoord c = site_coords(L, 0, r);
  Mat<double> a = link_at(U, L, c.x, c.y, c.z, c.t, 3);
  int t = c.t;
  for (int k = 1; k < L.nt; k++) {
    t = t + 1 == L.nt ? 0 : t + 1;
    a = mmul(a, link_at(U, L, c.x, c.y, c.z, t, 3));
  }
  vre[r] = -(a.e[0][0].x + a.e[1][1].x + a.e[2][2].x);
  vim[r] = -(a.e[0][0].y + a.e[1][1].y + a.e[2][2].y);
}

__device__ __forceinline__ uint4 philox4x32(uint4 ctr, uint2 key);

// chiralmeas.cc:18-205 (one noise vector, no imaginary chemical potential)
__global__ __launch_bounds__(kBlock) void k_axpby(double2 *out, const double2 *x, const double2 *y, double a,
                                                  double b, long n) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double2 u = x[i], v = y[i];
  out[i] = make_double2(a * u.x + b * v.x, a * u.y + b * v.y);
}

__global__ __launch_bounds__(kBlock) void k_z2noise(double2 *out, long n, unsigned long long seed,
                                                    unsigned long long stream) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint4 r = philox4x32(make_uint4((unsigned)i, (unsigned)(i >> 32), 7u, (unsigned)stream),
                             make_uint2((unsigned)seed, (unsigned)(seed >> 32)));
  out[i] = make_double2((r.x & 1u) ? -1.0 : 1.0, 0.0);
}

__device__ __forceinline__ double2 cdot3(const double2 *a, long as, const double2 *b, long bs) {
  double2 r = make_double2(0.0, 0.0);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double2 u = a[k * as], v = b[k * bs];
    r.x += u.x * v.x + u.y * v.y;  // conj(u) * v
    r.y += u.x * v.y - u.y * v.x;
  }
  return r;
}
// <w, U v> with U the link at (mu, par, h) and v the spinor at half index hn
__device__ __forceinline__ double2 dot_link(const double2 *w, long ws, const double2 *U, long Vh, int mu, int par,
                                            long h, const double2 *v, long hn) {
  const Mat<double> m = load_link_d18(U, Vh, mu, par, h);
  double2 r = make_double2(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double2 t = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < 3; k++) cfma(t, m.e[i][k], v[k * Vh + hn]);
    const double2 u = w[i * ws];
    r.x += u.x * t.x + u.y * t.y;
    r.y += u.x * t.y - u.y * t.x;
  }
  return r;
}
// <U w, v> = conj(<v, U w>)
__device__ __forceinline__ double2 dot_link_left(const double2 *U, long Vh, int mu, int par, long h,
                                                 const double2 *w, long hn, const double2 *v, long vs) {
  const double2 r = dot_link(v, vs, U, Vh, mu, par, h, w, hn);
  return make_double2(r.x, -r.y);
}

// per half index i: [chiral re, im, energy re, im, baryon re, im, pressure re, im]
__global__ __launch_bounds__(kBlock) void k_chiral_terms(LatticeDims L, const double2 *U, const double2 *re_,
                                                         const double2 *ro, const double2 *ce, const double2 *co,
                                                         double *w) {
  const long Vh = L.Vh;
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= Vh) return;
  const SiteCoord ce_ = site_coords(L, 0, i), co_ = site_coords(L, 1, i);
  const double2 a = cdot3(ro + i, Vh, co + i, Vh), b = cdot3(re_ + i, Vh, ce + i, Vh);
  double v[8];
  v[0] = a.x + b.x;
  v[1] = a.y + b.y;
  // temporal hops: even i -> odd neighbour, odd i -> even neighbour
  {
    const long ne = neighbour(L, ce_, 3, +1), no = neighbour(L, co_, 3, +1);
    double2 p1 = dot_link(re_ + i, Vh, U, Vh, 3, 0, i, co, ne);
    double2 p2 = dot_link_left(U, Vh, 3, 0, i, ro, ne, ce + i, Vh);
    double2 c1 = make_double2(p1.x - p2.x, p1.y - p2.y), c2 = make_double2(p1.x + p2.x, p1.y + p2.y);
    p1 = dot_link(ro + i, Vh, U, Vh, 3, 1, i, ce, no);
    p2 = dot_link_left(U, Vh, 3, 1, i, re_, no, co + i, Vh);
    c1.x += p1.x - p2.x; c1.y += p1.y - p2.y;
    c2.x += p1.x + p2.x; c2.y += p1.y + p2.y;
    v[2] = c1.x; v[3] = c1.y; v[4] = c2.x; v[5] = c2.y;
  }
  double2 c1 = make_double2(0.0, 0.0);
  for (int k = 0; k < 3; k++) {
    const long ne = neighbour(L, ce_, k, +1), no = neighbour(L, co_, k, +1);
    double2 p = dot_link(re_ + i, Vh, U, Vh, k, 0, i, co, ne);
    c1.x += p.x; c1.y += p.y;
    p = dot_link_left(U, Vh, k, 0, i, ro, ne, ce + i, Vh);
    c1.x -= p.x; c1.y -= p.y;
    p = dot_link(ro + i, Vh, U, Vh, k, 1, i, ce, no);
    c1.x += p.x; c1.y += p.y;
    p = dot_link_left(U, Vh, k, 1, i, re_, no, co + i, Vh);
    c1.x -= p.x; c1.y -= p.y;
  }
  v[6] = c1.x;
  v[7] = c1.y;
#pragma unroll
  for (int k = 0; k < 8; k++) w[k * Vh + i] = v[k];
}

int chiral_meas(Context &c, const double *rnd_e_h, const double *rnd_o_h, unsigned long long seed, double res,
                int max_cg, double n_flavours, double *out, int *iters) {
  const long Vh = c.Vh, n3 = 3 * Vh;
  double2 *buf = nullptr;
  double *w = nullptr;
  if (!hip_ok(c, hipMallocAsync((void **)&buf, sizeof(double2) * n3 * 6, c.stream), "chiral")) return -1;
  if (!hip_ok(c, hipMallocAsync((void **)&w, sizeof(double) * 8 * Vh, c.stream), "chiral")) return -1;
  double2 *re_ = buf, *ro = buf + n3, *phi = buf + 2 * n3, *ce = buf + 3 * n3, *co = buf + 4 * n3, *tmp = buf + 5 * n3;
  if (rnd_e_h && rnd_o_h) {
    if (upload_spinors(c, rnd_e_h, re_, 1) || upload_spinors(c, rnd_o_h, ro, 1)) return -1;
  } else {
    hipLaunchKernelGGL(k_z2noise, dim3(nblk(n3)), dim3(kBlock), 0, c.stream, re_, n3, seed, 0ull);
    hipLaunchKernelGGL(k_z2noise, dim3(nblk(n3)), dim3(kBlock), 0, c.stream, ro, n3, seed, 1ull);
  }
  launch_hop_d(c, 0, ro, tmp, 1, n3, n3);  // D_eo rnd_o
  hipLaunchKernelGGL(k_axpby, dim3(nblk(n3)), dim3(kBlock), 0, c.stream, phi, re_, tmp, c.mass, -1.0, n3);
  const double shift0 = 0.0;
  int it = 0;
  const int rc = multishift_solve(c, phi, 1, 1, &shift0, res, 0, PREC_FP64, max_cg, ce, &it);
  if (rc) {
    hipFreeAsync(buf, c.stream);
    hipFreeAsync(w, c.stream);
    hipStreamSynchronize(c.stream);
    return rc;
  }
  launch_hop_d(c, 1, ce, tmp, 1, n3, n3);  // D_oe chi_e
  const double im = 1.0 / c.mass;
  hipLaunchKernelGGL(k_axpby, dim3(nblk(n3)), dim3(kBlock), 0, c.stream, co, ro, tmp, im, -im, n3);
  hipLaunchKernelGGL(k_chiral_terms, dim3(nblk(Vh)), dim3(kBlock), 0, c.stream, c.L, c.U64, re_, ro, ce, co, w);
  for (int k = 0; k < 8; k++) pairwise_sum(c, w + k * Vh, Vh);
  double h[8];
  for (int k = 0; k < 8; k++) hipMemcpyAsync(&h[k], w + k * Vh, sizeof(double), hipMemcpyDeviceToHost, c.stream);
  hipFreeAsync(buf, c.stream);
  hipFreeAsync(w, c.stream);
  if (!hip_ok(c, hipStreamSynchronize(c.stream), "chiral")) return -1;
  const double inv_size = 1.0 / (double)c.V;
  const double f0 = n_flavours * 0.25 * inv_size, f1 = 0.5 * n_flavours * 0.25 * inv_size;
  out[0] = h[0] * f0; out[1] = h[1] * f0;
  for (int k = 2; k < 8; k++) out[k] = h[k] * f1;
  if (iters) *iters = it;
  return 0;
}

int gauge_meas(Context &c, double *pls, double *plt, double *pre, double *pim) {
  double *w = nullptr;
  const long vol3 = (long)c.L.nx * c.L.ny * c.L.nz;
  if (!hip_ok(c, hipMallocAsync((void **)&w, sizeof(double) * (2 * c.V + 2 * vol3), c.stream), "meas")) return -1;
  double *vt = w, *vs = w + c.V, *vre = w + 2 * c.V, *vim = vre + vol3;
  hipLaunchKernelGGL(k_plaq, dim3(nblk(c.Vh), 2), dim3(kBlock), 0, c.stream, c.L, c.U64, vt, vs);
  hipLaunchKernelGGL(k_poly, dim3(nblk(vol3)), dim3(kBlock), 0, c.stream, c.L, c.U64, vre, vim, vol3);
  pairwise_sum(c, vt, c.V);
  pairwise_sum(c, vs, c.V);
  pairwise_sum(c, vre, vol3);
  pairwise_sum(c, vim, vol3);
  double h[4];
  hipMemcpyAsync(&h[0], vt, sizeof(double), hipMemcpyDeviceToHost, c.stream);
  hipMemcpyAsync(&h[1], vs, sizeof(double), hipMemcpyDeviceToHost, c.stream);
  hipMemcpyAsync(&h[2], vre, sizeof(double), hipMemcpyDeviceToHost, c.stream);
  hipMemcpyAsync(&h[3], vim, sizeof(double), hipMemcpyDeviceToHost, c.stream);
  hipFreeAsync(w, c.stream);
  if (!hip_ok(c, hipStreamSynchronize(c.stream), "meas")) return -1;
  // the reference's normalisation order: sum * (1/V/3.) * one_by_three
  const double inv_size_by_three = 1.0 / (double)c.V / 3.;
  const double inv_vol3 = 1.0 / (double)vol3;
  const double one_by_three = 0.3333333333333333333333333333333333333333333333333333333333333333333;
  *plt = h[0] * inv_size_by_three * one_by_three;
  *pls = h[1] * inv_size_by_three * one_by_three;
  *pre = h[2] * inv_vol3 * one_by_three;
  *pim = h[3] * inv_vol3 * one_by_three;
  return 0;
}

// --------------------------------------------------------------------- RNG
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) counter-based generator;
// Box-Muller with radius sqrt(-ln u) gives N(0,1/2) components (gauss.cc:10-35).
__device__ __forceinline__ uint4 philox4x32(uint4 ctr, uint2 key) {
  const unsigned M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const unsigned long long p0 = (unsigned long long)M0 * ctr.x;
    const unsigned long long p1 = (unsigned long long)M1 * ctr.z;
    const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
    const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
    ctr = make_uint4(hi1 ^ ctr.y ^ key.x, lo1, hi0 ^ ctr.w ^ key.y, lo0);
    key.x += W0;
    key.y += W1;
  }
  return ctr;
}
__device__ __forceinline__ double u53(unsigned a, unsigned b) {
  const unsigned long long v = ((unsigned long long)a << 21) ^ (unsigned long long)b;
  return ((double)(v & ((1ull << 53) - 1)) + 0.5) * (1.0 / 9007199254740992.0);
}
__device__ __forceinline__ double2 gauss_pair(unsigned long long seed, unsigned long long stream, unsigned long long idx,
                                              unsigned k) {
  const uint4 r = philox4x32(make_uint4((unsigned)idx, (unsigned)(idx >> 32), k, (unsigned)stream),
                             make_uint2((unsigned)seed, (unsigned)(seed >> 32)));
  const double u1 = u53(r.x, r.y), u2 = u53(r.z, r.w);
  const double rad = sqrt(-log(u1));
  double sn, cs;
  sincos(6.283185307179586476925286766559 * u2, &sn, &cs);
  return make_double2(rad * cs, rad * sn);
}

// gauss.cc:63-86 layout of the gaussian su(3) matrix
__global__ __launch_bounds__(kBlock) void k_gauss_mom(double2 *P, long Vh, unsigned long long seed,
                                                      unsigned long long stream) {
  const long h = (long)blockIdx.x * kBlock + threadIdx.x;
  const int par = blockIdx.y, mu = blockIdx.z;
  if (h >= Vh) return;
  const unsigned long long link = ((unsigned long long)mu * 2 + par) * (unsigned long long)Vh + h;
  double r[8];
  for (int k = 0; k < 4; k++) {
    const double2 g = gauss_pair(seed, stream, link, k);
    r[2 * k] = g.x;
    r[2 * k + 1] = g.y;
  }
  const double s3 = 0.57735026918962576450914878050196;
  double2 *b = P + (long)(mu * 2 + par) * 4 * Vh + h;
  b[0] = make_double2(r[0], -r[1]);
  b[Vh] = make_double2(r[3], -r[4]);
  b[2 * Vh] = make_double2(r[5], -r[6]);
  b[3 * Vh] = make_double2(r[2] + s3 * r[7], -r[2] + s3 * r[7]);
}

int gauss_momenta(Context &c, unsigned long long seed, unsigned long long stream) {
  hipLaunchKernelGGL(k_gauss_mom, dim3(nblk(c.Vh), 2, 4), dim3(kBlock), 0, c.stream, c.P, c.Vh, seed, stream);
  return hip_ok(c, hipGetLastError(), "gauss momenta") ? 0 : -1;
}

__global__ __launch_bounds__(kBlock) void k_gauss_spinor(double2 *out, long Vh, unsigned long long seed,
                                                         unsigned long long stream) {
  const long h = (long)blockIdx.x * kBlock + threadIdx.x;
  const int ps = blockIdx.y;
  if (h >= Vh) return;
  for (int k = 0; k < 3; k++)
    out[((long)ps * 3 + k) * Vh + h] = gauss_pair(seed, stream, (unsigned long long)ps * Vh + h, k);
}

int gauss_spinor(Context &c, double2 *out, int n, unsigned long long seed, unsigned long long stream) {
  hipLaunchKernelGGL(k_gauss_spinor, dim3(nblk(c.Vh), n), dim3(kBlock), 0, c.stream, out, c.Vh, seed, stream);
  return hip_ok(c, hipGetLastError(), "gauss spinor") ? 0 : -1;
}

// ------------------------------------------------ reversibility distances
// revupdate.cc:66-92: sum over links of l2norm2(U - U_save) (kind 0: FP64
// 18-real links, n = 72 Vh complex entries) or of l2norm2(P + P_save) (kind
// 1: hermitian algebra fields stored as P01, P02, P12, (P00, P11), n = 32 Vh;
// the full 3x3 norm counts each off-diagonal entry twice and P22 = -(P00 +
// P11)).  One FP64 partial per workgroup, summed in a fixed order by one
// workgroup.
__global__ __launch_bounds__(kBlock) void k_pair_norm2(const double2 *a, const double2 *b, long Vh, int kind,
                                                       double *partial) {
  __shared__ double sh[kBlock / 64];
  const long stride = (long)gridDim.x * kBlock;
  double v = 0.0;
  if (kind == 0) {
    const long n = 72 * Vh;
    for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
      const double2 x = a[i], y = b[i];
      const double dx = x.x - y.x, dy = x.y - y.y;
      v += dx * dx + dy * dy;
    }
  } else {
    const long n = 8 * Vh;  // (mu, parity, h)
    for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
      const long blk = i / Vh, h = i - blk * Vh;
      double2 d[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const double2 x = a[(blk * 4 + k) * Vh + h], y = b[(blk * 4 + k) * Vh + h];
        d[k] = make_double2(x.x + y.x, x.y + y.y);
      }
      const double p22 = d[3].x + d[3].y;
      v += 2.0 * (d[0].x * d[0].x + d[0].y * d[0].y + d[1].x * d[1].x + d[1].y * d[1].y + d[2].x * d[2].x +
                  d[2].y * d[2].y) +
           d[3].x * d[3].x + d[3].y * d[3].y + p22 * p22;
    }
  }
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < kBlock / 64; k++) s += sh[k];
    partial[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(kBlock) void k_sum_partials(const double *partial, int n, double *out) {
  __shared__ double sh[kBlock / 64];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += kBlock) v += partial[i];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < kBlock / 64; k++) s += sh[k];
    *out = s;
  }
}

int pair_norm2(Context &c, const double2 *a, const double2 *b, int kind, double *out) {
  const int nb = 1024;
  double *part = nullptr;
  if (!hip_ok(c, hipMallocAsync((void **)&part, sizeof(double) * (nb + 1), c.stream), "pair norm")) return -1;
  hipLaunchKernelGGL(k_pair_norm2, dim3(nb), dim3(kBlock), 0, c.stream, a, b, c.Vh, kind, part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, c.stream, part, nb, part + nb);
  hipMemcpyAsync(out, part + nb, sizeof(double), hipMemcpyDeviceToHost, c.stream);
  hipFreeAsync(part, c.stream);
  return hip_ok(c, hipStreamSynchronize(c.stream), "pair norm") ? 0 : -1;
}

// ------------------------------------------------------- copy probe
// plain streaming copy of known size: calibrates the HBM byte counters and
// gives the achievable-bandwidth reference for the roofline
__global__ __launch_bounds__(kBlock) void k_copy_probe(const double4 *__restrict__ in, double4 *__restrict__ out,
                                                       long n) {
  const long stride = (long)gridDim.x * kBlock;
  for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) out[i] = in[i];
}

int time_copy(Context &c, size_t bytes, int reps, double *ms) {
  const long n = (long)(bytes / sizeof(double4));
  double4 *a = nullptr, *b = nullptr;
  if (!hip_ok(c, hipMalloc(&a, n * sizeof(double4)), "copy probe")) return -1;
  if (!hip_ok(c, hipMalloc(&b, n * sizeof(double4)), "copy probe")) { hipFree(a); return -1; }
  hipMemsetAsync(a, 0, n * sizeof(double4), c.stream);
  const unsigned nb = 256 * 8 * 4;
  hipLaunchKernelGGL(k_copy_probe, dim3(nb), dim3(kBlock), 0, c.stream, a, b, n);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, c.stream);
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL(k_copy_probe, dim3(nb), dim3(kBlock), 0, c.stream, (r & 1) ? b : a, (r & 1) ? a : b, n);
  hipEventRecord(e1, c.stream);
  hipEventSynchronize(e1);
  float t = 0.f;
  hipEventElapsedTime(&t, e0, e1);
  *ms = (double)t / reps;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(a);
  hipFree(b);
  return hip_ok(c, hipGetLastError(), "copy probe") ? 0 : -1;
}

// ----------------------------------------------------- host-layout transfers
// host links/algebra: [4V][3][3] complex at mu*V + par*Vh + h  ==  [(mu*2+par)*Vh + h]
__global__ __launch_bounds__(kBlock) void k_links_in(const double2 *aos, double2 *U, long Vh) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;  // over 8*Vh links
  if (i >= 8 * Vh) return;
  const long blk = i / Vh, h = i - blk * Vh;
#pragma unroll
  for (int k = 0; k < 9; k++) U[(blk * 9 + k) * Vh + h] = aos[i * 9 + k];
}
__global__ __launch_bounds__(kBlock) void k_links_out(const double2 *U, double2 *aos, long Vh) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= 8 * Vh) return;
  const long blk = i / Vh, h = i - blk * Vh;
#pragma unroll
  for (int k = 0; k < 9; k++) aos[i * 9 + k] = U[(blk * 9 + k) * Vh + h];
}
// algebra fields: full matrix -> 8-real packing (herm: real diag, aherm: imag diag)
__global__ __launch_bounds__(kBlock) void k_alg_in(const double2 *aos, double2 *A, long Vh, int herm) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= 8 * Vh) return;
  const long blk = i / Vh, h = i - blk * Vh;
  const double2 *m = aos + i * 9;
  double2 *b = A + blk * 4 * Vh + h;
  b[0] = m[1];
  b[Vh] = m[2];
  b[2 * Vh] = m[5];
  b[3 * Vh] = herm ? make_double2(m[0].x, m[4].x) : make_double2(m[0].y, m[4].y);
}
__global__ __launch_bounds__(kBlock) void k_alg_out(const double2 *A, double2 *aos, long Vh, int herm) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= 8 * Vh) return;
  const long blk = i / Vh, h = i - blk * Vh;
  const int mu = (int)(blk / 2), par = (int)(blk % 2);
  const Mat<double> m = herm ? load_herm(A, Vh, mu, par, h) : load_aherm(A, Vh, mu, par, h);
  double2 *o = aos + i * 9;
  for (int r = 0; r < 3; r++)
    for (int k = 0; k < 3; k++) o[r * 3 + k] = m.e[r][k];
}
// spinors: host [n][Vh][3] <-> device [n][3][Vh]
__global__ __launch_bounds__(kBlock) void k_spin_in(const double2 *aos, double2 *v, long Vh) {
  const long h = (long)blockIdx.x * kBlock + threadIdx.x;
  const int j = blockIdx.y;
  if (h >= Vh) return;
  for (int k = 0; k < 3; k++) v[((long)j * 3 + k) * Vh + h] = aos[((long)j * Vh + h) * 3 + k];
}
__global__ __launch_bounds__(kBlock) void k_spin_out(const double2 *v, double2 *aos, long Vh) {
  const long h = (long)blockIdx.x * kBlock + threadIdx.x;
  const int j = blockIdx.y;
  if (h >= Vh) return;
  for (int k = 0; k < 3; k++) aos[((long)j * Vh + h) * 3 + k] = v[((long)j * 3 + k) * Vh + h];
}

// Largest |row 2 - eta (row 0 x row 1)^*| over the links, per workgroup, on
// the host layout (link i = (2 mu + parity) Vh + h).  The 12-real copies
// (the FP32 MD operator, the FP64 solves, the gauge force) rebuild row 2
// this way, which reproduces the stored link only when it is unitary.  NaN
// entries are skipped (the solvers report them).
__global__ __launch_bounds__(kBlock) void k_unitarity(LatticeDims L, const double2 *aos, double *part) {
  __shared__ double red[kBlock];
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  double d = 0.0;
  if (i < 8 * L.Vh) {
    const long blk = i / L.Vh, h = i - blk * L.Vh;
    const int mu = (int)(blk / 2), par = (int)(blk % 2);
    Mat<double> m;
#pragma unroll
    for (int k = 0; k < 9; k++) m.e[k / 3][k % 3] = aos[i * 9 + k];
    const SiteCoord c = site_coords_i(L, par, (int)h);
    Mat<double> r = m;
    rebuild_row2(r, (double)eta_phase(L, c.x, c.y, c.z, c.t, mu));
#pragma unroll
    for (int k = 0; k < 3; k++) d = fmax(d, hypot(r.e[2][k].x - m.e[2][k].x, r.e[2][k].y - m.e[2][k].y));
  }
  red[threadIdx.x] = d;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Links from the host layout to dst (FP64 18-real).  *max_dev receives the
// largest deviation from unitarity (k_unitarity); above tol dst is left
// unchanged and 1 is returned.
int upload_links(Context &c, const double *host, double2 *dst, double tol, double *max_dev) {
  const size_t bytes = sizeof(double2) * 9 * 8 * c.Vh;
  const unsigned nb = nblk(8 * c.Vh);
  double2 *st = nullptr;
  double *part = nullptr;
  if (!hip_ok(c, hipMallocAsync((void **)&st, bytes, c.stream), "stage")) return -1;
  if (!hip_ok(c, hipMallocAsync((void **)&part, sizeof(double) * nb, c.stream), "stage")) {
    hipFreeAsync(st, c.stream);
    return -1;
  }
  hipMemcpyAsync(st, host, bytes, hipMemcpyHostToDevice, c.stream);
  hipLaunchKernelGGL(k_unitarity, dim3(nb), dim3(kBlock), 0, c.stream, c.L, st, part);
  std::vector<double> hp(nb);
  hipMemcpyAsync(hp.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, c.stream);
  hipFreeAsync(part, c.stream);
  if (!hip_ok(c, hipStreamSynchronize(c.stream), "upload links")) {
    hipFreeAsync(st, c.stream);
    return -1;
  }
  double m = 0.0;
  for (double v : hp) m = std::max(m, v);
  *max_dev = m;
  if (m > tol) {
    hipFreeAsync(st, c.stream);
    return hip_ok(c, hipStreamSynchronize(c.stream), "upload links") ? 1 : -1;
  }
  hipLaunchKernelGGL(k_links_in, dim3(nblk(8 * c.Vh)), dim3(kBlock), 0, c.stream, st, dst, c.Vh);
  hipFreeAsync(st, c.stream);
  return hip_ok(c, hipStreamSynchronize(c.stream), "upload links") ? 0 : -1;
}
int download_links(Context &c, const double2 *src, double *host) {
  const size_t bytes = sizeof(double2) * 9 * 8 * c.Vh;
  double2 *st = nullptr;
  if (!hip_ok(c, hipMallocAsync((void **)&st, bytes, c.stream), "stage")) return -1;
  hipLaunchKernelGGL(k_links_out, dim3(nblk(8 * c.Vh)), dim3(kBlock), 0, c.stream, src, st, c.Vh);
  hipMemcpyAsync(host, st, bytes, hipMemcpyDeviceToHost, c.stream);
  hipFreeAsync(st, c.stream);
  return hip_ok(c, hipStreamSynchronize(c.stream), "download links") ? 0 : -1;
}
int upload_alg(Context &c, const double *host, double2 *dst, int herm) {
  const size_t bytes = sizeof(double2) * 9 * 8 * c.Vh;
  double2 *st = nullptr;
  if (!hip_ok(c, hipMallocAsync((void **)&st, bytes, c.stream), "stage")) return -1;
  hipMemcpyAsync(st, host, bytes, hipMemcpyHostToDevice, c.stream);
  hipLaunchKernelGGL(k_alg_in, dim3(nblk(8 * c.Vh)), dim3(kBlock), 0, c.stream, st, dst, c.Vh, herm);
  hipFreeAsync(st, c.stream);
  return hip_ok(c, hipStreamSynchronize(c.stream), "upload algebra") ? 0 : -1;
}
int download_alg(Context &c, const double2 *src, double *host, int herm) {
  const size_t bytes = sizeof(double2) * 9 * 8 * c.Vh;
  double2 *st = nullptr;
  if (!hip_ok(c, hipMallocAsync((void **)&st, bytes, c.stream), "stage")) return -1;
  hipLaunchKernelGGL(k_alg_out, dim3(nblk(8 * c.Vh)), dim3(kBlock), 0, c.stream, src, st, c.Vh, herm);
  hipMemcpyAsync(host, st, bytes, hipMemcpyDeviceToHost, c.stream);
  hipFreeAsync(st, c.stream);
  return hip_ok(c, hipStreamSynchronize(c.stream), "download algebra") ? 0 : -1;
}
int upload_spinors(Context &c, const double *host, double2 *dst, int n) {
  const size_t bytes = sizeof(double2) * 3 * c.Vh * n;
  double2 *st = nullptr;
  if (!hip_ok(c, hipMallocAsync((void **)&st, bytes, c.stream), "stage")) return -1;
  hipMemcpyAsync(st, host, bytes, hipMemcpyHostToDevice, c.stream);
  hipLaunchKernelGGL(k_spin_in, dim3(nblk(c.Vh), n), dim3(kBlock), 0, c.stream, st, dst, c.Vh);
  hipFreeAsync(st, c.stream);
  return hip_ok(c, hipStreamSynchronize(c.stream), "upload spinors") ? 0 : -1;
}
int download_spinors(Context &c, const double2 *src, double *host, int n) {
  const size_t bytes = sizeof(double2) * 3 * c.Vh * n;
  double2 *st = nullptr;
  if (!hip_ok(c, hipMallocAsync((void **)&st, bytes, c.stream), "stage")) return -1;
  hipLaunchKernelGGL(k_spin_out, dim3(nblk(c.Vh), n), dim3(kBlock), 0, c.stream, src, st, c.Vh);
  hipMemcpyAsync(host, st, bytes, hipMemcpyDeviceToHost, c.stream);
  hipFreeAsync(st, c.stream);
  return hip_ok(c, hipStreamSynchronize(c.stream), "download spinors") ? 0 : -1;
}

}  // namespace rhmc
