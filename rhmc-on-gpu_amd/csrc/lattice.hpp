// lattice.hpp -- device-side geometry, complex/SU(3) helpers and HBM layouts
// for the staggered RHMC hot path on gfx950.
//
// Site numbering (kept identical to the reference so host arrays interchange
// without permutation, lib/Main/Geometry/geometry.cc:5-22,82-86):
//   full index i = parity*Vh + h,  h = lex/2,  lex = x + nx*(y + ny*(z + nz*t)).
// Neighbours and staggered phases are computed arithmetically from (parity,h)
// instead of being streamed from tables.
//
// HBM layouts (structure of arrays, one element per lane -> coalesced):
//   FP64 links, 18 reals : double2  U[(mu*2+par)*9 + (3*row+col)][Vh]
//   FP32 links, 12 reals : float4   U[(mu*2+par)*3 + k][Vh]
//                          k=0:(u00,u01) k=1:(u02,u10) k=2:(u11,u12); row 2 is
//                          rebuilt as eta*conj(row0 x row1)
//   spinor set           : Cx       psi[(vec*3 + color)][Vh]      (one parity)
//   su(3) algebra field  : double2  A[(mu*2+par)*4 + k][Vh]  (8 reals/link)
//                          hermitian P:   (p01)(p02)(p12)(p00,p11)
//                          antiherm. A:   (a01)(a02)(a12)(Im a00, Im a11)
//   per-site scalars     : double   s[par*Vh + h]
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rhmc {

struct LatticeDims {
  int nx, ny, nz, nt;
  int nxh;   // nx/2
  long Vh;   // half volume
  int antiperiodic_t;
  double inv_nxh, inv_ny, inv_nz;  // reciprocals for the 32-bit coordinate split
};

// q = n / d, rem = n % d for 0 <= n < 2^31: the FP64 product with the
// reciprocal is within one of the quotient, one exact correction fixes it
// (an integer division by a run-time divisor is a long instruction sequence).
__device__ __forceinline__ int fdivmod(int n, int d, double inv, int &rem) {
  int q = (int)((double)n * inv);
  int r = n - q * d;
  if (r < 0) { q -= 1; r += d; }
  else if (r >= d) { q += 1; r -= d; }
  rem = r;
  return q;
}

struct SiteCoord {
  int x, y, z, t;
};

__device__ __forceinline__ SiteCoord site_coords(const LatticeDims &L, int par, long h) {
  SiteCoord c;
  long r = h / L.nxh;
  int xh = (int)(h - r * L.nxh);
  c.y = (int)(r % L.ny);
  r /= L.ny;
  c.z = (int)(r % L.nz);
  c.t = (int)(r / L.nz);
  c.x = 2 * xh + ((par + c.y + c.z + c.t) & 1);
  return c;
}

__device__ __forceinline__ long half_index(const LatticeDims &L, int x, int y, int z, int t) {
  return (long)(x >> 1) + (long)L.nxh * (y + (long)L.ny * (z + (long)L.nz * t));
}

// 32-bit forms (Vh < 2^31, checked at context creation)
__device__ __forceinline__ SiteCoord site_coords_i(const LatticeDims &L, int par, int h) {
  SiteCoord c;
  int xh, y, z;
  const int r = fdivmod(h, L.nxh, L.inv_nxh, xh);
  const int r2 = fdivmod(r, L.ny, L.inv_ny, y);
  c.t = fdivmod(r2, L.nz, L.inv_nz, z);
  c.y = y;
  c.z = z;
  c.x = 2 * xh + ((par + y + z + c.t) & 1);
  return c;
}

__device__ __forceinline__ int half_index_i(const LatticeDims &L, int x, int y, int z, int t) {
  return (x >> 1) + L.nxh * (y + L.ny * (z + L.nz * t));
}

__device__ __forceinline__ int neighbour_i(const LatticeDims &L, SiteCoord c, int mu, int dir) {
  int x = c.x, y = c.y, z = c.z, t = c.t;
  if (mu == 0) x = dir > 0 ? (x + 1 == L.nx ? 0 : x + 1) : (x == 0 ? L.nx - 1 : x - 1);
  else if (mu == 1) y = dir > 0 ? (y + 1 == L.ny ? 0 : y + 1) : (y == 0 ? L.ny - 1 : y - 1);
  else if (mu == 2) z = dir > 0 ? (z + 1 == L.nz ? 0 : z + 1) : (z == 0 ? L.nz - 1 : z - 1);
  else t = dir > 0 ? (t + 1 == L.nt ? 0 : t + 1) : (t == 0 ? L.nt - 1 : t - 1);
  return half_index_i(L, x, y, z, t);
}

// half index of the neighbour x +/- mu (which has the opposite parity)
__device__ __forceinline__ long neighbour(const LatticeDims &L, SiteCoord c, int mu, int dir) {
  int x = c.x, y = c.y, z = c.z, t = c.t;
  if (mu == 0) x = dir > 0 ? (x + 1 == L.nx ? 0 : x + 1) : (x == 0 ? L.nx - 1 : x - 1);
  else if (mu == 1) y = dir > 0 ? (y + 1 == L.ny ? 0 : y + 1) : (y == 0 ? L.ny - 1 : y - 1);
  else if (mu == 2) z = dir > 0 ? (z + 1 == L.nz ? 0 : z + 1) : (z == 0 ? L.nz - 1 : z - 1);
  else t = dir > 0 ? (t + 1 == L.nt ? 0 : t + 1) : (t == 0 ? L.nt - 1 : t - 1);
  return half_index(L, x, y, z, t);
}

// staggered phase of the link (x,mu) incl. the antiperiodic temporal b.c.
// (geometry.cc:107-153)
__device__ __forceinline__ double eta_phase(const LatticeDims &L, int x, int y, int z, int t, int mu) {
  int s = 0;
  if (mu == 1) s = x;
  else if (mu == 2) s = x + y;
  else if (mu == 3) s = x + y + z + ((L.antiperiodic_t && t == L.nt - 1) ? 1 : 0);
  return (s & 1) ? -1.0 : 1.0;
}

// blockIdx remap so that each XCD (blocks are dealt round-robin over the 8
// XCDs) works on one contiguous slab of the lattice: neighbour gathers then
// hit in that XCD's own L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
  const unsigned nx = 8;
  if (nb % nx) return b;
  return (b % nx) * (nb / nx) + b / nx;
}

// ---------------------------------------------------------------- complex
template <typename R> struct CT;
template <> struct CT<double> { using t = double2; using q = double4; };
template <> struct CT<float> { using t = float2; using q = float4; };

template <typename R> using Cx = typename CT<R>::t;

template <typename R> __device__ __forceinline__ Cx<R> mk(R a, R b) { Cx<R> c; c.x = a; c.y = b; return c; }
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { a.x += b.x; a.y += b.y; return a; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { a.x -= b.x; a.y -= b.y; return a; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  C r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
// conj(a)*b
template <typename C> __device__ __forceinline__ C cmulc(C a, C b) {
  C r; r.x = a.x * b.x + a.y * b.y; r.y = a.x * b.y - a.y * b.x; return r;
}
template <typename C> __device__ __forceinline__ C cconj(C a) { a.y = -a.y; return a; }
template <typename C, typename R> __device__ __forceinline__ C cscale(C a, R s) { a.x *= s; a.y *= s; return a; }
__device__ __forceinline__ float fmar(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmar(double a, double b, double c) { return __builtin_fma(a, b, c); }
// Complex multiply-accumulates as chains of four fused multiply-adds
// (operand negation is free on the FMA).
// acc += a*b
template <typename C> __device__ __forceinline__ void cfma(C &acc, C a, C b) {
  acc.x = fmar(a.x, b.x, acc.x); acc.x = fmar(-a.y, b.y, acc.x);
  acc.y = fmar(a.x, b.y, acc.y); acc.y = fmar(a.y, b.x, acc.y);
}
// acc += conj(a)*b
template <typename C> __device__ __forceinline__ void cfmac(C &acc, C a, C b) {
  acc.x = fmar(a.x, b.x, acc.x); acc.x = fmar(a.y, b.y, acc.x);
  acc.y = fmar(a.x, b.y, acc.y); acc.y = fmar(-a.y, b.x, acc.y);
}
// acc -= conj(a)*b
template <typename C> __device__ __forceinline__ void cfmsc(C &acc, C a, C b) {
  acc.x = fmar(-a.x, b.x, acc.x); acc.x = fmar(-a.y, b.y, acc.x);
  acc.y = fmar(-a.x, b.y, acc.y); acc.y = fmar(a.y, b.x, acc.y);
}

template <typename R> struct Mat { Cx<R> e[3][3]; };
template <typename R> struct Vec { Cx<R> c[3]; };

template <typename R> __device__ __forceinline__ Mat<R> mmul(const Mat<R> &a, const Mat<R> &b) {
  Mat<R> r;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      Cx<R> s = cmul(a.e[i][0], b.e[0][j]);
      cfma(s, a.e[i][1], b.e[1][j]);
      cfma(s, a.e[i][2], b.e[2][j]);
      r.e[i][j] = s;
    }
  return r;
}
// a * b^dag
template <typename R> __device__ __forceinline__ Mat<R> mmul_nd(const Mat<R> &a, const Mat<R> &b) {
  Mat<R> r;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      Cx<R> s = cmul(a.e[i][0], cconj(b.e[j][0]));
      cfma(s, a.e[i][1], cconj(b.e[j][1]));
      cfma(s, a.e[i][2], cconj(b.e[j][2]));
      r.e[i][j] = s;
    }
  return r;
}
// a^dag * b
template <typename R> __device__ __forceinline__ Mat<R> mmul_dn(const Mat<R> &a, const Mat<R> &b) {
  Mat<R> r;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      Cx<R> s = cmulc(a.e[0][i], b.e[0][j]);
      cfmac(s, a.e[1][i], b.e[1][j]);
      cfmac(s, a.e[2][i], b.e[2][j]);
      r.e[i][j] = s;
    }
  return r;
}
template <typename R> __device__ __forceinline__ Mat<R> mdag(const Mat<R> &a) {
  Mat<R> r;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) r.e[i][j] = cconj(a.e[j][i]);
  return r;
}

// traceless anti-hermitian part (su3.cc:161-183)
template <typename R> __device__ __forceinline__ Mat<R> ta(const Mat<R> &a) {
  Mat<R> m;
  Cx<R> tr = mk<R>(0, 0);
#pragma unroll
  for (int i = 0; i < 3; i++) {
#pragma unroll
    for (int j = 0; j < 3; j++) m.e[i][j] = cscale(csub(a.e[i][j], cconj(a.e[j][i])), (R)0.5);
    tr = cadd(tr, m.e[i][i]);
  }
  tr = cscale(tr, (R)(1.0 / 3.0));
#pragma unroll
  for (int i = 0; i < 3; i++) m.e[i][i] = csub(m.e[i][i], tr);
  return m;
}

// Gram-Schmidt of rows 0,1, row 2 = conj(row0 x row1) (su3.cc:69-112)
template <typename R> __device__ __forceinline__ void sunitarize(Mat<R> &m) {
  R n = 0;
#pragma unroll
  for (int i = 0; i < 3; i++) n += m.e[0][i].x * m.e[0][i].x + m.e[0][i].y * m.e[0][i].y;
  n = (R)1.0 / sqrt(n);
#pragma unroll
  for (int i = 0; i < 3; i++) m.e[0][i] = cscale(m.e[0][i], n);
  Cx<R> p = mk<R>(0, 0);
#pragma unroll
  for (int i = 0; i < 3; i++) cfmac(p, m.e[0][i], m.e[1][i]);
#pragma unroll
  for (int i = 0; i < 3; i++) m.e[1][i] = csub(m.e[1][i], cmul(p, m.e[0][i]));
  n = 0;
#pragma unroll
  for (int i = 0; i < 3; i++) n += m.e[1][i].x * m.e[1][i].x + m.e[1][i].y * m.e[1][i].y;
  n = (R)1.0 / sqrt(n);
#pragma unroll
  for (int i = 0; i < 3; i++) m.e[1][i] = cscale(m.e[1][i], n);
  m.e[2][0] = cconj(csub(cmul(m.e[0][1], m.e[1][2]), cmul(m.e[0][2], m.e[1][1])));
  m.e[2][1] = cconj(csub(cmul(m.e[0][2], m.e[1][0]), cmul(m.e[0][0], m.e[1][2])));
  m.e[2][2] = cconj(csub(cmul(m.e[0][0], m.e[1][1]), cmul(m.e[0][1], m.e[1][0])));
}

// --------------------------------------------------------------- link I/O
// load at a workgroup-uniform base plus a 32-bit per-lane byte offset (the
// SGPR-base global load: one offset register serves many loads)
template <typename T>
__device__ __forceinline__ T ld_off(const T *base, unsigned byte_off) {
  return *(const T *)((const char *)base + byte_off);
}

__device__ __forceinline__ Mat<double> load_link_d18(const double2 *__restrict__ U, long Vh, int mu, int par, long h) {
  Mat<double> m;
  const double2 *b = U + (long)(mu * 2 + par) * 9 * Vh + h;
#pragma unroll
  for (int c = 0; c < 9; c++) m.e[c / 3][c % 3] = b[(long)c * Vh];
  return m;
}
__device__ __forceinline__ void store_link_d18(double2 *__restrict__ U, long Vh, int mu, int par, long h,
                                               const Mat<double> &m) {
  double2 *b = U + (long)(mu * 2 + par) * 9 * Vh + h;
#pragma unroll
  for (int c = 0; c < 9; c++) b[(long)c * Vh] = m.e[c / 3][c % 3];
}

// FP64 12-real: rows 0 and 1 as six double2 [(mu*2+par)*6 + k][Vh]; row 2 is
// rebuilt as eta * conj(row0 x row1), exact to rounding for the unitarised
// links (96 instead of 144 bytes per link)
__device__ __forceinline__ void store_link_d12(double2 *__restrict__ U, long Vh, int mu, int par, long h,
                                               const Mat<double> &m) {
  double2 *b = U + (long)(mu * 2 + par) * 6 * Vh + h;
#pragma unroll
  for (int c = 0; c < 6; c++) b[(long)c * Vh] = m.e[c / 3][c % 3];
}

template <typename R>
__device__ __forceinline__ void rebuild_row2(Mat<R> &m, R eta) {
  m.e[2][0] = cscale(cconj(csub(cmul(m.e[0][1], m.e[1][2]), cmul(m.e[0][2], m.e[1][1]))), eta);
  m.e[2][1] = cscale(cconj(csub(cmul(m.e[0][2], m.e[1][0]), cmul(m.e[0][0], m.e[1][2]))), eta);
  m.e[2][2] = cscale(cconj(csub(cmul(m.e[0][0], m.e[1][1]), cmul(m.e[0][1], m.e[1][0]))), eta);
}

__device__ __forceinline__ Mat<double> load_link_d12(const double2 *__restrict__ U, long Vh, int mu, int par, long h,
                                                     double eta) {
  Mat<double> m;
  const double2 *b = U + (long)(mu * 2 + par) * 6 * Vh + h;
#pragma unroll
  for (int c = 0; c < 6; c++) m.e[c / 3][c % 3] = b[(long)c * Vh];
  rebuild_row2(m, eta);
  return m;
}

__device__ __forceinline__ Mat<float> load_link_s12(const float4 *__restrict__ U, long Vh, int mu, int par, long h,
                                                    float eta) {
  Mat<float> m;
  const float4 *b = U + (long)(mu * 2 + par) * 3 * Vh + h;
  float4 a0 = b[0], a1 = b[Vh], a2 = b[2 * Vh];
  m.e[0][0] = make_float2(a0.x, a0.y);
  m.e[0][1] = make_float2(a0.z, a0.w);
  m.e[0][2] = make_float2(a1.x, a1.y);
  m.e[1][0] = make_float2(a1.z, a1.w);
  m.e[1][1] = make_float2(a2.x, a2.y);
  m.e[1][2] = make_float2(a2.z, a2.w);
  rebuild_row2(m, eta);
  return m;
}
__device__ __forceinline__ void store_link_s12(float4 *__restrict__ U, long Vh, int mu, int par, long h,
                                               const Mat<double> &m) {
  float4 *b = U + (long)(mu * 2 + par) * 3 * Vh + h;
  b[0] = make_float4((float)m.e[0][0].x, (float)m.e[0][0].y, (float)m.e[0][1].x, (float)m.e[0][1].y);
  b[Vh] = make_float4((float)m.e[0][2].x, (float)m.e[0][2].y, (float)m.e[1][0].x, (float)m.e[1][0].y);
  b[2 * Vh] = make_float4((float)m.e[1][1].x, (float)m.e[1][1].y, (float)m.e[1][2].x, (float)m.e[1][2].y);
}

// --------------------------------------------------------- algebra fields
// hermitian traceless P packed as (p01)(p02)(p12)(p00,p11)
__device__ __forceinline__ Mat<double> load_herm(const double2 *__restrict__ A, long Vh, int mu, int par, long h) {
  const double2 *b = A + (long)(mu * 2 + par) * 4 * Vh + h;
  double2 p01 = b[0], p02 = b[Vh], p12 = b[2 * Vh], d = b[3 * Vh];
  Mat<double> m;
  m.e[0][0] = make_double2(d.x, 0.0);
  m.e[1][1] = make_double2(d.y, 0.0);
  m.e[2][2] = make_double2(-d.x - d.y, 0.0);
  m.e[0][1] = p01; m.e[1][0] = cconj(p01);
  m.e[0][2] = p02; m.e[2][0] = cconj(p02);
  m.e[1][2] = p12; m.e[2][1] = cconj(p12);
  return m;
}
// anti-hermitian traceless A packed as (a01)(a02)(a12)(Im a00, Im a11)
__device__ __forceinline__ Mat<double> load_aherm(const double2 *__restrict__ A, long Vh, int mu, int par, long h) {
  const double2 *b = A + (long)(mu * 2 + par) * 4 * Vh + h;
  double2 a01 = b[0], a02 = b[Vh], a12 = b[2 * Vh], d = b[3 * Vh];
  Mat<double> m;
  m.e[0][0] = make_double2(0.0, d.x);
  m.e[1][1] = make_double2(0.0, d.y);
  m.e[2][2] = make_double2(0.0, -d.x - d.y);
  m.e[0][1] = a01; m.e[1][0] = make_double2(-a01.x, a01.y);
  m.e[0][2] = a02; m.e[2][0] = make_double2(-a02.x, a02.y);
  m.e[1][2] = a12; m.e[2][1] = make_double2(-a12.x, a12.y);
  return m;
}
__device__ __forceinline__ void store_aherm(double2 *__restrict__ A, long Vh, int mu, int par, long h,
                                            const Mat<double> &m) {
  double2 *b = A + (long)(mu * 2 + par) * 4 * Vh + h;
  b[0] = m.e[0][1];
  b[Vh] = m.e[0][2];
  b[2 * Vh] = m.e[1][2];
  b[3 * Vh] = make_double2(m.e[0][0].y, m.e[1][1].y);
}
__device__ __forceinline__ void store_herm(double2 *__restrict__ A, long Vh, int mu, int par, long h,
                                           const Mat<double> &m) {
  double2 *b = A + (long)(mu * 2 + par) * 4 * Vh + h;
  b[0] = m.e[0][1];
  b[Vh] = m.e[0][2];
  b[2 * Vh] = m.e[1][2];
  b[3 * Vh] = make_double2(m.e[0][0].x, m.e[1][1].x);
}

// ------------------------------------------------------------ reductions
template <typename T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace rhmc
