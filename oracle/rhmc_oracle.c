/*
 * rhmc_oracle.c -- FP64 CPU restatement of the reference RHMC hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  The
 * product library never links or calls this file.
 *
 * Every function restates the reference CPU semantics of
 * coppolachan/RHMC-on-GPU at the file:line given in its comment; loop orders
 * of reductions follow the reference so that results agree to rounding.
 */
#include "rhmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef orc_c su3[3][3];
typedef orc_c vec3[3];

/* ------------------------------------------------------------------ */
/* small SU(3) algebra (su3.h, vec3.h)                                 */

static void m_mul(su3 out, const su3 a, const su3 b) { /* out = a*b (su3.h:300-353) */
  su3 t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) t[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
  memcpy(out, t, sizeof(su3));
}
static void m_dag(su3 out, const su3 a) {
  su3 t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) t[i][j] = conj(a[j][i]);
  memcpy(out, t, sizeof(su3));
}
static double m_retr(const su3 a) { return creal(a[0][0] + a[1][1] + a[2][2]); }

static void mv(vec3 out, const su3 m, const vec3 v) { /* su3.h:480-500 */
  for (int i = 0; i < 3; i++) out[i] = v[0] * m[i][0] + v[1] * m[i][1] + v[2] * m[i][2];
}
static void mdv(vec3 out, const su3 m, const vec3 v) { /* (~m)*v */
  for (int i = 0; i < 3; i++) out[i] = v[0] * conj(m[0][i]) + v[1] * conj(m[1][i]) + v[2] * conj(m[2][i]);
}
static double v_norm2(const vec3 v) {
  double s = 0.0;
  for (int i = 0; i < 3; i++) s += creal(v[i]) * creal(v[i]) + cimag(v[i]) * cimag(v[i]);
  return s;
}
static double v_rdot(const vec3 a, const vec3 b) { /* r_scalprod, vec3.h:240-249 */
  orc_c s = 0.0;
  for (int i = 0; i < 3; i++) s += conj(a[i]) * b[i];
  return creal(s);
}

/* su3.cc:69-112 */
void orc_sunitarize(orc_c m[3][3]) {
  double norm = 0.0;
  for (int i = 0; i < 3; i++) norm += creal(m[0][i]) * creal(m[0][i]) + cimag(m[0][i]) * cimag(m[0][i]);
  norm = 1.0 / sqrt(norm);
  for (int i = 0; i < 3; i++) m[0][i] *= norm;
  orc_c prod = 0.0;
  for (int i = 0; i < 3; i++) prod += conj(m[0][i]) * m[1][i];
  for (int i = 0; i < 3; i++) m[1][i] -= prod * m[0][i];
  norm = 0.0;
  for (int i = 0; i < 3; i++) norm += creal(m[1][i]) * creal(m[1][i]) + cimag(m[1][i]) * cimag(m[1][i]);
  norm = 1.0 / sqrt(norm);
  for (int i = 0; i < 3; i++) m[1][i] *= norm;
  m[2][0] = conj(m[0][1] * m[1][2] - m[0][2] * m[1][1]);
  m[2][1] = conj(m[0][2] * m[1][0] - m[0][0] * m[1][2]);
  m[2][2] = conj(m[0][0] * m[1][1] - m[0][1] * m[1][0]);
}

/* su3.cc:161-183 */
void orc_ta(orc_c m[3][3]) {
  su3 a;
  memcpy(a, m, sizeof(su3));
  orc_c tr = 0.0;
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++) m[i][j] = (a[i][j] - conj(a[j][i])) * 0.5;
    tr += m[i][i];
  }
  tr *= (1.0 / 3.0);
  for (int i = 0; i < 3; i++) m[i][i] -= tr;
}

/* su3.cc:187-219 */
void orc_su3_exp(orc_c m[3][3]) {
  su3 x, aux, ris;
  memcpy(x, m, sizeof(su3));
  memset(ris, 0, sizeof(su3));
  for (int i = 0; i < 3; i++) ris[i][i] = 1.0;
  memcpy(aux, x, sizeof(su3));
  const double c[6] = {1.0, 0.5, 1.0 / 3.0, 0.25, 0.2, 1.0 / 6.0};
  for (int k = 0; k < 6; k++) {
    if (k > 0) {
      m_mul(aux, aux, x);
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) aux[i][j] *= c[k];
    }
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) ris[i][j] += aux[i][j];
  }
  orc_sunitarize(ris);
  memcpy(m, ris, sizeof(su3));
}

/* Threads of the per-site loops (OpenMP).  Every parallel loop writes
 * disjoint sites and every global sum stays serial, so results do not
 * depend on the thread count. */
#ifdef _OPENMP
#include <omp.h>
#endif
void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* geometry.cc:5-153                                                   */

static long snum(const orc_lattice *L, int x, int y, int z, int t) {
  return ((long)x + (long)L->nx * (y + (long)L->ny * (z + (long)L->nz * t))) / 2;
}

int orc_lattice_init(orc_lattice *L, int nx, int ny, int nz, int nt, int antiperiodic_t) {
  if (nx % 2 || ny % 2 || nz % 2 || nt % 2) return -1;
  L->nx = nx; L->ny = ny; L->nz = nz; L->nt = nt;
  L->antiperiodic_t = antiperiodic_t;
  L->V = (long)nx * ny * nz * nt;
  L->Vh = L->V / 2;
  L->nnp = (long *)malloc(sizeof(long) * 4 * L->V);
  L->nnm = (long *)malloc(sizeof(long) * 4 * L->V);
  L->eta = (int *)malloc(sizeof(int) * 4 * L->V);
  const long V = L->V, Vh = L->Vh;
  for (int t = 0; t < nt; t++)
    for (int z = 0; z < nz; z++)
      for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
          int xp = (x + 1) % nx, xm = (x + nx - 1) % nx;
          int yp = (y + 1) % ny, ym = (y + ny - 1) % ny;
          int zp = (z + 1) % nz, zm = (z + nz - 1) % nz;
          int tp = (t + 1) % nt, tm = (t + nt - 1) % nt;
          int par = (x + y + z + t) % 2;
          long num = par * Vh + snum(L, x, y, z, t);
          long o = (1 - par) * Vh;
          L->nnp[num * 4 + 0] = o + snum(L, xp, y, z, t);
          L->nnm[num * 4 + 0] = o + snum(L, xm, y, z, t);
          L->nnp[num * 4 + 1] = o + snum(L, x, yp, z, t);
          L->nnm[num * 4 + 1] = o + snum(L, x, ym, z, t);
          L->nnp[num * 4 + 2] = o + snum(L, x, y, zp, t);
          L->nnm[num * 4 + 2] = o + snum(L, x, y, zm, t);
          L->nnp[num * 4 + 3] = o + snum(L, x, y, z, tp);
          L->nnm[num * 4 + 3] = o + snum(L, x, y, z, tm);
          L->eta[num] = 1;
          L->eta[num + V] = (x % 2) ? -1 : 1;
          L->eta[num + 2 * V] = ((x + y) % 2) ? -1 : 1;
          int e3 = ((x + y + z) % 2) ? -1 : 1;
          if (antiperiodic_t && t == nt - 1) e3 = -e3;
          L->eta[num + 3 * V] = e3;
        }
  return 0;
}

void orc_lattice_free(orc_lattice *L) {
  free(L->nnp); free(L->nnm); free(L->eta);
  L->nnp = L->nnm = NULL; L->eta = NULL;
}

/* conf.cc:235-251 */
void orc_unitarize_with_eta(const orc_lattice *L, orc_c (*U)[3][3]) {
#pragma omp parallel for schedule(static)
  for (long r = 0; r < 4 * L->V; r++) {
    orc_sunitarize(U[r]);
    for (int j = 0; j < 3; j++) U[r][2][j] *= (double)L->eta[r];
  }
}

/* ------------------------------------------------------------------ */
/* fermionmatrix.cc:8-97                                               */

void orc_deo(const orc_lattice *L, const orc_c (*U)[3][3], orc_c (*out)[3], const orc_c (*in)[3]) {
  const long V = L->V, Vh = L->Vh;
#pragma omp parallel for schedule(static)
  for (long e = 0; e < Vh; e++) {
    vec3 acc = {0, 0, 0}, t;
    for (int mu = 0; mu < 4; mu++) {
      mv(t, U[e + mu * V], in[L->nnp[e * 4 + mu] - Vh]);
      for (int c = 0; c < 3; c++) acc[c] += t[c];
    }
    for (int mu = 0; mu < 4; mu++) {
      long n = L->nnm[e * 4 + mu];
      mdv(t, U[n + mu * V], in[n - Vh]);
      for (int c = 0; c < 3; c++) acc[c] -= t[c];
    }
    for (int c = 0; c < 3; c++) out[e][c] = acc[c] * 0.5;
  }
}

void orc_doe(const orc_lattice *L, const orc_c (*U)[3][3], orc_c (*out)[3], const orc_c (*in)[3]) {
  const long V = L->V, Vh = L->Vh;
#pragma omp parallel for schedule(static)
  for (long o = Vh; o < V; o++) {
    vec3 acc = {0, 0, 0}, t;
    for (int mu = 0; mu < 4; mu++) {
      mv(t, U[o + mu * V], in[L->nnp[o * 4 + mu]]);
      for (int c = 0; c < 3; c++) acc[c] += t[c];
    }
    for (int mu = 0; mu < 4; mu++) {
      long n = L->nnm[o * 4 + mu];
      mdv(t, U[n + mu * V], in[n]);
      for (int c = 0; c < 3; c++) acc[c] -= t[c];
    }
    for (int c = 0; c < 3; c++) out[o - Vh][c] = acc[c] * 0.5;
  }
}

void orc_mdagm(const orc_lattice *L, const orc_c (*U)[3][3], double mass, orc_c (*out)[3],
               const orc_c (*in)[3]) {
  const long Vh = L->Vh;
  vec3 *h = (vec3 *)malloc(sizeof(vec3) * Vh);
  orc_doe(L, U, h, in);
  orc_deo(L, U, out, (const orc_c(*)[3])h);
  const double m2 = mass * mass;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < Vh; i++)
    for (int c = 0; c < 3; c++) out[i][c] = m2 * in[i][c] - out[i][c];
  free(h);
}

/* include/tools.h:25-57 */
double orc_global_sum(double *vec, long n) {
  while (n > 1) {
    if (n % 2 == 0) {
      long j = n / 2;
      for (long k = 0; k < j; k++) vec[k] += vec[k + j];
      n = j;
    } else {
      long j1 = n / 2, j2 = j1 + 1;
      for (long k = 0; k < j1; k++) vec[k] += vec[k + j2];
      n = j2;
    }
  }
  return vec[0];
}

/* ------------------------------------------------------------------ */
/* inverter.cc:12-127 (CG), 142-339 (multi-shift CG)                   */

int orc_cg(const orc_lattice *L, const orc_c (*U)[3][3], double mass, const orc_c (*b)[3],
           double res, int max_cg, orc_c (*x)[3]) {
  const long Vh = L->Vh;
  vec3 *r = (vec3 *)malloc(sizeof(vec3) * Vh), *p = (vec3 *)malloc(sizeof(vec3) * Vh);
  vec3 *s = (vec3 *)malloc(sizeof(vec3) * Vh);
  double *dv = (double *)malloc(sizeof(double) * Vh);
  orc_mdagm(L, U, mass, s, (const orc_c(*)[3])x);
  for (long i = 0; i < Vh; i++) {
    for (int c = 0; c < 3; c++) r[i][c] = p[i][c] = b[i][c] - s[i][c];
    dv[i] = v_norm2(r[i]);
  }
  double delta = orc_global_sum(dv, Vh), lambda;
  int cg = 0;
  do {
    cg++;
    orc_mdagm(L, U, mass, s, (const orc_c(*)[3])p);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < Vh; i++) dv[i] = v_rdot(p[i], s[i]);
    double alpha = orc_global_sum(dv, Vh);
    double omega = delta / alpha;
    for (long i = 0; i < Vh; i++) {
      for (int c = 0; c < 3; c++) {
        x[i][c] += omega * p[i][c];
        r[i][c] -= omega * s[i][c];
      }
      dv[i] = v_norm2(r[i]);
    }
    lambda = orc_global_sum(dv, Vh);
    double g = lambda / delta;
    delta = lambda;
    for (long i = 0; i < Vh; i++)
      for (int c = 0; c < 3; c++) p[i][c] = r[i][c] + g * p[i][c];
  } while (sqrt(lambda) > res && cg < max_cg);
  free(r); free(p); free(s); free(dv);
  return cg;
}

int orc_multishift(const orc_lattice *L, const orc_c (*U)[3][3], double mass, const orc_c (*b)[3],
                   int ns, const double *shifts, double res, int max_cg, orc_c (*x)[3]) {
  const long Vh = L->Vh;
  vec3 *r = (vec3 *)malloc(sizeof(vec3) * Vh), *p = (vec3 *)malloc(sizeof(vec3) * Vh);
  vec3 *s = (vec3 *)malloc(sizeof(vec3) * Vh);
  vec3 *ps = (vec3 *)malloc(sizeof(vec3) * Vh * ns);
  double *dv = (double *)malloc(sizeof(double) * Vh);
  double *z1 = (double *)malloc(sizeof(double) * ns), *z2 = (double *)malloc(sizeof(double) * ns);
  double *z3 = (double *)malloc(sizeof(double) * ns), *om = (double *)malloc(sizeof(double) * ns);
  double *ga = (double *)malloc(sizeof(double) * ns);
  int *flag = (int *)malloc(sizeof(int) * ns);

  memset(x, 0, sizeof(vec3) * Vh * ns);
  for (int k = 0; k < ns; k++) flag[k] = 1;
  for (long i = 0; i < Vh; i++) {
    for (int c = 0; c < 3; c++) r[i][c] = p[i][c] = b[i][c];
    dv[i] = v_norm2(r[i]);
  }
  double delta = orc_global_sum(dv, Vh), lambda;
  double omega = 1.0, omega_save, gammag = 0.0;
  for (int k = 0; k < ns; k++) {
    memcpy(ps + (long)k * Vh, b, sizeof(vec3) * Vh);
    z1[k] = 1.0; z2[k] = 1.0; ga[k] = 0.0;
  }
  int cg = 0;
  do {
    cg++;
    orc_mdagm(L, U, mass, s, (const orc_c(*)[3])p);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < Vh; i++) dv[i] = v_rdot(p[i], s[i]);
    double alpha = orc_global_sum(dv, Vh);
    omega_save = omega;
    omega = -delta / alpha;
    for (int k = 0; k < ns; k++) {
      if (!flag[k]) continue;
      z3[k] = (z1[k] * z2[k] * omega_save) /
              (omega * gammag * (z1[k] - z2[k]) + z1[k] * omega_save * (1.0 - shifts[k] * omega));
      om[k] = omega * z3[k] / z2[k];
      vec3 *xk = x + (long)k * Vh, *pk = ps + (long)k * Vh;
#pragma omp parallel for schedule(static)
      for (long i = 0; i < Vh; i++)
        for (int c = 0; c < 3; c++) xk[i][c] = xk[i][c] - om[k] * pk[i][c];
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < Vh; i++) {
      for (int c = 0; c < 3; c++) r[i][c] = r[i][c] + omega * s[i][c];
      dv[i] = v_norm2(r[i]);
    }
    lambda = orc_global_sum(dv, Vh);
    gammag = lambda / delta;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < Vh; i++)
      for (int c = 0; c < 3; c++) p[i][c] = r[i][c] + gammag * p[i][c];
    for (int k = 0; k < ns; k++) {
      if (!flag[k]) continue;
      ga[k] = gammag * z3[k] * om[k] / (z2[k] * omega);
      vec3 *pk = ps + (long)k * Vh;
#pragma omp parallel for schedule(static)
      for (long i = 0; i < Vh; i++)
        for (int c = 0; c < 3; c++) pk[i][c] = z3[k] * r[i][c] + ga[k] * pk[i][c];
      double fact = sqrt(delta * z2[k] * z2[k]);
      if (fact < res) flag[k] = 0;
      z1[k] = z2[k];
      z2[k] = z3[k];
    }
    delta = lambda;
  } while (sqrt(lambda) > res && cg < max_cg);

  free(r); free(p); free(s); free(ps); free(dv);
  free(z1); free(z2); free(z3); free(om); free(ga); free(flag);
  return cg;
}

/* ------------------------------------------------------------------ */
/* fermionforce.cc:46-104                                              */

void orc_fermion_force_accum(const orc_lattice *L, const orc_c (*U)[3][3], int ns, const double *num,
                             const orc_c (*x)[3], orc_c (*F)[3][3]) {
  const long V = L->V, Vh = L->Vh;
  vec3 *h = (vec3 *)malloc(sizeof(vec3) * Vh);
  for (int k = 0; k < ns; k++) {
    const orc_c(*s)[3] = x + (long)k * Vh;
    orc_doe(L, U, h, s);
#pragma omp parallel for schedule(static)
    for (long e = 0; e < Vh; e++) {
      long o = e + Vh;
      for (int mu = 0; mu < 4; mu++) {
        long y = L->nnp[e * 4 + mu] - Vh;
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) F[mu * V + e][i][j] += (num[k] * h[y][i]) * conj(s[e][j]);
        long y2 = L->nnp[o * 4 + mu];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) F[mu * V + o][i][j] += (-num[k] * s[y2][i]) * conj(h[e][j]);
      }
    }
  }
  free(h);
}

/* staples.cc:16-74 */
void orc_staples(const orc_lattice *L, const orc_c (*U)[3][3], orc_c (*st)[3][3]) {
  const long V = L->V;
#pragma omp parallel for schedule(static)
  for (long pos = 0; pos < V; pos++)
    for (int mu = 0; mu < 4; mu++) {
      su3 acc, a, b;
      memset(acc, 0, sizeof(su3));
      for (int nu = 0; nu < 4; nu++) {
        if (nu == mu) continue;
        /* upper staple U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag */
        m_dag(b, U[mu * V + L->nnp[pos * 4 + nu]]);
        m_mul(a, U[nu * V + L->nnp[pos * 4 + mu]], b);
        m_dag(b, U[nu * V + pos]);
        m_mul(a, a, b);
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) acc[i][j] += a[i][j];
        /* lower staple U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu) */
        long hp = L->nnm[pos * 4 + nu];
        m_dag(a, U[nu * V + L->nnp[hp * 4 + mu]]);
        m_dag(b, U[mu * V + hp]);
        m_mul(a, a, b);
        m_mul(a, a, U[nu * V + hp]);
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) acc[i][j] += a[i][j];
      }
      memcpy(st[mu * V + pos], acc, sizeof(su3));
    }
}

/* ipdot.cc:25-79 */
void orc_ipdot_gauge(const orc_lattice *L, const orc_c (*U)[3][3], double beta, orc_c (*ipdot)[3][3]) {
  orc_ipdot_total(L, U, beta, NULL, 1, ipdot);
}

/* ipdot.cc:83-196 */
void orc_ipdot_total(const orc_lattice *L, const orc_c (*U)[3][3], double beta, const orc_c (*F)[3][3],
                     int with_gauge, orc_c (*ipdot)[3][3]) {
  const long NL = 4 * L->V;
  su3 *st = NULL;
  if (with_gauge) {
    st = (su3 *)malloc(sizeof(su3) * NL);
    orc_staples(L, U, st);
  }
  const double b3 = beta / 3.0;
#pragma omp parallel for schedule(static)
  for (long r = 0; r < NL; r++) {
    su3 m;
    memset(m, 0, sizeof(su3));
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        if (with_gauge) m[i][j] = b3 * st[r][i][j];
        if (F) m[i][j] += F[r][i][j];
      }
    m_mul(m, U[r], m);
    orc_ta(m);
    memcpy(ipdot[r], m, sizeof(su3));
  }
  free(st);
}

/* momenta.cc:42-78 */
void orc_conf_left_exp_multiply(const orc_lattice *L, orc_c (*U)[3][3], const orc_c (*P)[3][3], double step) {
  const long NL = 4 * L->V;
  const orc_c p = I * step;
#pragma omp parallel for schedule(static)
  for (long r = 0; r < NL; r++) {
    su3 a;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) a[i][j] = P[r][i][j] * p;
    orc_su3_exp(a);
    m_mul(U[r], a, U[r]);
  }
  orc_unitarize_with_eta(L, U);
}

/* momenta.cc:82-115 */
void orc_momenta_sum_multiply(const orc_lattice *L, orc_c (*P)[3][3], const orc_c (*ipdot)[3][3], double step) {
  const long NL = 4 * L->V;
  const orc_c p = I * step;
  for (long r = 0; r < NL; r++)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) P[r][i][j] += ipdot[r][i][j] * p;
}

/* action.cc:6-156 */
void orc_calc_action(const orc_lattice *L, const orc_c (*U)[3][3], const orc_c (*P)[3][3], double beta,
                     int n_pf, const orc_c (*phi)[3], const orc_c (*chi)[3], int with_fermions,
                     double *value, int init) {
  const long V = L->V, Vh = L->Vh;
  const double b3 = beta / 3.0;
  for (long r = 0; r < V; r++) {
    if (init == 0) value[r] = 0.0;
    for (int mu = 0; mu < 3; mu++)
      for (int nu = mu + 1; nu < 4; nu++) {
        su3 a, b;
        m_mul(a, U[mu * V + r], U[nu * V + L->nnp[r * 4 + mu]]);
        m_dag(b, U[mu * V + L->nnp[r * 4 + nu]]);
        m_mul(a, a, b);
        m_dag(b, U[nu * V + r]);
        m_mul(a, a, b);
        double p = b3 * m_retr(a);
        if (init == 0) value[r] += p;
        if (init == -1) value[r] -= p;
      }
  }
  for (long r = 0; r < V; r++)
    for (int i = 0; i < 4; i++) {
      su3 a;
      m_mul(a, P[r + i * V], P[r + i * V]);
      double p = 0.5 * m_retr(a);
      if (init == 0) value[r] += p;
      if (init == -1) value[r] -= p;
    }
  if (!with_fermions) return;
  if (init == 0)
    for (int ps = 0; ps < n_pf; ps++)
      for (long i = 0; i < Vh; i++) value[i] += v_rdot(phi[ps * Vh + i], phi[ps * Vh + i]);
  if (init == -1)
    for (int ps = 0; ps < n_pf; ps++)
      for (long i = 0; i < Vh; i++) value[i] -= v_rdot(phi[ps * Vh + i], chi[ps * Vh + i]);
}

/* gaugemeas.cc:6-66 (Conf::calc_plaq): -Re tr of the plaquettes, temporal
 * (nu = 3) and spatial, pairwise-summed, / (3 V) / 3 */
void orc_plaquette(const orc_lattice *L, const orc_c (*U)[3][3], double *pls, double *plt) {
  const long V = L->V;
  double *dt = (double *)malloc(sizeof(double) * V), *ds = (double *)malloc(sizeof(double) * V);
  for (long r = 0; r < V; r++) {
    dt[r] = 0.0;
    ds[r] = 0.0;
    for (int mu = 0; mu < 3; mu++)
      for (int nu = mu + 1; nu < 4; nu++) {
        su3 a, b;
        m_mul(a, U[mu * V + r], U[nu * V + L->nnp[r * 4 + mu]]);
        m_dag(b, U[mu * V + L->nnp[r * 4 + nu]]);
        m_mul(a, a, b);
        m_dag(b, U[nu * V + r]);
        m_mul(a, a, b);
        if (nu == 3) dt[r] -= m_retr(a);
        else ds[r] -= m_retr(a);
      }
  }
  const double inv_size_by_three = 1.0 / (double)V / 3.;
  const double one_by_three = 0.3333333333333333333333333333333333333333333333333333333333333333333;
  *plt = orc_global_sum(dt, V) * inv_size_by_three * one_by_three;
  *pls = orc_global_sum(ds, V) * inv_size_by_three * one_by_three;
  free(dt);
  free(ds);
}

/* gaugemeas.cc:70-104 (Conf::calc_poly): lines start at the sites r < vol3
 * (even parity, t = 0 or 1: one per spatial point), product of the t links
 * along t, minus the trace (antiperiodic phase), pairwise-summed, / vol3 / 3 */
void orc_polyakov(const orc_lattice *L, const orc_c (*U)[3][3], double *re, double *im) {
  const long V = L->V, vol3 = (long)L->nx * L->ny * L->nz;
  double *d1 = (double *)malloc(sizeof(double) * vol3), *d2 = (double *)malloc(sizeof(double) * vol3);
  for (long r = 0; r < vol3; r++) {
    su3 a;
    memset(a, 0, sizeof(a));
    a[0][0] = a[1][1] = a[2][2] = 1.0;
    long r2 = r;
    for (int t = 0; t < L->nt; t++) {
      m_mul(a, a, U[r2 + 3 * V]);
      r2 = L->nnp[r2 * 4 + 3];
    }
    const orc_c tr = a[0][0] + a[1][1] + a[2][2];
    d1[r] = -creal(tr);
    d2[r] = -cimag(tr);
  }
  const double inv_vol3 = 1.0 / (double)vol3;
  const double one_by_three = 0.3333333333333333333333333333333333333333333333333333333333333333333;
  *re = orc_global_sum(d1, vol3) * inv_vol3 * one_by_three;
  *im = orc_global_sum(d2, vol3) * inv_vol3 * one_by_three;
  free(d1);
  free(d2);
}

/* chiralmeas.cc:18-205 for one Z2 noise vector (rnd_e, rnd_o), without the
 * imaginary chemical potential: chi_e = (MdagM)^-1 (m rnd_e - D_eo rnd_o),
 * chi_o = (rnd_o - D_oe chi_e)/m; out = [chiral, energy density, baryon
 * density, pressure density] as (re, im) pairs.  Sums are the pairwise
 * global_sum over the half lattice on real and imaginary parts. */
static void c_sum_pair(double *re, double *im, long n, orc_c *out) {
  const double a = orc_global_sum(re, n), b = orc_global_sum(im, n);
  *out = a + I * b;
}
static orc_c c_dot(const vec3 a, const vec3 b) { /* c_scalprod, vec3.h:227-235 */
  orc_c r = 0.0;
  for (int c = 0; c < 3; c++) r += conj(a[c]) * b[c];
  return r;
}
int orc_chiral_meas(const orc_lattice *L, const orc_c (*U)[3][3], double mass, double n_flavours,
                    const orc_c (*rnd_e)[3], const orc_c (*rnd_o)[3], double res, int max_cg, double *out) {
  const long V = L->V, Vh = L->Vh;
  vec3 *phi = (vec3 *)malloc(sizeof(vec3) * Vh), *chi_e = (vec3 *)calloc(Vh, sizeof(vec3));
  vec3 *chi_o = (vec3 *)malloc(sizeof(vec3) * Vh), *tmp = (vec3 *)malloc(sizeof(vec3) * Vh);
  double *re = (double *)malloc(sizeof(double) * Vh), *im = (double *)malloc(sizeof(double) * Vh);
  double *re2 = (double *)malloc(sizeof(double) * Vh), *im2 = (double *)malloc(sizeof(double) * Vh);
  orc_deo(L, U, phi, rnd_o);
  for (long i = 0; i < Vh; i++)
    for (int c = 0; c < 3; c++) phi[i][c] = mass * rnd_e[i][c] - phi[i][c];
  const int iters = orc_cg(L, U, mass, (const orc_c(*)[3])phi, res, max_cg, chi_e);
  orc_doe(L, U, tmp, (const orc_c(*)[3])chi_e);
  const double one_by_mass = 1.0 / mass;
  for (long i = 0; i < Vh; i++)
    for (int c = 0; c < 3; c++) chi_o[i][c] = one_by_mass * (rnd_o[i][c] - tmp[i][c]);
  const double inv_size = 1.0 / (double)V;
  orc_c chiral, sum1, sum2, e, b, pd;
  for (long i = 0; i < Vh; i++) {
    const orc_c v = c_dot(rnd_o[i], chi_o[i]) + c_dot(rnd_e[i], chi_e[i]);
    re[i] = creal(v);
    im[i] = cimag(v);
  }
  c_sum_pair(re, im, Vh, &sum1);
  chiral = sum1 * (n_flavours * 0.25 * inv_size);
  /* energy and baryon density: temporal hops */
  for (long i = 0; i < Vh; i++) {
    const long j = i + Vh;
    vec3 ve, we, vo, wo;
    long i2 = L->nnp[i * 4 + 3] - Vh;
    mv(ve, U[3 * V + i], chi_o[i2]);
    mv(we, U[3 * V + i], rnd_o[i2]);
    orc_c p1 = c_dot(rnd_e[i], ve), p2 = c_dot(we, chi_e[i]);
    orc_c c1 = p1 - p2, c2 = p1 + p2;
    i2 = L->nnp[j * 4 + 3];
    mv(vo, U[3 * V + j], chi_e[i2]);
    mv(wo, U[3 * V + j], rnd_e[i2]);
    p1 = c_dot(rnd_o[i], vo);
    p2 = c_dot(wo, chi_o[i]);
    c1 += p1 - p2;
    c2 += p1 + p2;
    re[i] = creal(c1); im[i] = cimag(c1);
    re2[i] = creal(c2); im2[i] = cimag(c2);
  }
  c_sum_pair(re, im, Vh, &sum1);
  c_sum_pair(re2, im2, Vh, &sum2);
  e = sum1 * (0.5 * n_flavours * 0.25 * inv_size);
  b = sum2 * (0.5 * n_flavours * 0.25 * inv_size);
  /* pressure density: spatial hops */
  for (long i = 0; i < Vh; i++) {
    const long j = i + Vh;
    orc_c c1 = 0.0;
    for (int k = 0; k < 3; k++) {
      vec3 ve, we, vo, wo;
      long i2 = L->nnp[i * 4 + k] - Vh;
      mv(ve, U[k * V + i], chi_o[i2]);
      mv(we, U[k * V + i], rnd_o[i2]);
      c1 += c_dot(rnd_e[i], ve);
      c1 -= c_dot(we, chi_e[i]);
      i2 = L->nnp[j * 4 + k];
      mv(vo, U[k * V + j], chi_e[i2]);
      mv(wo, U[k * V + j], rnd_e[i2]);
      c1 += c_dot(rnd_o[i], vo);
      c1 -= c_dot(wo, chi_o[i]);
    }
    re[i] = creal(c1);
    im[i] = cimag(c1);
  }
  c_sum_pair(re, im, Vh, &sum1);
  pd = sum1 * (0.5 * n_flavours * 0.25 * inv_size);
  const orc_c o4[4] = {chiral, e, b, pd};
  for (int k = 0; k < 4; k++) {
    out[2 * k] = creal(o4[k]);
    out[2 * k + 1] = cimag(o4[k]);
  }
  free(phi); free(chi_e); free(chi_o); free(tmp); free(re); free(im); free(re2); free(im2);
  return iters;
}

/* findminmax.cc:12-72 */
int orc_findmaxeig(const orc_lattice *L, const orc_c (*U)[3][3], double mass, const orc_c (*start)[3],
                   double tol, double *max) {
  const long Vh = L->Vh;
  vec3 *p = (vec3 *)malloc(sizeof(vec3) * Vh), *r = (vec3 *)malloc(sizeof(vec3) * Vh);
  double *dv = (double *)malloc(sizeof(double) * Vh);
  memcpy(p, start, sizeof(vec3) * Vh);
  for (long i = 0; i < Vh; i++) dv[i] = v_norm2(p[i]);
  double norm = sqrt(orc_global_sum(dv, Vh)), rel;
  int loops = 0;
  do {
    double inorm = 1.0 / norm;
    for (long i = 0; i < Vh; i++)
      for (int c = 0; c < 3; c++) r[i][c] = p[i][c] = p[i][c] * inorm;
    double old = norm;
    orc_mdagm(L, U, mass, p, (const orc_c(*)[3])r);
    for (long i = 0; i < Vh; i++) dv[i] = v_norm2(p[i]);
    norm = sqrt(orc_global_sum(dv, Vh));
    rel = fabs(old - norm) / norm;
    loops++;
  } while (rel > tol);
  *max = norm;
  free(p); free(r); free(dv);
  return loops;
}

/* findminmax.cc:77-137 */
int orc_findmineig(const orc_lattice *L, const orc_c (*U)[3][3], double mass, const orc_c (*start)[3],
                   double tol, double max, double *min) {
  const long Vh = L->Vh;
  vec3 *p = (vec3 *)malloc(sizeof(vec3) * Vh), *r = (vec3 *)malloc(sizeof(vec3) * Vh);
  double *dv = (double *)malloc(sizeof(double) * Vh);
  memcpy(p, start, sizeof(vec3) * Vh);
  for (long i = 0; i < Vh; i++) dv[i] = v_norm2(p[i]);
  double norm = sqrt(orc_global_sum(dv, Vh)), rel;
  const double m2 = mass * mass;
  int loops = 0;
  do {
    double inorm = 1.0 / norm;
    for (long i = 0; i < Vh; i++)
      for (int c = 0; c < 3; c++) r[i][c] = p[i][c] = p[i][c] * inorm;
    double old = norm;
    orc_mdagm(L, U, mass, p, (const orc_c(*)[3])r);
    for (long i = 0; i < Vh; i++)
      for (int c = 0; c < 3; c++) p[i][c] = max * r[i][c] - p[i][c];
    (void)m2;
    for (long i = 0; i < Vh; i++) dv[i] = v_norm2(p[i]);
    norm = sqrt(orc_global_sum(dv, Vh));
    rel = fabs(old - norm) / norm;
    loops++;
  } while (rel > tol);
  *min = max - norm;
  free(p); free(r); free(dv);
  return loops;
}

/* rationalapprox.cc:134-263 */
void orc_rescale_approx(int order, double a0_in, const double *a_in, const double *b_in, double min,
                        double max, double exponent, int pad, double *a0, double *a, double *b) {
  if (pad) { min *= 0.95; max *= 1.05; }
  (void)min;
  double eps = pow(max, exponent);
  *a0 = a0_in * eps;
  for (int i = 0; i < order; i++) {
    a[i] = a_in[i] * (max * eps);
    b[i] = b_in[i] * max;
  }
}

/* ------------------------------------------------------------------ */
/* force, integrators, trajectory                                      */

int orc_full_force(const orc_lattice *L, const orc_c (*U)[3][3], double beta, double mass, int n_pf,
                   const orc_c (*chi)[3], int order, double a0, const double *a, const double *b,
                   double res, int max_cg, int with_gauge, orc_c (*ipdot)[3][3]) {
  (void)a0;
  const long Vh = L->Vh, NL = 4 * L->V;
  su3 *F = (su3 *)calloc(NL, sizeof(su3));
  vec3 *x = (vec3 *)malloc(sizeof(vec3) * Vh * order);
  int iters = 0;
  for (int ps = 0; ps < n_pf; ps++) {
    iters += orc_multishift(L, U, mass, chi + ps * Vh, order, b, res, max_cg, x);
    orc_fermion_force_accum(L, U, order, a, (const orc_c(*)[3])x, F);
  }
  orc_ipdot_total(L, U, beta, (const orc_c(*)[3][3])F, with_gauge, ipdot);
  free(F); free(x);
  return iters;
}

typedef struct {
  const orc_lattice *L;
  const orc_traj_params *prm;
  orc_c (*U)[3][3];
  orc_c (*P)[3][3];
  su3 *ipdot;
  const orc_c (*chi)[3];
  double md_a0, md_a[64], md_b[64];
} orc_md;

static void md_Q(orc_md *m, double c) { orc_conf_left_exp_multiply(m->L, m->U, (const orc_c(*)[3][3])m->P, c); }
static void md_P_full(orc_md *m, double c) {
  const orc_traj_params *q = m->prm;
  if (q->pure_gauge)
    orc_ipdot_gauge(m->L, (const orc_c(*)[3][3])m->U, q->beta, m->ipdot);
  else
    orc_full_force(m->L, (const orc_c(*)[3][3])m->U, q->beta, q->mass, q->n_pf, m->chi, q->order_md, m->md_a0,
                   m->md_a, m->md_b, q->res_md, q->max_cg, 1, m->ipdot);
  orc_momenta_sum_multiply(m->L, m->P, (const orc_c(*)[3][3])m->ipdot, c);
}
static void md_P_fermion(orc_md *m, double c) {
  const orc_traj_params *q = m->prm;
  orc_full_force(m->L, (const orc_c(*)[3][3])m->U, q->beta, q->mass, q->n_pf, m->chi, q->order_md, m->md_a0,
                 m->md_a, m->md_b, q->res_md, q->max_cg, 0, m->ipdot);
  orc_momenta_sum_multiply(m->L, m->P, (const orc_c(*)[3][3])m->ipdot, c);
}
static void md_P_gauge(orc_md *m, double c) {
  orc_ipdot_gauge(m->L, (const orc_c(*)[3][3])m->U, m->prm->beta, m->ipdot);
  orc_momenta_sum_multiply(m->L, m->P, (const orc_c(*)[3][3])m->ipdot, c);
}

static const double OMELYAN_LAMBDA = 0.1931833275037836;

/* minimum_norm2.cc:28-105 */
static void orc_minimum_norm2_A(orc_md *m) {
  const double eps = 1.0 / m->prm->no_md, l = OMELYAN_LAMBDA;
  md_Q(m, eps * l);
  for (int md = 1; md < m->prm->no_md; md++) {
    md_P_full(m, -0.5 * eps);
    md_Q(m, eps * (1.0 - 2.0 * l));
    md_P_full(m, -0.5 * eps);
    md_Q(m, eps * 2.0 * l);
  }
  md_P_full(m, -0.5 * eps);
  md_Q(m, eps * (1.0 - 2.0 * l));
  md_P_full(m, -0.5 * eps);
  md_Q(m, eps * l);
}

/* multistep_2MN.cc:31-103 */
static void orc_multistep_2MN_gauge(orc_md *m, double scale) {
  const double eps = 1.0 / m->prm->no_md, l = OMELYAN_LAMBDA;
  md_P_gauge(m, -eps * scale * l);
  for (int md = 1; md < m->prm->gauge_scale; md++) {
    md_Q(m, 0.5 * eps * scale);
    md_P_gauge(m, -eps * (1.0 - 2.0 * l) * scale);
    md_Q(m, 0.5 * eps * scale);
    md_P_gauge(m, -eps * 2.0 * l * scale);
  }
  md_Q(m, 0.5 * eps * scale);
  md_P_gauge(m, -eps * (1.0 - 2.0 * l) * scale);
  md_Q(m, 0.5 * eps * scale);
  md_P_gauge(m, -eps * l * scale);
}

/* multistep_2MN.cc:104-168 */
static void orc_multistep_2MN(orc_md *m) {
  const double eps = 1.0 / m->prm->no_md, l = OMELYAN_LAMBDA;
  const double gs = 0.5 / (double)m->prm->gauge_scale;
  md_P_fermion(m, -eps * l);
  for (int md = 1; md < m->prm->no_md; md++) {
    orc_multistep_2MN_gauge(m, gs);
    md_P_fermion(m, -eps * (1.0 - 2.0 * l));
    orc_multistep_2MN_gauge(m, gs);
    md_P_fermion(m, -eps * 2.0 * l);
  }
  orc_multistep_2MN_gauge(m, gs);
  md_P_fermion(m, -eps * (1.0 - 2.0 * l));
  orc_multistep_2MN_gauge(m, gs);
  md_P_fermion(m, -eps * l);
}

/* 4MN coefficients, multistep_4MN.cc:37-40 (Takaishi, de Forcrand) */
static const double MN4_THETA = 0.08398315262876693, MN4_RHO = 0.2539785108410595,
                    MN4_LAMBDA = 0.6822365335719091, MN4_MU = -0.03230286765269967;

/* multistep_4MN.cc:30-165: gauge-only 4MN at step eps*scale; the trailing
 * theta kick of one step merges with the leading one of the next */
static void orc_multistep_4MN_gauge(orc_md *m, double scale) {
  const double e = scale / m->prm->no_md;
  const double th = MN4_THETA, rh = MN4_RHO, la = MN4_LAMBDA, mu = MN4_MU;
  const double hf = 0.5 * (1.0 - 2.0 * (la + th)), mid = 1.0 - 2.0 * (mu + rh);
  md_P_gauge(m, -e * th);
  for (int md = 1; md < m->prm->gauge_scale; md++) {
    md_Q(m, e * rh); md_P_gauge(m, -e * la);
    md_Q(m, e * mu); md_P_gauge(m, -e * hf);
    md_Q(m, e * mid); md_P_gauge(m, -e * hf);
    md_Q(m, e * mu); md_P_gauge(m, -e * la);
    md_Q(m, e * rh); md_P_gauge(m, -e * 2.0 * th);
  }
  md_Q(m, e * rh); md_P_gauge(m, -e * la);
  md_Q(m, e * mu); md_P_gauge(m, -e * hf);
  md_Q(m, e * mid); md_P_gauge(m, -e * hf);
  md_Q(m, e * mu); md_P_gauge(m, -e * la);
  md_Q(m, e * rh); md_P_gauge(m, -e * th);
}

/* multistep_4MN.cc:167-300: fermion kicks around nested gauge 4MN with
 * gs = 1/gauge_scale */
static void orc_multistep_4MN(orc_md *m) {
  const double eps = 1.0 / m->prm->no_md, gs = 1.0 / (double)m->prm->gauge_scale;
  const double th = MN4_THETA, rh = MN4_RHO, la = MN4_LAMBDA, mu = MN4_MU;
  const double hf = 0.5 * (1.0 - 2.0 * (la + th)), mid = 1.0 - 2.0 * (mu + rh);
  md_P_fermion(m, -eps * th);
  for (int md = 1; md < m->prm->no_md; md++) {
    orc_multistep_4MN_gauge(m, rh * gs); md_P_fermion(m, -eps * la);
    orc_multistep_4MN_gauge(m, mu * gs); md_P_fermion(m, -eps * hf);
    orc_multistep_4MN_gauge(m, mid * gs); md_P_fermion(m, -eps * hf);
    orc_multistep_4MN_gauge(m, mu * gs); md_P_fermion(m, -eps * la);
    orc_multistep_4MN_gauge(m, rh * gs); md_P_fermion(m, -eps * 2.0 * th);
  }
  orc_multistep_4MN_gauge(m, rh * gs); md_P_fermion(m, -eps * la);
  orc_multistep_4MN_gauge(m, mu * gs); md_P_fermion(m, -eps * hf);
  orc_multistep_4MN_gauge(m, mid * gs); md_P_fermion(m, -eps * hf);
  orc_multistep_4MN_gauge(m, mu * gs); md_P_fermion(m, -eps * la);
  orc_multistep_4MN_gauge(m, rh * gs); md_P_fermion(m, -eps * th);
}

/* leapfrog.cc:18-62 with the momentum sign of the live integrators
 * (minimum_norm2.cc:52): P <- P - i eps ipdot. */
static void orc_leapfrog(orc_md *m) {
  const double eps = 1.0 / m->prm->no_md;
  md_Q(m, 0.5 * eps);
  for (int md = 1; md < m->prm->no_md; md++) {
    md_P_full(m, -eps);
    md_Q(m, eps);
  }
  md_P_full(m, -eps);
  md_Q(m, 0.5 * eps);
}

/* update.cc:58-152 without the Metropolis decision.
 * Rational functions: rationalapprox.cc:269-369. */
double orc_trajectory(const orc_lattice *L, const orc_traj_params *q, orc_c (*U)[3][3], orc_c (*P)[3][3],
                      const orc_c (*phi)[3], orc_c (*chi_out)[3], orc_c (*phi_new_out)[3]) {
  const long V = L->V, Vh = L->Vh, NL = 4 * V;
  const int npf = q->n_pf;
  double *value = (double *)malloc(sizeof(double) * V);
  vec3 *chi = (vec3 *)calloc((size_t)npf * Vh, sizeof(vec3));
  vec3 *phin = (vec3 *)calloc((size_t)npf * Vh, sizeof(vec3));
  int maxo = q->order_first > q->order_last ? q->order_first : q->order_last;
  vec3 *x = (vec3 *)malloc(sizeof(vec3) * Vh * (maxo > 0 ? maxo : 1));
  double a0, a[64], b[64];

  if (!q->pure_gauge) {
    /* first_inv_approx_calc: chi = r_first(M^dag M) phi */
    orc_rescale_approx(q->order_first, q->first_a0, q->first_a, q->first_b, q->eig_min, q->eig_max,
                       (double)q->n_flavours / 8.0 / (double)npf, 1, &a0, a, b);
    for (int ps = 0; ps < npf; ps++) {
      orc_multishift(L, (const orc_c(*)[3][3])U, q->mass, phi + ps * Vh, q->order_first, b, q->res_metro,
                     q->max_cg, x);
      for (long i = 0; i < Vh; i++)
        for (int c = 0; c < 3; c++) {
          orc_c v = a0 * phi[ps * Vh + i][c];
          for (int k = 0; k < q->order_first; k++) v += a[k] * x[(long)k * Vh + i][c];
          chi[ps * Vh + i][c] = v;
        }
    }
  }
  orc_calc_action(L, (const orc_c(*)[3][3])U, (const orc_c(*)[3][3])P, q->beta, npf, phi,
                  (const orc_c(*)[3])chi, !q->pure_gauge, value, 0);

  orc_md m;
  m.L = L; m.prm = q; m.U = U; m.P = P;
  m.ipdot = (su3 *)malloc(sizeof(su3) * NL);
  m.chi = (const orc_c(*)[3])chi;
  if (!q->pure_gauge)
    orc_rescale_approx(q->order_md, q->md_a0, q->md_a, q->md_b, q->eig_min, q->eig_max,
                       -(double)q->n_flavours / 4.0 / (double)npf, 1, &m.md_a0, m.md_a, m.md_b);
  if (q->integrator == ORC_INT_OMELYAN) orc_minimum_norm2_A(&m);
  else if (q->integrator == ORC_INT_MULTISTEP_2MN) orc_multistep_2MN(&m);
  else if (q->integrator == ORC_INT_MULTISTEP_4MN) orc_multistep_4MN(&m);
  else orc_leapfrog(&m);

  if (!q->pure_gauge) {
    /* last_inv_approx_calc: phi_new = r_last(M^dag M) chi */
    orc_rescale_approx(q->order_last, q->last_a0, q->last_a, q->last_b, q->eig_min, q->eig_max,
                       -(double)q->n_flavours / 4.0 / (double)npf, 1, &a0, a, b);
    for (int ps = 0; ps < npf; ps++) {
      orc_multishift(L, (const orc_c(*)[3][3])U, q->mass, (const orc_c(*)[3])(chi + ps * Vh), q->order_last, b,
                     q->res_metro, q->max_cg, x);
      for (long i = 0; i < Vh; i++)
        for (int c = 0; c < 3; c++) {
          orc_c v = a0 * chi[ps * Vh + i][c];
          for (int k = 0; k < q->order_last; k++) v += a[k] * x[(long)k * Vh + i][c];
          phin[ps * Vh + i][c] = v;
        }
    }
  }
  orc_calc_action(L, (const orc_c(*)[3][3])U, (const orc_c(*)[3][3])P, q->beta, npf, (const orc_c(*)[3])phin,
                  (const orc_c(*)[3])chi, !q->pure_gauge, value, -1);
  double dH = -orc_global_sum(value, V);
  if (chi_out) memcpy(chi_out, chi, sizeof(vec3) * Vh * npf);
  if (phi_new_out) memcpy(phi_new_out, phin, sizeof(vec3) * Vh * npf);
  free(value); free(chi); free(phin); free(x); free(m.ipdot);
  return dH;
}
