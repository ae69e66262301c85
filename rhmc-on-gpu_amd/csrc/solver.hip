// solver.hip -- staggered Dslash and the multi-shift CG of the RHMC path.
//
// Dslash: D_eo/D_oe of lib/Action/fermionmatrix.cc:8-97
//   (D psi)(x) = 1/2 sum_mu [ U_mu(x) psi(x+mu) - U_mu(x-mu)^dag psi(x-mu) ],
// with M^dag M = m^2 - D_eo D_oe (lib/Action/Inverter/inverter.cc:27-34).
// One thread per output site, NR right-hand sides per thread sharing every
// link load; neighbours are computed from the site index.
//
// Multi-shift CG: Jegerlehner recurrences of
// lib/Action/Inverter/inverter.cc:142-339, in two organisations:
// * stored residual basis (ms_cg_basis, the default): the base CG in its
//   single-reduction form, two launches per iteration -- the hop
//   h = D_oe r^(k-1) with |h|^2 partials (its first workgroups advance the
//   scalar recurrences), then M^dag M with the three-term r^(k) update and
//   |r^(k)|^2 partials; the shifted solutions are assembled from the basis
//   at the end; a full basis is continued from its last residual
//   (ms_basis_solve);
// * explicit shifts (ms_cg, the reference's organisation): hop, M^dag M with
//   <p, A p>, r update with |r|^2, p and per-shift x_s / p_s updates.
// Reductions are FP64: wave64 shuffles and one partial per workgroup, summed
// in a fixed order by the next launch (or a one-workgroup finalize), so no
// device-scope fence is needed and the host only polls a running-system
// count every few iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

#include "context.hpp"

namespace rhmc {

// ------------------------------------------------------------------ workspace
enum Slot {
  S_PARTIAL, S_SUMS, S_MS, S_CG, S_DIAG,
  S_R64, S_P64, S_S64, S_H64, S_PS64,
  S_B32, S_R32, S_P32, S_S32, S_H32, S_PS32, S_X32,
  S_CR64, S_CX32, S_CR32, S_CP32, S_CS32, S_CH32,
  S_PW1, S_PW2, S_PW3, S_PWN,
  S_FH64,
  S_PARTIAL2, S_BASIS, S_BSH, S_HIST, S_COEF, S_HPART,
  S_FORCE_H, S_FORCE_COEF, S_SOLSUM_COEF,
  S_RS_SRC, S_RS_Y, S_RS_IDX, S_GROUP_TAB,
  S_COUNT
};

// an instantiated graph of one solver chunk and the launch arguments it was
// captured with
struct GraphSlot {
  hipGraphExec_t exec = nullptr;
  size_t key_bytes = 0;
  unsigned char key[4096];
};

// per (rhs, iteration, shift) of the basis multi-shift solver: omega, zeta, gamma
struct BmsHist {
  double om, z, ga;
};

struct DsArgs;

struct Workspace {
  std::map<unsigned long long, int> tuned_nr;  // Dslash launch shape per (precision, links, mode, red, nvec)
  std::map<unsigned long long, double> tuned_ms;  // the chosen shape's time per launch when timed here
  std::string tune_path;                       // tuning cache file ("" = none)
  std::string arch;                            // GPU architecture, e.g. gfx950
  bool tune_init = false;                      // cache path from RHMC_TUNE_CACHE looked up and loaded
  GraphSlot bms_graph[2];                      // FP32 / FP64 stored-basis chunk
  GraphSlot group_graph[2];                    // the same for a group of lattices led by this context
  std::vector<DsArgs> group_tab;               // argument tables of the group's chunk (host copy)
  std::map<unsigned long long, int> group_tuned;  // lockstep launch shape per group composition
  long basis_kmax[2] = {0, 0};                 // most basis iterations a FP32 / FP64 solve needed
  void *buf[S_COUNT] = {};
  size_t cap[S_COUNT] = {};
  unsigned *counter = nullptr;
  int *h_flags = nullptr;     // pinned
  double *h_sums = nullptr;   // pinned
  // pinned staging of the solvers' small host <-> device transfers (a
  // pageable copy costs the runtime a staging pass, tens of microseconds to
  // a millisecond with the GPU idle): a ring for copies to the device (an
  // event behind the latest one; the ring is reused only after it has
  // executed) and a buffer for copies to the host (read after a sync)
  char *h2d_ring = nullptr;
  size_t h2d_cap = 0, h2d_off = 0;
  hipEvent_t h2d_ev = nullptr;
  bool h2d_used = false;
  char *d2h_buf = nullptr;
  size_t d2h_cap = 0;
  hipEvent_t poll_ev[2] = {}; // solver chunk polls, one chunk in flight ahead
};

Workspace *workspace_create(Context &c) {
  Workspace *w = new Workspace();
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c.device) == hipSuccess) {
    w->arch = prop.gcnArchName;
    const size_t colon = w->arch.find(':');
    if (colon != std::string::npos) w->arch.resize(colon);
  }
  if (!hip_ok(c, hipMalloc(&w->counter, 64 * sizeof(unsigned)), "hipMalloc counter")) return nullptr;
  if (!hip_ok(c, hipMemsetAsync(w->counter, 0, 64 * sizeof(unsigned), c.stream), "counter")) return nullptr;
  if (!hip_ok(c, hipHostMalloc(&w->h_flags, 4096 * sizeof(int), hipHostMallocDefault), "hipHostMalloc")) return nullptr;
  if (!hip_ok(c, hipHostMalloc(&w->h_sums, 4096 * sizeof(double), hipHostMallocDefault), "hipHostMalloc")) return nullptr;
  if (!hip_ok(c, hipEventCreateWithFlags(&w->h2d_ev, hipEventDisableTiming), "staging event")) return nullptr;
  for (auto &e : w->poll_ev)
    if (!hip_ok(c, hipEventCreateWithFlags(&e, hipEventDisableTiming), "poll event")) return nullptr;
  // reduction partials for the largest launch (kMaxRed vectors), allocated once
  const size_t nbx = (size_t)((c.Vh + kBlock - 1) / kBlock);
  w->cap[S_PARTIAL] = sizeof(double) * nbx * 256;
  if (!hip_ok(c, hipMalloc(&w->buf[S_PARTIAL], w->cap[S_PARTIAL]), "hipMalloc partials")) return nullptr;
  return w;
}

void workspace_destroy(Workspace *w) {
  if (!w) return;
  for (auto &g : w->bms_graph)
    if (g.exec) hipGraphExecDestroy(g.exec);
  for (auto &g : w->group_graph)
    if (g.exec) hipGraphExecDestroy(g.exec);
  for (int i = 0; i < S_COUNT; i++)
    if (w->buf[i]) hipFree(w->buf[i]);
  if (w->counter) hipFree(w->counter);
  if (w->h_flags) hipHostFree(w->h_flags);
  if (w->h_sums) hipHostFree(w->h_sums);
  for (auto e : w->poll_ev)
    if (e) hipEventDestroy(e);
  if (w->h2d_ev) hipEventDestroy(w->h2d_ev);
  if (w->h2d_ring) hipHostFree(w->h2d_ring);
  if (w->d2h_buf) hipHostFree(w->d2h_buf);
  delete w;
}

static void *ws_get(Context &c, int slot, size_t bytes) {
  Workspace *w = c.ws;
  if (w->cap[slot] >= bytes) return w->buf[slot];
  if (w->buf[slot]) {
    hipStreamSynchronize(c.stream);
    if (c.force_coef_dev == w->buf[slot]) c.force_coef_dev = nullptr;  // its held coefficients go with it
    hipFree(w->buf[slot]);
    w->buf[slot] = nullptr;
    w->cap[slot] = 0;
  }
  void *p = nullptr;
  if (!hip_ok(c, hipMalloc(&p, bytes), "hipMalloc workspace")) return nullptr;
  w->buf[slot] = p;
  w->cap[slot] = bytes;
  return p;
}

// host -> device through the pinned ring, enqueued on c.stream
bool h2d(Context &c, void *dst, const void *src, size_t bytes) {
  Workspace *w = c.ws;
  const size_t need = (bytes + 255) & ~(size_t)255;
  if (w->h2d_off + need > w->h2d_cap) {
    // wrap: every earlier copy from the ring has executed once the event
    // behind the latest one has
    if (w->h2d_used && !hip_ok(c, hipEventSynchronize(w->h2d_ev), "staging")) return false;
    w->h2d_off = 0;
    if (need > w->h2d_cap) {
      if (w->h2d_ring) hipHostFree(w->h2d_ring);
      w->h2d_ring = nullptr;
      w->h2d_cap = 0;
      const size_t cap = std::max<size_t>(need, 4u << 20);
      if (!hip_ok(c, hipHostMalloc((void **)&w->h2d_ring, cap, hipHostMallocDefault), "staging ring")) return false;
      w->h2d_cap = cap;
    }
  }
  char *stage = w->h2d_ring + w->h2d_off;
  w->h2d_off += need;
  memcpy(stage, src, bytes);
  if (!hip_ok(c, hipMemcpyAsync(dst, stage, bytes, hipMemcpyHostToDevice, c.stream), "staged copy")) return false;
  w->h2d_used = true;
  return hip_ok(c, hipEventRecord(w->h2d_ev, c.stream), "staging event");
}

// pinned host buffer of at least `bytes` for copies to the host; valid
// until the next call (the caller syncs the stream before reading it)
static void *d2h_stage(Context &c, size_t bytes) {
  Workspace *w = c.ws;
  if (w->d2h_cap < bytes) {
    if (w->d2h_buf) {
      hipStreamSynchronize(c.stream);
      hipHostFree(w->d2h_buf);
    }
    w->d2h_buf = nullptr;
    w->d2h_cap = 0;
    const size_t cap = std::max<size_t>(bytes, 1u << 20);
    if (!hip_ok(c, hipHostMalloc((void **)&w->d2h_buf, cap, hipHostMallocDefault), "staging buffer")) return nullptr;
    w->d2h_cap = cap;
  }
  return w->d2h_buf;
}

// device -> host through the pinned buffer; synchronises the stream
bool d2h_sync(Context &c, void *dst, const void *src, size_t bytes) {
  void *stage = d2h_stage(c, bytes);
  if (!stage) return false;
  if (!hip_ok(c, hipMemcpyAsync(stage, src, bytes, hipMemcpyDeviceToHost, c.stream), "staged copy")) return false;
  if (!hip_ok(c, hipStreamSynchronize(c.stream), "staged copy")) return false;
  memcpy(dst, stage, bytes);
  return true;
}

// workspace buffers of the gauge-field side (fields.hip), kept for the
// context's lifetime instead of an allocation per MD force
void *scratch_get(Context &c, int use, size_t bytes) {
  static const int slot[SCR_COUNT] = {S_FORCE_H, S_FORCE_COEF, S_SOLSUM_COEF};
  return use >= 0 && use < SCR_COUNT ? ws_get(c, slot[use], bytes) : nullptr;
}

// ------------------------------------------------------------- reductions
constexpr int kMaxRed = 256;  // vectors reduced in one launch

// Partials are written with plain stores and consumed by a separate
// one-workgroup finalize launch: the kernel boundary makes them visible
// without a device-scope fence per workgroup (such a fence writes back the
// XCD's L2 and serialises tens of thousands of workgroups).
__device__ __forceinline__ void store_partial(double *p, double v) { *p = v; }
__device__ __forceinline__ double load_partial(const double *p) { return *p; }

// ----------------------------------------------------- work decomposition
// Reducing launches use `nbr` workgroups per vector, each owning a contiguous
// range of sites (a multiple of kBlock); fewer partials keep the finalize
// short while the grid still holds >= 8 workgroups per CU.
__device__ __forceinline__ void site_range(long Vh, unsigned bx, unsigned nbr, long &h0, long &h1) {
  const long nb = (Vh + kBlock - 1) / kBlock;
  const long per = (nb + nbr - 1) / nbr * kBlock;
  h0 = (long)bx * per;
  h1 = h0 + per < Vh ? h0 + per : Vh;
}

// 1-D grid of nsb site blocks x nch vector chunks, flattened so that the
// workgroups dispatched consecutively to one XCD (hardware ids b, b+8, ...)
// cover the same site block for every chunk: the shared links then stay in
// that XCD's L2.  Each XCD owns a contiguous slab of site blocks.  With 8
// chunks or more (the force's hop over every pole and pseudofermion) the
// chunks' vectors no longer fit the caches together, and the grid runs
// chunk by chunk instead, each chunk over the XCDs' slabs (128 right-hand
// sides at 32^3x8: 343 -> 315 us; with 4 chunks of 2 it costs 16 us).
__device__ __forceinline__ void xcd_split(unsigned b, unsigned nsb, unsigned nch, unsigned &sb, unsigned &ch) {
  if (nsb % 8 == 0 && nch >= 8) {
    const unsigned xcd = b % 8, l = b / 8, per = nsb / 8;
    sb = xcd * per + l % per;
    ch = l / per;
  } else if (nsb % 8 == 0) {
    const unsigned xcd = b % 8, l = b / 8;
    sb = xcd * (nsb / 8) + l / nch;
    ch = l % nch;
  } else {
    sb = b % nsb;
    ch = b / nsb;
  }
}

// block-reduce NV values per thread and store one partial per vector
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double *partial, int jbase, int nvec, int nbx,
                                               int bx) {
  __shared__ double sh[kBlock / 64][NV];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < NV; r++) {
    double s = wave_sum(v[r]);
    if (lane == 0) sh[w][r] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int r = threadIdx.x, j = jbase + r;
    if (j < nvec) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < kBlock / 64; k++) s += sh[k][r];
      partial[(long)j * nbx + bx] = s;
    }
  }
}

// deterministic final sums: wave w reduces vectors w, w+4, ...
__device__ __forceinline__ void final_sums(const double *partial, int nvec, int nbx, double *out_sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int j = w; j < nvec; j += kBlock / 64) {
    double s = 0.0;
    for (int b = lane; b < nbx; b += 64) s += partial[(long)j * nbx + b];
    s = wave_sum(s);
    if (lane == 0) out_sh[j] = s;
  }
  __syncthreads();
}

// ------------------------------------------------------------- epilogues
enum Epi { EPI_STORE = 0, EPI_MS_INIT, EPI_MS_B, EPI_MS_C, EPI_CG_INIT, EPI_CG_B, EPI_CG_C };

struct RedArgs {
  double *partial;
  unsigned *counter;
  double *sums;
  void *state;
  int epi;
  int nbr;  // partials per vector
};

// Multi-shift recurrences of inverter.cc:186-284, one thread per right-hand
// side for the shared scalars and one per (rhs, shift) for the shifted ones.
__device__ void ms_epilogue(int epi, MsState *st, const double *sums, int nvec) {
  __shared__ int live[kMaxRed];
  __shared__ double g_new[kMaxRed], delta_old[kMaxRed];
  const int ns = st[0].nshift;
  for (int j = threadIdx.x; j < nvec; j += blockDim.x) {
    MsState &s = st[j];
    const double v = sums[j];
    if (epi == EPI_MS_INIT) {
      s.delta = v;
      s.omega = 1.0;
      s.omega_prev = 1.0;
      s.gammag = 0.0;
      s.iter = 0;
      s.nact = s.nshift;
      s.done = (v == 0.0) ? 1 : 0;  // zero source: x = 0
      s.stop = s.done;
      live[j] = 1;
    } else if (epi == EPI_MS_B) {
      if (s.stop) s.done = 1;  // the iteration that set stop has completed
      live[j] = !s.done;
      if (live[j]) {
        s.omega_prev = s.omega;
        s.omega = -s.delta / v;  // omega = -(r,r)/(p,Ap)
      }
    } else {  // EPI_MS_C
      live[j] = !s.done;
      delta_old[j] = s.delta;
      g_new[j] = v / s.delta;  // gamma = (r_new,r_new)/(r,r)
    }
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < nvec * ns; idx += blockDim.x) {
    const int j = idx / ns, k = idx - j * ns;
    if (!live[j]) continue;
    MsState &s = st[j];
    if (epi == EPI_MS_INIT) {
      s.z1[k] = 1.0; s.z2[k] = 1.0; s.z3[k] = 1.0;
      s.om[k] = 0.0; s.ga[k] = 0.0;
      s.flag[k] = 1; s.flag_next[k] = 1; s.its[k] = 0;
    } else if (epi == EPI_MS_B) {
      const int f = s.flag_next[k];
      s.flag[k] = f;
      if (!f) continue;
      const double z1 = s.z1[k], z2 = s.z2[k], om = s.omega, omp = s.omega_prev;
      const double z3 = (z1 * z2 * omp) / (om * s.gammag * (z1 - z2) + z1 * omp * (1.0 - s.shift[k] * om));
      s.z3[k] = z3;
      s.om[k] = om * z3 / z2;
    } else {
      if (!s.flag[k]) continue;
      const double z2 = s.z2[k], z3 = s.z3[k];
      s.ga[k] = g_new[j] * z3 * s.om[k] / (z2 * s.omega);
      const double fact = sqrt(delta_old[j] * z2 * z2);
      if (fact < s.res) {
        s.flag_next[k] = 0;
        s.its[k] = s.iter + 1;
      }
      s.z1[k] = z2;
      s.z2[k] = z3;
    }
  }
  if (epi != EPI_MS_C) return;
  __syncthreads();
  for (int j = threadIdx.x; j < nvec; j += blockDim.x) {
    if (!live[j]) continue;
    MsState &s = st[j];
    int nact = 0;
    for (int k = 0; k < ns; k++) nact += s.flag_next[k];
    s.nact = nact;
    s.gammag = g_new[j];
    s.delta = sums[j];
    s.iter++;
    s.stop = !(sqrt(sums[j]) > s.res && s.iter < s.max_iter) || nact == 0;
    if (!isfinite(sums[j]) || !isfinite(s.omega)) s.nonfinite = s.stop = 1;
  }
}

__device__ void cg_epilogue(int epi, CgState *st, const double *sums, int nvec) {
  for (int j = threadIdx.x; j < nvec; j += blockDim.x) {
    CgState &s = st[j];
    const double v = sums[j];
    if (epi == EPI_CG_INIT) {
      s.delta = v;
      s.iter = 0;
      s.done = (v <= s.tol2) ? 1 : 0;
      s.stop = s.done;
      if (!isfinite(v)) s.nonfinite = s.done = s.stop = 1;
    } else if (epi == EPI_CG_B) {
      if (s.stop) s.done = 1;
      if (s.done) continue;
      s.alpha = s.delta / v;
    } else {
      if (s.done) continue;
      s.beta = v / s.delta;
      s.delta = v;
      s.iter++;
      s.stop = (v <= s.tol2) || s.iter >= s.max_iter;
      if (!isfinite(v) || !isfinite(s.alpha)) s.nonfinite = s.stop = 1;
    }
  }
}

// one workgroup: fixed-order sums of the per-workgroup partials, then the
// scalar recurrences; *active receives the number of systems still running.
__global__ __launch_bounds__(kBlock) void k_finalize(RedArgs ra, int nvec, int *active) {
  __shared__ double sums[kMaxRed];
  final_sums(ra.partial, nvec, ra.nbr, sums);
  const bool ms = ra.epi == EPI_MS_INIT || ra.epi == EPI_MS_B || ra.epi == EPI_MS_C;
  const bool cg = ra.epi == EPI_CG_INIT || ra.epi == EPI_CG_B || ra.epi == EPI_CG_C;
  if (ms) ms_epilogue(ra.epi, (MsState *)ra.state, sums, nvec);
  else if (cg) cg_epilogue(ra.epi, (CgState *)ra.state, sums, nvec);
  else
    for (int j = threadIdx.x; j < nvec; j += blockDim.x) ra.sums[j] = sums[j];
  if (active) {
    __syncthreads();
    const int t = threadIdx.x;
    const int run = t < nvec ? (ms ? !((MsState *)ra.state)[t].stop : !((CgState *)ra.state)[t].stop) : 0;
    const int n = __syncthreads_count(run);
    if (t == 0) *active = n;
  }
}

// fixed-order workgroup sum of n partials
// per-lane sums over b = threadIdx.x + m kBlock (b < n) of NV arrays, in
// order of m, with the loads of four strides of every array in flight
// together (adding the 0 of a stride past n changes nothing)
template <int NV>
__device__ __forceinline__ void lane_sums(const double *const (&p)[NV], int n, double (&v)[NV]) {
  for (int b0 = threadIdx.x; b0 < n; b0 += 4 * kBlock) {
    double x[4][NV];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int r = 0; r < NV; r++) x[u][r] = b0 + u * kBlock < n ? p[r][b0 + u * kBlock] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int r = 0; r < NV; r++) v[r] += x[u][r];
  }
}

__device__ __forceinline__ double block_sum_partials(const double *p, int n) {
  __shared__ double sh[kBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const double *pp[1] = {p};
  double vv[1] = {0.0};
  lane_sums<1>(pp, n, vv);
  double v = wave_sum(vv[0]);
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kBlock / 64; k++) s += sh[k];
  __syncthreads();
  return s;
}

// ------------------------------------------------------------------ Dslash
// DS_BMS: the stored-basis solver's M^dag M with the residual update
enum DsMode { DS_HOP = 0, DS_MDAGM = 1, DS_RESID = 2, DS_BMS = 3 };
enum DsRed { RED_NONE = 0, RED_DOT = 1, RED_NORM = 2 };
enum LinkFmt { LINK_D18 = 0, LINK_S12 = 1, LINK_D12 = 2 };

// scalars of one system of the stored-basis multi-shift CG shared between
// its launches (the algorithm is described at ms_cg_basis)
struct BmsShared {
  double rho[2];    // [k&1] at the hop of iteration k: rho_{k-2}
  double alpha[2];  // [k&1] at iteration k: the alpha of iteration k-1
  int stop[2];      // [k&1]: system stopped before iteration k
  double rho1;      // rho_{k-1}, written by the hop of iteration k (rho_0 by the set-up)
  double beta;      // rho_{k-1} / rho_{k-2} (0 at k = 1), written with rho1
};

struct DsArgs {
  LatticeDims L;
  int par_out;
  const void *U;
  const void *in; long in_vs;
  void *out; long out_vs;
  const void *pin; long pin_vs;        // MDAGM/RESID: vector at the output site
  const double2 *b; long b_vs; int b_group;  // RESID
  const double *diag;                  // per vector m^2 + shift
  const int *skip; int skip_stride;    // stop flags (ints), may be null
  int nvec;
  unsigned nsb, nch;                   // site blocks x vector chunks
  const int *tiles; int ntiles;        // traversal order of the kBlock-site tiles
  // iteration-indexed operand: in (iter_in) and pin (iter_pin) advance by
  // (*iterp + iter_off) * iter_stride elements (a basis slot chosen on the device)
  const int *iterp; int iter_off; long iter_stride; int iter_in, iter_pin;
  // live kernel timing (solver hop, device wall clock): slot s = (*iterp +
  // iter_off) of clk holds the end of the launch before the hop and the end
  // of the hop.  A timed launch stores each workgroup's end in clk_wg (plain
  // stores, no contention); the first workgroup of the launch after it takes
  // the maximum of those clk_nprev stamps (clk_prev) into field clk_field of
  // slot s.
  unsigned long long *clk; int clk_cap;
  unsigned long long *clk_wg;
  const unsigned long long *clk_prev; int clk_nprev, clk_field;
  // stored-basis CG (iteration k = *iterp + iter_off + 1): the hop (bms_ms
  // set) has workgroup j < nvec advance system j's scalars from the rho
  // partials of the previous M^dag M launch (bms_nbr per system); the
  // DS_BMS launch forms alpha from the hop's |h|^2 partials (bms_hnbr per
  // system) and writes r^(k) into the basis slot after pin's, below
  // bms_cap, from pin's slot and the one before it (out is unused), with
  // |r^(k)|^2 partials in red
  BmsShared *bms; const double *bms_part; int bms_nbr;
  MsState *bms_ms; BmsHist *bms_hist; int *bms_active; int bms_cap, bms_nshift;
  const double *bms_hpart; int bms_hnbr;
  RedArgs red;
};

// Hop of iteration k, workgroup j mod the grid: system j's scalars.  rho_{k-1}
// is the fixed-order sum of the previous M^dag M launch's partials (at k = 1
// the set-up's rho_0), beta = rho_{k-1} / rho_{k-2}; then the Jegerlehner
// recurrences of iteration k-1 (inverter.cc:224-287, they need alpha and
// beta of the base CG's step k-2) and the stopping test, whose result
// applies from iteration k+1: the stop decision lags one iteration.  The
// workgroup of system 0 also clears the running-system count of iteration k
// (slot k&1) while the others add to that of k+1.  Values the other
// launches of iteration k read (stop, rho, alpha) are double-buffered by
// the parity of k.
__device__ __forceinline__ void bms_step(const DsArgs &a, int j) {
  const int k = *a.iterp + a.iter_off + 1;
  const int sl = k & 1, sn = sl ^ 1;
  const int t = threadIdx.x;
  BmsShared &sh = a.bms[j];
  if (j == 0 && t == 0) a.bms_active[sl] = 0;
  if (sh.stop[sl]) {
    if (t == 0) {
      sh.stop[sn] = 1;
      sh.rho[sn] = sh.rho[sl];
      sh.alpha[sn] = sh.alpha[sl];
    }
    return;
  }
  const double rho1 = k == 1 ? sh.rho1 : block_sum_partials(a.bms_part + (long)j * a.bms_nbr, a.bms_nbr);
  const double rho2 = sh.rho[sl], al2 = sh.alpha[sl];  // rho_{k-2}, alpha of iteration k-1
  const double be = k >= 2 ? rho1 / rho2 : 0.0;
  MsState &s = a.bms_ms[j];
  int stop = 0;
  if (k >= 2) {
    // recurrences of iteration kk = k-1: omega = -alpha, gamma = beta
    const int kk = k - 1;
    const double omp = s.omega, gold = s.gammag;
    const double om = -al2, gam = be;
    if (t < s.nshift && s.flag[t]) {
      const double z1 = s.z1[t], z2 = s.z2[t];
      const double z3 = (z1 * z2 * omp) / (om * gold * (z1 - z2) + z1 * omp * (1.0 - s.shift[t] * om));
      const double oms = om * z3 / z2;
      const double gas = gam * z3 * oms / (z2 * om);
      BmsHist &hh = a.bms_hist[((long)j * a.bms_cap + kk) * a.bms_nshift + t];
      hh.om = oms;
      hh.z = z3;
      hh.ga = gas;
      if (sqrt(rho2 * z2 * z2) < s.res) {
        s.flag[t] = 0;
        s.its[t] = kk;
      }
      s.z1[t] = z2;
      s.z2[t] = z3;
    }
    const int nact = __syncthreads_count(t < s.nshift && s.flag[t]);
    if (t == 0) {
      s.omega_prev = omp;
      s.omega = om;
      s.gammag = gam;
      s.delta = rho1;
      s.iter = kk;
      s.nact = nact;
      // r^(k) stays inside the basis: the last M^dag M launch that writes
      // runs at k = cap-1
      stop = !(sqrt(rho1) > s.res && kk < s.max_iter) || nact == 0 || kk + 2 >= a.bms_cap;
      // NaN/Inf in the reductions or the shifted recurrences: stop and report
      if (!isfinite(rho1) || !isfinite(be) || !isfinite(om)) s.nonfinite = stop = 1;
      s.stop = s.done = stop;
    }
  }
  if (t == 0) {
    sh.rho1 = rho1;
    sh.beta = be;
    sh.rho[sn] = rho1;
    sh.stop[sn] = stop;
    if (!stop) atomicAdd(&a.bms_active[sn], 1);
  }
}

// M^dag M of iteration k (DS_BMS), every workgroup: alpha and beta of its NR
// systems.  mu = <r, M^dag M r> = diag rho_{k-1} + |D_oe r|^2 (the staggered
// D is anti-hermitian, D_eo^+ = -D_oe), from the hop's |h|^2 partials summed
// in a fixed order, so every workgroup forms the same value; <p, A p> = mu -
// beta rho_{k-1} / alpha_prev (single-reduction CG), alpha = rho_{k-1} /
// <p, A p>, into ab (workgroup memory: [0] alpha, [1] the three-term
// coefficient c1 = alpha beta / alpha_prev, read by the epilogue).  The first site block's workgroup of each chunk stores alpha
// for iteration k+1.  A system that is stopped is not used (its output is
// not stored).
template <typename R, int NR>
__device__ __forceinline__ void bms_alpha(const DsArgs &a, int jbase, unsigned sb, R (&ab)[2][NR]) {
  const int k = *a.iterp + a.iter_off + 1;
  const int sl = k & 1, sn = sl ^ 1;
  const int n = a.bms_hnbr;
  __shared__ double sm[kBlock / 64][NR];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double v[NR];
  const double *pr[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) {
    v[r] = 0.0;
    const int j = jbase + r < a.nvec ? jbase + r : a.nvec - 1;
    pr[r] = a.bms_hpart + (long)j * n;
  }
  lane_sums<NR>(pr, n, v);
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const double q = wave_sum(v[r]);
    if (lane == 0) sm[w][r] = q;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int j = jbase + r < a.nvec ? jbase + r : a.nvec - 1;
    double h2 = 0.0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; q++) h2 += sm[q][r];
    const BmsShared &sh = a.bms[j];
    const double rho1 = sh.rho1, b = sh.beta;
    const double mu = a.diag[j] * rho1 + h2;
    const double pap = k >= 2 ? mu - b * rho1 / sh.alpha[sl] : mu;
    const double alpha = rho1 / pap;
    // three-term residual update r^(k) = r^(k-1) - alpha w - c1 (r^(k-2) -
    // r^(k-1)): c1 = alpha beta / alpha_prev carries beta s_{k-2} = beta
    // (r^(k-2) - r^(k-1)) / alpha_prev
    const double ap = sh.alpha[sl];
    const double c1 = k >= 2 && ap != 0.0 ? alpha * b / ap : 0.0;
    if (threadIdx.x == 0) {
      ab[0][r] = (R)alpha;
      ab[1][r] = (R)c1;
      if (sb == 0 && jbase + r < a.nvec && !sh.stop[sl]) a.bms[j].alpha[sn] = alpha;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ unsigned long long *clk_slot(const DsArgs &a) {
  if (!a.clk || !a.iterp) return nullptr;
  const int s = *a.iterp + a.iter_off;
  return s >= 0 && s < a.clk_cap ? a.clk + 2 * (long)s : nullptr;
}
// workgroup maximum of n end stamps, stored by thread 0 into *dst
__device__ __forceinline__ void clk_max(const unsigned long long *p, int n, unsigned long long *dst) {
  __shared__ unsigned long long sm[kBlock / 64];
  unsigned long long m = 0;
  for (int i = threadIdx.x; i < n; i += kBlock) m = p[i] > m ? p[i] : m;
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long q = __shfl_xor(m, o);
    m = q > m ? q : m;
  }
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < kBlock / 64; k++) m = sm[k] > m ? sm[k] : m;
    *dst = m;
  }
}
// first workgroup: the previous launch's end into this launch's slot
__device__ __forceinline__ void clk_begin(const DsArgs &a, unsigned bid) {
  if (!a.clk_prev || bid != 0) return;
  unsigned long long *t = clk_slot(a);
  if (t) clk_max(a.clk_prev, a.clk_nprev, t + a.clk_field);
}
// every workgroup: its end stamp
__device__ __forceinline__ void clk_end(const DsArgs &a, unsigned bid) {
  if (!a.clk_wg) return;
  __syncthreads();
  if (threadIdx.x == 0) a.clk_wg[bid] = (unsigned long long)wall_clock64();
}

// element offset of the iteration-indexed operand
__device__ __forceinline__ long ds_iter_shift(const DsArgs &a) {
  return a.iterp ? (long)(*a.iterp + a.iter_off) * a.iter_stride : 0;
}

// A link as stored: rows 0-1 for the 12-real formats (row 2 is rebuilt when
// the link is used, so the loads can be issued a hop ahead of the FMAs),
// all nine entries for FP64-18.
template <typename R, int LINK>
struct LinkRaw {
  static constexpr int N = LINK == LINK_D18 ? 9 : LINK == LINK_D12 ? 6 : 3;
  using T = typename std::conditional<LINK == LINK_S12, float4, double2>::type;
  T e[N];
};

template <typename R, int LINK>
__device__ __forceinline__ void link_raw_load(const DsArgs &a, int mu, int par, int h, LinkRaw<R, LINK> &w) {
  using T = typename LinkRaw<R, LINK>::T;
  constexpr int N = LinkRaw<R, LINK>::N;
  const T *b = (const T *)a.U + (long)(mu * 2 + par) * N * a.L.Vh;
  const unsigned off = (unsigned)h * (unsigned)sizeof(T);
#pragma unroll
  for (int k = 0; k < N; k++) w.e[k] = ld_off(b + (long)k * a.L.Vh, off);
}

template <typename R, int LINK>
__device__ __forceinline__ Mat<R> link_raw_mat(const LinkRaw<R, LINK> &w, R eta) {
  Mat<R> m;
  if constexpr (LINK == LINK_S12) {
    m.e[0][0] = make_float2(w.e[0].x, w.e[0].y);
    m.e[0][1] = make_float2(w.e[0].z, w.e[0].w);
    m.e[0][2] = make_float2(w.e[1].x, w.e[1].y);
    m.e[1][0] = make_float2(w.e[1].z, w.e[1].w);
    m.e[1][1] = make_float2(w.e[2].x, w.e[2].y);
    m.e[1][2] = make_float2(w.e[2].z, w.e[2].w);
    rebuild_row2(m, eta);
  } else if constexpr (LINK == LINK_D12) {
#pragma unroll
    for (int k = 0; k < 6; k++) m.e[k / 3][k % 3] = w.e[k];
    rebuild_row2(m, eta);
  } else {
#pragma unroll
    for (int k = 0; k < 9; k++) m.e[k / 3][k % 3] = w.e[k];
  }
  return m;
}

// One hop's operands: the link and the neighbour's colour vectors of every
// right-hand side of the chunk.
template <typename R, int LINK, int NR>
struct HopIn {
  LinkRaw<R, LINK> U;
  Cx<R> v[NR][3];
};

// hop k = 2 mu + (0 forward, 1 backward) of the site c (half index h):
// forward  U_mu(x) psi(x+mu), link at x (output parity);
// backward -U_mu(x-mu)^dag psi(x-mu), link at x-mu (the other parity)
template <typename R, int LINK, int NR>
__device__ __forceinline__ void hop_load(const DsArgs &a, const Cx<R> *const *inr, const SiteCoord &c, int h, int k,
                                         HopIn<R, LINK, NR> &d, bool spinors = true) {
  const int mu = k >> 1;
  const bool fwd = (k & 1) == 0;
  const int hn = neighbour_i(a.L, c, mu, fwd ? +1 : -1);
  link_raw_load<R, LINK>(a, mu, fwd ? a.par_out : 1 - a.par_out, fwd ? h : hn, d.U);
  if (!spinors) return;
  const long Vh = a.L.Vh;
  const unsigned off = (unsigned)hn * (unsigned)sizeof(Cx<R>);
#pragma unroll
  for (int r = 0; r < NR; r++) {
    d.v[r][0] = ld_off(inr[r], off);
    d.v[r][1] = ld_off(inr[r] + Vh, off);
    d.v[r][2] = ld_off(inr[r] + 2 * Vh, off);
  }
}

// Lane i of each 16-lane row receives lane (i-1) mod 16 (DPP row_ror:1).
__device__ __forceinline__ float row_ror1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x121, 0xF, 0xF, false));
}
__device__ __forceinline__ double row_ror1(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), 0x121, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x121, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// With 16 half-sites per x-row (nx = 32) a wave holds whole rows, one per
// 16-lane DPP row, and the x-1 neighbour of a site is the x+1 neighbour of
// the site one lane to the left in the same row (cyclically, which is the
// periodic wrap in x): the backward x hop takes its spinors from the
// forward hop's registers instead of loading them.
template <typename R, int LINK, int NR>
__device__ __forceinline__ void xm_from_xp(const HopIn<R, LINK, NR> &xp, HopIn<R, LINK, NR> &xm) {
#pragma unroll
  for (int r = 0; r < NR; r++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      xm.v[r][i].x = row_ror1(xp.v[r][i].x);
      xm.v[r][i].y = row_ror1(xp.v[r][i].y);
    }
}

// Ends a hop: the accumulators pass through an empty asm statement (in-out
// operands, memory clobber), so the hop's multiply-adds are issued before
// the loads that follow it in program order instead of being sunk to the
// end of the site with every hop's operands live at once.
template <typename R, int NR>
__device__ __forceinline__ void hop_fence(Cx<R> (&acc)[NR][3]) {
#pragma unroll
  for (int r = 0; r < NR; r++)
#pragma unroll
    for (int i = 0; i < 3; i++) asm volatile("" : "+v"(acc[r][i].x), "+v"(acc[r][i].y) : : "memory");
}

template <typename R, int LINK, int NR>
__device__ __forceinline__ void hop_apply(const LatticeDims &L, const SiteCoord &c, int k, const HopIn<R, LINK, NR> &d,
                                          Cx<R> (&acc)[NR][3]) {
  const int mu = k >> 1;
  if ((k & 1) == 0) {
    const Mat<R> U = link_raw_mat<R, LINK>(d.U, (R)eta_phase(L, c.x, c.y, c.z, c.t, mu));
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        cfma(acc[r][i], U.e[i][0], d.v[r][0]);
        cfma(acc[r][i], U.e[i][1], d.v[r][1]);
        cfma(acc[r][i], U.e[i][2], d.v[r][2]);
      }
  } else {
    const int tm = (mu == 3) ? (c.t == 0 ? L.nt - 1 : c.t - 1) : c.t;
    const Mat<R> U = link_raw_mat<R, LINK>(d.U, (R)eta_phase(L, c.x, c.y, c.z, tm, mu));
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        cfmsc(acc[r][i], U.e[0][i], d.v[r][0]);
        cfmsc(acc[r][i], U.e[1][i], d.v[r][1]);
        cfmsc(acc[r][i], U.e[2][i], d.v[r][2]);
      }
  }
}

// device link array for a link format
template <int LINK>
static const void *link_ptr(Context &c) {
  return LINK == LINK_D18 ? (const void *)c.U64 : LINK == LINK_D12 ? (const void *)c.U64c : (const void *)c.U32;
}

// Waves per SIMD the Dslash kernels are compiled for: the registers of two
// hops in flight (link + NR neighbour vectors each), the NR accumulators and
// ~48 for addresses, coordinates and the x-row path must fit the wave's
// share of the 512.
template <typename R, int LINK, int NR>
constexpr int ds_waves() {
  constexpr int rs = (int)sizeof(R) / 4;
  constexpr int link = LINK == LINK_D18 ? 36 : LINK == LINK_D12 ? 24 : 12;
  constexpr int hop = link + NR * 6 * rs;
  constexpr int need = 2 * hop + NR * 6 * rs + 48;
  constexpr int w = 512 / need;
  return w < 1 ? 1 : (w > 8 ? 8 : w);
}

// Output of one site for the chunk's right-hand sides: the hop result (times
// 1/2), or m^2 + shift times the input minus it (M^dag M), or the residual
// b - M^dag M x, with the dot/norm partials of the reducing modes (the
// stored-basis solver's DS_BMS output is ds_site_out_bms).
template <typename R, int NR, int MODE, int RED>
__device__ __forceinline__ void ds_site_out(const DsArgs &a, const Cx<R> (&acc)[NR][3], const bool (&act)[NR],
                                            double (&red)[NR], int jbase, int h, const Cx<R> *pinb) {
  using C = Cx<R>;
  const long Vh = a.L.Vh;
#pragma unroll
  for (int r = 0; r < NR; r++) {
    if (!act[r]) continue;
    const int j = jbase + r;
    C *o = (C *)a.out + (long)j * a.out_vs + h;
    C res[3];
    if (MODE == DS_HOP) {
#pragma unroll
      for (int i = 0; i < 3; i++) res[i] = cscale(acc[r][i], (R)0.5);
    } else {
      const R d = (R)a.diag[j];
      const C *p = pinb + (long)j * a.pin_vs + h;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const C pv = p[(long)i * Vh];
        res[i].x = fmar(d, pv.x, (R)-0.5 * acc[r][i].x);
        res[i].y = fmar(d, pv.y, (R)-0.5 * acc[r][i].y);
        if (MODE == DS_RESID) {
          const double2 bv = a.b[(long)(j / a.b_group) * a.b_vs + (long)i * Vh + h];
          res[i] = csub(mk<R>((R)bv.x, (R)bv.y), res[i]);
        }
        if (RED == RED_DOT) red[r] += (double)pv.x * (double)res[i].x + (double)pv.y * (double)res[i].y;
      }
    }
#pragma unroll
    for (int i = 0; i < 3; i++) o[(long)i * Vh] = res[i];
    if (RED == RED_NORM) {
      // the site's |v|^2 in the working precision, accumulated in FP64
      R n2 = (R)0;
#pragma unroll
      for (int i = 0; i < 3; i++) n2 = fmar(res[i].x, res[i].x, fmar(res[i].y, res[i].y, n2));
      red[r] += (double)n2;
    }
  }
}

// DS_BMS output of one site, the three-term form of the CG residual update:
// w = m^2 r^(k-1) - D_eo h (= M^dag M r^(k-1)), r^(k) = r^(k-1) - alpha w +
// c1 (r^(k-1) - r^(k-2)) into basis slot k (when rk_ok), |r^(k)|^2 partials.
// It equals r^(k-1) - alpha s_{k-1} with s_{k-1} = w + beta s_{k-2} and
// s_{k-2} = (r^(k-2) - r^(k-1)) / alpha_prev, so no search-direction vector
// is stored.  bms_site_loads issues the site's loads (r^(k-1) and, from the
// second iteration, r^(k-2) of every right-hand side of the chunk; a slot
// past nvec reads the last vector): in FP32 before the site's hops, so that
// they are in flight with the first hop's (38.3 against 39.4 us per launch
// at 32^3x8, 8 right-hand sides), in FP64 after them (the registers they
// would hold through the hops cost the FP64 launch a wave per SIMD).
template <typename R, int NR>
__device__ __forceinline__ void bms_site_loads(const DsArgs &a, int jbase, int h, const Cx<R> *pinb, bool use_prev,
                                               Cx<R> (&pv)[NR][3], Cx<R> (&qv)[NR][3]) {
  using C = Cx<R>;
  const long Vh = a.L.Vh;
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int j = jbase + r < a.nvec ? jbase + r : a.nvec - 1;
    const C *p = pinb + (long)j * a.pin_vs + h;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      pv[r][i] = p[(long)i * Vh];
      qv[r][i] = use_prev ? p[(long)i * Vh - a.iter_stride] : pv[r][i];
    }
  }
}

template <typename R, int NR>
__device__ __forceinline__ void ds_site_out_bms(const DsArgs &a, const Cx<R> (&acc)[NR][3], const bool (&act)[NR],
                                                double (&red)[NR], int jbase, int h, const Cx<R> *pinb,
                                                const R (&ab)[2][NR], bool rk_ok, const Cx<R> (&pv)[NR][3],
                                                const Cx<R> (&qv)[NR][3]) {
  using C = Cx<R>;
  const long Vh = a.L.Vh;
#pragma unroll
  for (int r = 0; r < NR; r++) {
    if (!act[r]) continue;
    const int j = jbase + r;
    C *rk = (C *)(pinb + a.iter_stride) + (long)j * a.pin_vs + h;
    const R d = (R)a.diag[j];
    const R al = ab[0][r], c1 = ab[1][r];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      C w, rn;
      w.x = fmar(d, pv[r][i].x, (R)-0.5 * acc[r][i].x);
      w.y = fmar(d, pv[r][i].y, (R)-0.5 * acc[r][i].y);
      rn.x = fmar(c1, pv[r][i].x - qv[r][i].x, fmar(-al, w.x, pv[r][i].x));
      rn.y = fmar(c1, pv[r][i].y - qv[r][i].y, fmar(-al, w.y, pv[r][i].y));
      if (rk_ok) rk[(long)i * Vh] = rn;
      red[r] += (double)rn.x * (double)rn.x + (double)rn.y * (double)rn.y;
    }
  }
}

// One workgroup (bid of nblk) of the one-lane Dslash: the body of k_dslash
// and of the multi-lattice k_dslash_group.
template <typename R, int LINK, int NR, int MODE, int RED>
__device__ __forceinline__ void ds_body(const DsArgs &a, unsigned bid, unsigned nblk) {
  using C = Cx<R>;
  if (MODE == DS_HOP && a.bms_ms)
    for (int j = bid; j < a.nvec; j += nblk) bms_step(a, j);  // a small grid takes several
  clk_begin(a, bid);
  const LatticeDims &L = a.L;
  const long Vh = L.Vh;
  unsigned sb, ch;
  xcd_split(bid, a.nsb, a.nch, sb, ch);
  const int jbase = ch * NR;
  __shared__ R bab[2][NR];
  if (MODE == DS_BMS) bms_alpha<R, NR>(a, jbase, sb, bab);
  const bool rk_ok = MODE == DS_BMS && *a.iterp + a.iter_off + 1 < a.bms_cap;
  // iteration k >= 2 of the stored-basis solver reads basis slot k-2 too
  const bool use_prev = MODE == DS_BMS && *a.iterp + a.iter_off + 1 >= 2;
  // this workgroup's run of tiles in the traversal order
  const int per = (a.ntiles + (int)a.nsb - 1) / (int)a.nsb;
  const int q0 = (int)sb * per;
  const int q1 = q0 + per < a.ntiles ? q0 + per : a.ntiles;

  bool act[NR];
  bool any = false;
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int j = jbase + r;
    act[r] = (j < a.nvec) && (a.skip == nullptr || !a.skip[(long)j * a.skip_stride]);
    any |= act[r];
  }
  double red[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) red[r] = 0.0;

  const int Vh32 = (int)Vh;
  // every right-hand side of the chunk is loaded and multiplied without a
  // branch, so the loads of a whole hop are in flight together; a slot past
  // nvec reads the last vector and a stopped system is computed but not
  // stored
  const long ish = ds_iter_shift(a);
  const C *inb = (const C *)a.in + (a.iter_in ? ish : 0);
  const C *pinb = (const C *)a.pin + (a.iter_pin ? ish : 0);
  const C *inr[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int j = jbase + r < a.nvec ? jbase + r : a.nvec - 1;
    inr[r] = inb + (long)j * a.in_vs;
  }
  const bool xrow = L.nxh == 16;
  // the hop's grid has one workgroup per tile: a single pass, so that the
  // reduction's accumulators are not carried through a loop around the hops
  const int qend = MODE == DS_HOP ? (q0 + 1 < q1 ? q0 + 1 : q1) : q1;
  for (int q = q0; any && q < qend; q++) {
    const int h = (a.tiles ? a.tiles[q] : q) * kBlock + (int)threadIdx.x;
    if (h >= Vh32) continue;
    const SiteCoord c = site_coords_i(L, a.par_out, h);
    C acc[NR][3];
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
      for (int k = 0; k < 3; k++) acc[r][k] = mk<R>(0, 0);
    // software pipeline over the eight hops: the operands of hop k+1 are in
    // flight while hop k is multiplied
    constexpr bool kBmsEarly = MODE == DS_BMS && sizeof(R) == 4;
    C pv[MODE == DS_BMS ? NR : 1][3], qv[MODE == DS_BMS ? NR : 1][3];
    if constexpr (kBmsEarly) bms_site_loads<R, NR>(a, jbase, h, pinb, use_prev, pv, qv);
    HopIn<R, LINK, NR> buf[2];
    hop_load<R, LINK, NR>(a, inr, c, h, 0, buf[0]);
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (k + 1 < 8) hop_load<R, LINK, NR>(a, inr, c, h, k + 1, buf[(k + 1) & 1], k + 1 != 1 || !xrow);
      hop_apply<R, LINK, NR>(L, c, k, buf[k & 1], acc);
      if (k == 0 && xrow) xm_from_xp<R, LINK, NR>(buf[0], buf[1]);
      hop_fence<R, NR>(acc);
    }
    if constexpr (MODE == DS_BMS) {
      if constexpr (!kBmsEarly) bms_site_loads<R, NR>(a, jbase, h, pinb, use_prev, pv, qv);
      ds_site_out_bms<R, NR>(a, acc, act, red, jbase, h, pinb, bab, rk_ok, pv, qv);
    }
    else ds_site_out<R, NR, MODE, RED>(a, acc, act, red, jbase, h, pinb);
  }
  if (RED != RED_NONE) block_partials<NR>(red, a.red.partial, jbase, a.nvec, a.nsb, sb);
  clk_end(a, bid);
}

template <typename R, int LINK, int NR, int MODE, int RED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(ds_waves<R, LINK, NR>()))) void k_dslash(DsArgs a) {
  ds_body<R, LINK, NR, MODE, RED>(a, blockIdx.x, gridDim.x);
}

// Dslash traversal order.  The hardware hands consecutive workgroups to the
// eight XCDs in turn and xcd_split gives each XCD one contiguous eighth of the
// site blocks, so each eighth of the tile list is one XCD's share.  Listing
// the kBlock-site tiles by sub-block of a (y, z, t) decomposition into eight
// equal sub-blocks makes each share compact: the neighbours a share reads
// from other XCDs' L2 slices shrink from two full t-slices (share = one
// t-slice) to the sub-block surface.  The split with the smallest surface
// among those that give every sub-block the same number of whole tiles is
// taken; without one the identity order is kept.
int build_tile_order(Context &c) {
  const LatticeDims &L = c.L;
  const int T = (int)((c.Vh + kBlock - 1) / kBlock);
  std::vector<int> order(T);
  for (int q = 0; q < T; q++) order[q] = q;
  int best[3] = {1, 1, 1};
  double best_surface = 1e300;
  const int dims[3] = {L.ny, L.nz, L.nt};
  if (c.Vh % kBlock == 0 && T % 8 == 0) {
    for (int sy = 1; sy <= 8; sy *= 2)
      for (int sz = 1; sy * sz <= 8; sz *= 2) {
        const int st = 8 / (sy * sz);
        const int sp[3] = {sy, sz, st};
        bool ok = true;
        double surface = 0.0;
        for (int d = 0; d < 3; d++) {
          if (dims[d] % sp[d]) ok = false;
          if (sp[d] > 1) surface += 2.0 * sp[d] / dims[d];
        }
        if (!ok || surface >= best_surface) continue;
        // every tile inside one sub-block, T/8 tiles per sub-block
        std::vector<int> count(8, 0);
        for (int q = 0; q < T && ok; q++) {
          int id0 = -1;
          for (long h = (long)q * kBlock; h < (long)(q + 1) * kBlock; h += L.nxh) {
            const long row = h / L.nxh;
            const int y = (int)(row % L.ny), z = (int)(row / L.ny % L.nz), t = (int)(row / L.ny / L.nz);
            const int id = y / (L.ny / sy) + sy * (z / (L.nz / sz) + sz * (t / (L.nt / st)));
            if (id0 < 0) id0 = id;
            else if (id != id0) { ok = false; break; }
          }
          if (ok) count[id0]++;
        }
        for (int k = 0; k < 8 && ok; k++) ok = count[k] == T / 8;
        if (!ok) continue;
        best_surface = surface;
        best[0] = sy; best[1] = sz; best[2] = st;
      }
  }
  if (best_surface < 1e300) {
    const int sy = best[0], sz = best[1], st = best[2];
    std::vector<int> key(T);
    for (int q = 0; q < T; q++) {
      const long row = (long)q * kBlock / L.nxh;
      const int y = (int)(row % L.ny), z = (int)(row / L.ny % L.nz), t = (int)(row / L.ny / L.nz);
      key[q] = y / (L.ny / sy) + sy * (z / (L.nz / sz) + sz * (t / (L.nt / st)));
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
  }
  c.tile_split[0] = best[0]; c.tile_split[1] = best[1]; c.tile_split[2] = best[2];
  c.tiles_host = order;
  c.ntiles = T;
  if (getenv("RHMC_TILE_IDENTITY")) {
    c.tiles = nullptr;
    return 0;
  }
  c.tiles = (int *)dev_alloc(c, sizeof(int) * T);
  if (!c.tiles) return -1;
  hipMemcpyAsync(c.tiles, order.data(), sizeof(int) * T, hipMemcpyHostToDevice, c.stream);
  return hip_ok(c, hipStreamSynchronize(c.stream), "tile order") ? 0 : -1;
}

// ---- lane-split Dslash: four lanes per site, lane q takes direction mu = q
// (forward and backward hop), the partial sums are combined across the quad
// with DPP moves and lanes 0-2 write one colour each.  Four times the waves
// of the one-lane form at the same right-hand sides per lane, with each link
// still loaded once per site for all NR right-hand sides.
__device__ __forceinline__ int dpp_quad_xor1(int v) { return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false); }
__device__ __forceinline__ int dpp_quad_xor2(int v) { return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false); }
__device__ __forceinline__ float quad_sum(float v) {
  v += __int_as_float(dpp_quad_xor1(__float_as_int(v)));
  v += __int_as_float(dpp_quad_xor2(__float_as_int(v)));
  return v;
}
__device__ __forceinline__ double quad_xor(double v, bool two) {
  const long long b = __double_as_longlong(v);
  int lo = (int)(b & 0xffffffffLL), hi = (int)(b >> 32);
  lo = two ? dpp_quad_xor2(lo) : dpp_quad_xor1(lo);
  hi = two ? dpp_quad_xor2(hi) : dpp_quad_xor1(hi);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double quad_sum(double v) {
  v += quad_xor(v, false);
  v += quad_xor(v, true);
  return v;
}

constexpr int kLsSites = kBlock / 4;  // sites per workgroup pass

// value select among three (by value, so that the accumulators stay in
// registers instead of an indexed private array)
template <typename C>
__device__ __forceinline__ C pick3(C a0, C a1, C a2, int k) {
  C r;
  r.x = k == 0 ? a0.x : (k == 1 ? a1.x : a2.x);
  r.y = k == 0 ? a0.y : (k == 1 ? a1.y : a2.y);
  return r;
}

// one workgroup (bid of nblk) of the lane-split Dslash (k_dslash_ls and the
// multi-lattice k_dslash_ls_group)
template <typename R, int LINK, int NR, int MODE, int RED>
__device__ __forceinline__ void ds_ls_body(const DsArgs &a, unsigned bid, unsigned nblk) {
  using C = Cx<R>;
  if (MODE == DS_HOP && a.bms_ms)
    for (int j = bid; j < a.nvec; j += nblk) bms_step(a, j);  // a small grid takes several
  clk_begin(a, bid);
  const LatticeDims &L = a.L;
  const long Vh = L.Vh;
  const int Vh32 = (int)Vh;
  unsigned sb, ch;
  xcd_split(bid, a.nsb, a.nch, sb, ch);
  const int jbase = ch * NR;
  __shared__ R bab[2][NR];
  if (MODE == DS_BMS) bms_alpha<R, NR>(a, jbase, sb, bab);
  const bool rk_ok = MODE == DS_BMS && *a.iterp + a.iter_off + 1 < a.bms_cap;
  const bool use_prev = MODE == DS_BMS && *a.iterp + a.iter_off + 1 >= 2;
  const int nunits = a.ntiles * (kBlock / kLsSites);
  const int per = (nunits + (int)a.nsb - 1) / (int)a.nsb;
  const int q0 = (int)sb * per;
  const int q1 = q0 + per < nunits ? q0 + per : nunits;
  const int mu = threadIdx.x & 3;
  const int slot = threadIdx.x >> 2;

  bool act[NR];
  bool any = false;
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int j = jbase + r;
    act[r] = (j < a.nvec) && (a.skip == nullptr || !a.skip[(long)j * a.skip_stride]);
    any |= act[r];
  }
  double red[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) red[r] = 0.0;
  // every right-hand side of the chunk is loaded and multiplied without a
  // branch, so the loads of a whole hop are in flight together; a slot past
  // nvec reads the last vector and a stopped system is computed but not
  // stored
  const long ish = ds_iter_shift(a);
  const C *inb = (const C *)a.in + (a.iter_in ? ish : 0);
  const C *pinb = (const C *)a.pin + (a.iter_pin ? ish : 0);
  const C *inr[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int j = jbase + r < a.nvec ? jbase + r : a.nvec - 1;
    inr[r] = inb + (long)j * a.in_vs;
  }

  // one workgroup per tile quarter for the hop: a single pass (as k_dslash)
  const int qend = MODE == DS_HOP ? (q0 + 1 < q1 ? q0 + 1 : q1) : q1;
  for (int q = q0; any && q < qend; q++) {
    const int tile = a.tiles ? a.tiles[q >> 2] : (q >> 2);
    const int h = tile * kBlock + (q & 3) * kLsSites + slot;
    const bool valid = h < Vh32;
    const int hs = valid ? h : 0;
    C acc[NR][3];
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
      for (int k = 0; k < 3; k++) acc[r][k] = mk<R>(0, 0);
    if (valid) {
      const SiteCoord c = site_coords_i(L, a.par_out, hs);
      HopIn<R, LINK, NR> fw, bw;
      hop_load<R, LINK, NR>(a, inr, c, hs, 2 * mu, fw);
      hop_load<R, LINK, NR>(a, inr, c, hs, 2 * mu + 1, bw);
      hop_apply<R, LINK, NR>(L, c, 2 * mu, fw, acc);
      hop_apply<R, LINK, NR>(L, c, 2 * mu + 1, bw, acc);
    }
    // all four lanes of every quad take part in the exchange
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        acc[r][i].x = quad_sum(acc[r][i].x);
        acc[r][i].y = quad_sum(acc[r][i].y);
      }
    if (!valid || mu == 3) continue;
    const int col = mu;  // lanes 0-2: one colour each
#pragma unroll
    for (int r = 0; r < NR; r++) {
      if (!act[r]) continue;
      const int j = jbase + r;
      const C av = pick3(acc[r][0], acc[r][1], acc[r][2], col);
      C res;
      if (MODE == DS_HOP) {
        res = cscale(av, (R)0.5);
      } else {
        const C pv = (pinb + (long)j * a.pin_vs + h)[(long)col * Vh];
        const R d = (R)a.diag[j];
        res.x = fmar(d, pv.x, (R)-0.5 * av.x);
        res.y = fmar(d, pv.y, (R)-0.5 * av.y);
        if (MODE == DS_RESID) {
          const double2 bv = a.b[(long)(j / a.b_group) * a.b_vs + (long)col * Vh + h];
          res = csub(mk<R>((R)bv.x, (R)bv.y), res);
        }
        if (RED == RED_DOT) red[r] += (double)pv.x * (double)res.x + (double)pv.y * (double)res.y;
        if (MODE == DS_BMS) {
          // three-term residual update (ds_site_out_bms); w = res is not stored
          const R al = bab[0][r], c1 = bab[1][r];
          const C qv = use_prev ? (pinb + (long)j * a.pin_vs + h)[(long)col * Vh - a.iter_stride] : pv;
          C rn;
          rn.x = fmar(c1, pv.x - qv.x, fmar(-al, res.x, pv.x));
          rn.y = fmar(c1, pv.y - qv.y, fmar(-al, res.y, pv.y));
          if (rk_ok) ((C *)(pinb + a.iter_stride) + (long)j * a.pin_vs + h)[(long)col * Vh] = rn;
          red[r] += (double)rn.x * (double)rn.x + (double)rn.y * (double)rn.y;
          continue;
        }
      }
      ((C *)a.out + (long)j * a.out_vs + h)[(long)col * Vh] = res;
      if (RED == RED_NORM && MODE != DS_BMS) red[r] += (double)res.x * (double)res.x + (double)res.y * (double)res.y;
    }
  }
  if (RED != RED_NONE) block_partials<NR>(red, a.red.partial, jbase, a.nvec, a.nsb, sb);
  clk_end(a, bid);
}

template <typename R, int LINK, int NR, int MODE, int RED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(ds_waves<R, LINK, NR>()))) void k_dslash_ls(DsArgs a) {
  ds_ls_body<R, LINK, NR, MODE, RED>(a, blockIdx.x, gridDim.x);
}

// ---- several lattices in one launch.  The grid is the concatenation of
// every lattice's own grid, each padded to a multiple of 8 workgroups so that
// a lattice's local workgroup index keeps the XCD mapping of xcd_split
// (the hardware deals workgroup ids to the 8 XCDs in turn); workgroup b runs
// lattice l with start[l] <= b < start[l+1] on its arguments tab[l] (read
// from device memory, uniform over the workgroup).  Padding workgroups exit.
constexpr int kMaxGroup = 16;
struct GroupArgs {
  const DsArgs *tab;
  unsigned start[kMaxGroup + 1];
  int n;
};

__device__ __forceinline__ int group_member(const GroupArgs &g, unsigned b) {
  int l = 0;
  while (l + 1 < g.n && b >= g.start[l + 1]) l++;
  return l;
}

template <typename R, int LINK, int NR, int MODE, int RED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(ds_waves<R, LINK, NR>()))) void k_dslash_group(GroupArgs g) {
  const int l = group_member(g, blockIdx.x);
  const DsArgs &a = g.tab[l];
  const unsigned bid = blockIdx.x - g.start[l], nblk = a.nsb * a.nch;
  if (bid >= nblk) return;
  ds_body<R, LINK, NR, MODE, RED>(a, bid, nblk);
}

template <typename R, int LINK, int NR, int MODE, int RED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(ds_waves<R, LINK, NR>()))) void k_dslash_ls_group(GroupArgs g) {
  const int l = group_member(g, blockIdx.x);
  const DsArgs &a = g.tab[l];
  const unsigned bid = blockIdx.x - g.start[l], nblk = a.nsb * a.nch;
  if (bid >= nblk) return;
  ds_ls_body<R, LINK, NR, MODE, RED>(a, bid, nblk);
}

// number of right-hand sides handled per thread for a launch with nvec vectors
static int pick_nr(int nvec) {
  if (nvec >= 8) return 8;
  if (nvec >= 4) return 4;
  if (nvec >= 2) return 2;
  return 1;
}

static unsigned nbx_of(Context &c) { return (unsigned)((c.Vh + kBlock - 1) / kBlock); }

// workgroups per vector for a reducing launch with `ny` vectors/chunks:
// at least 2048 workgroups overall, at most one per kBlock sites
static unsigned reduce_blocks(Context &c, int ny) {
  const unsigned nbx = nbx_of(c);
  unsigned want = (2048u + ny - 1) / ny;
  if (want < 64) want = 64;
  want = (want + 7) / 8 * 8;
  return want < nbx ? want : nbx;
}

// Launch shape code:
//   bits 0-3   right-hand sides per lane (1, 2, 4, 8)
//   bits 4-7   lanes per site (1, or 4 with the lane-split kernel)
//   bit  8     tile order: 0 the XCD sub-block order, 1 lattice order
//   bits 9-11  reducing M^dag M launches: log2 of the kBlock-site tiles per
//              workgroup (0: reduce_blocks' default of one tile per
//              workgroup down to 2048 workgroups overall)
static inline int shape_nr(int code) { return code & 15; }
static inline int shape_lanes(int code) { return (code >> 4) & 15; }
static inline int shape_identity(int code) { return (code >> 8) & 1; }
static inline int shape_tpw_log2(int code) { return (code >> 9) & 7; }
static inline int make_shape(int nr, int lanes, int identity, int tpw_log2) {
  return nr | (lanes << 4) | (identity << 8) | (tpw_log2 << 9);
}

static unsigned long long tune_key(int real_bytes, int link, int mode, int red, int nvec) {
  return ((unsigned long long)real_bytes << 40) | ((unsigned long long)link << 36) |
         ((unsigned long long)mode << 32) | ((unsigned long long)red << 28) | (unsigned long long)nvec;
}

// ---- tuning cache file: a JSON object with one entry per line,
//   {"arch": "gfx950", "dims": [nx, ny, nz, nt], "real_bytes": 4, "links": 1,
//    "mode": 3, "reduction": 2, "nvec": 8, "shape": 530, ...readable fields...}
// keyed by (arch, dims, real_bytes, links, mode, reduction, nvec).
static bool json_num(const std::string &s, const char *field, double &v) {
  const std::string f = std::string("\"") + field + "\":";
  const size_t p = s.find(f);
  if (p == std::string::npos) return false;
  return sscanf(s.c_str() + p + f.size(), " %lf", &v) == 1;
}
static std::string json_str(const std::string &s, const char *field) {
  const std::string f = std::string("\"") + field + "\": \"";
  const size_t p = s.find(f);
  if (p == std::string::npos) return "";
  const size_t e = s.find('"', p + f.size());
  return e == std::string::npos ? "" : s.substr(p + f.size(), e - p - f.size());
}
static std::string dims_json(const LatticeDims &L) {
  char b[96];
  snprintf(b, sizeof b, "[%d, %d, %d, %d]", L.nx, L.ny, L.nz, L.nt);
  return b;
}
static std::string tune_entry_json(Context &c, unsigned long long k, int code, double ms) {
  char b[512];
  snprintf(b, sizeof b,
           "{\"arch\": \"%s\", \"dims\": %s, \"real_bytes\": %llu, \"links\": %llu, \"mode\": %llu, "
           "\"reduction\": %llu, \"nvec\": %llu, \"shape\": %d, \"rhs_per_lane\": %d, \"lanes_per_site\": %d, "
           "\"tile_order\": \"%s\", \"tiles_per_workgroup\": %d, \"ms\": %.6g}",
           c.ws->arch.c_str(), dims_json(c.L).c_str(), k >> 40, (k >> 36) & 15, (k >> 32) & 15, (k >> 28) & 15,
           k & 0xfffffff, code, shape_nr(code), shape_lanes(code), shape_identity(code) ? "lattice" : "xcd",
           1 << shape_tpw_log2(code), ms);
  return b;
}
static bool entry_key(const std::string &line, std::string &arch, std::string &dims, unsigned long long &k,
                      int &code) {
  double rb, li, mo, re, nv, sh;
  if (!json_num(line, "real_bytes", rb) || !json_num(line, "links", li) || !json_num(line, "mode", mo) ||
      !json_num(line, "reduction", re) || !json_num(line, "nvec", nv) || !json_num(line, "shape", sh))
    return false;
  arch = json_str(line, "arch");
  const size_t p = line.find("\"dims\": [");
  const size_t e = p == std::string::npos ? p : line.find(']', p);
  dims = p == std::string::npos || e == std::string::npos ? "" : line.substr(p + 8, e - p - 7);
  k = tune_key((int)rb, (int)li, (int)mo, (int)re, (int)nv);
  code = (int)sh;
  return true;
}
static std::vector<std::string> read_lines(const std::string &path) {
  std::vector<std::string> out;
  FILE *f = fopen(path.c_str(), "r");
  if (!f) return out;
  char buf[1024];
  while (fgets(buf, sizeof buf, f)) out.emplace_back(buf);
  fclose(f);
  return out;
}

// shapes stored for this context's GPU architecture and lattice extents
void tune_cache_load(Context &c) {
  Workspace *w = c.ws;
  if (w->tune_path.empty()) return;
  const std::string mydims = dims_json(c.L);
  for (const std::string &line : read_lines(w->tune_path)) {
    std::string arch, dims;
    unsigned long long k;
    int code;
    if (!entry_key(line, arch, dims, k, code) || arch != w->arch || dims != mydims) continue;
    if (shape_nr(code) < 1 || shape_nr(code) > 8 || (shape_lanes(code) != 1 && shape_lanes(code) != 4)) continue;
    if (shape_identity(code) == 0 && !c.tiles) continue;
    w->tuned_nr.emplace(k, code);
  }
}

// this context's tuned shapes merged into the cache file (entries of other
// architectures or extents kept), written to path.tmp and renamed
void tune_cache_store(Context &c) {
  Workspace *w = c.ws;
  if (w->tune_path.empty()) return;
  const std::string mydims = dims_json(c.L);
  std::vector<std::string> keep;
  for (const std::string &line : read_lines(w->tune_path)) {
    std::string arch, dims;
    unsigned long long k;
    int code;
    if (!entry_key(line, arch, dims, k, code)) continue;
    if (arch == w->arch && dims == mydims && w->tuned_nr.count(k)) continue;
    std::string l = line;
    while (!l.empty() && (l.back() == '\n' || l.back() == ',' || l.back() == ' ')) l.pop_back();
    keep.push_back(l);
  }
  for (auto &kv : w->tuned_nr) {
    auto m = w->tuned_ms.find(kv.first);
    keep.push_back(tune_entry_json(c, kv.first, kv.second, m == w->tuned_ms.end() ? 0.0 : m->second));
  }
  const std::string tmp = w->tune_path + ".tmp";
  FILE *f = fopen(tmp.c_str(), "w");
  if (!f) return;
  fprintf(f, "{\"rhmc_tuning_cache\": 1, \"entries\": [\n");
  for (size_t i = 0; i < keep.size(); i++) fprintf(f, "%s%s\n", keep[i].c_str(), i + 1 < keep.size() ? "," : "");
  fprintf(f, "]}\n");
  fclose(f);
  rename(tmp.c_str(), w->tune_path.c_str());
}

int tune_cache_set(Context &c, const char *path) {
  c.ws->tune_path = path ? path : "";
  tune_cache_load(c);
  return 0;
}

// workgroups per vector chunk of a reducing launch with 2^k tiles per
// workgroup (k = 0: reduce_blocks), a multiple of 8 so that the XCD
// mapping of xcd_split holds
static unsigned reduce_blocks_tpw(Context &c, int nch, int k) {
  if (k == 0) return reduce_blocks(c, nch);
  const unsigned nbx = nbx_of(c);
  unsigned n = (nbx + (1u << k) - 1) >> k;
  n = (n + 7) / 8 * 8;
  return n < nbx ? n : nbx;
}

// the lattice and grid fields of a launch with shape code
template <int MODE, int RED>
static void dslash_prep(Context &c, DsArgs &a, int code) {
  const int nr = shape_nr(code), lanes = shape_lanes(code) == 4 ? 4 : 1;
  a.L = c.L;
  a.tiles = shape_identity(code) ? nullptr : c.tiles;
  a.ntiles = c.ntiles;
  a.nch = (unsigned)((a.nvec + nr - 1) / nr);
  // one workgroup per tile for the hop (its reduction, when any, then has a
  // partial per tile); fewer, longer workgroups for the other reducing modes
  a.nsb = RED == RED_NONE || MODE == DS_HOP ? nbx_of(c) * lanes : reduce_blocks_tpw(c, a.nch, shape_tpw_log2(code));
  a.red.nbr = (int)a.nsb;
}

template <typename R, int LINK, int MODE, int RED>
static void dslash_launch(Context &c, DsArgs &a, int code) {
#define DS_LAUNCH(KERNEL) hipLaunchKernelGGL(KERNEL, grid, dim3(kBlock), 0, c.stream, a)
  const int nr = shape_nr(code), lanes = shape_lanes(code) == 4 ? 4 : 1;
  dslash_prep<MODE, RED>(c, a, code);
  dim3 grid(a.nsb * a.nch);
  if (lanes == 4) {
    switch (nr) {
      case 8: DS_LAUNCH((k_dslash_ls<R, LINK, 8, MODE, RED>)); break;
      case 4: DS_LAUNCH((k_dslash_ls<R, LINK, 4, MODE, RED>)); break;
      case 2: DS_LAUNCH((k_dslash_ls<R, LINK, 2, MODE, RED>)); break;
      default: DS_LAUNCH((k_dslash_ls<R, LINK, 1, MODE, RED>)); break;
    }
    return;
  }
  switch (nr) {
    case 8: DS_LAUNCH((k_dslash<R, LINK, 8, MODE, RED>)); break;
    case 4: DS_LAUNCH((k_dslash<R, LINK, 4, MODE, RED>)); break;
    case 2: DS_LAUNCH((k_dslash<R, LINK, 2, MODE, RED>)); break;
    default: DS_LAUNCH((k_dslash<R, LINK, 1, MODE, RED>)); break;
  }
#undef DS_LAUNCH
}

// Autotuner.  The launch shape trades link reuse in registers against waves
// in flight and cache locality:
// * right-hand sides per lane (1, 2, 4, 8) and lanes per site (1, or 4 with
//   the lane-split kernel): a single 8-RHS lane per site leaves a 32^3x8
//   half lattice at ~2 waves per SIMD;
// * tile order: the XCD sub-block order (build_tile_order) or lattice order;
// * for the reducing M^dag M launches, the tiles per workgroup (1, 2, 4:
//   fewer, longer workgroups and fewer reduction partials).
// On the first launch of each (precision, links, mode, reduction, nvec)
// every candidate is timed in place with HIP events and the fastest is kept
// for the lifetime of the context -- and, when a tuning cache file is set
// (RHMC_TUNE_CACHE or gpu_tune_cache), stored there under (kernel, lattice
// extents, precision, nvec, GPU architecture) and reused by later contexts
// and processes without timing.  The launches are idempotent, except the
// stored-basis solver's hop, which advances the systems' scalar state (the
// candidates skip that step); the r^(k) the candidates of its M^dag M
// launch write is overwritten by the final launch.  RHMC_DSLASH_NR ("4", "4x4" or "2t4")
// forces the right-hand sides per lane, lanes per site and tiles per workgroup.
constexpr int kTuneRounds = 3;

template <typename R, int LINK, int MODE, int RED>
static void dslash_dispatch(Context &c, DsArgs &a) {
  const unsigned long long key = tune_key(sizeof(R), LINK, MODE, RED, a.nvec);
  if (!c.ws->tune_init) {
    c.ws->tune_init = true;
    if (const char *p = getenv("RHMC_TUNE_CACHE"))
      if (c.ws->tune_path.empty()) tune_cache_set(c, p);
  }
  auto &tab = c.ws->tuned_nr;
  auto it = tab.find(key);
  int code;
  if (it != tab.end()) {
    code = it->second;
  } else if (const char *env = getenv("RHMC_DSLASH_NR")) {
    // "<rhs per thread>", "<rhs per thread>x4" for four lanes per site, and
    // "...t<2|4>" for tiles per workgroup in the reducing non-hop launches
    int v = std::max(1, std::min(8, atoi(env)));
    v = v >= 8 ? 8 : v >= 4 ? 4 : v >= 2 ? 2 : 1;
    const char *tp = strchr(env, 't');
    const int tiles = tp && RED != RED_NONE && MODE != DS_HOP ? atoi(tp + 1) : 1;
    code = make_shape(v, strchr(env, 'x') ? 4 : 1, 0, tiles >= 4 ? 2 : tiles >= 2 ? 1 : 0);
    tab[key] = code;
  } else {
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    float best = 1e30f;
    code = make_shape(pick_nr(a.nvec), 1, 0, 0);
    DsArgs t = a;
    t.bms_ms = nullptr;  // the hop's scalar step runs once, in the final launch
    const bool reducing = RED != RED_NONE && MODE != DS_HOP;
    std::vector<int> cands;
    for (int lanes = 1; lanes <= 4; lanes *= 4)
      for (int v = 1; v <= 8 && v <= std::max(1, a.nvec); v *= 2)
        for (int ident = c.tiles ? 0 : 1; ident <= 1; ident++)
          for (int tpw = 0; tpw <= (reducing ? 2 : 0); tpw++) cands.push_back(make_shape(v, lanes, ident, tpw));
    // every candidate launched once, then the best of kTuneRounds interleaved
    // rounds of 3 launches each (one round alone picked the FP32 8-RHS hop's
    // shape differently from run to run: 26.2 or 28.6 us per launch)
    for (const int cand : cands) dslash_launch<R, LINK, MODE, RED>(c, t, cand);
    std::vector<float> cand_ms(cands.size(), 1e30f);
    for (int round = 0; round < kTuneRounds; round++)
      for (size_t i = 0; i < cands.size(); i++) {
        hipEventRecord(e0, c.stream);
        for (int k = 0; k < 3; k++) dslash_launch<R, LINK, MODE, RED>(c, t, cands[i]);
        hipEventRecord(e1, c.stream);
        hipEventSynchronize(e1);
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        cand_ms[i] = std::min(cand_ms[i], ms);
      }
    for (size_t i = 0; i < cands.size(); i++)
      if (cand_ms[i] < best) {
        best = cand_ms[i];
        code = cands[i];
      }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    tab[key] = code;
    c.ws->tuned_ms[key] = best / 3.0;
    tune_cache_store(c);
  }
  dslash_launch<R, LINK, MODE, RED>(c, a, code);
}

constexpr int kChunk = 8;  // stored-basis solver iterations between host polls

// Fold the device-clock slots of one solve (iterations 0..n-1 of the hop
// input) into the context's totals and clear them for the next solve.
__global__ void k_clk_clear(unsigned long long *t, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 2 * n) t[i] = 0ULL;
}

void kernel_timing_harvest(Context &c, int n) {
  if (!c.kt_clk || n <= 0) return;
  n = std::min(n, c.kt_clk_cap);
  const unsigned long long *t = (const unsigned long long *)d2h_stage(c, sizeof(unsigned long long) * 2 * (size_t)n);
  if (!t) return;
  hipMemcpyAsync((void *)t, c.kt_clk, sizeof(unsigned long long) * 2 * (size_t)n, hipMemcpyDeviceToHost, c.stream);
  hipLaunchKernelGGL(k_clk_clear, dim3((2 * n + 255) / 256), dim3(256), 0, c.stream, c.kt_clk, n);
  hipStreamSynchronize(c.stream);
  // Slot i holds the end of the M^dag M launch of iteration i (the launch
  // before the hop of iteration i+1, stored by that hop) and the end of that
  // hop (stored by the M^dag M launch after it).  The hop's interval, end of
  // M^dag M(i) to end of hop(i+1), is what a per-dispatch trace records for
  // the hop (dispatch, execution, end of kernel); the first hop of a chunk
  // follows the chunk's bookkeeping launches and the host poll instead, so
  // slots that are multiples of the chunk length are left out.  The M^dag M
  // launch of iteration i+1 directly follows that hop: its interval runs to
  // the end stored in slot i+1.
  for (int i = 1; i < n; i++)
    if (i % kChunk != 0 && t[2 * i] != 0 && t[2 * i + 1] > t[2 * i]) {
      c.kt_sum_ms += (double)(t[2 * i + 1] - t[2 * i]) / c.wall_khz;
      c.kt_count++;
    }
  for (int i = 0; i + 1 < n; i++)
    if (t[2 * i + 1] != 0 && t[2 * i + 2] > t[2 * i + 1]) {
      c.kt_sum2_ms += (double)(t[2 * i + 2] - t[2 * i + 1]) / c.wall_khz;
      c.kt_count2++;
    }
}

void kernel_timing_release(Context &c) {
  if (c.kt_clk) hipFree(c.kt_clk);
  if (c.kt_wg) hipFree(c.kt_wg);
  c.kt_clk = nullptr;
  c.kt_wg = nullptr;
  c.kt_clk_cap = 0;
  c.kt_wg_cap = 0;
}

// per-workgroup end stamps of the timed launches (grown on demand)
int kernel_timing_stamps(Context &c, size_t n) {
  if (c.kt_wg_cap >= n) return 0;
  if (c.kt_wg) {
    hipStreamSynchronize(c.stream);
    hipFree(c.kt_wg);
  }
  c.kt_wg = nullptr;
  c.kt_wg_cap = 0;
  if (!hip_ok(c, hipMalloc((void **)&c.kt_wg, sizeof(unsigned long long) * n), "timing stamps")) return -1;
  c.kt_wg_cap = n;
  return 0;
}

// device-clock slots of the live kernel timing, one per solver iteration
// (allocated and cleared when timing is switched on)
int kernel_timing_clock(Context &c) {
  if (c.kt_clk) return 0;
  int khz = 0;
  if (!hip_ok(c, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c.device), "wall clock rate") ||
      khz <= 0)
    return -1;
  c.wall_khz = (double)khz;
  const int cap = 16384;
  if (!hip_ok(c, hipMalloc((void **)&c.kt_clk, sizeof(unsigned long long) * 2 * cap), "timing slots")) return -1;
  c.kt_clk_cap = cap;
  hipLaunchKernelGGL(k_clk_clear, dim3((2 * cap + 255) / 256), dim3(256), 0, c.stream, c.kt_clk, cap);
  return hip_ok(c, hipStreamSynchronize(c.stream), "timing slots") ? 0 : -1;
}

std::string tuning_table(Context &c) {
  std::string out = "[";
  bool first = true;
  for (auto &kv : c.ws->tuned_nr) {
    char buf[320];
    const unsigned long long k = kv.first;
    snprintf(buf, sizeof buf,
             "%s{\"real_bytes\":%llu,\"links\":\"%s\",\"mode\":%llu,\"reduction\":%llu,\"nvec\":%llu,"
             "\"rhs_per_thread\":%d,\"lanes_per_site\":%d,\"tile_order\":\"%s\",\"tiles_per_workgroup\":%d}",
             first ? "" : ",", k >> 40,
             ((k >> 36) & 15) == LINK_S12 ? "fp32_12" : ((k >> 36) & 15) == LINK_D12 ? "fp64_12" : "fp64_18",
             (k >> 32) & 15, (k >> 28) & 15, k & 0xfffffff, shape_nr(kv.second), shape_lanes(kv.second),
             shape_identity(kv.second) ? "lattice" : "xcd", 1 << shape_tpw_log2(kv.second));
    out += buf;
    first = false;
  }
  return out + "]";
}

static RedArgs make_red(Context &c, int nvec, int epi, void *state, double *sums) {
  RedArgs r;
  const int nbx = (int)nbx_of(c);
  r.partial = (double *)ws_get(c, S_PARTIAL, sizeof(double) * (size_t)nbx * std::max(nvec, 1));
  r.counter = c.ws->counter;
  r.sums = sums;
  r.state = state;
  r.epi = epi;
  r.nbr = nbx;
  return r;
}

void launch_hop_d(Context &c, int par_out, const double2 *in, double2 *out, int nvec, long vs_in, long vs_out) {
  DsArgs a = {};
  a.par_out = par_out;
  a.U = c.U64;
  a.in = in; a.in_vs = vs_in;
  a.out = out; a.out_vs = vs_out;
  a.nvec = nvec;
  dslash_dispatch<double, LINK_D18, DS_HOP, RED_NONE>(c, a);
}

// The same launch without the autotuner: the shape already tuned for this
// (mode, nvec) when there is one, else the default shape.  Nothing is timed,
// created or allocated (for the allocation-free entry points).
void launch_hop_d_fixed(Context &c, int par_out, const double2 *in, double2 *out, int nvec, long vs_in, long vs_out) {
  DsArgs a = {};
  a.par_out = par_out;
  a.U = c.U64;
  a.in = in; a.in_vs = vs_in;
  a.out = out; a.out_vs = vs_out;
  a.nvec = nvec;
  const auto &tab = c.ws->tuned_nr;
  const auto it = tab.find(tune_key(sizeof(double), LINK_D18, DS_HOP, RED_NONE, nvec));
  const int code = it != tab.end() ? it->second : make_shape(pick_nr(nvec), 1, 0, 0);
  dslash_launch<double, LINK_D18, DS_HOP, RED_NONE>(c, a, code);
}

void launch_hop_d12(Context &c, int par_out, const double2 *in, double2 *out, int nvec, long vs_in, long vs_out) {
  DsArgs a = {};
  a.par_out = par_out;
  a.U = c.U64c;
  a.in = in; a.in_vs = vs_in;
  a.out = out; a.out_vs = vs_out;
  a.nvec = nvec;
  dslash_dispatch<double, LINK_D12, DS_HOP, RED_NONE>(c, a);
}

void launch_hop_s(Context &c, int par_out, const float2 *in, float2 *out, int nvec, long vs_in, long vs_out) {
  DsArgs a = {};
  a.par_out = par_out;
  a.U = c.U32;
  a.in = in; a.in_vs = vs_in;
  a.out = out; a.out_vs = vs_out;
  a.nvec = nvec;
  dslash_dispatch<float, LINK_S12, DS_HOP, RED_NONE>(c, a);
}

// --------------------------------------------------------------- BLAS kernels
// grid: (nbr site blocks, one y-block per vector); reducing kernels loop over
// their contiguous site range
struct BlasArgs {
  long Vh;
  int nvec, nshift;
  const void *b;      // init: FP64 sources / residuals
  void *r, *p, *s, *ps, *x;
  MsState *ms;
  CgState *cg;
  RedArgs red;
};

// r = p = b, p_s = b, x_s = 0, |b|^2
template <typename R>
__global__ __launch_bounds__(kBlock) void k_ms_init(BlasArgs a) {
  using C = Cx<R>;
  const unsigned bx = xcd_block(blockIdx.x, gridDim.x);
  const int j = blockIdx.y;
  const long Vh = a.Vh;
  long h0, h1;
  site_range(Vh, bx, gridDim.x, h0, h1);
  double v[1] = {0.0};
  for (long h = h0 + threadIdx.x; h < h1; h += kBlock) {
    const double2 *b = (const double2 *)a.b + (long)j * 3 * Vh + h;
    C *r = (C *)a.r + (long)j * 3 * Vh + h, *p = (C *)a.p + (long)j * 3 * Vh + h;
    for (int i = 0; i < 3; i++) {
      const double2 bv = b[(long)i * Vh];
      const C cv = mk<R>((R)bv.x, (R)bv.y);
      r[(long)i * Vh] = cv;
      p[(long)i * Vh] = cv;
      v[0] += (double)cv.x * (double)cv.x + (double)cv.y * (double)cv.y;
      for (int k = 0; k < a.nshift; k++) {
        const long o = ((long)(j * a.nshift + k) * 3 + i) * Vh + h;
        ((C *)a.ps)[o] = cv;
        ((C *)a.x)[o] = mk<R>(0, 0);
      }
    }
  }
  block_partials<1>(v, a.red.partial, j, a.nvec, gridDim.x, bx);
}

// step C: r += omega s; lambda = |r|^2
template <typename R>
__global__ __launch_bounds__(kBlock) void k_ms_update_r(BlasArgs a) {
  using C = Cx<R>;
  const unsigned bx = xcd_block(blockIdx.x, gridDim.x);
  const int j = blockIdx.y;
  const long Vh = a.Vh;
  long h0, h1;
  site_range(Vh, bx, gridDim.x, h0, h1);
  double v[1] = {0.0};
  const MsState &st = a.ms[j];
  if (!st.done) {
    const R om = (R)st.omega;
    for (long h = h0 + threadIdx.x; h < h1; h += kBlock) {
      C *r = (C *)a.r + (long)j * 3 * Vh + h;
      const C *s = (const C *)a.s + (long)j * 3 * Vh + h;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        C rv = r[(long)i * Vh];
        const C sv = s[(long)i * Vh];
        rv.x += om * sv.x;
        rv.y += om * sv.y;
        r[(long)i * Vh] = rv;
        v[0] += (double)rv.x * (double)rv.x + (double)rv.y * (double)rv.y;
      }
    }
  }
  block_partials<1>(v, a.red.partial, j, a.nvec, gridDim.x, bx);
}

// step D: p = r + gamma p ; for active shifts x_s -= om_s p_s, p_s = zeta_s r + ga_s p_s
template <typename R>
__global__ __launch_bounds__(kBlock) void k_ms_update_p(BlasArgs a) {
  using C = Cx<R>;
  const int bx = xcd_block(blockIdx.x, gridDim.x);
  const long h = (long)bx * kBlock + threadIdx.x;
  const int j = blockIdx.y;
  const long Vh = a.Vh;
  const MsState &st = a.ms[j];
  if (h < Vh && !st.done) {
    const R g = (R)st.gammag;
    C *p = (C *)a.p + (long)j * 3 * Vh + h;
    const C *r = (const C *)a.r + (long)j * 3 * Vh + h;
    C rv[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      rv[i] = r[(long)i * Vh];
      C pv = p[(long)i * Vh];
      pv.x = rv[i].x + g * pv.x;
      pv.y = rv[i].y + g * pv.y;
      p[(long)i * Vh] = pv;
    }
    for (int k = 0; k < a.nshift; k++) {
      if (!st.flag[k]) continue;
      const R om = (R)st.om[k], z = (R)st.z2[k], ga = (R)st.ga[k];
      C *ps = (C *)a.ps + (long)(j * a.nshift + k) * 3 * Vh + h;
      C *x = (C *)a.x + (long)(j * a.nshift + k) * 3 * Vh + h;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        C pv = ps[(long)i * Vh];
        C xv = x[(long)i * Vh];
        xv.x -= om * pv.x;
        xv.y -= om * pv.y;
        x[(long)i * Vh] = xv;
        pv.x = z * rv[i].x + ga * pv.x;
        pv.y = z * rv[i].y + ga * pv.y;
        ps[(long)i * Vh] = pv;
      }
    }
  }
}

// ---- batched single-shift CG (refinement of every shifted system)
// init: r = r0 (FP64 -> R), p = r, x = 0, delta = |r|^2
template <typename R>
__global__ __launch_bounds__(kBlock) void k_cg_init(BlasArgs a) {
  using C = Cx<R>;
  const unsigned bx = xcd_block(blockIdx.x, gridDim.x);
  const int j = blockIdx.y;
  const long Vh = a.Vh;
  long h0, h1;
  site_range(Vh, bx, gridDim.x, h0, h1);
  double v[1] = {0.0};
  for (long h = h0 + threadIdx.x; h < h1; h += kBlock) {
    const double2 *b = (const double2 *)a.b + (long)j * 3 * Vh + h;
    C *r = (C *)a.r + (long)j * 3 * Vh + h, *p = (C *)a.p + (long)j * 3 * Vh + h;
    C *x = (C *)a.x + (long)j * 3 * Vh + h;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const double2 bv = b[(long)i * Vh];
      const C cv = mk<R>((R)bv.x, (R)bv.y);
      r[(long)i * Vh] = cv;
      p[(long)i * Vh] = cv;
      x[(long)i * Vh] = mk<R>(0, 0);
      v[0] += (double)cv.x * (double)cv.x + (double)cv.y * (double)cv.y;
    }
  }
  block_partials<1>(v, a.red.partial, j, a.nvec, gridDim.x, bx);
}

// x += alpha p ; r -= alpha s ; lambda = |r|^2
template <typename R>
__global__ __launch_bounds__(kBlock) void k_cg_update_xr(BlasArgs a) {
  using C = Cx<R>;
  const unsigned bx = xcd_block(blockIdx.x, gridDim.x);
  const int j = blockIdx.y;
  const long Vh = a.Vh;
  long h0, h1;
  site_range(Vh, bx, gridDim.x, h0, h1);
  double v[1] = {0.0};
  const CgState &st = a.cg[j];
  if (!st.done) {
    const R al = (R)st.alpha;
    for (long h = h0 + threadIdx.x; h < h1; h += kBlock) {
      C *x = (C *)a.x + (long)j * 3 * Vh + h, *r = (C *)a.r + (long)j * 3 * Vh + h;
      const C *p = (const C *)a.p + (long)j * 3 * Vh + h, *s = (const C *)a.s + (long)j * 3 * Vh + h;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        C xv = x[(long)i * Vh], rv = r[(long)i * Vh];
        const C pv = p[(long)i * Vh], sv = s[(long)i * Vh];
        xv.x += al * pv.x; xv.y += al * pv.y;
        rv.x -= al * sv.x; rv.y -= al * sv.y;
        x[(long)i * Vh] = xv;
        r[(long)i * Vh] = rv;
        v[0] += (double)rv.x * (double)rv.x + (double)rv.y * (double)rv.y;
      }
    }
  }
  block_partials<1>(v, a.red.partial, j, a.nvec, gridDim.x, bx);
}

// p = r + beta p
template <typename R>
__global__ __launch_bounds__(kBlock) void k_cg_update_p(BlasArgs a) {
  using C = Cx<R>;
  const int bx = xcd_block(blockIdx.x, gridDim.x);
  const long h = (long)bx * kBlock + threadIdx.x;
  const int j = blockIdx.y;
  const long Vh = a.Vh;
  const CgState &st = a.cg[j];
  if (h < Vh && !st.done) {
    const R be = (R)st.beta;
    C *p = (C *)a.p + (long)j * 3 * Vh + h;
    const C *r = (const C *)a.r + (long)j * 3 * Vh + h;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      C pv = p[(long)i * Vh];
      const C rv = r[(long)i * Vh];
      pv.x = rv.x + be * pv.x;
      pv.y = rv.y + be * pv.y;
      p[(long)i * Vh] = pv;
    }
  }
}

// x64[j] += (double) x32[j]  (n elements of double2)
__global__ __launch_bounds__(kBlock) void k_accum_f2d(double2 *x64, const float2 *x32, long n) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    const float2 v = x32[i];
    double2 o = x64[i];
    o.x += (double)v.x;
    o.y += (double)v.y;
    x64[i] = o;
  }
}
__global__ __launch_bounds__(kBlock) void k_copy_f2d(double2 *x64, const float2 *x32, long n) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    const float2 v = x32[i];
    x64[i] = make_double2((double)v.x, (double)v.y);
  }
}

// ------------------------------------------- multi-shift CG on a stored basis
// The residuals r^(0..K) of the base CG are the (scaled) Lanczos vectors of
// MdagM; every shifted solution is a combination of them,
//   x_s = sum_{i<K_s} d_{s,i} r^(i),
// with d_{s,i} from the same Jegerlehner scalar recurrences.  Keeping the
// basis resident in HBM (one vector per iteration and right-hand side) makes
// the per-iteration traffic independent of the number of shifts (the
// explicit-shift form adds 4 vectors per active shift); the shifted
// solutions are assembled in one pass at the end.
//
// Per iteration k two launches, with the base CG in its single-reduction
// form (Chronopoulos and Gear, J. Comput. Appl. Math. 25 (1989) 153): the
// operator is applied to the residual, w = A r^(k-1), and s = A p follows
// its own recurrence, so p is never formed.  With rho_i = |r_i|^2 and
// mu_i = <r_i, A r_i>:
//   beta_{i-1} = rho_i / rho_{i-1},
//   <p_i, A p_i> = mu_i - beta_{i-1} rho_i / alpha_{i-1},  alpha_i = rho_i / <p_i, A p_i>,
//   s_i = w_i + beta_{i-1} s_{i-1},  r_{i+1} = r_i - alpha_i s_i
// (alpha, beta are those of the textbook CG; the shifted solutions come
// from the basis).  Since s_{i-1} = (r_{i-1} - r_i) / alpha_{i-1}, and both
// residuals are basis slots, s is not stored either: the update is the
// three-term recurrence
//   r_{i+1} = r_i - alpha_i w_i + c_i (r_i - r_{i-1}),  c_i = alpha_i beta_{i-1} / alpha_{i-1}.  With A = m^2 - D_eo D_oe and the staggered D
// anti-hermitian, mu = m^2 rho + |D_oe r|^2, which the hop can reduce, so
// alpha is known when M^dag M runs and the residual update joins it.
// Launches of iteration k:
// A h = D_oe r^(k-1) with |h|^2 partials; workgroup j < nrhs also sums the
//   partials of rho_{k-1} in a fixed order, forms beta, advances the
//   Jegerlehner recurrences of rhs j for iteration k-1 and the stopping
//   test (bms_step);
// B every workgroup forms alpha and c from the |h|^2 partials (bms_alpha),
//   then w = m^2 r^(k-1) - D_eo h and r^(k) = r^(k-1) - alpha w +
//   c (r^(k-1) - r^(k-2)) into basis slot k, with partials of |r^(k)|^2
//   (w is never stored).
// Vector traffic per iteration and rhs: 2 (A) + 4 (B) passes, independent
// of the number of shifts.  The stop decision lags by one iteration: one
// extra application of the operator per solve, identical recurrences and
// iteration counts.  Scalars read by every workgroup of a launch (rho_{k-2},
// alpha, stop) are double-buffered by the parity of k, so no workgroup
// reads a value another one writes in the same launch.

// set-up launches of the stored-basis solver
struct BmsArgs {
  long Vh;
  int nvec, nshift, cap;
  const double2 *b;
  void *basis;          // [cap][rhs][3][Vh]: slot k of every rhs contiguous
  double *partC;        // partials of rho_0
  int nbrC;
  BmsShared *sh;
  MsState *ms;
  int *active;          // [2]
};

struct BmsGraphKey {
  BmsArgs a;
  DsArgs hop, mdm;
  unsigned nbr;
};
static_assert(sizeof(BmsGraphKey) <= sizeof(GraphSlot::key), "graph key");

template <typename R>
__global__ __launch_bounds__(kBlock) void k_bms_init(BmsArgs a) {
  using C = Cx<R>;
  const unsigned bx = xcd_block(blockIdx.x, gridDim.x);
  const int j = blockIdx.y;
  const long Vh = a.Vh;
  long h0, h1;
  site_range(Vh, bx, gridDim.x, h0, h1);
  double v[1] = {0.0};
  C *r0 = (C *)a.basis + (long)j * 3 * Vh;
  for (long h = h0 + threadIdx.x; h < h1; h += kBlock)
    for (int i = 0; i < 3; i++) {
      const double2 bv = a.b[((long)j * 3 + i) * Vh + h];
      const C cv = mk<R>((R)bv.x, (R)bv.y);
      r0[(long)i * Vh + h] = cv;
      v[0] += (double)cv.x * (double)cv.x + (double)cv.y * (double)cv.y;
    }
  block_partials<1>(v, a.partC, j, a.nvec, gridDim.x, bx);  // rho_0
}

// scalar state before iteration 1: inverter.cc:166-198
__global__ __launch_bounds__(kBlock) void k_bms_init_fin(BmsArgs a) {
  __shared__ double sums[kMaxRed];
  final_sums(a.partC, a.nvec, a.nbrC, sums);
  for (int j = threadIdx.x; j < a.nvec; j += blockDim.x) {
    MsState &s = a.ms[j];
    s.omega = 1.0;
    s.omega_prev = 1.0;
    s.gammag = 0.0;
    s.iter = 0;
    s.delta = sums[j];
    s.bnorm2 = sums[j];
    if (s.res < 0.0) s.res = fmax(s.tol_abs, s.tol_rel * sqrt(sums[j]));
    s.nact = s.nshift;
    for (int k = 0; k < s.nshift; k++) {
      s.z1[k] = 1.0; s.z2[k] = 1.0; s.z3[k] = 1.0;
      s.om[k] = 0.0; s.ga[k] = 0.0;
      s.flag[k] = 1; s.flag_next[k] = 1; s.its[k] = 0;
    }
    a.sh[j].rho[1] = 1.0;
    a.sh[j].alpha[1] = 1.0;
    a.sh[j].rho1 = sums[j];            // rho_0 for the hop of iteration 1
    a.sh[j].stop[1] = sums[j] == 0.0;  // zero source: x = 0
    s.done = s.stop = a.sh[j].stop[1];
  }
  // the hop of iteration k counts running systems into active[(k+1)&1]
  if (threadIdx.x == 0) a.active[0] = 0;
}

// x[j][s] (FP64) = sum_{i<K} coef[j][s][i] basis[j][i] for NS shifts per
// pass; klen[j] = basis vectors of rhs j (slots beyond were never written).
// A tall-skinny product (3 Vh complex rows x K) . (K x NS real): one thread
// per (colour, site) element streams its K basis values once, while the
// workgroup stages the coefficients of kTileK basis indices in LDS, where
// every lane reads the same entry (broadcast).  Coefficients of shifts at or
// beyond nshift are staged as zero, so the inner loop has no per-shift
// branches.
constexpr int kTileK = 32;

template <typename R, int NS, typename XO>
__global__ __launch_bounds__(kBlock) void k_bms_assemble(const void *basis, const double *coef, const int *klen,
                                                         XO *x, long Vh, int cstride, int nshift, int s0, int nvec) {
  using C = Cx<R>;
  __shared__ double2 cs[kTileK][NS / 2];
  const long n3 = 3 * Vh;
  const long e = (long)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const int j = blockIdx.y;
  const int kmax = klen[j];
  const int ns = nshift - s0 < NS ? nshift - s0 : NS;
  const long ss = (long)nvec * n3;  // slot stride
  const C *bj = (const C *)basis + (long)j * n3 + (e < n3 ? e : 0);
  const double *cj = coef + ((long)j * nshift + s0) * cstride;
  double2 acc[NS];
#pragma unroll
  for (int q = 0; q < NS; q++) acc[q] = make_double2(0.0, 0.0);
  for (int i0 = 0; i0 < kmax; i0 += kTileK) {
    const int nk = kmax - i0 < kTileK ? kmax - i0 : kTileK;
    __syncthreads();
    for (int idx = threadIdx.x; idx < kTileK * NS; idx += kBlock) {
      const int u = idx / NS, q = idx - u * NS;
      ((double *)cs[u])[q] = (u < nk && q < ns) ? cj[(long)q * cstride + i0 + u] : 0.0;
    }
    __syncthreads();
    if (e >= n3) continue;
    int u = 0;
    for (; u + 8 <= nk; u += 8) {
      C v[8];
#pragma unroll
      for (int w = 0; w < 8; w++) v[w] = bj[(long)(i0 + u + w) * ss];
#pragma unroll
      for (int w = 0; w < 8; w++) {
        const double vx = (double)v[w].x, vy = (double)v[w].y;
#pragma unroll
        for (int q2 = 0; q2 < NS / 2; q2++) {
          const double2 d = cs[u + w][q2];
          acc[2 * q2].x += d.x * vx;
          acc[2 * q2].y += d.x * vy;
          acc[2 * q2 + 1].x += d.y * vx;
          acc[2 * q2 + 1].y += d.y * vy;
        }
      }
    }
    for (; u < nk; u++) {
      const C v = bj[(long)(i0 + u) * ss];
      const double vx = (double)v.x, vy = (double)v.y;
#pragma unroll
      for (int q2 = 0; q2 < NS / 2; q2++) {
        const double2 d = cs[u][q2];
        acc[2 * q2].x += d.x * vx;
        acc[2 * q2].y += d.x * vy;
        acc[2 * q2 + 1].x += d.y * vx;
        acc[2 * q2 + 1].y += d.y * vy;
      }
    }
  }
  if (e >= n3) return;
#pragma unroll
  for (int q = 0; q < NS; q++)
    if (q < ns) {
      XO v;
      v.x = acc[q].x;
      v.y = acc[q].y;
      x[((long)j * nshift + s0 + q) * n3 + e] = v;
    }
}

// FP32 basis, FP32 accumulation on the matrix cores (the FP32 solver, whose
// solutions carry FP32 accuracy anyway).  The 3Vh complex rows are 6Vh real
// rows, so x = basis^T . coef is a real (rows x K) . (K x 16 shifts) product,
// run as v_mfma_f32_16x16x4f32 on 16 rows x 4 basis slots x 16 shifts.  The
// FP64 form above spends its time in FP64 FMAs (16 flop per basis byte);
// here the product runs at the matrix-core rate and the basis stream bounds
// the kernel.
// A wave owns 32 consecutive complex elements.  Lane l (col = l%16, kq =
// l/16) loads elements 2col and 2col+1 of slot i0+kq as one float4: its four
// words are the A operands of four products (product c takes logical row m
// as element 2m + c/2, component c%2; A[m][k] sits in lane m + 16k), and the
// coefficient of shift col at slot i0+kq is the B operand (B[k][n] in lane n
// + 16k).  Accumulator c holds D[m][n] in lane n + 16(m/4), register m%4, so
// lane l ends with shift col of elements 8kq .. 8kq+7 and stores them as four
// 16-byte words.  coef: FP32 [rhs][shift group][kp slots][16], zero beyond
// each shift's length and for absent shifts; slots at or beyond klen[j] are
// never read from the basis (they may hold stale values).
constexpr int kMxUnroll = 4;  // K steps (4 slots each) per loop trip: kp is a multiple of 16

template <typename XO>
__global__ __launch_bounds__(kBlock) void k_bms_assemble_mx(const float2 *basis, const float *coef, const int *klen,
                                                            XO *x, long Vh, int kp, int nshift, int nvec) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  const long n3 = 3 * Vh;
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, kq = lane >> 4;
  const int j = blockIdx.y, g = blockIdx.z;
  const long e0 = ((long)xcd_block(blockIdx.x, gridDim.x) * (kBlock / 64) + (threadIdx.x >> 6)) * 32;
  const int kmax = klen[j];
  const long ea = e0 + 2 * col;
  const bool in = ea < n3;  // n3 is even: ea and ea+1 are both inside or both outside
  const long ss = (long)nvec * n3;
  const float2 *bp = basis + (long)j * n3 + (in ? ea : 0);
  const float *cp = coef + (long)(j * (int)gridDim.z + g) * kp * 16 + lane;
  f4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; c++) acc[c] = (f4){0.f, 0.f, 0.f, 0.f};
  for (int i0 = 0; i0 < kmax; i0 += 4 * kMxUnroll) {
    float4 v[kMxUnroll];
    float b[kMxUnroll];
#pragma unroll
    for (int u = 0; u < kMxUnroll; u++) {
      const int i = i0 + 4 * u + kq;
      v[u] = in && i < kmax ? *(const float4 *)(bp + (long)i * ss) : make_float4(0.f, 0.f, 0.f, 0.f);
      b[u] = cp[(long)(i0 + 4 * u) * 16];
    }
#pragma unroll
    for (int u = 0; u < kMxUnroll; u++) {
      acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(v[u].x, b[u], acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(v[u].y, b[u], acc[1], 0, 0, 0);
      acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(v[u].z, b[u], acc[2], 0, 0, 0);
      acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(v[u].w, b[u], acc[3], 0, 0, 0);
    }
  }
  const int q = g * 16 + col;
  if (q >= nshift) return;
  XO *xq = x + ((long)j * nshift + q) * n3;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const long e = e0 + 8 * kq + 2 * r;
    if (e < n3) {
      if constexpr (sizeof(XO) == 8) {
        // (j nshift + q) n3 + e is even: one 16-byte store
        *(float4 *)(xq + e) = make_float4(acc[0][r], acc[1][r], acc[2][r], acc[3][r]);
      } else {
        XO w0, w1;
        w0.x = acc[0][r]; w0.y = acc[1][r];
        w1.x = acc[2][r]; w1.y = acc[3][r];
        xq[e] = w0;
        xq[e + 1] = w1;
      }
    }
  }
}

// set the device iteration base (a kernel, not a memset, so that it can sit
// in a captured graph)
__global__ void k_set_int(int *p, int v) {
  if (threadIdx.x == 0) *p = v;
}

// End of a chunk of solver steps: advances the device iteration base by
// `step` and reports every run's running-system count (for the new
// iteration's parity) into pinned host memory, slot (chunk number & 1) of
// [2][n], with system-scope stores; the host reads them after an event
// behind this launch, so polling needs no copy per run.
constexpr int kMaxPoll = 16;
struct ChunkPoll {
  const int *active[kMaxPoll];  // per run: running-system counters, [iteration & 1]
  int *host;                    // pinned, [2][n]
  int n;
};
__global__ void k_chunk_end(int *iterc, int step, ChunkPoll p) {
  const int k = *iterc + step;
  __syncthreads();
  const int slot = ((k - 1) / step - 1) & 1;
  const int i = threadIdx.x;
  if (i < p.n) __hip_atomic_store(p.host + slot * p.n + i, p.active[i][k & 1], __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_SYSTEM);
  if (i == 0) *iterc = k;
  __threadfence_system();
}

// ------------------------------------------------------------- host drivers
// systems still running, written by every solver finalize launch
static int *active_ptr(Context &c) { return (int *)(c.ws->counter + 1); }
// device iteration base of the stored-basis solver
static int *iter_ptr(Context &c) { return (int *)(c.ws->counter + 8); }

static void finalize(Context &c, const RedArgs &ra, int nvec, bool count_active) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, c.stream, ra, nvec,
                     count_active ? active_ptr(c) : (int *)nullptr);
}

// read the active count after a chunk of iterations (one 4-byte copy)
static bool none_active(Context &c) {
  int *hf = c.ws->h_flags;
  hipMemcpyAsync(hf, active_ptr(c), sizeof(int), hipMemcpyDeviceToHost, c.stream);
  hipStreamSynchronize(c.stream);
  return hf[0] == 0;
}

// multi-shift CG in precision R with link format LINK on sources b (FP64).
// x (precision R) receives [nrhs][nshift] solutions.  res_abs[j] per rhs.
template <typename R, int LINK>
static int ms_cg(Context &c, const double2 *b, int nrhs, int nshift, const double *shifts, const double *res_abs,
                 int max_iter, void *x, int *iters_out) {
  using C = Cx<R>;
  const long Vh = c.Vh;
  const size_t vb = sizeof(C) * 3 * Vh;
  const bool dbl = sizeof(R) == 8;
  void *r = ws_get(c, dbl ? S_R64 : S_R32, vb * nrhs);
  void *p = ws_get(c, dbl ? S_P64 : S_P32, vb * nrhs);
  void *s = ws_get(c, dbl ? S_S64 : S_S32, vb * nrhs);
  void *hh = ws_get(c, dbl ? S_H64 : S_H32, vb * nrhs);
  void *ps = ws_get(c, dbl ? S_PS64 : S_PS32, vb * nrhs * nshift);
  MsState *ms = (MsState *)ws_get(c, S_MS, sizeof(MsState) * nrhs);
  double *diag = (double *)ws_get(c, S_DIAG, sizeof(double) * 256);
  if (!r || !p || !s || !hh || !ps || !ms || !diag) return -1;

  std::vector<MsState> hs(nrhs);
  for (int j = 0; j < nrhs; j++) {
    MsState &st = hs[j];
    memset(&st, 0, sizeof(st));
    st.nshift = nshift;
    st.res = res_abs[j];
    st.max_iter = max_iter;
    for (int k = 0; k < nshift; k++) st.shift[k] = shifts[k];
  }
  std::vector<double> hd(nrhs, c.mass * c.mass);
  if (!h2d(c, ms, hs.data(), sizeof(MsState) * nrhs) || !h2d(c, diag, hd.data(), sizeof(double) * nrhs)) return -1;

  const unsigned nbx = nbx_of(c);
  BlasArgs ba = {};
  ba.Vh = Vh; ba.nvec = nrhs; ba.nshift = nshift;
  ba.b = b; ba.r = r; ba.p = p; ba.s = s; ba.ps = ps; ba.x = x; ba.ms = ms;
  ba.red = make_red(c, nrhs, EPI_MS_INIT, ms, nullptr);
  const unsigned nbr = reduce_blocks(c, nrhs);
  ba.red.nbr = (int)nbr;
  hipLaunchKernelGGL(k_ms_init<R>, dim3(nbr, nrhs), dim3(kBlock), 0, c.stream, ba);
  finalize(c, ba.red, nrhs, true);
  RedArgs red_c = ba.red;
  red_c.epi = EPI_MS_C;

  DsArgs hop = {};
  hop.par_out = 1; hop.U = link_ptr<LINK>(c);
  hop.in = p; hop.in_vs = 3 * Vh; hop.out = hh; hop.out_vs = 3 * Vh; hop.nvec = nrhs;
  hop.skip = &ms[0].stop; hop.skip_stride = sizeof(MsState) / sizeof(int);
  DsArgs mdm = hop;
  mdm.par_out = 0; mdm.in = hh; mdm.out = s; mdm.pin = p; mdm.pin_vs = 3 * Vh; mdm.diag = diag;
  mdm.red = make_red(c, nrhs, EPI_MS_B, ms, nullptr);

  const int chunk = 8;
  int it = 0;
  while (!none_active(c)) {
    for (int k = 0; k < chunk; k++) {
      dslash_dispatch<R, LINK, DS_HOP, RED_NONE>(c, hop);
      dslash_dispatch<R, LINK, DS_MDAGM, RED_DOT>(c, mdm);
      finalize(c, mdm.red, nrhs, false);
      hipLaunchKernelGGL(k_ms_update_r<R>, dim3(nbr, nrhs), dim3(kBlock), 0, c.stream, ba);
      finalize(c, red_c, nrhs, true);
      hipLaunchKernelGGL(k_ms_update_p<R>, dim3(nbx, nrhs), dim3(kBlock), 0, c.stream, ba);
    }
    it += chunk;
    if (it > max_iter + chunk) break;
  }
  if (!d2h_sync(c, hs.data(), ms, sizeof(MsState) * nrhs)) return -1;
  int maxit = 0;
  for (int j = 0; j < nrhs; j++) maxit = std::max(maxit, hs[j].iter);
  *iters_out = maxit;
  for (int j = 0; j < nrhs; j++)
    if (hs[j].nonfinite) {
      c.last_error = "multishift: non-finite value in the solver reductions";
      return kStatusNonFinite;
    }
  for (int j = 0; j < nrhs; j++)
    if (hs[j].iter >= max_iter && hs[j].nact > 0) {
      c.last_error = "multishift: max_cg reached";
      return kStatusMaxCg;
    }
  return 0;
}

// Basis capacity policy.  The slots a solve needs are not known in advance,
// so the capacity starts at max(kMinSlots, twice the most iterations a solve
// of this precision has needed, the slots already allocated) and is bounded
// by half of the device memory not held elsewhere and by kMaxBasisBytes (a
// 32^3x8 FP64 slot of 8 right-hand sides is 50 MB).  A solve that fills it
// is continued from its last residual (ms_basis_solve) instead of being
// restarted from zero, so a bounded basis loses no iterations; growing a
// basis of ~100 GB stalled the host for seconds at 48^3x12.
constexpr long kMinSlots = 128;
constexpr long kMaxSlots = 16384;
constexpr size_t kMaxBasisBytes = 32ull << 30;
constexpr int kBasisFull = 100;     // not even kMinBasis slots fit: the explicit-shift solver runs
constexpr int kBasisRestart = 101;  // the basis filled: x holds the solutions so far (see BasisRestart)
constexpr long kMinBasis = 16;

// Systems still running when the basis filled.  Their shifted residuals
// are collinear with the base residual, r_s^(K) = zeta_s^(K) r^(K)
// (inverter.cc:224-284), so the remaining equations (MdagM + s) e_s = r_s
// are one multi-shift solve with source r^(K) (basis slot K), whose
// solutions are added with the factors zeta_s.
struct BasisRestart {
  const void *basis = nullptr;  // [cap][nrhs][3][Vh] of the call that filled
  int nrhs = 0;
  std::vector<int> rhs;         // systems to continue (indices into that call's right-hand sides)
  std::vector<int> slot;        // their K
  std::vector<double> zeta;     // [rhs.size()][nshift]: zeta_s^(K) of the running shifts, 0 for frozen ones
  std::vector<double> res;      // per right-hand side of that call: the stopping threshold on |r|
  std::vector<double> bnorm2;   // and |b|^2 as the solver formed it
};

// The first pass of a stored-basis solve when it was already run (a
// multi-lattice solve): its status, continuation state and iterations.
struct FirstPass {
  int rc = 0, iters = 0;
  BasisRestart rs;
};

// multi-shift CG with the stored residual basis; x (FP64) receives
// [nrhs][nshift] solutions.  Returns kBasisFull when not even kMinBasis
// slots fit (the caller then runs the explicit-shift algorithm) and
// kBasisRestart, with rs filled, when systems were still running when the
// basis filled.
// One stored-basis solve on one lattice, split into phases so that several
// lattices can run their iterations in lockstep (bms_group_iterate): set-up
// (buffers, the right-hand sides into basis slot 0, the scalar state), the
// iterations (chunks of kChunk, replayed as HIP graphs) and the assembly.
struct BasisRun {
  int nrhs = 0, nshift = 0, max_iter = 0;
  long cap = 0;
  void *basis = nullptr, *hh = nullptr;
  MsState *ms = nullptr;
  BmsShared *bsh = nullptr;
  BmsHist *hist = nullptr;
  double *diag = nullptr, *partC = nullptr, *partH = nullptr;
  std::vector<MsState> hs;
  BmsArgs a = {};
  DsArgs hop = {}, mdm = {};
  unsigned nbr = 0;
  bool timed = false;
};

// buffers, scalar state and the first launches; the iteration counter
// iterc (device) is read by every launch of the iterations
template <typename R, int LINK>
static int bms_setup(Context &c, BasisRun &br, const double2 *b, int nrhs, int nshift, const double *shifts,
                     const double *res_abs, int max_iter, int *iterc, double tol_abs = 0.0, double tol_rel = 0.0) {
  using C = Cx<R>;
  const long Vh = c.Vh;
  const size_t vb = sizeof(C) * 3 * Vh;
  const bool dbl = sizeof(R) == 8;
  const size_t slot_bytes = vb * nrhs;
  const size_t have = c.ws->cap[S_BASIS];
  long want = std::max<long>(kMinSlots, 2L * c.ws->basis_kmax[dbl ? 1 : 0]);
  long budget = (long)(have / slot_bytes);
  if (budget < want) {
    // growing: up to half of the free device memory (with the current basis),
    // at most kMaxBasisBytes
    size_t fr = 0, tot = 0;
    hipMemGetInfo(&fr, &tot);
    budget = (long)(std::min((fr + have) / 2, std::max(have, kMaxBasisBytes)) / slot_bytes);
  }
  want = std::max<long>(want, (long)(have / slot_bytes));
  long cap = std::min<long>(std::min<long>((long)max_iter + 2, kMaxSlots), std::min<long>(budget, want));
  if (const char *e = getenv("RHMC_BASIS_CAP")) cap = std::min<long>(cap, std::max(atol(e), kMinBasis));
  if (cap < kMinBasis) return kBasisFull;
  br.nrhs = nrhs;
  br.nshift = nshift;
  br.max_iter = max_iter;
  br.cap = cap;
  br.basis = ws_get(c, S_BASIS, slot_bytes * cap);
  br.hh = ws_get(c, dbl ? S_H64 : S_H32, vb * nrhs);
  br.ms = (MsState *)ws_get(c, S_MS, sizeof(MsState) * nrhs);
  br.bsh = (BmsShared *)ws_get(c, S_BSH, sizeof(BmsShared) * nrhs);
  br.hist = (BmsHist *)ws_get(c, S_HIST, sizeof(BmsHist) * nrhs * cap * nshift);
  br.diag = (double *)ws_get(c, S_DIAG, sizeof(double) * 256);
  // partials of |r|^2 (M^dag M: at most nbx per rhs) and of |h|^2 (hop: one
  // per tile, at most 4 nbx per rhs with four lanes per site)
  br.partC = (double *)ws_get(c, S_PARTIAL2, sizeof(double) * nbx_of(c) * std::max(nrhs, 1));
  br.partH = (double *)ws_get(c, S_HPART, sizeof(double) * 4 * nbx_of(c) * std::max(nrhs, 1));
  if (!br.basis || !br.hh || !br.ms || !br.bsh || !br.hist || !br.diag || !br.partC || !br.partH)
    return -1;

  br.hs.assign(nrhs, MsState());
  for (int j = 0; j < nrhs; j++) {
    MsState &st = br.hs[j];
    memset(&st, 0, sizeof(st));
    st.nshift = nshift;
    // without res_abs: max(tol_abs, tol_rel |b|), set on the device (k_bms_init_fin)
    st.res = res_abs ? res_abs[j] : -1.0;
    st.tol_abs = tol_abs;
    st.tol_rel = tol_rel;
    st.max_iter = max_iter;
    for (int k = 0; k < nshift; k++) st.shift[k] = shifts[k];
  }
  std::vector<double> hd(nrhs, c.mass * c.mass);
  if (!h2d(c, br.ms, br.hs.data(), sizeof(MsState) * nrhs) || !h2d(c, br.diag, hd.data(), sizeof(double) * nrhs))
    return -1;

  BmsArgs &a = br.a;
  a = BmsArgs{};
  a.Vh = Vh; a.nvec = nrhs; a.nshift = nshift; a.cap = (int)cap;
  a.b = b; a.basis = br.basis;
  a.partC = br.partC;
  a.sh = br.bsh; a.ms = br.ms;
  a.active = active_ptr(c);
  br.nbr = reduce_blocks(c, nrhs);
  a.nbrC = (int)br.nbr;
  hipLaunchKernelGGL(k_bms_init<R>, dim3(br.nbr, nrhs), dim3(kBlock), 0, c.stream, a);
  hipLaunchKernelGGL(k_bms_init_fin, dim3(1), dim3(kBlock), 0, c.stream, a);

  // iteration k applies the operator to basis slot k-1 (chosen on the device)
  DsArgs &hop = br.hop;
  hop = DsArgs{};
  hop.par_out = 1; hop.U = link_ptr<LINK>(c);
  hop.in = br.basis; hop.in_vs = 3 * Vh; hop.out = br.hh; hop.out_vs = 3 * Vh; hop.nvec = nrhs;
  hop.skip_stride = sizeof(BmsShared) / sizeof(int);
  hop.iterp = iterc; hop.iter_stride = (long)nrhs * 3 * Vh; hop.iter_in = 1;
  hop.bms = br.bsh; hop.bms_part = br.partC; hop.bms_nbr = 0;  // set from M^dag M's grid (unused at k = 1)
  hop.bms_ms = br.ms; hop.bms_hist = br.hist; hop.bms_active = active_ptr(c);
  hop.bms_cap = (int)cap; hop.bms_nshift = nshift;
  hop.red = make_red(c, nrhs, EPI_STORE, nullptr, nullptr);
  hop.red.partial = br.partH;
  // live timing of the MD solver's hop (FP32, all pseudofermions)
  br.timed = c.time_kernels && c.kt_clk && sizeof(R) == 4 && nrhs == c.n_pf && iterc == iter_ptr(c);
  DsArgs &mdm = br.mdm;
  mdm = hop;
  mdm.par_out = 0; mdm.in = br.hh; mdm.in_vs = 3 * Vh; mdm.out = nullptr; mdm.pin = br.basis; mdm.pin_vs = 3 * Vh;
  mdm.diag = br.diag; mdm.iter_in = 0; mdm.iter_pin = 1;
  mdm.bms_ms = nullptr; mdm.bms_part = nullptr; mdm.bms_hpart = br.partH;
  mdm.red.partial = br.partC;
  if (br.timed) {
    // end stamps of M^dag M (at most nbx x nrhs workgroups) and of the hop
    // (at most 4 nbx x nrhs): each launch reads the other's
    const size_t nm = (size_t)nbx_of(c) * nrhs;
    if (kernel_timing_stamps(c, nm + 4 * nm)) return -1;
    hop.clk = mdm.clk = c.kt_clk;
    hop.clk_cap = mdm.clk_cap = c.kt_clk_cap;
    mdm.clk_wg = c.kt_wg;
    hop.clk_wg = c.kt_wg + nm;
    hop.clk_prev = mdm.clk_wg; hop.clk_field = 0; hop.clk_nprev = 0;
    mdm.clk_prev = hop.clk_wg; mdm.clk_field = 1;
  }
  return 0;
}

// A chunk is a fixed launch sequence: the first chunk of a solve runs
// eagerly (the Dslash autotuner may time candidates there), later chunks
// replay a captured graph of it, re-captured when a launch argument (key)
// changes.  An executable graph is destroyed only with nothing left in
// flight on the stream, so no earlier launch of it can still be pending.
template <typename Launch>
static bool chunk_run(Context &c, GraphSlot &gs, bool eager, const void *key, size_t key_bytes, Launch &&launch) {
  if (eager || !c.use_graphs) {
    launch();
    return true;
  }
  if (key_bytes > sizeof(gs.key)) {
    c.last_error = "solver graph key too large";
    return false;
  }
  if (!gs.exec || gs.key_bytes != key_bytes || memcmp(key, gs.key, key_bytes) != 0) {
    if (gs.exec) {
      if (!hip_ok(c, hipStreamSynchronize(c.stream), "graph retire")) return false;
      hipGraphExecDestroy(gs.exec);
    }
    gs.exec = nullptr;
    c.graph_captures++;
    hipGraph_t g = nullptr;
    if (!hip_ok(c, hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal), "capture")) return false;
    launch();
    if (!hip_ok(c, hipStreamEndCapture(c.stream, &g), "capture end")) return false;
    const bool ok = hip_ok(c, hipGraphInstantiate(&gs.exec, g, nullptr, nullptr, 0), "instantiate");
    hipGraphDestroy(g);
    if (!ok) {
      gs.exec = nullptr;
      return false;
    }
    memcpy(gs.key, key, key_bytes);
    gs.key_bytes = key_bytes;
  }
  return hip_ok(c, hipGraphLaunch(gs.exec, c.stream), "graph launch");
}

// Chunks until every system of every run has stopped or the iteration
// count passes the largest capacity.  The host reads each chunk's
// running-system counts (written by the chunk's last launch into pinned
// memory, before the next chunk resets them) while the next chunk already
// runs, so the poll's
// round trip overlaps a chunk instead of idling the GPU.  A chunk launched
// after a system has stopped does no work for it (stop flags are checked by
// every launch; nothing is written at or past a basis capacity).
// The chunk ends with k_chunk_end (chunk_poll below), which writes the
// counts into the pinned slot (chunk number & 1).
static ChunkPoll chunk_poll(Context &c, const std::vector<const int *> &active) {
  ChunkPoll p = {};
  p.n = (int)active.size();
  for (int l = 0; l < p.n && l < kMaxPoll; l++) p.active[l] = active[l];
  p.host = c.ws->h_flags;
  return p;
}

template <typename Chunk>
static int bms_iterate(Context &c, int *iterc, const std::vector<const int *> &active, long cap_max, Chunk &&chunk) {
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, c.stream, iterc, 1);
  const int n = (int)active.size();
  if (n > kMaxPoll) return -1;
  int *hf = c.ws->h_flags;
  hipEvent_t *pe = c.ws->poll_ev;
  int k = 1;
  auto launch_and_poll = [&](int slot) -> bool {
    if (!chunk(k == 1)) return false;
    k += kChunk;
    return hip_ok(c, hipEventRecord(pe[slot], c.stream), "poll event");
  };
  if (!launch_and_poll(0)) return -1;
  for (int it = 1;; it++) {
    const bool more = k < cap_max;
    if (more && !launch_and_poll(it & 1)) return -1;
    if (!hip_ok(c, hipEventSynchronize(pe[(it - 1) & 1]), "poll")) return -1;
    int running = 0;
    for (int l = 0; l < n; l++) running += hf[((it - 1) & 1) * n + l];
    if (running == 0 || !more) break;
  }
  return 0;
}

// the chunk of one lattice: kChunk iterations of the two launches
template <typename R, int LINK>
static void bms_chunk(Context &c, BasisRun &br, int *iterc, const ChunkPoll &poll) {
  DsArgs &hop = br.hop, &mdm = br.mdm;
  for (int q = 0; q < kChunk; q++) {
    hop.skip = &br.bsh[0].stop[(1 + q) & 1];
    mdm.skip = hop.skip;
    hop.iter_off = q - 1;
    mdm.iter_off = q - 1;
    dslash_dispatch<R, LINK, DS_HOP, RED_NORM>(c, hop);
    mdm.bms_hnbr = hop.red.nbr;
    mdm.clk_nprev = (int)(hop.nsb * hop.nch);
    dslash_dispatch<R, LINK, DS_BMS, RED_NORM>(c, mdm);
    hop.bms_nbr = mdm.red.nbr;
    hop.clk_nprev = (int)(mdm.nsb * mdm.nch);
  }
  hipLaunchKernelGGL(k_chunk_end, dim3(1), dim3(64), 0, c.stream, iterc, kChunk, poll);
}

// Coefficients of the assembly, on the device:  d_{s,i} = -z_s^(i) S_i with
// S_{K-1} = om^(K), S_i = om^(i+1) + ga^(i+1) S_{i+1}, K the iteration at
// which shift s froze.  S_i = f_i(f_{i+1}(... f_{K-1}(0))) for the affine maps
// f_i(S) = om^(i+1) + ga^(i+1) S.  One wave per (rhs, shift slot): lane l
// composes the maps of its contiguous segment of 0..K-1, a suffix scan over
// the wave's lanes composes the segments after it (the value entering the
// segment), and the lane runs its segment's recurrence from there.  Rows of
// slots K..klen[rhs]-1 (rounded up to 16 for the matrix-core layout) and
// absent shifts are written as zero; klen[rhs] = the largest K of its
// shifts.  Layouts: FP64 [rhs][shift][cst], or with c32 the matrix-core
// assembly's FP32 [rhs][shift group][kp][16].
__global__ __launch_bounds__(64) void k_bms_coef(const MsState *ms, const BmsHist *hist, long cap, int nshift,
                                                 double *c64, long cst, float *c32, long kp, int *klen) {
  const int sft = blockIdx.x, j = blockIdx.y, lane = threadIdx.x;
  const MsState &st = ms[j];
  int kmaxj = 0;
  for (int q = 0; q < nshift; q++) kmaxj = max(kmaxj, st.flag[q] ? st.iter : st.its[q]);
  const int K = sft < nshift ? (st.flag[sft] ? st.iter : st.its[sft]) : 0;
  if (sft == 0 && lane == 0) klen[j] = kmaxj;
  const BmsHist *hj = hist + (long)j * cap * nshift + sft;  // entry i at hj[i * nshift]
  const int seg = (K + 63) / 64;
  const int i0 = min(lane * seg, K), i1 = min(i0 + seg, K);
  // this segment's map f_{i0} o ... o f_{i1-1}: S -> A + B S
  double A = 0.0, B = 1.0;
  for (int i = i1 - 1; i >= i0; i--) {
    const BmsHist n1 = hj[(long)(i + 1) * nshift];
    A = n1.om + n1.ga * A;
    B = n1.ga * B;
  }
  // inclusive suffix composition P_l = M_l o M_{l+1} o ... o M_63
  for (int off = 1; off < 64; off *= 2) {
    const double a2 = __shfl_down(A, off), b2 = __shfl_down(B, off);
    if (lane + off < 64) {
      A = A + B * a2;
      B = B * b2;
    }
  }
  // S entering the segment (S_{i1}) = P_{l+1}(0); S_K = 0
  double S = __shfl_down(A, 1);
  if (lane == 63) S = 0.0;
  for (int i = i1 - 1; i >= i0; i--) {
    const BmsHist n1 = hj[(long)(i + 1) * nshift];
    S = n1.om + n1.ga * S;
    const double d = -(i == 0 ? 1.0 : hj[(long)i * nshift].z) * S;
    if (c32) c32[((long)(j * (int)gridDim.x / 16 + sft / 16) * kp + i) * 16 + sft % 16] = (float)d;
    else c64[((long)j * nshift + sft) * cst + i] = d;
  }
  const long k16 = (long)(kmaxj + 15) / 16 * 16;
  const int kz = c32 ? (int)(k16 < kp ? k16 : kp) : kmaxj;
  for (int i = K + lane; i < kz; i += 64) {
    if (c32) c32[((long)(j * (int)gridDim.x / 16 + sft / 16) * kp + i) * 16 + sft % 16] = 0.f;
    else c64[((long)j * nshift + sft) * cst + i] = 0.0;
  }
}

// Assembly of the shifted solutions (coefficients on the device, queued
// behind the last chunk), then the host state after the iterations:
// statuses and continuation bookkeeping.
template <typename R, int LINK>
static int bms_finish(Context &c, BasisRun &br, double2 *x, int *iters_out, float2 *x32, bool asm_fp32,
                      BasisRestart *rs) {
  const long Vh = c.Vh;
  const bool dbl = sizeof(R) == 8;
  const int nrhs = br.nrhs, nshift = br.nshift;
  const long cap = br.cap;
  // every coefficient index i < K <= iterations run < cap
  const long cst = cap + 1;
  if (asm_fp32 && sizeof(R) == 4) {
    // matrix-core assembly: FP32 coefficients [rhs][shift group][kp][16]
    const int ng = (nshift + 15) / 16;
    const long kp = (cst + 15) / 16 * 16;
    const size_t ncf = (size_t)nrhs * ng * kp * 16;
    float *dcf = (float *)ws_get(c, S_COEF, sizeof(float) * ncf + sizeof(int) * 256);
    if (!dcf) return -1;
    int *dkl = (int *)(dcf + ncf);
    hipLaunchKernelGGL(k_bms_coef, dim3(16 * ng, nrhs), dim3(64), 0, c.stream, br.ms, br.hist, cap, nshift,
                       (double *)nullptr, cst, dcf, kp, dkl);
    const dim3 grid((unsigned)((3 * Vh + 32 * (kBlock / 64) - 1) / (32 * (kBlock / 64))), nrhs, ng);
    if (x32)
      hipLaunchKernelGGL(k_bms_assemble_mx<float2>, grid, dim3(kBlock), 0, c.stream, (const float2 *)br.basis, dcf,
                         dkl, x32, Vh, (int)kp, nshift, nrhs);
    else
      hipLaunchKernelGGL(k_bms_assemble_mx<double2>, grid, dim3(kBlock), 0, c.stream, (const float2 *)br.basis, dcf,
                         dkl, x, Vh, (int)kp, nshift, nrhs);
  } else {
    const size_t ncf = (size_t)nrhs * nshift * cst;
    double *dcoef = (double *)ws_get(c, S_COEF, sizeof(double) * ncf + sizeof(int) * 256);
    if (!dcoef) return -1;
    int *dklen = (int *)(dcoef + ncf);
    hipLaunchKernelGGL(k_bms_coef, dim3(nshift, nrhs), dim3(64), 0, c.stream, br.ms, br.hist, cap, nshift, dcoef,
                       cst, (float *)nullptr, 0L, dklen);
    const dim3 grid((unsigned)((3 * Vh + kBlock - 1) / kBlock), nrhs);
    for (int s0 = 0; s0 < nshift; s0 += 16)
      if (x32)
        hipLaunchKernelGGL((k_bms_assemble<R, 16, float2>), grid, dim3(kBlock), 0, c.stream, br.basis, dcoef, dklen,
                           x32, Vh, (int)cst, nshift, s0, nrhs);
      else
        hipLaunchKernelGGL((k_bms_assemble<R, 16, double2>), grid, dim3(kBlock), 0, c.stream, br.basis, dcoef, dklen,
                           x, Vh, (int)cst, nshift, s0, nrhs);
  }
  if (!hip_ok(c, hipGetLastError(), "basis assemble")) return -1;
  // the systems' final state (after the assembly: one wait for both)
  std::vector<MsState> &hs = br.hs;
  if (!d2h_sync(c, hs.data(), br.ms, sizeof(MsState) * nrhs)) return -1;
  if (rs) {
    rs->res.resize(nrhs);
    rs->bnorm2.resize(nrhs);
    for (int j = 0; j < nrhs; j++) {
      rs->res[j] = hs[j].res;
      rs->bnorm2[j] = hs[j].bnorm2;
    }
  }
  int kmax = 0;
  for (int j = 0; j < nrhs; j++) kmax = std::max(kmax, hs[j].iter);
  long &kseen = c.ws->basis_kmax[dbl ? 1 : 0];
  kseen = std::max<long>(kseen, kmax);
  // the hop ran with work at iterations 1..kmax+1 (input slots 0..kmax);
  // slot kmax+1 holds the end of M^dag M(kmax+1) when a hop followed it
  if (br.timed) kernel_timing_harvest(c, kmax + 1);
  for (int j = 0; j < nrhs; j++)
    if (hs[j].nonfinite) {
      c.last_error = "multishift: non-finite value in the solver reductions";
      return kStatusNonFinite;
    }
  // basis filled before convergence: the solutions so far are assembled
  // above and the caller continues from r^(K)
  if (rs) {
    rs->rhs.clear();
    rs->slot.clear();
    rs->zeta.clear();
    rs->basis = br.basis;
    rs->nrhs = nrhs;
  }
  bool restart = false;
  for (int j = 0; j < nrhs; j++)
    if (hs[j].nact > 0 && hs[j].iter + 2 >= cap && hs[j].iter < br.max_iter) {
      restart = true;
      if (rs) {
        rs->rhs.push_back(j);
        rs->slot.push_back(hs[j].iter);
        for (int sft = 0; sft < nshift; sft++) rs->zeta.push_back(hs[j].flag[sft] ? hs[j].z2[sft] : 0.0);
      }
    }
  if (restart && !rs) {
    c.last_error = "multishift: stored basis filled";
    return kStatusMaxCg;
  }
  *iters_out = kmax;
  if (restart) return kBasisRestart;
  for (int j = 0; j < nrhs; j++)
    if (hs[j].iter >= br.max_iter && hs[j].nact > 0) {
      c.last_error = "multishift: max_cg reached";
      return kStatusMaxCg;
    }
  return 0;
}

// multi-shift CG with the stored residual basis on one lattice; x (FP64)
// receives [nrhs][nshift] solutions (or x32 in FP32).  Returns kBasisFull
// when not even kMinBasis slots fit (the caller then runs the
// explicit-shift algorithm) and kBasisRestart, with rs filled, when systems
// were still running when the basis filled.
template <typename R, int LINK>
static int ms_cg_basis(Context &c, const double2 *b, int nrhs, int nshift, const double *shifts,
                       const double *res_abs, int max_iter, double2 *x, int *iters_out, float2 *x32,
                       bool asm_fp32, BasisRestart *rs, double tol_abs = 0.0, double tol_rel = 0.0) {
  BasisRun br;
  int *iterc = iter_ptr(c);
  const int rc = bms_setup<R, LINK>(c, br, b, nrhs, nshift, shifts, res_abs, max_iter, iterc, tol_abs, tol_rel);
  if (rc) return rc;
  GraphSlot &gs = c.ws->bms_graph[sizeof(R) == 8 ? 1 : 0];
  BmsGraphKey key;
  const std::vector<const int *> active = {br.a.active};
  const ChunkPoll poll = chunk_poll(c, active);
  auto chunk = [&](bool eager) -> bool {
    memset(&key, 0, sizeof(key));
    key.a = br.a; key.hop = br.hop; key.mdm = br.mdm; key.nbr = br.nbr;
    key.a.b = nullptr;  // the right-hand sides are read only by k_bms_init, outside the chunk
    return chunk_run(c, gs, eager, &key, sizeof(key), [&]() { bms_chunk<R, LINK>(c, br, iterc, poll); });
  };
  if (bms_iterate(c, iterc, active, br.cap, chunk)) return -1;
  return bms_finish<R, LINK>(c, br, x, iters_out, x32, asm_fp32, rs);
}

// ---- several lattices in lockstep
// One solve per lattice (each its own right-hand sides, shifts and
// targets), all advanced together: each solver step is one hop launch and
// one M^dag M launch covering every lattice's tiles (k_dslash_group, a
// per-lattice argument table in device memory), chunks of kChunk steps are
// captured as one HIP graph, and the host polls every lattice's running
// count per chunk.  Lattices whose systems have stopped cost only their
// workgroups' early exit.  Every context must be on one stream (the
// caller's group stream).  The launch shape is timed once per group
// composition (group_shape).
struct GroupJob {
  const double2 *b;
  int nrhs, nshift;
  const double *shifts;
  const double *res_abs;
  int max_iter;
  double2 *x;
  float2 *x32;
  bool asm_fp32;
  FirstPass out;  // status, continuation state, iterations of this lattice
};

static unsigned long long fnv1a(const void *p, size_t n, unsigned long long h = 1469598103934665603ULL) {
  const unsigned char *b = (const unsigned char *)p;
  for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ULL;
  return h;
}

template <typename R, int LINK, int MODE, int RED>
static void dslash_group_launch(hipStream_t st, const DsArgs *dtab, const DsArgs *htab, int n, int code) {
  GroupArgs g;
  g.tab = dtab;
  g.n = n;
  unsigned s = 0;
  for (int l = 0; l < n; l++) {
    g.start[l] = s;
    s += (htab[l].nsb * htab[l].nch + 7) / 8 * 8;
  }
  g.start[n] = s;
  const dim3 grid(s);
  const int nr = shape_nr(code);
#define DSG(KERNEL) hipLaunchKernelGGL(KERNEL, grid, dim3(kBlock), 0, st, g)
  if (shape_lanes(code) == 4) {
    switch (nr) {
      case 8: DSG((k_dslash_ls_group<R, LINK, 8, MODE, RED>)); break;
      case 4: DSG((k_dslash_ls_group<R, LINK, 4, MODE, RED>)); break;
      case 2: DSG((k_dslash_ls_group<R, LINK, 2, MODE, RED>)); break;
      default: DSG((k_dslash_ls_group<R, LINK, 1, MODE, RED>)); break;
    }
  } else {
    switch (nr) {
      case 8: DSG((k_dslash_group<R, LINK, 8, MODE, RED>)); break;
      case 4: DSG((k_dslash_group<R, LINK, 4, MODE, RED>)); break;
      case 2: DSG((k_dslash_group<R, LINK, 2, MODE, RED>)); break;
      default: DSG((k_dslash_group<R, LINK, 1, MODE, RED>)); break;
    }
  }
#undef DSG
}

// The lockstep launch shape of a group (right-hand sides per lane, lanes per
// site, tile order), timed on its first solve: every candidate runs the first
// step's hop (without its scalar step) and M^dag M over all lattices, once to
// warm up and three times under HIP events; the r^(1), alpha and partials the
// candidates write are rewritten by the real first step.  Kept per group
// composition (precision, every lattice's extents and systems) in the lead
// context's workspace.
template <typename R, int LINK>
static int group_shape(std::vector<Context *> &cs, std::vector<BasisRun> &runs, int nmin) {
  const int n = (int)cs.size();
  Context &c0 = *cs[0];
  unsigned long long key = fnv1a(&n, sizeof n);
  const int rb = (int)sizeof(R), lk = LINK;
  key = fnv1a(&rb, sizeof rb, key);
  key = fnv1a(&lk, sizeof lk, key);
  for (int l = 0; l < n; l++) {
    key = fnv1a(&cs[l]->L, sizeof(LatticeDims), key);
    key = fnv1a(&runs[l].nrhs, sizeof(int), key);
  }
  auto &gt = c0.ws->group_tuned;
  const auto it = gt.find(key);
  if (it != gt.end()) return it->second;
  std::vector<DsArgs> tab(2 * (size_t)n);
  DsArgs *dt = (DsArgs *)ws_get(c0, S_GROUP_TAB, sizeof(DsArgs) * tab.size());
  if (!dt) return -1;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  float best = 1e30f;
  int code = make_shape(pick_nr(nmin), 1, 0, 0);
  for (int lanes = 1; lanes <= 4; lanes *= 4)
    for (int nr = 1; nr <= 8 && nr <= std::max(1, nmin); nr *= 2)
      for (int ident = 0; ident <= 1; ident++) {
        const int cand = make_shape(nr, lanes, ident, 0);
        for (int l = 0; l < n; l++) {
          DsArgs hop = runs[l].hop, mdm = runs[l].mdm;
          const int cl = cs[l]->tiles ? cand : make_shape(nr, lanes, 1, 0);
          hop.bms_ms = nullptr;  // no scalar step while timing
          hop.skip = &runs[l].bsh[0].stop[1];
          mdm.skip = hop.skip;
          hop.iter_off = -1;
          mdm.iter_off = -1;
          dslash_prep<DS_HOP, RED_NORM>(*cs[l], hop, cl);
          dslash_prep<DS_BMS, RED_NORM>(*cs[l], mdm, cl);
          mdm.bms_hnbr = hop.red.nbr;
          hop.bms_nbr = mdm.red.nbr;
          tab[l] = hop;
          tab[n + l] = mdm;
        }
        if (!h2d(c0, dt, tab.data(), sizeof(DsArgs) * tab.size())) return -1;
        for (int rep = 0; rep < 4; rep++) {
          if (rep == 1) hipEventRecord(e0, c0.stream);
          dslash_group_launch<R, LINK, DS_HOP, RED_NORM>(c0.stream, dt, tab.data(), n, cand);
          dslash_group_launch<R, LINK, DS_BMS, RED_NORM>(c0.stream, dt + n, tab.data() + n, n, cand);
        }
        hipEventRecord(e1, c0.stream);
        hipEventSynchronize(e1);
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) {
          best = ms;
          code = cand;
        }
      }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (!hip_ok(c0, hipGetLastError(), "group shape")) return -1;
  gt[key] = code;
  return code;
}

// Returns kBasisFull when some lattice's basis does not fit (nothing has run
// then beyond set-up; the caller solves the lattices one by one), -1 on a
// HIP error, else 0 with each job's first-pass status in jobs[l].out.
template <typename R, int LINK>
static int ms_cg_basis_group(std::vector<Context *> &cs, std::vector<GroupJob> &jobs) {
  const int n = (int)cs.size();
  if (n < 1 || n > kMaxGroup) return kBasisFull;
  Context &c0 = *cs[0];
  int *iterc = iter_ptr(c0);
  std::vector<BasisRun> runs(n);
  int nmin = 1 << 30;
  long cap_max = 0;
  for (int l = 0; l < n; l++) {
    GroupJob &j = jobs[l];
    const int rc = bms_setup<R, LINK>(*cs[l], runs[l], j.b, j.nrhs, j.nshift, j.shifts, j.res_abs, j.max_iter, iterc);
    if (rc) return rc;
    BasisRun &br = runs[l];
    br.timed = false;
    for (DsArgs *d : {&br.hop, &br.mdm}) {
      d->clk = nullptr;
      d->clk_wg = nullptr;
      d->clk_prev = nullptr;
    }
    nmin = std::min(nmin, j.nrhs);
    cap_max = std::max(cap_max, br.cap);
  }
  const int code = group_shape<R, LINK>(cs, runs, nmin);
  if (code < 0) return -1;
  // argument tables of the kChunk steps: [q][hop, M^dag M][lattice]
  std::vector<DsArgs> &htab = c0.ws->group_tab;
  htab.assign((size_t)kChunk * 2 * n, DsArgs{});
  for (int q = 0; q < kChunk; q++)
    for (int l = 0; l < n; l++) {
      BasisRun &br = runs[l];
      DsArgs hop = br.hop, mdm = br.mdm;
      const int cl = cs[l]->tiles ? code : make_shape(shape_nr(code), shape_lanes(code), 1, 0);
      hop.skip = &br.bsh[0].stop[(1 + q) & 1];
      mdm.skip = hop.skip;
      hop.iter_off = q - 1;
      mdm.iter_off = q - 1;
      dslash_prep<DS_HOP, RED_NORM>(*cs[l], hop, cl);
      dslash_prep<DS_BMS, RED_NORM>(*cs[l], mdm, cl);
      mdm.bms_hnbr = hop.red.nbr;
      hop.bms_nbr = mdm.red.nbr;  // the hop of step k reads the |r|^2 partials of M^dag M at k-1
      htab[((size_t)q * 2 + 0) * n + l] = hop;
      htab[((size_t)q * 2 + 1) * n + l] = mdm;
    }
  DsArgs *dtab = (DsArgs *)ws_get(c0, S_GROUP_TAB, sizeof(DsArgs) * htab.size());
  if (!dtab) return -1;
  if (!h2d(c0, dtab, htab.data(), sizeof(DsArgs) * htab.size())) return -1;
  GraphSlot &gs = c0.ws->group_graph[sizeof(R) == 8 ? 1 : 0];
  unsigned long long key[3];
  key[0] = fnv1a(htab.data(), sizeof(DsArgs) * htab.size());
  key[1] = (unsigned long long)(size_t)dtab;
  key[2] = ((unsigned long long)n << 32) | (unsigned)code;
  std::vector<const int *> active(n);
  for (int l = 0; l < n; l++) active[l] = runs[l].a.active;
  const ChunkPoll poll = chunk_poll(c0, active);
  auto launch = [&]() {
    for (int q = 0; q < kChunk; q++) {
      const size_t h0 = ((size_t)q * 2 + 0) * n, m0 = ((size_t)q * 2 + 1) * n;
      dslash_group_launch<R, LINK, DS_HOP, RED_NORM>(c0.stream, dtab + h0, htab.data() + h0, n, code);
      dslash_group_launch<R, LINK, DS_BMS, RED_NORM>(c0.stream, dtab + m0, htab.data() + m0, n, code);
    }
    hipLaunchKernelGGL(k_chunk_end, dim3(1), dim3(64), 0, c0.stream, iterc, kChunk, poll);
  };
  auto chunk = [&](bool) -> bool { return chunk_run(c0, gs, false, key, sizeof(key), launch); };
  if (bms_iterate(c0, iterc, active, cap_max, chunk)) return -1;
  for (int l = 0; l < n; l++) {
    GroupJob &j = jobs[l];
    j.out.iters = 0;
    j.out.rc = bms_finish<R, LINK>(*cs[l], runs[l], j.x, &j.out.iters, j.x32, j.asm_fp32, &j.out.rs);
    if (j.out.rc < 0) return -1;
  }
  return 0;
}

// r^(K) of a continued system (basis slot K, precision R) as an FP64 source
template <typename R>
__global__ __launch_bounds__(kBlock) void k_slot_to_src(const Cx<R> *v, double2 *out, long n) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = make_double2((double)v[i].x, (double)v[i].y);
}

// x[orig[q]][s] += scale[q][s] y[q][s] for the continued systems
template <typename XO>
__global__ __launch_bounds__(kBlock) void k_add_scaled(XO *x, const XO *y, const int *orig, const double *scale,
                                                       int nshift, long n3) {
  const long e = (long)blockIdx.x * kBlock + threadIdx.x;
  const int qs = blockIdx.y, q = qs / nshift, s = qs - q * nshift;
  const double f = scale[qs];
  if (e >= n3 || f == 0.0) return;
  const XO v = y[(long)qs * n3 + e];
  XO &o = x[((long)orig[q] * nshift + s) * n3 + e];
  o.x += f * v.x;
  o.y += f * v.y;
}

// Stored-basis multi-shift solve continued past a full basis: each
// continuation solves (MdagM + s) y_s = r^(K) for the systems still running,
// with the target divided by the largest |zeta_s| (so every shift reaches
// res_abs), and adds zeta_s y_s to their solutions; factors compound over
// continuations.  Returns 0, kStatusMaxCg, kStatusNonFinite, kBasisFull (the
// first call did not fit) or -1.
// first: when not null, the first pass was already run (multi-lattice solve).
template <typename R, int LINK>
// res_abs null: the first pass's thresholds are max(tol_abs, tol_rel |b|),
// formed on the device; bn2_out (if given) receives the |b|^2 it used.
static int ms_basis_solve(Context &c, const double2 *b, int nrhs, int nshift, const double *shifts,
                          const double *res_abs, int max_iter, double2 *x, int *iters_out, float2 *x32 = nullptr,
                          bool asm_fp32 = false, FirstPass *first = nullptr, double tol_abs = 0.0,
                          double tol_rel = 0.0, std::vector<double> *bn2_out = nullptr) {
  using C = Cx<R>;
  BasisRestart rs;
  int it = 0;
  int rc;
  if (first) {
    rc = first->rc;
    it = first->iters;
    rs = first->rs;
  } else {
    rc = ms_cg_basis<R, LINK>(c, b, nrhs, nshift, shifts, res_abs, max_iter, x, &it, x32, asm_fp32, &rs, tol_abs,
                              tol_rel);
  }
  // the first pass's thresholds (continuations divide them by the largest |zeta|)
  std::vector<double> res_first;
  if (!res_abs) {
    if (rc < 0 || rc == kBasisFull) return rc;
    res_first = rs.res;
    res_abs = res_first.data();
    if (bn2_out) *bn2_out = rs.bnorm2;
  }
  int total = it;
  std::vector<int> orig(nrhs);
  for (int j = 0; j < nrhs; j++) orig[j] = j;
  std::vector<double> scale((size_t)nrhs * nshift, 1.0);
  const long n3 = 3 * c.Vh;
  while (rc == kBasisRestart) {
    c.last_restarts++;
    const int n2 = (int)rs.rhs.size();
    std::vector<int> orig2(n2);
    std::vector<double> scale2((size_t)n2 * nshift), res2(n2);
    for (int q = 0; q < n2; q++) {
      const int j = rs.rhs[q];
      orig2[q] = orig[j];
      double zmax = 0.0;
      for (int s = 0; s < nshift; s++) {
        const double f = scale[(size_t)j * nshift + s] * rs.zeta[(size_t)q * nshift + s];
        scale2[(size_t)q * nshift + s] = f;
        zmax = std::max(zmax, std::fabs(f));
      }
      res2[q] = zmax > 0.0 ? res_abs[orig2[q]] / zmax : res_abs[orig2[q]];
    }
    double2 *src = (double2 *)ws_get(c, S_RS_SRC, sizeof(double2) * n3 * n2);
    const size_t yb = (x32 ? sizeof(float2) : sizeof(double2)) * n3 * (size_t)n2 * nshift;
    void *y = ws_get(c, S_RS_Y, yb);
    int *dorig = (int *)ws_get(c, S_RS_IDX, sizeof(int) * 256 + sizeof(double) * 256 * kMaxShifts);
    if (!src || !y || !dorig || n2 * nshift > 256 * kMaxShifts) return -1;
    double *dscale = (double *)(dorig + 256);
    for (int q = 0; q < n2; q++) {
      const C *v = (const C *)rs.basis + ((size_t)rs.slot[q] * rs.nrhs + rs.rhs[q]) * n3;
      hipLaunchKernelGGL(k_slot_to_src<R>, dim3((unsigned)((n3 + kBlock - 1) / kBlock)), dim3(kBlock), 0, c.stream,
                         v, src + (size_t)q * n3, n3);
    }
    if (!h2d(c, dorig, orig2.data(), sizeof(int) * n2) || !h2d(c, dscale, scale2.data(), sizeof(double) * scale2.size()))
      return -1;
    it = 0;
    rc = ms_cg_basis<R, LINK>(c, src, n2, nshift, shifts, res2.data(), std::max(1, max_iter - total),
                              x32 ? nullptr : (double2 *)y, &it, x32 ? (float2 *)y : nullptr, asm_fp32, &rs);
    total += it;
    if (rc < 0 || rc == kStatusNonFinite || rc == kBasisFull) return rc == kBasisFull ? -1 : rc;
    const dim3 grid((unsigned)((n3 + kBlock - 1) / kBlock), (unsigned)(n2 * nshift));
    if (x32)
      hipLaunchKernelGGL(k_add_scaled<float2>, grid, dim3(kBlock), 0, c.stream, x32, (const float2 *)y, dorig,
                         dscale, nshift, n3);
    else
      hipLaunchKernelGGL(k_add_scaled<double2>, grid, dim3(kBlock), 0, c.stream, x, (const double2 *)y, dorig,
                         dscale, nshift, n3);
    orig = orig2;
    scale = scale2;
    if (rc == kBasisRestart && total >= max_iter) rc = kStatusMaxCg;
  }
  *iters_out = total;
  if (!hip_ok(c, hipGetLastError(), "basis continuation")) return -1;
  return rc;
}

// batched CG on (MdagM + shift_j) e_j = r_j (FP32 iterations, FP32 links);
// x32 receives e_j.
static int block_cg32(Context &c, const double2 *r64, int nsys, const double *shifts, const double *tol,
                      int max_iter, float2 *x32, int *iters_out) {
  const long Vh = c.Vh;
  const size_t vb = sizeof(float2) * 3 * Vh;
  void *r = ws_get(c, S_CR32, vb * nsys);
  void *p = ws_get(c, S_CP32, vb * nsys);
  void *s = ws_get(c, S_CS32, vb * nsys);
  void *hh = ws_get(c, S_CH32, vb * nsys);
  CgState *cg = (CgState *)ws_get(c, S_CG, sizeof(CgState) * nsys);
  double *diag = (double *)ws_get(c, S_DIAG, sizeof(double) * 256);
  if (!r || !p || !s || !hh || !cg || !diag) return -1;
  std::vector<CgState> hs(nsys);
  std::vector<double> hd(nsys);
  for (int j = 0; j < nsys; j++) {
    memset(&hs[j], 0, sizeof(CgState));
    hs[j].tol2 = tol[j] * tol[j];
    hs[j].shift = shifts[j];
    hs[j].max_iter = max_iter;
    hd[j] = c.mass * c.mass + shifts[j];
  }
  if (!h2d(c, cg, hs.data(), sizeof(CgState) * nsys) || !h2d(c, diag, hd.data(), sizeof(double) * nsys)) return -1;
  const unsigned nbx = nbx_of(c);
  BlasArgs ba = {};
  ba.Vh = Vh; ba.nvec = nsys; ba.nshift = 1;
  ba.b = r64; ba.r = r; ba.p = p; ba.s = s; ba.x = x32; ba.cg = cg;
  ba.red = make_red(c, nsys, EPI_CG_INIT, cg, nullptr);
  const unsigned nbr = reduce_blocks(c, nsys);
  ba.red.nbr = (int)nbr;
  hipLaunchKernelGGL(k_cg_init<float>, dim3(nbr, nsys), dim3(kBlock), 0, c.stream, ba);
  finalize(c, ba.red, nsys, true);
  RedArgs red_c = ba.red;
  red_c.epi = EPI_CG_C;
  DsArgs hop = {};
  hop.par_out = 1; hop.U = c.U32;
  hop.in = p; hop.in_vs = 3 * Vh; hop.out = hh; hop.out_vs = 3 * Vh; hop.nvec = nsys;
  hop.skip = &cg[0].stop; hop.skip_stride = sizeof(CgState) / sizeof(int);
  DsArgs mdm = hop;
  mdm.par_out = 0; mdm.in = hh; mdm.out = s; mdm.pin = p; mdm.pin_vs = 3 * Vh; mdm.diag = diag;
  mdm.red = make_red(c, nsys, EPI_CG_B, cg, nullptr);
  const int chunk = 8;
  int it = 0;
  while (!none_active(c)) {
    for (int k = 0; k < chunk; k++) {
      dslash_dispatch<float, LINK_S12, DS_HOP, RED_NONE>(c, hop);
      dslash_dispatch<float, LINK_S12, DS_MDAGM, RED_DOT>(c, mdm);
      finalize(c, mdm.red, nsys, false);
      hipLaunchKernelGGL(k_cg_update_xr<float>, dim3(nbr, nsys), dim3(kBlock), 0, c.stream, ba);
      finalize(c, red_c, nsys, true);
      hipLaunchKernelGGL(k_cg_update_p<float>, dim3(nbx, nsys), dim3(kBlock), 0, c.stream, ba);
    }
    it += chunk;
    if (it > max_iter + chunk) break;
  }
  if (!d2h_sync(c, hs.data(), cg, sizeof(CgState) * nsys)) return -1;
  int maxit = 0;
  for (int j = 0; j < nsys; j++) maxit = std::max(maxit, hs[j].iter);
  *iters_out = maxit;
  for (int j = 0; j < nsys; j++)
    if (hs[j].nonfinite) {
      c.last_error = "defect correction: non-finite value in the solver reductions";
      return kStatusNonFinite;
    }
  return 0;
}

// r = b - ((m^2+shift_j) x_j - D_eo D_oe x_j) in FP64 with FP64 links; norms2[j]
int residual_d(Context &c, const double2 *b, int b_group, const double2 *x, const double *shifts, int nvec,
               double2 *r, double *norms2) {
  const long Vh = c.Vh;
  double2 *hh = (double2 *)ws_get(c, S_FH64, sizeof(double2) * 3 * Vh * nvec);
  double *diag = (double *)ws_get(c, S_DIAG, sizeof(double) * 256);
  double *sums = (double *)ws_get(c, S_SUMS, sizeof(double) * 256);
  if (!hh || !diag || !sums || nvec > kMaxRed) return -1;
  std::vector<double> hd(nvec);
  for (int j = 0; j < nvec; j++) hd[j] = c.mass * c.mass + shifts[j];
  if (!h2d(c, diag, hd.data(), sizeof(double) * nvec)) return -1;
  launch_hop_d(c, 1, x, hh, nvec, 3 * Vh, 3 * Vh);
  DsArgs a = {};
  a.par_out = 0; a.U = c.U64;
  a.in = hh; a.in_vs = 3 * Vh;
  a.out = r; a.out_vs = 3 * Vh;
  a.pin = x; a.pin_vs = 3 * Vh;
  a.b = b; a.b_vs = 3 * Vh; a.b_group = b_group;
  a.diag = diag; a.nvec = nvec;
  a.red = make_red(c, nvec, EPI_STORE, nullptr, sums);
  dslash_dispatch<double, LINK_D18, DS_RESID, RED_NORM>(c, a);
  finalize(c, a.red, nvec, false);
  return d2h_sync(c, norms2, sums, sizeof(double) * nvec) ? 0 : -1;
}

// FP64 defect correction of the shifted solutions x ([nrhs][nshift], FP64):
// every system's true residual r = b - (MdagM + s) x is computed in FP64 with
// the FP64 links; the systems above their target solve (MdagM + s) e = r
// as one batched FP32 CG (to ~1e-5 of |r|, never below half the target) and
// x += e, until every system meets its target (at most 8 cycles).  ref_iters
// receives the FP32 iterations spent; returns kStatusMaxCg when the targets
// are not reached, kStatusNonFinite on NaN/Inf.
static int refine_fp64(Context &c, const double2 *b, int nrhs, int nshift, const double *shifts,
                       const double *res_abs, int max_iter, double2 *x, int *ref_iters) {
  const long Vh = c.Vh;
  const int nsys = nrhs * nshift;
  const long n = 3 * Vh * (long)nsys;
  std::vector<double> sh(nsys), nr2(nsys), tol(nsys);
  for (int j = 0; j < nsys; j++) sh[j] = shifts[j % nshift];
  double2 *r64 = (double2 *)ws_get(c, S_CR64, sizeof(double2) * 3 * Vh * nsys);
  if (!r64) return -1;
  float2 *e32 = nullptr;
  int total_ref = 0;
  double worst = 0.0;
  *ref_iters = 0;
  for (int cyc = 0;; cyc++) {
    if (residual_d(c, b, nshift, x, sh.data(), nsys, r64, nr2.data())) return -1;
    bool ok = true;
    worst = 0.0;
    for (int j = 0; j < nsys; j++) {
      if (!std::isfinite(nr2[j])) {
        c.last_error = "multishift: non-finite true residual";
        return kStatusNonFinite;
      }
      const double rn = sqrt(nr2[j]);
      const double target = res_abs[j / nshift];
      worst = std::max(worst, rn / target);
      // inner FP32 solve: reduce this residual by ~1e-5, never below 0.5*target;
      // systems already converged get a huge tolerance (done at the first test)
      tol[j] = rn <= target ? 1e150 : std::max(0.5 * target, 1e-5 * rn);
      if (rn > target) ok = false;
    }
    c.last_true_res_max = worst;
    if (ok || cyc == 8) break;
    if (!e32 && !(e32 = (float2 *)ws_get(c, S_CX32, sizeof(float2) * 3 * Vh * nsys))) return -1;
    int itc = 0;
    const int rc = block_cg32(c, r64, nsys, sh.data(), tol.data(), max_iter, e32, &itc);
    if (rc) return rc;
    total_ref += itc;
    hipLaunchKernelGGL(k_accum_f2d, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, c.stream, x,
                       e32, n);
    c.last_refine_cycles = cyc + 1;
  }
  *ref_iters = total_ref;
  if (!hip_ok(c, hipStreamSynchronize(c.stream), "defect correction")) return -1;
  if (worst > 1.0) {
    c.last_error = "multishift: defect correction did not reach the target";
    return kStatusMaxCg;
  }
  if (c.want_xr) {
    // r64 holds the true residuals of the final x (the last check passed)
    const size_t n0 = c.last_xr.size();
    c.last_xr.resize(n0 + nsys);
    if (vec_dots(c, x, r64, nsys, 3 * Vh, c.last_xr.data() + n0)) return -1;
  }
  return 0;
}

// first: the stored-basis first pass already run by a multi-lattice solve
// (for the precision's basis solver), or null
static int ms_solve(Context &c, const double2 *b, int nrhs, int nshift, const double *shifts, double res, int rel,
                    int prec, int max_iter, double2 *x, int *iters, float2 *x32, FirstPass *first) {
  if (x32 && prec != PREC_FP32) {
    c.last_error = "multishift: FP32 solutions only from the FP32 solver";
    return 1;
  }
  if (nshift < 1 || nshift > kMaxShifts || nrhs < 1 || nrhs * nshift > kMaxRed) {
    c.last_error = "multishift: unsupported number of shifts/rhs";
    return 1;
  }
  const long Vh = c.Vh;
  c.last_refine_cycles = 0;
  c.last_iters_refine = 0;
  c.last_restarts = 0;
  c.last_true_res_max = 0.0;
  const bool explicit_shifts = c.ms_algorithm == 1 || getenv("RHMC_MS_EXPLICIT") != nullptr;
  // |b| per rhs (relative residuals, the FP32 stage target, the FP64 checks).
  // The stored-basis pass forms it on the device and sets its own
  // thresholds there, max(tol_abs, tol_rel |b|), so the host waits for no
  // norm before the solve; the explicit-shift solver and a first pass run
  // by a multi-lattice solve take host norms.
  std::vector<double> bn2(nrhs), res_abs(nrhs);
  bool have_norms = false;
  auto set_norms = [&](const std::vector<double> &n2) {
    bn2 = n2;
    for (int j = 0; j < nrhs; j++) res_abs[j] = rel ? res * sqrt(bn2[j]) : res;
    have_norms = true;
  };
  auto host_norms = [&]() -> bool {
    if (have_norms) return true;
    std::vector<double> n2(nrhs);
    if (vec_norms2(c, b, nrhs, 3 * Vh, n2.data())) return false;
    set_norms(n2);
    return true;
  };
  if ((explicit_shifts || first) && !host_norms()) return -1;
  std::vector<double> dev_bn2;
  if (prec == PREC_FP64) {
    int rc = explicit_shifts ? kBasisFull
                             : ms_basis_solve<double, LINK_D12>(c, b, nrhs, nshift, shifts,
                                                                have_norms ? res_abs.data() : nullptr, max_iter, x,
                                                                iters, nullptr, false, first, rel ? 0.0 : res,
                                                                rel ? res : 0.0, &dev_bn2);
    if (rc < 0) return rc;
    if (!have_norms && rc != kBasisFull) set_norms(dev_bn2);
    const bool basis = rc != kBasisFull;
    if (rc == kBasisFull) {
      if (!host_norms()) return -1;
      rc = ms_cg<double, LINK_D12>(c, b, nrhs, nshift, shifts, res_abs.data(), max_iter, x, iters);
    }
    c.last_iters = *iters;
    c.last_iters_fp32 = 0;
    if (rc != 0 || !basis) return rc;
    // The stored-basis solver stops on its recursive residual, which in
    // the single-reduction form can drift above the true one: every
    // shifted system's FP64 true residual is checked and, where it is above
    // the target, reduced by defect correction.
    int ref = 0;
    rc = refine_fp64(c, b, nrhs, nshift, shifts, res_abs.data(), max_iter, x, &ref);
    c.last_iters_refine = ref;
    c.last_iters = *iters + ref;
    *iters = c.last_iters;
    return rc;
  }
  // mixed: FP32 multi-shift to ~1e-6 relative, then FP64 defect correction;
  // FP32 only: the FP32 recursion to the requested residual
  // the FP32 stage's threshold max(res_abs, rel32 |b|)
  const double rel32 = prec == PREC_FP32 ? 0.0 : 2e-6;
  std::vector<double> res32(nrhs);
  if (have_norms)
    for (int j = 0; j < nrhs; j++) res32[j] = std::max(res_abs[j], rel32 * sqrt(bn2[j]));
  int it32 = 0;
  const long n = 3 * Vh * (long)nrhs * nshift;
  int rc = explicit_shifts ? kBasisFull
                           : ms_basis_solve<float, LINK_S12>(c, b, nrhs, nshift, shifts,
                                                             have_norms ? res32.data() : nullptr, max_iter, x, &it32,
                                                             x32, prec == PREC_FP32, first, rel ? 0.0 : res,
                                                             rel ? std::max(res, rel32) : rel32, &dev_bn2);
  if (rc < 0) return rc;
  if (!have_norms && rc != kBasisFull) set_norms(dev_bn2);
  if (rc == kBasisFull) {
    if (!host_norms()) return -1;
    for (int j = 0; j < nrhs; j++) res32[j] = std::max(res_abs[j], rel32 * sqrt(bn2[j]));
    float2 *xs = x32 ? x32 : (float2 *)ws_get(c, S_X32, sizeof(float2) * 3 * Vh * nrhs * nshift);
    if (!xs) return -1;
    rc = ms_cg<float, LINK_S12>(c, b, nrhs, nshift, shifts, res32.data(), max_iter, xs, &it32);
    if (rc < 0 || rc == kStatusNonFinite) return rc;
    if (!x32)
      hipLaunchKernelGGL(k_copy_f2d, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, c.stream, x, xs, n);
  }
  if (rc < 0 || rc == kStatusNonFinite) return rc;
  if (prec == PREC_FP32) {
    if (!hip_ok(c, hipGetLastError(), "fp32 solve")) return -1;
    c.last_iters_fp32 = it32;
    c.last_iters = it32;
    *iters = it32;
    return rc;
  }
  int ref = 0;
  const int rr = refine_fp64(c, b, nrhs, nshift, shifts, res_abs.data(), max_iter, x, &ref);
  c.last_iters_fp32 = it32;
  c.last_iters_refine = ref;
  c.last_iters = it32 + ref;
  *iters = c.last_iters;
  return rr ? rr : rc;
}

int multishift_solve(Context &c, const double2 *b, int nrhs, int nshift, const double *shifts, double res, int rel,
                     int prec, int max_iter, double2 *x, int *iters, float2 *x32) {
  return ms_solve(c, b, nrhs, nshift, shifts, res, rel, prec, max_iter, x, iters, x32, nullptr);
}

// Several lattices' solves of one precision in lockstep (ms_cg_basis_group
// for the stored-basis pass), then each lattice's continuation and FP64
// defect correction as in multishift_solve.  Falls back to one lattice
// after another when the precisions differ, the explicit-shift solver is
// selected, or a basis does not fit.  All contexts on one stream.
int multishift_solve_group(std::vector<Context *> &cs, std::vector<MsSpec> &sp) {
  const int n = (int)cs.size();
  bool group = n > 1 && n <= kMaxGroup && getenv("RHMC_MS_EXPLICIT") == nullptr;
  for (int l = 0; l < n && group; l++) {
    const MsSpec &q = sp[l];
    group = cs[l]->ms_algorithm == 0 && q.prec == sp[0].prec && !(q.x32 && q.prec != PREC_FP32) &&
            q.nshift >= 1 && q.nshift <= kMaxShifts && q.nrhs >= 1 && q.nrhs * q.nshift <= kMaxRed;
  }
  std::vector<FirstPass> first(n);
  bool have_first = false;
  if (group) {
    std::vector<GroupJob> jobs(n);
    std::vector<std::vector<double>> res_abs(n);
    const int prec = sp[0].prec;
    for (int l = 0; l < n; l++) {
      Context &c = *cs[l];
      MsSpec &q = sp[l];
      std::vector<double> bn2(q.nrhs);
      if (vec_norms2(c, q.b, q.nrhs, 3 * c.Vh, bn2.data())) return -1;
      res_abs[l].resize(q.nrhs);
      // the basis pass's target (ms_solve: FP32 stage of the mixed solver at 2e-6 |b|)
      for (int j = 0; j < q.nrhs; j++) {
        const double ra = q.rel ? q.res * sqrt(bn2[j]) : q.res;
        res_abs[l][j] = prec == PREC_MIXED ? std::max(ra, 2e-6 * sqrt(bn2[j])) : ra;
      }
      jobs[l] = GroupJob{q.b, q.nrhs, q.nshift, q.shifts, res_abs[l].data(), q.max_iter, q.x,
                         prec == PREC_FP32 ? q.x32 : nullptr, prec == PREC_FP32, FirstPass()};
    }
    const int rc = prec == PREC_FP64 ? ms_cg_basis_group<double, LINK_D12>(cs, jobs)
                                     : ms_cg_basis_group<float, LINK_S12>(cs, jobs);
    if (rc < 0) return rc;
    if (rc == 0) {
      have_first = true;
      for (int l = 0; l < n; l++) first[l] = jobs[l].out;
    }
  }
  int worst = 0;
  for (int l = 0; l < n; l++) {
    MsSpec &q = sp[l];
    q.iters = 0;
    q.rc = ms_solve(*cs[l], q.b, q.nrhs, q.nshift, q.shifts, q.res, q.rel, q.prec, q.max_iter, q.x, &q.iters, q.x32,
                    have_first ? &first[l] : nullptr);
    if (q.rc < 0) return -1;
    if (q.rc && !worst) worst = q.rc;
  }
  return worst;
}

// ---------------------------------------------------------- norms / scaling
__global__ __launch_bounds__(kBlock) void k_norms(const double2 *v, long vs, long Vh, int nvec, RedArgs red) {
  const unsigned bx = xcd_block(blockIdx.x, gridDim.x);
  const int j = blockIdx.y;
  long h0, h1;
  site_range(Vh, bx, gridDim.x, h0, h1);
  double s[1] = {0.0};
  for (long h = h0 + threadIdx.x; h < h1; h += kBlock)
    for (int i = 0; i < 3; i++) {
      const double2 a = v[(long)j * vs + (long)i * Vh + h];
      s[0] += a.x * a.x + a.y * a.y;
    }
  block_partials<1>(s, red.partial, j, nvec, gridDim.x, bx);
}

__global__ __launch_bounds__(kBlock) void k_dots(const double2 *a, const double2 *b, long vs, long Vh, int nvec,
                                                 RedArgs red) {
  const unsigned bx = xcd_block(blockIdx.x, gridDim.x);
  const int j = blockIdx.y;
  long h0, h1;
  site_range(Vh, bx, gridDim.x, h0, h1);
  double s[1] = {0.0};
  for (long h = h0 + threadIdx.x; h < h1; h += kBlock)
    for (int i = 0; i < 3; i++) {
      const double2 u = a[(long)j * vs + (long)i * Vh + h], v = b[(long)j * vs + (long)i * Vh + h];
      s[0] += u.x * v.x + u.y * v.y;
    }
  block_partials<1>(s, red.partial, j, nvec, gridDim.x, bx);
}

int vec_dots(Context &c, const double2 *a, const double2 *b, int nvec, long vs, double *out) {
  if (nvec > kMaxRed) return -1;
  double *sums = (double *)ws_get(c, S_SUMS, sizeof(double) * 256);
  RedArgs red = make_red(c, nvec, EPI_STORE, nullptr, sums);
  const unsigned nbr = reduce_blocks(c, nvec);
  red.nbr = (int)nbr;
  hipLaunchKernelGGL(k_dots, dim3(nbr, nvec), dim3(kBlock), 0, c.stream, a, b, vs, c.Vh, nvec, red);
  finalize(c, red, nvec, false);
  return d2h_sync(c, out, sums, sizeof(double) * nvec) ? 0 : -1;
}

// Re<a1,b1> - Re<a2,b2> per vector, differenced per element before the
// sum: for two nearly equal dot products of large magnitude the result
// carries the rounding of the (small) differences, not one ulp of each
// product.
__global__ __launch_bounds__(kBlock) void k_dot_diffs(const double2 *a1, const double2 *b1, const double2 *a2,
                                                      const double2 *b2, long vs, long Vh, int nvec, RedArgs red) {
  const unsigned bx = xcd_block(blockIdx.x, gridDim.x);
  const int j = blockIdx.y;
  long h0, h1;
  site_range(Vh, bx, gridDim.x, h0, h1);
  double s[1] = {0.0};
  for (long h = h0 + threadIdx.x; h < h1; h += kBlock)
    for (int i = 0; i < 3; i++) {
      const long e = (long)j * vs + (long)i * Vh + h;
      const double2 u1 = a1[e], v1 = b1[e], u2 = a2[e], v2 = b2[e];
      s[0] += (u1.x * v1.x + u1.y * v1.y) - (u2.x * v2.x + u2.y * v2.y);
    }
  block_partials<1>(s, red.partial, j, nvec, gridDim.x, bx);
}

int vec_dot_diffs(Context &c, const double2 *a1, const double2 *b1, const double2 *a2, const double2 *b2, int nvec,
                  long vs, double *out) {
  if (nvec > kMaxRed) return -1;
  double *sums = (double *)ws_get(c, S_SUMS, sizeof(double) * 256);
  RedArgs red = make_red(c, nvec, EPI_STORE, nullptr, sums);
  const unsigned nbr = reduce_blocks(c, nvec);
  red.nbr = (int)nbr;
  hipLaunchKernelGGL(k_dot_diffs, dim3(nbr, nvec), dim3(kBlock), 0, c.stream, a1, b1, a2, b2, vs, c.Vh, nvec, red);
  finalize(c, red, nvec, false);
  return d2h_sync(c, out, sums, sizeof(double) * nvec) ? 0 : -1;
}

int vec_norms2(Context &c, const double2 *v, int nvec, long vs, double *out) {
  if (nvec > kMaxRed) return -1;
  double *sums = (double *)ws_get(c, S_SUMS, sizeof(double) * 256);
  RedArgs red = make_red(c, nvec, EPI_STORE, nullptr, sums);
  const unsigned nbr = reduce_blocks(c, nvec);
  red.nbr = (int)nbr;
  hipLaunchKernelGGL(k_norms, dim3(nbr, nvec), dim3(kBlock), 0, c.stream, v, vs, c.Vh, nvec, red);
  finalize(c, red, nvec, false);
  return d2h_sync(c, out, sums, sizeof(double) * nvec) ? 0 : -1;
}

// p = a*r - p
__global__ __launch_bounds__(kBlock) void k_axmy(double2 *p, const double2 *r, double a, long n) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    const double2 rv = r[i];
    double2 pv = p[i];
    pv.x = a * rv.x - pv.x;
    pv.y = a * rv.y - pv.y;
    p[i] = pv;
  }
}

// r = p / sqrt(*n2) (the norm of the previous iteration, read on the device)
__global__ __launch_bounds__(kBlock) void k_scale_by_norm(double2 *r, const double2 *p, const double *n2, long n) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    const double s = 1.0 / sqrt(*n2);
    const double2 v = p[i];
    r[i] = make_double2(s * v.x, s * v.y);
  }
}

// |v|^2 of one vector into *dst on the device (no host round trip)
static void norm2_to(Context &c, const double2 *v, long n, double *dst) {
  RedArgs red = make_red(c, 1, EPI_STORE, nullptr, dst);
  const unsigned nbr = reduce_blocks(c, 1);
  red.nbr = (int)nbr;
  hipLaunchKernelGGL(k_norms, dim3(nbr, 1), dim3(kBlock), 0, c.stream, v, n, n / 3, 1, red);
  finalize(c, red, 1, false);
}

// Power iterations of findminmax.cc:12-137 in FP64: p <- A r with r =
// p / |p|, A = M^dag M (largest eigenvalue |A r| at convergence) or
// max - M^dag M (then min = max - |A r|), until the relative change of |p|
// is at most tol.  Every iteration stores |p|^2 on the device and
// normalises with it there; the host reads the norms after each chunk of
// kPowChunk iterations and stops at the first iteration that met the test,
// whose value it returns (the chunk's later iterations are discarded), so
// the result is the one a test after every iteration gives, without a
// host round trip per iteration.
constexpr int kPowChunk = 8;
constexpr int kPowMax = 20000;
static int power_iter(Context &c, const double2 *start, double tol, int is_min, double mx, double *out, int *loops) {
  const long Vh = c.Vh, n = 3 * Vh;
  double2 *p = (double2 *)ws_get(c, S_PW1, sizeof(double2) * n);
  double2 *r = (double2 *)ws_get(c, S_PW2, sizeof(double2) * n);
  double2 *hh = (double2 *)ws_get(c, S_PW3, sizeof(double2) * n);
  double *diag = (double *)ws_get(c, S_DIAG, sizeof(double) * 256);
  double *pn = (double *)ws_get(c, S_PWN, sizeof(double) * (kPowMax + kPowChunk + 1));
  if (!p || !r || !hh || !diag || !pn) return -1;
  hipMemcpyAsync(p, start, sizeof(double2) * n, hipMemcpyDeviceToDevice, c.stream);
  const double m2 = c.mass * c.mass;
  if (!h2d(c, diag, &m2, sizeof(double))) return -1;
  norm2_to(c, p, n, pn);  // pn[0] = |start|^2
  const unsigned nb = (unsigned)((n + kBlock - 1) / kBlock);
  double prev = -1.0, norm = 0.0;
  double h[kPowChunk + 1];
  int done = 0;
  for (int i0 = 1; !done && i0 <= kPowMax; i0 += kPowChunk) {
    for (int i = i0; i < i0 + kPowChunk; i++) {
      hipLaunchKernelGGL(k_scale_by_norm, dim3(nb), dim3(kBlock), 0, c.stream, r, p, pn + i - 1, n);
      launch_hop_d(c, 1, r, hh, 1, n, n);
      DsArgs a = {};
      a.par_out = 0; a.U = c.U64; a.in = hh; a.in_vs = n; a.out = p; a.out_vs = n; a.pin = r; a.pin_vs = n;
      a.diag = diag; a.nvec = 1;
      if (!is_min) {
        a.red = make_red(c, 1, EPI_STORE, nullptr, pn + i);
        dslash_dispatch<double, LINK_D18, DS_MDAGM, RED_NORM>(c, a);
        finalize(c, a.red, 1, false);
      } else {
        dslash_dispatch<double, LINK_D18, DS_MDAGM, RED_NONE>(c, a);
        hipLaunchKernelGGL(k_axmy, dim3(nb), dim3(kBlock), 0, c.stream, p, r, mx, n);
        norm2_to(c, p, n, pn + i);
      }
    }
    if (!d2h_sync(c, h, pn + i0 - 1, sizeof(double) * (kPowChunk + 1))) return -1;
    if (prev < 0.0) prev = sqrt(h[0]);
    for (int q = 1; q <= kPowChunk && i0 + q - 1 <= kPowMax; q++) {
      norm = sqrt(h[q]);
      const double rel = fabs(prev - norm) / norm;
      prev = norm;
      if (!(rel > tol) || i0 + q - 1 == kPowMax) {
        *loops = i0 + q - 1;
        done = 1;
        break;
      }
    }
  }
  *out = is_min ? mx - norm : norm;
  return hip_ok(c, hipStreamSynchronize(c.stream), "power") ? 0 : -1;
}

int power_max(Context &c, const double2 *start, double tol, double *out, int *loops) {
  return power_iter(c, start, tol, 0, 0.0, out, loops);
}
int power_min(Context &c, const double2 *start, double tol, double mx, double *out, int *loops) {
  return power_iter(c, start, tol, 1, mx, out, loops);
}

// ----------------------------------------------------- device ops for tests
// out (FP32, [nvec][3][Vh]) = 0.5 hop(in) with FP32 12-real links
int hop_s12(Context &c, int par_out, const float2 *in, float2 *out, int nvec) {
  DsArgs a = {};
  a.par_out = par_out; a.U = c.U32;
  a.in = in; a.in_vs = 3 * c.Vh; a.out = out; a.out_vs = 3 * c.Vh; a.nvec = nvec;
  dslash_dispatch<float, LINK_S12, DS_HOP, RED_NONE>(c, a);
  return 0;
}

}  // namespace rhmc
