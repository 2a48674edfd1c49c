// coulomb.hip -- Coulomb gauge fixing with every time slice batched on the
// device: FACG (or FASD) per slice with Fourier acceleration over the three
// spatial dimensions, per-slice scalars and convergence flags, so slices
// that converge early stop while the others continue in the same launches.
//
// GLU fixes the slices one after another (Gfix/CFACG.c:532-636); they are
// independent problems (each uses only its slice's spatial links), so the
// batched result is the same slice by slice.  Per slice this follows
// steep_step_FACG (Gfix/CFACG.c:286-382), steep_step_FASD (:385-430),
// steep_fix with continuation and random restarts (:460-530), the line
// search of Gfix/line_search.c:101-271 and the slice transforms of
// Gfix/gtrans.c:242-264.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <utility>
#include <vector>

#include "../../include/glu_gfix.h"
#include "gfix_ctx.h"

using namespace gfix;

// per-slice control arrays (device)
struct SliceCtrl {
  double *theta, *inold, *insum, *sconj, *beta, *alpha, *probe;  // probe: 3 per slice
  double *dev;                                                   // unitarity deviation
  long long *iters;
  int *done, *did_ls;
  // per-slice trajectory rows (theta_k, step, beta_k) of the current fix,
  // kSliceTraceCap per slice: row 0 the SD start, then one per iteration,
  // restarts and continuations appended
  double *trace;
  long long *trace_n;
  // row -> time slice of the launches of a batch: only the slices still
  // running at the batch's start get blocks
  int *map;
  double acc, gf_alpha, cg_tol, norm;
  long long max_iters;
  int T;
};

constexpr int kSliceTraceCap = 64;

__device__ __forceinline__ void slice_rec(const SliceCtrl &sc, int t, double theta, double alpha, double beta) {
  const long long n = sc.trace_n[t];
  if (n >= kSliceTraceCap) return;
  double *r = sc.trace + ((long long)t * kSliceTraceCap + n) * 3;
  r[0] = theta;
  r[1] = alpha;
  r[2] = beta;
  sc.trace_n[t] = n + 1;
}

struct CoulombState {
  void *G = nullptr;      // slice gauge transforms, one per site (column planes)
  void *arena = nullptr;  // SliceCtrl arrays
  SliceCtrl sc{};
  long long nbs = 0;      // blocks per slice
  std::vector<size_t> slice_iters;
  std::vector<double> slice_theta;
};

namespace {

constexpr double kCgTol = 5e-11;
constexpr double PC1 = 0.17, PC2 = 0.34;  // Gfix/line_search.c:30-31

__device__ __forceinline__ long long site_of(const Geo &g, long long t, long long xl) { return t * g.Lcu + xl; }

// (slice, block in slice) of this workgroup of a (blocks per slice, T)
// grid.  The dispatcher deals workgroups to the 8 XCDs round-robin.  With
// g.xcd set (and 8 | blocks per slice) XCD x is given the contiguous eighth
// x of every slice instead, so a block's +-z neighbours (Lx Ly / 256 blocks
// away) were fetched into the same L2 just before, while every XCD keeps a
// share of every slice (slices stop at different iterations; whole slices
// per XCD would leave some XCDs idle).  Partials are indexed by (slice,
// block), so results do not depend on the order.
__device__ __forceinline__ void slice_block(const Geo &g, const SliceCtrl &sc, int &t, long long &bx) {
  const unsigned l = blockIdx.x + blockIdx.y * gridDim.x, nbs = gridDim.x;
  int row;
  if (g.xcd && (nbs & 7u) == 0) {
    const unsigned e = nbs >> 3, x = l & 7u, k = l >> 3;
    row = (int)(k / e);
    bx = (long long)(x * e + k % e);
  } else {
    row = (int)(l / nbs);
    bx = (long long)(l % nbs);
  }
  t = sc.map[row];
}

__device__ __forceinline__ void herm_short(const su3<double> &U, double s[10]) {
  // Matrix_Ops/exactQ.c:400-428: {Q00, Q01, Q02, Q11, Q12}
  const double i0 = U.e[0].im, i4 = U.e[4].im, i8 = U.e[8].im;
  s[0] = (2.0 * i0 - i4 - i8) / 3.0;
  s[1] = 0.0;
  // i/2 (conj(U10) - U01)
  s[2] = -0.5 * (-U.e[3].im - U.e[1].im);
  s[3] = 0.5 * (U.e[3].re - U.e[1].re);
  s[4] = -0.5 * (-U.e[6].im - U.e[2].im);
  s[5] = 0.5 * (U.e[6].re - U.e[2].re);
  s[6] = (2.0 * i4 - i0 - i8) / 3.0;
  s[7] = 0.0;
  s[8] = -0.5 * (-U.e[7].im - U.e[5].im);
  s[9] = 0.5 * (U.e[7].re - U.e[5].re);
}

// derivative of the gauge-rotated spatial links: Gfix/CFACG.c:92-163
__global__ __launch_bounds__(kBlock) void k_cderiv(const double2 *__restrict__ U, const double2 *__restrict__ G,
                                                  double2 *__restrict__ F, Geo g, SliceCtrl sc) {
  int t;
  long long bx;
  slice_block(g, sc, t, bx);
  if (sc.done[t]) return;
  const long long xl = bx * blockDim.x + threadIdx.x;
  if (xl >= g.Lcu) return;
  const long long i = site_of(g, t, xl), V = g.V;
  int x[4];
  coords(g, i, x);
  const su3<double> Gi = load_link<double>(G, 0, i, V);
  double sum[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  // one direction at a time, each rotated link projected as soon as it is
  // formed (the same arithmetic, fewer registers held)
#pragma unroll 1
  for (int mu = 0; mu < 3; mu++) {
    const long long it = nb_fwd(g, i, x, mu), bk = nb_bwd(g, i, x, mu);
    double a[10], b[10];
    {
      const su3<double> C = gtransform<double>(Gi, load_link<double>(U, mu, i, V), load_link<double>(G, 0, it, V));
      herm_short(C, b);
    }
    {
      const su3<double> D = gtransform<double>(load_link<double>(G, 0, bk, V), load_link<double>(U, mu, bk, V), Gi);
      herm_short(D, a);
    }
#pragma unroll
    for (int k = 0; k < 10; k++) sum[k] += a[k] - b[k];
  }
  F[i] = make_double2(-sum[6], sum[0]);
  F[V + i] = make_double2(-sum[3], sum[2]);
  F[2 * V + i] = make_double2(-sum[5], sum[4]);
  F[3 * V + i] = make_double2(-sum[9], sum[8]);
}

// trace-metric sums per slice (Gfix/CFACG.c:166-260)
__global__ __launch_bounds__(kBlock) void k_cpr(const double2 *__restrict__ F, const double2 *__restrict__ O,
                                               double *__restrict__ part, Geo g, SliceCtrl sc, long long nbs) {
  int t;
  long long bx;
  slice_block(g, sc, t, bx);
  if (sc.done[t]) return;
  const long long xl = bx * blockDim.x + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (xl < g.Lcu) {
    const long long i = site_of(g, t, xl), V = g.V;
    const double2 z0 = F[i], o0 = O[i];
    const double tr = z0.x - o0.x, ti = z0.y - o0.y;
    double s1 = 2.0 * (z0.x * z0.x + z0.y * z0.y + z0.x * z0.y);
    double s2 = 2.0 * (z0.x * tr + z0.y * ti) + z0.x * ti + z0.y * tr;
    for (int f = 1; f < 4; f++) {
      const double2 z = F[f * V + i], o = O[f * V + i];
      s1 += 2.0 * (z.x * z.x + z.y * z.y);
      s2 += 2.0 * (z.x * (z.x - o.x) + z.y * (z.y - o.y));
    }
    v[0] = s1;
    v[1] = s2;
  }
  block_sum_store<2>(v, part + 2 * (t * nbs + bx));
}

__device__ double block_reduce_1024(double v, double *sh) {
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  v = wave_sum(v);
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int j = 0; j < (int)(blockDim.x / kWave); j++) s += sh[j];
  __syncthreads();
  return s;
}

// per-slice finaliser of the PR sums.  mode 0: SD start (inold, theta);
// mode 1: loop iteration (theta, beta, inold) -- Gfix/CFACG.c:311-339
__global__ void k_cfin_pr(const double *__restrict__ part, long long nbs, SliceCtrl sc, int mode) {
  const int t = blockIdx.x;
  if (sc.done[t]) return;
  __shared__ double sh[16];
  double a = 0.0, b = 0.0;
  for (long long k = threadIdx.x; k < nbs; k += blockDim.x) {
    a += part[2 * (t * nbs + k)];
    b += part[2 * (t * nbs + k) + 1];
  }
  const double s1 = block_reduce_1024(a, sh), s2 = block_reduce_1024(b, sh);
  if (threadIdx.x == 0) {
    sc.theta[t] = s1 * sc.norm;
    if (mode == 0) {
      sc.inold[t] = s1;
      sc.beta[t] = 0.0;
    } else {
      double beta = s2 / sc.inold[t];
      if (beta < 0.0) beta = 0.0;
      if (sc.theta[t] < sc.cg_tol) beta = s1 / sc.inold[t];
      sc.inold[t] = s1;
      sc.insum[t] = s1;
      sc.sconj[t] = s2;
      sc.beta[t] = beta;
    }
  }
}

// sn = in + beta sn (mode 1, Gfix/CFACG.c:359-365) or sn = in (mode 0, the
// SD start, :323).  GLU's copy old = in is not made: the host swaps its in
// and in_old pointers after every CG iteration, so the next derivative
// lands in the old buffer and this iteration's in becomes in_old in place.
__global__ __launch_bounds__(kBlock) void k_cupdate(const double2 *__restrict__ F, double2 *__restrict__ sn, Geo g,
                                                   SliceCtrl sc, int mode) {
  int t;
  long long bx;
  slice_block(g, sc, t, bx);
  if (sc.done[t]) return;
  const long long xl = bx * blockDim.x + threadIdx.x;
  if (xl >= g.Lcu) return;
  const long long i = site_of(g, t, xl), V = g.V;
  const double beta = mode ? sc.beta[t] : 0.0;
  for (int f = 0; f < 4; f++) {
    const double2 a = F[f * V + i];
    if (mode) {
      const double2 b = sn[f * V + i];
      sn[f * V + i] = make_double2(a.x + beta * b.x, a.y + beta * b.y);
    } else {
      sn[f * V + i] = a;
    }
  }
}

__device__ __forceinline__ su3<double> cgauge_at(const double2 *__restrict__ F, long long V, long long i, double al) {
  auto f = [&](long long k) { return mk<double>(F[k].x, F[k].y); };
  return gauge_matrix<double>(f(i), f(V + i), f(2 * V + i), f(3 * V + i), al);
}

// probes at 0, PC1, PC2 over every spatial link of the slice once (GLU's
// two-colour form visits colour-0 sites with forward and backward links,
// Gfix/line_search.c:194-271; the sum is the same)
__global__ __launch_bounds__(kBlock) void k_cline(const double2 *__restrict__ U, const double2 *__restrict__ G,
                                                 const double2 *__restrict__ F, double *__restrict__ part, Geo g,
                                                 SliceCtrl sc, long long nbs, int force) {
  int t;
  long long bx;
  slice_block(g, sc, t, bx);
  if (sc.done[t]) return;
  if (!force && !(sc.theta[t] > sc.cg_tol)) return;
  const long long xl = bx * blockDim.x + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (xl < g.Lcu) {
    const long long i = site_of(g, t, xl), V = g.V;
    int x[4];
    coords(g, i, x);
    const su3<double> Gi = load_link<double>(G, 0, i, V);
    const su3<double> A1 = cgauge_at(F, V, i, PC1), A2 = cgauge_at(F, V, i, PC2);
    // one direction at a time, rows 0, 1 of A C formed before each probe's
    // neighbour matrix is built: the same arithmetic as retrace_abc_dag
    // with fewer registers held
#pragma unroll 1
    for (int mu = 0; mu < 3; mu++) {
      const long long fw = nb_fwd(g, i, x, mu);
      cx<double> ac1[6], ac2[6];
      {
        const su3<double> C = gtransform<double>(Gi, load_link<double>(U, mu, i, V), load_link<double>(G, 0, fw, V));
        v[0] += C.e[0].re + C.e[4].re + C.e[8].re;
        ab_rows01<double>(A1, C, ac1);
        ab_rows01<double>(A2, C, ac2);
      }
      v[1] += retrace_rows01_dag<double>(ac1, cgauge_at(F, V, fw, PC1));
      v[2] += retrace_rows01_dag<double>(ac2, cgauge_at(F, V, fw, PC2));
    }
  }
  block_sum_store<3>(v, part + 3 * (t * nbs + bx));
}

__device__ __forceinline__ double approx_min(const double al[3], const double f[3]) {
  const double A = f[0] / ((al[0] - al[1]) * (al[0] - al[2]));
  const double B = f[1] / ((al[1] - al[0]) * (al[1] - al[2]));
  const double C = f[2] / ((al[2] - al[0]) * (al[2] - al[1]));
  const double sol = (al[0] * (B + C) + al[1] * (A + C) + al[2] * (A + B)) / (2 * (A + B + C));
  if (sol > al[2]) return al[2];
  if (isnan(sol) || isinf(sol) || sol < al[0]) return 0.08;
  return sol;
}

// step of each slice: parabola vertex (Gfix/line_search.c:118-138) or
// GF_TUNE below CG_TOL / for FASD
__global__ void k_cfin_ls(const double *__restrict__ part, long long nbs, SliceCtrl sc, int force, int fixed) {
  const int t = blockIdx.x;
  if (sc.done[t]) return;
  const int do_ls = !fixed && (force || (sc.theta[t] > sc.cg_tol));
  __shared__ double sh[16];
  double s[3] = {0, 0, 0};
  if (do_ls) {
    double a[3] = {0, 0, 0};
    for (long long k = threadIdx.x; k < nbs; k += blockDim.x)
      for (int j = 0; j < 3; j++) a[j] += part[3 * (t * nbs + k) + j];
    for (int j = 0; j < 3; j++) s[j] = block_reduce_1024(a[j], sh);
  }
  if (threadIdx.x == 0) {
    if (do_ls) {
      const double al[3] = {0.0, PC1, PC2};
      const double val[3] = {-s[0], -s[1], -s[2]};
      for (int j = 0; j < 3; j++) sc.probe[3 * t + j] = s[j];
      sc.alpha[t] = approx_min(al, val);
      sc.did_ls[t] = 1;
    } else {
      sc.alpha[t] = sc.gf_alpha;
      sc.did_ls[t] = 0;
    }
    // FACG row: theta measured before the step, the step, beta (0 at the
    // SD start); FASD rows are written once theta is known (k_cfin_der)
    if (!fixed) slice_rec(sc, t, sc.theta[t], sc.alpha[t], force ? 0.0 : sc.beta[t]);
  }
}

// G(x) <- P(1 + alpha_t X(F(x))) G(x): Gfix/line_search.c:101-115
__global__ __launch_bounds__(kBlock) void k_cexp(double2 *__restrict__ G, const double2 *__restrict__ F, Geo g,
                                                SliceCtrl sc) {
  int t;
  long long bx;
  slice_block(g, sc, t, bx);
  if (sc.done[t]) return;
  const long long xl = bx * blockDim.x + threadIdx.x;
  if (xl >= g.Lcu) return;
  const long long i = site_of(g, t, xl), V = g.V;
  const su3<double> a = cgauge_at(F, V, i, sc.alpha[t]);
  store_link<double>(G, 0, i, V, mul_cols<double>(a, load_link<double>(G, 0, i, V)));
}

// FASD: theta from the accelerated field after the step (Gfix/CFACG.c:410-425)
__global__ void k_cfin_der(const double *__restrict__ part, long long nbs, SliceCtrl sc) {
  const int t = blockIdx.x;
  if (sc.done[t]) return;
  __shared__ double sh[16];
  double a = 0.0;
  for (long long k = threadIdx.x; k < nbs; k += blockDim.x) a += part[2 * (t * nbs + k)];
  const double s1 = block_reduce_1024(a, sh);
  if (threadIdx.x == 0) {
    sc.theta[t] = s1 * sc.norm;
    slice_rec(sc, t, sc.theta[t], sc.gf_alpha, 0.0);
  }
}

__global__ void k_citer_end(SliceCtrl sc) {
  for (int t = threadIdx.x; t < sc.T; t += blockDim.x) {
    if (sc.done[t]) continue;
    sc.iters[t]++;
    if (!(sc.theta[t] > sc.acc && sc.iters[t] < sc.max_iters)) sc.done[t] = 1;
  }
}

// reunitarise the slice transforms and record the largest |sum |g|^2 - NC|
// per block (Gfix/CFACG.c:433-457); only slices flagged in `sel`
__global__ __launch_bounds__(kBlock) void k_cadequate(double2 *__restrict__ G, const int *__restrict__ sel,
                                                     double *__restrict__ part, Geo g, long long nbs) {
  const int t = blockIdx.y;
  if (!sel[t]) return;
  const long long xl = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ double sh[kBlock];
  double dev = 0.0;
  if (xl < g.Lcu) {
    const long long i = site_of(g, t, xl), V = g.V;
    su3<double> m = load_link<double>(G, 0, i, V);
    gram_reunit<double>(m);
    store_link<double>(G, 0, i, V, m);
    double s = 0.0;
    for (int k = 0; k < 9; k++) s += cnorm2(m.e[k]);
    dev = fabs(s - 3.0);
    if (s != s) dev = 1e300;
  }
  sh[threadIdx.x] = dev;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int k = 0; k < (int)blockDim.x; k++) m = sh[k] > m ? sh[k] : m;
    part[t * nbs + blockIdx.x] = m;
  }
}


__global__ __launch_bounds__(kBlock) void k_cidentity(double2 *__restrict__ G, long long V) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  su3<double> m;
  for (int k = 0; k < 9; k++) m.e[k] = mk<double>(0, 0);
  m.e[0].re = m.e[4].re = m.e[8].re = 1.0;
  store_link<double>(G, 0, i, V, m);
}

// U_i(x,t) <- g_t(x) U_i g_t(x+i)^dag ; U_t(x,t) <- g_t(x) U_t g_{t+1}(x)^dag
// (Gfix/gtrans.c:242-264, the last slice pairs with slice 0)
__global__ __launch_bounds__(kBlock) void k_cfinal(double2 *__restrict__ U, const double2 *__restrict__ G, Geo g) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.V) return;
  int x[4];
  coords(g, i, x);
  const su3<double> A = load_link<double>(G, 0, i, g.V);
  for (int mu = 0; mu < 4; mu++) {
    const su3<double> B = load_link<double>(G, 0, nb_fwd(g, i, x, mu), g.V);
    store_link<double>(U, mu, i, g.V, gtransform<double>(A, load_link<double>(U, mu, i, g.V), B));
  }
}

unsigned nbl(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

int coulomb_alloc(glu_gfix_ctx *c) {
  CoulombState *cs = new CoulombState();
  c->cs = cs;
  const int T = c->g.L[3];
  cs->nbs = (c->g.Lcu + kBlock - 1) / kBlock;
  if (hipMalloc(&cs->G, (size_t)c->g.V * 6 * sizeof(double2)) != hipSuccess)
    return gfix_fail(-2, "slice transform allocation failed");
  const size_t nd = (11 + 3 * (size_t)kSliceTraceCap) * (size_t)T, nl = 2 * (size_t)T, ni = 3 * (size_t)T;
  if (hipMalloc(&cs->arena, nd * sizeof(double) + nl * sizeof(long long) + ni * sizeof(int) + 256) != hipSuccess)
    return gfix_fail(-2, "slice control allocation failed");
  double *d = (double *)cs->arena;
  SliceCtrl &s = cs->sc;
  s.theta = d;
  s.inold = d + T;
  s.insum = d + 2 * T;
  s.sconj = d + 3 * T;
  s.beta = d + 4 * T;
  s.alpha = d + 5 * T;
  s.probe = d + 6 * T;  // 3T
  s.dev = d + 9 * T;    // 2T spare
  s.trace = d + 11 * T;  // 3 kSliceTraceCap T
  s.iters = (long long *)(d + nd);
  s.trace_n = s.iters + T;
  s.done = (int *)(s.iters + nl);
  s.did_ls = s.done + T;
  s.map = s.did_ls + T;
  s.T = T;
  cs->slice_iters.assign(T, 0);
  cs->slice_theta.assign(T, 0.0);
  return 0;
}

void coulomb_free(glu_gfix_ctx *c) {
  if (!c->cs) return;
  if (c->cs->G) (void)hipFree(c->cs->G);  // teardown: nothing to do on failure
  if (c->cs->arena) (void)hipFree(c->cs->arena);
  delete c->cs;
  c->cs = nullptr;
}

extern "C" int glu_gfix_slice_trace(glu_gfix_ctx *c, size_t slice, double *rows, size_t cap, size_t *n) {
  if (!c || !rows || !n) return gfix_fail(-1, "null argument");
  if (!c->cs) return gfix_fail(-1, "not a Coulomb context");
  if (slice >= (size_t)c->g.L[3]) return gfix_fail(-1, "slice %zu beyond the %d time slices", slice, c->g.L[3]);
  long long m = 0;
  std::vector<double> h(3 * kSliceTraceCap);
  const SliceCtrl &sc = c->cs->sc;
  if (hipMemcpyAsync(&m, sc.trace_n + slice, sizeof(long long), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipMemcpyAsync(h.data(), sc.trace + slice * 3 * kSliceTraceCap, h.size() * sizeof(double),
                     hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return gfix_fail(-2, "slice trace read failed");
  const size_t k = std::min((size_t)m, cap);
  std::copy(h.begin(), h.begin() + 3 * k, rows);
  *n = k;
  return 0;
}

extern "C" int glu_gfix_slice_info(glu_gfix_ctx *c, size_t *iters, double *theta) {
  if (!c || !iters || !theta) return gfix_fail(-1, "null argument");
  if (!c->cs) return gfix_fail(-1, "not a Coulomb context");
  for (int t = 0; t < c->g.L[3]; t++) {
    iters[t] = c->cs->slice_iters[t];
    theta[t] = c->cs->slice_theta[t];
  }
  return 0;
}

// the batched per-slice loop until every selected slice stops
static int run_slices(glu_gfix_ctx *c, int method, const std::vector<int> &active) {
  CoulombState *cs = c->cs;
  SliceCtrl sc = cs->sc;
  const int T = c->g.L[3];
  hipStream_t s = c->stream;
  Geo g = c->g;
  g.xcd = c->ord_cl ? 1 : 0;
  std::vector<int> done(T), rows;
  std::vector<long long> zero(T, 0);
  for (int t = 0; t < T; t++) done[t] = active[t] ? 0 : 1;
  // the launches of a batch get blocks for the slices running at its start
  // only (rows -> slices through sc.map); slices that stop inside a batch
  // return at once until the next one drops them
  auto compact = [&]() {
    rows.clear();
    for (int t = 0; t < T; t++)
      if (!done[t]) rows.push_back(t);
    if (rows.empty()) return 0;
    if (hipMemcpyAsync(sc.map, rows.data(), rows.size() * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return gfix_fail(-2, "Coulomb slice map upload failed");
    return 0;
  };
  if (hipMemcpyAsync(sc.done, done.data(), T * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(sc.iters, zero.data(), T * sizeof(long long), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return gfix_fail(-2, "Coulomb slice state upload failed");
  if (int rc = compact()) return rc;
  if (rows.empty()) return 0;
  dim3 grid((unsigned)cs->nbs, (unsigned)rows.size());
  const dim3 blk(kBlock);
  double2 *U = (double2 *)c->U64, *G = (double2 *)cs->G, *F = (double2 *)c->F, *sn = (double2 *)c->sn,
          *O = (double2 *)c->old;
  double *part = c->part;
  if (method == GLU_GFIX_CG) {
    // SD start: derivative, FA, theta and inold, line search, sn = in; the
    // pointer swap makes in the in_old of the first CG iteration.  The CG
    // state lives within this call (every call starts with an SD step), so
    // a slice that stopped may keep its fields in either buffer.
    GFIX_PROF(c, PK_CDERIV, hipLaunchKernelGGL(k_cderiv, grid, blk, 0, s, U, G, F, g, sc));
    enq_fa<double>(c, F, nullptr, sc.done, sc.map, (int)rows.size());
    GFIX_PROF(c, PK_CPR, hipLaunchKernelGGL(k_cpr, grid, blk, 0, s, F, F, part, g, sc, cs->nbs));
    hipLaunchKernelGGL(k_cfin_pr, dim3(T), dim3(1024), 0, s, part, cs->nbs, sc, 0);
    GFIX_PROF(c, PK_CLINE, hipLaunchKernelGGL(k_cline, grid, blk, 0, s, U, G, F, part, g, sc, cs->nbs, 1));
    hipLaunchKernelGGL(k_cfin_ls, dim3(T), dim3(1024), 0, s, part, cs->nbs, sc, 1, 0);
    GFIX_PROF(c, PK_CEXP, hipLaunchKernelGGL(k_cexp, grid, blk, 0, s, G, F, g, sc));
    GFIX_PROF(c, PK_CUPD, hipLaunchKernelGGL(k_cupdate, grid, blk, 0, s, F, sn, g, sc, 0));
    std::swap(F, O);
  }
  for (;;) {
    for (int b = 0; b < 8; b++) {
      GFIX_PROF(c, PK_CDERIV, hipLaunchKernelGGL(k_cderiv, grid, blk, 0, s, U, G, F, g, sc));
      enq_fa<double>(c, F, nullptr, sc.done, sc.map, (int)rows.size());
      if (method == GLU_GFIX_CG) {
        GFIX_PROF(c, PK_CPR, hipLaunchKernelGGL(k_cpr, grid, blk, 0, s, F, O, part, g, sc, cs->nbs));
        hipLaunchKernelGGL(k_cfin_pr, dim3(T), dim3(1024), 0, s, part, cs->nbs, sc, 1);
        GFIX_PROF(c, PK_CUPD, hipLaunchKernelGGL(k_cupdate, grid, blk, 0, s, F, sn, g, sc, 1));
        GFIX_PROF(c, PK_CLINE, hipLaunchKernelGGL(k_cline, grid, blk, 0, s, U, G, sn, part, g, sc, cs->nbs, 0));
        hipLaunchKernelGGL(k_cfin_ls, dim3(T), dim3(1024), 0, s, part, cs->nbs, sc, 0, 0);
        GFIX_PROF(c, PK_CEXP, hipLaunchKernelGGL(k_cexp, grid, blk, 0, s, G, sn, g, sc));
        std::swap(F, O);
      } else {
        hipLaunchKernelGGL(k_cfin_ls, dim3(T), dim3(1024), 0, s, part, cs->nbs, sc, 0, 1);
        GFIX_PROF(c, PK_CEXP, hipLaunchKernelGGL(k_cexp, grid, blk, 0, s, G, F, g, sc));
        GFIX_PROF(c, PK_CPR, hipLaunchKernelGGL(k_cpr, grid, blk, 0, s, F, F, part, g, sc, cs->nbs));
        hipLaunchKernelGGL(k_cfin_der, dim3(T), dim3(1024), 0, s, part, cs->nbs, sc);
      }
      hipLaunchKernelGGL(k_citer_end, dim3(1), dim3(256), 0, s, sc);
    }
    if (hipMemcpyAsync(done.data(), sc.done, T * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return gfix_fail(-2, "Coulomb iteration failed: %s", hipGetErrorString(hipGetLastError()));
    if (c->prof) prof_harvest(c);
    bool all = true;
    for (int t = 0; t < T; t++) all = all && done[t];
    if (all) break;
    if (int rc = compact()) return rc;
    grid.y = (unsigned)rows.size();
  }
  return 0;
}

int coulomb_fix(glu_gfix_ctx *c, int method, double gf, double acc, size_t max_iters, double *theta, size_t *iters) {
  CoulombState *cs = c->cs;
  const int T = c->g.L[3];
  hipStream_t s = c->stream;
  SliceCtrl &sc = cs->sc;
  sc.acc = acc;
  sc.gf_alpha = gf;
  sc.cg_tol = kCgTol;
  sc.norm = 1.0 / (3.0 * (double)c->g.Lcu);  // GFNORM_COULOMB
  sc.max_iters = (long long)max_iters;
  hipLaunchKernelGGL(k_cidentity, dim3(nbl(c->g.V)), dim3(kBlock), 0, s, (double2 *)cs->G, c->g.V);
  {
    const std::vector<long long> zero(T, 0);
    if (hipMemcpyAsync(sc.trace_n, zero.data(), T * sizeof(long long), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return gfix_fail(-2, "Coulomb trace reset failed");
  }
  std::vector<int> active(T, 1), control(T, 0), sel(T, 0);
  std::vector<size_t> tot(T, 0);
  std::vector<double> th(T);
  std::vector<long long> it(T);
  int *dsel = sc.did_ls;  // reused as a device selection mask between rounds
  // slices waiting for their random restart: GLU fixes the slices one
  // after another, so the restarts of slice t draw from the streams after
  // every draw of the slices before it; a slice that needs one waits until
  // all slices before it have finished (Gfix/CFACG.c:505-521, 593-614)
  std::vector<int> pending(T, 0);
  int rc_all = 0;
  for (int round = 0; round < 4096 && !rc_all; round++) {
    bool any = false, anyp = false;
    for (int t = 0; t < T; t++) {
      any = any || active[t];
      anyp = anyp || pending[t];
    }
    if (!any && !anyp) break;
    if (any)
      if (int rc = run_slices(c, method, active)) return rc;
    // every copy below is ordered on the context stream (a non-blocking
    // stream does not wait for the null stream, nor the null stream for it)
    if (hipMemcpyAsync(th.data(), sc.theta, T * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(it.data(), sc.iters, T * sizeof(long long), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return gfix_fail(-2, "Coulomb slice results read failed");
    // unitarity test of the converged slices (reunitarises them)
    std::vector<double> devs(T, 0.0);
    for (int t = 0; t < T; t++) sel[t] = active[t] && th[t] < acc;
    if (!any) std::fill(sel.begin(), sel.end(), 0);
    if (hipMemcpyAsync(dsel, sel.data(), T * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess)
      return gfix_fail(-2, "Coulomb selection upload failed");
    hipLaunchKernelGGL(k_cadequate, dim3((unsigned)cs->nbs, T), dim3(kBlock), 0, s, (double2 *)cs->G, dsel, c->part,
                       c->g, cs->nbs);
    std::vector<double> pm((size_t)cs->nbs * T);
    if (hipMemcpyAsync(pm.data(), c->part, pm.size() * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return gfix_fail(-2, "unitarity check failed");
    for (int t = 0; t < T; t++)
      if (sel[t])
        for (long long b = 0; b < cs->nbs; b++) devs[t] = std::max(devs[t], pm[t * cs->nbs + b]);
    for (int t = 0; t < T; t++) {
      if (!active[t]) continue;
      const size_t loc = (size_t)it[t];
      bool trans = false;
      if (th[t] < acc && devs[t] > (double)(loc > 0 ? loc : 1) * 3e-14) trans = true;
      if (loc >= max_iters) {
        if (th[t] < 1e4 * acc && std::isfinite(th[t])) {  // continuation
          tot[t] += loc;
          continue;
        }
        if (th[t] > acc || !std::isfinite(th[t])) trans = true;
      }
      if (control[t] < 8 && trans) {
        if (c->verbose) fprintf(stdout, "[CG] random gauge transform %d :: %e (%zu)\n", control[t], th[t], loc);
        pending[t] = 1;
        active[t] = 0;
        tot[t] += loc;
        control[t]++;
        continue;
      }
      tot[t] += loc;
      active[t] = 0;
      cs->slice_theta[t] = th[t];
    }
    // draw, in slice order, for the waiting slices whose predecessors have
    // all finished; Sunitary_gen sets the slice transform itself
    std::vector<long long> first;
    bool blocked = false;
    for (int t = 0; t < T && !blocked; t++) {
      if (active[t]) blocked = true;
      if (!pending[t]) continue;
      if (blocked) break;
      first.push_back((long long)t * c->g.Lcu);
      pending[t] = 0;
      active[t] = 1;
      blocked = true;  // its own new attempt must end before later slices draw
    }
    if (!first.empty()) rc_all = mwc_sunitary(c, cs->G, c->g.V, c->g.Lcu, first, true);
  }
  c->mwc_live = 0;  // Coulomb_FA frees GLU's streams on return (CFACG.c:630)
  if (rc_all) return rc_all;
  hipLaunchKernelGGL(k_cfinal, dim3(nbl(c->g.V)), dim3(kBlock), 0, s, (double2 *)c->U64, (const double2 *)cs->G,
                     c->g);
  if (hipStreamSynchronize(s) != hipSuccess)
    return gfix_fail(-2, "Coulomb final transform failed: %s", hipGetErrorString(hipGetLastError()));
  size_t sum = 0;
  double mx = 0.0;
  for (int t = 0; t < T; t++) {
    cs->slice_iters[t] = tot[t];
    sum += tot[t];
    mx = std::max(mx, cs->slice_theta[t]);
  }
  *iters = sum;
  *theta = mx;
  return 0;
}
