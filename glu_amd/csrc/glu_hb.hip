// glu_hb.hip -- GLU's own quenched update (MODE = HEATBATH,
// Update/KPHB.c:121-257) on the device links: per sweep one heatbath pass
// (Update/hb.c:156-194) and `nor` overrelaxation passes
// (Update/relax.c:76-114) over GLU's draughtboard colouring
// (Geometry/draughtboard.c:48-123), with Wilson staples
// (Smear/staples.c:63-76), the SU(2) subgroup algebra
// (Update/SU2_rotate_SSE.c:75-394) and the Kennedy-Pendleton SU(2)
// heatbath (Update/hb.c:53-127), drawing from the context's copy of GLU's
// MWC_4096 streams.
//
// GLU's heatbath pass walks, for each direction and colour, the colour's
// site list with `omp for`: thread t takes libgomp's static block of it and
// draws from its own stream in site order.  Here one lane walks one
// stream's block, so a heatbath launch has as many lanes as GLU has
// threads; every site of a colour reads only links of the other colour or
// of other directions, so the lanes are independent.  Overrelaxation draws
// nothing, and runs one thread per site.  The file is compiled without
// floating-point contraction, so every operation rounds as GLU's x86 build
// does; cos, sin and log come from the device library and may differ from
// the host's in the last bit, which over many sweeps decorrelates the chain
// from a host run (as two GLU builds on different libm would).
// (the pragma comes first so that the inline SU(3) helpers the headers
// define are compiled without contraction in this file too)
#pragma clang fp contract(off)

#include <vector>

#include "../../include/glu_gfix.h"
#include "gfix_ctx.h"

using namespace gfix;

namespace {

constexpr int kNhbMax = 42;  // NHBMAX, Update/hb.c:44
constexpr double kTwoPi = 6.283185307179586;  // TWOPI, Headers/GLU_definitions.h:139

__device__ __forceinline__ double mwc_dbl(uint32_t &x, uint32_t &carry) {
  const uint64_t p = 18782ULL * (uint64_t)x + carry;  // Utils/par_MWC_4096.c:33-46
  carry = (uint32_t)(p >> 32);
  uint32_t v = (uint32_t)p + carry;
  if (v < carry) {
    v++;
    carry++;
  }
  x = 0xfffffffeu - v;
  return (double)x * 2.3283064365386963e-10;
}

// r = a b c^dag and r = a^dag b c, sums left to right
__device__ su3<double> mul_abcdag(const su3<double> &a, const su3<double> &b, const su3<double> &c) {
  su3<double> t, r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      t.e[3 * i + j] =
          cadd(cadd(cmul(a.e[3 * i], b.e[j]), cmul(a.e[3 * i + 1], b.e[3 + j])), cmul(a.e[3 * i + 2], b.e[6 + j]));
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      r.e[3 * i + j] = cadd(cadd(cmulc(t.e[3 * i], c.e[3 * j]), cmulc(t.e[3 * i + 1], c.e[3 * j + 1])),
                            cmulc(t.e[3 * i + 2], c.e[3 * j + 2]));
  return r;
}
__device__ su3<double> mul_adagbc(const su3<double> &a, const su3<double> &b, const su3<double> &c) {
  su3<double> t, r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      t.e[3 * i + j] = cadd(cadd(cmul(cconj(a.e[i]), b.e[j]), cmul(cconj(a.e[3 + i]), b.e[3 + j])),
                            cmul(cconj(a.e[6 + i]), b.e[6 + j]));
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      r.e[3 * i + j] =
          cadd(cadd(cmul(t.e[3 * i], c.e[j]), cmul(t.e[3 * i + 1], c.e[3 + j])), cmul(t.e[3 * i + 2], c.e[6 + j]));
  return r;
}

// Smear/staples.c:63-76: upper U_nu(x) U_mu(x+nu) U_nu(x+mu)^dag and lower
// U_nu(x-nu)^dag U_mu(x-nu) U_nu(x-nu+mu) over nu != mu
__device__ su3<double> glu_staple(const double2 *__restrict__ U, const Geo &g, long long i, const int x[4], int mu) {
  const long long V = g.V;
  su3<double> s;
  for (int k = 0; k < 9; k++) s.e[k] = mk<double>(0.0, 0.0);
  const long long t2 = nb_fwd(g, i, x, mu);
  for (int j = 0; j < 3; j++) {
    const int nu = j < mu ? j : j + 1;
    const long long t1 = nb_fwd(g, i, x, nu), b1 = nb_bwd(g, i, x, nu);
    int xb[4];
    coords(g, b1, xb);
    const long long b2 = nb_fwd(g, b1, xb, mu);
    su3<double> a =
        mul_abcdag(load_link<double>(U, nu, i, V), load_link<double>(U, mu, t1, V), load_link<double>(U, nu, t2, V));
    for (int k = 0; k < 9; k++) s.e[k] = cadd(s.e[k], a.e[k]);
    a = mul_adagbc(load_link<double>(U, nu, b1, V), load_link<double>(U, mu, b1, V), load_link<double>(U, nu, b2, V));
    for (int k = 0; k < 9; k++) s.e[k] = cadd(s.e[k], a.e[k]);
  }
  return s;
}

// rows (a, b) of subgroup sg: (0, 1), (1, 2), (2, 0)
__device__ __forceinline__ int sg_a(int sg) { return sg; }
__device__ __forceinline__ int sg_b(int sg) { return sg == 2 ? 0 : sg + 1; }

// Update/SU2_rotate_SSE.c:75-200: s0, s1 of the subgroup of U S^dag,
// normalised, each three-term sum nested as the intrinsics nest it
__device__ void only_subgroup(cx<double> &s0, cx<double> &s1, double &scale, const su3<double> &U,
                              const su3<double> &S, int sg) {
  const int a = 3 * sg_a(sg), b = 3 * sg_b(sg);
  auto sum3 = [](cx<double> e0, cx<double> e1, cx<double> e2) { return cadd(e0, cadd(e1, e2)); };
  const cx<double> t0 = sum3(cmulc(U.e[a], S.e[a]), cmulc(U.e[a + 1], S.e[a + 1]), cmulc(U.e[a + 2], S.e[a + 2]));
  const cx<double> t3 = sum3(cmul(cconj(U.e[b]), S.e[b]), cmul(cconj(U.e[b + 1]), S.e[b + 1]),
                             cmul(cconj(U.e[b + 2]), S.e[b + 2]));
  const cx<double> t1 = sum3(cmulc(U.e[a], S.e[b]), cmulc(U.e[a + 1], S.e[b + 1]), cmulc(U.e[a + 2], S.e[b + 2]));
  const cx<double> t2 = sum3(cmul(cconj(U.e[b]), S.e[a]), cmul(cconj(U.e[b + 1]), S.e[a + 1]),
                             cmul(cconj(U.e[b + 2]), S.e[a + 2]));
  const cx<double> m0 = cadd(t0, t3), m1 = csub(t1, t2);
  const double z = 1.0 / sqrt((m0.re * m0.re + m1.re * m1.re) + (m0.im * m0.im + m1.im * m1.im));
  s0 = mk<double>(m0.re * z, m0.im * z);
  s1 = mk<double>(m1.re * z, m1.im * z);
  scale = z;
}

// U <- s U on the rows (a, b) of subgroup sg (SU2_rotate_SSE.c:287-394)
__device__ void su2_rotate(su3<double> &U, cx<double> s0, cx<double> s1, int sg) {
  const int a = 3 * sg_a(sg), b = 3 * sg_b(sg);
  for (int k = 0; k < 3; k++) {
    const cx<double> ua = U.e[a + k], ub = U.e[b + k];
    U.e[a + k] = cadd(cmul(s0, ua), cmul(s1, ub));
    U.e[b + k] = csub(cmul(cconj(s0), ub), cmul(cconj(s1), ua));
  }
}

// Kennedy-Pendleton trial (Update/hb.c:53-62), the draws taken in order
__device__ __forceinline__ int kp(double &d, double norm, uint32_t &x, uint32_t &c) {
  const double x4 = mwc_dbl(x, c);
  const double x3 = cos(kTwoPi * mwc_dbl(x, c));
  const double l1 = log(mwc_dbl(x, c));
  const double l2 = log(mwc_dbl(x, c));
  d = -(l1 + l2 * x3 * x3) * norm;
  return (1. - 0.5 * d) < x4 * x4;
}

// generate_SU2 (Update/hb.c:81-127): s <- (heatbath SU(2) element) s
__device__ void generate_su2(cx<double> &s0, cx<double> &s1, double norm, uint32_t &x, uint32_t &c) {
  double d = 0.0;
  int iters = 1;
  if (kp(d, norm, x, c))
    while (kp(d, norm, x, c) && iters < kNhbMax) iters++;
  const double a0 = 1.0 - d;
  const double ar2 = fabs(1.0 - a0 * a0);
  const double a3 = sqrt(ar2) * (2.0 * mwc_dbl(x, c) - 1.0);
  const double rho = sqrt(fabs(ar2 - a3 * a3));
  const double xr2 = kTwoPi * mwc_dbl(x, c);
  const double a1 = cos(xr2) * rho, a2 = sin(xr2) * rho;
  const double rS0 = s0.re, iS0 = s0.im, rS1 = s1.re, iS1 = s1.im;
  s0 = mk<double>(a0 * rS0 + a3 * iS0 + a2 * rS1 + a1 * iS1, -a0 * iS0 + rS0 * a3 - a2 * iS1 + rS1 * a1);
  s1 = mk<double>(-a0 * rS1 + a3 * iS1 + a2 * rS0 - a1 * iS0, -a0 * iS1 - a3 * rS1 + a2 * iS0 + a1 * rS0);
}

// Gram-Schmidt with plain divisions (Matrix_Ops/gramschmidt.c:73-165)
__device__ void glu_reunit(su3<double> &m) {
  double n = 0.0;
  n += m.e[0].re * m.e[0].re + m.e[0].im * m.e[0].im;
  n += m.e[3].re * m.e[3].re + m.e[3].im * m.e[3].im;
  n += m.e[6].re * m.e[6].re + m.e[6].im * m.e[6].im;
  n = 1.0 / sqrt(n);
  const cx<double> c0 = cscale(m.e[0], n), c1 = cscale(m.e[3], n), c2 = cscale(m.e[6], n);
  const cx<double> p = cadd(cadd(cmul(cconj(c0), m.e[1]), cmul(cconj(c1), m.e[4])), cmul(cconj(c2), m.e[7]));
  cx<double> d0 = csub(m.e[1], cmul(p, c0)), d1 = csub(m.e[4], cmul(p, c1)), d2 = csub(m.e[7], cmul(p, c2));
  double q = d0.re * d0.re + d0.im * d0.im + d1.re * d1.re + d1.im * d1.im + d2.re * d2.re + d2.im * d2.im;
  q = 1.0 / sqrt(q);
  d0 = cscale(d0, q);
  d1 = cscale(d1, q);
  d2 = cscale(d2, q);
  m.e[0] = c0;
  m.e[3] = c1;
  m.e[6] = c2;
  m.e[1] = d0;
  m.e[4] = d1;
  m.e[7] = d2;
  complete_col2(m);
}

// heatbath of direction mu over one colour: lane t walks stream t's static
// block of the colour's n sites (Update/hb.c:174-188)
__global__ __launch_bounds__(64) void k_glu_hb(double2 *__restrict__ U, Geo g, const int *__restrict__ sites,
                                               long long n, int mu, uint32_t *__restrict__ state, int nth,
                                               double invbeta) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nth) return;
  uint32_t x = state[t], c = state[nth + t];
  long long q = n / nth, r = n % nth;
  if (t < r) {
    q++;
    r = 0;
  }
  const long long lo = q * t + r, hi = lo + q;
  for (long long k = lo; k < hi; k++) {
    const long long it = sites[k];
    int xc[4];
    coords(g, it, xc);
    const su3<double> S = glu_staple(U, g, it, xc, mu);
    su3<double> u = load_link<double>(U, mu, it, g.V);
    for (int sg = 0; sg < 3; sg++) {
      cx<double> s0, s1;
      double scale;
      only_subgroup(s0, s1, scale, u, S, sg);
      generate_su2(s0, s1, invbeta * scale, x, c);
      su2_rotate(u, s0, s1, sg);
    }
    glu_reunit(u);
    store_link<double>(U, mu, it, g.V, u);
  }
  state[t] = x;
  state[nth + t] = c;
}

// overrelaxation of direction mu over one colour (Update/relax.c:47-69,
// microcanonical reflection :33-44), one thread per site
__global__ __launch_bounds__(kBlock) void k_glu_or(double2 *__restrict__ U, Geo g, const int *__restrict__ sites,
                                                   long long n, int mu) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const long long it = sites[k];
  int xc[4];
  coords(g, it, xc);
  const su3<double> S = glu_staple(U, g, it, xc, mu);
  su3<double> u = load_link<double>(U, mu, it, g.V);
  for (int sg = 0; sg < 3; sg++) {
    cx<double> s0, s1;
    double scale;
    only_subgroup(s0, s1, scale, u, S, sg);
    const double z = s0.re * s0.re - s0.im * s0.im - s1.re * s1.re - s1.im * s1.im;
    s1 = cscale(s1, -2. * s0.re);
    s0 = mk<double>(z, -(2. * s0.re * s0.im));
    su2_rotate(u, s0, s1, sg);
  }
  store_link<double>(U, mu, it, g.V, u);
}

__global__ __launch_bounds__(kBlock) void k_glu_cold(double2 *__restrict__ U, long long V) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  su3<double> m;
  for (int k = 0; k < 9; k++) m.e[k] = mk<double>(0.0, 0.0);
  m.e[0].re = m.e[4].re = m.e[8].re = 1.0;
  for (int mu = 0; mu < 4; mu++) store_link<double>(U, mu, i, V, m);
}

// draughtboard colour of site i: Geometry/draughtboard.c:48-68
int glu_colour(const size_t dims[4], size_t i) {
  size_t even = 0, odd = 0, sub = 1;
  for (int mu = 0; mu < 4; mu++) {
    const size_t xm = (i / sub) % dims[mu];
    sub *= dims[mu];
    if ((dims[mu] & 1) == 0)
      even += xm;
    else
      odd += xm;
  }
  return (int)((even % 2 + odd) % 3);
}

}  // namespace

extern "C" int glu_gfix_glu_heatbath(glu_gfix_ctx *c, double beta, int sweeps, int nor, int cold_start) {
  if (!c) return gfix_fail(-1, "null context");
  if (!(beta > 0.0) || sweeps < 0 || nor < 0) return gfix_fail(-1, "glu_gfix_glu_heatbath: bad beta, sweeps or nor");
  if (int rc = mwc_ensure(c)) return rc;
  const long long V = c->g.V;
  // the colour lists, each in increasing site order (draughtboard.c:86-123)
  std::vector<int> list;
  list.reserve(V);
  long long off[4] = {0, 0, 0, 0};
  for (int col = 0; col < 3; col++) {
    off[col] = (long long)list.size();
    for (long long i = 0; i < V; i++)
      if (glu_colour(c->dims, (size_t)i) == col) list.push_back((int)i);
  }
  off[3] = V;
  const int ncol = off[3] > off[2] ? 3 : 2;
  int *dl = nullptr;
  if (hipMalloc(&dl, (size_t)V * sizeof(int)) != hipSuccess) return gfix_fail(-2, "colour list allocation failed");
  int rc = 0;
  hipStream_t s = c->stream;
  if (hipMemcpyAsync(dl, list.data(), (size_t)V * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    rc = gfix_fail(-2, "colour list upload failed");
  } else {
    double2 *U = (double2 *)c->U64;
    const Geo g = c->g;
    const int nth = c->mwc_nth;
    const double invbeta = 3.0 / beta;  // NC / beta, Update/KPHB.c:173
    c->deriv_ready[0] = c->deriv_ready[1] = 0;
    if (cold_start) hipLaunchKernelGGL(k_glu_cold, dim3((unsigned)((V + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, U, V);
    for (int sw = 0; sw < sweeps && !rc; sw++) {
      for (int pass = 0; pass <= nor; pass++) {
        for (int cmu = 0; cmu < 4 * ncol; cmu++) {
          const int mu = cmu / ncol, col = cmu % ncol;
          const long long n = off[col + 1] - off[col];
          if (pass == 0)
            hipLaunchKernelGGL(k_glu_hb, dim3((unsigned)((nth + 63) / 64)), dim3(64), 0, s, U, g, dl + off[col], n,
                               mu, (uint32_t *)c->mwc, nth, invbeta);
          else
            hipLaunchKernelGGL(k_glu_or, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, U, g,
                               dl + off[col], n, mu);
        }
      }
      if (hipStreamSynchronize(s) != hipSuccess)
        rc = gfix_fail(-2, "GLU heatbath sweep %d failed: %s", sw, hipGetErrorString(hipGetLastError()));
    }
    if (!rc && hipStreamSynchronize(s) != hipSuccess)
      rc = gfix_fail(-2, "GLU heatbath failed: %s", hipGetErrorString(hipGetLastError()));
  }
  (void)hipFree(dl);  // nothing to undo if the free fails
  return rc;
}
