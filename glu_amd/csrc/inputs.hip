// inputs.hip -- device-side input generation for the gauge-fixing runs:
// a quenched Wilson-action heatbath (Cabibbo-Marinari over the three SU(2)
// subgroups with Kennedy-Pendleton sampling, plus microcanonical
// overrelaxation) and a seeded Haar-like random gauge transform.
//
// These play the role of GLU's own input generators (Update/KPHB.c:124-257,
// Geometry/random_config.c:61-95) for lattices too large to generate on the
// host; they are not on the gauge-fixing path and are not bit-compatible
// with GLU's MWC_4096 streams.  Random numbers are counter-based
// (splitmix64 keyed on seed, sweep, pass, direction and site), so a config
// depends only on its seed.
#include "../../include/glu_gfix.h"
#include "gfix_ctx.h"

namespace gfix {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
struct Rng {
  uint64_t s;
  __device__ __forceinline__ double u() {  // uniform in (0, 1]
    s += 0x9E3779B97F4A7C15ULL;
    return ((double)(mix64(s) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
  }
};

// Gram-Schmidt with plain divisions: the generator keeps its own arithmetic
// so that its output does not move when the solver's reunitarisation is
// tuned (the heatbath is chaotic: one ulp anywhere changes the config)
__device__ void hb_reunit(su3<double> &m) {
  double n = cnorm2(m.e[0]) + cnorm2(m.e[3]) + cnorm2(m.e[6]);
  n = 1.0 / sqrt(n);
  m.e[0] = cscale(m.e[0], n);
  m.e[3] = cscale(m.e[3], n);
  m.e[6] = cscale(m.e[6], n);
  const cx<double> p = cadd(cadd(cmulc(m.e[1], m.e[0]), cmulc(m.e[4], m.e[3])), cmulc(m.e[7], m.e[6]));
  m.e[1] = csub(m.e[1], cmul(p, m.e[0]));
  m.e[4] = csub(m.e[4], cmul(p, m.e[3]));
  m.e[7] = csub(m.e[7], cmul(p, m.e[6]));
  double q = cnorm2(m.e[1]) + cnorm2(m.e[4]) + cnorm2(m.e[7]);
  q = 1.0 / sqrt(q);
  m.e[1] = cscale(m.e[1], q);
  m.e[4] = cscale(m.e[4], q);
  m.e[7] = cscale(m.e[7], q);
  complete_col2(m);
}

// staple sum A with the local action Re tr(U_mu(x) A)
__device__ su3<double> staple(const double2 *__restrict__ U, const Geo &g, long long i, const int x[4], int mu) {
  su3<double> A;
#pragma unroll
  for (int k = 0; k < 9; k++) A.e[k] = mk<double>(0, 0);
  const long long V = g.V;
  const long long xp = nb_fwd(g, i, x, mu);
  int xpc[4];
  coords(g, xp, xpc);
  for (int nu = 0; nu < 4; nu++) {
    if (nu == mu) continue;
    const long long xn = nb_fwd(g, i, x, nu), xm = nb_bwd(g, i, x, nu), xpm = nb_bwd(g, xp, xpc, nu);
    const su3<double> a = load_link<double>(U, nu, xp, V), b = load_link<double>(U, mu, xn, V),
                      c = load_link<double>(U, nu, i, V);
    // U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
    su3<double> t;
    for (int r = 0; r < 3; r++)
      for (int s = 0; s < 3; s++)
        t.e[r * 3 + s] = cadd(cadd(cmulc(a.e[r * 3], b.e[s * 3]), cmulc(a.e[r * 3 + 1], b.e[s * 3 + 1])),
                              cmulc(a.e[r * 3 + 2], b.e[s * 3 + 2]));
    for (int r = 0; r < 3; r++)
      for (int s = 0; s < 3; s++)
        A.e[r * 3 + s] =
            cadd(A.e[r * 3 + s], cadd(cadd(cmulc(t.e[r * 3], c.e[s * 3]), cmulc(t.e[r * 3 + 1], c.e[s * 3 + 1])),
                                      cmulc(t.e[r * 3 + 2], c.e[s * 3 + 2])));
    // U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu)
    const su3<double> d = load_link<double>(U, nu, xpm, V), e = load_link<double>(U, mu, xm, V),
                      f = load_link<double>(U, nu, xm, V);
    for (int r = 0; r < 3; r++)
      for (int s = 0; s < 3; s++) {
        cx<double> z = mk<double>(0, 0);
        for (int k = 0; k < 3; k++) z = cadd(z, cconj(cmul(d.e[k * 3 + r], e.e[s * 3 + k])));
        t.e[r * 3 + s] = z;
      }
    for (int r = 0; r < 3; r++)
      for (int s = 0; s < 3; s++)
        A.e[r * 3 + s] = cadd(A.e[r * 3 + s], cadd(cadd(cmul(t.e[r * 3], f.e[s]), cmul(t.e[r * 3 + 1], f.e[3 + s])),
                                                   cmul(t.e[r * 3 + 2], f.e[6 + s])));
  }
  return A;
}

__global__ __launch_bounds__(kBlock) void k_heatbath(double2 *__restrict__ U, Geo g, int mu, int parity, double beta,
                                                    uint64_t seed, int sweep, int pass) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.V) return;
  int x[4];
  coords(g, i, x);
  if (((x[0] + x[1] + x[2] + x[3]) & 1) != parity) return;
  Rng r;
  r.s = mix64(seed ^ mix64(((uint64_t)sweep << 42) ^ ((uint64_t)pass << 38) ^ ((uint64_t)mu << 36) ^ (uint64_t)i));
  const su3<double> A = staple(U, g, i, x, mu);
  su3<double> u = load_link<double>(U, mu, i, g.V);
  const int SUB[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int sg = 0; sg < 3; sg++) {
    const int p = SUB[sg][0], q = SUB[sg][1];
    cx<double> w[4];
    const int idx[2] = {p, q};
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) {
        const int ra = idx[a], cb = idx[b];
        w[a * 2 + b] = cadd(cadd(cmul(u.e[ra * 3], A.e[cb]), cmul(u.e[ra * 3 + 1], A.e[3 + cb])),
                            cmul(u.e[ra * 3 + 2], A.e[6 + cb]));
      }
    const double a0 = 0.5 * (w[0].re + w[3].re), a3 = 0.5 * (w[0].im - w[3].im);
    const double a1 = 0.5 * (w[1].im + w[2].im), a2 = 0.5 * (w[1].re - w[2].re);
    const double k = sqrt(a0 * a0 + a1 * a1 + a2 * a2 + a3 * a3);
    if (k < 1e-14) continue;
    const double v0 = a0 / k, v1 = a1 / k, v2 = a2 / k, v3 = a3 / k;
    double x0, x1, x2, x3;
    if (pass == 0) {
      const double al = 2.0 * beta * k / 3.0;
      double lam2 = 0.0;
      for (int tries = 0; tries < 1000; tries++) {
        const double r1 = r.u(), r2 = r.u(), r3 = r.u(), r4 = r.u();
        const double c = cos(6.283185307179586 * r2);
        lam2 = -(log(r1) + c * c * log(r3)) / (2.0 * al);
        if (r4 * r4 <= 1.0 - lam2) break;
      }
      x0 = 1.0 - 2.0 * lam2;
      const double rad = sqrt(fmax(0.0, 1.0 - x0 * x0));
      const double cth = 2.0 * r.u() - 1.0, phi = 6.283185307179586 * r.u();
      const double sth = sqrt(fmax(0.0, 1.0 - cth * cth));
      x1 = rad * sth * cos(phi);
      x2 = rad * sth * sin(phi);
      x3 = rad * cth;
    } else {
      x0 = v0;
      x1 = -v1;
      x2 = -v2;
      x3 = -v3;
    }
    const double R0 = x0 * v0 + x1 * v1 + x2 * v2 + x3 * v3;
    const double R1 = -x0 * v1 + x1 * v0 + (x2 * v3 - x3 * v2);
    const double R2 = -x0 * v2 + x2 * v0 + (x3 * v1 - x1 * v3);
    const double R3 = -x0 * v3 + x3 * v0 + (x1 * v2 - x2 * v1);
    const cx<double> m00 = mk<double>(R0, R3), m01 = mk<double>(R2, R1), m10 = mk<double>(-R2, R1),
                     m11 = mk<double>(R0, -R3);
    for (int s = 0; s < 3; s++) {
      const cx<double> up = u.e[p * 3 + s], uq = u.e[q * 3 + s];
      u.e[p * 3 + s] = cadd(cmul(m00, up), cmul(m01, uq));
      u.e[q * 3 + s] = cadd(cmul(m10, up), cmul(m11, uq));
    }
  }
  // rows p, q were rotated; rebuild an exact SU(3) matrix from columns 0, 1
  hb_reunit(u);
  store_link<double>(U, mu, i, g.V, u);
}

__global__ __launch_bounds__(kBlock) void k_unit(double2 *__restrict__ U, long long V) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  su3<double> m;
  for (int k = 0; k < 9; k++) m.e[k] = mk<double>(0, 0);
  m.e[0].re = m.e[4].re = m.e[8].re = 1.0;
  for (int mu = 0; mu < 4; mu++) store_link<double>(U, mu, i, V, m);
}

// G(x): 9 complex Gaussians (Box-Muller), Gram-Schmidt to SU(3)
__global__ __launch_bounds__(kBlock) void k_random_gauge(double2 *__restrict__ G, long long V, uint64_t seed) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  Rng r;
  r.s = mix64(seed ^ mix64(0x5eedULL << 40 ^ (uint64_t)i));
  su3<double> m;
  for (int k = 0; k < 9; k++) {
    const double a = r.u(), b = r.u();
    const double rad = sqrt(-2.0 * log(a));
    m.e[k] = mk<double>(rad * cos(6.283185307179586 * b), rad * sin(6.283185307179586 * b));
  }
  hb_reunit(m);
  store_link<double>(G, 0, i, V, m);
}

// U_mu(x) <- G(x) U_mu(x) G(x+mu)^dagger (Gfix/gtrans.c:169-189)
__global__ __launch_bounds__(kBlock) void k_gtransform(double2 *__restrict__ U, const double2 *__restrict__ G, Geo g) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.V) return;
  int x[4];
  coords(g, i, x);
  const su3<double> A = load_link<double>(G, 0, i, g.V);
  for (int mu = 0; mu < 4; mu++) {
    const su3<double> B = load_link<double>(G, 0, nb_fwd(g, i, x, mu), g.V);
    const su3<double> u = load_link<double>(U, mu, i, g.V);
    store_link<double>(U, mu, i, g.V, gtransform<double>(A, u, B));
  }
}

}  // namespace gfix

using namespace gfix;

static unsigned nb(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

extern "C" int glu_gfix_heatbath(glu_gfix_ctx *c, double beta, int sweeps, int nor, uint64_t seed, int cold_start) {
  if (!c || sweeps < 0 || nor < 0 || !(beta > 0)) return gfix_fail(-1, "glu_gfix_heatbath: bad argument");
  if (cold_start) hipLaunchKernelGGL(k_unit, dim3(nb(c->g.V)), dim3(kBlock), 0, c->stream, (double2 *)c->U64, c->g.V);
  for (int sw = 0; sw < sweeps; sw++)
    for (int pass = 0; pass <= nor; pass++)
      for (int mu = 0; mu < 4; mu++)
        for (int par = 0; par < 2; par++)
          hipLaunchKernelGGL(k_heatbath, dim3(nb(c->g.V)), dim3(kBlock), 0, c->stream, (double2 *)c->U64, c->g, mu,
                             par, beta, seed, sw, pass);
  if (hipStreamSynchronize(c->stream) != hipSuccess)
    return gfix_fail(-2, "heatbath failed: %s", hipGetErrorString(hipGetLastError()));
  return 0;
}

extern "C" int glu_gfix_random_transform(glu_gfix_ctx *c, uint64_t seed) {
  if (!c) return gfix_fail(-1, "null context");
  void *G = nullptr;
  if (hipMallocAsync(&G, (size_t)c->g.V * 6 * sizeof(double2), c->stream) != hipSuccess)
    return gfix_fail(-2, "gauge allocation failed");
  hipLaunchKernelGGL(k_random_gauge, dim3(nb(c->g.V)), dim3(kBlock), 0, c->stream, (double2 *)G, c->g.V, seed);
  hipLaunchKernelGGL(k_gtransform, dim3(nb(c->g.V)), dim3(kBlock), 0, c->stream, (double2 *)c->U64,
                     (const double2 *)G, c->g);
  (void)hipFreeAsync(G, c->stream);  // a failure shows in the sync below
  if (hipStreamSynchronize(c->stream) != hipSuccess)
    return gfix_fail(-2, "random transform failed: %s", hipGetErrorString(hipGetLastError()));
  return 0;
}
