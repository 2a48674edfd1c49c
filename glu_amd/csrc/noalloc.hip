// noalloc.hip -- glu_gfix_measure_noalloc: the observables of
// glu_gfix_measure and glu_gfix_theta (plaquette, link traces, functional,
// theta_test_lin) of caller-owned device links, with no device allocation.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops).  The per-block partial sums live in a
// fixed file-scope device array; its address on each device is looked up
// once, with that device current, and only that address is handed to the
// kernels.  Every call on one device shares the array: calls that overlap
// on one device (two streams, two threads) race and must be serialised by
// the caller.  Nothing in the array needs clearing: each call's grid
// writes every partial row the finishing kernel reads.
#include <mutex>

#include "../../include/glu_gfix.h"
#include "gfix_ctx.h"

using namespace gfix;

namespace {

// blocks of the measuring grid (each walks the lattice grid-stride) and
// partial-sum slots per block: plaquette, link trace, temporal trace,
// tr Delta^2, largest link trace
constexpr int kNaBlocks = 1024;
constexpr int kNaSlots = 8;
constexpr int kMaxDevices = 64;

// 64 KiB per device, held until the process ends
static __device__ __align__(256) double scratch[kNaBlocks * kNaSlots];

std::mutex g_na_mu;
double *g_na_ptr[kMaxDevices] = {};

// the two stored columns of a link in GLU's layout (double[V][4][18],
// row-major 3x3, Geometry/readers.c:51-58), third column completed as for
// the library's own links
__device__ __forceinline__ su3<double> load_glu_link(const double2 *__restrict__ L, long long i, int mu) {
  const double2 *b = L + (i * 4 + mu) * 9;
  su3<double> m;
  m.e[0] = ld<double>(b + 0);
  m.e[1] = ld<double>(b + 1);
  m.e[3] = ld<double>(b + 3);
  m.e[4] = ld<double>(b + 4);
  m.e[6] = ld<double>(b + 6);
  m.e[7] = ld<double>(b + 7);
  complete_col2(m);
  return m;
}

// per-block partials of the plaquette (Field/plaqs_links.c:243-265), link
// traces (:399-448) and tr Delta^2 of the derivative (Gfix/lin_derivs.c:
// 75-193); the arithmetic per site is that of k_measure and k_deriv
__global__ __launch_bounds__(kBlock) void k_na_measure(const double2 *__restrict__ L, Geo g,
                                                      double *__restrict__ part) {
  const long long V = g.V;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  double mx = -1e300;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < V; i += (long long)gridDim.x * kBlock) {
    int x[4];
    coords(g, i, x);
    su3<double> u[4];
#pragma unroll
    for (int mu = 0; mu < 4; mu++) u[mu] = load_glu_link(L, i, mu);
    DerivAcc<double> d;
    for (int mu = 0; mu < 4; mu++) {
      const double r = retrace<double>(u[mu]);
      v[1] += r;
      if (mu == 3) v[2] += r;
      mx = r > mx ? r : mx;
      const long long t = nb_fwd(g, i, x, mu);
      for (int nu = 0; nu < mu; nu++) {
        const long long s = nb_fwd(g, i, x, nu);
        const su3<double> ab = mul_su3<double>(u[mu], load_glu_link(L, t, nu));
        const su3<double> dc = mul_su3<double>(u[nu], load_glu_link(L, s, mu));
        double t2 = 0.0;
        for (int k = 0; k < 9; k++) t2 += ab.e[k].re * dc.e[k].re + ab.e[k].im * dc.e[k].im;
        v[0] += t2;
      }
      d.add(u[mu], load_glu_link(L, nb_bwd(g, i, x, mu), mu));
    }
    v[3] += d.tr2();
  }
  __shared__ double mxs[kBlock];
  mxs[threadIdx.x] = mx;
  block_sum_store<4>(v, part + kNaSlots * blockIdx.x);
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = -1e300;
    for (int k = 0; k < kBlock; k++) m = mxs[k] > m ? mxs[k] : m;
    part[kNaSlots * blockIdx.x + 4] = m;
  }
}

// the partials summed in block order (reproducible), normalised as
// glu_gfix_measure and glu_gfix_theta do
__global__ void k_na_finish(const double *__restrict__ part, int nb, double V, double *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double p = 0, t = 0, tt = 0, th = 0, m = -1e300;
  for (int b = 0; b < nb; b++) {
    const double *r = part + kNaSlots * b;
    p += r[0];
    t += r[1];
    tt += r[2];
    th += r[3];
    m = r[4] > m ? r[4] : m;
  }
  out[0] = 2.0 * p / (3.0 * 4 * 3 * V);  // average plaquette
  out[1] = t / (4 * 3 * V);              // average link trace
  out[2] = m / 3.0;                      // largest link trace / NC
  out[3] = 1.0 - t / (4 * 3 * V);        // gauge functional
  out[4] = tt / (3 * V);                 // temporal link trace
  out[5] = 2.0 * th / (3.0 * V);         // theta_test_lin
}

// this device's copy of the scratch array, looked up on the first call
int na_scratch(double **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return gfix_fail(-2, "glu_gfix_measure_noalloc: no current HIP device");
  if (dev < 0 || dev >= kMaxDevices) return gfix_fail(-1, "glu_gfix_measure_noalloc: device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    double *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return gfix_fail(-2, "glu_gfix_measure_noalloc: scratch address on device %d: %s", dev, hipGetErrorString(e));
    g_na_ptr[dev] = ptr;
  }
  *out = g_na_ptr[dev];
  return 0;
}

}  // namespace

extern "C" int glu_gfix_measure_noalloc(const double *links_dev, const size_t dims[4], double *out_dev,
                                        void *stream) {
  if (!links_dev || !dims || !out_dev) return gfix_fail(-1, "glu_gfix_measure_noalloc: null argument");
  long double vol = 1;
  for (int mu = 0; mu < 4; mu++) {
    if (dims[mu] < 2 || dims[mu] > 4096)
      return gfix_fail(-1, "glu_gfix_measure_noalloc: lattice extent %zu outside 2..4096", dims[mu]);
    vol *= (long double)dims[mu];
  }
  if (vol >= (long double)(1ll << 31))
    return gfix_fail(-1, "glu_gfix_measure_noalloc: %.0Lf sites exceed the 2^31 - 1 supported", vol);
  double *part = nullptr;
  if (int rc = na_scratch(&part)) return rc;
  Geo g{};
  for (int mu = 0; mu < 4; mu++) g.L[mu] = (int)dims[mu];
  g.S[0] = 1;
  for (int mu = 1; mu < 4; mu++) g.S[mu] = g.S[mu - 1] * (long long)dims[mu - 1];
  g.V = g.S[3] * (long long)dims[3];
  g.Lcu = g.S[3];
  geo_set_div(g);
  g.xcd = 0;
  // every block gets at least one site, so every partial row read is written
  const int nb = (int)std::min<long long>(kNaBlocks, (g.V + kBlock - 1) / kBlock);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_na_measure, dim3(nb), dim3(kBlock), 0, s, (const double2 *)links_dev, g, part);
  hipLaunchKernelGGL(k_na_finish, dim3(1), dim3(64), 0, s, (const double *)part, nb, (double)g.V, out_dev);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return gfix_fail(-2, "glu_gfix_measure_noalloc: launch failed: %s", hipGetErrorString(e));
  return 0;
}
