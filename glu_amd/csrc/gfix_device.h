// gfix_device.h -- lattice geometry, device control block and reductions
// shared by the gauge-fixing kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "su3.h"

namespace gfix {

constexpr int kNC = 3;
constexpr int kND = 4;
constexpr int kBlock = 256;  // threads per block of the streaming kernels
constexpr int kWave = 64;

// periodic 4D lattice, x fastest (Geometry/geometry.c:28-36)
struct Geo {
  int L[4];
  long long S[4];  // strides 1, Lx, LxLy, LxLyLz
  long long V;     // < 2^31 (checked at context creation)
  long long Lcu;
  // n / L[mu] = umulhi(n, dm[mu]) >> ds[mu] for 0 <= n < 2^31
  // (round-up multiplier of Granlund and Montgomery, "Division by
  // invariant integers using multiplication", 1994, Thm 4.2 with N = 31)
  unsigned dm[4];
  int ds[4];
  int xcd;  // stencil block order: 0 plain, 1 XCD-contiguous, 2 XCD t-slabs, t fastest
  int bpt;  // 256-site blocks per time slice (order 2)
  int t8;   // time slices per XCD slab (order 2)
};

// block order 2 needs whole 256-site blocks per time slice and 8 | T
inline bool geo_slab_order_ok(const Geo &g) { return g.S[3] % 256 == 0 && g.L[3] % 8 == 0; }

inline void geo_set_div(Geo &g) {
  if (geo_slab_order_ok(g)) {
    g.bpt = (int)(g.S[3] / 256);
    g.t8 = g.L[3] / 8;
  } else {
    g.bpt = g.t8 = 1;
  }
  for (int mu = 0; mu < 4; mu++) {
    const unsigned long long d = (unsigned long long)g.L[mu];
    int l = 0;
    while ((1ull << l) < d) l++;  // ceil(log2 d), d >= 2
    g.dm[mu] = (unsigned)(((1ull << (31 + l)) + d - 1) / d);
    g.ds[mu] = l - 1;
  }
}

__device__ __forceinline__ unsigned geo_div(const Geo &g, int mu, unsigned n) {
  return __umulhi(n, g.dm[mu]) >> g.ds[mu];
}

// Logical block of this workgroup.  A dispatch hands workgroups to the 8
// XCDs round-robin, and each XCD has its own L2.  With g.xcd set, XCD k
// gets a contiguous run of logical blocks, so the neighbour sites a
// stencil reads (+-x, +-y in the block, +-z a few blocks away) were
// fetched into that XCD's L2 by a block just before.  Partial sums are
// stored by logical block, so the reduction order -- and every result --
// does not depend on the mapping.
// Order 2 gives XCD k the time slices [k T/8, (k+1) T/8) and walks them
// with t fastest, so a block's +-t neighbours are the blocks dispatched
// just before and after it on the same XCD, and its +-z neighbours are
// (T/8) Lx Ly / 256 blocks away (usable when 256 | Lx Ly Lz and 8 | T).
__device__ __forceinline__ unsigned logical_block(const Geo &g) {
  const unsigned b = blockIdx.x;
  if (!g.xcd) return b;
  const unsigned x = b & 7u, k = b >> 3;
  if (g.xcd == 2) {
    const unsigned t8 = (unsigned)g.t8, tl = k % t8, r = k / t8;
    return (x * t8 + tl) * (unsigned)g.bpt + r;
  }
  const unsigned nb = gridDim.x, q = nb >> 3, rm = nb & 7u;
  return x * q + (x < rm ? x : rm) + k;
}

__device__ __forceinline__ void coords(const Geo &g, long long i, int x[4]) {
  unsigned r = (unsigned)i;
  unsigned q = geo_div(g, 0, r);
  x[0] = (int)(r - q * (unsigned)g.L[0]);
  r = q;
  q = geo_div(g, 1, r);
  x[1] = (int)(r - q * (unsigned)g.L[1]);
  r = q;
  q = geo_div(g, 2, r);
  x[2] = (int)(r - q * (unsigned)g.L[2]);
  x[3] = (int)q;
}
__device__ __forceinline__ long long nb_bwd(const Geo &g, long long i, const int x[4], int mu) {
  return x[mu] > 0 ? i - g.S[mu] : i + (long long)(g.L[mu] - 1) * g.S[mu];
}
__device__ __forceinline__ long long nb_fwd(const Geo &g, long long i, const int x[4], int mu) {
  return x[mu] < g.L[mu] - 1 ? i + g.S[mu] : i - (long long)(g.L[mu] - 1) * g.S[mu];
}

// ------------------------------------------------------------ derivative --
// Gfix/lin_derivs.c:75-193 with the packing of Gfix/FACG.c:38-73:
// Delta(x) = sum_mu [A_mu(x - mu) - A_mu(x)] from the own link q = U_mu(x)
// and the backward link s = U_mu(x - mu), accumulated direction by
// direction, then packed into the 4 fields with the local tr Delta^2
template <typename T> struct DerivAcc {
  T R0 = 0, R1 = 0, I1 = 0, R2 = 0, I2 = 0, R3 = 0, R4 = 0, I4 = 0;
  __device__ __forceinline__ void add(const su3<T> &q, const su3<T> &s) {
    R0 += T(2) * s.e[0].im - s.e[4].im - s.e[8].im;
    R0 += T(-2) * q.e[0].im + q.e[4].im + q.e[8].im;
    R1 += s.e[1].re - s.e[3].re - q.e[1].re + q.e[3].re;
    I1 += s.e[1].im + s.e[3].im - q.e[1].im - q.e[3].im;
    R2 += s.e[2].re - s.e[6].re - q.e[2].re + q.e[6].re;
    I2 += s.e[2].im + s.e[6].im - q.e[2].im - q.e[6].im;
    R3 += T(2) * s.e[4].im - s.e[0].im - s.e[8].im;
    R3 += T(-2) * q.e[4].im + q.e[0].im + q.e[8].im;
    R4 += s.e[5].re - s.e[7].re - q.e[5].re + q.e[7].re;
    I4 += s.e[5].im + s.e[7].im - q.e[5].im - q.e[7].im;
  }
  // the two halves of add, for a caller that has the site's own links
  // (q) and its backward neighbours' (s) at different times
  __device__ __forceinline__ void add_own(const su3<T> &q) {
    R0 += T(-2) * q.e[0].im + q.e[4].im + q.e[8].im;
    R1 += q.e[3].re - q.e[1].re;
    I1 += -q.e[1].im - q.e[3].im;
    R2 += q.e[6].re - q.e[2].re;
    I2 += -q.e[2].im - q.e[6].im;
    R3 += T(-2) * q.e[4].im + q.e[0].im + q.e[8].im;
    R4 += q.e[7].re - q.e[5].re;
    I4 += -q.e[5].im - q.e[7].im;
  }
  __device__ __forceinline__ void add_bwd(const su3<T> &s) {
    R0 += T(2) * s.e[0].im - s.e[4].im - s.e[8].im;
    R1 += s.e[1].re - s.e[3].re;
    I1 += s.e[1].im + s.e[3].im;
    R2 += s.e[2].re - s.e[6].re;
    I2 += s.e[2].im + s.e[6].im;
    R3 += T(2) * s.e[4].im - s.e[0].im - s.e[8].im;
    R4 += s.e[5].re - s.e[7].re;
    I4 += s.e[5].im + s.e[7].im;
  }
  // local tr Delta^2 (lin_derivs.c:180-183)
  __device__ __forceinline__ double tr2() const {
    const double r0 = R0, r3 = R3;
    return 0.4444444444444444 * (r0 * r0 + r3 * (r3 + r0)) +
           ((double)R1 * R1 + (double)I1 * I1 + (double)R2 * R2 + (double)I2 * I2 + (double)R4 * R4 +
            (double)I4 * I4);
  }
  // packing of Gfix/FACG.c:57-60; returns tr Delta^2
  __device__ __forceinline__ double store(vec2<T> *__restrict__ F, long long V, long long i) const {
    st<T>(F + i, mk<T>(-R3 / T(3), R0 / T(3)));
    st<T>(F + V + i, mk<T>(T(0.5) * R1, T(0.5) * I1));
    st<T>(F + 2 * V + i, mk<T>(T(0.5) * R2, T(0.5) * I2));
    st<T>(F + 3 * V + i, mk<T>(T(0.5) * R4, T(0.5) * I4));
    return tr2();
  }
};

// parabola vertex through the three probes, clamped: Gfix/line_search.c:38-57
__device__ __forceinline__ double approx_minimum(const double al[3], const double f[3]) {
  const double A = f[0] / ((al[0] - al[1]) * (al[0] - al[2]));
  const double B = f[1] / ((al[1] - al[0]) * (al[1] - al[2]));
  const double C = f[2] / ((al[2] - al[0]) * (al[2] - al[1]));
  const double sol = (al[0] * (B + C) + al[1] * (A + C) + al[2] * (A + B)) / (2 * (A + B + C));
  if (sol > al[2]) return al[2];
  if (isnan(sol) || isinf(sol) || sol < al[0]) return 0.08;
  return sol;
}

// Device-resident scalars of one FACG/FASD run.  Every per-iteration kernel
// reads `done` first and returns when it is set, so a captured batch of
// iterations becomes a no-op past convergence.
struct Ctrl {
  double theta;      // theta_k of the current iteration (GFNORM_LANDAU)
  double inold;      // Polak-Ribiere denominator
  double insum;      // <in, in>
  double sum_conj;   // <in, in - in_old>
  double beta;
  double alpha;      // step of the current update
  double probe[3];   // line-search functionals at 0, PL1, PL2
  double acc;        // stopping accuracy
  double gf_alpha;   // GF_TUNE
  double cg_tol;     // PR -> FR switch (CG_TOL)
  double switch_theta;  // mixed precision: FP32 -> FP64 switch
  double stale_probe[3];  // GLU's red[] slots left by the previous segment
  double stale_insum;
  double stale_theta;
  long long iters;   // iterations in the current segment
  long long max_iters;
  long long total;   // iterations over all segments (monitoring)
  // fused gauge updates that ran (counted by the line-search finaliser
  // that precedes each one); the host compares it with the updates it
  // enqueued to put its buffer pointers back after a stop inside a batch
  long long upd;
  int done;          // set after the update of the iteration that stopped
  int did_ls;        // current iteration used the line search
  int sd_start;      // currently inside the SD start
  int switched;      // mixed precision: 1 reached switch_theta, 2 stalled
  long long trace_cap;
  long long trace_n;
  // safeguards of the FP32 phase of a mixed run (active while
  // switch_theta > 0): bit 0 restarts CG (beta = 0) after an uphill
  // direction, bit 1 caps beta at beta_cap, bit 2 takes no step along an
  // uphill direction
  int sg_mode;
  int sg_uphill;  // the last line search found the direction uphill
  double beta_cap;
  // FP32-phase stall exit: once theta_min < stall_theta, a theta above
  // stall_factor * theta_min ends the phase early (switched), as at the
  // switch point; stall_factor 0 disables it
  double theta_min;
  double stall_theta;
  double stall_factor;
};

// ---------------------------------------------------------------- reduce --
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_down(v, o, kWave);
  return v;
}

// Sum NV values over the block; thread 0 writes them to out[0..NV).
// The order of additions is fixed, so results are reproducible.
template <int NV>
__device__ __forceinline__ void block_sum_store(double (&v)[NV], double *out) {
  __shared__ double sh[NV][kBlock / kWave];
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const double s = wave_sum(v[k]);
    if (lane == 0) sh[k][w] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = 0.0;
      for (int j = 0; j < (int)(blockDim.x / kWave); j++) s += sh[k][j];
      out[k] = s;
    }
  }
}

}  // namespace gfix
