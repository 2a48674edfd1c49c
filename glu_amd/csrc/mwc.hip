// mwc.hip -- GLU's parallel MWC_4096 random streams, on the host (the
// process stream the drop-in entry points share, as GLU's global one) and
// on the device (one copy per context), and the random SU(3) matrices drawn
// from them (Sunitary_gen) for the random gauge transform and the restarts.
//
// GLU keeps one generator per OpenMP thread (Utils/par_rng.c:43-132,
// Utils/par_MWC_4096.c:33-112).  Its index is reset to 4095 on every call
// (par_MWC_4096.c:35-36), so only table[0] is ever used: a lag-1
// complementary multiply-with-carry with multiplier 18782 on 2^32 - 1.
// Stream t is seeded with SEED + t through MT19937's init_genrand recursion
// (par_MWC_4096.c:72-89), its carry is table[4095] mod 809430660, and it is
// warmed up by 10000 draws (par_rng.c:113-120).  A double is x 2^-32
// (par_MWC_4096.c:110).  A loop over n sites hands stream t libgomp's static
// block of them (the first n mod nthreads streams take one extra site), and
// each stream draws its sites in order.
//
// On the device one lane walks one stream: the draws are inherently
// sequential per stream (the polar method rejects a variable number of
// pairs), so a launch has as many lanes as GLU has threads.  The accept test
// runs in exactly the host's arithmetic (no fused multiply-add), so every
// stream consumes the same uint32 sequence as GLU's; the accepted Gaussians
// agree with the host's to the last bit or two of log().
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/glu_gfix.h"
#include "gfix_ctx.h"

using namespace gfix;

// ------------------------------------------------------------------ host --
// one step of stream (x, carry): Utils/par_MWC_4096.c:33-46
static inline uint32_t mwc_step(uint32_t &x, uint32_t &carry) {
  const uint64_t p = 18782ULL * (uint64_t)x + carry;
  carry = (uint32_t)(p >> 32);
  uint32_t v = (uint32_t)p + carry;
  if (v < carry) {
    v++;
    carry++;
  }
  x = 0xfffffffeu - v;
  return x;
}

void mwc_seed_states(uint32_t seed, int nth, uint32_t *x, uint32_t *carry) {
  std::vector<uint32_t> tab(4096);
  for (int t = 0; t < nth; t++) {
    tab[0] = seed + (uint32_t)t;
    for (uint32_t j = 1; j < 4096; j++) tab[j] = (uint32_t)(1812433253UL * (tab[j - 1] ^ (tab[j - 1] >> 30)) + j);
    x[t] = tab[0];
    carry[t] = tab[4095] % 809430660UL;
    for (int j = 0; j < 10000; j++) mwc_step(x[t], carry[t]);
  }
}

// GLU's process-wide stream (RNG_inited, par_rng.c:43-132), shared by the
// drop-in entry points: Landau()'s restarts, LUXURY_GAUGE copies and
// Coulomb_FA()'s restarts draw from it as GLU's do
MwcHost g_mwc;

int mwc_host_init(uint32_t seed, int nth) {
  if (g_mwc.live) return 0;
  if (nth < 1) nth = 1;
  if (seed == 0) {  // SEED = 0: one 32-bit seed from the entropy pool
    FILE *f = fopen("/dev/urandom", "r");
    if (!f || fread(&seed, sizeof(seed), 1, f) != 1) {
      if (f) fclose(f);
      return gfix_fail(-1, "[RNG] /dev/urandom not read");
    }
    fclose(f);
  }
  g_mwc.x.assign(nth, 0);
  g_mwc.carry.assign(nth, 0);
  mwc_seed_states(seed, nth, g_mwc.x.data(), g_mwc.carry.data());
  g_mwc.live = true;
  return 0;
}

void mwc_host_free() {
  g_mwc.live = false;
  g_mwc.x.clear();
  g_mwc.carry.clear();
}

double mwc_host_dbl(int t) { return (double)(mwc_step(g_mwc.x[t], g_mwc.carry[t]) * 2.3283064365386963e-10); }

extern "C" int glu_gfix_rng_draws(uint32_t seed, int nthreads, int thread, size_t n, double *out) {
  if (nthreads < 1 || thread < 0 || thread >= nthreads || (n && !out) || seed == 0)
    return gfix_fail(-1, "glu_gfix_rng_draws: bad argument");
  std::vector<uint32_t> x(nthreads), c(nthreads);
  mwc_seed_states(seed, nthreads, x.data(), c.data());
  for (size_t k = 0; k < n; k++) out[k] = (double)(mwc_step(x[thread], c[thread]) * 2.3283064365386963e-10);
  return 0;
}

// ---------------------------------------------------------------- device --
namespace {

__device__ __forceinline__ uint32_t dmwc(uint32_t &x, uint32_t &carry) {
  const uint64_t p = 18782ULL * (uint64_t)x + carry;
  carry = (uint32_t)(p >> 32);
  uint32_t v = (uint32_t)p + carry;
  if (v < carry) {
    v++;
    carry++;
  }
  x = 0xfffffffeu - v;
  return x;
}
__device__ __forceinline__ double ddbl(uint32_t &x, uint32_t &c) {
  return (double)dmwc(x, c) * 2.3283064365386963e-10;
}

// Marsaglia's polar method as GLU has it (Utils/par_rng.c:180-186: no
// factor 2 under the root); keep = 0 only advances the stream
__device__ __forceinline__ cx<double> polar(uint32_t &x, uint32_t &c, bool keep) {
  for (;;) {
    const double u = __dadd_rn(2.0 * ddbl(x, c), -1.0);
    const double v = __dadd_rn(2.0 * ddbl(x, c), -1.0);
    const double s = __dadd_rn(__dmul_rn(u, u), __dmul_rn(v, v));
    if (s < 1.0) {
      if (!keep) return mk<double>(0.0, 0.0);
      const double r = sqrt(__ddiv_rn(-log(s), s));
      return mk<double>(__dmul_rn(r, u), __dmul_rn(r, v));
    }
  }
}

// Sunitary_gen's nine Gaussians for every site of `nblk` blocks of n sites
// (block b starts at site first[b] of the G planes): stream t draws its
// static block of each block in turn.  Columns 0 and 1 go to the planes
// (the third is rebuilt by the Gram-Schmidt that follows); the third's
// Gaussians are drawn and dropped, as GLU discards them.
__global__ __launch_bounds__(64) void k_mwc_sunitary(uint32_t *__restrict__ state, int nth, long long n,
                                                     const long long *__restrict__ first, int nblk,
                                                     double2 *__restrict__ G, long long Vp) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nth) return;
  uint32_t x = state[t], c = state[nth + t];
  long long q = n / nth, r = n % nth;
  if (t < r) {
    q++;
    r = 0;
  }
  const long long lo = q * t + r, hi = lo + q;
  // row-major entry k of the 3x3 matrix -> its column plane (k % 3 < 2)
  constexpr int plane[9] = {0, 3, -1, 1, 4, -1, 2, 5, -1};
  for (int b = 0; b < nblk; b++) {
    const long long base = first[b];
    for (long long i = lo; i < hi; i++) {
#pragma unroll 1
      for (int k = 0; k < 9; k++) {
        const cx<double> z = polar(x, c, plane[k] >= 0);
        if (plane[k] >= 0) st<double>(G + (long long)plane[k] * Vp + base + i, z);
      }
    }
  }
  state[t] = x;
  state[nth + t] = c;
}

// Gram-Schmidt of the drawn sites (Matrix_Ops/gramschmidt.c:73-165)
__global__ __launch_bounds__(kBlock) void k_mwc_reunit(double2 *__restrict__ G, long long Vp, long long n,
                                                       const long long *__restrict__ first, int nblk) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n * nblk) return;
  const long long i = first[j / n] + j % n;
  su3<double> m = load_link<double>(G, 0, i, Vp);
  gram_reunit<double>(m);
  store_link<double>(G, 0, i, Vp, m);
}

}  // namespace

static int mwc_upload(glu_gfix_ctx *c, const uint32_t *x, const uint32_t *carry, int nth) {
  if (c->mwc && c->mwc_nth != nth) {
    (void)hipFree(c->mwc);  // a stream count change: the state buffer is reshaped
    c->mwc = nullptr;
  }
  if (!c->mwc && hipMalloc(&c->mwc, 2 * (size_t)nth * sizeof(uint32_t)) != hipSuccess)
    return gfix_fail(-2, "random stream allocation failed");
  c->mwc_nth = nth;
  if (hipMemcpyAsync(c->mwc, x, nth * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync((uint32_t *)c->mwc + nth, carry, nth * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream) !=
          hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return gfix_fail(-2, "random stream upload failed");
  c->mwc_live = 1;
  return 0;
}

int mwc_seed(glu_gfix_ctx *c, uint32_t seed, int nth) {
  if (seed == 0) return gfix_fail(-1, "random stream seed must be nonzero");
  if (nth < 1 || nth > 65536) return gfix_fail(-1, "random stream count %d out of range", nth);
  std::vector<uint32_t> x(nth), carry(nth);
  mwc_seed_states(seed, nth, x.data(), carry.data());
  c->rng_seed = seed;
  c->rng_threads = nth;
  return mwc_upload(c, x.data(), carry.data(), nth);
}

// GLU's initialise_par_rng: seed from the context's seed unless live
int mwc_ensure(glu_gfix_ctx *c) { return c->mwc_live ? 0 : mwc_seed(c, c->rng_seed, c->rng_threads); }

int mwc_from_host(glu_gfix_ctx *c) {
  if (!g_mwc.live) return gfix_fail(-1, "the process random stream is not initialised");
  return mwc_upload(c, g_mwc.x.data(), g_mwc.carry.data(), (int)g_mwc.x.size());
}

int mwc_to_host(glu_gfix_ctx *c) {
  if (!c->mwc_live) return 0;
  const int nth = c->mwc_nth;
  g_mwc.x.resize(nth);
  g_mwc.carry.resize(nth);
  if (hipMemcpyAsync(g_mwc.x.data(), c->mwc, nth * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipMemcpyAsync(g_mwc.carry.data(), (uint32_t *)c->mwc + nth, nth * sizeof(uint32_t), hipMemcpyDeviceToHost,
                     c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return gfix_fail(-2, "random stream download failed");
  g_mwc.live = true;
  return 0;
}

// Sunitary_gen (Matrix_Ops/gramschmidt_SSE.c:175-181) for blocks of n sites
// starting at first[0..nblk) of the 2-column planes G (plane stride Vp),
// drawn from the context's streams in GLU's order: block by block, each
// stream its static share of a block.  reunit: do the Gram-Schmidt here
// (else the planes keep the raw Gaussians of columns 0 and 1).
// Synchronous.
int mwc_sunitary(glu_gfix_ctx *c, void *G, long long Vp, long long n, const std::vector<long long> &first,
                 bool reunit) {
  if (int rc = mwc_ensure(c)) return rc;
  if (first.empty() || n <= 0) return 0;
  const int nb = (int)first.size();
  long long *dfirst = nullptr;
  if (hipMalloc(&dfirst, nb * sizeof(long long)) != hipSuccess) return gfix_fail(-2, "draw list allocation failed");
  int rc = 0;
  if (hipMemcpyAsync(dfirst, first.data(), nb * sizeof(long long), hipMemcpyHostToDevice, c->stream) != hipSuccess) {
    rc = gfix_fail(-2, "draw list upload failed");
  } else {
    const int nth = c->mwc_nth;
    hipLaunchKernelGGL(k_mwc_sunitary, dim3((unsigned)((nth + 63) / 64)), dim3(64), 0, c->stream, (uint32_t *)c->mwc,
                       nth, n, dfirst, nb, (double2 *)G, Vp);
    const long long tot = n * nb;
    if (reunit)
      hipLaunchKernelGGL(k_mwc_reunit, dim3((unsigned)((tot + kBlock - 1) / kBlock)), dim3(kBlock), 0, c->stream,
                         (double2 *)G, Vp, n, dfirst, nb);
    if (hipStreamSynchronize(c->stream) != hipSuccess)
      rc = gfix_fail(-2, "random matrices failed: %s", hipGetErrorString(hipGetLastError()));
  }
  (void)hipFree(dfirst);  // a failed free leaves nothing to undo
  return rc;
}

// random_transform of Geometry/random_config.c:108-120 on the device links
int mwc_random_transform(glu_gfix_ctx *c) {
  // the second link buffer is free between runs: its first six planes hold
  // the Gaussians, and the transform's load does their Gram-Schmidt
  void *G = c->U64b;
  if (int rc = mwc_sunitary(c, G, c->g.V, c->g.V, std::vector<long long>{0}, false)) return rc;
  launch_apply_gauge(c->U64, G, c->g, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess)
    return gfix_fail(-2, "random transform failed: %s", hipGetErrorString(hipGetLastError()));
  return 0;
}

extern "C" int glu_gfix_rng_seed(glu_gfix_ctx *c, uint32_t seed, int nthreads) {
  if (!c) return gfix_fail(-1, "null context");
  return mwc_seed(c, seed, nthreads);
}

extern "C" int glu_gfix_glu_random_transform(glu_gfix_ctx *c) {
  if (!c) return gfix_fail(-1, "null context");
  if (!c->mwc_live) return gfix_fail(-1, "glu_gfix_glu_random_transform: seed the streams first (glu_gfix_rng_seed)");
  return mwc_random_transform(c);
}
