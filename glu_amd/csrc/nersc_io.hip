// nersc_io.hip -- NERSC gauge-file output of a context's device links.
//
// The NERSC branch of GLU's write_lat (IO/writers.c:330-422): the links are
// reunitarised in place, the checksum and the plaquette / link trace of the
// header are taken on the device, and the links leave the device in chunks
// through the staging buffer, each chunk byte-swapped to the chosen
// endianness and appended to the file behind the header.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/glu_gfix.h"
#include "gfix_ctx.h"

using namespace gfix;

namespace {

// NERSC checksum (IO/writers.c:49-69, 183-203): the uint32 sum of the two
// 32-bit halves of every stored double, over rows 0 and 1 of each link
// (nentry = 6, 4D_SU3_GAUGE) or all nine entries (3x3), of the links as the
// unpacking kernel completes them.  Addition mod 2^32 does not depend on
// the order, so the word is the host's for any reduction order.
__global__ __launch_bounds__(kBlock) void k_nersc_checksum(const double2 *__restrict__ U, long long V, int nentry,
                                                          unsigned *__restrict__ part) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned s = 0;
  if (i < V) {
#pragma unroll 1
    for (int mu = 0; mu < kND; mu++) {
      const su3<double> m = load_link<double>(U, mu, i, V);
#pragma unroll
      for (int k = 0; k < 9; k++) {
        const unsigned long long a = (unsigned long long)__double_as_longlong(m.e[k].re);
        const unsigned long long b = (unsigned long long)__double_as_longlong(m.e[k].im);
        const unsigned w = (unsigned)a + (unsigned)(a >> 32) + (unsigned)b + (unsigned)(b >> 32);
        s += k < nentry ? w : 0u;
      }
    }
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_down(s, o, kWave);
  __shared__ unsigned sh[kBlock / kWave];
  if (threadIdx.x % kWave == 0) sh[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned t = 0;
    for (int w = 0; w < kBlock / kWave; w++) t += sh[w];
    part[blockIdx.x] = t;
  }
}

// UTC date as GLU prints it (Utils/GLU_timer.c:79-86): asctime without
// its newline
std::string utc_date() {
  const time_t now = time(nullptr);
  struct tm t;
  char buf[64];
  if (!gmtime_r(&now, &t) || !asctime_r(&t, buf)) return "Mon Jan  0 00:00:00 1970";
  std::string d(buf);
  while (!d.empty() && (d.back() == '\n' || d.back() == '\r')) d.pop_back();
  return d;
}

}  // namespace

extern "C" int glu_gfix_write_nersc(glu_gfix_ctx *c, const char *path, int two_rows, int big_endian,
                                    const char *label, size_t sequence) {
  if (!c || !path) return gfix_fail(-1, "glu_gfix_write_nersc: null argument");
  if (c->run.inflight) return gfix_fail(-1, "a run is in flight on this context");
  // the lattice is reunitarised before it is written (IO/writers.c:333)
  if (int rc = glu_gfix_reunit(c)) return rc;
  double plaq = 0.0, tr = 0.0;
  if (int rc = glu_gfix_measure(c, &plaq, &tr, nullptr, nullptr, nullptr)) return rc;
  const long long V = c->g.V;
  const int nentry = two_rows ? 6 : 9;
  hipLaunchKernelGGL(k_nersc_checksum, dim3((unsigned)c->nblk), dim3(kBlock), 0, c->stream,
                     (const double2 *)c->U64, V, nentry, (unsigned *)c->part);
  std::vector<unsigned> part(c->nblk);
  if (hipMemcpyAsync(part.data(), c->part, part.size() * sizeof(unsigned), hipMemcpyDeviceToHost, c->stream) !=
          hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return gfix_fail(-2, "checksum failed: %s", hipGetErrorString(hipGetLastError()));
  unsigned sum = 0;
  for (unsigned p : part) sum += p;
  if (int rc = ensure_stage(c)) return rc;

  FILE *f = fopen(path, "wb");
  if (!f) return gfix_fail(-3, "cannot open %s for writing", path);
  // write_header_NERSC, IO/write_headers.c:77-138
  const char *user = getenv("USER");
  fprintf(f, "BEGIN_HEADER\nHDR_VERSION = 1.0\nDATATYPE = %s\nSTORAGE_FORMAT = 1.0\n",
          two_rows ? "4D_SU3_GAUGE" : "4D_SU3_GAUGE_3x3");
  for (int mu = 0; mu < 4; mu++) fprintf(f, "DIMENSION_%d = %zu\n", mu + 1, c->dims[mu]);
  fprintf(f, "CHECKSUM = %x\nLINK_TRACE = %1.15f\nPLAQUETTE = %1.15f\n", sum, tr, plaq);
  for (int mu = 0; mu < 4; mu++) fprintf(f, "BOUNDARY_%d = PERIODIC\n", mu + 1);
  fprintf(f,
          "ENSEMBLE_ID = ukqcd\nSEQUENCE_NUMBER = %zu\nENSEMBLE_LABEL = %s\nCREATOR_MACHINE = %s\n"
          "CREATION_DATE = %s\nFLOATING_POINT = %s\nEND_HEADER\n",
          sequence, label ? label : "", user ? user : "(null)", utc_date().c_str(),
          big_endian ? "IEEE64BIG" : "IEEE64LITTLE");

  // data: site-major, then direction, then the stored entries as re/im
  // pairs (IO/writers.c:110-140, 242-300)
  const int per = 2 * nentry;
  const bool swap = big_endian != (__BYTE_ORDER__ == __ORDER_BIG_ENDIAN__);
  std::vector<double> h((size_t)kChunk * 72);
  std::vector<uint64_t> out((size_t)kChunk * 4 * per);
  int rc = 0;
  for (long long s0 = 0; s0 < V && !rc; s0 += kChunk) {
    const long long n = std::min(kChunk, V - s0);
    launch_unpack<double>(c->U64, (double *)c->stage, V, 4, s0, n, c->stream);
    if (hipMemcpyAsync(h.data(), c->stage, (size_t)n * 72 * sizeof(double), hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
      rc = gfix_fail(-2, "link download failed");
      break;
    }
    size_t w = 0;
    for (long long l = 0; l < n * 4; l++)
      for (int e = 0; e < per; e++) {
        uint64_t u;
        memcpy(&u, &h[(size_t)l * 18 + e], 8);
        out[w++] = swap ? __builtin_bswap64(u) : u;
      }
    if (fwrite(out.data(), 8, w, f) != w) rc = gfix_fail(-3, "write to %s failed", path);
  }
  if (fclose(f) != 0 && !rc) rc = gfix_fail(-3, "closing %s failed", path);
  return rc;
}
