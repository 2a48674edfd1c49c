/*
 * or_nersc.c -- oracle: NERSC gauge-file reader and writer.
 * TEST INFRASTRUCTURE ONLY (see glu_oracle.h).
 *
 * Header: IO/read_headers.c:229-362 (DIMENSION_1..4, FLOATING_POINT,
 * DATATYPE 4D_SU3_GAUGE = top two rows / 4D_SU3_GAUGE_3x3, CHECKSUM,
 * PLAQUETTE, LINK_TRACE; SEQUENCE_NUMBER must be non-zero).
 * Data: site-major (x fastest), then direction, then the stored rows,
 * re/im pairs (IO/readers.c:170-302).  The third row of a two-row file is
 * conj(row0 x row1) (IO/readers.c:51-58).  The NERSC checksum is the uint32
 * sum of the two 32-bit halves of every stored double, in host order
 * (IO/readers.c:256-263).
 */
#include "or_internal.h"

#include <stdint.h>

static uint64_t bswap64(uint64_t v) { return __builtin_bswap64(v); }
static uint32_t bswap32(uint32_t v) { return __builtin_bswap32(v); }

static int header_value(const char *hdr, const char *key, char *out, size_t n)
{
  const char *p = hdr;
  const size_t kl = strlen(key);
  while ((p = strstr(p, key)) != NULL) {
    const int at_line_start = (p == hdr) || p[-1] == '\n';
    const char *q = p + kl;
    while (*q == ' ' || *q == '\t') q++;
    if (at_line_start && *q == '=') {
      q++;
      while (*q == ' ' || *q == '\t') q++;
      size_t k = 0;
      while (*q && *q != '\n' && *q != '\r' && k + 1 < n) out[k++] = *q++;
      while (k > 0 && (out[k - 1] == ' ' || out[k - 1] == '\t')) k--;
      out[k] = 0;
      return 0;
    }
    p += kl;
  }
  return -1;
}

int or_read_nersc(const char *path, or_nersc_info *info, double *U)
{
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  char *hdr = calloc(1, 65536);
  size_t hl = 0;
  char line[1024];
  int ok = 0;
  while (fgets(line, sizeof(line), f)) {
    const size_t l = strlen(line);
    if (hl + l + 1 < 65536) { memcpy(hdr + hl, line, l); hl += l; }
    if (strncmp(line, "END_HEADER", 10) == 0) { ok = 1; break; }
  }
  if (!ok) { free(hdr); fclose(f); return -2; }
  info->data_offset = ftell(f);
  char v[256];
  for (int mu = 0; mu < OR_ND; mu++) {
    char key[32];
    snprintf(key, sizeof(key), "DIMENSION_%d", mu + 1);
    if (header_value(hdr, key, v, sizeof(v))) { free(hdr); fclose(f); return -3; }
    info->dims[mu] = (size_t)strtoull(v, NULL, 10);
  }
  if (header_value(hdr, "SEQUENCE_NUMBER", v, sizeof(v)) ||
      strtoull(v, NULL, 10) == 0) { free(hdr); fclose(f); return -4; }
  if (header_value(hdr, "FLOATING_POINT", v, sizeof(v))) { free(hdr); fclose(f); return -5; }
  if (!strcmp(v, "IEEE64BIG")) { info->double_prec = 1; info->big_endian = 1; }
  else if (!strcmp(v, "IEEE64LITTLE")) { info->double_prec = 1; info->big_endian = 0; }
  else if (!strcmp(v, "IEEE32BIG") || !strcmp(v, "IEEE32")) { info->double_prec = 0; info->big_endian = 1; }
  else if (!strcmp(v, "IEEE32LITTLE")) { info->double_prec = 0; info->big_endian = 0; }
  else { free(hdr); fclose(f); return -6; }
  if (header_value(hdr, "DATATYPE", v, sizeof(v))) { free(hdr); fclose(f); return -7; }
  if (!strcmp(v, "4D_SU3_GAUGE")) info->two_rows = 1;
  else if (!strcmp(v, "4D_SU3_GAUGE_3x3")) info->two_rows = 0;
  else { free(hdr); fclose(f); return -8; }
  info->checksum_header = 0;
  if (!header_value(hdr, "CHECKSUM", v, sizeof(v)))
    info->checksum_header = (uint32_t)strtoul(v, NULL, 16);
  info->plaquette = header_value(hdr, "PLAQUETTE", v, sizeof(v)) ? 0.0 : atof(v);
  info->link_trace = header_value(hdr, "LINK_TRACE", v, sizeof(v)) ? 0.0 : atof(v);
  free(hdr);
  if (!U) { fclose(f); return 0; }

  size_t V = 1;
  for (int mu = 0; mu < OR_ND; mu++) V *= info->dims[mu];
  const size_t per = info->two_rows ? 12 : 18; /* reals per link */
  const size_t nreal = V * OR_ND * per;
  const size_t wsz = info->double_prec ? 8 : 4;
  unsigned char *buf = malloc(nreal * wsz);
  if (fread(buf, wsz, nreal, f) != nreal) { free(buf); fclose(f); return -9; }
  fclose(f);
  uint32_t sum = 0;
  const int swap = info->big_endian != (__BYTE_ORDER__ == __ORDER_BIG_ENDIAN__);
  for (size_t k = 0; k < nreal; k++) {
    double d;
    if (info->double_prec) {
      uint64_t w;
      memcpy(&w, buf + 8 * k, 8);
      if (swap) w = bswap64(w);
      sum += (uint32_t)(w & 0xffffffffu) + (uint32_t)(w >> 32);
      memcpy(&d, &w, 8);
    } else {
      uint32_t w;
      float fl;
      memcpy(&w, buf + 4 * k, 4);
      if (swap) w = bswap32(w);
      sum += w;
      memcpy(&fl, &w, 4);
      d = fl;
    }
    const size_t link = k / per, e = k % per;
    U[link * 18 + e] = d;
  }
  free(buf);
  info->checksum_computed = sum;
  if (info->two_rows) {
    for (size_t l = 0; l < V * OR_ND; l++) {
      cplx *O = (cplx *)(U + l * 18);
      O[6] = conj(O[1] * O[5] - O[2] * O[4]);
      O[7] = conj(O[2] * O[3] - O[0] * O[5]);
      O[8] = conj(O[0] * O[4] - O[1] * O[3]);
    }
  }
  return 0;
}

/* NERSC writer, the NERSC branch of write_lat (IO/writers.c:330-422): the
   links are reunitarised in place first (:333, skipped with reunit = 0);
   the checksum is the uint32 sum of the 32-bit halves of every stored
   double in host order (:49-69, :183-203); the header keys follow
   write_header_NERSC (IO/write_headers.c:77-138); the data are rows 0 and 1
   of each link (4D_SU3_GAUGE, LOOP_GAUGE) or all nine entries (3x3), in
   IEEE64 of the chosen endianness (GLU's default is big,
   Headers/GLU_definitions.h:651-662).  creator and date stand for $USER
   and the UTC asctime() GLU prints. */
int or_write_nersc(const char *path, const or_lat *L, double *U, int two_rows,
                   int big_endian, int reunit, const char *label,
                   size_t sequence, const char *creator, const char *date)
{
  if (reunit) or_latt_reunit(L, U);
  const size_t per = two_rows ? 12 : 18, nreal = L->V * OR_ND * per;
  const int swap = big_endian != (__BYTE_ORDER__ == __ORDER_BIG_ENDIAN__);
  uint64_t *buf = malloc(nreal * 8);
  uint32_t sum = 0;
  for (size_t l = 0; l < L->V * OR_ND; l++)
    for (size_t e = 0; e < per; e++) {
      uint64_t w;
      memcpy(&w, U + l * 18 + e, 8);
      sum += (uint32_t)(w & 0xffffffffu) + (uint32_t)(w >> 32);
      buf[l * per + e] = swap ? bswap64(w) : w;
    }
  FILE *f = fopen(path, "wb");
  if (!f) { free(buf); return -1; }
  double tmax;
  fprintf(f, "BEGIN_HEADER\nHDR_VERSION = 1.0\nDATATYPE = %s\n"
             "STORAGE_FORMAT = 1.0\n",
          two_rows ? "4D_SU3_GAUGE" : "4D_SU3_GAUGE_3x3");
  for (int mu = 0; mu < OR_ND; mu++)
    fprintf(f, "DIMENSION_%d = %zu\n", mu + 1, L->dims[mu]);
  fprintf(f, "CHECKSUM = %x\n", sum);
  fprintf(f, "LINK_TRACE = %1.15f\n", or_links(L, U, &tmax));
  fprintf(f, "PLAQUETTE = %1.15f\n", or_plaquette(L, U));
  for (int mu = 0; mu < OR_ND; mu++) fprintf(f, "BOUNDARY_%d = PERIODIC\n", mu + 1);
  fprintf(f, "ENSEMBLE_ID = ukqcd\nSEQUENCE_NUMBER = %zu\n"
             "ENSEMBLE_LABEL = %s\nCREATOR_MACHINE = %s\n"
             "CREATION_DATE = %s\nFLOATING_POINT = %s\nEND_HEADER\n",
          sequence, label ? label : "", creator ? creator : "(null)",
          date ? date : "Mon Jan  0 00:00:00 1970",
          big_endian ? "IEEE64BIG" : "IEEE64LITTLE");
  const size_t wrote = fwrite(buf, 8, nreal, f);
  fclose(f);
  free(buf);
  return wrote == nreal ? 0 : -2;
}
