// gfix_kernels.h -- launch interface of the gauge-fixing kernels (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "gfix_device.h"

namespace gfix {

// one FFT axis pass: pencils of length N at stride S
struct FftAxis {
  int N;
  int P;        // pencils per workgroup
  long long S;  // element stride along the axis
  long long npencil;
  int nst;
  int rad[8];
  const double2 *tw;  // N entries exp(-2 pi i k / N)
  int pv;             // pencils-per-workgroup variant of the LDS kernel (tuned)
};

// closed-form p^2 preconditioner of the fused pass
struct PsqInfo {
  int D;             // 4 Landau, 3 Coulomb
  int L[4];
  int off[4];        // offsets of each axis' cos table in cosv
  const double *cosv;
  double max;        // MAX_LANDAU = 4 ND / V or MAX_COULOMB = 4 (ND-1) / Lcu
};

// nslice > 1: independent transforms of slice_len elements each (per field),
// ax.npencil counting the pencils of one slice; with slice_map, batch row z
// transforms slice slice_map[z] (nslice rows); slices flagged in slice_done
// are skipped
template <typename T>
void launch_fft_axis(void *F, long long fstride, int nfield, const FftAxis &ax, int mode, const PsqInfo &ps,
                     const Ctrl *ctrl, hipStream_t s, int nslice = 1, long long slice_len = 0,
                     const int *slice_done = nullptr, const int *slice_map = nullptr);

// both in-plane axes (x, y) in one pass when the plane fits a workgroup;
// false when this shape has no plane kernel (the caller runs axis passes)
template <typename T>
bool launch_fft_xy(void *F, long long fstride, int nfield, const FftAxis &ax, const FftAxis &ay, int mode,
                   const Ctrl *ctrl, hipStream_t s, long long nplane, int nslice, long long slice_len,
                   const int *slice_done, const int *slice_map = nullptr);

// the z and t axes of the 4D transform with p^2 in between (forward z, t,
// x p^2, backward t, z) in one pass over the fields; false when this shape
// has no z-t kernel (the caller runs axis passes); launch = false only
// answers whether it has one; variant: x values per workgroup (tuned)
template <typename T>
bool launch_fft_zt(void *F, long long V, int nfield, const FftAxis &ax, const FftAxis &ay, const FftAxis &az,
                   const FftAxis &at, const PsqInfo &ps, const Ctrl *ctrl, hipStream_t s, bool launch = true,
                   int variant = 0);

template <typename T>
void launch_deriv(const void *U, void *F, double *part, const Geo &g, const Ctrl *ctrl, hipStream_t s);
template <typename T>
void launch_pr_sums(const void *F, const void *O, double *part, long long V, const Ctrl *ctrl, hipStream_t s);
template <typename T>
void launch_cg_update(const void *F, void *sn, long long V, const Ctrl *ctrl, hipStream_t s);
template <typename T>
void launch_copy(const void *F, void *a, long long V, const Ctrl *ctrl, hipStream_t s);
// the line search on 256-site tiles: variant k (1..kLsTiles) has the tile
// kLsTileShape[k - 1] = (TX, TY, TZ, TT), the link columns held in
// registers for both probes (k <= 4) or loaded per probe (k > 4); 0, or a
// shape that does not divide the lattice, runs the one-site-per-thread kernel
constexpr int kLsTiles = 8;
constexpr int kLsTileShape[kLsTiles][4] = {{8, 4, 4, 2}, {4, 4, 4, 4}, {8, 8, 2, 2}, {16, 4, 2, 2},
                                           {8, 4, 4, 2}, {4, 4, 4, 4}, {8, 8, 2, 2}, {16, 4, 2, 2}};
bool ls_tile_fits(const Geo &g, int tile);
template <typename T>
void launch_line_search(const void *U, const void *F, double *part, const Geo &g, const Ctrl *ctrl, int force,
                        hipStream_t s, int tile = 0);
template <typename T>
void launch_egauge(void *U, const void *F, void *G, const Geo &g, const Ctrl *ctrl, int fixed, hipStream_t s);
template <typename T>
void launch_egauge_deriv(const void *Uin, void *Uout, const void *D, void *G, void *Fin, void *Fout, double *part,
                         const Geo &g, const Ctrl *ctrl, hipStream_t s);
template <typename T>
void launch_pack(const double *aos, void *U, long long V, int nd, long long site0, long long n, hipStream_t s);
template <typename T>
void launch_unpack(const void *U, double *aos, long long V, int nd, long long site0, long long n, hipStream_t s);
template <typename T> void launch_identity(void *G, long long V, hipStream_t s);

void launch_fin_theta(const double *part, long long nb, Ctrl *ctrl, double norm, hipStream_t s);
void launch_fin_pr(const double *part, long long nb, Ctrl *ctrl, hipStream_t s);
void launch_fin_sd(const double *part, long long nb, Ctrl *ctrl, hipStream_t s);
void launch_fin_ls(const double *part, long long nb, Ctrl *ctrl, double *trace, int force, double pl1, double pl2,
                   hipStream_t s);
void launch_iter_end(Ctrl *ctrl, double *trace, int fasd, hipStream_t s);
void launch_sum_part(const double *part, long long nb, double *out, hipStream_t s);
void launch_fin_iter(const double *part, long long nb, Ctrl *ctrl, double *trace, double norm, hipStream_t s);
void launch_begin_segment(Ctrl *ctrl, int sd, hipStream_t s);
void launch_demote(const void *U, void *W, long long n, hipStream_t s);
void launch_apply_gauge(void *U, const void *G, const Geo &g, hipStream_t s);
void launch_promote(const void *a, void *b, long long n, hipStream_t s);
void launch_reunit(void *U, long long V, int nd, hipStream_t s);
void launch_measure(const void *U, double *part, const Geo &g, hipStream_t s);

}  // namespace gfix
