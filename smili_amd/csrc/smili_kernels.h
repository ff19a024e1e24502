// Device-side data structures shared by the kernels and the host plan
// builder.  Every evaluation is expressed as work lists over "units"
// (one image + its uv points: a problem of a batch, or a frame of a movie)
// and "problems" (one cost: a set of chi-square data blocks over the units'
// visibilities plus the units' regularizers).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace smili {

constexpr int kThreads = 256;        // work-group size of every kernel
constexpr int kW = 7;                // ES kernel width (eps 1e-6)
constexpr int kMaxRadixStages = 12;
constexpr int kRegTerms = 5;         // l1, tv, tsv, kl, gs partial sums per item
constexpr int kMomTerms = 10;        // moments for tfd / cen / sm
constexpr double kZeroEps = 1e-10;   // image.f90:7

struct FftPlanDev {
  int n;
  int nstages;
  int radix[kMaxRadixStages];
  int tw_off;                        // offset of exp(+2 pi i t / n), t < n
};

struct UnitDev {
  int nx, ny, nf1, nf2;
  int npix, identity;                // identity: x is the full image, x fastest
  int m, pt_off;                     // points [pt_off, pt_off + m)
  int x_off;                         // offset of this unit's pixels in x / grad
  int fft1, fft2;                    // plan index for nf1, nf2
  long long grid_off;                // float2 offset into the fine grids
  long long img_off;                 // double offset into I2d / grad2d (nx*ny)
  int cx_off, cy_off;                // deconvolution factors (float)
  int pix_off;                       // window pixel list offset (pix = iy*nx+ix)
  int prob;                          // problem (cost) this unit belongs to
  double nxref, nyref;
  double reg_scale;                  // multiplies this unit's reg cost and gradient
  // regularizers (lambda <= 0 disables)
  double l1_l, tv_l, tsv_l, kl_l, gs_l;
  const double *l1_w, *tv_w, *tsv_w, *kl_w, *gs_w;
  double tfd_l, tfd_tgtfd, cen_l, cen_alpha, sm_l, sm_maj, sm_min, sm_phi;
  int need_moments;
};

// chi-square data items; idx are signed 1-based indices into the global
// (sorted) point arrays; w = fnorm / variance
struct FcvItem { double vr, vi, w; int idx, pad; };
struct AmpItem { double amp, w; int idx, pad; };
struct CpItem  { int idx[4]; double cp, w; };
struct CaItem  { int idx[4]; double ca, w; };

enum { DT_FCV = 0, DT_AMP = 1, DT_CP = 2, DT_CA = 3 };

struct Chunk { int type, prob, start, count; };
struct RowItem { int unit, first, count, pad; };   // rows or columns
struct SubProb { int unit, bin_x, bin_y, start, count, pad; };
struct RegItem { int unit, start, count, pad; };

struct ProbDev {
  int chunk_begin[4], chunk_end[4];  // chi-square chunk ranges per data type
  int reg_begin, reg_end;            // reg item range
  int unit_begin, unit_end;
  int n[4];                          // data counts
  double w[4];                       // normalised weights
  int terms_off;
};

struct UnitMoments {
  double S, A, Sx, Sy, Sxx, Syy, Sxy, P, Px, Py;
};

}  // namespace smili
