// Device-side data structures shared by the kernels and the host plan
// builder.  Every evaluation is expressed as work lists over "units"
// (one image + its uv points: a problem of a batch, or a frame of a movie)
// and "problems" (one cost: a set of chi-square data blocks over the units'
// visibilities plus the units' regularizers).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace smili {

constexpr int kThreads = 256;        // work-group size of every kernel but the chi-square ones
constexpr int kChisqThreadsMax = 512;  // chi-square (+ regularizer) work-groups: 256 or 512
constexpr int kW = 7;                // ES kernel width (eps 1e-6)
constexpr int kMaxRadixStages = 12;
constexpr int kRegTerms = 7;         // l1, tv, tsv, kl, gs, rt, ri partial sums per item
constexpr int kMomTerms = 10;        // moments for tfd / cen / sm
constexpr double kZeroEps = 1e-10;   // image.f90:7

// host-side FFT plan of one length (the kernels derive the radices)
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
  int movie, frame;                  // movie index (-1: none) and frame number
  double nxref, nyref;
  double reg_scale;                  // multiplies this unit's reg cost and gradient
  // regularizers (lambda <= 0 disables)
  double l1_l, tv_l, tsv_l, kl_l, gs_l;
  // per-pixel weights; a null pointer means every pixel's weight is the
  // matching *_u (uniform weights: the flat-prior TV / TSV of the
  // reference's set-up), which then is not read per pixel
  const double *l1_w, *tv_w, *tsv_w, *kl_w, *gs_w;
  double l1_u, tv_u, tsv_u, kl_u, gs_u;
  double tfd_l, tfd_tgtfd, cen_l, cen_alpha, sm_l, sm_maj, sm_min, sm_phi;
  int need_moments;
};

// chi-square data items.  A work item ("chunk") of consecutive data stages
// the model visibilities of the distinct uv points its data reference into
// LDS; each datum refers to them by chunk-local slot numbers.  A slot field
// holds the slot in bits 0..14 and the sign of the uv index (1 = the datum
// sits at (-u, -v)) in bit 15.  w = fnorm / variance.
#ifndef SMILI_MAX_SLOTS
#define SMILI_MAX_SLOTS 1024
#endif
constexpr int kMaxSlots = SMILI_MAX_SLOTS;      // distinct uv points per chunk
constexpr int kSlotNeg = 0x8000;
// pt: the datum's global point (the direct path reads V[pt] without the
// slot table)
struct FcvItem { double vr, vi, w; unsigned short slot, pad0; int pt; };
struct AmpItem { double amp, w; unsigned short slot, pad0; int pt; };
struct CpItem  { double cc, sc, w; unsigned short slot[3], pad0; };   // cos, sin of the closure phase
struct CaItem  { double ca, w; unsigned short slot[4]; };

enum { DT_FCV = 0, DT_AMP = 1, DT_CP = 2, DT_CA = 3 };

// direct: every datum of the chunk has its own uv point (amplitudes and
// full visibilities at distinct points), so it skips the LDS staging
struct Chunk { int type, prob, start, count, slot_off, nslots, direct, pad1; };
// FFT work item: `count` rows or columns starting at `first` of one unit,
// with the unit's geometry embedded so a work-group needs one descriptor load
struct FftItem {
  int first, count, nx, ny, nf1, nf2;
  int tw1, tw2;                      // twiddle tables of length nf1, nf2
  int cx_off, cy_off, x_off, identity;
  int maxr, pad;                     // largest power-of-two radix of the LDS FFT
  long long grid_off, img_off;
};
// Pruned-DFT product work item: one tile of C = A B (A complex, M x K,
// leading dimension lda; B real or complex, K x N, ldb; C, ldc), rows
// [m0, m0 + tile rows), columns [n0, n0 + 64).  bsel picks the real B
// operand (0: x, 1: the scattered 2-D image).  The gradient epilogue writes
// Re C[m][n] as image pixel (ix = n, iy = m) of the unit at x_off / img_off
// and clears the zn cells of the unit's spreading window at z_off.
struct DftItem {
  int m0, n0, M, N, K, lda, ldb, ldc;
  long long a_off, b_off, c_off, img_off, z_off;
  int x_off, identity, nx, bsel, zn, pad;
};
// Grid of a unit as the interpolation and spreading kernels see it: nf1 x
// nf2 cells with periodic wrap (the full fine grid of the FFT path, or the
// compact window of fine-grid cells the unit's points touch in the pruned
// DFT path, where no tap reaches past the window's edge).
struct SubProb { int bin_x, bin_y, start, count, nf1, nf2; long long grid_off; };
struct RegItem { int unit, start, count, pad; };
// spreading: up to kSpreadPoints consecutive points (sorted by cell) of one
// 8x8 block of first-tap cells; the block's 14x14 grid footprint is
// accumulated in LDS and added to the grid once per work item
constexpr int kCellBlock = 8;
constexpr int kSpreadPoints = 256;   // = work-group size: one staging round per wave
// mask: the item's footprint cells its points can reach (bit ty * 14 + tx of
// the 14 x 14 footprint); slot_base: where its cells' slot indices start in
// the plan's slot table (one per reachable cell, in bit order)
struct SpreadItem {
  int gx0, gy0, pt_begin, pt_end, nf1, nf2;
  long long grid_off;
  int slot_base, pad0;
  unsigned mask[7];
  int pad1;
};

// temporal regularizers of one movie (frames = consecutive units with
// consecutive x ranges of npix pixels each)
struct MovieDev {
  int nz, npix;
  int x_off;                         // offset of frame 0 in x / grad
  int pad;
  double rt_l, ri_l;
  const double *rt_w, *ri_w;         // per window pixel; null: uniform, rt_u / ri_u
  const double *lc;                  // light-curve normalisation per frame
  double *iavg;                      // mean normalised frame (npix)
  double rt_u, ri_u;
};

struct ProbDev {
  int chunk_begin[4], chunk_end[4];  // chi-square chunk ranges per data type
  int reg_begin, reg_end;            // reg item range
  int unit_begin, unit_end;
  int n[4];                          // data counts
  double w[4];                       // normalised weights
  int terms_off;
  int nterms;                        // 15, or 18 for a movie
  int lc_mode;                       // movie in light-curve mode: tfd terms are the light curve
  int any_moments;                   // some unit of the problem has tfd / cen / sm
};

// the unit's image moments (k_moments) and, from them, its tfd / cen / sm
// costs (image.f90:111-135) scaled by reg_scale (k_moments_reduce)
struct UnitMoments {
  double S, A, Sx, Sy, Sxx, Syy, Sxy, P, Px, Py;
  double tfd_c, cen_c, sm_c;
};

}  // namespace smili
