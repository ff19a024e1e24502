// Pruned-DFT form of the NUFFT's grid transforms for gfx950.
//
// The ES-kernel NUFFT (finufft2d2 / finufft2d1 with the +i sign,
// src/wrapfinufft/wrapfinufft.f90:94-128, 172-206) needs the oversampled
// grid only at the cells its points' 7 x 7 taps touch.  For VLBI imaging
// the pixel is a fraction of the beam, so the uv points fill a small
// window of the fine grid (a 256^2 image at 2 uas with baselines up to
// 10 Glambda touches 110 x 60 of 512 x 512 cells, v >= 0 after the
// conjugate flip of tools.py:20-27).  On that window the transforms are
// two small dense products each, with the deconvolution factors folded
// into the DFT matrices:
//
//   type 2:  S = Ey I          (gh x Ny)(Ny x Nx), I real
//            G = S Ex          (gh x Nx)(Nx x gw)         -> interpolation
//   type 1:  P = Gs ExT        (gh x gw)(gw x Nx)         <- spreading
//            grad = Re(EyT P)  (Ny x gh)(gh x Nx)
//
// Ex[k1][a] = c_x(k1) exp(+2 pi i (k1 - Nx/2)(x0 + a) / nf1),
// Ey[b][k2] = c_y(k2) exp(+2 pi i (k2 - Ny/2)(y0 + b) / nf2),
// (x0, y0) the window's first fine-grid cell.  These are the values the
// FFT path's zero-padded transforms produce at those cells, so the result
// is the same NUFFT; the products replace 4 latency-bound LDS FFT passes
// over the whole grid with four launches of a few hundred work-groups of
// independent FMA chains.
#pragma once
#include "smili_kernels.h"
#include "smili_nufft.hip"

namespace smili {

enum { DFT_STORE = 0, DFT_GRAD = 1 };

// One tile of C = A B per work-group of NW waves.  Lane = output column;
// the waves split K into KS slices (reduced through LDS) and the remaining
// NW / KS wave groups take RB rows each.  A thread's slice is short (K / KS
// <= ~16 steps) and its B loads are issued in batches of U before the FMAs,
// so a wave keeps all its loads in flight at once: the products are a few
// hundred thousand FMAs spread over thousands of waves, bound by load
// latency, not by the FMA rate.  A's element is uniform across a wave
// (scalar loads), B's row is read coalesced.  Rows past M are clamped for
// the loads and not stored.
//   STORE:  C = A B (complex)
//   GRAD:   Re(A B) -> grad2d, or + gradreg -> grad (gradient assembly,
//           fftim2d.f90:553-554, through the window's inverse pixel map);
//           the spreading window Z that the first product read is cleared
//           for the next type-1 transform (each of the Nx Ny outputs
//           clears cells p, p + Nx Ny, ... of it)
template <typename R, int NW, int RB, int KS, bool BREAL, int EPI>
__global__ void __launch_bounds__(64 * NW)
k_dft(const DftItem* __restrict__ items, int nitems, const typename Cx<R>::T* __restrict__ A,
      const typename Cx<R>::T* __restrict__ Bc, const double* __restrict__ Bx,
      const double* __restrict__ Bimg, typename Cx<R>::T* __restrict__ Cout,
      typename Cx<R>::T* __restrict__ Z, double* __restrict__ grad2d, double* __restrict__ grad,
      const double* __restrict__ greg, const int* __restrict__ pixinv, int assemble) {
  using V = typename Cx<R>::T;
  static_assert(NW % KS == 0, "KS must divide the wave count");
  constexpr int MG = NW / KS;
  constexpr int U = BREAL ? 16 : 8;        // B loads in flight per batch
  const int item = xcd_item(nitems);   // uniform per work-group
  if (item < 0) return;
  const DftItem it = items[item];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ks = w % KS, mg = w / KS;
  const int n = it.n0 + lane;
  const int nn = n < it.N ? n : it.N - 1;
  const int mb = it.m0 + mg * RB;
  const int kper = (it.K + KS - 1) / KS;
  const int kb = ks * kper;
  const int ke = min(it.K, kb + kper);
  const V* arow[RB];
#pragma unroll
  for (int r = 0; r < RB; ++r) arow[r] = A + it.a_off + (long long)min(mb + r, it.M - 1) * it.lda;
  // two accumulators per output component (even / odd k): independent FMA chains
  R ar[RB][2], ai[RB][2];
#pragma unroll
  for (int r = 0; r < RB; ++r) ar[r][0] = ar[r][1] = ai[r][0] = ai[r][1] = (R)0;
  const double* Br = (it.bsel ? Bimg : Bx) + it.b_off + nn;
  const V* Bv = Bc + it.b_off + nn;
  auto fmas = [&](int k, int h, const V& bv, R br) {
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const V a = arow[r][k];
      if constexpr (BREAL) {
        ar[r][h] = fma(a.x, br, ar[r][h]);
        ai[r][h] = fma(a.y, br, ai[r][h]);
      } else if constexpr (EPI == DFT_GRAD) {
        // Re(a b) = a.x b.x - a.y b.y: the two halves in the two chains
        ar[r][0] = fma(a.x, bv.x, ar[r][0]);
        ar[r][1] = fma(a.y, bv.y, ar[r][1]);
      } else {
        ar[r][h] = fma(a.x, bv.x, ar[r][h]);
        ar[r][h] = fma(-a.y, bv.y, ar[r][h]);
        ai[r][h] = fma(a.x, bv.y, ai[r][h]);
        ai[r][h] = fma(a.y, bv.x, ai[r][h]);
      }
    }
  };
  int k = kb;
  for (; k + U <= ke; k += U) {
    V bv[U];
    R br[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if constexpr (BREAL) br[u] = (R)Br[(long long)(k + u) * it.ldb];
      else bv[u] = Bv[(long long)(k + u) * it.ldb];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) fmas(k + u, u & 1, bv[u], br[u]);
  }
  for (; k < ke; ++k) {
    V bv{};
    R br = 0;
    if constexpr (BREAL) br = (R)Br[(long long)k * it.ldb];
    else bv = Bv[(long long)k * it.ldb];
    fmas(k, k & 1, bv, br);
  }
  R cr[RB], ci[RB];
#pragma unroll
  for (int r = 0; r < RB; ++r) {
    if constexpr (EPI == DFT_GRAD) {
      cr[r] = ar[r][0] - ar[r][1];
      ci[r] = (R)0;
    } else {
      cr[r] = ar[r][0] + ar[r][1];
      ci[r] = ai[r][0] + ai[r][1];
    }
  }
  if constexpr (KS > 1) {
    __shared__ V red[NW][RB][64];
#pragma unroll
    for (int r = 0; r < RB; ++r) red[w][r][lane] = cmk<V>(cr[r], ci[r]);
    __syncthreads();
    if (ks != 0) return;
#pragma unroll
    for (int s = 1; s < KS; ++s)
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const V p = red[w + s][r][lane];
        cr[r] += p.x;
        ci[r] += p.y;
      }
  }
  if (n >= it.N) return;
#pragma unroll
  for (int r = 0; r < RB; ++r) {
    const int m = mb + r;
    if (m >= it.M) break;
    if constexpr (EPI == DFT_GRAD) {
      const long long p = (long long)m * it.nx + n;
      for (long long z = p; z < it.zn; z += (long long)it.M * it.nx) Z[it.z_off + z] = cmk<V>((R)0, (R)0);
      const double val = (double)cr[r];
      if (!assemble) {
        grad2d[it.img_off + p] = val;
      } else {
        const int k = it.identity ? (int)p : pixinv[it.img_off + p];
        if (k >= 0) grad[it.x_off + k] = val + greg[it.x_off + k];
      }
    } else {
      Cout[it.c_off + (long long)m * it.ldc + n] = cmk<V>(cr[r], ci[r]);
    }
  }
}

// Shapes of the four products: waves per work-group NW, rows per thread
// RB, K slices KS (rows per work-group = RB * NW / KS).  K is Ny, Nx, the
// window width and the window height in turn.
#ifndef SMILI_DFT_T2Y
#define SMILI_DFT_T2Y 16, 1, 16      // S = Ey I
#endif
#ifndef SMILI_DFT_T2X
#define SMILI_DFT_T2X 16, 1, 16      // G = S Ex
#endif
#ifndef SMILI_DFT_T1X
#define SMILI_DFT_T1X 8, 1, 8        // P = Gs ExT (256^2 image: 3.9 us vs 4.5 with 16, 1, 8)
#endif
#ifndef SMILI_DFT_T1Y
#define SMILI_DFT_T1Y 16, 1, 4       // grad = Re(EyT P)
#endif

// The shapes above suit a plan with few products (one 256^2 image: a few
// hundred work-groups, so each splits K many ways).  A plan with many
// (movies, surveys: thousands of work-groups) takes these, with more rows
// and fewer K slices per work-group and (for three of them) 8 waves
// (swept per product over 30 variants; per evaluation: 64-frame movie 166
// -> ~97 us, 512-problem survey 2.53 -> ~1.42 ms; the 256^2 image is 53 us
// with the shapes above, 73 us with the first of these sets).
#ifndef SMILI_DFT_B_T2Y
#define SMILI_DFT_B_T2Y 8, 4, 4
#endif
#ifndef SMILI_DFT_B_T2X
#define SMILI_DFT_B_T2X 8, 2, 4
#endif
#ifndef SMILI_DFT_B_T1X
#define SMILI_DFT_B_T1X 8, 4, 4
#endif
#ifndef SMILI_DFT_B_T1Y
#define SMILI_DFT_B_T1Y 8, 4, 1      // 64-frame movie 92.2 us (92.8 with 4, 4, 1)
#endif
// plans with at least this many work-groups in their first product (with
// the few-product shape) take the batch shapes
#ifndef SMILI_DFT_BATCH_ITEMS
#define SMILI_DFT_BATCH_ITEMS 1024
#endif

// Plans far larger still (hundreds of images: the 512-problem survey has
// ~43,000 first-product work-groups, the 64-frame movie ~3,500) take 4-wave
// work-groups with fewer K slices: survey 1.188 -> 1.124 ms against the
// batch shapes, which stay the movie's (97.0 vs 91.8 us with these).
#ifndef SMILI_DFT_LARGE_ITEMS
#define SMILI_DFT_LARGE_ITEMS 16384
#endif
#ifndef SMILI_DFT_L_T2Y
#define SMILI_DFT_L_T2Y 4, 4, 4
#endif
#ifndef SMILI_DFT_L_T2X
#define SMILI_DFT_L_T2X 4, 4, 2
#endif
#ifndef SMILI_DFT_L_T1X
#define SMILI_DFT_L_T1X 4, 4, 2
#endif
#ifndef SMILI_DFT_L_T1Y
#define SMILI_DFT_L_T1Y 4, 8, 1      // survey 1.134 -> 1.124 ms vs 4, 4, 1
#endif

template <int NW, int RB, int KS> struct DftShape {
  static constexpr int nw = NW, rb = RB, ks = KS, rows = RB * (NW / KS), threads = 64 * NW;
};
using ShapeT2Y = DftShape<SMILI_DFT_T2Y>;
using ShapeT2X = DftShape<SMILI_DFT_T2X>;
using ShapeT1X = DftShape<SMILI_DFT_T1X>;
using ShapeT1Y = DftShape<SMILI_DFT_T1Y>;
using ShapeBT2Y = DftShape<SMILI_DFT_B_T2Y>;
using ShapeBT2X = DftShape<SMILI_DFT_B_T2X>;
using ShapeBT1X = DftShape<SMILI_DFT_B_T1X>;
using ShapeBT1Y = DftShape<SMILI_DFT_B_T1Y>;
using ShapeLT2Y = DftShape<SMILI_DFT_L_T2Y>;
using ShapeLT2X = DftShape<SMILI_DFT_L_T2X>;
using ShapeLT1X = DftShape<SMILI_DFT_L_T1X>;
using ShapeLT1Y = DftShape<SMILI_DFT_L_T1Y>;

}  // namespace smili
