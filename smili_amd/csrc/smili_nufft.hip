// ES-kernel NUFFT kernels for gfx950, templated on the real type R of the
// spreading / interpolation / FFT arithmetic (float: the throughput path;
// double: the reference-precision path).  Type 2 (image -> visibilities,
// finufft2d2 with iflag=0, src/wrapfinufft/wrapfinufft.f90:94-128) and type 1
// (residuals -> image, finufft2d1 with iflag=1, wrapfinufft.f90:172-206) both
// use the +i sign.  Method: Barnett, Magland & af Klinteberg, SIAM J. Sci.
// Comput. 41(5), 2019 (exponential-of-semicircle kernel, width 7, upsampling
// 2, beta = 2.30 * 7).
#pragma once
#include "smili_kernels.h"

namespace smili {

template <typename R> struct Cx;
template <> struct Cx<float> { using T = float2; };
template <> struct Cx<double> { using T = double2; };

template <typename V> __device__ __forceinline__ V cmk(decltype(V::x) a, decltype(V::x) b) {
  V r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return cmk<V>(a.x + b.x, a.y + b.y); }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return cmk<V>(a.x - b.x, a.y - b.y); }
template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
  return cmk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename V, typename S> __device__ __forceinline__ V cscale(V a, S s) { return cmk<V>(a.x * s, a.y * s); }
template <typename V> __device__ __forceinline__ V mul_i(V a) { return cmk<V>(-a.y, a.x); }

// ---------------------------------------------------------------------------
// small DFTs with the +i sign: y_m = sum_q v_q exp(+2 pi i q m / r)
// ---------------------------------------------------------------------------
template <typename V> __device__ __forceinline__ void dft2(V* v) {
  V a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <typename V> __device__ __forceinline__ void dft4(V& a0, V& a1, V& a2, V& a3) {
  V t0 = cadd(a0, a2), t1 = csub(a0, a2);
  V t2 = cadd(a1, a3), t3 = mul_i(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  V y1 = cadd(t1, t3), y3 = csub(t1, t3);
  a1 = y1;
  a3 = y3;
}

template <typename V> __device__ __forceinline__ void dft3(V* v) {
  using R = decltype(V::x);
  const R s3 = (R)0.86602540378443864676;
  V t = cadd(v[1], v[2]);
  V m = cmk<V>(v[0].x - (R)0.5 * t.x, v[0].y - (R)0.5 * t.y);
  V d = mul_i(cscale(csub(v[1], v[2]), s3));
  v[0] = cadd(v[0], t);
  v[1] = cadd(m, d);
  v[2] = csub(m, d);
}

template <typename V> __device__ __forceinline__ void dft5(V* v) {
  using R = decltype(V::x);
  const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
  const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;
  V t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  V t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  V a0 = v[0];
  V b1 = cmk<V>(a0.x + c1 * t1.x + c2 * t2.x, a0.y + c1 * t1.y + c2 * t2.y);
  V b2 = cmk<V>(a0.x + c2 * t1.x + c1 * t2.x, a0.y + c2 * t1.y + c1 * t2.y);
  V d1 = mul_i(cmk<V>(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
  V d2 = mul_i(cmk<V>(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
  v[0] = cadd(a0, cadd(t1, t2));
  v[1] = cadd(b1, d1);
  v[4] = csub(b1, d1);
  v[2] = cadd(b2, d2);
  v[3] = csub(b2, d2);
}

template <typename V> __device__ __forceinline__ void dft8(V* v) {
  using R = decltype(V::x);
  const R s = (R)0.70710678118654752440;
  V e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  V o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4(e0, e1, e2, e3);
  dft4(o0, o1, o2, o3);
  // twiddles W8^m = exp(+i pi m / 4)
  V w1 = cmk<V>(s * (o1.x - o1.y), s * (o1.x + o1.y));
  V w2 = mul_i(o2);
  V w3 = cmk<V>(-s * (o3.x + o3.y), s * (o3.x - o3.y));
  v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
  v[1] = cadd(e1, w1); v[5] = csub(e1, w1);
  v[2] = cadd(e2, w2); v[6] = csub(e2, w2);
  v[3] = cadd(e3, w3); v[7] = csub(e3, w3);
}

// ---------------------------------------------------------------------------
// LDS Stockham FFT (autosort, +i sign) of nb transforms of length n stored
// at a + b*stride.  Uses scratch s of the same shape and the twiddle table
// tw[t] = exp(+2 pi i t / n) staged in LDS.  Radices: the largest power of
// two up to maxr that divides what is left, then 3, 5 -- a smaller maxr
// means more stages but more butterflies per stage (more threads busy on a
// short transform).  Returns the buffer that holds the result.  All
// threads of the block must call it.
// ---------------------------------------------------------------------------
template <typename V>
__device__ V* lds_fft(V* a, V* s, int nb, int stride, int n, const V* tw, int maxr) {
  int Ns = 1, rem = n;
  while (rem > 1) {
    const int r = (maxr >= 8 && rem % 8 == 0) ? 8 : (maxr >= 4 && rem % 4 == 0) ? 4 : (rem % 2 == 0) ? 2
                : (rem % 3 == 0) ? 3 : 5;
    rem /= r;
    const int nbf = n / r;
    const int total = nbf * nb;
    const int twstep = n / (Ns * r);
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
      const int b = t / nbf;
      const int i = t - b * nbf;
      const int k = i % Ns;
      const V* in = a + b * stride;
      V* out = s + b * stride;
      V v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (q < r) v[q] = in[i + q * nbf];
      if (Ns > 1) {
        const int step = k * twstep;
#pragma unroll
        for (int q = 1; q < 8; ++q)
          if (q < r) v[q] = cmul(v[q], tw[q * step]);
      }
      switch (r) {
        case 2: dft2(v); break;
        case 3: dft3(v); break;
        case 4: dft4(v[0], v[1], v[2], v[3]); break;
        case 5: dft5(v); break;
        default: dft8(v); break;
      }
      const int base = (i / Ns) * Ns * r + k;
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (q < r) out[base + q * Ns] = v[q];
    }
    __syncthreads();
    V* tmp = a; a = s; s = tmp;
    Ns *= r;
  }
  return a;
}

template <typename V>
__device__ __forceinline__ void stage_twiddles(V* dst, const V* __restrict__ tw, int off, int n) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) dst[t] = tw[off + t];
}

// ES kernel weights.  The points are fixed for a plan's lifetime, so their
// ES weights are computed once on the host (FP64) and stored per tap:
// wts[q * npt + j] = phi_x,q of point j, wts[(kW + q) * npt + j] = phi_y,q.
// Both transforms read them (coalesced over consecutive points) instead of
// evaluating 14 exp / sqrt / divisions per point each.
template <typename R>
__device__ __forceinline__ void load_weights(const R* __restrict__ wts, long long npt, int j, R* wx, R* wy) {
#pragma unroll
  for (int q = 0; q < kW; ++q) {
    wx[q] = wts[q * npt + j];
    wy[q] = wts[(kW + q) * npt + j];
  }
}

// Residual of each point for the type-1 transform: the sum of the
// chi-square chunks' slot shares that reference the point, in a fixed
// order, or, for the bare adjoint, the caller's Vres[j].  pref[j] = (first
// slot, second slot, number of slots, offset of the others in refs): the
// common one- and two-slot points need one load round trip after pref, the
// others load their further slots four at a time.  A null pref: point j's
// residual is the single share sres[j].
struct ResidSrc {
  const double2* sres;
  const int4* pref;
  const int* refs;
  const double2* Vres;
};

__device__ __forceinline__ double2 point_resid(const ResidSrc& S, int j) {
  if (S.Vres) return S.Vres[j];
  if (!S.pref) return S.sres[j];
  const int4 p = S.pref[j];
  double2 r = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
  if (p.z > 0) r = S.sres[p.x];
  if (p.z > 1) b = S.sres[p.y];
  r.x += b.x;
  r.y += b.y;
  for (int e = 0; e < p.z - 2; e += 4) {
    int id[4];
    double2 t[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (e + k < p.z - 2) id[k] = S.refs[p.w + e + k];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (e + k < p.z - 2) t[k] = S.sres[id[k]];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (e + k < p.z - 2) {
        r.x += t[k].x;
        r.y += t[k].y;
      }
  }
  return r;
}

// ---------------------------------------------------------------------------
// Type-2 step 1: image rows -> deconvolve, zero-pad, FFT along x.
// ---------------------------------------------------------------------------
// Work item of this work-group for kernels whose neighbouring work items
// share cache lines (the column passes: a 128-byte line holds 4-8 columns).
// Work-groups are dispatched round-robin over the 8 XCDs, each with its own
// L2, so work-group b runs on XCD b % 8; XCD x is given the contiguous item
// range [x * per, (x + 1) * per) so that line-sharing items meet in one L2.
// The grid is 8 * per work-groups (per = ceil(nitems / 8)); -1 past the end.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_item(int nitems) {
  const int per = gridDim.x / kXcds;
  const int b = blockIdx.x;
  const int item = (b % kXcds) * per + b / kXcds;
  return item < nitems ? item : -1;
}

template <typename R>
__global__ void __launch_bounds__(kThreads)
k_t2_rows(const FftItem* __restrict__ items, const typename Cx<R>::T* __restrict__ tw,
          const R* __restrict__ corr, const double* __restrict__ x,
          const double* __restrict__ I2dbuf, typename Cx<R>::T* __restrict__ gridA) {
  using V = typename Cx<R>::T;
  extern __shared__ __align__(16) unsigned char smraw[];
  V* sm = reinterpret_cast<V*>(smraw);
  const FftItem it = items[blockIdx.x];
  const int n = it.nf1, nb = it.count;
  V* A = sm;
  V* B = sm + nb * n;
  V* W = B + nb * n;
  stage_twiddles(W, tw, it.tw1, n);
  const double* img = it.identity ? (x + it.x_off) : (I2dbuf + it.img_off);
  const R* cx = corr + it.cx_off;
  const R* cy = corr + it.cy_off;
  const int hx = it.nx / 2;
  for (int t = threadIdx.x; t < nb * n; t += blockDim.x) {
    const int b = t / n, c = t - b * n;
    // grid column c holds mode k1 = c (c < nf1 - nx/2) or c - nf1
    const int k1 = (c < n - hx) ? c : c - n;
    V val = cmk<V>(0, 0);
    if (k1 >= -hx && k1 < it.nx - hx) {
      const int ix = k1 + hx;
      const int iy = it.first + b;
      val.x = (R)(img[(long long)iy * it.nx + ix] * (double)cx[ix] * (double)cy[iy]);
    }
    A[t] = val;
  }
  __syncthreads();
  V* Rv = lds_fft(A, B, nb, n, n, W, it.maxr);
  const int hy = it.ny / 2;
  for (int t = threadIdx.x; t < nb * n; t += blockDim.x) {
    const int b = t / n, c = t - b * n;
    int row = it.first + b - hy;
    if (row < 0) row += it.nf2;
    gridA[it.grid_off + (long long)row * n + c] = Rv[t];
  }
}

// ---------------------------------------------------------------------------
// Type-2 step 2: FFT along y of `count` grid columns; only the ny rows that
// hold image modes are read (the others are zero by construction).
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(kThreads)
k_t2_cols(const FftItem* __restrict__ items, int nitems, const typename Cx<R>::T* __restrict__ tw,
          typename Cx<R>::T* __restrict__ gridA) {
  using V = typename Cx<R>::T;
  extern __shared__ __align__(16) unsigned char smraw[];
  V* sm = reinterpret_cast<V*>(smraw);
  const int item = xcd_item(nitems);
  if (item < 0) return;
  const FftItem it = items[item];
  const int n = it.nf2, nb = it.count, stride = n + 1;
  V* A = sm;
  V* B = sm + nb * stride;
  V* W = B + nb * stride;
  stage_twiddles(W, tw, it.tw2, n);
  const int hy = it.ny / 2;
  V* g = gridA + it.grid_off;
  for (int t = threadIdx.x; t < nb * n; t += blockDim.x) {
    const int r = t / nb, b = t - r * nb;
    const int k2 = (r < n - hy) ? r : r - n;
    V val = cmk<V>(0, 0);
    if (k2 >= -hy && k2 < it.ny - hy) val = g[(long long)r * it.nf1 + it.first + b];
    A[b * stride + r] = val;
  }
  __syncthreads();
  V* Rv = lds_fft(A, B, nb, stride, n, W, it.maxr);
  for (int t = threadIdx.x; t < nb * n; t += blockDim.x) {
    const int r = t / nb, b = t - r * nb;
    g[(long long)r * it.nf1 + it.first + b] = Rv[b * stride + r];
  }
}

// ---------------------------------------------------------------------------
// Type-2 step 3: interpolation from an LDS-staged tile of the grid.  One work
// item = up to sub_points points of one bin.  The bin edge is a template
// argument, so the tile is a static LDS array (a captured graph replays a
// kernel with dynamic LDS after a gap of ~18 us on this stack; static LDS
// kernels follow their predecessor directly).
// ---------------------------------------------------------------------------
template <typename R, int BIN>
__global__ void __launch_bounds__(kThreads)
k_interp(const SubProb* __restrict__ sps, const typename Cx<R>::T* __restrict__ gridA,
         const int* __restrict__ pt_cell, const R* __restrict__ wts, long long npt,
         typename Cx<R>::T* __restrict__ Vout) {
  using V = typename Cx<R>::T;
  constexpr int T = BIN + kW - 1;
  constexpr int bin = BIN;
  __shared__ __align__(16) V tile[T * T];
  const SubProb sp = sps[blockIdx.x];
  const V* g = gridA + sp.grid_off;
  const int gx0 = sp.bin_x * bin, gy0 = sp.bin_y * bin;
  // the first point's cell and weights are loaded before the tile, so both
  // loads are in flight together
  int cell = 0;
  R wx[kW], wy[kW];
  if ((int)threadIdx.x < sp.count) {
    cell = pt_cell[sp.start + threadIdx.x];
    load_weights(wts, npt, sp.start + threadIdx.x, wx, wy);
  }
  for (int c = threadIdx.x; c < T * T; c += blockDim.x) {
    const int ty = c / T, tx = c - ty * T;
    const int gy = (gy0 + ty) % sp.nf2;
    const int gx = (gx0 + tx) % sp.nf1;
    tile[c] = g[(long long)gy * sp.nf1 + gx];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < sp.count; i += blockDim.x) {
    const int j = sp.start + i;
    if (i != (int)threadIdx.x) {
      cell = pt_cell[j];
      load_weights(wts, npt, j, wx, wy);
    }
    const int lx = (cell & 0xffff) - gx0;
    const int ly = (cell >> 16) - gy0;
    R ar = 0, ai = 0;
#pragma unroll
    for (int qy = 0; qy < kW; ++qy) {
      const V* row = tile + (ly + qy) * T + lx;
      R rr = 0, ri = 0;
#pragma unroll
      for (int qx = 0; qx < kW; ++qx) {
        const V gv = row[qx];
        rr = fma(wx[qx], gv.x, rr);
        ri = fma(wx[qx], gv.y, ri);
      }
      ar = fma(wy[qy], rr, ar);
      ai = fma(wy[qy], ri, ai);
    }
    Vout[j] = cmk<V>(ar, ai);
  }
}

// ---------------------------------------------------------------------------
// Type-1 step 1: spread residuals c = Vres_re - i Vres_im onto the grid.
// One work item = up to 256 points (sorted by first-tap cell) of one 8x8
// block of cells.  Each wave stages the ES weights of its 64 points in LDS
// (lane = point: 7 + 7 kernel values and the residual), then switches roles
// (lane = one of the 49 taps) and walks the points in order, accumulating
// each run of points of one cell in registers and adding the run to the
// 14x14 LDS footprint when the cell changes.  Two ways to finish:
//  * default: one footprint for the work-group (LDS atomics of its four
//    waves), added to the grid with one float atomic per touched cell
//    (work items of one block and of neighbouring blocks overlap); the grid
//    is zero on entry (its consumer clears what it reads);
//  * DET (deterministic): one footprint per wave, summed in wave order, and
//    every cell the item's points can reach stored with a plain store into
//    that cell's slot of the slot array (a cell's slots are contiguous, in
//    item order); k_spread_gather then adds each cell's slots in order into
//    the grid.  No float atomics across waves: the spread grid is the same
//    bit for bit from run to run.
// ---------------------------------------------------------------------------
#ifndef SMILI_SPREAD_BATCH
#define SMILI_SPREAD_BATCH 4
#endif
constexpr int kSpreadBatch = SMILI_SPREAD_BATCH;

template <typename R, bool DET>
__device__ void spread_block(int bid, const SpreadItem* __restrict__ items, const int* __restrict__ pt_cell,
                             const R* __restrict__ wts, long long npt, const ResidSrc& res,
                             typename Cx<R>::T* __restrict__ gridB, const int* __restrict__ slot_idx,
                             typename Cx<R>::T* __restrict__ slots) {
  using V = typename Cx<R>::T;
  constexpr int T = kCellBlock + kW - 1;
  constexpr int NW = kThreads / 64;
  constexpr int NT = DET ? NW : 1;                 // footprints: one per wave, or one shared
  __shared__ V tile[NT][T * T];
  __shared__ R wxs[NW][kW][65], wys[NW][kW][65];   // padded: the 7 rows sit in distinct banks
  __shared__ V cs[NW][64];
  __shared__ int cl[NW][64];
  const SpreadItem it = items[bid];
  for (int c = threadIdx.x; c < NT * T * T; c += blockDim.x) (&tile[0][0])[c] = cmk<V>(0, 0);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int base = it.pt_begin + w * 64;
  const int np = max(0, min(64, it.pt_end - base));
  if (lane < np) {
    const int j = base + lane;
    R wx[kW], wy[kW];
    load_weights(wts, npt, j, wx, wy);
#pragma unroll
    for (int q = 0; q < kW; ++q) {
      wxs[w][q][lane] = wx[q];
      wys[w][q][lane] = wy[q];
    }
    const double2 r = point_resid(res, j);
    cs[w][lane] = cmk<V>((R)r.x, (R)(-r.y));
    const int cell = pt_cell[j];
    cl[w][lane] = ((cell >> 16) - it.gy0) * T + ((cell & 0xffff) - it.gx0);
  }
  __syncthreads();
  if (lane < kW * kW && np > 0) {
    // lanes hold distinct taps, so a wave's adds of one run go to distinct
    // cells, and its runs are added in point order (a wave's LDS operations
    // complete in issue order)
    V* tw = tile[DET ? w : 0];
    const int a = lane / kW, b = lane - a * kW;
    const int off = a * T + b;
    R ar = 0, ai = 0;
    int cur = cl[w][0];
    auto step = [&](int c, R wt, V v) {
      if (c != cur) {
        atomicAdd(&tw[cur + off].x, ar);
        atomicAdd(&tw[cur + off].y, ai);
        ar = 0;
        ai = 0;
        cur = c;
      }
      ar = fma(wt, v.x, ar);
      ai = fma(wt, v.y, ai);
    };
    // kSpreadBatch points' LDS values are read together, then walked in
    // order (the walk is otherwise one LDS latency per point)
    int p = 0;
    for (; p + kSpreadBatch <= np; p += kSpreadBatch) {
      int c[kSpreadBatch];
      R wt[kSpreadBatch];
      V v[kSpreadBatch];
#pragma unroll
      for (int k = 0; k < kSpreadBatch; ++k) {
        c[k] = cl[w][p + k];
        wt[k] = wxs[w][b][p + k] * wys[w][a][p + k];
        v[k] = cs[w][p + k];
      }
#pragma unroll
      for (int k = 0; k < kSpreadBatch; ++k) step(c[k], wt[k], v[k]);
    }
    for (; p < np; ++p) step(cl[w][p], wxs[w][b][p] * wys[w][a][p], cs[w][p]);
    atomicAdd(&tw[cur + off].x, ar);
    atomicAdd(&tw[cur + off].y, ai);
  }
  __syncthreads();
  if constexpr (DET) {
    // reachable cells in bit order: the slot index of cell c is entry
    // slot_base + (number of reachable cells before c)
    for (int c = threadIdx.x; c < T * T; c += blockDim.x) {
      const int wd = c >> 5;
      const unsigned bit = 1u << (c & 31);
      if (!(it.mask[wd] & bit)) continue;
      int rank = __popc(it.mask[wd] & (bit - 1));
      for (int k = 0; k < wd; ++k) rank += __popc(it.mask[k]);
      V val = tile[0][c];
#pragma unroll
      for (int k = 1; k < NW; ++k) {
        val.x += tile[k][c].x;
        val.y += tile[k][c].y;
      }
      slots[slot_idx[it.slot_base + rank]] = val;
    }
  } else {
    V* g = gridB + it.grid_off;
    for (int c = threadIdx.x; c < T * T; c += blockDim.x) {
      const V val = tile[0][c];
      if (val.x != 0 || val.y != 0) {
        const int ty = c / T, tx = c - ty * T;
        const int gy = (it.gy0 + ty) % it.nf2;
        const int gx = (it.gx0 + tx) % it.nf1;
        V* dst = g + (long long)gy * it.nf1 + gx;
        unsafeAtomicAdd(&dst->x, val.x);
        unsafeAtomicAdd(&dst->y, val.y);
      }
    }
  }
}

template <typename R, bool DET>
__global__ void __launch_bounds__(kThreads)
k_spread(const SpreadItem* __restrict__ items, const int* __restrict__ pt_cell,
         const R* __restrict__ wts, long long npt, ResidSrc res, typename Cx<R>::T* __restrict__ gridB,
         const int* __restrict__ slot_idx, typename Cx<R>::T* __restrict__ slots) {
  spread_block<R, DET>(blockIdx.x, items, pt_cell, wts, npt, res, gridB, slot_idx, slots);
}

// Type-1 step 1b: every grid cell some spreading item can reach is the sum
// of its slots, added in item order (cell c's slots are [off[c], off[c+1])).
// Cells no item reaches are never written and stay zero.
template <typename V>
__global__ void __launch_bounds__(kThreads)
k_spread_gather(const long long* __restrict__ cell, const int* __restrict__ off, int ncells,
                const V* __restrict__ slots, V* __restrict__ grid) {
  constexpr int U = 8;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x) {
    const int b = off[c], e = off[c + 1];
    V s = slots[b];
    int k = b + 1;
    for (; k + U <= e; k += U) {
      V v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = slots[k + u];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        s.x += v[u].x;
        s.y += v[u].y;
      }
    }
    for (; k < e; ++k) {
      const V v = slots[k];
      s.x += v.x;
      s.y += v.y;
    }
    grid[cell[c]] = s;
  }
}

// ---------------------------------------------------------------------------
// Type-1 step 2: FFT along y of `count` full grid columns; keeps the ny
// output rows that are image modes (compact, into H = gridA) and clears the
// consumed columns of gridB for the next evaluation.
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(kThreads)
k_t1_cols(const FftItem* __restrict__ items, int nitems, const typename Cx<R>::T* __restrict__ tw,
          typename Cx<R>::T* __restrict__ gridB, typename Cx<R>::T* __restrict__ H) {
  using V = typename Cx<R>::T;
  extern __shared__ __align__(16) unsigned char smraw[];
  V* sm = reinterpret_cast<V*>(smraw);
  const int item = xcd_item(nitems);
  if (item < 0) return;
  const FftItem it = items[item];
  const int n = it.nf2, nb = it.count, stride = n + 1;
  V* A = sm;
  V* B = sm + nb * stride;
  V* W = B + nb * stride;
  stage_twiddles(W, tw, it.tw2, n);
  V* g = gridB + it.grid_off;
  for (int t = threadIdx.x; t < nb * n; t += blockDim.x) {
    const int r = t / nb, b = t - r * nb;
    V* src = g + (long long)r * it.nf1 + it.first + b;
    A[b * stride + r] = *src;
    *src = cmk<V>(0, 0);
  }
  __syncthreads();
  V* Rv = lds_fft(A, B, nb, stride, n, W, it.maxr);
  const int hy = it.ny / 2;
  V* h = H + it.grid_off;
  for (int t = threadIdx.x; t < nb * it.ny; t += blockDim.x) {
    const int iy = t / nb, b = t - iy * nb;
    int r = iy - hy;
    if (r < 0) r += n;
    h[(long long)iy * it.nf1 + it.first + b] = Rv[b * stride + r];
  }
}

// ---------------------------------------------------------------------------
// Type-1 step 3: FFT along x of `count` compact rows, keep the nx modes,
// deconvolve, real part -> the chi-square gradient (FP64): either stored as
// a 2-D image (grad2d) or added to the regularizer gradient into grad.
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(kThreads)
k_t1_rows(const FftItem* __restrict__ items, const typename Cx<R>::T* __restrict__ tw,
          const R* __restrict__ corr, const typename Cx<R>::T* __restrict__ H,
          double* __restrict__ grad2d, double* __restrict__ grad, const double* __restrict__ greg,
          const int* __restrict__ pixinv, int assemble) {
  using V = typename Cx<R>::T;
  extern __shared__ __align__(16) unsigned char smraw[];
  V* sm = reinterpret_cast<V*>(smraw);
  const FftItem it = items[blockIdx.x];
  const int n = it.nf1, nb = it.count;
  V* A = sm;
  V* B = sm + nb * n;
  V* W = B + nb * n;
  stage_twiddles(W, tw, it.tw1, n);
  const V* h = H + it.grid_off + (long long)it.first * n;
  for (int t = threadIdx.x; t < nb * n; t += blockDim.x) A[t] = h[t];
  __syncthreads();
  V* Rv = lds_fft(A, B, nb, n, n, W, it.maxr);
  const R* cx = corr + it.cx_off;
  const R* cy = corr + it.cy_off;
  const int hx = it.nx / 2;
  for (int t = threadIdx.x; t < nb * it.nx; t += blockDim.x) {
    const int b = t / it.nx, ix = t - b * it.nx;
    const int iy = it.first + b;
    int c = ix - hx;
    if (c < 0) c += n;
    const double val = (double)Rv[b * n + c].x * (double)cx[ix] * (double)cy[iy];
    const long long p = (long long)iy * it.nx + ix;
    if (!assemble) {
      grad2d[it.img_off + p] = val;
    } else {
      // gradient assembly: grad = gradchisq + gradreg (fftim2d.f90:553-554),
      // gradreg from k_reg; window pixels through the inverse pixel map
      const int k = it.identity ? (int)p : pixinv[it.img_off + p];
      if (k >= 0) grad[it.x_off + k] = val + greg[it.x_off + k];
    }
  }
}

}  // namespace smili
