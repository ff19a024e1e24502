// gfx950 kernels of the RML imaging cost: the fused chi-square terms with
// FP64 accumulation, the pixel regularizers and the final reductions (the
// ES-kernel NUFFT kernels are in smili_nufft.hip).
//
// Reference algorithm: fftim2d.calc_cost (src/fortlib/fftim2d.f90:387-607),
// fftlib.calc_chisq (src/fortlib/fftlib.f90:250-655), image.calc_cost_reg
// (src/fortlib/image.f90:12-200, 245-754, 760-1158) and the FINUFFT calls of
// src/wrapfinufft/wrapfinufft.f90:94-206 (both transforms use the +i sign).
#include "smili_kernels.h"
#include "smili_nufft.hip"
#include "smili_dft.hip"

namespace smili {

// ---------------------------------------------------------------------------
// block reduction of doubles (deterministic: fixed shuffle tree + fixed
// wave order)
// ---------------------------------------------------------------------------
// red must hold NV * NW doubles; blockDim.x is a multiple of 64 and at
// most 64 NW.  The result is valid in thread 0.
template <int NV, int NW = kThreads / 64>
__device__ __forceinline__ void block_sum(double* v, double* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  // the NV shuffle trees advance together: one LDS round trip per level
  // for all values instead of one per value and level
  double s[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = v[k];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double t[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) t[k] = __shfl_down(s[k], off, 64);
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k] += t[k];
  }
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) red[k * NW + wid] = s[k];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += red[k * NW + w];
      v[k] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// chi-square terms and residual vectors (fftlib.f90:384-655).  One work item
// = `count` consecutive data of one type of one problem that reference
// `nslots` distinct uv points.  Their model visibilities are staged in LDS
// (FP64) and each point's share of the residual vector is accumulated in
// LDS, then written with plain stores to the chunk's own slot range of
// `sres` (slot_off + slot).  No two chunks write the same slot, so there are
// no global atomics; the spreading kernel sums a point's slots (its
// references, in a fixed order) when it reads the residual.  The
// chi-square of the item is written to partial[blockIdx.x].
//
// Every datum's contribution to one of its points is that point's fixed
// vector times a datum scalar, so a point needs one real accumulator:
//   amp:  (-2 r w / |V|) V                      (fftlib.f90:484-489)
//   cp:   (-2 sin r w sgn / |V|^2) (-Im V, Re V) (fftlib.f90:644-652)
//   ca:   (+-2 r w / |V|^2) V                    (fftlib.f90:556-571)
// Full complex visibilities need two (fftlib.f90:425-433).
// ---------------------------------------------------------------------------

// Closure data are stored by timestamp, and consecutive closures of one
// timestamp share baselines: 64 of them in one wave would hit the same few
// uv points, and LDS atomics on one address serialise.  Lane l of every wave
// therefore takes the closures l Q, l Q + 1, ... (Q = ceil(count / 64)) of
// the chunk, so a wave's lanes hold closures about Q apart.
#ifndef SMILI_CHISQ_INTERLEAVE
#define SMILI_CHISQ_INTERLEAVE 1
#endif
__device__ __forceinline__ int closure_item(int j, int q) {
  return SMILI_CHISQ_INTERLEAVE ? (j & 63) * q + (j >> 6) : j;
}

// Natural logarithm for the cost's per-point and per-pixel logs (log closure
// amplitudes, MEM, GS): x = m 2^k with m in [1/sqrt 2, sqrt 2), and
// log m = 2 atanh s = 2 s (1 + z/3 + z^2/5 + ... + z^10/21), s = (m - 1) /
// (m + 1), z = s^2 <= 0.0295 (the next term is below 1e-18 relative); k ln 2
// in two parts (the high part has trailing zero bits, so k ln2_hi is exact
// for |k| < 2^10; the split is fdlibm's).  About 35 instructions against
// ~150 for the library's, within 2 ulp of glibc's log over 2e7 arguments
// from subnormal to the largest double; zero gives -inf, +inf gives +inf,
// negative and NaN arguments give NaN (selects, no second code path).
__device__ __forceinline__ double log_s(double x) {
  int k;
  double m = frexp(x, &k);   // m in [0.5, 1)
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    --k;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double p = 1.0 / 21.0;
  p = fma(p, z, 1.0 / 19.0);
  p = fma(p, z, 1.0 / 17.0);
  p = fma(p, z, 1.0 / 15.0);
  p = fma(p, z, 1.0 / 13.0);
  p = fma(p, z, 1.0 / 11.0);
  p = fma(p, z, 1.0 / 9.0);
  p = fma(p, z, 1.0 / 7.0);
  p = fma(p, z, 1.0 / 5.0);
  p = fma(p, z, 1.0 / 3.0);
  const double t = 2.0 * s;
  const double lm = fma(t, p * z, t);
  const double dk = (double)k;
  double r = fma(dk, 6.93147180369123816490e-01, fma(dk, 1.90821492927058770002e-10, lm));
  r = x == 0.0 ? -__builtin_inf() : r;
  r = x == __builtin_inf() ? x : r;
  r = x < 0.0 ? __builtin_nan("") : r;
  return r;
}

// The logarithm of the closure-amplitude points, out of line for the LEAN
// kernels (below).
__device__ __noinline__ double log_nl(double x) { return log_s(x); }

// Each thread loads kChisqBatch of its data before it processes them, so
// a wave keeps that many item loads in flight (measured neutral at 1, 2, 4
// and 8 on the 256^2 workload; kept at 4).
#ifndef SMILI_CHISQ_BATCH
#define SMILI_CHISQ_BATCH 4
#endif
constexpr int kChisqBatch = SMILI_CHISQ_BATCH;

// LEAN: the rare or heavy math (FP64 transcendentals here, the tfd / cen /
// sm gradients of the regularizers) is called out of line, so the kernel
// needs about half the registers and a CU holds twice the work-groups.
// Measured: faster for the standalone chi-square and regularizer launches
// of large plans (movie 0.258 -> 0.252 ms); slower inside the fused
// chi-square + regularizer launch of a single problem (C2 0.1045 ->
// 0.1055 ms), which keeps the inline code.

// A chunk's data, kChisqBatch per thread per round trip.  Closure chunks
// are stored in the interleaved order (position j holds closure
// closure_item(j, Q), arranged so on the host), so the loads of a wave are
// contiguous; j >= lim or a position past the chunk's count is not loaded
// (ok = false).
template <typename Item>
__device__ __forceinline__ void load_items(Item (&d)[kChisqBatch], bool (&ok)[kChisqBatch],
                                           const Item* __restrict__ src, int j0, int lim, int count, int q,
                                           bool closure) {
#pragma unroll
  for (int b = 0; b < kChisqBatch; ++b) {
    const int j = j0 + b * (int)blockDim.x;
    const int i = closure ? closure_item(j, q) : j;
    ok[b] = j < lim && i < count;
    if (ok[b]) d[b] = src[closure && SMILI_CHISQ_INTERLEAVE ? j : i];
  }
}

// One datum of a staged chunk: its chi-square, and its scalar(s) added to
// the accumulators of the points it references.
template <bool LEAN>
__device__ __forceinline__ double datum(const FcvItem& d, const double2* Vl, const double2*, double (*Sl)[kMaxSlots]) {
  const int s = d.slot & 0x7fff;
  const double sg = (d.slot & kSlotNeg) ? -1.0 : 1.0;
  const double2 m = Vl[s];
  const double rr = d.vr - m.x;
  const double ri = d.vi - sg * m.y;           // conj for negative index
  const double f = -2.0 * d.w;
  atomicAdd(&Sl[0][s], f * rr);
  atomicAdd(&Sl[1][s], f * ri * sg);
  return (rr * rr + ri * ri) * d.w;
}
template <bool LEAN>
__device__ __forceinline__ double datum(const AmpItem& d, const double2*, const double2* Al, double (*Sl)[kMaxSlots]) {
  const int s = d.slot & 0x7fff;
  const double2 al = Al[s];
  const double r = d.amp - al.x;
  atomicAdd(&Sl[0][s], -2.0 * r * d.w * al.y);
  return r * r * d.w;
}
// Closure phase: with B = V1 V2 V3 (each conjugated for a negative
// index), arg B is the model closure phase and |B| = |V1||V2||V3|, so the
// residual r = cp - arg B has cos r = (cos cp Re B + sin cp Im B) / |B| and
// sin r = (sin cp Re B - cos cp Im B) / |B|: the terms 2 (1 - cos r) and
// -2 sin r of fftlib.f90:617-653 with no transcendental per datum (cos cp
// and sin cp come with the datum, 1/|V| with the staged point).
template <bool LEAN>
__device__ __forceinline__ double datum(const CpItem& d, const double2* Vl, const double2* Al, double (*Sl)[kMaxSlots]) {
  double sg[3], br = 1.0, bi = 0.0, inv = 1.0;
  int s[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    s[k] = d.slot[k] & 0x7fff;
    sg[k] = (d.slot[k] & kSlotNeg) ? -1.0 : 1.0;
    const double2 m = Vl[s[k]];
    const double mi = sg[k] * m.y;
    const double tr = br * m.x - bi * mi;
    bi = br * mi + bi * m.x;
    br = tr;
    inv *= Al[s[k]].x;
  }
  const double cr = (d.cc * br + d.sc * bi) * inv;
  const double sr = (d.sc * br - d.cc * bi) * inv;
  const double f = -2.0 * sr * d.w;
#pragma unroll
  for (int k = 0; k < 3; ++k) atomicAdd(&Sl[0][s[k]], f * sg[k]);
  return 2.0 * (1.0 - cr) * d.w;
}
template <bool LEAN>
__device__ __forceinline__ double datum(const CaItem& d, const double2*, const double2* Al, double (*Sl)[kMaxSlots]) {
  double l[4];
  int s[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    s[k] = d.slot[k] & 0x7fff;     // signs are ignored (fftlib.f90:549-553)
    l[k] = Al[s[k]].x;
  }
  // log|V| = 0.5 log|V|^2
  const double model = 0.5 * (l[0] + l[1] - l[2] - l[3]);
  const double r = d.ca - model;
  const double f = -2.0 * r * d.w;
#pragma unroll
  for (int k = 0; k < 4; ++k) atomicAdd(&Sl[0][s[k]], k < 2 ? f : -f);
  return r * r * d.w;
}

// A chunk whose data share uv points.  Every step is one round trip to
// memory at most, and the data loads overlap the rest: the first batch of
// data is requested before the points are staged, and each later batch
// while the one before it is processed.
//   stage: the chunk's point indices, then their model visibilities
//   (kMaxSlots / kThreads per thread, all requested at once), and once per
//   point rather than once per reference: amp |V| and 1/|V|; closure phase
//   1/|V| and 1/|V|^2; closure amplitude log|V|^2 and 1/|V|^2
//   data: chi-square, and the datum's scalars into the points' accumulators
//   out: each point's share, its accumulator(s) times its fixed vector
template <int TYPE, bool LEAN, typename Item, typename VT>
__device__ __forceinline__ double staged_chunk(const Chunk& ch, const int* __restrict__ sl,
                                               const Item* __restrict__ src, const VT* __restrict__ V,
                                               double2* __restrict__ out, double2* Vl, double2* Al,
                                               double (*Sl)[kMaxSlots]) {
  constexpr int NB = kChisqBatch;
  constexpr int SB = kMaxSlots / kThreads;
  constexpr bool closure = TYPE == DT_CP || TYPE == DT_CA;
  const int nt = blockDim.x;
  const int q = (ch.count + 63) >> 6;
  const int lim = closure ? 64 * q : ch.count;
  src += ch.start;
  Item cur[NB];
  bool ok[NB];
  load_items(cur, ok, src, threadIdx.x, lim, ch.count, q, closure);
  for (int s0 = threadIdx.x; s0 < ch.nslots; s0 += nt * SB) {
    int g[SB];
    VT mv[SB];
#pragma unroll
    for (int b = 0; b < SB; ++b)
      if (s0 + b * nt < ch.nslots) g[b] = sl[s0 + b * nt];
#pragma unroll
    for (int b = 0; b < SB; ++b)
      if (s0 + b * nt < ch.nslots) mv[b] = V[g[b]];
#pragma unroll
    for (int b = 0; b < SB; ++b) {
      const int s = s0 + b * nt;
      if (s >= ch.nslots) break;
      const double2 m = make_double2((double)mv[b].x, (double)mv[b].y);
      Vl[s] = m;
      Sl[0][s] = 0.0;
      Sl[1][s] = 0.0;
      const double a2 = m.x * m.x + m.y * m.y;
      if (TYPE == DT_AMP) {
        const double mod = sqrt(a2);
        Al[s] = make_double2(mod, 1.0 / mod);
      } else if (TYPE == DT_CP) {
        Al[s] = make_double2(1.0 / sqrt(a2), 1.0 / a2);
      } else if (TYPE == DT_CA) {
        Al[s] = make_double2(LEAN ? log_nl(a2) : log_s(a2), 1.0 / a2);
      }
    }
  }
  __syncthreads();
  double acc = 0.0;
  for (int j0 = threadIdx.x; j0 < lim; j0 += nt * NB) {
    Item nxt[NB];
    bool nok[NB];
    load_items(nxt, nok, src, j0 + nt * NB, lim, ch.count, q, closure);
#pragma unroll
    for (int b = 0; b < NB; ++b)
      if (ok[b]) acc += datum<LEAN>(cur[b], Vl, Al, Sl);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      cur[b] = nxt[b];
      ok[b] = nok[b];
    }
  }
  __syncthreads();
  for (int s = threadIdx.x; s < ch.nslots; s += nt) {
    const double2 m = Vl[s];
    const double a = Sl[0][s];
    double2 r;
    if (TYPE == DT_FCV) r = make_double2(a, Sl[1][s]);
    else if (TYPE == DT_AMP) r = make_double2(a * m.x, a * m.y);
    else if (TYPE == DT_CP) r = make_double2(-a * Al[s].y * m.y, a * Al[s].y * m.x);
    else r = make_double2(a * Al[s].y * m.x, a * Al[s].y * m.y);
    out[s] = r;
  }
  return acc;
}

template <typename VT, bool LEAN>
__device__ void chisq_block(int bid, const Chunk* __restrict__ chunks, const int* __restrict__ slots,
                            const FcvItem* __restrict__ fcv, const AmpItem* __restrict__ amp,
                            const CpItem* __restrict__ cp, const CaItem* __restrict__ ca,
                            const VT* __restrict__ V, double2* __restrict__ sres,
                            double* __restrict__ partial) {
  constexpr int NB = kChisqBatch;
  __shared__ double2 Vl[kMaxSlots];
  __shared__ double2 Al[kMaxSlots];      // per-point quantities shared by the chunk's data
  __shared__ double Sl[2][kMaxSlots];    // per-point accumulators
  __shared__ double red[kChisqThreadsMax / 64];
  const Chunk ch = chunks[bid];
  const int* sl = slots + ch.slot_off;
  double2* out = sres + ch.slot_off;
  const int nt = blockDim.x;
  double acc = 0.0;
  if (ch.direct) {
    // one datum per uv point: no staging, the datum's share goes straight
    // to its slot (the same arithmetic as the staged path below)
    for (int i0 = threadIdx.x; i0 < ch.count; i0 += nt * NB) {
      if (ch.type == DT_FCV) {
        FcvItem d[NB];
        VT mv[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b)
          if (i0 + b * nt < ch.count) d[b] = fcv[ch.start + i0 + b * nt];
#pragma unroll
        for (int b = 0; b < NB; ++b)
          if (i0 + b * nt < ch.count) mv[b] = V[d[b].pt];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (i0 + b * nt >= ch.count) break;
          const double sg = (d[b].slot & kSlotNeg) ? -1.0 : 1.0;
          const double rr = d[b].vr - (double)mv[b].x;
          const double ri = d[b].vi - sg * (double)mv[b].y;
          acc += (rr * rr + ri * ri) * d[b].w;
          const double f = -2.0 * d[b].w;
          out[d[b].slot & 0x7fff] = make_double2(f * rr, f * ri * sg);
        }
      } else {
        AmpItem d[NB];
        VT mv[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b)
          if (i0 + b * nt < ch.count) d[b] = amp[ch.start + i0 + b * nt];
#pragma unroll
        for (int b = 0; b < NB; ++b)
          if (i0 + b * nt < ch.count) mv[b] = V[d[b].pt];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (i0 + b * nt >= ch.count) break;
          const double mx = (double)mv[b].x, my = (double)mv[b].y;
          const double mod = sqrt(mx * mx + my * my);
          const double r = d[b].amp - mod;
          acc += r * r * d[b].w;
          const double f = -2.0 * r * d[b].w * (1.0 / mod);
          out[d[b].slot & 0x7fff] = make_double2(f * mx, f * my);
        }
      }
    }
    double v[1] = {acc};
    block_sum<1, kChisqThreadsMax / 64>(v, red);
    if (threadIdx.x == 0) partial[bid] = v[0];
    return;
  }
  double v[1] = {0.0};
  if (ch.type == DT_FCV) v[0] = staged_chunk<DT_FCV, LEAN>(ch, sl, fcv, V, out, Vl, Al, Sl);
  else if (ch.type == DT_AMP) v[0] = staged_chunk<DT_AMP, LEAN>(ch, sl, amp, V, out, Vl, Al, Sl);
  else if (ch.type == DT_CP) v[0] = staged_chunk<DT_CP, LEAN>(ch, sl, cp, V, out, Vl, Al, Sl);
  else v[0] = staged_chunk<DT_CA, LEAN>(ch, sl, ca, V, out, Vl, Al, Sl);
  block_sum<1, kChisqThreadsMax / 64>(v, red);
  if (threadIdx.x == 0) partial[bid] = v[0];
}

template <typename VT>
__global__ void __launch_bounds__(kChisqThreadsMax)
k_chisq(const Chunk* __restrict__ chunks, const int* __restrict__ slots,
        const FcvItem* __restrict__ fcv, const AmpItem* __restrict__ amp,
        const CpItem* __restrict__ cp, const CaItem* __restrict__ ca,
        const VT* __restrict__ V, double2* __restrict__ sres,
        double* __restrict__ partial) {
  chisq_block<VT, true>(blockIdx.x, chunks, slots, fcv, amp, cp, ca, V, sres, partial);
}

// ---------------------------------------------------------------------------
// window pixels -> 2-D image (I1d_I2d_fwd, image.f90:1167-1186).  Pixels
// outside the window stay zero from allocation.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
k_scatter(const RegItem* __restrict__ items, const UnitDev* __restrict__ units,
          const int* __restrict__ pix, const double* __restrict__ x,
          double* __restrict__ I2d) {
  const RegItem it = items[blockIdx.x];
  const UnitDev U = units[it.unit];
  if (U.identity) return;
  for (int i = threadIdx.x; i < it.count; i += blockDim.x) {
    const int k = it.start + i;
    I2d[U.img_off + pix[U.pix_off + k]] = x[U.x_off + k];
  }
}

// ---------------------------------------------------------------------------
// pixel regularizers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double smabs(double v) { return sqrt(v * v + kZeroEps); }

// moments for tfd / cen / sm (image.f90:774-1138)
__global__ void __launch_bounds__(kThreads)
k_moments(const RegItem* __restrict__ items, const UnitDev* __restrict__ units,
          const int* __restrict__ pix, const double* __restrict__ x,
          double* __restrict__ mpart) {
  __shared__ double red[kMomTerms * (kThreads / 64)];
  const RegItem it = items[blockIdx.x];
  const UnitDev U = units[it.unit];
  double v[kMomTerms];
#pragma unroll
  for (int k = 0; k < kMomTerms; ++k) v[k] = 0.0;
  if (U.need_moments) {
    const bool alpha1 = fabs(U.cen_alpha - 1.0) < kZeroEps;
    for (int i = threadIdx.x; i < it.count; i += blockDim.x) {
      const int k = it.start + i;
      const int p = U.identity ? k : pix[U.pix_off + k];
      const double xi = (double)(p % U.nx + 1), yi = (double)(p / U.nx + 1);
      const double I = x[U.x_off + k];
      v[0] += I;
      v[1] += fabs(I);
      v[2] += xi * I;
      v[3] += yi * I;
      v[4] += xi * xi * I;
      v[5] += yi * yi * I;
      v[6] += xi * yi * I;
      const double Ip = alpha1 ? smabs(I) : pow(smabs(I), U.cen_alpha);
      v[7] += Ip;
      v[8] += Ip * (xi - U.nxref);
      v[9] += Ip * (yi - U.nyref);
    }
  }
  block_sum<kMomTerms>(v, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < kMomTerms; ++k) mpart[(long long)blockIdx.x * kMomTerms + k] = v[k];
}

// one block per unit: sum its items' moments
__device__ __noinline__ void unit_moment_costs(const UnitDev& U, const UnitMoments& M, double& tfd, double& cen,
                                               double& smc);

// one block per unit: sum its items' moments, then its tfd / cen / sm costs
__global__ void k_moments_reduce(const UnitDev* __restrict__ units, const int* __restrict__ unit_reg_begin,
                                 const double* __restrict__ mpart, UnitMoments* __restrict__ mom) {
  __shared__ double m[kMomTerms];
  const int u = blockIdx.x;
  const int b = unit_reg_begin[u], e = unit_reg_begin[u + 1];
  if (threadIdx.x < kMomTerms) {
    double s = 0.0;
    for (int i = b; i < e; ++i) s += mpart[(long long)i * kMomTerms + threadIdx.x];
    m[threadIdx.x] = s;
    reinterpret_cast<double*>(&mom[u])[threadIdx.x] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    UnitMoments M{};
    double* mp = reinterpret_cast<double*>(&M);
    for (int k = 0; k < kMomTerms; ++k) mp[k] = m[k];
    double tfd = 0.0, cen = 0.0, smc = 0.0;
    if (units[u].need_moments) unit_moment_costs(units[u], M, tfd, cen, smc);
    mom[u].tfd_c = tfd;
    mom[u].cen_c = cen;
    mom[u].sm_c = smc;
  }
}

struct SmState {
  double Isum, xc, yc, Sg[3], Sp[3], mom[2];
};

__device__ __forceinline__ SmState sm_state(const UnitDev& U, const UnitMoments& M) {
  SmState s;
  s.Isum = M.A;
  s.xc = M.Sx / M.A;
  s.yc = M.Sy / M.A;
  s.Sg[0] = (M.Sxx - 2.0 * s.xc * M.Sx + s.xc * s.xc * M.S) / s.Isum;
  s.Sg[1] = (M.Syy - 2.0 * s.yc * M.Sy + s.yc * s.yc * M.S) / s.Isum;
  s.Sg[2] = (M.Sxy - s.xc * M.Sy - s.yc * M.Sx + s.xc * s.yc * M.S) / s.Isum;
  s.mom[0] = M.Sx - s.xc * M.S;
  s.mom[1] = M.Sy - s.yc * M.S;
  double sp, cp;
  sincos(U.sm_phi, &sp, &cp);
  s.Sp[0] = U.sm_min * cp * cp + U.sm_maj * sp * sp;
  s.Sp[1] = U.sm_min * sp * sp + U.sm_maj * cp * cp;
  s.Sp[2] = (U.sm_maj - U.sm_min) * sp * cp;
  return s;
}

// The pixel and the six neighbours the TV and TSV gradients read, loaded
// together; a neighbour outside the image is 0 and never used.
struct Nbr { double c, r, u, l, lu, d, rd; };   // (ix, iy) and +x, +y, -x, (-x, +y), -y, (+x, -y)

__device__ __forceinline__ Nbr load_nbr(const double* I, int nx, int ny, int ix, int iy) {
  Nbr n{};
  const long long o = (long long)iy * nx + ix;
  n.c = I[o];
  if (ix + 1 < nx) n.r = I[o + 1];
  if (iy + 1 < ny) n.u = I[o + nx];
  if (ix > 0) n.l = I[o - 1];
  if (ix > 0 && iy + 1 < ny) n.lu = I[o + nx - 1];
  if (iy > 0) n.d = I[o - nx];
  if (iy > 0 && ix + 1 < nx) n.rd = I[o - nx + 1];
  return n;
}

// tv_e (image.f90:477-508), tv_grade with the reference's reset quirk
// (image.f90:512-612) and tsv_grade (image.f90:713-754) at (ix, iy),
// 0-based, on the loaded neighbourhood
__device__ __forceinline__ void nbr_fwd_diff(const Nbr& n, int nx, int ny, int ix, int iy, double& dx, double& dy) {
  dx = (ix + 1 < nx) ? n.r - n.c : 0.0;
  dy = (iy + 1 < ny) ? n.u - n.c : 0.0;
}

__device__ __forceinline__ double nbr_tv_grade(const Nbr& n, int nx, int ny, int ix, int iy) {
  double g = 0.0, dx, dy;
  nbr_fwd_diff(n, nx, ny, ix, iy, dx, dy);
  double e = sqrt(dx * dx + dy * dy);
  if (e > kZeroEps) g -= (dx + dy) / e;
  if (ix > 0) {
    dx = n.c - n.l;
    dy = (iy + 1 < ny) ? n.lu - n.l : 0.0;
    e = sqrt(dx * dx + dy * dy);
    if (e > kZeroEps) g += dx / e; else g = 0.0;
  }
  if (iy > 0) {
    dy = n.c - n.d;
    dx = (ix + 1 < nx) ? n.rd - n.d : 0.0;
    e = sqrt(dx * dx + dy * dy);
    if (e > kZeroEps) g += dy / e; else g = 0.0;
  }
  return g;
}

__device__ __forceinline__ double nbr_tsv_grade(const Nbr& n, int nx, int ny, int ix, int iy) {
  double g = 0.0;
  if (ix + 1 < nx) g -= 2.0 * (n.r - n.c);
  if (iy + 1 < ny) g -= 2.0 * (n.u - n.c);
  if (ix > 0) g += 2.0 * (n.c - n.l);
  if (iy > 0) g += 2.0 * (n.c - n.d);
  return g;
}

// Mean normalised frame of a movie, <I>_k = sum_z (x_zk / lc_z) / nz in
// frame order (image3d.f90:344-348); one thread per window pixel,
// blockIdx.y = movie.
__global__ void __launch_bounds__(kThreads)
k_frame_mean(const MovieDev* __restrict__ movies, const double* __restrict__ x) {
  const MovieDev Mv = movies[blockIdx.y];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= Mv.npix || !(Mv.ri_l > 0)) return;
  double a = 0.0;
  for (int z = 0; z < Mv.nz; ++z) a += x[Mv.x_off + (long long)z * Mv.npix + k] / Mv.lc[z] / (double)Mv.nz;
  Mv.iavg[k] = a;
}

// Per-pixel regularizers: gradreg = reg_scale * gradient (reg_scale = 1/nz
// for a movie frame, calc_cost_reg3d image3d.f90:255) plus, for a movie
// frame, the Rt / Ri terms of calc_cost_dynamical (image3d.f90:358-388).
// It depends on x only, so it runs beside the transforms; k_t1_rows forms
// grad = gradchisq + gradreg (fftim2d.f90:553-554).
struct RegArgs {
  const RegItem* items;
  const UnitDev* units;
  const int* pix;
  const double* x;
  const double* I2dbuf;
  const UnitMoments* mom;
  const MovieDev* movies;
  double* greg;
  double* grad;
  double* rpart;
};

// Gradient of the tfd, cen and sm terms at one pixel (image.f90:111-135,
// 760-1138) from the unit's moments, for the LEAN regularizer kernel (the
// per-unit quantities are recomputed per pixel; these terms are rare).
__device__ __noinline__ double moment_grad(const UnitDev& U, const UnitMoments& M, int ix, int iy, double I) {
  double g = 0.0;
  if (U.tfd_l > 0) g += U.tfd_l * 2.0 * (M.S - U.tfd_tgtfd);
  if (U.cen_l > 0) {
    const double cen_sIp = kZeroEps + M.P;
    const double cen_sx = M.Px / cen_sIp, cen_sy = M.Py / cen_sIp;
    const double cen_cost = sqrt(cen_sx * cen_sx + cen_sy * cen_sy + kZeroEps);
    const bool alpha1 = fabs(U.cen_alpha - 1.0) < kZeroEps;
    const double dix = (ix + 1) - U.nxref, diy = (iy + 1) - U.nyref;
    const double sa = smabs(I);
    const double gIp = alpha1 ? I / sa : I / sa * U.cen_alpha * pow(sa, U.cen_alpha - 1.0);
    g += U.cen_l * (cen_sx * (dix - cen_sx) + cen_sy * (diy - cen_sy)) / cen_sIp / cen_cost * gIp;
  }
  if (U.sm_l > 0) {
    const SmState S = sm_state(U, M);
    const double xi = ix + 1, yi = iy + 1;
    const double xg = (xi - S.xc) / S.Isum, yg = (yi - S.yc) / S.Isum;
    const double g1 = (((xi - S.xc) * (xi - S.xc) - 2.0 * xg * S.mom[0]) - S.Sg[0]) / S.Isum;
    const double g2 = (((yi - S.yc) * (yi - S.yc) - 2.0 * yg * S.mom[1]) - S.Sg[1]) / S.Isum;
    const double g3 = ((xi - S.xc) * (yi - S.yc) - xg * S.mom[1] - yg * S.mom[0] - S.Sg[2]) / S.Isum;
    g += U.sm_l * (2.0 * (S.Sg[0] - S.Sp[0]) * g1 + 2.0 * (S.Sg[1] - S.Sp[1]) * g2 +
                   4.0 * (S.Sg[2] - S.Sp[2]) * g3);
  }
  return g;
}

template <bool LEAN>
__device__ __forceinline__ void reg_block_unit(int bid, const RegArgs& A, const RegItem& it, const UnitDev& U) {
  const int* __restrict__ pix = A.pix;
  const double* __restrict__ x = A.x;
  const double* __restrict__ I2dbuf = A.I2dbuf;
  const UnitMoments* __restrict__ mom = A.mom;
  const MovieDev* __restrict__ movies = A.movies;
  double* __restrict__ greg = A.greg;
  double* __restrict__ grad = A.grad;
  double* __restrict__ rpart = A.rpart;
  __shared__ double red[kRegTerms * (kChisqThreadsMax / 64)];
  const double* I2d = U.identity ? (x + U.x_off) : (I2dbuf + U.img_off);
  // (LEAN: the moments' gradient is called per pixel instead)
  const bool moments = !LEAN && U.need_moments;
  double v[kRegTerms];
#pragma unroll
  for (int k = 0; k < kRegTerms; ++k) v[k] = 0.0;
  // global quantities of tfd / cen / sm
  double tfd_g = 0.0, cen_sx = 0.0, cen_sy = 0.0, cen_sIp = 1.0, cen_cost = 1.0;
  SmState S;
  if (moments) {
    const UnitMoments M = mom[it.unit];
    if (U.tfd_l > 0) tfd_g = U.tfd_l * 2.0 * (M.S - U.tfd_tgtfd);
    if (U.cen_l > 0) {
      cen_sIp = kZeroEps + M.P;
      cen_sx = M.Px / cen_sIp;
      cen_sy = M.Py / cen_sIp;
      cen_cost = sqrt(cen_sx * cen_sx + cen_sy * cen_sy + kZeroEps);
    }
    if (U.sm_l > 0) S = sm_state(U, M);
  }
  const bool alpha1 = fabs(U.cen_alpha - 1.0) < kZeroEps;
  MovieDev Mv{};
  bool rt_on = false, ri_on = false;
  double lz = 1.0, lu = 1.0, ll = 1.0;
  if (U.movie >= 0) {
    Mv = movies[U.movie];
    lz = Mv.lc[U.frame];
    rt_on = Mv.rt_l > 0 && U.frame < Mv.nz - 1;
    ri_on = Mv.ri_l > 0;
    if (rt_on) lu = Mv.lc[U.frame + 1];
    if (rt_on && U.frame > 0) ll = Mv.lc[U.frame - 1];
  }
  // the unit's fields the pixel loop reads, in registers: read in place
  // (LEAN) they would be loaded again after every gradient store, which
  // may alias the descriptor -- a round trip to memory per field and pixel
  const struct {
    int nx, ny, x_off, pix_off, identity, m, frame, need_moments;
    const double *l1_w, *tv_w, *tsv_w, *kl_w, *gs_w;
    double cen_alpha, cen_l, gs_l, gs_u, kl_l, kl_u, l1_l, l1_u, nxref, nyref, reg_scale, sm_l, tfd_l, tsv_l, tsv_u, tv_l, tv_u;
  } L = {U.nx, U.ny, U.x_off, U.pix_off, U.identity, U.m, U.frame, U.need_moments, U.l1_w, U.tv_w, U.tsv_w, U.kl_w, U.gs_w, U.cen_alpha, U.cen_l, U.gs_l, U.gs_u, U.kl_l, U.kl_u, U.l1_l, U.l1_u, U.nxref, U.nyref, U.reg_scale, U.sm_l, U.tfd_l, U.tsv_l, U.tsv_u, U.tv_l, U.tv_u};
  for (int i = threadIdx.x; i < it.count; i += blockDim.x) {
    const int k = it.start + i;
    const int p = L.identity ? k : pix[L.pix_off + k];
    const int ix = p % L.nx, iy = p / L.nx;
    // every load of the pixel first, all in flight together (issued inside
    // each term's branch they would be one round trip to memory per term)
    const double I = x[L.x_off + k];
    double wl1 = 0.0, Pkl = 1.0, Pgs = 1.0, wtv = 0.0, wtsv = 0.0;
    if (L.l1_l > 0) wl1 = L.l1_w ? L.l1_w[k] : L.l1_u;
    if (L.kl_l > 0) Pkl = L.kl_w ? L.kl_w[k] : L.kl_u;
    if (L.gs_l > 0) Pgs = L.gs_w ? L.gs_w[k] : L.gs_u;
    if (L.tv_l > 0) wtv = L.tv_w ? L.tv_w[k] : L.tv_u;
    if (L.tsv_l > 0) wtsv = L.tsv_w ? L.tsv_w[k] : L.tsv_u;
    Nbr nb{};
    if (L.tv_l > 0 || L.tsv_l > 0) nb = load_nbr(I2d, L.nx, L.ny, ix, iy);
    double Iu = 0.0, Il = 0.0, avg = 0.0, wrt = 0.0, wri = 0.0;
    if (rt_on) {
      wrt = Mv.rt_w ? Mv.rt_w[k] : Mv.rt_u;
      Iu = x[L.x_off + Mv.npix + k];
      if (L.frame > 0) Il = x[L.x_off - Mv.npix + k];
    }
    if (ri_on) {
      wri = Mv.ri_w ? Mv.ri_w[k] : Mv.ri_u;
      avg = Mv.iavg[k];
    }
    double g = 0.0;
    if (LEAN && L.need_moments) g += moment_grad(U, mom[it.unit], ix, iy, I);
    if (moments && L.tfd_l > 0) g += tfd_g;
    if (moments && L.cen_l > 0) {
      const double dix = (ix + 1) - L.nxref, diy = (iy + 1) - L.nyref;
      const double sa = smabs(I);
      const double gIp = alpha1 ? I / sa : I / sa * L.cen_alpha * pow(sa, L.cen_alpha - 1.0);
      g += L.cen_l * (cen_sx * (dix - cen_sx) + cen_sy * (diy - cen_sy)) / cen_sIp / cen_cost * gIp;
    }
    if (L.l1_l > 0) {
      const double w = L.l1_l * wl1;
      const double sa = smabs(I);
      v[0] += w * sa;
      g += w * (I / sa);
    }
    if (L.kl_l > 0) {
      const double sa = smabs(I);
      const double lg = log_s(sa / Pkl);
      v[3] += L.kl_l * (sa * lg + Pkl / 2.71828182845904523536);
      g += L.kl_l * (I / sa) * (lg + 1.0);
    }
    if (L.gs_l > 0) {
      const double sa = smabs(I);
      const double lg = log_s(sa / Pgs);
      v[4] += L.gs_l * (sa * (lg - 1.0) + Pgs);
      g += L.gs_l * (I / sa) * lg;
    }
    if (L.tv_l > 0) {
      const double w = L.tv_l * wtv;
      double dx, dy;
      nbr_fwd_diff(nb, L.nx, L.ny, ix, iy, dx, dy);
      v[1] += w * sqrt(dx * dx + dy * dy);
      g += w * nbr_tv_grade(nb, L.nx, L.ny, ix, iy);
    }
    if (L.tsv_l > 0) {
      const double w = L.tsv_l * wtsv;
      double dx, dy;
      nbr_fwd_diff(nb, L.nx, L.ny, ix, iy, dx, dy);
      v[2] += w * (dx * dx + dy * dy);
      g += w * nbr_tsv_grade(nb, L.nx, L.ny, ix, iy);
    }
    if (moments && L.sm_l > 0) {
      const double xi = ix + 1, yi = iy + 1;
      const double xg = (xi - S.xc) / S.Isum, yg = (yi - S.yc) / S.Isum;
      const double g1 = (((xi - S.xc) * (xi - S.xc) - 2.0 * xg * S.mom[0]) - S.Sg[0]) / S.Isum;
      const double g2 = (((yi - S.yc) * (yi - S.yc) - 2.0 * yg * S.mom[1]) - S.Sg[1]) / S.Isum;
      const double g3 = ((xi - S.xc) * (yi - S.yc) - xg * S.mom[1] - yg * S.mom[0] - S.Sg[2]) / S.Isum;
      g += L.sm_l * (2.0 * (S.Sg[0] - S.Sp[0]) * g1 + 2.0 * (S.Sg[1] - S.Sp[1]) * g2 +
                     4.0 * (S.Sg[2] - S.Sp[2]) * g3);
    }
    double gt = 0.0;
    if (rt_on || ri_on) {
      const double II = I / lz;
      if (rt_on) {
        const double w = Mv.rt_l * wrt;
        const double Ju = Iu / lu;
        v[5] += w * ((Ju - II) * (Ju - II));
        double gr;
        if (L.frame == 0) {
          gr = 2.0 * (II - Ju);
        } else {
          const double Jl = Il / ll;
          gr = 2.0 * (2.0 * II - Jl - Ju);
        }
        gt += w * gr / lz;
      }
      if (ri_on) {
        const double w = Mv.ri_l * wri;
        const double d = II - avg;
        v[6] += w * (d * d);
        gt += w * (2.0 * d) / lz;
      }
    }
    // units with uv points: k_t1_rows adds the chi-square gradient; units
    // without (movie frames with no data) have none
    const double gr = L.reg_scale * g + gt;
    if (L.m > 0) greg[L.x_off + k] = gr;
    else grad[L.x_off + k] = 0.0 + gr;
  }
  block_sum<kRegTerms, kChisqThreadsMax / 64>(v, red);
  // l1 .. gs partials scaled here by reg_scale (the sums read them as is);
  // Rt and Ri are the movie's terms, not scaled
  if (threadIdx.x == 0)
    for (int k = 0; k < kRegTerms; ++k) rpart[(long long)bid * kRegTerms + k] = k < 5 ? U.reg_scale * v[k] : v[k];
}

// The unit's descriptor is read in place (LEAN) or copied to registers
// first (the fused launch, measured faster so).
template <bool LEAN>
__device__ void reg_block(int bid, const RegArgs& A) {
  const RegItem it = A.items[bid];
  if (LEAN) {
    reg_block_unit<true>(bid, A, it, A.units[it.unit]);
  } else {
    const UnitDev U = A.units[it.unit];
    reg_block_unit<false>(bid, A, it, U);
  }
}

__global__ void __launch_bounds__(kThreads) k_reg(RegArgs A) { reg_block<true>(blockIdx.x, A); }

// The chi-square chunks and the regularizer items in one launch: the two
// are independent (V vs x), so the regularizer work-groups fill the CUs the
// chunks leave idle instead of adding a launch to the dependent chain.
template <typename VT>
__global__ void __launch_bounds__(kChisqThreadsMax)
k_chisq_reg(int nchunks, const Chunk* __restrict__ chunks, const int* __restrict__ slots,
            const FcvItem* __restrict__ fcv, const AmpItem* __restrict__ amp,
            const CpItem* __restrict__ cp, const CaItem* __restrict__ ca,
            const VT* __restrict__ V, double2* __restrict__ sres,
            double* __restrict__ partial, RegArgs A) {
  if ((int)blockIdx.x < nchunks) chisq_block<VT, false>(blockIdx.x, chunks, slots, fcv, amp, cp, ca, V, sres, partial);
  else reg_block<false>(blockIdx.x - nchunks, A);
}

// ---------------------------------------------------------------------------
// fixed-order sums of one problem's partials -> its 15 terms
// ---------------------------------------------------------------------------

// tfd, cen and sm costs of one unit (image.f90:111-135), added to tfd,
// cen, smc.  Not inlined: the moments path is rare, and inlined its
// registers would count against the spreading work that shares a launch
// with the final sums.
__device__ __noinline__ void unit_moment_costs(const UnitDev& U, const UnitMoments& M, double& tfd, double& cen,
                                               double& smc) {
  if (U.tfd_l > 0) tfd += U.reg_scale * U.tfd_l * (M.S - U.tfd_tgtfd) * (M.S - U.tfd_tgtfd);
  if (U.cen_l > 0) {
    const double sIp = kZeroEps + M.P;
    const double sx = M.Px / sIp, sy = M.Py / sIp;
    cen += U.reg_scale * U.cen_l * sqrt(sx * sx + sy * sy + kZeroEps);
  }
  if (U.sm_l > 0) {
    const SmState S = sm_state(U, M);
    const double c = (S.Sg[0] - S.Sp[0]) * (S.Sg[0] - S.Sp[0]) + (S.Sg[1] - S.Sp[1]) * (S.Sg[1] - S.Sp[1]) +
                     2.0 * (S.Sg[2] - S.Sp[2]) * (S.Sg[2] - S.Sp[2]);
    smc += U.reg_scale * U.sm_l * c;
  }
}

// The problem's chunks are one contiguous range in type order (fcv, amp,
// cp, ca) and its regularizer items another; each thread sums a strided
// share of both, then one block reduction.  Every load depends only on the
// problem descriptor, so the sums are two memory round trips deep.
__device__ void finalize_problem(const ProbDev& P, const double* cpart, const double* rpart,
                                 const UnitMoments* __restrict__ mom, double* __restrict__ terms, double* red) {
  double v[14];
#pragma unroll
  for (int k = 0; k < 14; ++k) v[k] = 0.0;
  for (int i = P.chunk_begin[0] + threadIdx.x; i < P.chunk_end[3]; i += blockDim.x) {
    const int t = i < P.chunk_end[0] ? 0 : i < P.chunk_end[1] ? 1 : i < P.chunk_end[2] ? 2 : 3;
    const double c = cpart[i];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (k == t) v[k] += c;
  }
  for (int i = P.reg_begin + threadIdx.x; i < P.reg_end; i += blockDim.x) {
    const double* r = rpart + (long long)i * kRegTerms;
    for (int k = 0; k < 5; ++k) v[4 + k] += r[k];
    v[12] += r[5];
    v[13] += r[6];
  }
  // tfd / cen / sm costs per unit (k_moments_reduce)
  if (P.any_moments)
    for (int u = P.unit_begin + threadIdx.x; u < P.unit_end; u += blockDim.x) {
      v[9] += mom[u].tfd_c;
      v[10] += mom[u].cen_c;
      v[11] += mom[u].sm_c;
    }
  block_sum<14>(v, red);
  if (threadIdx.x == 0) {
    double* T = terms + P.terms_off;
    const double cf = v[0], ca_ = v[1], cc = v[2], cq = v[3];
    // order of accumulation of calc_chisq: fcv, amp, ca, cp (fftlib.f90:338-367)
    double chisq = 0.0;
    if (P.n[DT_FCV] > 0) chisq += cf;
    if (P.n[DT_AMP] > 0) chisq += ca_;
    if (P.n[DT_CA] > 0) chisq += cq;
    if (P.n[DT_CP] > 0) chisq += cc;
    T[0] = chisq;
    T[1] = P.n[DT_FCV] > 0 ? cf / P.w[DT_FCV] / (2.0 * P.n[DT_FCV]) : 0.0;
    T[2] = P.n[DT_AMP] > 0 ? ca_ / P.w[DT_AMP] / P.n[DT_AMP] : 0.0;
    T[3] = P.n[DT_CP] > 0 ? cc / P.w[DT_CP] / P.n[DT_CP] : 0.0;
    T[4] = P.n[DT_CA] > 0 ? cq / P.w[DT_CA] / P.n[DT_CA] : 0.0;
    const double l1 = v[4], tv = v[5], tsv = v[6], kl = v[7], gs = v[8];
    const double tfd = v[9], cen = v[10], smc = v[11], rt = v[12], ri = v[13];
    // image.f90:199 order, then calc_cost_dynamical (image3d.f90:390)
    const double reg = (l1 + tv + tsv + kl + gs + tfd + cen + smc) + (ri + rt);
    T[5] = reg;
    T[6] = l1; T[7] = tv; T[8] = tsv; T[9] = kl; T[10] = gs;
    T[11] = P.lc_mode ? 0.0 : tfd; T[12] = cen; T[13] = smc;
    T[14] = chisq + reg;
    if (P.nterms > 15) {                  // movie: Rt, Ri, light curve
      T[15] = rt;
      T[16] = ri;
      T[17] = P.lc_mode ? tfd : 0.0;
    }
  }
}

struct FinArgs {
  const ProbDev* probs;
  const double* cpart;
  const double* rpart;
  const UnitMoments* mom;
  double* terms;
};

// one block per problem
__global__ void __launch_bounds__(kThreads) k_finalize(FinArgs F) {
  __shared__ double red[16 * (kThreads / 64)];
  const ProbDev P = F.probs[blockIdx.x];
  finalize_problem(P, F.cpart, F.rpart, F.mom, F.terms, red);
}

// Spreading work items followed by one finalize work-group per problem: the
// sums need only the chi-square and regularizer partials, complete before
// this launch, so they overlap the spreading instead of ending the chain.
template <typename R, bool DET>
__global__ void __launch_bounds__(kThreads)
k_spread_fin(int nspread, const SpreadItem* __restrict__ items, const int* __restrict__ pt_cell,
             const R* __restrict__ wts, long long npt, ResidSrc res, typename Cx<R>::T* __restrict__ gridB,
             const int* __restrict__ slot_idx, typename Cx<R>::T* __restrict__ slots, FinArgs F) {
  __shared__ double red[16 * (kThreads / 64)];
  // the finalize work-groups first: dispatched with the launch, they
  // finish under the spreading work instead of trailing it
  const int nfin = (int)gridDim.x - nspread;
  if ((int)blockIdx.x < nfin) {
    const ProbDev P = F.probs[blockIdx.x];
    finalize_problem(P, F.cpart, F.rpart, F.mom, F.terms, red);
  } else
    spread_block<R, DET>(blockIdx.x - nfin, items, pt_cell, wts, npt, res, gridB, slot_idx, slots);
}

}  // namespace smili
