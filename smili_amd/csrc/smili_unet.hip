// Warm-start U-Net on gfx950: BF16 activations and weights in HBM (NHWC),
// FP32 accumulation on the matrix cores (v_mfma_f32_16x16x32_bf16), forward
// and backward kernels and an Adam optimiser, all on one HIP stream.
//
// The network has no counterpart in the reference; its definition (layers,
// rounding points, loss) is stated in include/smili_unet.h.
//
// Convolution kernels (implicit GEMM, 3x3, zero padding):
//  * k_conv3x3: output tile of 8x16 pixels x 32/64 output channels per
//    256-thread work-group; each 32-channel slice of the input is staged
//    once in LDS as a 10x18 halo tile and read by all 9 taps.  The fused
//    gather reads the nearest-x2 upsample of a coarser tensor and
//    concatenates a skip tensor (decoder inputs), so neither is
//    materialised.  Weights are the MFMA A operand (rows = output
//    channels), the halo the B operand (columns = pixels): each lane then
//    holds 4 consecutive channels of one pixel for the store.  Used for the
//    forward pass (bias + ReLU, bf16 out) and for the input gradient
//    (weights flipped and transposed, fp32 out).
//  * k_conv3x3_wgrad: dW[co][tap][ci] = sum_px dP[px][co] X[px+off][ci];
//    pixels are the reduction axis, so dP and the input halo are staged
//    transposed ([channel][pixel]) in LDS (8 x 8 blocks transposed in
//    registers); a tap's 8-pixel fragment, shifted by dx columns, is cut
//    from one aligned 16-byte read and the next word.
//    Work-groups split the pixel tiles; each writes its partial sums to a
//    row of a scratch array that k_sum_rows reduces (float atomics of every
//    group on the same addresses would serialise).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/smili_unet.h"

typedef uint16_t bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kBase = 32;
constexpr int kThreads = 256;
constexpr int kTileH = 8, kTileW = 16;
constexpr int kHaloH = kTileH + 2, kHaloW = kTileW + 2;

__device__ __forceinline__ float bf2f(uint32_t h) { return __uint_as_float(h << 16); }
// float -> bfloat16, round to nearest even
__device__ __forceinline__ uint32_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

// Item of a block so that consecutive items share an XCD (blocks are
// dealt round-robin over the 8 XCDs); -1 past the end.
__device__ __forceinline__ int xcd_item(int nitems) {
  const int b = blockIdx.x, per = gridDim.x >> 3;
  const int item = (b & 7) * per + (b >> 3);
  return item < nitems ? item : -1;
}
inline int round8(int n) { return (n + 7) & ~7; }

// ---------------------------------------------------------------------------
// 3x3 convolution, forward / input gradient
// ---------------------------------------------------------------------------
struct ConvArgs {
  const bf16_t* x1;  // channels [0, c1): full resolution, or half if up1
  const bf16_t* x2;  // channels [c1, c1 + c2): full resolution (may be null)
  const bf16_t* w;   // [cout][9][c1 + c2]
  const float* bias; // [cout] (mode 0)
  bf16_t* ybf;       // mode 0 / 3 output
  float* yf;         // mode 1 / 2 output
  const bf16_t* mask;  // mode 3: activation whose sign gates the result
  float* part;         // mode 3: per-tile channel sums [tile][cout]
  int c1, c2, up1;
  int n, h, w_, cout;
  int mode;          // 0: bias + ReLU -> bf16; 1: fp32 store; 2: fp32 add;
                     // 3: input gradient through the ReLU of `mask` -> bf16 dP of
                     //    the layer below, with its bias-gradient partial sums
  int tiles_x, tiles_y, ncob;
};

__device__ __forceinline__ int halo_off(int p, int g) {
  // 64 bytes per halo pixel; the 16-byte channel groups are swizzled so
  // that 8 consecutive pixels of one group (the lanes a 16-byte LDS read
  // serves together) cover all 32 banks
  return p * 32 + ((g ^ ((p >> 1) & 3)) << 3);
}

__device__ __forceinline__ const bf16_t* gather_src(const ConvArgs& a, int b, int yy, int xx, int c) {
  if (c < a.c1) {
    if (a.up1)
      return a.x1 + ((size_t)(b * (a.h >> 1) + (yy >> 1)) * (a.w_ >> 1) + (xx >> 1)) * a.c1 + c;
    return a.x1 + ((size_t)(b * a.h + yy) * a.w_ + xx) * a.c1 + c;
  }
  return a.x2 + ((size_t)(b * a.h + yy) * a.w_ + xx) * a.c2 + (c - a.c1);
}

// Work-group tile: 8 rows x TW columns of output pixels (TW = 16 or 32) x
// NCB * 16 output channels; wave w computes rows 2w, 2w + 1 (TW / 8 blocks
// of 16 pixels) for all NCB channel blocks.  Per 32-channel slice of the
// input, the halo (10 x (TW + 2) pixels) and the slice's weights
// (NCB * 16 x 9 taps x 32) sit in LDS; the next slice's halo is loaded into
// registers while the current one feeds the matrix cores (the weights,
// L2-resident, are copied at staging; a second work-group per CU covers
// that latency).
constexpr int kWStride = 9 * 32 + 8;   // bf16 per output channel in LDS: 592 bytes, bank-spread

#ifndef SMILI_UNET_TW16_MAXCIN
#define SMILI_UNET_TW16_MAXCIN (1 << 30)   // measured: 8 x 16 tiles everywhere are fastest
#endif
#ifndef SMILI_UNET_WGRAD_TW
#define SMILI_UNET_WGRAD_TW 32
#endif
#ifndef SMILI_UNET_NCB4_MIN_GROUPS
#define SMILI_UNET_NCB4_MIN_GROUPS 1536   // measured: 0 (always 64) 1.753 ms per step, 1536 1.711, all-32 1.723
#endif
#ifndef SMILI_UNET_CONV_MIN_BLOCKS
#define SMILI_UNET_CONV_MIN_BLOCKS 2
#endif
template <int TW, int NCB>
__global__ __launch_bounds__(kThreads, SMILI_UNET_CONV_MIN_BLOCKS) void k_conv3x3(ConvArgs a) {
  constexpr int HW = TW + 2;                      // halo width
  constexpr int NHALO = kHaloH * HW * 4;          // 16-byte halo pieces
  constexpr int NCO = NCB * 16;
  constexpr int NWT = NCO * 9 * 4;                // 16-byte weight pieces
  constexpr int RH = (NHALO + kThreads - 1) / kThreads;
  constexpr int RW = (NWT + kThreads - 1) / kThreads;
  constexpr int NPB = TW / 8;                     // 16-pixel blocks per wave (2 rows of TW)
  __shared__ __attribute__((aligned(16))) bf16_t halo[kHaloH * HW * 32];
  __shared__ __attribute__((aligned(16))) bf16_t wl[NCO * kWStride];
  const int item = xcd_item(a.n * a.tiles_y * a.tiles_x * a.ncob);
  if (item < 0) return;
  const int cob = item % a.ncob;
  int t = item / a.ncob;
  const int tx = t % a.tiles_x;
  t /= a.tiles_x;
  const int ty = t % a.tiles_y, b = t / a.tiles_y;
  const int y0 = ty * kTileH, x0 = tx * TW, co0 = cob * NCO;
  const int cin = a.c1 + a.c2;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ln = lane & 15, lg = lane >> 4;

  uint4 ph[RH];
  auto fetch = [&](int c0) {
#pragma unroll
    for (int r = 0; r < RH; ++r) {
      const int e = threadIdx.x + r * kThreads;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (e < NHALO) {
        const int p = e >> 2, g = e & 3;
        const int hy = p / HW, hx = p - hy * HW;
        const int yy = y0 + hy - 1, xx = x0 + hx - 1;
        if (yy >= 0 && yy < a.h && xx >= 0 && xx < a.w_)
          v = *reinterpret_cast<const uint4*>(gather_src(a, b, yy, xx, c0 + g * 8));
      }
      ph[r] = v;
    }
  };
  auto stage = [&](int c0) {
#pragma unroll
    for (int r = 0; r < RH; ++r) {
      const int e = threadIdx.x + r * kThreads;
      if (e < NHALO) *reinterpret_cast<uint4*>(halo + halo_off(e >> 2, e & 3)) = ph[r];
    }
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int e = threadIdx.x + r * kThreads;
      if (e < NWT) {
        const int q = e & 3, ct = e >> 2, co = ct / 9, tap = ct - co * 9;   // ct = co * 9 + tap
        *reinterpret_cast<uint4*>(wl + co * kWStride + tap * 32 + q * 8) =
            *reinterpret_cast<const uint4*>(a.w + ((size_t)(co0 * 9 + ct)) * cin + c0 + q * 8);
      }
    }
  };

  f32x4 acc[NPB][NCB];
#pragma unroll
  for (int i = 0; i < NPB; ++i)
#pragma unroll
    for (int j = 0; j < NCB; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  fetch(0);
  for (int c0 = 0; c0 < cin; c0 += 32) {
    __syncthreads();
    stage(c0);
    __syncthreads();
    if (c0 + 32 < cin) fetch(c0 + 32);
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int dy = tap / 3, dx = tap - 3 * (tap / 3);
      bf16x8 aw[NCB];
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb)
        aw[cb] = *reinterpret_cast<const bf16x8*>(wl + (cb * 16 + ln) * kWStride + tap * 32 + lg * 8);
#pragma unroll
      for (int pb = 0; pb < NPB; ++pb) {
        // block pb: row 2 wv + pb / (TW / 16), columns 16 (pb % (TW / 16)) ..
        const int row = wv * 2 + pb / (TW / 16), col = (pb % (TW / 16)) * 16;
        const int p = (row + dy) * HW + col + ln + dx;
        const bf16x8 bx = *reinterpret_cast<const bf16x8*>(halo + halo_off(p, lg));
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb)
          acc[pb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aw[cb], bx, acc[pb][cb], 0, 0, 0);
      }
    }
  }
  // D[row = channel 4*lg + i][col = pixel ln]
  float csum[NCB][4];
#pragma unroll
  for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
    for (int i = 0; i < 4; ++i) csum[cb][i] = 0.f;
#pragma unroll
  for (int pb = 0; pb < NPB; ++pb) {
    const int y = y0 + wv * 2 + pb / (TW / 16), x = x0 + (pb % (TW / 16)) * 16 + ln;
    const size_t o = ((size_t)(b * a.h + y) * a.w_ + x) * a.cout;
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
      const int co = co0 + cb * 16 + lg * 4;
      const f32x4 v = acc[pb][cb];
      if (a.mode == 3) {
        const uint2 m = *reinterpret_cast<const uint2*>(a.mask + o + co);
        const uint32_t mw[4] = {m.x & 0xFFFFu, m.x >> 16, m.y & 0xFFFFu, m.y >> 16};
        float r[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          r[i] = bf2f(mw[i]) > 0.f ? v[i] : 0.f;
          csum[cb][i] += r[i];
        }
        uint2 pk;
        pk.x = f2bf(r[0]) | (f2bf(r[1]) << 16);
        pk.y = f2bf(r[2]) | (f2bf(r[3]) << 16);
        *reinterpret_cast<uint2*>(a.ybf + o + co) = pk;
      } else if (a.mode == 0) {
        float r[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) r[i] = fmaxf(v[i] + a.bias[co + i], 0.f);
        uint2 pk;
        pk.x = f2bf(r[0]) | (f2bf(r[1]) << 16);
        pk.y = f2bf(r[2]) | (f2bf(r[3]) << 16);
        *reinterpret_cast<uint2*>(a.ybf + o + co) = pk;
      } else {
        float4* dst = reinterpret_cast<float4*>(a.yf + o + co);
        float4 r = make_float4(v[0], v[1], v[2], v[3]);
        if (a.mode == 2) {
          const float4 q = *dst;
          r.x += q.x; r.y += q.y; r.z += q.z; r.w += q.w;
        }
        *dst = r;
      }
    }
  }
  if (a.mode == 3) {
    // channel sums over the tile: the 16 lanes of a channel group hold
    // different pixels, the 4 waves different rows
    __shared__ float red[4][NCO];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float s = csum[cb][i];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        s += __shfl_xor(s, 4);
        s += __shfl_xor(s, 8);
        if (ln == 0) red[wv][cb * 16 + lg * 4 + i] = s;
      }
    __syncthreads();
    if (threadIdx.x < NCO) {
      const float s = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
      a.part[(size_t)(item / a.ncob) * a.cout + co0 + threadIdx.x] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// 3x3 convolution, weight gradient
// ---------------------------------------------------------------------------
// 8 x 8 transpose of bfloat16: v[j] holds channels 0..7 of pixel j, w[i]
// receives pixels 0..7 of channel i
__device__ __forceinline__ void transpose8(const uint4 (&v)[8], uint4 (&w)[8]) {
  uint32_t s[8][4];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    s[j][0] = v[j].x; s[j][1] = v[j].y; s[j][2] = v[j].z; s[j][3] = v[j].w;
  }
  uint32_t o[8][4];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t lo = s[2 * k][i >> 1], hi = s[2 * k + 1][i >> 1];
      o[i][k] = (i & 1) ? ((lo >> 16) | (hi & 0xFFFF0000u)) : ((lo & 0xFFFFu) | (hi << 16));
    }
#pragma unroll
  for (int i = 0; i < 8; ++i) w[i] = make_uint4(o[i][0], o[i][1], o[i][2], o[i][3]);
}

struct WgradArgs {
  const bf16_t* dp;  // [n][h][w][cout]
  const bf16_t* x1;
  const bf16_t* x2;
  float* dw;         // [ngroups][cout][9][c1 + c2]: one row of partial sums per pixel group
  int c1, c2, up1;
  int n, h, w_, cout;
  int tiles_x, tiles_y, ncob, nchunk, tpg, ngroups;
};

// Tile: 8 rows x TW columns (TW = 16 or 32); 8 TW / 32 k-steps of 32
// pixels per staging round.
template <int NCB, int TW>
__global__ __launch_bounds__(kThreads) void k_conv3x3_wgrad(WgradArgs a) {
  constexpr int SXW = TW + 8;        // halo row in LDS: TW + 2 columns padded to a multiple of 8
  constexpr int NCG = SXW / 8;       // 8-column groups of a halo row
  // rows padded by 16 bytes so the 8 channel rows a 16-byte fragment read
  // serves together start on different banks (strides of 4 or 12 dwords
  // modulo 32)
  constexpr int SDP = kTileH * TW + 8;          // bf16 per dP channel row
  constexpr int SXC = kHaloH * SXW + 8;         // bf16 per input channel
  __shared__ __attribute__((aligned(16))) bf16_t sdp_[NCB * 16 * SDP];
  __shared__ __attribute__((aligned(16))) bf16_t sx_[32 * SXC];
  auto sdp = [&](int c, int px) -> bf16_t* { return sdp_ + c * SDP + px; };
  auto sx = [&](int c, int hy, int col) -> bf16_t* { return sx_ + c * SXC + hy * SXW + col; };
  const int item = xcd_item(a.ncob * a.nchunk * a.ngroups);
  if (item < 0) return;
  const int grp = item % a.ngroups;
  const int rest = item / a.ngroups;
  const int chunk = rest % a.nchunk, cob = rest / a.nchunk;
  const int co0 = cob * NCB * 16, ci0 = chunk * 32;
  const int cin = a.c1 + a.c2;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ln = lane & 15, lg = lane >> 4;
  const int cib = wv & 1, tap0 = wv >> 1;   // this wave: ci block, taps tap0 + 2j
  const int ntap = tap0 == 0 ? 5 : 4;
  ConvArgs g;
  g.x1 = a.x1; g.x2 = a.x2; g.c1 = a.c1; g.c2 = a.c2; g.up1 = a.up1;
  g.h = a.h; g.w_ = a.w_;

  f32x4 acc[5][NCB];
#pragma unroll
  for (int i = 0; i < 5; ++i)
#pragma unroll
    for (int j = 0; j < NCB; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  const int ntiles = a.n * a.tiles_y * a.tiles_x;
  const int tb = grp * a.tpg, te = min(ntiles, tb + a.tpg);
  for (int tile = tb; tile < te; ++tile) {
    const int tx = tile % a.tiles_x;
    const int ty = (tile / a.tiles_x) % a.tiles_y, b = tile / (a.tiles_x * a.tiles_y);
    const int y0 = ty * kTileH, x0 = tx * TW;
    __syncthreads();
    // dP tile, transposed to [co][px]: a unit is 8 pixels (pixel group pg:
    // row pg / (TW / 8), columns 8 (pg % (TW / 8)) .. + 7) x 8 channels,
    // transposed in registers
    for (int u = threadIdx.x; u < TW * NCB * 2; u += kThreads) {
      const int pg = u % TW, q = u / TW;
      const int y = y0 + pg / (TW / 8), xb = x0 + (pg % (TW / 8)) * 8;
      uint4 v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        v[j] = *reinterpret_cast<const uint4*>(a.dp + ((size_t)(b * a.h + y) * a.w_ + xb + j) * a.cout + co0 + q * 8);
      uint4 w[8];
      transpose8(v, w);
#pragma unroll
      for (int i = 0; i < 8; ++i) *reinterpret_cast<uint4*>(sdp(q * 8 + i, pg * 8)) = w[i];
    }
    // input halo slice, transposed to [ci][row][col]: a unit is (halo row,
    // 8 columns, 8 channels); columns TW + 2 .. SXW - 1 are padding.
    // Consecutive lanes take consecutive column groups, then rows, so the
    // 16-byte stores of 8 lanes cover all banks.
    for (int u = threadIdx.x; u < kHaloH * NCG * 4; u += kThreads) {
      const int cg = u % NCG, hy = (u / NCG) % kHaloH, q = u / (NCG * kHaloH);
      const int yy = y0 + hy - 1;
      uint4 v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int hx = cg * 8 + j, xx = x0 + hx - 1;
        v[j] = make_uint4(0, 0, 0, 0);
        if (hx < TW + 2 && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w_)
          v[j] = *reinterpret_cast<const uint4*>(gather_src(g, b, yy, xx, ci0 + q * 8));
      }
      uint4 w[8];
      transpose8(v, w);
#pragma unroll
      for (int i = 0; i < 8; ++i) *reinterpret_cast<uint4*>(sx(q * 8 + i, hy, cg * 8)) = w[i];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kTileH * TW / 32; ++s) {
      // k = 32 pixels: group lg covers pixels 32 s + 8 lg .. + 7 of the tile
      const int p0 = 32 * s + 8 * lg, row = p0 / TW, cb0 = p0 % TW;
      bf16x8 ad[NCB];
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb)
        ad[cb] = *reinterpret_cast<const bf16x8*>(sdp(cb * 16 + ln, p0));
#pragma unroll
      for (int j = 0; j < 5; ++j) {
        if (j < ntap) {
          const int tap = tap0 + 2 * j;
          const int dy = tap / 3, dx = tap - 3 * (tap / 3);
          // columns cb0 + dx .. cb0 + dx + 7 from the aligned words at cb0
          const bf16_t* rp = sx(cib * 16 + ln, row + dy, cb0);
          const uint4 s0 = *reinterpret_cast<const uint4*>(rp);
          const uint32_t s4 = *reinterpret_cast<const uint32_t*>(rp + 8);
          uint4 sel;
          if (dx == 0) {
            sel = s0;
          } else if (dx == 2) {
            sel = make_uint4(s0.y, s0.z, s0.w, s4);
          } else {
            sel = make_uint4((s0.x >> 16) | (s0.y << 16), (s0.y >> 16) | (s0.z << 16),
                             (s0.z >> 16) | (s0.w << 16), (s0.w >> 16) | (s4 << 16));
          }
          const bf16x8 bx = __builtin_bit_cast(bf16x8, sel);
#pragma unroll
          for (int cb = 0; cb < NCB; ++cb)
            acc[j][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ad[cb], bx, acc[j][cb], 0, 0, 0);
        }
      }
    }
  }
  // D[row = co 4*lg + i][col = ci ln]
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    if (j < ntap) {
      const int tap = tap0 + 2 * j;
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int co = co0 + cb * 16 + lg * 4 + i, ci = ci0 + cib * 16 + ln;
          a.dw[(size_t)grp * a.cout * 9 * cin + ((size_t)co * 9 + tap) * cin + ci] = acc[j][cb][i];
        }
    }
  }
}

// ---------------------------------------------------------------------------
// first convolution (2 -> 32 channels, fp32 input and weights)
// ---------------------------------------------------------------------------
#ifndef SMILI_UNET_CONVIN_PRELOAD
#define SMILI_UNET_CONVIN_PRELOAD 1
#endif
__global__ __launch_bounds__(kThreads) void k_conv_in(const float* __restrict__ x, const float* __restrict__ w,
                                                      const float* __restrict__ bias, bf16_t* __restrict__ y,
                                                      int n, int h, int wd) {
  // one thread per (pixel, group of 8 output channels); weights transposed
  // in LDS to [tap * 2 + ci][co] so the 8 channels of one input value are
  // two 16-byte reads
  __shared__ __attribute__((aligned(16))) float sw[18 * kBase];
  __shared__ float sb[kBase];
  for (int i = threadIdx.x; i < kBase * 18; i += kThreads) sw[(i % 18) * kBase + i / 18] = w[i];
  for (int i = threadIdx.x; i < kBase; i += kThreads) sb[i] = bias[i];
  __syncthreads();
  const size_t work = (size_t)n * h * wd * (kBase / 8);
  for (size_t e = (size_t)blockIdx.x * kThreads + threadIdx.x; e < work; e += (size_t)gridDim.x * kThreads) {
    const int q = e % (kBase / 8);
    const size_t p = e / (kBase / 8);
    const int xx = p % wd, yy = (p / wd) % h, b = p / ((size_t)wd * h);
    float acc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = sb[q * 8 + k];
#if SMILI_UNET_CONVIN_PRELOAD
    // the 9 input pixels are loaded together (one latency, not nine); the
    // weight offset is made opaque so that the 144 weights are read from LDS
    // per pixel rather than hoisted into registers for the whole loop
    float2 vin[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int sy = yy + t / 3 - 1, sx = xx + t % 3 - 1;
      const bool ok = sy >= 0 && sy < h && sx >= 0 && sx < wd;
      vin[t] = ok ? *reinterpret_cast<const float2*>(x + ((size_t)(b * h + sy) * wd + sx) * 2) : make_float2(0.f, 0.f);
    }
    int wq = q * 8;
    asm volatile("" : "+v"(wq));
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const float2 v = vin[t];
#else
    const int wq = q * 8;
#pragma unroll 1
    for (int t = 0; t < 9; ++t) {
      const int sy = yy + t / 3 - 1, sx = xx + t % 3 - 1;
      const bool ok = sy >= 0 && sy < h && sx >= 0 && sx < wd;
      const float2 v = ok ? *reinterpret_cast<const float2*>(x + ((size_t)(b * h + sy) * wd + sx) * 2)
                          : make_float2(0.f, 0.f);
#endif
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const float xv = c ? v.y : v.x;
        const float4 w0 = *reinterpret_cast<const float4*>(&sw[(2 * t + c) * kBase + wq]);
        const float4 w1 = *reinterpret_cast<const float4*>(&sw[(2 * t + c) * kBase + wq + 4]);
        acc[0] = fmaf(w0.x, xv, acc[0]);
        acc[1] = fmaf(w0.y, xv, acc[1]);
        acc[2] = fmaf(w0.z, xv, acc[2]);
        acc[3] = fmaf(w0.w, xv, acc[3]);
        acc[4] = fmaf(w1.x, xv, acc[4]);
        acc[5] = fmaf(w1.y, xv, acc[5]);
        acc[6] = fmaf(w1.z, xv, acc[6]);
        acc[7] = fmaf(w1.w, xv, acc[7]);
      }
    }
    uint32_t out[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = f2bf(fmaxf(acc[2 * k], 0.f)) | (f2bf(fmaxf(acc[2 * k + 1], 0.f)) << 16);
    *reinterpret_cast<uint4*>(y + p * kBase + q * 8) = make_uint4(out[0], out[1], out[2], out[3]);
  }
}

// dW[co][tap][ci] of the first convolution over the block's pixels,
// straight from global memory: thread t takes the 4 output channels
// 4 ((t % 16) >> 1) and input channel t % 2 for all 9 taps (36 sums) over
// pixels t / 16, + 16, ... of the block's range (the 16 threads of a pixel
// share its loads through L1); the 16 pixel lanes of each (channels, input)
// pair are reduced with two shuffles and over the waves in LDS, and the
// block writes part[block][576], reduced by k_sum_rows.
__global__ __launch_bounds__(kThreads) void k_conv_in_wgrad(const bf16_t* __restrict__ dp,
                                                            const float* __restrict__ x, float* __restrict__ part,
                                                            int n, int h, int wd, int px_per_block) {
  static_assert(kBase == 32 && kThreads % 64 == 0, "16 (channels, input) pairs per pixel");
  __shared__ float red[kThreads / 64][16][37];
  const int pair = threadIdx.x & 15, cg = pair >> 1, ci = pair & 1;
  float acc[4][9];
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int tp = 0; tp < 9; ++tp) acc[k][tp] = 0.f;
  const size_t npx = (size_t)n * h * wd;
  const size_t pb = (size_t)blockIdx.x * px_per_block, pe = min(npx, pb + px_per_block);
  for (size_t p = pb + (threadIdx.x >> 4); p < pe; p += kThreads / 16) {
    const uint2 g2 = *reinterpret_cast<const uint2*>(dp + p * kBase + cg * 4);
    const float g[4] = {bf2f(g2.x & 0xFFFFu), bf2f(g2.x >> 16), bf2f(g2.y & 0xFFFFu), bf2f(g2.y >> 16)};
    const int xx = p % wd, yy = (p / wd) % h, b = p / ((size_t)wd * h);
#pragma unroll
    for (int tp = 0; tp < 9; ++tp) {
      const int sy = yy + tp / 3 - 1, sx = xx + tp % 3 - 1;
      const bool ok = sy >= 0 && sy < h && sx >= 0 && sx < wd;
      const float v = ok ? x[((size_t)(b * h + sy) * wd + sx) * 2 + ci] : 0.f;
#pragma unroll
      for (int k = 0; k < 4; ++k) acc[k][tp] = fmaf(g[k], v, acc[k][tp]);
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int tp = 0; tp < 9; ++tp) {
      float v = acc[k][tp];
      v += __shfl_xor(v, 16);
      v += __shfl_xor(v, 32);
      if (lane < 16) red[wv][pair][k * 9 + tp] = v;
    }
  __syncthreads();
  for (int i = threadIdx.x; i < kBase * 18; i += kThreads) {
    // i = co * 18 + tap * 2 + ci
    const int co = i / 18, j = i - co * 18, tp = j >> 1, c = j & 1;
    const int pr = (co >> 2) * 2 + c, r = (co & 3) * 9 + tp;
    float s = 0.f;
    for (int w = 0; w < kThreads / 64; ++w) s += red[w][pr][r];
    part[(size_t)blockIdx.x * kBase * 18 + i] = s;
  }
}

// ---------------------------------------------------------------------------
// elementwise kernels
// ---------------------------------------------------------------------------
// 2x2 average pool, bf16 -> bf16 (8 channels per thread)
__global__ void k_pool(const bf16_t* __restrict__ x, bf16_t* __restrict__ y, int n, int h, int wd, int c) {
  const int ho = h >> 1, wo = wd >> 1, cg = c >> 3;
  const size_t total = (size_t)n * ho * wo * cg;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int q = e % cg;
    const size_t p = e / cg;
    const int xo = p % wo, yo = (p / wo) % ho, b = p / ((size_t)wo * ho);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int yy = 2 * yo + (k >> 1), xx = 2 * xo + (k & 1);
      const uint4 v = *reinterpret_cast<const uint4*>(x + ((size_t)(b * h + yy) * wd + xx) * c + q * 8);
      const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int i = 0; i < 8; ++i) s[i] += bf2f((w4[i >> 1] >> ((i & 1) * 16)) & 0xFFFFu);
    }
    uint32_t o[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) o[i] = f2bf(0.25f * s[2 * i]) | (f2bf(0.25f * s[2 * i + 1]) << 16);
    *reinterpret_cast<uint4*>(y + p * c + q * 8) = make_uint4(o[0], o[1], o[2], o[3]);
  }
}

// Partial sums of a bias gradient: each thread holds s[4] for its channels
// 4 q .. 4 q + 3 (q = threadIdx.x % cg, fixed because the grid stride is a
// multiple of cg); the block's row part[blockIdx.x][c] (c = 4 cg) sums the
// threads of each q.  Every block writes its row.
__device__ __forceinline__ void block_row_sums(const float* s, float* part, int cg) {
  __shared__ float red[kThreads][5];
#pragma unroll
  for (int i = 0; i < 4; ++i) red[threadIdx.x][i] = s[i];
  __syncthreads();
  if ((int)threadIdx.x < cg) {
    float t[4] = {0.f, 0.f, 0.f, 0.f};
    for (int r = threadIdx.x; r < kThreads; r += cg)
#pragma unroll
      for (int i = 0; i < 4; ++i) t[i] += red[r][i];
#pragma unroll
    for (int i = 0; i < 4; ++i) part[(size_t)blockIdx.x * 4 * cg + 4 * threadIdx.x + i] = t[i];
  }
}

// (dA * (Y > 0)) of 4 channels: bf16 into *dst, fp32 added to s
__device__ __forceinline__ void mask4(float4 g, uint2 y, uint2* dst, float* s) {
  const float gv[4] = {g.x, g.y, g.z, g.w};
  const uint32_t yw[4] = {y.x & 0xFFFFu, y.x >> 16, y.y & 0xFFFFu, y.y >> 16};
  float r[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    r[i] = bf2f(yw[i]) > 0.f ? gv[i] : 0.f;
    s[i] += r[i];
  }
  *dst = make_uint2(f2bf(r[0]) | (f2bf(r[1]) << 16), f2bf(r[2]) | (f2bf(r[3]) << 16));
}

// Gradient of a pooled activation X (the output of a layer, bf16 y): its
// skip gradient dskip (fp32) plus the pooling adjoint dY[y][x] / 4 of every
// 2x2 block, through the layer's ReLU (y > 0) into the layer's dP (bf16), and
// the per-block partial sums part[block][c] of its bias gradient.  4 channels
// per thread.
__global__ __launch_bounds__(kThreads) void k_pool_bwd_mask(const float* __restrict__ dy,
                                                            const float* __restrict__ dskip,
                                                            const bf16_t* __restrict__ y, bf16_t* __restrict__ dp,
                                                            float* __restrict__ part, int n, int h, int wd, int c) {
  const int ho = h >> 1, wo = wd >> 1, cg = c >> 2;
  const size_t total = (size_t)n * ho * wo * cg;
  float s[4] = {0.f, 0.f, 0.f, 0.f};
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int q = e % cg;
    const size_t p = e / cg;
    const int xo = p % wo, yo = (p / wo) % ho, b = p / ((size_t)wo * ho);
    const float4 g = *reinterpret_cast<const float4*>(dy + p * c + q * 4);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int yy = 2 * yo + (k >> 1), xx = 2 * xo + (k & 1);
      const size_t o = ((size_t)(b * h + yy) * wd + xx) * c + q * 4;
      float4 a = *reinterpret_cast<const float4*>(dskip + o);
      a.x += 0.25f * g.x; a.y += 0.25f * g.y; a.z += 0.25f * g.z; a.w += 0.25f * g.w;
      mask4(a, *reinterpret_cast<const uint2*>(y + o), reinterpret_cast<uint2*>(dp + o), s);
    }
  }
  block_row_sums(s, part, cg);
}

// Adjoint of the decoder input [up(x1) ++ x2] (dcat fp32, c1 + c2
// channels): the x1 part sums each 2x2 block (nearest upsampling) and goes
// through x1's ReLU (y1 > 0) into the dP of x1's layer (bf16 dp1) with the
// per-block partial sums part[block][c1] of its bias gradient; the x2 part
// (the skip connection) is copied to dx2 (fp32; the pooling adjoint adds to
// it later).  4 channels per thread.
__global__ __launch_bounds__(kThreads) void k_split_up(const float* __restrict__ dcat, const bf16_t* __restrict__ y1,
                                                       bf16_t* __restrict__ dp1, float* __restrict__ part,
                                                       float* __restrict__ dx2, int n, int h, int wd, int c1, int c2) {
  const int ho = h >> 1, wo = wd >> 1, ct = c1 + c2;
  const size_t n1 = (size_t)n * ho * wo * (c1 >> 2), n2 = (size_t)n * h * wd * (c2 >> 2);
  float s[4] = {0.f, 0.f, 0.f, 0.f};
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n1 + n2; e += (size_t)gridDim.x * blockDim.x) {
    if (e < n1) {
      const int q = e % (c1 >> 2);
      const size_t p = e / (c1 >> 2);
      const int xo = p % wo, yo = (p / wo) % ho, b = p / ((size_t)wo * ho);
      float4 a = make_float4(0, 0, 0, 0);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int yy = 2 * yo + (k >> 1), xx = 2 * xo + (k & 1);
        const float4 v = *reinterpret_cast<const float4*>(dcat + ((size_t)(b * h + yy) * wd + xx) * ct + q * 4);
        a.x += v.x; a.y += v.y; a.z += v.z; a.w += v.w;
      }
      mask4(a, *reinterpret_cast<const uint2*>(y1 + p * c1 + q * 4), reinterpret_cast<uint2*>(dp1 + p * c1 + q * 4),
            s);
    } else {
      const size_t f = e - n1;
      const int q = f % (c2 >> 2);
      const size_t p = f / (c2 >> 2);
      *reinterpret_cast<float4*>(dx2 + p * c2 + q * 4) =
          *reinterpret_cast<const float4*>(dcat + p * ct + c1 + q * 4);
    }
  }
  block_row_sums(s, part, c1 >> 2);
}


// out[j] += sum_r part[r][j]: a block takes CB columns (consecutive
// threads, coalesced), its 256 / CB row slices sum rows r = slice,
// slice + 256 / CB, ... and are combined in LDS
template <int CB>
__global__ __launch_bounds__(256) void k_sum_rows(const float* __restrict__ part, int nrow, int ncol,
                                                  float* __restrict__ out) {
  constexpr int NS = 256 / CB;
  __shared__ float red[NS][CB];
  const int c = threadIdx.x % CB, sl = threadIdx.x / CB;
  const size_t j = (size_t)blockIdx.x * CB + c;
  float s = 0.f;
  if (j < (size_t)ncol) {
#pragma unroll 8
    for (int r = sl; r < nrow; r += NS) s += part[(size_t)r * ncol + j];
  }
  red[sl][c] = s;
  __syncthreads();
  if (threadIdx.x < CB && j < (size_t)ncol) {
    float v = 0.f;
    for (int k = 0; k < NS; ++k) v += red[k][threadIdx.x];
    out[j] += v;
  }
}

// Head z = w . d0 + b.  Four threads per pixel, 8 channels each (the dot
// product is finished with two lane shuffles).  Inference: out = max(z, 0).
// Training (tgt != null): dz = 2 (z - t) / npx; the gradient dz * w of the
// last 3x3 layer's output goes through that layer's ReLU (d0 > 0) straight
// into its dP (bf16) and per-block sums of its bias gradient part[block][c]
// (fp32, before rounding); the head's own gradient (33 values) and the loss
// are reduced per block and added atomically.
__global__ __launch_bounds__(kThreads) void k_head(const bf16_t* __restrict__ d0, const float* __restrict__ hw,
                                                   float* __restrict__ out, const float* __restrict__ tgt,
                                                   bf16_t* __restrict__ dp, float* __restrict__ part,
                                                   float* __restrict__ ghw, double* __restrict__ loss, size_t npx,
                                                   float scale) {
  static_assert(kBase == 32, "four threads of 8 channels per pixel");
  __shared__ float red[kThreads / 64][2 * kBase + 1];
  __shared__ double lred[kThreads / 64];
  const int g = threadIdx.x & 3, c0 = 8 * g;
  float sw[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) sw[i] = hw[c0 + i];
  const float hb = hw[kBase];
  float gw[8], bs[8], gb = 0.f;
#pragma unroll
  for (int i = 0; i < 8; ++i) gw[i] = bs[i] = 0.f;
  double lsum = 0.0;
  const size_t work = npx * 4;
  for (size_t e = (size_t)blockIdx.x * kThreads + threadIdx.x; e < work; e += (size_t)gridDim.x * kThreads) {
    const size_t p = e >> 2;
    const uint4 v = *reinterpret_cast<const uint4*>(d0 + p * kBase + c0);
    const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
    float a[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = bf2f((w4[i >> 1] >> ((i & 1) * 16)) & 0xFFFFu);
    float z = 0.f;
#pragma unroll
    for (int i = 0; i < 8; ++i) z = fmaf(sw[i], a[i], z);
    z += __shfl_xor(z, 1);
    z += __shfl_xor(z, 2);
    z += hb;
    if (out && g == 0) out[p] = fmaxf(z, 0.f);
    if (tgt) {
      const float diff = z - tgt[p];
      if (g == 0) {
        lsum += (double)diff * diff;
      }
      const float dz = 2.f * diff * scale;
      uint32_t o4[4];
#pragma unroll
      for (int i = 0; i < 8; i += 2) {
        const float g0 = a[i] > 0.f ? dz * sw[i] : 0.f;
        const float g1 = a[i + 1] > 0.f ? dz * sw[i + 1] : 0.f;
        bs[i] += g0;
        bs[i + 1] += g1;
        o4[i >> 1] = f2bf(g0) | (f2bf(g1) << 16);
      }
      *reinterpret_cast<uint4*>(dp + p * kBase + c0) = make_uint4(o4[0], o4[1], o4[2], o4[3]);
#pragma unroll
      for (int i = 0; i < 8; ++i) gw[i] = fmaf(dz, a[i], gw[i]);
      if (g == 0) gb += dz;
    }
  }
  if (!tgt) return;
  // sums over the lanes of one channel group (lane % 4), then over waves
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    float x = gw[i], y = bs[i];
#pragma unroll
    for (int off = 4; off < 64; off <<= 1) {
      x += __shfl_xor(x, off);
      y += __shfl_xor(y, off);
    }
    if (lane < 4) {
      red[wv][c0 + i] = x;
      red[wv][kBase + 1 + c0 + i] = y;
    }
  }
  float xb = gb;
  double lv = lsum;
  for (int off = 32; off > 0; off >>= 1) {
    xb += __shfl_xor(xb, off);
    lv += __shfl_xor(lv, off);
  }
  if (lane == 0) {
    red[wv][kBase] = xb;
    lred[wv] = lv;
  }
  __syncthreads();
  if (threadIdx.x < 2 * kBase + 1) {
    float x = 0.f;
    for (int w = 0; w < kThreads / 64; ++w) x += red[w][threadIdx.x];
    if (threadIdx.x <= kBase)
      unsafeAtomicAdd(ghw + threadIdx.x, x);
    else
      part[(size_t)blockIdx.x * kBase + threadIdx.x - kBase - 1] = x;
  }
  if (threadIdx.x == 0) {
    double x = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) x += lred[w];
    unsafeAtomicAdd(loss, x);
  }
}

__global__ void k_fill(float* __restrict__ p, size_t n, float v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_d(double* __restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

// Adam with decoupled weight decay over the flat parameter vector
__global__ void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, size_t n, const float* __restrict__ lr_arr, int s,
                       float b1, float b2, float eps, float wd, float bc1, float bc2) {
  const float lr = lr_arr[s];
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / bc1, vh = vi / bc2;
    p[i] = p[i] - lr * (mh / (sqrtf(vh) + eps) + wd * p[i]);
  }
}

// bf16 copies of the 3x3 weight tensors (layer blockIdx.y + 1 of the
// table): W[co][tap][ci] and the flipped, transposed Wf[ci][8 - tap][co]
// the input gradient convolves with.
struct CastTable {
  int n;
  int cout[16], cin[16];
  long long src[16], dst[16];   // offsets into the fp32 parameters / the bf16 arenas
};
__global__ void k_cast_w(const float* __restrict__ params, bf16_t* __restrict__ wbf, bf16_t* __restrict__ wfl,
                         CastTable tab) {
  const int l = blockIdx.y;
  const int cout = tab.cout[l], cin = tab.cin[l];
  const int total = cout * 9 * cin;
  const float* src = params + tab.src[l];
  bf16_t* w = wbf + tab.dst[l];
  bf16_t* wf = wfl + tab.dst[l];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const int ci = i % cin, tap = (i / cin) % 9, co = i / (9 * cin);
    const bf16_t hv = (bf16_t)f2bf(src[i]);
    w[i] = hv;
    wf[((size_t)ci * 9 + (8 - tap)) * cout + co] = hv;
  }
}

// Batch gather from the device-resident training set.  idx = 8 sample + k
// with k a symmetry of the square (k >= 4: transpose, then k % 4 quarter
// turns, out[i][j] = in[j][n-1-i] per turn; square images when k != 0),
// applied to the input and the target alike.
__global__ void k_gather(const float* __restrict__ din, const float* __restrict__ dtg,
                         const int64_t* __restrict__ idx, float* __restrict__ xin, float* __restrict__ tgt,
                         int nb, int h, int wd) {
  const int npx = h * wd;
  const size_t total = (size_t)nb * npx;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int b = e / npx, p = e - (size_t)b * npx;
    const int64_t s = idx[b] >> 3;
    const int k = (int)(idx[b] & 7);
    int i = p / wd, j = p - (p / wd) * wd;
    for (int r = 0; r < (k & 3); ++r) {  // undo the quarter turns
      const int ni = j, nj = h - 1 - i;
      i = ni;
      j = nj;
    }
    if (k >= 4) {
      const int q = i;
      i = j;
      j = q;
    }
    const size_t src = (size_t)s * npx + (size_t)i * wd + j;
    reinterpret_cast<float2*>(xin)[e] = reinterpret_cast<const float2*>(din)[src];
    tgt[e] = dtg[src];
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
thread_local std::string g_err;
int fail(int code, const std::string& m) {
  g_err = m;
  return code;
}
#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) return fail(-2, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct Layer {
  const char* name;
  int cin, cout;
  size_t woff, boff;     // into the flat parameters
  size_t bfoff;          // into the bf16 weight arena (3x3 convs except conv_in)
};

struct Act {
  bf16_t* y = nullptr;   // activation (bf16)
  float* g = nullptr;    // its gradient (fp32)
  int h = 0, w = 0, c = 0;
};

int layer_count() { return 14; }

void layer_table(std::vector<Layer>& L) {
  const int c0 = kBase, c1 = 2 * kBase, c2 = 4 * kBase, c3 = 8 * kBase;
  const struct { const char* n; int ci, co; } spec[14] = {
      {"conv_in", 2, c0},     {"enc0b", c0, c0},    {"enc1a", c0, c1},    {"enc1b", c1, c1},
      {"enc2a", c1, c2},      {"enc2b", c2, c2},    {"enc3a", c2, c3},    {"enc3b", c3, c3},
      {"dec2a", c3 + c2, c2}, {"dec2b", c2, c2},    {"dec1a", c2 + c1, c1}, {"dec1b", c1, c1},
      {"dec0a", c1 + c0, c0}, {"dec0b", c0, c0}};
  L.clear();
  size_t o = 0, bo = 0;
  for (int i = 0; i < 14; ++i) {
    Layer l;
    l.name = spec[i].n;
    l.cin = spec[i].ci;
    l.cout = spec[i].co;
    l.woff = o;
    o += (size_t)l.cout * 9 * l.cin;
    l.boff = o;
    o += l.cout;
    l.bfoff = bo;
    if (i > 0) bo += (size_t)l.cout * 9 * l.cin;
    L.push_back(l);
  }
}

size_t nparams_total() {
  std::vector<Layer> L;
  layer_table(L);
  return L.back().boff + L.back().cout + kBase + 1;
}
size_t bf_total() {
  std::vector<Layer> L;
  layer_table(L);
  const Layer& l = L.back();
  return l.bfoff + (size_t)l.cout * 9 * l.cin;
}

}  // namespace

struct smili_unet {
  int B, H, W;
  hipStream_t s = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
  std::vector<Layer> L;
  size_t np = 0, nbf = 0;
  float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
  bf16_t *wbf = nullptr, *wfl = nullptr;
  float *xin = nullptr, *tgt = nullptr, *out = nullptr;
  float *dtmp = nullptr;          // input gradient of decoder convolutions
  bf16_t* dp = nullptr;           // masked output gradient (two buffers: a
  bf16_t* dp2 = nullptr;          // layer's dP and the dP it makes for the layer below)
  float* part = nullptr;          // per-block partial sums of bias / first-layer reductions
  float* wpart = nullptr;         // per-pixel-group partial sums of 3x3 weight gradients
  size_t nwpart = 0;
  double* loss = nullptr;         // per step
  int64_t nloss = 0;
  float* lr = nullptr;
  int64_t nlr = 0;
  int64_t* idx = nullptr;
  int64_t nidx = 0;
  float *ds_in = nullptr, *ds_tg = nullptr;
  int64_t ds_n = 0;
  int64_t step = 0;
  // activations
  Act e0a, e0, p0, e1a, e1, p1, e2a, e2, p2, e3a, e3, d2a, d2, d1a, d1, d0a, d0;
  std::vector<void*> allocs;
};

namespace {

int dalloc(smili_unet* n, void** p, size_t bytes) {
  HIPCHK(hipMalloc(p, bytes ? bytes : 16));
  n->allocs.push_back(*p);
  return 0;
}

// An activation; `grad`: also an fp32 gradient buffer (the pooled inputs
// and the skip sources, whose gradients have two parts)
int act_alloc(smili_unet* n, Act& a, int h, int w, int c, bool grad) {
  a.h = h; a.w = w; a.c = c;
  const size_t e = (size_t)n->B * h * w * c;
  int r = dalloc(n, (void**)&a.y, e * sizeof(bf16_t));
  if (r || !grad) return r;
  return dalloc(n, (void**)&a.g, e * sizeof(float));
}

inline int grid_for(size_t work, int per_block = kThreads, int cap = 8192) {
  size_t b = (work + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > (size_t)cap) b = cap;
  return (int)b;
}

const bf16_t* wbf_of(smili_unet* n, int li) { return n->wbf + n->L[li].bfoff; }
const bf16_t* wfl_of(smili_unet* n, int li) { return n->wfl + n->L[li].bfoff; }

// Launches one 3x3 convolution; returns its number of pixel tiles (the rows
// of mode 3's partial sums).
int launch_conv(smili_unet* n, int nb, const Act* x1, int up1, const Act* x2, int li, int mode, bf16_t* ybf,
                float* yf, int h, int w, int cout, const bf16_t* wt, const float* bias,
                const bf16_t* mask = nullptr, float* part = nullptr) {
  ConvArgs a;
  a.mask = mask;
  a.part = part;
  a.x1 = x1->y;
  a.c1 = x1->c;
  a.up1 = up1;
  a.x2 = x2 ? x2->y : nullptr;
  a.c2 = x2 ? x2->c : 0;
  a.w = wt;
  a.bias = bias;
  a.ybf = ybf;
  a.yf = yf;
  a.n = nb;
  a.h = h;
  a.w_ = w;
  a.cout = cout;
  a.mode = mode;
  // 8 x 16 tiles where the width needs it, or where the input has at most
  // SMILI_UNET_TW16_MAXCIN channels (few slices: more, smaller work-groups
  // hide the staging latency better than larger tiles)
  const int tw = (w % 32 == 0 && x1->c + (x2 ? x2->c : 0) > SMILI_UNET_TW16_MAXCIN) ? 32 : 16;
  a.tiles_x = w / tw;
  a.tiles_y = h / kTileH;
  // 64 output channels per work-group where that still gives enough
  // work-groups to fill the CUs, else 32 (more, smaller work-groups: the
  // deep, small levels)
  const int tiles = nb * a.tiles_x * a.tiles_y;
  const int ncb = (cout % 64 == 0 && (long long)tiles * (cout / 64) >= SMILI_UNET_NCB4_MIN_GROUPS) ? 4 : 2;
  a.ncob = cout / (ncb * 16);
  const dim3 grid(round8(nb * a.tiles_x * a.tiles_y * a.ncob));
  if (tw == 32 && ncb == 4)
    hipLaunchKernelGGL((k_conv3x3<32, 4>), grid, dim3(kThreads), 0, n->s, a);
  else if (tw == 32)
    hipLaunchKernelGGL((k_conv3x3<32, 2>), grid, dim3(kThreads), 0, n->s, a);
  else if (ncb == 4)
    hipLaunchKernelGGL((k_conv3x3<16, 4>), grid, dim3(kThreads), 0, n->s, a);
  else
    hipLaunchKernelGGL((k_conv3x3<16, 2>), grid, dim3(kThreads), 0, n->s, a);
  (void)li;
  return tiles;
}

// forward 3x3 conv of layer li: y = relu(conv(x1 [up] ++ x2) + b)
void conv_fwd(smili_unet* n, int nb, int li, const Act* x1, int up1, const Act* x2, Act* y) {
  launch_conv(n, nb, x1, up1, x2, li, 0, y->y, nullptr, y->h, y->w, y->c, wbf_of(n, li), n->p + n->L[li].boff);
}

// Split of a weight gradient over pixel groups: about kWgradBlocks
// work-groups in all, each (output block, input slice, pixel group).
#ifndef SMILI_UNET_WGRAD_BLOCKS
#define SMILI_UNET_WGRAD_BLOCKS 512
#endif
constexpr int kWgradBlocks = SMILI_UNET_WGRAD_BLOCKS;
struct WgradPlan {
  int ncb, tw, ncob, nchunk, ntiles, tpg, ngroups;
  int want;   // upper bound of ngroups for this layer at any batch up to this one
  size_t row() const { return (size_t)ncob * ncb * 16 * 9 * nchunk * 32; }
};
WgradPlan wgrad_plan(int nb, int h, int w, int cout, int cin) {
  WgradPlan p;
  p.ncb = (cout % 64 == 0) ? 4 : 2;
  p.tw = (w % 32 == 0 && SMILI_UNET_WGRAD_TW == 32) ? 32 : 16;
  p.ncob = cout / (p.ncb * 16);
  p.nchunk = cin / 32;
  p.ntiles = nb * (h / kTileH) * (w / p.tw);
  const int base = p.ncob * p.nchunk;
  int want = std::max(1, (kWgradBlocks + base / 2) / base);
  want = std::min(want, p.ntiles);
  p.want = want;
  p.tpg = (p.ntiles + want - 1) / want;
  p.ngroups = (p.ntiles + p.tpg - 1) / p.tpg;
  return p;
}

void sum_rows(smili_unet* n, const float* part, int nrow, size_t ncol, float* out) {
  if (ncol >= 4096)
    hipLaunchKernelGGL(k_sum_rows<64>, dim3((unsigned)((ncol + 63) / 64)), dim3(256), 0, n->s, part, nrow,
                       (int)ncol, out);
  else
    hipLaunchKernelGGL(k_sum_rows<4>, dim3((unsigned)((ncol + 3) / 4)), dim3(256), 0, n->s, part, nrow, (int)ncol,
                       out);
}

void conv_wgrad(smili_unet* n, int nb, int li, const Act* x1, int up1, const Act* x2, int h, int w, int cout,
                const bf16_t* dp) {
  WgradArgs a;
  a.dp = dp;
  a.x1 = x1->y;
  a.c1 = x1->c;
  a.up1 = up1;
  a.x2 = x2 ? x2->y : nullptr;
  a.c2 = x2 ? x2->c : 0;
  a.dw = n->wpart;
  a.n = nb;
  a.h = h;
  a.w_ = w;
  a.cout = cout;
  const WgradPlan pl = wgrad_plan(nb, h, w, cout, a.c1 + a.c2);
  a.tiles_x = w / pl.tw;
  a.tiles_y = h / kTileH;
  a.ncob = pl.ncob;
  a.nchunk = pl.nchunk;
  a.tpg = pl.tpg;
  a.ngroups = pl.ngroups;
  const dim3 grid(round8(pl.ncob * pl.nchunk * pl.ngroups));
  if (pl.ncb == 4 && pl.tw == 32)
    hipLaunchKernelGGL((k_conv3x3_wgrad<4, 32>), grid, dim3(kThreads), 0, n->s, a);
  else if (pl.ncb == 4)
    hipLaunchKernelGGL((k_conv3x3_wgrad<4, 16>), grid, dim3(kThreads), 0, n->s, a);
  else if (pl.tw == 32)
    hipLaunchKernelGGL((k_conv3x3_wgrad<2, 32>), grid, dim3(kThreads), 0, n->s, a);
  else
    hipLaunchKernelGGL((k_conv3x3_wgrad<2, 16>), grid, dim3(kThreads), 0, n->s, a);
  sum_rows(n, n->wpart, pl.ngroups, pl.row(), n->g + n->L[li].woff);
}

#ifndef SMILI_UNET_MASK_BLOCKS
#define SMILI_UNET_MASK_BLOCKS 1024
#endif
constexpr int kMaskBlocks = SMILI_UNET_MASK_BLOCKS;   // rows of the partial-sum scratch
constexpr int kInPxPerBlock = 512;


// Backward through ReLU + conv li (input x1 [up] ++ x2, output y with
// gradient da).  The input gradient goes to dx (fp32, [n][h][w][cin] at
// the output's resolution) unless dx is null.
// Backward of 3x3 layer li (input x1 [up] ++ x2, output y) from its dP:
// the weight gradient, then the input gradient -- to dx (fp32), or, with
// fuse (the input activation, when this layer is its only consumer),
// through fuse's ReLU straight into the dP of layer li - 1 (in the other dP
// buffer) with that layer's bias gradient; that dP is returned.
bf16_t* conv_bwd(smili_unet* n, int nb, int li, const Act* y, bf16_t* dp, const Act* x1, int up1, const Act* x2,
                 float* dx, const Act* fuse = nullptr) {
  conv_wgrad(n, nb, li, x1, up1, x2, y->h, y->w, y->c, dp);
  Act dpa;
  dpa.y = dp;
  dpa.c = y->c;
  if (fuse) {
    bf16_t* out = (dp == n->dp) ? n->dp2 : n->dp;
    const int ntiles = launch_conv(n, nb, &dpa, 0, nullptr, li, 3, out, nullptr, y->h, y->w, n->L[li].cin,
                                   wfl_of(n, li), nullptr, fuse->y, n->part);
    sum_rows(n, n->part, ntiles, n->L[li].cin, n->g + n->L[li - 1].boff);
    return out;
  }
  launch_conv(n, nb, &dpa, 0, nullptr, li, 1, nullptr, dx, y->h, y->w, n->L[li].cin, wfl_of(n, li), nullptr);
  return nullptr;
}

// Splits the input gradient dcat of decoder layer li + 1 (input [up(x1) ++
// x2], h x w): x1's part becomes the dP of layer li (x1's producer) in
// n->dp, with that layer's bias gradient; x2's part goes to x2->g.
void split_up(smili_unet* n, int nb, int li, const float* dcat, int h, int w, const Act* x1, Act* x2) {
  const size_t work = (size_t)nb * (h / 2) * (w / 2) * (x1->c / 4) + (size_t)nb * h * w * (x2->c / 4);
  const int blocks = grid_for(work, kThreads, kMaskBlocks);
  hipLaunchKernelGGL(k_split_up, dim3(blocks), dim3(kThreads), 0, n->s, dcat, x1->y, n->dp, n->part, x2->g, nb, h, w,
                     x1->c, x2->c);
  sum_rows(n, n->part, blocks, x1->c, n->g + n->L[li].boff);
}

void pool_fwd(smili_unet* n, int nb, const Act* x, Act* y) {
  const size_t work = (size_t)nb * y->h * y->w * (y->c / 8);
  hipLaunchKernelGGL(k_pool, dim3(grid_for(work)), dim3(kThreads), 0, n->s, x->y, y->y, nb, x->h, x->w, x->c);
}

// dP of layer li, whose output x is pooled (gradient dy) and also feeds a
// skip connection (gradient x->g), into n->dp, with that layer's bias
// gradient.
void pool_bwd(smili_unet* n, int nb, int li, const float* dy, const Act* x) {
  const size_t work = (size_t)nb * (x->h / 2) * (x->w / 2) * (x->c / 4);
  const int blocks = grid_for(work, kThreads, kMaskBlocks);
  hipLaunchKernelGGL(k_pool_bwd_mask, dim3(blocks), dim3(kThreads), 0, n->s, dy, x->g, x->y, n->dp, n->part, nb,
                     x->h, x->w, x->c);
  sum_rows(n, n->part, blocks, x->c, n->g + n->L[li].boff);
}

void forward(smili_unet* n, int nb) {
  const Layer& l0 = n->L[0];
  hipLaunchKernelGGL(k_conv_in, dim3(grid_for((size_t)nb * n->H * n->W * (kBase / 8))), dim3(kThreads), 0, n->s, n->xin,
                     n->p + l0.woff, n->p + l0.boff, n->e0a.y, nb, n->H, n->W);
  conv_fwd(n, nb, 1, &n->e0a, 0, nullptr, &n->e0);
  pool_fwd(n, nb, &n->e0, &n->p0);
  conv_fwd(n, nb, 2, &n->p0, 0, nullptr, &n->e1a);
  conv_fwd(n, nb, 3, &n->e1a, 0, nullptr, &n->e1);
  pool_fwd(n, nb, &n->e1, &n->p1);
  conv_fwd(n, nb, 4, &n->p1, 0, nullptr, &n->e2a);
  conv_fwd(n, nb, 5, &n->e2a, 0, nullptr, &n->e2);
  pool_fwd(n, nb, &n->e2, &n->p2);
  conv_fwd(n, nb, 6, &n->p2, 0, nullptr, &n->e3a);
  conv_fwd(n, nb, 7, &n->e3a, 0, nullptr, &n->e3);
  conv_fwd(n, nb, 8, &n->e3, 1, &n->e2, &n->d2a);
  conv_fwd(n, nb, 9, &n->d2a, 0, nullptr, &n->d2);
  conv_fwd(n, nb, 10, &n->d2, 1, &n->e1, &n->d1a);
  conv_fwd(n, nb, 11, &n->d1a, 0, nullptr, &n->d1);
  conv_fwd(n, nb, 12, &n->d1, 1, &n->e0, &n->d0a);
  conv_fwd(n, nb, 13, &n->d0a, 0, nullptr, &n->d0);
}

size_t head_off(smili_unet* n) { return n->L.back().boff + n->L.back().cout; }

// Inference: out = max(z, 0).  Training: the loss, the head's gradient, and
// the dP (in n->dp) and bias gradient of the last 3x3 layer.
void head(smili_unet* n, int nb, bool train, double* loss) {
  const size_t npx = (size_t)nb * n->H * n->W;
  const float scale = 1.f / (float)npx;
  const int blocks = grid_for(npx * 4, kThreads, 1024);
  hipLaunchKernelGGL(k_head, dim3(blocks), dim3(kThreads), 0, n->s, n->d0.y, n->p + head_off(n),
                     train ? nullptr : n->out, train ? n->tgt : nullptr, train ? n->dp : nullptr,
                     train ? n->part : nullptr, train ? n->g + head_off(n) : nullptr, loss, npx, scale);
  if (train) sum_rows(n, n->part, blocks, kBase, n->g + n->L[13].boff);
}

void backward(smili_unet* n, int nb) {
  // Every dP is made by the step that completes its gradient: the head
  // (layer 13), the input gradient of the second convolution of a level
  // (the only consumer of the first one's output: layers 12, 10, 8, 6, 4,
  // 2, 0), the upsampling adjoint (layers 11, 9, 7) and the pooling adjoint
  // (layers 5, 3, 1, whose outputs also feed skip connections).  Those of
  // the first kind go to n->dp2, the others to n->dp.
  // decoder
  bf16_t* dp = conv_bwd(n, nb, 13, &n->d0, n->dp, &n->d0a, 0, nullptr, nullptr, &n->d0a);
  conv_bwd(n, nb, 12, &n->d0a, dp, &n->d1, 1, &n->e0, n->dtmp);
  split_up(n, nb, 11, n->dtmp, n->d0a.h, n->d0a.w, &n->d1, &n->e0);
  dp = conv_bwd(n, nb, 11, &n->d1, n->dp, &n->d1a, 0, nullptr, nullptr, &n->d1a);
  conv_bwd(n, nb, 10, &n->d1a, dp, &n->d2, 1, &n->e1, n->dtmp);
  split_up(n, nb, 9, n->dtmp, n->d1a.h, n->d1a.w, &n->d2, &n->e1);
  dp = conv_bwd(n, nb, 9, &n->d2, n->dp, &n->d2a, 0, nullptr, nullptr, &n->d2a);
  conv_bwd(n, nb, 8, &n->d2a, dp, &n->e3, 1, &n->e2, n->dtmp);
  split_up(n, nb, 7, n->dtmp, n->d2a.h, n->d2a.w, &n->e3, &n->e2);
  // encoder
  dp = conv_bwd(n, nb, 7, &n->e3, n->dp, &n->e3a, 0, nullptr, nullptr, &n->e3a);
  conv_bwd(n, nb, 6, &n->e3a, dp, &n->p2, 0, nullptr, n->p2.g);
  pool_bwd(n, nb, 5, n->p2.g, &n->e2);
  dp = conv_bwd(n, nb, 5, &n->e2, n->dp, &n->e2a, 0, nullptr, nullptr, &n->e2a);
  conv_bwd(n, nb, 4, &n->e2a, dp, &n->p1, 0, nullptr, n->p1.g);
  pool_bwd(n, nb, 3, n->p1.g, &n->e1);
  dp = conv_bwd(n, nb, 3, &n->e1, n->dp, &n->e1a, 0, nullptr, nullptr, &n->e1a);
  conv_bwd(n, nb, 2, &n->e1a, dp, &n->p0, 0, nullptr, n->p0.g);
  pool_bwd(n, nb, 1, n->p0.g, &n->e0);
  // layer 1's input gradient is the first convolution's dP (and bias gradient)
  dp = conv_bwd(n, nb, 1, &n->e0, n->dp, &n->e0a, 0, nullptr, nullptr, &n->e0a);
  // first convolution: weight gradient only
  const size_t npx = (size_t)nb * n->H * n->W;
  const int blocks = (int)((npx + kInPxPerBlock - 1) / kInPxPerBlock);
  hipLaunchKernelGGL(k_conv_in_wgrad, dim3(blocks), dim3(kThreads), 0, n->s, dp, n->xin, n->part, nb, n->H,
                     n->W, kInPxPerBlock);
  sum_rows(n, n->part, blocks, kBase * 18, n->g + n->L[0].woff);
}

void cast_weights(smili_unet* n) {
  CastTable tab;
  tab.n = layer_count() - 1;
  size_t most = 0;
  for (int li = 1; li < layer_count(); ++li) {
    const Layer& l = n->L[li];
    tab.cout[li - 1] = l.cout;
    tab.cin[li - 1] = l.cin;
    tab.src[li - 1] = (long long)l.woff;
    tab.dst[li - 1] = (long long)l.bfoff;
    most = std::max(most, (size_t)l.cout * 9 * l.cin);
  }
  hipLaunchKernelGGL(k_cast_w, dim3(grid_for(most, kThreads, 256), tab.n), dim3(kThreads), 0, n->s, n->p, n->wbf,
                     n->wfl, tab);
}

void zero_grad(smili_unet* n) {
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n->np)), dim3(kThreads), 0, n->s, n->g, n->np, 0.f);
}

int ensure_loss(smili_unet* n, int64_t count) {
  if (count <= n->nloss) return 0;
  if (n->loss) {
    HIPCHK(hipFree(n->loss));
    for (auto& a : n->allocs)
      if (a == n->loss) a = nullptr;
  }
  n->loss = nullptr;
  int r = dalloc(n, (void**)&n->loss, count * sizeof(double));
  if (r) return r;
  n->nloss = count;
  return 0;
}

double conv_flops_fwd(smili_unet* n) {
  double f = 0.0;
  const int hs[14] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0};
  for (int li = 0; li < 14; ++li) {
    const double px = (double)n->B * (n->H >> hs[li]) * (n->W >> hs[li]);
    f += 2.0 * px * n->L[li].cout * 9.0 * n->L[li].cin;
  }
  return f;
}

}  // namespace

extern "C" {

const char* smili_unet_last_error(void) { return g_err.c_str(); }

int64_t smili_unet_nparams(void) { return (int64_t)nparams_total(); }

void smili_unet_destroy(smili_unet* n) {
  if (!n) return;
  if (n->s) (void)hipStreamSynchronize(n->s);
  for (void* a : n->allocs)
    if (a) (void)hipFree(a);
  if (n->ev0) (void)hipEventDestroy(n->ev0);
  if (n->ev1) (void)hipEventDestroy(n->ev1);
  if (n->ev2) (void)hipEventDestroy(n->ev2);
  if (n->s) (void)hipStreamDestroy(n->s);
  delete n;
}

int smili_unet_create(int32_t batch, int32_t height, int32_t width, smili_unet** out) {
  if (!out) return fail(-1, "out is null");
  *out = nullptr;
  if (batch < 1 || batch > 4096) return fail(-1, "batch must be in 1..4096");
  if (height < 64 || height % 64 || width < 128 || width % 128)
    return fail(-1, "height must be a multiple of 64 and width a multiple of 128");
  if ((int64_t)batch * height * width > (int64_t)1 << 28) return fail(-1, "batch x height x width too large");
  smili_unet* n = new smili_unet();
  n->B = batch;
  n->H = height;
  n->W = width;
  layer_table(n->L);
  n->np = nparams_total();
  n->nbf = bf_total();
  int r = 0;
  auto chk = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && !r) r = fail(-2, std::string(what) + ": " + hipGetErrorString(e));
  };
  chk(hipStreamCreateWithFlags(&n->s, hipStreamNonBlocking), "hipStreamCreate");
  chk(hipEventCreate(&n->ev0), "hipEventCreate");
  chk(hipEventCreate(&n->ev1), "hipEventCreate");
  chk(hipEventCreate(&n->ev2), "hipEventCreate");
  const int c0 = kBase, c1 = 2 * kBase, c2 = 4 * kBase, c3 = 8 * kBase;
  const int H = height, W = width;
  const size_t B = batch;
  if (!r) r = dalloc(n, (void**)&n->p, n->np * 4);
  if (!r) r = dalloc(n, (void**)&n->m, n->np * 4);
  if (!r) r = dalloc(n, (void**)&n->v, n->np * 4);
  if (!r) r = dalloc(n, (void**)&n->g, n->np * 4);
  if (!r) r = dalloc(n, (void**)&n->wbf, n->nbf * 2);
  if (!r) r = dalloc(n, (void**)&n->wfl, n->nbf * 2);
  if (!r) r = dalloc(n, (void**)&n->xin, B * H * W * 2 * 4);
  if (!r) r = dalloc(n, (void**)&n->tgt, B * H * W * 4);
  if (!r) r = dalloc(n, (void**)&n->out, B * H * W * 4);
  // largest decoder input gradient: level 0, c1 + c0 channels
  size_t dt = B * H * W * (c1 + c0);
  dt = std::max(dt, B * (H / 2) * (W / 2) * (size_t)(c2 + c1));
  dt = std::max(dt, B * (H / 4) * (W / 4) * (size_t)(c3 + c2));
  if (!r) r = dalloc(n, (void**)&n->dtmp, dt * 4);
  if (!r) r = dalloc(n, (void**)&n->dp, B * H * W * c0 * 2);
  if (!r) r = dalloc(n, (void**)&n->dp2, B * H * W * c0 * 2);
  // partial sums: the larger of the bias reductions (kMaskBlocks rows of at
  // most c3 channels) and the first convolution's weight gradient
  const size_t nin_blocks = (B * H * W + kInPxPerBlock - 1) / kInPxPerBlock;
  // (and the fused input gradients' per-tile sums: at most one row of c0
  // channels per 8 x 16 tile of the full-resolution level)
  const size_t fused_rows = B * (H / kTileH) * (W / 16) * (size_t)c0;
  if (!r)
    r = dalloc(n, (void**)&n->part,
               std::max(std::max((size_t)kMaskBlocks * c3, nin_blocks * kBase * 18), fused_rows) * 4);
  // weight-gradient partial rows: the largest split over the 3x3 layers
  // (ngroups <= min(want, tiles) <= its value at the full batch)
  const int lev[14] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0};
  for (int li = 1; li < layer_count(); ++li) {
    const WgradPlan pl = wgrad_plan(batch, H >> lev[li], W >> lev[li], n->L[li].cout, n->L[li].cin);
    n->nwpart = std::max(n->nwpart, (size_t)pl.want * pl.row());
  }
  if (!r) r = dalloc(n, (void**)&n->wpart, n->nwpart * 4);
  const struct { Act* a; int s, c; bool g; } acts[17] = {
      {&n->e0a, 0, c0, false}, {&n->e0, 0, c0, true},   {&n->p0, 1, c0, true},   {&n->e1a, 1, c1, false},
      {&n->e1, 1, c1, true},   {&n->p1, 2, c1, true},   {&n->e2a, 2, c2, false}, {&n->e2, 2, c2, true},
      {&n->p2, 3, c2, true},   {&n->e3a, 3, c3, false}, {&n->e3, 3, c3, false},  {&n->d2a, 2, c2, false},
      {&n->d2, 2, c2, false},  {&n->d1a, 1, c1, false}, {&n->d1, 1, c1, false},  {&n->d0a, 0, c0, false},
      {&n->d0, 0, c0, false}};
  for (int i = 0; i < 17 && !r; ++i)
    r = act_alloc(n, *acts[i].a, H >> acts[i].s, W >> acts[i].s, acts[i].c, acts[i].g);
  if (!r) r = ensure_loss(n, 16);
  if (!r) {
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n->np)), dim3(kThreads), 0, n->s, n->p, n->np, 0.f);
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n->np)), dim3(kThreads), 0, n->s, n->m, n->np, 0.f);
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n->np)), dim3(kThreads), 0, n->s, n->v, n->np, 0.f);
    cast_weights(n);
    chk(hipGetLastError(), "launch");
    chk(hipStreamSynchronize(n->s), "hipStreamSynchronize");
  }
  if (r) {
    smili_unet_destroy(n);
    return r;
  }
  *out = n;
  return 0;
}

int smili_unet_set_params(smili_unet* n, const float* params) {
  if (!n || !params) return fail(-1, "null argument");
  for (size_t i = 0; i < n->np; ++i)
    if (!std::isfinite(params[i])) return fail(-1, "non-finite parameter at index " + std::to_string(i));
  HIPCHK(hipMemcpyAsync(n->p, params, n->np * 4, hipMemcpyHostToDevice, n->s));
  cast_weights(n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(n->s));
  return 0;
}

int smili_unet_get_params(smili_unet* n, float* params) {
  if (!n || !params) return fail(-1, "null argument");
  HIPCHK(hipMemcpyAsync(params, n->p, n->np * 4, hipMemcpyDeviceToHost, n->s));
  HIPCHK(hipStreamSynchronize(n->s));
  return 0;
}

int smili_unet_forward(smili_unet* n, const float* input, int32_t count, float* output) {
  if (!n || !input || !output) return fail(-1, "null argument");
  if (count < 1) return fail(-1, "n must be >= 1");
  const size_t npx = (size_t)n->H * n->W;
  for (int32_t b0 = 0; b0 < count; b0 += n->B) {
    const int nb = std::min<int>(n->B, count - b0);
    HIPCHK(hipMemcpyAsync(n->xin, input + (size_t)b0 * npx * 2, nb * npx * 2 * 4, hipMemcpyHostToDevice, n->s));
    forward(n, nb);
    head(n, nb, false, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(output + (size_t)b0 * npx, n->out, nb * npx * 4, hipMemcpyDeviceToHost, n->s));
  }
  HIPCHK(hipStreamSynchronize(n->s));
  return 0;
}

int smili_unet_loss_grad(smili_unet* n, const float* input, const float* target, int32_t count, double* loss,
                         float* grad) {
  if (!n || !input || !target || !loss) return fail(-1, "null argument");
  if (count < 1 || count > n->B) return fail(-1, "n must be in 1..batch");
  const size_t npx = (size_t)n->H * n->W;
  HIPCHK(hipMemcpyAsync(n->xin, input, count * npx * 2 * 4, hipMemcpyHostToDevice, n->s));
  HIPCHK(hipMemcpyAsync(n->tgt, target, count * npx * 4, hipMemcpyHostToDevice, n->s));
  zero_grad(n);
  hipLaunchKernelGGL(k_fill_d, dim3(1), dim3(64), 0, n->s, n->loss, (size_t)1);
  forward(n, count);
  head(n, count, true, n->loss);
  backward(n, count);
  HIPCHK(hipGetLastError());
  double l = 0.0;
  HIPCHK(hipMemcpyAsync(&l, n->loss, sizeof(double), hipMemcpyDeviceToHost, n->s));
  if (grad) HIPCHK(hipMemcpyAsync(grad, n->g, n->np * 4, hipMemcpyDeviceToHost, n->s));
  HIPCHK(hipStreamSynchronize(n->s));
  *loss = l / ((double)count * npx);
  return 0;
}

int smili_unet_set_data(smili_unet* n, const float* inputs, const float* targets, int64_t count) {
  if (!n || !inputs || !targets) return fail(-1, "null argument");
  if (count < 1) return fail(-1, "count must be >= 1");
  const size_t npx = (size_t)n->H * n->W;
  HIPCHK(hipStreamSynchronize(n->s));
  for (void** p : {(void**)&n->ds_in, (void**)&n->ds_tg}) {
    if (*p) {
      HIPCHK(hipFree(*p));
      for (auto& a : n->allocs)
        if (a == *p) a = nullptr;
      *p = nullptr;
    }
  }
  int r = dalloc(n, (void**)&n->ds_in, count * npx * 2 * 4);
  if (!r) r = dalloc(n, (void**)&n->ds_tg, count * npx * 4);
  if (r) return r;
  HIPCHK(hipMemcpyAsync(n->ds_in, inputs, count * npx * 2 * 4, hipMemcpyHostToDevice, n->s));
  HIPCHK(hipMemcpyAsync(n->ds_tg, targets, count * npx * 4, hipMemcpyHostToDevice, n->s));
  HIPCHK(hipStreamSynchronize(n->s));
  n->ds_n = count;
  return 0;
}

int smili_unet_reset_optimizer(smili_unet* n) {
  if (!n) return fail(-1, "null argument");
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n->np)), dim3(kThreads), 0, n->s, n->m, n->np, 0.f);
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n->np)), dim3(kThreads), 0, n->s, n->v, n->np, 0.f);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(n->s));
  n->step = 0;
  return 0;
}

static int train_steps(smili_unet* n, int64_t nsteps, const float* hyper) {
  const size_t npx = (size_t)n->H * n->W;
  const float b1 = hyper[0], b2 = hyper[1], eps = hyper[2], wd = hyper[3];
  for (int64_t s = 0; s < nsteps; ++s) {
    hipLaunchKernelGGL(k_gather, dim3(grid_for((size_t)n->B * npx)), dim3(kThreads), 0, n->s, n->ds_in, n->ds_tg,
                       n->idx + s * n->B, n->xin, n->tgt, n->B, n->H, n->W);
    zero_grad(n);
    forward(n, n->B);
    head(n, n->B, true, n->loss + s);
    backward(n, n->B);
    n->step += 1;
    const float bc1 = 1.f - powf(b1, (float)n->step), bc2 = 1.f - powf(b2, (float)n->step);
    hipLaunchKernelGGL(k_adam, dim3(grid_for(n->np)), dim3(kThreads), 0, n->s, n->p, n->m, n->v, n->g, n->np,
                       n->lr, (int)s, b1, b2, eps, wd, bc1, bc2);
    cast_weights(n);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

static int upload_steps(smili_unet* n, const int64_t* index, int64_t nsteps, const float* lr) {
  const int64_t ni = nsteps * n->B;
  for (int64_t i = 0; i < ni; ++i) {
    if (index[i] < 0 || (index[i] >> 3) >= n->ds_n) return fail(-1, "index out of range at " + std::to_string(i));
    if ((index[i] & 7) != 0 && n->H != n->W)
      return fail(-1, "symmetries other than the identity need square images (index " + std::to_string(i) + ")");
  }
  HIPCHK(hipStreamSynchronize(n->s));
  if (ni > n->nidx) {
    if (n->idx) {
      HIPCHK(hipFree(n->idx));
      for (auto& a : n->allocs)
        if (a == n->idx) a = nullptr;
      n->idx = nullptr;
    }
    int r = dalloc(n, (void**)&n->idx, ni * 8);
    if (r) return r;
    n->nidx = ni;
  }
  if (nsteps > n->nlr) {
    if (n->lr) {
      HIPCHK(hipFree(n->lr));
      for (auto& a : n->allocs)
        if (a == n->lr) a = nullptr;
      n->lr = nullptr;
    }
    int r = dalloc(n, (void**)&n->lr, nsteps * 4);
    if (r) return r;
    n->nlr = nsteps;
  }
  int r = ensure_loss(n, nsteps);
  if (r) return r;
  HIPCHK(hipMemcpyAsync(n->idx, index, ni * 8, hipMemcpyHostToDevice, n->s));
  HIPCHK(hipMemcpyAsync(n->lr, lr, nsteps * 4, hipMemcpyHostToDevice, n->s));
  hipLaunchKernelGGL(k_fill_d, dim3(grid_for(nsteps)), dim3(kThreads), 0, n->s, n->loss, (size_t)nsteps);
  return 0;
}

int smili_unet_train(smili_unet* n, const int64_t* index, int64_t nsteps, const float* lr, const float* hyper,
                     double* losses) {
  if (!n || !index || !lr || !hyper) return fail(-1, "null argument");
  if (nsteps < 1) return fail(-1, "nsteps must be >= 1");
  if (!n->ds_n) return fail(-1, "no training set uploaded");
  for (int64_t s = 0; s < nsteps; ++s)
    if (!std::isfinite(lr[s]) || lr[s] < 0.f) return fail(-1, "invalid learning rate at step " + std::to_string(s));
  if (!(hyper[0] >= 0.f && hyper[0] < 1.f) || !(hyper[1] >= 0.f && hyper[1] < 1.f) || !(hyper[2] > 0.f) ||
      !(hyper[3] >= 0.f) || !std::isfinite(hyper[2]) || !std::isfinite(hyper[3]))
    return fail(-1, "hyper must be {beta1, beta2 in [0, 1), eps > 0, weight_decay >= 0}");
  int r = upload_steps(n, index, nsteps, lr);
  if (r) return r;
  r = train_steps(n, nsteps, hyper);
  if (r) return r;
  if (losses) HIPCHK(hipMemcpyAsync(losses, n->loss, nsteps * sizeof(double), hipMemcpyDeviceToHost, n->s));
  HIPCHK(hipStreamSynchronize(n->s));
  if (losses) {
    const double norm = (double)n->B * n->H * n->W;
    for (int64_t s = 0; s < nsteps; ++s) losses[s] /= norm;
  }
  return 0;
}

int smili_unet_time_steps(smili_unet* n, int32_t reps, float* ms) {
  if (!n || !ms) return fail(-1, "null argument");
  if (reps < 1) return fail(-1, "reps must be >= 1");
  if (n->ds_n < n->B) return fail(-1, "the uploaded set must hold at least one batch");
  std::vector<int64_t> index((size_t)reps * n->B);
  for (size_t i = 0; i < index.size(); ++i) index[i] = (int64_t)(i % n->B) * 8;
  std::vector<float> lr(reps, 0.f);
  int r = upload_steps(n, index.data(), reps, lr.data());
  if (r) return r;
  // the timed steps run with learning rate 0 (and no weight decay), so the
  // parameters stay; the Adam moments and step count they advance are
  // saved here and restored below, so a training run can continue as if
  // the timing had not happened
  float* saved = nullptr;
  HIPCHK(hipMalloc(&saved, 2 * n->np * sizeof(float)));
  const int64_t step0 = n->step;
  auto restore = [&]() {
    (void)hipStreamSynchronize(n->s);
    (void)hipMemcpy(n->m, saved, n->np * sizeof(float), hipMemcpyDeviceToDevice);
    (void)hipMemcpy(n->v, saved + n->np, n->np * sizeof(float), hipMemcpyDeviceToDevice);
    (void)hipFree(saved);
    n->step = step0;
  };
  if (hipMemcpyAsync(saved, n->m, n->np * sizeof(float), hipMemcpyDeviceToDevice, n->s) != hipSuccess ||
      hipMemcpyAsync(saved + n->np, n->v, n->np * sizeof(float), hipMemcpyDeviceToDevice, n->s) != hipSuccess) {
    (void)hipFree(saved);
    return fail(-2, "copy of the optimizer state failed");
  }
  const float hyper[4] = {0.9f, 0.999f, 1e-8f, 0.f};
  if (hipStreamSynchronize(n->s) != hipSuccess) {
    restore();
    return fail(-2, "stream synchronisation failed");
  }
  float t_train = 0.f, t_fwd = 0.f;
  hipError_t e = hipEventRecord(n->ev0, n->s);
  r = e == hipSuccess ? train_steps(n, reps, hyper) : -2;
  if (r == 0) {
    (void)hipEventRecord(n->ev1, n->s);
    // forward passes alone
    for (int i = 0; i < reps; ++i) forward(n, n->B);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(n->ev2, n->s);
    if (e == hipSuccess) e = hipEventSynchronize(n->ev2);
    if (e == hipSuccess) e = hipEventElapsedTime(&t_train, n->ev0, n->ev1);
    if (e == hipSuccess) e = hipEventElapsedTime(&t_fwd, n->ev1, n->ev2);
  }
  restore();
  if (r) return r;
  if (e != hipSuccess) return fail(-2, hipGetErrorString(e));
  ms[0] = t_train / reps;
  ms[1] = t_fwd / reps;
  return 0;
}

int smili_unet_time_layer(smili_unet* n, int32_t layer, int32_t pass, int32_t reps, float* ms, double* flop) {
  if (!n || !ms || !flop) return fail(-1, "null argument");
  if (layer < 1 || layer >= layer_count()) return fail(-1, "layer must be in 1..13");
  if (pass < 0 || pass > 2) return fail(-1, "pass must be 0, 1 or 2");
  if (reps < 1) return fail(-1, "reps must be >= 1");
  // inputs and output of each 3x3 layer, as in forward()
  const Act* io[14][3] = {{nullptr, nullptr, nullptr},
                          {&n->e0a, nullptr, &n->e0},  {&n->p0, nullptr, &n->e1a}, {&n->e1a, nullptr, &n->e1},
                          {&n->p1, nullptr, &n->e2a},  {&n->e2a, nullptr, &n->e2}, {&n->p2, nullptr, &n->e3a},
                          {&n->e3a, nullptr, &n->e3},  {&n->e3, &n->e2, &n->d2a},  {&n->d2a, nullptr, &n->d2},
                          {&n->d2, &n->e1, &n->d1a},   {&n->d1a, nullptr, &n->d1}, {&n->d1, &n->e0, &n->d0a},
                          {&n->d0a, nullptr, &n->d0}};
  const Act* x1 = io[layer][0];
  const Act* x2 = io[layer][1];
  const Act* y = io[layer][2];
  const int up1 = x2 ? 1 : 0;
  Act dpa;
  dpa.y = n->dp;
  dpa.c = y->c;
  forward(n, n->B);
  HIPCHK(hipEventRecord(n->ev0, n->s));
  for (int r = 0; r < reps; ++r) {
    if (pass == 0)
      conv_fwd(n, n->B, layer, x1, up1, x2, const_cast<Act*>(y));
    else if (pass == 1 && (layer & 1))  // fused with the ReLU of its input, as in training
      launch_conv(n, n->B, &dpa, 0, nullptr, layer, 3, n->dp2, nullptr, y->h, y->w, n->L[layer].cin,
                  wfl_of(n, layer), nullptr, x1->y, n->part);
    else if (pass == 1)
      launch_conv(n, n->B, &dpa, 0, nullptr, layer, 1, nullptr, n->dtmp, y->h, y->w, n->L[layer].cin,
                  wfl_of(n, layer), nullptr);
    else
      conv_wgrad(n, n->B, layer, x1, up1, x2, y->h, y->w, y->c, n->dp);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(n->ev1, n->s));
  HIPCHK(hipEventSynchronize(n->ev1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, n->ev0, n->ev1));
  *ms = t / reps;
  *flop = 2.0 * n->B * y->h * y->w * (double)n->L[layer].cout * 9.0 * n->L[layer].cin;
  return 0;
}

int smili_unet_flops(smili_unet* n, double* flop) {
  if (!n || !flop) return fail(-1, "null argument");
  const double f = conv_flops_fwd(n);
  // training: forward + input gradient (all but conv_in) + weight gradient
  const double f_in = 2.0 * n->B * n->H * n->W * kBase * 9.0 * 2;
  flop[0] = 3.0 * f - f_in;
  flop[1] = f;
  return 0;
}

}  // extern "C"
