// Kernels of the pressure warm-start network: BF16 MFMA implicit-GEMM 3x3
// convolutions (forward, input gradient, weight gradient), GELU, 2x2
// average pooling, nearest upsampling, 1x1 head, MSE, Adam.
//
// Activations are NHWC bf16 (channels contiguous) with an explicit channel
// stride, so a decoder input is one buffer whose first channels come from
// the upsampled coarser level and whose last channels are the encoder skip
// written there directly (no concatenation copy).  Convolutions are
// "same" 3x3 with zero padding.
#pragma once

#include "aj_common.h"

namespace aj {
namespace nn {

typedef unsigned short bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ bf16_t f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (bf16_t)(u >> 16);
}

__device__ __forceinline__ float gelu(float x) {
  return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * __expf(-0.5f * x * x);
  return cdf + x * pdf;
}

__device__ __forceinline__ bf16x8 load8(const bf16_t* p) {
  const uint4 v = *reinterpret_cast<const uint4*>(p);
  return __builtin_bit_cast(bf16x8, v);
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0)); }

// ---------------------------------------------------------------------------
// 3x3 convolution, implicit GEMM on v_mfma_f32_16x16x32_bf16.
// D[pixel][co] = sum_{tap, ci} in[pixel + tap][ci] * w[co][tap][ci]
// A operand (16 pixels x 32 ci): lane l holds pixel l%16, ci 8*(l/16)..+7.
// B operand (32 ci x 16 co):     lane l holds co l%16,    ci 8*(l/16)..+7.
// D (16 pixels x 16 co):         lane l holds co l%16, pixels 4*(l/16)+i.
//
// A workgroup (4 waves) computes a PH x PW pixel patch of one image for NT
// output channels.  Per 32-channel slice of the input it stages the
// (PH+2) x (PW+2) input patch and the 9 x NT x 32 weights in LDS, then each
// wave runs its PH/4 pixel rows: per tap, NT/16 B fragments are read once
// and reused across the wave's (PH/4)(PW/16) pixel sub-tiles.  LDS rows are
// 32 bf16 + 8 pad (80 B): the 16 lanes of a quarter-wave then hit 16
// distinct 4-bank groups, so the 16-byte fragment reads are conflict-free.
// The next slice's global loads are issued before the current slice's
// MFMAs, so their latency overlaps the arithmetic; all loads are
// unconditional (padding is applied when staging), which keeps the
// compiler's counter waits to the slice being staged.
//
// Input modes (inference, where the pooled / upsampled tensors are not
// needed afterwards): src = 1 reads the input as the 2x2 average of `in`
// (2H x 2W, the arithmetic of k_pool); src = 2 reads channels < cu as the
// nearest upsampling of in2 (H/2 x W/2, stride ld2, as k_up) and channels
// >= cu from `in` (the concatenation's skip part); src = 3 builds the
// warm-start network's input row from the restricted rhs R and T =
// tanh(sdf_r / sdf_ref) (FP32, B x H x W): channel 0 = R / s, channel 1 =
// T, the others 0, with s the instance's RMS of R from the restriction's
// partial sums (the arithmetic of k_ws_input, whose row it replaces).
//
// HEAD epilogue (inference, NT = co): the activations of the workgroup's
// pixels go through LDS instead of memory and each pixel's 1x1 head (bias
// + the channel sum in channel order, as k_ws_head) times s pscale is
// written.
// ---------------------------------------------------------------------------
struct ConvArgs {
  const bf16_t* in;
  int ldi, ci;          // input channel stride, channels (multiple of 32)
  const bf16_t* w;      // [co][9][ci]
  const float* bias;    // [co] or null
  int co;               // output channels (multiple of NT)
  int B, H, W;
  bf16_t* out;          // post-activation bf16 (may be null)
  int ldo;
  bf16_t* pre;          // pre-activation bf16 (may be null)
  int ldp;
  float* outf;          // fp32 output (input-gradient mode, may be null)
  int ldf, accum;
  int act;              // 1 = GELU
  int src;              // 0 = in, 1 = 2x2 average of in, 2 = [up(in2), in]
  const bf16_t* in2;    // src = 2: lower-resolution tensor
  int ld2, cu;          // its channel stride, channels taken from it
  const float* inR;     // src = 3: restricted rhs and tanh(sdf_r / sdf_ref)
  const float* inT;
  const double* part;   // src = 3: per instance nblk partial sums of R^2
  int nblk;
  float* s_out;         // src = 3: per-instance s, written by the instance's first workgroup
  const float* hw;      // HEAD: head weights [co], bias, per-instance s, pscale, output
  const float* hb;
  const float* hs;
  float hscale;
  float* hout;          // B x H x W
};

// one term of the 1x1 head's channel sum (the same operation in the fused
// epilogue and in k_ws_head)
__device__ __forceinline__ float head_acc(float v, float x, float w) { return fmaf(x, w, v); }

// 2x2 average of four bf16x8 rows, channel by channel (k_pool's order)
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ u32x4 pool4(u32x4 a, u32x4 b, u32x4 c, u32x4 d) {
  const u16x8 x00 = __builtin_bit_cast(u16x8, a), x01 = __builtin_bit_cast(u16x8, b);
  const u16x8 x10 = __builtin_bit_cast(u16x8, c), x11 = __builtin_bit_cast(u16x8, d);
  u16x8 o;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float s = bf2f(x00[k]) + bf2f(x01[k]) + bf2f(x10[k]) + bf2f(x11[k]);
    o[k] = f2bf(0.25f * s);
  }
  return __builtin_bit_cast(u32x4, o);
}

constexpr int kLdsRow = 40;  // bf16 per LDS row (32 channels + pad)

template <int PH, int PW, int NT, int SRC, bool HEAD = false>
__global__ __launch_bounds__(256) void k_conv3x3(ConvArgs a) {
  static_assert(PH % 4 == 0 && PW % 16 == 0 && NT % 16 == 0, "conv tile");
  static_assert(SRC >= 0 && SRC <= 3, "conv input mode");
  constexpr int RW = PW + 2, RP = (PH + 2) * RW;
  constexpr int WR = PH / 4;         // pixel rows per wave
  constexpr int SX = PW / 16;        // 16-pixel sub-tiles per row
  constexpr int MS = WR * SX, NS = NT / 16;
  constexpr int NXL = (RP * 4 + 255) / 256;      // 16-byte input elements per thread and slice
  constexpr int NWL = (9 * NT * 4 + 255) / 256;  // 16-byte weight loads per thread and slice
  constexpr int NQ = SRC == 1 ? 4 : 1;           // loads per input element (pooling reads 2x2)
  static_assert(NXL <= 32, "conv input patch");
  // one spare row each: a thread without an element stages into it, so
  // that staging has no branches (and the counter waits no joins)
  __shared__ __attribute__((aligned(16))) bf16_t sx[(RP + 1) * kLdsRow];
  __shared__ __attribute__((aligned(16))) bf16_t sw[(9 * NT + 1) * kLdsRow];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int tilesW = (a.W + PW - 1) / PW, tilesH = (a.H + PH - 1) / PH;
  int t = blockIdx.x;
  const int tw = t % tilesW;
  t /= tilesW;
  const int th = t % tilesH;
  const int b = t / tilesH;
  const int h0 = th * PH, w0 = tw * PW;
  const int co0 = blockIdx.y * NT;
  const int kq = (lane >> 4) * 8;
  const int ns = a.ci / 32;
  floatx4 acc[MS][NS];
#pragma unroll
  for (int m = 0; m < MS; ++m)
#pragma unroll
    for (int j = 0; j < NS; ++j) acc[m][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  // Element offsets of this thread's loads, the same for every slice (the
  // slice adds its first channel c0).  Elements outside the patch or the
  // image load from offset 0 of the image (a valid address) and are zeroed
  // when stashed, so every load is unconditional and the counter wait
  // before a stash does not cover more than that slice's loads.
  const bf16_t* img = SRC == 1 ? a.in + (long)b * 4 * a.H * a.W * a.ldi   // 2H x 2W input
                               : a.in + (long)b * a.H * a.W * a.ldi;
  const bf16_t* img2 = SRC == 2 ? a.in2 + (long)b * (a.H >> 1) * (a.W >> 1) * a.ld2 : nullptr;
  const float* Rin = SRC == 3 ? a.inR + (long)b * a.H * a.W : nullptr;
  const float* Tin = SRC == 3 ? a.inT + (long)b * a.H * a.W : nullptr;
  int xo[NXL][NQ], xo2[NXL];
  unsigned okx = 0;
#pragma unroll
  for (int k = 0; k < NXL; ++k) {
    const int e = tid + k * 256;
    const int px = e >> 2, q = e & 3;
    const int rr = px / RW, cc = px - rr * RW;
    const int hh = h0 + rr - 1, ww = w0 + cc - 1;
    const bool ok = e < RP * 4 && hh >= 0 && hh < a.H && ww >= 0 && ww < a.W;
    okx |= ok ? (1u << k) : 0u;
    if (SRC == 1) {
      const int W2 = 2 * a.W;
      const int f = ok ? ((2 * hh) * W2 + 2 * ww) * a.ldi + q * 8 : 0;
      xo[k][0] = f;
      xo[k][NQ > 1 ? 1 : 0] = ok ? f + a.ldi : 0;
      xo[k][NQ > 2 ? 2 : 0] = ok ? f + W2 * a.ldi : 0;
      xo[k][NQ > 3 ? 3 : 0] = ok ? f + (W2 + 1) * a.ldi : 0;
    } else if (SRC == 3) {
      xo[k][0] = ok ? hh * a.W + ww : 0;  // pixel of the instance
    } else {
      xo[k][0] = ok ? (hh * a.W + ww) * a.ldi + q * 8 : 0;
    }
    xo2[k] = (SRC == 2 && ok) ? ((hh >> 1) * (a.W >> 1) + (ww >> 1)) * a.ld2 + q * 8 : 0;
  }
  int wo[NWL];
#pragma unroll
  for (int k = 0; k < NWL; ++k) {
    const int e = tid + k * 256;
    const int q = e & 3, r = e >> 2;
    const int col = r / 9, tap = r - col * 9;
    wo[k] = e < 9 * NT * 4 ? ((co0 + col) * 9 + tap) * a.ci + q * 8 : 0;
  }
  // global -> registers: slice at channel c0 (input patch with a one-pixel
  // halo, and the slice's weights, row (tap, co) holding 32 channels)
  auto fetch = [&](u32x4 (&X)[NXL][NQ], u32x4 (&Wt)[NWL], int c0) __attribute__((always_inline)) {
    const bool up = SRC == 2 && c0 < a.cu;  // uniform: upsampled part of the concatenation
#pragma unroll
    for (int k = 0; k < NXL; ++k) {
      if constexpr (SRC == 3) {
        X[k][0] = u32x4{__float_as_uint(Rin[xo[k][0]]), __float_as_uint(Tin[xo[k][0]]), 0u, 0u};
      } else {
#pragma unroll
        for (int j = 0; j < NQ; ++j)
          X[k][j] = *reinterpret_cast<const u32x4*>((up ? img2 + xo2[k] : img + xo[k][j]) + c0);
      }
    }
#pragma unroll
    for (int k = 0; k < NWL; ++k) Wt[k] = *reinterpret_cast<const u32x4*>(a.w + wo[k] + c0);
    // keep the slice's loads together, in issue order
    __builtin_amdgcn_sched_barrier(0);
  };
  float s_in = 1.0f;  // src = 3: the instance's s
  // registers -> LDS (pooling and zero padding applied here)
  auto stash = [&](const u32x4 (&X)[NXL][NQ], const u32x4 (&Wt)[NWL]) __attribute__((always_inline)) {
#pragma unroll
    for (int k = 0; k < NXL; ++k) {
      const int e = tid + k * 256;
      u32x4 v;
      if constexpr (SRC == 3) {
        // channels 0..7 of the pixel: R / s, T, 0 ...; channels 8.. zero
        const unsigned lo = f2bf(__uint_as_float(X[k][0][0]) / s_in);
        const unsigned hi = f2bf(__uint_as_float(X[k][0][1]));
        v = (e & 3) == 0 ? u32x4{lo | (hi << 16), 0u, 0u, 0u} : u32x4{0u, 0u, 0u, 0u};
      } else {
        v = SRC == 1 ? pool4(X[k][0], X[k][NQ > 1 ? 1 : 0], X[k][NQ > 2 ? 2 : 0],
                             X[k][NQ > 3 ? 3 : 0])
                     : X[k][0];
      }
      if (!((okx >> k) & 1u)) v = u32x4{0u, 0u, 0u, 0u};
      const int row = e < RP * 4 ? (e >> 2) : RP;
      *reinterpret_cast<u32x4*>(sx + row * kLdsRow + (e & 3) * 8) = v;
    }
#pragma unroll
    for (int k = 0; k < NWL; ++k) {
      const int e = tid + k * 256;
      const int r = e >> 2;
      const int col = r / 9, tap = r - col * 9;
      const int row = e < 9 * NT * 4 ? tap * NT + col : 9 * NT;
      *reinterpret_cast<u32x4*>(sw + row * kLdsRow + (e & 3) * 8) = Wt[k];
    }
  };
  // one slice's MFMAs from LDS
  auto compute = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int dy = tap / 3, dx = tap - (tap / 3) * 3;
      bf16x8 Bf[NS];
#pragma unroll
      for (int j = 0; j < NS; ++j)
        Bf[j] = *reinterpret_cast<const bf16x8*>(sw + (tap * NT + 16 * j + (lane & 15)) * kLdsRow + kq);
#pragma unroll
      for (int m = 0; m < MS; ++m) {
        const int row = wv * WR + m / SX;
        const int col = (m % SX) * 16 + (lane & 15);
        const bf16x8 A =
            *reinterpret_cast<const bf16x8*>(sx + ((row + dy) * RW + col + dx) * kLdsRow + kq);
#pragma unroll
        for (int j = 0; j < NS; ++j)
          acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, Bf[j], acc[m][j], 0, 0, 0);
      }
    }
  };
  // the next slice is loaded while the current one is multiplied (none
  // after the last: a wave retires only once its loads have landed)
  u32x4 rx[NXL][NQ], rw[NWL];
  fetch(rx, rw, 0);
  if constexpr (SRC == 3) {
    // s = sqrt(mean R^2) + 1e-30 of the instance, its partial sums in the
    // order and reduction tree of k_ws_input
    __shared__ double ssum[16];
    __shared__ float s_sh;
    double acc0 = 0.0;
    for (int k = tid; k < a.nblk; k += 256) acc0 += a.part[(long)b * a.nblk + k];
    const double tot = block_sum(acc0, ssum);
    if (tid == 0) {
      s_sh = (float)(sqrt(tot / (double)((long)a.H * a.W)) + 1e-30);
      if (tw == 0 && th == 0 && blockIdx.y == 0) a.s_out[b] = s_sh;
    }
    __syncthreads();
    s_in = s_sh;
  }
  for (int s = 0; s < ns; ++s) {
    __syncthreads();  // every wave is done with the previous slice
    stash(rx, rw);
    __syncthreads();
    if (s + 1 < ns) fetch(rx, rw, (s + 1) * 32);
    compute();
  }
  if constexpr (HEAD) {
    // the workgroup's activations (NT = co channels) into LDS, then one
    // thread per pixel: head bias + the channel sum in channel order
    static_assert(PH * PW * NT <= (RP + 1) * kLdsRow, "head tile in LDS");
    __syncthreads();  // every wave is done reading the last slice
    bf16_t* ht = sx;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const int c = 16 * j + (lane & 15);
      const float bv = a.bias ? a.bias[co0 + c] : 0.0f;
#pragma unroll
      for (int m = 0; m < MS; ++m)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int lp = (wv * WR + m / SX) * PW + (m % SX) * 16 + (lane >> 4) * 4 + i;
          const float v = acc[m][j][i] + bv;
          ht[lp * NT + c] = f2bf(a.act ? gelu(v) : v);
        }
    }
    __syncthreads();
    for (int lp = tid; lp < PH * PW; lp += 256) {
      const int hh = h0 + lp / PW, ww = w0 + lp % PW;
      if (hh >= a.H || ww >= a.W) continue;
      float v = a.hb[0];
      for (int c = 0; c < NT; ++c) v = head_acc(v, bf2f(ht[lp * NT + c]), a.hw[c]);
      a.hout[((long)b * a.H + hh) * a.W + ww] = v * a.hs[b] * a.hscale;
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const int co = co0 + 16 * j + (lane & 15);
    const float bv = a.bias ? a.bias[co] : 0.0f;
#pragma unroll
    for (int m = 0; m < MS; ++m) {
      const int hh = h0 + wv * WR + m / SX;
      if (hh >= a.H) continue;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int ww = w0 + (m % SX) * 16 + (lane >> 4) * 4 + i;
        if (ww >= a.W) continue;
        const long p = ((long)b * a.H + hh) * a.W + ww;
        const float v = acc[m][j][i] + bv;
        if (a.pre) a.pre[p * a.ldp + co] = f2bf(v);
        if (a.out) a.out[p * a.ldo + co] = f2bf(a.act ? gelu(v) : v);
        if (a.outf) {
          float* o = a.outf + p * a.ldf + co;
          *o = a.accum ? *o + v : v;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[co][ci][kh][kw] += sum_p g[p][co] * x[p + tap][ci]
// (fp32, torch OIHW layout, atomics across pixel chunks).
// A (16 co x 32 pixels) and B (32 pixels x 16 ci) are staged transposed in
// LDS from 16-byte loads: sg[co][px], and sx[dw][ci][row][px], the 3 x 34
// input pixels the taps of 32 pixels of one image row touch, stored once per
// tap column offset dw so that every fragment is one aligned 16-byte read.
// The next segment's loads are issued before the current segment's MFMAs.
// Wave wv owns output channels co0 + 16 wv .. +15 and keeps 9 taps x 2 ci
// sub-tiles of accumulators.
// ---------------------------------------------------------------------------
struct DwArgs {
  const bf16_t* g;   // dPre, NHWC, stride ldg, co channels
  int ldg, co;
  const bf16_t* x;   // conv input, NHWC, stride ldx, ci_real channels read
  int ldx, ci;       // ci: channels of the weight (real, not padded)
  float* dw;         // [co][ci][3][3]
  int B, H, W;
  int rows_per_block;
};

__global__ __launch_bounds__(256) void k_conv_dw(DwArgs a) {
  // sg[co][px]: 32 pixels of dPre per output channel (rows of 40: 16-byte
  // aligned fragments); sx[dw][ci][dh][px]: the input row segment shifted by
  // the tap's column offset dw, so every B fragment is one aligned read
  __shared__ __attribute__((aligned(16))) bf16_t sg[64][40];
  __shared__ __attribute__((aligned(16))) bf16_t sx[3][32][3][40];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int co0 = blockIdx.y * 64, ci0 = blockIdx.z * 32;
  const int nseg = (a.W + 31) / 32;
  const long nrows = (long)a.B * a.H;
  floatx4 acc[9][2];
#pragma unroll
  for (int t = 0; t < 9; ++t) acc[t][0] = acc[t][1] = floatx4{0.f, 0.f, 0.f, 0.f};
  const long r0 = (long)blockIdx.x * a.rows_per_block;
  const long r1 = r0 + a.rows_per_block < nrows ? r0 + a.rows_per_block : nrows;
  const long nsegs = (r1 > r0) ? (r1 - r0) * nseg : 0;
  // per segment and thread: one 8-channel dPre load, two 8-channel input loads
  uint4 rg, rx0, rx1;
  auto fetch = [&](long k) {
    const long row = r0 + k / nseg;
    const int w0 = (int)(k % nseg) * 32;
    const int b = (int)(row / a.H), h = (int)(row - (long)b * a.H);
    {
      const int px = tid >> 3, cg = (tid & 7) * 8;
      const int ww = w0 + px, co = co0 + cg;
      rg = make_uint4(0, 0, 0, 0);
      if (ww < a.W && co < a.co)
        rg = *reinterpret_cast<const uint4*>(a.g + ((long)(b * a.H + h) * a.W + ww) * a.ldg + co);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int e = tid + u * 256;  // 3 rows x 34 px x 4 channel groups
      uint4 v = make_uint4(0, 0, 0, 0);
      if (e < 3 * 34 * 4) {
        const int cg = (e & 3) * 8, q = e >> 2;
        const int dh = q / 34, px = q - dh * 34;
        const int hh = h + dh - 1, ww = w0 + px - 1, ci = ci0 + cg;
        if (hh >= 0 && hh < a.H && ww >= 0 && ww < a.W && ci < a.ci)
          v = *reinterpret_cast<const uint4*>(a.x + ((long)(b * a.H + hh) * a.W + ww) * a.ldx + ci);
      }
      if (u == 0) rx0 = v; else rx1 = v;
    }
  };
  auto stash = [&]() {
    {
      const int px = tid >> 3, cg = (tid & 7) * 8;
      const bf16_t* s = reinterpret_cast<const bf16_t*>(&rg);
#pragma unroll
      for (int k = 0; k < 8; ++k) sg[cg + k][px] = s[k];
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int e = tid + u * 256;
      if (e < 3 * 34 * 4) {
        const int cg = (e & 3) * 8, q = e >> 2;
        const int dh = q / 34, px = q - dh * 34;  // px = column + 1 of the segment
        const uint4 v = (u == 0) ? rx0 : rx1;
        const bf16_t* s = reinterpret_cast<const bf16_t*>(&v);
        // copy dw holds column p + dw - 1 at position p (p = px - dw)
#pragma unroll
        for (int dw = 0; dw < 3; ++dw) {
          const int p = px - dw;
          if (p >= 0 && p < 32)
#pragma unroll
            for (int k = 0; k < 8; ++k) sx[dw][cg + k][dh][p] = s[k];
        }
      }
    }
  };
  if (nsegs > 0) fetch(0);
  const int kq = (lane >> 4) * 8;
  const bool active = co0 + wv * 16 < a.co;
  for (long k = 0; k < nsegs; ++k) {
    if (k > 0) __syncthreads();  // the previous segment's fragments are read
    stash();
    __syncthreads();
    if (k + 1 < nsegs) fetch(k + 1);
    if (!active) continue;
    const bf16x8 A = *reinterpret_cast<const bf16x8*>(&sg[wv * 16 + (lane & 15)][kq]);
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int dh = t / 3, dw = t % 3;
#pragma unroll
      for (int nsub = 0; nsub < 2; ++nsub) {
        const bf16x8 Bm =
            *reinterpret_cast<const bf16x8*>(&sx[dw][nsub * 16 + (lane & 15)][dh][kq]);
        acc[t][nsub] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, Bm, acc[t][nsub], 0, 0, 0);
      }
    }
  }
  if (co0 + wv * 16 >= a.co) return;
  // D (16 co x 16 ci): lane holds ci lane%16, co rows 4*(lane/16)..+3
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int nsub = 0; nsub < 2; ++nsub) {
      const int ci = ci0 + nsub * 16 + (lane & 15);
      if (ci >= a.ci) continue;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int co = co0 + wv * 16 + (lane >> 4) * 4 + i;
        if (co >= a.co) continue;
        atomicAdd(&a.dw[((long)co * a.ci + ci) * 9 + t], acc[t][nsub][i]);
      }
    }
}

// bias gradient db[c] += sum_p g[p][c]
__global__ __launch_bounds__(256) void k_bias_grad(const bf16_t* g, int ldg, int C, long P,
                                                   float* db, int chunk) {
  const int c = threadIdx.x % C;  // C <= 256 divides 256 or C == 256
  const int lanes_per_c = 256 / C;
  const int sub = threadIdx.x / C;
  const long p0 = (long)blockIdx.x * chunk;
  const long p1 = p0 + chunk < P ? p0 + chunk : P;
  float acc = 0.f;
  for (long p = p0 + sub; p < p1; p += lanes_per_c) acc += bf2f(g[p * ldg + c]);
  atomicAdd(&db[c], acc);
}

// dPre = dPost * gelu'(pre)  (fp32 in, bf16 out)
__global__ __launch_bounds__(256) void k_gelu_back(const float* dpost, int ldd,
                                                   const bf16_t* pre, int ldp, bf16_t* dpre,
                                                   int ldo, int C, long P) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= P * C) return;
  const long p = e / C;
  const int c = (int)(e - p * C);
  const float x = bf2f(pre[p * ldp + c]);
  dpre[p * ldo + c] = f2bf(dpost[p * ldd + c] * gelu_grad(x));
}

// 2x2 average pool of a channel slice: out[B][H/2][W/2][C]
__global__ __launch_bounds__(256) void k_pool(const bf16_t* in, int ldi, bf16_t* out, int ldo,
                                              int C, int B, int H, int W) {
  const int Ho = H / 2, Wo = W / 2;
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long)B * Ho * Wo * C) return;
  const long po = e / C;
  const int c = (int)(e - po * C);
  const int b = (int)(po / (Ho * Wo));
  const int r = (int)(po - (long)b * Ho * Wo);
  const int ho = r / Wo, wo = r - ho * Wo;
  const long p00 = ((long)(b * H + 2 * ho) * W + 2 * wo);
  const float s = bf2f(in[p00 * ldi + c]) + bf2f(in[(p00 + 1) * ldi + c]) +
                  bf2f(in[(p00 + W) * ldi + c]) + bf2f(in[(p00 + W + 1) * ldi + c]);
  out[po * ldo + c] = f2bf(0.25f * s);
}

// pool backward: din[2ho+a][2wo+b][c] += dout[ho][wo][c] / 4  (fp32)
__global__ __launch_bounds__(256) void k_pool_back(const float* dout, int ldo, float* din,
                                                   int ldi, int C, int B, int H, int W) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long)B * H * W * C) return;
  const long p = e / C;
  const int c = (int)(e - p * C);
  const int b = (int)(p / ((long)H * W));
  const int r = (int)(p - (long)b * H * W);
  const int h = r / W, w = r - h * W;
  const long po = ((long)(b * (H / 2) + h / 2) * (W / 2) + w / 2);
  din[p * ldi + c] += 0.25f * dout[po * ldo + c];
}

// nearest 2x upsample: out[h][w][c] = in[h/2][w/2][c]
__global__ __launch_bounds__(256) void k_up(const bf16_t* in, int ldi, bf16_t* out, int ldo,
                                            int C, int B, int H, int W) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long)B * H * W * C) return;
  const long p = e / C;
  const int c = (int)(e - p * C);
  const int b = (int)(p / ((long)H * W));
  const int r = (int)(p - (long)b * H * W);
  const int h = r / W, w = r - h * W;
  const long pi = ((long)(b * (H / 2) + h / 2) * (W / 2) + w / 2);
  out[p * ldo + c] = in[pi * ldi + c];
}

// upsample backward: din[ho][wo][c] = sum of the 2x2 children (fp32)
__global__ __launch_bounds__(256) void k_up_back(const float* dout, int ldo, float* din, int ldi,
                                                 int C, int B, int Ho, int Wo) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long)B * Ho * Wo * C) return;
  const long po = e / C;
  const int c = (int)(e - po * C);
  const int b = (int)(po / ((long)Ho * Wo));
  const int r = (int)(po - (long)b * Ho * Wo);
  const int ho = r / Wo, wo = r - ho * Wo;
  const int W = 2 * Wo;
  const long p00 = ((long)(b * 2 * Ho + 2 * ho) * W + 2 * wo);
  din[po * ldi + c] = dout[p00 * ldo + c] + dout[(p00 + 1) * ldo + c] +
                      dout[(p00 + W) * ldo + c] + dout[(p00 + W + 1) * ldo + c];
}

// 1x1 head: out[p] = sum_c x[p][c] w[c] + b
__global__ __launch_bounds__(256) void k_head(const bf16_t* x, int ldx, int C, const float* w,
                                              const float* b, float* out, long P) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= P) return;
  float s = b[0];
  for (int c = 0; c < C; ++c) s += bf2f(x[p * ldx + c]) * w[c];
  out[p] = s;
}

// head backward: dx[p][c] = dout[p] w[c] (fp32), dw[c] += sum_p dout[p] x[p][c], db += sum dout
__global__ __launch_bounds__(256) void k_head_back(const bf16_t* x, int ldx, int C, const float* w,
                                                   const float* dout, float* dx, int ldd,
                                                   float* dw, float* db, long P) {
  __shared__ float sw[256];
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  const float g = p < P ? dout[p] : 0.f;
  for (int c = 0; c < C; ++c) {
    if (p < P) dx[p * ldd + c] = g * w[c];
    float v = p < P ? g * bf2f(x[p * ldx + c]) : 0.f;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&dw[c], sw[0] + sw[1] + sw[2] + sw[3]);
    __syncthreads();
  }
  float v = g;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(db, v);
}

}  // namespace nn
}  // namespace aj
