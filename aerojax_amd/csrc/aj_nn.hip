// Pressure warm-start network: a 4-level convolutional encoder-decoder
// (U-Net, 3x3 convs, GELU, 2x2 average pooling, nearest upsampling, skip
// connections, 1x1 head) mapping (rhs, sdf) to p.  BF16 weights and
// activations, FP32 accumulation (MFMA), FP32 master weights and Adam.
//
// The reference's learned pressure operators (neural_operators/*.py,
// trained in neural_operators/training_utils.py:84-253) replace the solve;
// here the network only provides the initial guess of the multigrid solve,
// which then runs to the same tolerance (and falls back to a zero guess
// when the network's guess has the larger residual).  Conventions kept from
// the reference training loop: MSE loss, elementwise gradient clipping to
// +-1, Adam.
#include "aj_nn_kernels.h"

#include <algorithm>
#include <cstdio>

#include <cmath>
#include <cstring>
#include <vector>

using namespace aj;
using namespace aj::nn;

namespace {

struct Layer {
  int ci, ci_pad, co;
  long w_off, b_off;  // offsets into the flat parameter vector (OIHW, bias)
  bf16_t* wpk;        // [co][9][ci_pad] bf16
  bf16_t* wtk;        // [ci][9 (flipped)][co] bf16, input-gradient weights
};

struct Act {  // activations and gradients of one batch shape
  int B = 0, H = 0, W = 0;
  void* mem = nullptr;
  size_t bytes = 0;
  bf16_t* X0 = nullptr;
  std::vector<bf16_t*> A, PA, PB, CAT, POOL, DA, DPA, DB, DPB;
  bf16_t* Ebot = nullptr;
  float* out = nullptr;
  float* target = nullptr;
  float* dout = nullptr;
  bf16_t* dpre = nullptr;
  std::vector<float*> dA, dCAT, dPOOL, dDA, dDB;
  float* dEbot = nullptr;
  float* loss = nullptr;
};

}  // namespace

struct aj_nn {
  int c0 = 32, L = 4, cin = 2;
  int device = 0;
  int fuse_io = 1;  // AJ_NN_OPT_FUSED_IO
  std::vector<Layer> layers;  // enc (2 per level), dec (2 per level above the bottom)
  long head_w = 0, head_b = 0, nparams = 0;
  float* params = nullptr;
  float* grads = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  void* wmem = nullptr;
  long step = 0;
  Act act;
  // warm-start scratch (per fine-grid shape)
  void* ws_mem = nullptr;
  size_t ws_bytes = 0;
  // captured warm starts, one per call signature (buffers, shape, scales):
  // the first call with a signature runs eagerly (the launch tuner decides
  // every shape), the second is captured, later ones replay the graph
  struct WsGraph {
    std::vector<uint64_t> key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<WsGraph> ws_graphs;
  std::vector<std::vector<uint64_t>> ws_seen;
};

namespace {

constexpr int kCinPad = 32;

int chan(const aj_nn* n, int l) { return n->c0 << l; }

// layer indices
int enc_a(int l) { return 2 * l; }
int enc_b(int l) { return 2 * l + 1; }
int dec_a(const aj_nn* n, int l) { return 2 * n->L + 2 * (n->L - 2 - l); }
int dec_b(const aj_nn* n, int l) { return dec_a(n, l) + 1; }

unsigned long long splitmix(unsigned long long& s) {
  unsigned long long z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__global__ void k_pack_w(const float* params, long w_off, int co, int ci, int ci_pad, bf16_t* wpk,
                         bf16_t* wtk) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;  // over co*ci_pad*9
  if (e >= (long)co * ci_pad * 9) return;
  const int t = (int)(e % 9);
  const long q = e / 9;
  const int c = (int)(q % ci_pad);
  const int o = (int)(q / ci_pad);
  const float v = c < ci ? params[w_off + ((long)o * ci + c) * 9 + t] : 0.0f;
  wpk[((long)o * 9 + t) * ci_pad + c] = f2bf(v);
  if (wtk && c < ci) wtk[((long)c * 9 + (8 - t)) * co + o] = f2bf(v);
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1,
                       float b2, float eps, float bc1, float bc2, float clip) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  gi = gi > clip ? clip : (gi < -clip ? -clip : gi);
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// gather samples idx[0..B) of x_all [N][H][W][2] / t_all [N][H][W] into the
// padded bf16 input and the fp32 target
__global__ void k_gather(const float* x_all, const float* t_all, const int* idx, int B, long HW,
                         bf16_t* X0, float* target) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long)B * HW) return;
  const int b = (int)(e / HW);
  const long p = e - (long)b * HW;
  const long s = (long)(idx ? idx[b] : b) * HW + p;
  bf16_t* o = X0 + e * kCinPad;
  o[0] = f2bf(x_all[2 * s]);
  o[1] = f2bf(x_all[2 * s + 1]);
  for (int c = 2; c < kCinPad; ++c) o[c] = 0;
  if (target) target[e] = t_all ? t_all[s] : 0.f;
}

__global__ void k_mse(const float* out, const float* target, float* dout, long P, float* loss) {
  __shared__ double sh[16];
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  double acc = 0.0;
  if (p < P) {
    const float d = out[p] - target[p];
    dout[p] = 2.f * d / (float)P;
    acc = (double)d * d;
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) atomicAdd(loss, (float)(t / (double)P));
}

__global__ void k_zero(float* p, long n) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

unsigned blocks(long n) { return (unsigned)((n + 255) / 256); }

template <typename X>
X* carve(char*& base, size_t& off, size_t count, bool alloc) {
  X* r = alloc ? reinterpret_cast<X*>(base + off) : nullptr;
  off += (count * sizeof(X) + 255) & ~size_t(255);
  return r;
}

// activation buffers for a (B, H, W) input, carved from one allocation that
// only grows (callers alternating between input shapes reuse it)
int ensure_act(aj_nn* n, int B, int H, int W) {
  Act& a = n->act;
  if (a.B == B && a.H == H && a.W == W && a.mem) return AJ_OK;
  if (H % (1 << (n->L - 1)) || W % (1 << (n->L - 1)))
    return fail(AJ_ERR_INVALID, "network input H, W must be divisible by 2^(levels-1)");
  const int L = n->L;
  for (int pass = 0; pass < 2; ++pass) {
    const bool al = pass == 1;
    char* base = (char*)a.mem;
    size_t off = 0;
    auto P = [&](int l) { return (size_t)B * (H >> l) * (W >> l); };
    a.A.assign(L, nullptr); a.PA.assign(L, nullptr); a.PB.assign(L, nullptr);
    a.CAT.assign(L, nullptr); a.POOL.assign(L, nullptr);
    a.DA.assign(L, nullptr); a.DPA.assign(L, nullptr); a.DB.assign(L, nullptr);
    a.DPB.assign(L, nullptr);
    a.dA.assign(L, nullptr); a.dCAT.assign(L, nullptr); a.dPOOL.assign(L, nullptr);
    a.dDA.assign(L, nullptr); a.dDB.assign(L, nullptr);
    a.X0 = carve<bf16_t>(base, off, P(0) * kCinPad, al);
    for (int l = 0; l < L; ++l) {
      const int c = chan(n, l);
      a.A[l] = carve<bf16_t>(base, off, P(l) * c, al);
      a.PA[l] = carve<bf16_t>(base, off, P(l) * c, al);
      a.PB[l] = carve<bf16_t>(base, off, P(l) * c, al);
      a.dA[l] = carve<float>(base, off, P(l) * c, al);
      if (l > 0) {
        a.POOL[l] = carve<bf16_t>(base, off, P(l) * chan(n, l - 1), al);
        a.dPOOL[l] = carve<float>(base, off, P(l) * chan(n, l - 1), al);
      }
      if (l < L - 1) {
        const int cc = chan(n, l + 1) + c;
        a.CAT[l] = carve<bf16_t>(base, off, P(l) * cc, al);
        a.dCAT[l] = carve<float>(base, off, P(l) * cc, al);
        a.DA[l] = carve<bf16_t>(base, off, P(l) * c, al);
        a.DPA[l] = carve<bf16_t>(base, off, P(l) * c, al);
        a.DB[l] = carve<bf16_t>(base, off, P(l) * c, al);
        a.DPB[l] = carve<bf16_t>(base, off, P(l) * c, al);
        a.dDA[l] = carve<float>(base, off, P(l) * c, al);
        a.dDB[l] = carve<float>(base, off, P(l) * c, al);
      }
    }
    a.Ebot = carve<bf16_t>(base, off, P(L - 1) * chan(n, L - 1), al);
    a.dEbot = carve<float>(base, off, P(L - 1) * chan(n, L - 1), al);
    a.out = carve<float>(base, off, P(0), al);
    a.target = carve<float>(base, off, P(0), al);
    a.dout = carve<float>(base, off, P(0), al);
    a.dpre = carve<bf16_t>(base, off, P(0) * 4 * n->c0, al);
    a.loss = carve<float>(base, off, 64, al);
    if (!al && (!a.mem || off > a.bytes)) {
      if (a.mem) (void)hipFree(a.mem);
      a.mem = nullptr;
      a.bytes = 0;
      AJ_HIP(hipMalloc(&a.mem, off));
      a.bytes = off;
    }
  }
  a.B = B;
  a.H = H;
  a.W = W;
  return AJ_OK;
}

// input mode of a forward convolution (ConvArgs::src and its operands)
struct ConvSrc {
  int mode = 0;
  const bf16_t* in2 = nullptr;
  int ld2 = 0, cu = 0;
  // mode 3: the network input row from R, T and the partial sums of R^2
  const float* inR = nullptr;
  const float* inT = nullptr;
  const double* part = nullptr;
  int nblk = 0;
  float* s_out = nullptr;
};

// the 1x1 head applied in a convolution's epilogue (its activation is then
// not stored): weights, bias, per-instance s, pscale, output
struct HeadOut {
  const float* w;
  const float* b;
  const float* s;
  float scale;
  float* out;
};

// tile choice: 8x32-pixel patches where the layer has enough of them to
// fill the chip, 4x16 otherwise; NT = 64 / 32 / 16 output channels
// the head epilogue stages the tile's activations in the input patch's LDS
template <int PH, int PW, int NT>
constexpr bool kHeadFits = PH * PW * NT <= ((PH + 2) * (PW + 2) + 1) * kLdsRow;

template <int PH, int PW, int NT, int SRC, bool HEAD>
static int conv_launch_nt(const ConvArgs& c, dim3 grid, hipStream_t st) {
  if constexpr (HEAD && !kHeadFits<PH, PW, NT>) {
    (void)c, (void)grid, (void)st;
    return fail(AJ_ERR_INVALID, "conv: head tile larger than the LDS patch");
  } else {
    hipLaunchKernelGGL((k_conv3x3<PH, PW, NT, SRC, HEAD>), grid, dim3(256), 0, st, c);
    AJ_LAUNCH_CHECK("k_conv3x3");
    return AJ_OK;
  }
}

template <int PH, int PW, int SRC, bool HEAD = false>
static int conv_launch_src(const ConvArgs& c, int nt, hipStream_t st) {
  const long tiles = (long)c.B * ((c.H + PH - 1) / PH) * ((c.W + PW - 1) / PW);
  const dim3 grid((unsigned)tiles, c.co / nt);
  if (nt == 64) return conv_launch_nt<PH, PW, 64, SRC, HEAD>(c, grid, st);
  if (nt == 32) return conv_launch_nt<PH, PW, 32, SRC, HEAD>(c, grid, st);
  return conv_launch_nt<PH, PW, 16, SRC, HEAD>(c, grid, st);
}

template <int PH, int PW>
static int conv_launch_t(const ConvArgs& c, int nt, hipStream_t st) {
  if (c.hout) return conv_launch_src<PH, PW, 0, true>(c, nt, st);  // plain input only
  if (c.src == 1) return conv_launch_src<PH, PW, 1>(c, nt, st);
  if (c.src == 2) return conv_launch_src<PH, PW, 2>(c, nt, st);
  if (c.src == 3) return conv_launch_src<PH, PW, 3>(c, nt, st);
  return conv_launch_src<PH, PW, 0>(c, nt, st);
}

// candidates: pixel tiles 8x32 and 4x16 with the widest output-channel
// block NT that divides co, then 4x16 with narrower blocks (more workgroups
// for layers with few pixel tiles), then 8x16 with the widest and the next
// block (two 16-pixel rows per wave: each LDS fragment feeds two MFMAs,
// where the 4x16 tiles' A fragments feed NT / 16); tuned per layer shape.
// Launches that accumulate into their output (input gradients of a
// concatenated input) reuse the choice of the plain launch of the same
// shape.  A launch with the head epilogue needs every output channel in one
// workgroup: its candidates are the three pixel tiles with NT = co.
int conv_launch(const ConvArgs& c, hipStream_t st) {
  if (c.ci % 32 || c.co % 16) return fail(AJ_ERR_INVALID, "conv: ci % 32 or co % 16 != 0");
  if (c.hout && (c.src != 0 || (c.co != 16 && c.co != 32 && c.co != 64)))
    return fail(AJ_ERR_INVALID, "conv: the head epilogue needs a plain input and co of 16, 32 or 64");
  const int nt = (c.co % 64 == 0) ? 64 : (c.co % 32 == 0 ? 32 : 16);
  enum { kT8x32 = 0, kT4x16 = 1, kT8x16 = 2 };
  int cands[6][2] = {{kT8x32, nt}, {kT4x16, nt}};
  int ncand = 2;
  if (c.hout) {
    cands[ncand][0] = kT8x16;
    cands[ncand++][1] = nt;
  } else {
    for (int w = nt / 2; w >= 16; w /= 2) {
      cands[ncand][0] = kT4x16;
      cands[ncand++][1] = w;
    }
    for (int w = nt; w >= 16 && w >= nt / 2; w /= 2) {
      cands[ncand][0] = kT8x16;
      cands[ncand++][1] = w;
    }
  }
  auto launch = [&](int k, hipStream_t s) -> int {
    switch (cands[k][0]) {
      case kT4x16: return conv_launch_t<4, 16>(c, cands[k][1], s);
      case kT8x16: return conv_launch_t<8, 16>(c, cands[k][1], s);
      default: return conv_launch_t<8, 32>(c, cands[k][1], s);
    }
  };
  char key[96];
  std::snprintf(key, sizeof(key), "conv/%dx%dx%d/%d->%d/%s%s%s", c.B, c.H, c.W, c.ci, c.co,
                c.outf ? "f32" : "bf16",
                c.src == 1 ? "/pool" : (c.src == 2 ? "/up" : (c.src == 3 ? "/in" : "")),
                c.hout ? "/head" : "");
  int pick = 0;
  return tune_pick(key, c.accum ? 1 : ncand, launch, st, &pick);
}

// forward conv + GELU: out (post) and pre both stored
int conv_fwd(aj_nn* n, int li, const bf16_t* in, int ldi, bf16_t* out, int ldo, bf16_t* pre,
             int B, int H, int W, hipStream_t st, const ConvSrc& src = ConvSrc(),
             const HeadOut* ho = nullptr) {
  const Layer& Ly = n->layers[li];
  ConvArgs c{};
  c.in = in; c.ldi = ldi; c.ci = Ly.ci_pad;
  c.w = Ly.wpk; c.bias = n->params + Ly.b_off; c.co = Ly.co;
  c.B = B; c.H = H; c.W = W;
  c.out = out; c.ldo = ldo; c.pre = pre; c.ldp = Ly.co;
  c.outf = nullptr; c.ldf = 0; c.accum = 0; c.act = 1;
  c.src = src.mode; c.in2 = src.in2; c.ld2 = src.ld2; c.cu = src.cu;
  c.inR = src.inR; c.inT = src.inT; c.part = src.part; c.nblk = src.nblk; c.s_out = src.s_out;
  if (ho) {
    c.out = nullptr;
    c.hw = ho->w; c.hb = ho->b; c.hs = ho->s; c.hscale = ho->scale; c.hout = ho->out;
  }
  ProfScope prof(AJ_PROF_NN_CONV, 2.0 * 9 * Ly.ci * Ly.co * (double)B * H * W, st);
  return conv_launch(c, st);
}

// backward of one conv: gpost (fp32, gradient of the post-activation) ->
// dPre, dW, db, and (if dx) the input gradient written/accumulated to dx
int conv_bwd(aj_nn* n, int li, const float* gpost, int ldg, const bf16_t* pre, const bf16_t* in,
             int ldi, float* dx, int ldx, int accum, int B, int H, int W, hipStream_t st) {
  const Layer& Ly = n->layers[li];
  const long P = (long)B * H * W;
  hipLaunchKernelGGL(k_gelu_back, dim3(blocks(P * Ly.co)), dim3(256), 0, st, gpost, ldg, pre,
                     Ly.co, n->act.dpre, Ly.co, Ly.co, P);
  AJ_LAUNCH_CHECK("k_gelu_back");
  {
    DwArgs d{};
    d.g = n->act.dpre; d.ldg = Ly.co; d.co = Ly.co;
    d.x = in; d.ldx = ldi; d.ci = Ly.ci;
    d.dw = n->grads + Ly.w_off;
    d.B = B; d.H = H; d.W = W;
    // about 2048 workgroups: enough to fill the chip, few enough that the
    // per-workgroup atomic adds of the 9 co x ci partial sums stay cheap
    const long rows = (long)B * H;
    const long yz = (long)((Ly.co + 63) / 64) * ((Ly.ci + 31) / 32);
    const long nrb = std::max<long>(1, std::min<long>(rows, 2048 / yz));
    d.rows_per_block = (int)((rows + nrb - 1) / nrb);
    dim3 grid((unsigned)((rows + d.rows_per_block - 1) / d.rows_per_block), (Ly.co + 63) / 64,
              (Ly.ci + 31) / 32);
    ProfScope prof(AJ_PROF_NN_DW, 2.0 * 9 * Ly.ci * Ly.co * (double)P, st);
    hipLaunchKernelGGL(k_conv_dw, grid, dim3(256), 0, st, d);
    AJ_LAUNCH_CHECK("k_conv_dw");
  }
  {
    const int chunk = 4096;
    hipLaunchKernelGGL(k_bias_grad, dim3((unsigned)((P + chunk - 1) / chunk)), dim3(256), 0, st,
                       (const bf16_t*)n->act.dpre, Ly.co, Ly.co, P, n->grads + Ly.b_off, chunk);
    AJ_LAUNCH_CHECK("k_bias_grad");
  }
  if (dx) {
    ConvArgs c{};
    c.in = n->act.dpre; c.ldi = Ly.co; c.ci = Ly.co;
    c.w = Ly.wtk; c.bias = nullptr; c.co = Ly.ci;
    c.B = B; c.H = H; c.W = W;
    c.out = nullptr; c.pre = nullptr;
    c.outf = dx; c.ldf = ldx; c.accum = accum; c.act = 0;
    ProfScope prof(AJ_PROF_NN_CONV, 2.0 * 9 * Ly.ci * Ly.co * (double)P, st);
    if (int rc = conv_launch(c, st)) return rc;
  }
  return AJ_OK;
}

// fused = true (inference): pooling and upsampling are read by the
// convolutions directly instead of being materialised, and the
// pre-activations are not stored (the backward pass needs both, so
// training uses fused = false)
// in0: the first convolution's input mode instead of the staged input row
// (warm start: the row is built in its loads); ho: the 1x1 head in the last
// convolution's epilogue (warm start; its activation is then not stored)
int forward(aj_nn* n, int B, int H, int W, hipStream_t st, bool head = true,
            bool fused = false, const ConvSrc* in0 = nullptr, const HeadOut* ho = nullptr) {
  Act& a = n->act;
  const int L = n->L;
  auto pre = [&](bf16_t* buf) -> bf16_t* { return fused ? nullptr : buf; };
  for (int l = 0; l < L; ++l) {
    const int Hl = H >> l, Wl = W >> l, c = chan(n, l);
    const long P = (long)B * Hl * Wl;
    const bf16_t* in;
    int ldi;
    ConvSrc src{};
    if (l == 0) {
      in = a.X0;
      ldi = kCinPad;
      if (in0) src = *in0;
    }
    else if (fused) {
      // 2x2 average of the skip of level l-1 (channels chan(l) .. of CAT[l-1])
      const int cp = chan(n, l - 1);
      in = a.CAT[l - 1] + chan(n, l);
      ldi = chan(n, l) + cp;
      src.mode = 1;
    } else {
      // pool the skip of level l-1 (channels chan(l) .. of CAT[l-1])
      const int cp = chan(n, l - 1);
      const int cc = chan(n, l) + cp;
      hipLaunchKernelGGL(k_pool, dim3(blocks(P * cp)), dim3(256), 0, st,
                         (const bf16_t*)(a.CAT[l - 1] + chan(n, l)), cc, a.POOL[l], cp, cp, B,
                         H >> (l - 1), W >> (l - 1));
      AJ_LAUNCH_CHECK("k_pool");
      in = a.POOL[l];
      ldi = cp;
    }
    if (int rc = conv_fwd(n, enc_a(l), in, ldi, a.A[l], c, pre(a.PA[l]), B, Hl, Wl, st, src)) return rc;
    bf16_t* ob;
    int ldo;
    if (l < L - 1) { ob = a.CAT[l] + chan(n, l + 1); ldo = chan(n, l + 1) + c; }
    else { ob = a.Ebot; ldo = c; }
    if (int rc = conv_fwd(n, enc_b(l), a.A[l], c, ob, ldo, pre(a.PB[l]), B, Hl, Wl, st)) return rc;
  }
  for (int l = L - 2; l >= 0; --l) {
    const int Hl = H >> l, Wl = W >> l, c = chan(n, l), cu = chan(n, l + 1);
    const long P = (long)B * Hl * Wl;
    const bf16_t* lower = (l == L - 2) ? a.Ebot : a.DB[l + 1];
    ConvSrc src{};
    if (fused) {
      src.mode = 2;  // channels < cu: nearest upsampling of `lower`
      src.in2 = lower;
      src.ld2 = cu;
      src.cu = cu;
    } else {
      hipLaunchKernelGGL(k_up, dim3(blocks(P * cu)), dim3(256), 0, st, lower, cu, a.CAT[l],
                         cu + c, cu, B, Hl, Wl);
      AJ_LAUNCH_CHECK("k_up");
    }
    if (int rc = conv_fwd(n, dec_a(n, l), a.CAT[l], cu + c, a.DA[l], c, pre(a.DPA[l]), B, Hl, Wl, st,
                          src))
      return rc;
    if (int rc = conv_fwd(n, dec_b(n, l), a.DA[l], c, a.DB[l], c, pre(a.DPB[l]), B, Hl, Wl, st,
                          ConvSrc(), l == 0 ? ho : nullptr))
      return rc;
  }
  if (!head) return AJ_OK;
  const long P0 = (long)B * H * W;
  hipLaunchKernelGGL(k_head, dim3(blocks(P0)), dim3(256), 0, st, (const bf16_t*)a.DB[0], n->c0,
                     n->c0, (const float*)(n->params + n->head_w),
                     (const float*)(n->params + n->head_b), a.out, P0);
  AJ_LAUNCH_CHECK("k_head");
  return AJ_OK;
}

int backward(aj_nn* n, int B, int H, int W, hipStream_t st) {
  Act& a = n->act;
  const int L = n->L;
  const long P0 = (long)B * H * W;
  hipLaunchKernelGGL(k_zero, dim3(blocks(n->nparams)), dim3(256), 0, st, n->grads, n->nparams);
  AJ_LAUNCH_CHECK("k_zero");
  hipLaunchKernelGGL(k_head_back, dim3(blocks(P0)), dim3(256), 0, st, (const bf16_t*)a.DB[0],
                     n->c0, n->c0, (const float*)(n->params + n->head_w), (const float*)a.dout,
                     a.dDB[0], n->c0, n->grads + n->head_w, n->grads + n->head_b, P0);
  AJ_LAUNCH_CHECK("k_head_back");
  for (int l = 0; l <= L - 2; ++l) {
    const int Hl = H >> l, Wl = W >> l, c = chan(n, l), cu = chan(n, l + 1);
    if (int rc = conv_bwd(n, dec_b(n, l), a.dDB[l], c, a.DPB[l], a.DA[l], c, a.dDA[l], c, 0, B,
                          Hl, Wl, st))
      return rc;
    if (int rc = conv_bwd(n, dec_a(n, l), a.dDA[l], c, a.DPA[l], a.CAT[l], cu + c, a.dCAT[l],
                          cu + c, 0, B, Hl, Wl, st))
      return rc;
    // upsample backward into the gradient of the coarser decoder output
    float* dsrc = (l == L - 2) ? a.dEbot : a.dDB[l + 1];
    const long Pc = (long)B * (Hl / 2) * (Wl / 2);
    hipLaunchKernelGGL(k_up_back, dim3(blocks(Pc * cu)), dim3(256), 0, st,
                       (const float*)a.dCAT[l], cu + c, dsrc, cu, cu, B, Hl / 2, Wl / 2);
    AJ_LAUNCH_CHECK("k_up_back");
  }
  for (int l = L - 1; l >= 0; --l) {
    const int Hl = H >> l, Wl = W >> l, c = chan(n, l);
    const float* gb;
    int ldg;
    if (l == L - 1) { gb = a.dEbot; ldg = c; }
    else { gb = a.dCAT[l] + chan(n, l + 1); ldg = chan(n, l + 1) + c; }
    const bf16_t* ob = (l < L - 1) ? a.CAT[l] + chan(n, l + 1) : a.Ebot;
    (void)ob;
    if (int rc = conv_bwd(n, enc_b(l), gb, ldg, a.PB[l], a.A[l], c, a.dA[l], c, 0, B, Hl, Wl, st))
      return rc;
    if (l == 0) {
      if (int rc = conv_bwd(n, enc_a(0), a.dA[0], c, a.PA[0], a.X0, kCinPad, nullptr, 0, 0, B, Hl,
                            Wl, st))
        return rc;
    } else {
      const int cp = chan(n, l - 1);
      if (int rc = conv_bwd(n, enc_a(l), a.dA[l], c, a.PA[l], a.POOL[l], cp, a.dPOOL[l], cp, 0, B,
                            Hl, Wl, st))
        return rc;
      // pool backward adds into the skip slice of the finer level's gradient
      const int Hf = H >> (l - 1), Wf = W >> (l - 1), cc = c + cp;
      const long Pf = (long)B * Hf * Wf;
      hipLaunchKernelGGL(k_pool_back, dim3(blocks(Pf * cp)), dim3(256), 0, st,
                         (const float*)a.dPOOL[l], cp, a.dCAT[l - 1] + c, cc, cp, B, Hf, Wf);
      AJ_LAUNCH_CHECK("k_pool_back");
    }
  }
  return AJ_OK;
}

int pack_all(aj_nn* n, hipStream_t st) {
  for (size_t i = 0; i < n->layers.size(); ++i) {
    const Layer& Ly = n->layers[i];
    const long cnt = (long)Ly.co * Ly.ci_pad * 9;
    hipLaunchKernelGGL(k_pack_w, dim3(blocks(cnt)), dim3(256), 0, st, (const float*)n->params,
                       Ly.w_off, Ly.co, Ly.ci, Ly.ci_pad, Ly.wpk, Ly.wtk);
    AJ_LAUNCH_CHECK("k_pack_w");
  }
  return AJ_OK;
}

// ---------------------------------------------------------------------------
// warm start: fine (rhs, sdf) -> restricted, normalised -> network -> p0
// ---------------------------------------------------------------------------
__global__ void k_restrict2(const float* f, int n, int m, float* c) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  const int mc = m / 2;
  if (e >= (long)(n / 2) * mc) return;
  const int I = (int)(e / mc), J = (int)(e - (long)I * mc);
  const long q = (long)(2 * I) * m + 2 * J;
  c[e] = 0.25f * ((f[q] + f[q + m]) + (f[q + 1] + f[q + m + 1]));
}

// per-instance sum of squares: instance blockIdx.y, n values each
__global__ void k_sumsq(const float* f, long n, double* out) {
  __shared__ double sh[16];
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  f += (long)blockIdx.y * n;
  double v = 0.0;
  if (i < n) v = (double)f[i] * f[i];
  const double t = block_sum(v, sh);
  if (threadIdx.x == 0) atomicAdd(out + blockIdx.y, t);
}

__global__ void k_zero_d(double* p, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// x[p] = (rhs/s, tanh(sdf/sdf_ref)), s = sqrt(mean rhs^2) + 1e-30 per
// instance (ni values each, n in total)
__global__ void k_normalise(const float* rhs, const float* sdf, long n, long ni,
                            const double* sumsq, float sdf_ref, float* x, float* scale_out) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i == 0 && scale_out) *scale_out = (float)(sqrt(sumsq[0] / (double)ni) + 1e-30);
  if (i >= n) return;
  const float s = (float)(sqrt(sumsq[i / ni] / (double)ni) + 1e-30);
  x[2 * i] = rhs[i] / s;
  x[2 * i + 1] = tanhf(sdf[i] / sdf_ref);
}

__global__ void k_scale_target(const float* p, long n, const double* sumsq, float pscale,
                               float* t) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float s = (float)(sqrt(*sumsq / (double)n) + 1e-30);
  t[i] = p[i] / (s * pscale);
}

__global__ void k_unscale(const float* out, long n, long ni, const double* sumsq, float pscale,
                          float* p) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float s = (float)(sqrt(sumsq[i / ni] / (double)ni) + 1e-30);
  p[i] = out[i] * s * pscale;
}

// cell-centred bilinear prolongation of values (coarse level l+1 -> l); the
// fine outlet row of level 0 is the Dirichlet row (0); beyond a coarse
// outlet the ghost is gam * last (linear through the fine Dirichlet point)
__global__ void k_prolong_values(const float* e, int nc, int mc, float gamc, float* f, int nf,
                                 int mf, int fine_is_level0) {
  long q = (long)blockIdx.x * 256 + threadIdx.x;
  if (q >= (long)nf * mf) return;
  e += (long)blockIdx.y * nc * mc;  // instance blockIdx.y
  f += (long)blockIdx.y * nf * mf;
  const int i = (int)(q / mf), j = (int)(q - (long)i * mf);
  if (fine_is_level0 && i == nf - 1) { f[q] = 0.f; return; }
  const int I = min(i >> 1, nc - 1), J = min(j >> 1, mc - 1);
  int In = (i & 1) ? I + 1 : I - 1;
  int Jn = (j & 1) ? J + 1 : J - 1;
  float sx = 1.0f;
  if (In < 0) In = I;
  else if (In >= nc) { In = I; sx = gamc; }
  if (Jn < 0) Jn = J;
  else if (Jn >= mc) Jn = J;
  const float own = e[I * mc + J], xn = e[In * mc + J] * sx, yn = e[I * mc + Jn];
  const float dn = e[In * mc + Jn] * sx;
  f[q] = 0.5625f * own + 0.1875f * (xn + yn) + 0.0625f * dn;
}

double gamma_level(int l) { return 1.0 - 1.0 / (0.5 - std::ldexp(1.0, -l - 1)); }

// ---- warm start in two launches on each side of the network ---------------
// (the arithmetic of the separate restrict / sum-of-squares / normalise /
// gather and head / unscale / prolong passes, operation for operation)

// value of cell (I, J) of the grid restricted R times from f (row length m):
// R nested 2x2 averages, each 0.25 ((a + b) + (c + d)) as k_restrict2
template <int R>
__device__ __forceinline__ float nested_avg(const float* f, int m, int I, int J) {
  if constexpr (R == 0) {
    return f[(long)I * m + J];
  } else {
    const float a = nested_avg<R - 1>(f, m, 2 * I, 2 * J);
    const float b = nested_avg<R - 1>(f, m, 2 * I + 1, 2 * J);
    const float c = nested_avg<R - 1>(f, m, 2 * I, 2 * J + 1);
    const float d = nested_avg<R - 1>(f, m, 2 * I + 1, 2 * J + 1);
    return 0.25f * ((a + b) + (c + d));
  }
}

// a / d for 0 <= a < 2^53 from the FP64 product with rd = 1/d and one
// correction step (a 64-bit integer division is a long instruction sequence
// on this target)
__device__ __forceinline__ long ws_quot(long a, long d, double rd) {
  long q = (long)((double)a * rd);
  const long r = a - q * d;
  if (r < 0) --q;
  else if (r >= d) ++q;
  return q;
}

// per coarse cell of instance blockIdx.y: restricted rhs R, tanh(sdf_r /
// sdf_ref) T, and the block's sum of R^2 (fixed-order partials)
template <int RR>
__global__ __launch_bounds__(256) void k_ws_prep(const float* __restrict__ rhs,
                                                 const float* __restrict__ sdf, int nx, int ny,
                                                 float sdf_ref, float* __restrict__ Rout,
                                                 float* __restrict__ Tout,
                                                 double* __restrict__ part) {
  __shared__ double sh[16];
  const int H = nx >> RR, W = ny >> RR;
  const long ni = (long)H * W;
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  const long b = blockIdx.y;
  double v = 0.0;
  if (e < ni) {
    const int I = (int)ws_quot(e, W, 1.0 / W), J = (int)(e - (long)I * W);
    const float r = nested_avg<RR>(rhs + b * nx * ny, ny, I, J);
    const float sd = nested_avg<RR>(sdf + b * nx * ny, ny, I, J);
    Rout[b * ni + e] = r;
    Tout[b * ni + e] = tanhf(sd / sdf_ref);
    v = (double)r * r;
  }
  const double t = block_sum(v, sh);
  if (threadIdx.x == 0) part[b * gridDim.x + blockIdx.x] = t;
}

// s = sqrt(mean R^2) + 1e-30 of the instance (every block sums the
// instance's partials in the same order), network input row (R / s, T, 0..)
__global__ __launch_bounds__(256) void k_ws_input(const float* __restrict__ R,
                                                  const float* __restrict__ T,
                                                  const double* __restrict__ part, int nblk,
                                                  long ni, bf16_t* __restrict__ X0,
                                                  float* __restrict__ s_out) {
  __shared__ double sh[16];
  __shared__ float s_sh;
  const long b = blockIdx.y;
  double acc = 0.0;
  for (int k = threadIdx.x; k < nblk; k += 256) acc += part[b * nblk + k];
  const double tot = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    s_sh = (float)(sqrt(tot / (double)ni) + 1e-30);
    if (blockIdx.x == 0) s_out[b] = s_sh;
  }
  __syncthreads();
  const float s = s_sh;
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= ni) return;
  const long q = b * ni + e;
  bf16_t* o = X0 + q * kCinPad;
  o[0] = f2bf(R[q] / s);
  o[1] = f2bf(T[q]);
  for (int c = 2; c < kCinPad; ++c) o[c] = 0;
}

// head (1x1 conv C -> 1 + bias, as k_head) times s pscale (as k_unscale);
// vec: rows of x 16-byte aligned and C a multiple of 8, read 8 channels per
// load (the same channel order)
__global__ __launch_bounds__(256) void k_ws_head(const bf16_t* x, int ldx, int C, const float* w,
                                                 const float* bias, const float* __restrict__ s,
                                                 long ni, float pscale, float* out, long P,
                                                 int vec) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= P) return;
  float v = bias[0];
  if (vec) {
    const u32x4* xv = reinterpret_cast<const u32x4*>(x + p * ldx);
    for (int c8 = 0; c8 < C / 8; ++c8) {
      const u32x4 q = xv[c8];
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const unsigned u = q[h];
        v = head_acc(v, bf2f((bf16_t)(u & 0xffffu)), w[8 * c8 + 2 * h]);
        v = head_acc(v, bf2f((bf16_t)(u >> 16)), w[8 * c8 + 2 * h + 1]);
      }
    }
  } else {
    for (int c = 0; c < C; ++c) v = head_acc(v, bf2f(x[p * ldx + c]), w[c]);
  }
  out[p] = v * s[ws_quot(p, ni, 1.0 / (double)ni)] * pscale;
}

// value at (i, j) of the grid K prolongations finer than c (hc x wc), the
// arithmetic of K k_prolong_values passes; the finest of the r levels is the
// fine grid (outlet row 0); gam[l] = outlet ghost factor of level l
template <int K>
__device__ __forceinline__ float up_val(const float* c, int hc, int wc, int r, const float* gam,
                                        int i, int j) {
  if constexpr (K == 0) {
    return c[(long)i * wc + j];
  } else {
    const int nc = hc << (K - 1), mc = wc << (K - 1), nf = 2 * nc;
    const int l = r - K;  // level of the value computed here
    if (l == 0 && i == nf - 1) return 0.f;
    const int I = min(i >> 1, nc - 1), J = min(j >> 1, mc - 1);
    int In = (i & 1) ? I + 1 : I - 1;
    int Jn = (j & 1) ? J + 1 : J - 1;
    float sx = 1.0f;
    if (In < 0) In = I;
    else if (In >= nc) { In = I; sx = gam[l + 1]; }
    if (Jn < 0) Jn = J;
    else if (Jn >= mc) Jn = J;
    const float own = up_val<K - 1>(c, hc, wc, r, gam, I, J);
    const float xn = up_val<K - 1>(c, hc, wc, r, gam, In, J) * sx;
    const float yn = up_val<K - 1>(c, hc, wc, r, gam, I, Jn);
    const float dn = up_val<K - 1>(c, hc, wc, r, gam, In, Jn) * sx;
    return 0.5625f * own + 0.1875f * (xn + yn) + 0.0625f * dn;
  }
}

struct Gam4 {
  float g[4];
};

template <int RR>
__global__ __launch_bounds__(256) void k_ws_prolong(const float* __restrict__ c, int H, int W,
                                                    Gam4 gam, float* __restrict__ p0, int nx,
                                                    int ny) {
  const long q = (long)blockIdx.x * 256 + threadIdx.x;
  if (q >= (long)nx * ny) return;
  const long b = blockIdx.y;
  const int i = (int)ws_quot(q, ny, 1.0 / ny), j = (int)(q - (long)i * ny);
  p0[b * nx * ny + q] = up_val<RR>(c + b * H * W, H, W, RR, gam.g, i, j);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int aj_nn_create(aj_nn** out, int base_channels, int levels, int seed) {
  if (!out) return fail(AJ_ERR_INVALID, "aj_nn_create: null");
  *out = nullptr;
  if (base_channels < 16 || base_channels % 32 != 0 || levels < 2 || levels > 6)
    return fail(AJ_ERR_INVALID, "aj_nn_create: base_channels must be a multiple of 32, levels 2..6");
  aj_nn* n = new aj_nn();
  n->c0 = base_channels;
  n->L = levels;
  (void)hipGetDevice(&n->device);
  long off = 0;
  const int L = levels;
  n->layers.resize(4 * L - 2);
  auto def = [&](int idx, int ci, int co) {
    Layer& Ly = n->layers[idx];
    Ly.ci = ci;
    Ly.ci_pad = ((ci + 31) / 32) * 32;
    Ly.co = co;
    Ly.w_off = off;
    off += (long)co * ci * 9;
    Ly.b_off = off;
    off += co;
  };
  for (int l = 0; l < L; ++l) {
    def(enc_a(l), l == 0 ? n->cin : chan(n, l - 1), chan(n, l));
    def(enc_b(l), chan(n, l), chan(n, l));
  }
  for (int l = L - 2; l >= 0; --l) {
    def(dec_a(n, l), chan(n, l + 1) + chan(n, l), chan(n, l));
    def(dec_b(n, l), chan(n, l), chan(n, l));
  }
  n->head_w = off;
  off += n->c0;
  n->head_b = off;
  off += 1;
  n->nparams = off;
  // host init: uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) (torch Conv2d default bound)
  std::vector<float> hp(off, 0.0f);
  unsigned long long s = 0x5eed0000ULL + (unsigned long long)seed;
  auto uni = [&](float bound) {
    const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
    return (float)((2.0 * u - 1.0) * bound);
  };
  for (auto& Ly : n->layers) {
    const float bound = 1.0f / std::sqrt((float)(Ly.ci * 9));
    for (long i = 0; i < (long)Ly.co * Ly.ci * 9; ++i) hp[Ly.w_off + i] = uni(bound);
    for (int i = 0; i < Ly.co; ++i) hp[Ly.b_off + i] = uni(bound);
  }
  {
    const float bound = 1.0f / std::sqrt((float)n->c0);
    for (int i = 0; i < n->c0; ++i) hp[n->head_w + i] = uni(bound);
    hp[n->head_b] = 0.0f;
  }
  // device memory: params, grads, m, v + packed weights
  size_t wbytes = 0;
  for (auto& Ly : n->layers) {
    wbytes += (((size_t)Ly.co * Ly.ci_pad * 9 * 2) + 255) & ~size_t(255);
    wbytes += (((size_t)Ly.co * Ly.ci * 9 * 2) + 255) & ~size_t(255);
  }
  const size_t pbytes = (((size_t)off * 4) + 255) & ~size_t(255);
  if (hipMalloc(&n->wmem, 4 * pbytes + wbytes) != hipSuccess) {
    delete n;
    return fail(AJ_ERR_HIP, "aj_nn_create: hipMalloc failed");
  }
  char* b = (char*)n->wmem;
  n->params = (float*)b; b += pbytes;
  n->grads = (float*)b; b += pbytes;
  n->m = (float*)b; b += pbytes;
  n->v = (float*)b; b += pbytes;
  for (auto& Ly : n->layers) {
    Ly.wpk = (bf16_t*)b;
    b += (((size_t)Ly.co * Ly.ci_pad * 9 * 2) + 255) & ~size_t(255);
    Ly.wtk = (bf16_t*)b;
    b += (((size_t)Ly.co * Ly.ci * 9 * 2) + 255) & ~size_t(255);
  }
  if (hipMemcpy(n->params, hp.data(), off * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(n->m, 0, pbytes * 3) != hipSuccess) {
    (void)hipFree(n->wmem);
    delete n;
    return fail(AJ_ERR_HIP, "aj_nn_create: upload failed");
  }
  if (int rc = pack_all(n, nullptr)) {
    (void)hipFree(n->wmem);
    delete n;
    return rc;
  }
  if (hipDeviceSynchronize() != hipSuccess) {
    (void)hipFree(n->wmem);
    delete n;
    return fail(AJ_ERR_HIP, "aj_nn_create: pack failed");
  }
  *out = n;
  return AJ_OK;
}

extern "C" int aj_nn_destroy(aj_nn* n) {
  if (!n) return AJ_OK;
  if (n->act.mem) (void)hipFree(n->act.mem);
  for (auto& g : n->ws_graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
  }
  if (n->ws_mem) (void)hipFree(n->ws_mem);
  if (n->wmem) (void)hipFree(n->wmem);
  delete n;
  return AJ_OK;
}

extern "C" long aj_nn_num_params(const aj_nn* n) { return n ? n->nparams : 0; }

extern "C" int aj_nn_set_option(aj_nn* n, int option, int value) {
  if (!n) return fail(AJ_ERR_INVALID, "aj_nn_set_option: null network");
  switch (option) {
    case AJ_NN_OPT_FUSED_IO:
      n->fuse_io = value ? 1 : 0;
      return AJ_OK;
    default:
      return fail(AJ_ERR_INVALID, "aj_nn_set_option: unknown option");
  }
}

extern "C" int aj_nn_get_params(aj_nn* n, float* host, long count, int which) {
  if (!n || !host || count != n->nparams) return fail(AJ_ERR_INVALID, "aj_nn_get_params: bad args");
  const float* src = which == 1 ? n->grads : n->params;
  AJ_HIP(hipDeviceSynchronize());
  AJ_HIP(hipMemcpy(host, src, count * 4, hipMemcpyDeviceToHost));
  return AJ_OK;
}

extern "C" int aj_nn_set_params(aj_nn* n, const float* host, long count) {
  if (!n || !host || count != n->nparams) return fail(AJ_ERR_INVALID, "aj_nn_set_params: bad args");
  AJ_HIP(hipDeviceSynchronize());
  AJ_HIP(hipMemcpy(n->params, host, count * 4, hipMemcpyHostToDevice));
  AJ_HIP(hipMemset(n->m, 0, (size_t)count * 4));
  AJ_HIP(hipMemset(n->v, 0, (size_t)count * 4));
  n->step = 0;
  if (int rc = pack_all(n, nullptr)) return rc;
  AJ_HIP(hipDeviceSynchronize());
  return AJ_OK;
}

extern "C" int aj_nn_forward(aj_nn* n, const float* x, const int* idx, float* out, int B, int H,
                             int W, void* stream) {
  if (!n || !x || !out || B <= 0) return fail(AJ_ERR_INVALID, "aj_nn_forward: bad args");
  hipStream_t st = (hipStream_t)stream;
  if (int rc = ensure_act(n, B, H, W)) return rc;
  const long HW = (long)H * W, P = (long)B * HW;
  hipLaunchKernelGGL(k_gather, dim3(blocks(P)), dim3(256), 0, st, x, (const float*)nullptr, idx, B,
                     HW, n->act.X0, (float*)nullptr);
  AJ_LAUNCH_CHECK("k_gather");
  if (int rc = forward(n, B, H, W, st)) return rc;
  AJ_HIP(hipMemcpyAsync(out, n->act.out, P * 4, hipMemcpyDeviceToDevice, st));
  return AJ_OK;
}

// forward + MSE + backward (gradients in the network's buffer); loss (device
// float, may be null) receives the batch loss
extern "C" int aj_nn_loss_grad(aj_nn* n, const float* x, const float* t, const int* idx, int B,
                               int H, int W, float* loss, void* stream) {
  if (!n || !x || !t || B <= 0) return fail(AJ_ERR_INVALID, "aj_nn_loss_grad: bad args");
  hipStream_t st = (hipStream_t)stream;
  if (int rc = ensure_act(n, B, H, W)) return rc;
  Act& a = n->act;
  const long HW = (long)H * W, P = (long)B * HW;
  hipLaunchKernelGGL(k_gather, dim3(blocks(P)), dim3(256), 0, st, x, t, idx, B, HW, a.X0, a.target);
  AJ_LAUNCH_CHECK("k_gather");
  if (int rc = forward(n, B, H, W, st)) return rc;
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(256), 0, st, a.loss, 1L);
  hipLaunchKernelGGL(k_mse, dim3(blocks(P)), dim3(256), 0, st, (const float*)a.out,
                     (const float*)a.target, a.dout, P, a.loss);
  AJ_LAUNCH_CHECK("k_mse");
  if (int rc = backward(n, B, H, W, st)) return rc;
  if (loss) AJ_HIP(hipMemcpyAsync(loss, a.loss, 4, hipMemcpyDeviceToDevice, st));
  return AJ_OK;
}

extern "C" int aj_nn_adam(aj_nn* n, float lr, float clip, void* stream) {
  if (!n) return fail(AJ_ERR_INVALID, "aj_nn_adam: null");
  hipStream_t st = (hipStream_t)stream;
  n->step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float bc1 = 1.0f - std::pow(b1, (float)n->step);
  const float bc2 = 1.0f - std::pow(b2, (float)n->step);
  hipLaunchKernelGGL(k_adam, dim3(blocks(n->nparams)), dim3(256), 0, st, n->params,
                     (const float*)n->grads, n->m, n->v, n->nparams, lr, b1, b2, 1e-8f, bc1, bc2,
                     clip);
  AJ_LAUNCH_CHECK("k_adam");
  return pack_all(n, st);
}

extern "C" int aj_nn_train_step(aj_nn* n, const float* x, const float* t, const int* idx, int B,
                                int H, int W, float lr, float* loss, void* stream) {
  if (int rc = aj_nn_loss_grad(n, x, t, idx, B, H, W, loss, stream)) return rc;
  return aj_nn_adam(n, lr, 1.0f, stream);
}

namespace {

// scratch of the sample/warm-start pipeline: 8 fine-size float slots and
// the per-instance sums of squares
int ensure_ws(aj_nn* n, long nf, int B = 1) {
  const size_t need = (size_t)(8 * nf + 64) * 4 + (size_t)B * 8;
  if (n->ws_bytes >= need) return AJ_OK;
  if (n->ws_mem) (void)hipFree(n->ws_mem);
  n->ws_mem = nullptr;
  AJ_HIP(hipMalloc(&n->ws_mem, need));
  n->ws_bytes = need;
  return AJ_OK;
}

// restrict `src` (nx x ny) r times by 2x2 averaging into `dst` using `tmp`
// (both >= nx*ny/4 floats, r >= 1) or copy (r == 0)
int restrict_r(const float* src, int nx, int ny, int r, float* dst, float* tmp, hipStream_t st) {
  if (r == 0) {
    AJ_HIP(hipMemcpyAsync(dst, src, (size_t)nx * ny * 4, hipMemcpyDeviceToDevice, st));
    return AJ_OK;
  }
  const float* cur = src;
  int cn = nx, cm = ny;
  for (int s = 0; s < r; ++s) {
    float* out = ((r - 1 - s) % 2 == 0) ? dst : tmp;
    hipLaunchKernelGGL(k_restrict2, dim3(blocks((long)(cn / 2) * (cm / 2))), dim3(256), 0, st, cur,
                       cn, cm, out);
    cur = out;
    cn /= 2;
    cm /= 2;
  }
  AJ_LAUNCH_CHECK("k_restrict2");
  return AJ_OK;
}

}  // namespace

// one training pair from solver fields: restrict (rhs, sdf, p) `r` times and
// normalise: x = (rhs/s, tanh(sdf/sdf_ref)) interleaved, t = p/(s*pscale),
// s = RMS of the restricted rhs
extern "C" int aj_nn_make_sample(aj_nn* n, const float* rhs, const float* sdf, const float* p,
                                 int nx, int ny, int r, float sdf_ref, float pscale, float* x_out,
                                 float* t_out, void* stream) {
  if (!n || !rhs || !sdf || !x_out || r < 0) return fail(AJ_ERR_INVALID, "aj_nn_make_sample: bad args");
  if ((nx >> r) << r != nx || (ny >> r) << r != ny)
    return fail(AJ_ERR_INVALID, "aj_nn_make_sample: grid not divisible by 2^r");
  hipStream_t st = (hipStream_t)stream;
  const long nf = (long)nx * ny, nc = nf >> (2 * r);
  if (int rc = ensure_ws(n, nf)) return rc;
  float* w = (float*)n->ws_mem;
  float *R = w, *S = w + nf, *Pc = w + 2 * nf, *tmp = w + 3 * nf;
  double* sumsq = (double*)(w + 8 * nf);
  if (int rc = restrict_r(rhs, nx, ny, r, R, tmp, st)) return rc;
  if (int rc = restrict_r(sdf, nx, ny, r, S, tmp, st)) return rc;
  if (p && t_out)
    if (int rc = restrict_r(p, nx, ny, r, Pc, tmp, st)) return rc;
  hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(256), 0, st, sumsq, 1);
  hipLaunchKernelGGL(k_sumsq, dim3(blocks(nc)), dim3(256), 0, st, (const float*)R, nc, sumsq);
  hipLaunchKernelGGL(k_normalise, dim3(blocks(nc)), dim3(256), 0, st, (const float*)R,
                     (const float*)S, nc, nc, (const double*)sumsq, sdf_ref, x_out,
                     (float*)nullptr);
  if (p && t_out)
    hipLaunchKernelGGL(k_scale_target, dim3(blocks(nc)), dim3(256), 0, st, (const float*)Pc, nc,
                       (const double*)sumsq, pscale, t_out);
  AJ_LAUNCH_CHECK("k_normalise");
  return AJ_OK;
}

// warm start: p0 (fine, nx x ny) = P^r( s*pscale * net(rhs_r/s, tanh(sdf_r/sdf_ref)) )
// with the multigrid's cell-centred bilinear prolongation (outlet row 0);
// B instances stored one after the other, each normalised by its own s.
// Stacked instances restrict as one (B nx) x ny array: nx is divisible by
// 2^r, so no 2x2 average straddles two instances.
static int warm_start_launches(aj_nn* n, const float* rhs, const float* sdf, float* p0, int B,
                               int nx, int ny, int r, float sdf_ref, float pscale,
                               hipStream_t st) {
  const int H = nx >> r, W = ny >> r;
  const long ni = (long)H * W, nf = (long)B * nx * ny, nc = (long)B * ni;
  if (int rc = ensure_ws(n, nf, B)) return rc;
  float* w = (float*)n->ws_mem;
  float *R = w, *S = w + nf, *X = w + 2 * nf, *A = w + 4 * nf, *Bf = w + 5 * nf, *tmp = w + 6 * nf;
  double* sumsq = (double*)(w + 8 * nf + 64);
  if (r >= 1 && r <= 3) {
    // two launches before the network, two after
    if (int rc = ensure_act(n, B, H, W)) return rc;
    const int nblk = (int)blocks(ni);
    double* part = (double*)(w + 6 * nf);  // B * nblk doubles (<= nf / 128 floats)
    float* sv = (float*)(w + 7 * nf);      // per-instance s
    const dim3 g2(nblk, B);
    switch (r) {
      case 1: hipLaunchKernelGGL(k_ws_prep<1>, g2, dim3(256), 0, st, rhs, sdf, nx, ny, sdf_ref, R, S, part); break;
      case 2: hipLaunchKernelGGL(k_ws_prep<2>, g2, dim3(256), 0, st, rhs, sdf, nx, ny, sdf_ref, R, S, part); break;
      default: hipLaunchKernelGGL(k_ws_prep<3>, g2, dim3(256), 0, st, rhs, sdf, nx, ny, sdf_ref, R, S, part); break;
    }
    AJ_LAUNCH_CHECK("k_ws_prep");
    const float* hw = n->params + n->head_w;
    const float* hb = n->params + n->head_b;
    if (n->fuse_io && (n->c0 == 32 || n->c0 == 64)) {
      // the input row in the first convolution's loads, the head in the
      // last one's epilogue (the arithmetic of k_ws_input and k_ws_head)
      ConvSrc in0{};
      in0.mode = 3;
      in0.inR = R;
      in0.inT = S;
      in0.part = part;
      in0.nblk = nblk;
      in0.s_out = sv;
      const HeadOut ho{hw, hb, sv, pscale, A};
      if (int rc = forward(n, B, H, W, st, false, true, &in0, &ho)) return rc;
    } else {
      hipLaunchKernelGGL(k_ws_input, g2, dim3(256), 0, st, (const float*)R, (const float*)S,
                         (const double*)part, nblk, ni, n->act.X0, sv);
      AJ_LAUNCH_CHECK("k_ws_input");
      if (int rc = forward(n, B, H, W, st, false, true)) return rc;
      hipLaunchKernelGGL(k_ws_head, dim3(blocks(nc)), dim3(256), 0, st, (const bf16_t*)n->act.DB[0],
                         n->c0, n->c0, hw, hb, (const float*)sv, ni, pscale, A, nc,
                         (((uintptr_t)n->act.DB[0] % 16) == 0 && n->c0 % 8 == 0) ? 1 : 0);
      AJ_LAUNCH_CHECK("k_ws_head");
    }
    Gam4 gam{};
    for (int l = 0; l <= r; ++l) gam.g[l] = (float)gamma_level(l);
    const dim3 gp(blocks((long)nx * ny), B);
    switch (r) {
      case 1: hipLaunchKernelGGL(k_ws_prolong<1>, gp, dim3(256), 0, st, (const float*)A, H, W, gam, p0, nx, ny); break;
      case 2: hipLaunchKernelGGL(k_ws_prolong<2>, gp, dim3(256), 0, st, (const float*)A, H, W, gam, p0, nx, ny); break;
      default: hipLaunchKernelGGL(k_ws_prolong<3>, gp, dim3(256), 0, st, (const float*)A, H, W, gam, p0, nx, ny); break;
    }
    AJ_LAUNCH_CHECK("k_ws_prolong");
    return AJ_OK;
  }
  if (int rc = restrict_r(rhs, B * nx, ny, r, R, tmp, st)) return rc;
  if (int rc = restrict_r(sdf, B * nx, ny, r, S, tmp, st)) return rc;
  hipLaunchKernelGGL(k_zero_d, dim3(blocks(B)), dim3(256), 0, st, sumsq, B);
  hipLaunchKernelGGL(k_sumsq, dim3(blocks(ni), B), dim3(256), 0, st, (const float*)R, ni, sumsq);
  hipLaunchKernelGGL(k_normalise, dim3(blocks(nc)), dim3(256), 0, st, (const float*)R,
                     (const float*)S, nc, ni, (const double*)sumsq, sdf_ref, X, (float*)nullptr);
  AJ_LAUNCH_CHECK("k_normalise");
  if (int rc = ensure_act(n, B, H, W)) return rc;
  hipLaunchKernelGGL(k_gather, dim3(blocks(nc)), dim3(256), 0, st, (const float*)X,
                     (const float*)nullptr, (const int*)nullptr, B, ni, n->act.X0,
                     (float*)nullptr);
  AJ_LAUNCH_CHECK("k_gather");
  if (int rc = forward(n, B, H, W, st)) return rc;
  float* cur = (r == 0) ? p0 : A;
  hipLaunchKernelGGL(k_unscale, dim3(blocks(nc)), dim3(256), 0, st, (const float*)n->act.out, nc,
                     ni, (const double*)sumsq, pscale, cur);
  AJ_LAUNCH_CHECK("k_unscale");
  int cn = H, cm = W;
  for (int l = r - 1; l >= 0; --l) {
    const int fn = cn * 2, fm = cm * 2;
    float* dst = (l == 0) ? p0 : (cur == A ? Bf : A);
    hipLaunchKernelGGL(k_prolong_values, dim3(blocks((long)fn * fm), B), dim3(256), 0, st,
                       (const float*)cur, cn, cm, (float)gamma_level(l + 1), dst, fn, fm,
                       l == 0 ? 1 : 0);
    AJ_LAUNCH_CHECK("k_prolong_values");
    cur = dst;
    cn = fn;
    cm = fm;
  }
  return AJ_OK;
}

extern "C" int aj_nn_warm_start_batched(aj_nn* n, const float* rhs, const float* sdf, float* p0,
                                        int B, int nx, int ny, int r, float sdf_ref, float pscale,
                                        void* stream) {
  if (!n || !rhs || !sdf || !p0 || r < 0 || B < 1 || B > 65535)
    return fail(AJ_ERR_INVALID, "aj_nn_warm_start: bad args");
  if ((nx >> r) << r != nx || (ny >> r) << r != ny)
    return fail(AJ_ERR_INVALID, "aj_nn_warm_start: grid not divisible by 2^r");
  hipStream_t st = (hipStream_t)stream;
  // the buffers a replay uses are fixed once the scratch and activations
  // exist for this shape; allocate them first (never during a capture)
  const int H = nx >> r, W = ny >> r;
  if (int rc = ensure_ws(n, (long)B * nx * ny, B)) return rc;
  if (int rc = ensure_act(n, B, H, W)) return rc;
  auto bits = [](float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (uint64_t)u;
  };
  const std::vector<uint64_t> key = {(uint64_t)(uintptr_t)rhs, (uint64_t)(uintptr_t)sdf,
                                     (uint64_t)(uintptr_t)p0, (uint64_t)B, (uint64_t)nx,
                                     (uint64_t)ny, (uint64_t)r, bits(sdf_ref), bits(pscale),
                                     (uint64_t)(uintptr_t)n->ws_mem,
                                     (uint64_t)(uintptr_t)n->act.mem, (uint64_t)n->fuse_io};
  if (prof_on() || !(r >= 1 && r <= 3))
    return warm_start_launches(n, rhs, sdf, p0, B, nx, ny, r, sdf_ref, pscale, st);
  for (auto& g : n->ws_graphs)
    if (g.key == key) {
      AJ_HIP(hipGraphLaunch(g.exec, st));
      return AJ_OK;
    }
  if (std::find(n->ws_seen.begin(), n->ws_seen.end(), key) == n->ws_seen.end()) {
    // first call with this signature: eager (tunes the launch shapes)
    if (n->ws_seen.size() >= 16) n->ws_seen.erase(n->ws_seen.begin());
    n->ws_seen.push_back(key);
    return warm_start_launches(n, rhs, sdf, p0, B, nx, ny, r, sdf_ref, pscale, st);
  }
  if (n->ws_graphs.size() >= 8) {  // keep the eight most recent signatures
    auto& old = n->ws_graphs.front();
    if (old.exec) (void)hipGraphExecDestroy(old.exec);
    if (old.graph) (void)hipGraphDestroy(old.graph);
    n->ws_graphs.erase(n->ws_graphs.begin());
  }
  aj_nn::WsGraph wg;
  wg.key = key;
  hipStream_t cs;
  AJ_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  int rc = AJ_OK;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    rc = warm_start_launches(n, rhs, sdf, p0, B, nx, ny, r, sdf_ref, pscale, cs);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    if (e == hipSuccess && rc == AJ_OK) {
      e = hipGraphInstantiate(&wg.exec, g, nullptr, nullptr, 0);
      wg.graph = g;
    } else if (g) {
      (void)hipGraphDestroy(g);
    }
  }
  (void)hipStreamDestroy(cs);
  if (rc != AJ_OK || e != hipSuccess) {
    if (wg.exec) (void)hipGraphExecDestroy(wg.exec);
    if (wg.graph) (void)hipGraphDestroy(wg.graph);
    if (rc != AJ_OK) return rc;
    return hip_fail(e, "capture of the warm start");
  }
  n->ws_graphs.push_back(wg);
  AJ_HIP(hipGraphLaunch(wg.exec, st));
  return AJ_OK;
}

extern "C" int aj_nn_warm_start(aj_nn* n, const float* rhs, const float* sdf, float* p0, int nx,
                                int ny, int r, float sdf_ref, float pscale, void* stream) {
  return aj_nn_warm_start_batched(n, rhs, sdf, p0, 1, nx, ny, r, sdf_ref, pscale, stream);
}
