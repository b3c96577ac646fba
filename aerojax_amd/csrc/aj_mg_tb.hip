// Textbook geometric multigrid for the consistent outlet-Dirichlet operator,
// FP32 V(2,2)-cycles inside an FP64 defect correction.
//
// Discrete problem (rows 0..nx-2 carry equations, p[nx-1, :] = 0):
//   ax (p[i+1] + p[i-1] - 2p) + ay (p[j+1] + p[j-1] - 2p) = b
// with edge (Neumann) ghosts at the inlet row and both walls.
//
// Hierarchy: cell-centred 2x coarsening down to <= 65536 cells (or exactly
// the requested number of levels); coarse level
// l has coefficients ax/4^l, ay/4^l and an outlet ghost gamma_l * last that
// puts its Dirichlet point where the fine grid has it
// (gamma_l = 1 - 1/(1/2 - 2^-(l+1))).  Smoother: red-black Gauss-Seidel
// with the exact row solve, over-relaxed by omega (default 1.2: red-black
// SOR smoothing, about 5x the residual reduction per outer iteration of
// omega = 1 here).  Restriction: 2x2 average.  Prolongation:
// cell-centred bilinear.  Coarsest level: exact FP64 solve (cosine
// transform along y, one tridiagonal solve per mode along x, three kernels).
//
// Kernels per V-cycle on every level above the coarsest (temporal blocking
// in LDS):
//   pre  : 2 RB-GS sweeps (4 half-sweeps, halo 5) + residual + restriction,
//          writes the smoothed iterate Y and the coarse rhs
//   post : Y + prolong(E_coarse), 2 RB-GS sweeps (halo 4), writes E
// Defect correction per outer iteration:
//   resid: p64' = p64 + e (on the fly), r = b - A p64' in FP64, writes
//          r32 and p64', FP64 partial sums of r^2
//   final: norm = sqrt(sum r^2 dx dy), convergence gate, buffer flip
//
// A plan may hold `batch` independent instances of the same grid, stored one
// after the other (instance z at z * cells of every buffer): every kernel
// takes the instance from blockIdx.z (the per-instance reductions from
// blockIdx.x), and each instance has its own convergence gate, guess test,
// tolerance and live FP64 buffer.
#include "aj_mg.h"

#include <cmath>
#include <cstdio>
#include <vector>

namespace aj {

constexpr int kTbCoarseCells = 65536;  // coarsening stops at or below this many cells
constexpr int kMaxLevels = 24;

static double outlet_gamma(int level) {
  return 1.0 - 1.0 / (0.5 - std::ldexp(1.0, -level - 1));
}

struct TbLevel {
  int n, m;        // rows (x), cols (y)
  int level;       // 0 = fine
  float axl, ayl;  // coefficients
  float gam;       // outlet ghost factor (level >= 1)
  float dE_last;   // (1 - gamma) * axl
  float inv_int;   // 1 / (2 axl + 2 ayl): inverse diagonal of interior rows
  float omega;     // over-relaxation of the red-black sweeps (1 = Gauss-Seidel)
};

// The per-cell arithmetic of the smoothers, written out operation by
// operation (explicit fused multiply-adds): every smoother kernel (LDS
// tiles and register columns) evaluates a cell with these functions, so
// any of them gives the same bits.
//
// Weight of a red-black update: omega / diagonal of the cell's row of the
// level operator (omega * inv_int for interior cells).  Absent neighbours
// (outside the domain) leave the diagonal; a coarse level's outlet ghost
// enters it as dE_last.  The diagonal sums dE, W, S, N in that order.
__device__ __forceinline__ float tb_weight(const TbLevel& L, bool hasW, bool last, bool hasS,
                                           bool hasN) {
#pragma clang fp contract(off)
  const bool inr = hasW && hasS && hasN && !last;
  const float dE = (L.level > 0 && last) ? L.dE_last : L.axl;
  const float dg = ((dE + (hasW ? L.axl : 0.0f)) + (hasS ? L.ayl : 0.0f)) + (hasN ? L.ayl : 0.0f);
  return L.omega * (inr ? L.inv_int : 1.0f / dg);
}

// Over-relaxed exact row solve p + omega (p_GS - p) as p (1 - omega) + t wr,
// t = ax (E + W) + ay (N + S) - b; xs = E + W and ys = N + S with absent
// neighbours 0, omw = 1 - omega.
__device__ __forceinline__ float tb_update(const TbLevel& L, float xs, float ys, float bv,
                                           float po, float wr, float omw) {
#pragma clang fp contract(off)
  const float t = fmaf(L.axl, xs, fmaf(L.ayl, ys, -bv));
  return fmaf(t, wr, po * omw);
}

// r = b - A p from the cell and its four neighbours (an absent neighbour is
// the cell itself, the coarse outlet ghost gamma * p)
__device__ __forceinline__ float tb_res(const TbLevel& L, float pc, float bv, float W, float E,
                                        float S, float N) {
#pragma clang fp contract(off)
  const float Ap = fmaf(L.axl, (E - pc) + (W - pc), L.ayl * ((N - pc) + (S - pc)));
  return bv - Ap;
}

// r = b - A p at (i, j) (level operator; level 0 outlet row -> 0)
template <typename F>
__device__ __forceinline__ float tb_resid(const TbLevel& L, int i, int j, float pc, float bv,
                                          F&& nb) {
  if (L.level == 0 && i == L.n - 1) return 0.0f;
  const float W = i > 0 ? nb(-1, 0) : pc;
  float E;
  if (i < L.n - 1) E = nb(1, 0);
  else E = L.gam * pc;  // level >= 1 only (level 0 returned above)
  const float S = j > 0 ? nb(0, -1) : pc;
  const float N = j < L.m - 1 ? nb(0, 1) : pc;
  return tb_res(L, pc, bv, W, E, S, N);
}

// bilinear weights of the cell-centred prolongation: the own coarse cell,
// its x and y neighbours towards the fine cell, the diagonal one
__device__ __forceinline__ float tb_prolong_val(float own, float xn, float yn, float dn) {
#pragma clang fp contract(off)
  return fmaf(0.0625f, dn, fmaf(0.1875f, xn + yn, 0.5625f * own));
}

// cell-centred bilinear prolongation of coarse e (nc x mc) at fine (i, j);
// gamc = outlet ghost factor of the coarse level
__device__ __forceinline__ float tb_prolong(const float* __restrict__ e, int nc, int mc,
                                            float gamc, int i, int j) {
  const int I = i >> 1, J = j >> 1;
  int In = (i & 1) ? I + 1 : I - 1;
  int Jn = (j & 1) ? J + 1 : J - 1;
  float sx = 1.0f;
  if (In < 0) In = I;
  else if (In >= nc) { In = I; sx = gamc; }
  if (Jn < 0) Jn = J;
  else if (Jn >= mc) Jn = J;
  const float own = e[I * mc + J];
  const float xn = e[In * mc + J] * sx;
  const float yn = e[I * mc + Jn];
  const float dn = e[In * mc + Jn] * sx;
  return tb_prolong_val(own, xn, yn, dn);
}

// LDS layout of a smoother tile (RH x RW cells, RW even): the two colours
// of the checkerboard live in separate planes, cell (lx, ly) of colour
// c = (bp + lx + ly) & 1 at plane c, row lx, slot ly >> 1, where bp is the
// parity of the tile origin.  A half-sweep then reads its four neighbours
// (all in the other plane: slot ly >> 1 of rows lx -+ 1, slots
// (ly -+ 1) >> 1 of row lx) and b and writes its cell at unit stride
// across the lanes.
template <int RH, int RW>
struct TbTile {
  static_assert(RW % 2 == 0, "tile width");
  static constexpr int RW2 = RW / 2;
  static constexpr int PS = ((RH * RW2 + 31) / 32) * 32 + 16;  // plane stride
  static constexpr int N = 2 * PS;
  __device__ static __forceinline__ int at(int bp, int lx, int ly) {
    return ((bp + lx + ly) & 1) * PS + lx * RW2 + (ly >> 1);
  }
};

// Thread layout of a pass over a tile region (RH x RW cells): each thread
// keeps one column and steps over the rows RG at a time, so the row part of
// an address is uniform per step and the column part is fixed (threads
// past RG * RW idle).
template <int RH, int RW, int NTH>
struct TbPass {
  static constexpr int RG = NTH / RW;              // rows per step
  static constexpr int NST = (RH + RG - 1) / RG;   // steps
  static_assert(RG >= 1, "tile wider than the workgroup");
};

// Four red-black half-sweeps on an LDS tile whose (0, 0) is global
// (gx0, gy0), each over the cells of its colour within MG0 - 1 of the tile.
// (A half-sweep spoils one more ring of the halo than the one before, so
// after the four the cells within MG0 - 4 of the tile hold the exact
// RB-GS iterate, the same as with regions shrinking by one per half-sweep.)
// The region is the same for every half-sweep, so each thread's cells,
// their LDS slots, which neighbours take part and the diagonal are fixed
// once for both colours.  The update is the exact solve of the cell's row
// of the level operator: neighbours outside the domain are absent (the
// level-0 outlet row is not updated, a coarse level's outlet ghost enters
// the diagonal as dE_last).
template <int TX, int TY, int H, int MG0, int NTH>
__device__ __forceinline__ void tb_tile_sweeps(const TbLevel& L, float* s_p, const float* s_b,
                                               int gx0, int gy0) {
  constexpr int RW = TY + 2 * H;
  using T = TbTile<TX + 2 * H, RW>;
  constexpr int MG = MG0 - 1;                      // margin of every half-sweep
  constexpr int W2 = (TY + 2 * MG) / 2, HH = TX + 2 * MG;
  constexpr int NIT = (W2 * HH + NTH - 1) / NTH;   // cells per thread and colour
  const int tid = threadIdx.x;
  const int bp = (gx0 + gy0) & 1;
  const int active_rows = (L.level == 0) ? L.n - 1 : L.n;
  // per cell: own slot, the other plane's slot of (lx, ly) and of
  // (lx, ly + 1), all-ones masks of the neighbours that take part (E, W, S,
  // N), whether it is updated and its update weight
  int own[NIT][2], oth[NIT][2], onx[NIT][2];
  unsigned msk[NIT][2][4];
  bool act[NIT][2];
  float wr[NIT][2];
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int idx = tid + it * NTH;
    const int q = idx < W2 * HH ? idx : 0;  // idle lanes read a valid cell
    const int lx = H - MG + q / W2;
    const int gx = gx0 + lx;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int ly0 = H - MG;
      const int par = (c - bp - lx - ly0) & 1;
      const int ly = ly0 + par + 2 * (q % W2);
      const int gy = gy0 + ly;
      const bool hasW = gx > 0, hasS = gy > 0, hasN = gy < L.m - 1, last = gx == L.n - 1;
      act[it][c] = idx < W2 * HH && gx >= 0 && gx < active_rows && gy >= 0 && gy < L.m;
      own[it][c] = c * T::PS + lx * T::RW2 + (ly >> 1);
      oth[it][c] = (1 - c) * T::PS + lx * T::RW2 + (ly >> 1);
      onx[it][c] = oth[it][c] + (ly & 1);
      msk[it][c][0] = last ? 0u : ~0u;
      msk[it][c][1] = hasW ? ~0u : 0u;
      msk[it][c][2] = hasS ? ~0u : 0u;
      msk[it][c][3] = hasN ? ~0u : 0u;
      wr[it][c] = tb_weight(L, hasW, last, hasS, hasN);
    }
  }
  const float omw = 1.0f - L.omega;
  // neighbour values are read unconditionally (every slot lies in the
  // tile) and masked to 0 where the neighbour is outside the domain
  auto nb = [](float v, unsigned m) { return __uint_as_float(__float_as_uint(v) & m); };
  // a half-sweep writes only its colour's plane and reads only the other
  // plane (and b), so every cell's reads are issued before any write
#pragma unroll
  for (int k = 1; k <= 4; ++k) {
    const int c = (k - 1) & 1;
    float xs[NIT], ys[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int o = oth[it][c], on = onx[it][c];
      const float E = nb(s_p[o + T::RW2], msk[it][c][0]);
      const float W = nb(s_p[o - T::RW2], msk[it][c][1]);
      const float S = nb(s_p[on - 1], msk[it][c][2]);
      const float N = nb(s_p[on], msk[it][c][3]);
      xs[it] = E + W;
      ys[it] = N + S;
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      if (!act[it][c]) continue;
      const int a = own[it][c];
      s_p[a] = tb_update(L, xs[it], ys[it], s_b[a], s_p[a], wr[it][c], omw);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// pre: 2 RB-GS sweeps + residual + restriction (tile TX x TY, both even)
// ---------------------------------------------------------------------------
template <int TX, int TY, int NTH>
__global__ __launch_bounds__(NTH) void k_tb_pre(TbLevel L, const float* __restrict__ p_in,
                                                const float* __restrict__ b,
                                                float* __restrict__ Y, float* __restrict__ Bc,
                                                const aj_mg_stats* gate) {
  constexpr int H = 5, RH = TX + 2 * H, RW = TY + 2 * H;
  using T = TbTile<RH, RW>;
  __shared__ float s_p[T::N];
  __shared__ float s_b[T::N];
  __shared__ float s_r[TX * TY];
  const int n = L.n, m = L.m;
  {
    const long z = blockIdx.z;
    gate += z;
    if (p_in) p_in += z * n * m;
    b += z * n * m;
    Y += z * n * m;
    Bc += z * (n >> 1) * (m >> 1);
  }
  const int gated = gate->converged;
  const int gx0 = blockIdx.y * TX - H, gy0 = blockIdx.x * TY - H;
  const int bp = (gx0 + gy0) & 1;
  const int tid = threadIdx.x;
  {
    using P = TbPass<RH, RW, NTH>;
    const int col = tid % RW, rg = tid / RW;
    const int gyc = clampi(gy0 + col, 0, m - 1);
    float rp[P::NST], rb[P::NST];
#pragma unroll
    for (int st = 0; st < P::NST; ++st) {
      const int lx = rg + st * P::RG;
      const int g = clampi(gx0 + lx, 0, n - 1) * m + gyc;
      rp[st] = p_in ? p_in[g] : 0.0f;
      rb[st] = b[g];
    }
    if (gated) return;
#pragma unroll
    for (int st = 0; st < P::NST; ++st) {
      const int lx = rg + st * P::RG;
      if (rg < P::RG && lx < RH) {
        const int a = T::at(bp, lx, col);
        s_p[a] = rp[st];
        s_b[a] = rb[st];
      }
    }
  }
  __syncthreads();
  tb_tile_sweeps<TX, TY, H, H, NTH>(L, s_p, s_b, gx0, gy0);
  // residual on the tile, write Y (the four neighbours of a cell are in the
  // other plane: slot ly >> 1 of rows lx -+ 1, slots (ly -+ 1) >> 1 of row lx)
  for (int idx = tid; idx < TX * TY; idx += NTH) {
    const int tx = idx / TY, ty = idx % TY;
    const int lx = H + tx, ly = H + ty;
    const int gx = gx0 + lx, gy = gy0 + ly;
    if (gx >= n || gy >= m) continue;
    const int slot = lx * T::RW2 + (ly >> 1);
    const bool odd = (bp + lx + ly) & 1;
    const int a = odd ? slot + T::PS : slot, o = odd ? slot : slot + T::PS;
    const int on = o + (ly & 1);
    const float pc = s_p[a];
    Y[(long)gx * m + gy] = pc;
    s_r[idx] = tb_resid(L, gx, gy, pc, s_b[a], [&](int di, int dj) {
      return di > 0 ? s_p[o + T::RW2] : (di < 0 ? s_p[o - T::RW2] : (dj > 0 ? s_p[on] : s_p[on - 1]));
    });
  }
  __syncthreads();
  const int mc = m >> 1;
  for (int idx = tid; idx < (TX / 2) * (TY / 2); idx += NTH) {
    const int cx = idx / (TY / 2), cy = idx % (TY / 2);
    const int gx = gx0 + H + 2 * cx, gy = gy0 + H + 2 * cy;
    if (gx >= n || gy >= m) continue;
    const int t = (2 * cx) * TY + 2 * cy;
    const float r00 = s_r[t], r10 = s_r[t + TY], r01 = s_r[t + 1], r11 = s_r[t + TY + 1];
    Bc[(long)(gx >> 1) * mc + (gy >> 1)] = 0.25f * ((r00 + r10) + (r01 + r11));
  }
}

// ---------------------------------------------------------------------------
// post: Y + prolong(E_coarse), 2 RB-GS sweeps
// ---------------------------------------------------------------------------
template <int TX, int TY, int NTH>
__global__ __launch_bounds__(NTH) void k_tb_post(TbLevel L, const float* __restrict__ Yin,
                                                 const float* __restrict__ Ec, int nc, int mc,
                                                 float gamc, const float* __restrict__ b,
                                                 float* __restrict__ E,
                                                 const aj_mg_stats* gate) {
  constexpr int H = 4, RH = TX + 2 * H, RW = TY + 2 * H;
  using T = TbTile<RH, RW>;
  __shared__ float s_p[T::N];
  __shared__ float s_b[T::N];
  const int n = L.n, m = L.m;
  {
    const long z = blockIdx.z;
    gate += z;
    Yin += z * n * m;
    Ec += z * nc * mc;
    b += z * n * m;
    E += z * n * m;
  }
  const int gated = gate->converged;
  const int gx0 = blockIdx.y * TX - H, gy0 = blockIdx.x * TY - H;
  const int bp = (gx0 + gy0) & 1;
  const int tid = threadIdx.x;
  {
    using P = TbPass<RH, RW, NTH>;
    const int col = tid % RW, rg = tid / RW;
    const int gy = clampi(gy0 + col, 0, m - 1);
    float rp[P::NST], rb[P::NST];
#pragma unroll
    for (int st = 0; st < P::NST; ++st) {
      const int lx = rg + st * P::RG;
      const int gx = clampi(gx0 + lx, 0, n - 1);
      const int g = gx * m + gy;
      float pv = Yin[g];
      if (!(L.level == 0 && gx == n - 1)) pv = pv + tb_prolong(Ec, nc, mc, gamc, gx, gy);
      rp[st] = pv;
      rb[st] = b[g];
    }
    if (gated) return;
#pragma unroll
    for (int st = 0; st < P::NST; ++st) {
      const int lx = rg + st * P::RG;
      if (rg < P::RG && lx < RH) {
        const int a = T::at(bp, lx, col);
        s_p[a] = rp[st];
        s_b[a] = rb[st];
      }
    }
  }
  __syncthreads();
  tb_tile_sweeps<TX, TY, H, H, NTH>(L, s_p, s_b, gx0, gy0);
  for (int idx = tid; idx < TX * TY; idx += NTH) {
    const int lx = H + idx / TY, ly = H + idx % TY;
    const int gx = gx0 + lx, gy = gy0 + ly;
    if (gx >= n || gy >= m) continue;
    E[(long)gx * m + gy] = s_p[T::at(bp, lx, ly)];
  }
}

#include "aj_mg_tb_coarse.h"
#include "aj_mg_tb_rc.h"

// ---------------------------------------------------------------------------
// FP64 defect correction
// ---------------------------------------------------------------------------
// Row i and column j of the linear cell index c < 2^31 with ny columns,
// from the FP64 product c * (1/ny) and one correction step (a 64-bit
// integer division is a long instruction sequence on this target).
__device__ __forceinline__ int tb_row(long c, int ny, double rny, int* j) {
  const int ci = (int)c;
  int i = (int)((double)ci * rny);
  int r = ci - i * ny;
  if (r < 0) { --i; r += ny; } else if (r >= ny) { ++i; r -= ny; }
  *j = r;
  return i;
}

// Reset of a solve's device state (statistics, live-buffer parity, the
// tolerance), made by workgroup 0 of each instance of the solve's first
// launch; every later launch of the solve reads it.
struct TbBegin {
  aj_mg_stats* st;
  int* state;
  double* dtol;
  double tol;
};

__device__ __forceinline__ void tb_begin(const TbBegin& bg) {
  if (blockIdx.x != 0) return;
  const int z = blockIdx.z;
  aj_mg_stats* st = bg.st + z;
  if (threadIdx.x == 0) {
    bg.dtol[z] = bg.tol;
    st->cycles = 0;
    st->outer = 0;
    st->converged = 0;
    st->nonfinite = 0;
    st->norm = 0.0;
    st->guess_norm = 0.0;
    st->zero_norm = 0.0;
    st->used_guess = 0;
    bg.state[2 * z] = 0;
    bg.state[2 * z + 1] = 0;
  }
  for (int k = threadIdx.x; k < AJ_MG_MAX_NORMS; k += blockDim.x) st->norms[k] = 0.0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_tb_init(const T* __restrict__ p0, const T* __restrict__ rhs,
                                                 double* __restrict__ P, T* __restrict__ bcopy,
                                                 int nx, int ny, TbBegin bg) {
  tb_begin(bg);
  const long c = (long)blockIdx.x * 256 + threadIdx.x;
  if (c >= (long)nx * ny) return;
  int j;
  const int i = tb_row(c, ny, 1.0 / ny, &j);
  const long q = (long)blockIdx.z * nx * ny + c;
  P[q] = (p0 && i < nx - 1) ? (double)p0[q] : 0.0;
  bcopy[q] = rhs[q];
}

// Temporal initial guess (AJ_MG_OPT_TEMPORAL): k_tb_out also writes the
// solve's FP64 result into the history ring (every output of a solve that
// is topped up rewrites the same slot); the next solve starts from the last
// one (order 1) or the linear extrapolation 2 H_last - H_prev (order 2),
// formed by k_tb_start_guess.

// p' = P[cur] + e (e may be null), r = b - A p' in FP64 on rows < nx-1
// (blocks renumbered by xcd_block).

// r = b - A (src + e) in FP64 on rows < nx-1 (r32 <- r, FP64 partial sums
// of r^2); with e, dst <- src + e.  src / dst are fixed per captured outer
// iteration (even / odd), so no device state decides them.  Every load is
// issued before the convergence gate is tested.
template <typename T>
__global__ __launch_bounds__(256) void k_tb_resid(const double* __restrict__ src,
                                                  double* __restrict__ dst,
                                                  const float* __restrict__ e,
                                                  const T* __restrict__ b, float* __restrict__ r32,
                                                  double* __restrict__ partial, int nx, int ny,
                                                  double rny, double ax, double ay,
                                                  const aj_mg_stats* gate) {
  __shared__ double sh[16];
  const long n = (long)nx * ny;
  {
    const long z = blockIdx.z;
    gate += z;
    src += z * n;
    if (dst) dst += z * n;
    if (e) e += z * n;
    b += z * n;
    r32 += z * n;
    partial += z * gridDim.x;
  }
  const int gated = gate->converged;
  const long blk = xcd_block(blockIdx.x, gridDim.x);
  const long c = blk * 256 + threadIdx.x;
  double r = 0.0, pc = 0.0;
  const bool in = c < n;
  int j = 0;
  const int i = in ? tb_row(c, ny, rny, &j) : 0;
  const bool eq = in && i < nx - 1;
  if (eq) {
    auto P = [&](long q) { return e ? src[q] + (double)e[q] : src[q]; };
    pc = P(c);
    const double W = i > 0 ? P(c - ny) : pc;
    const double E = (i + 1 < nx - 1) ? P(c + ny) : 0.0;
    const double S = j > 0 ? P(c - 1) : pc;
    const double N = j < ny - 1 ? P(c + 1) : pc;
    const double Ap = ax * ((E - pc) + (W - pc)) + ay * ((N - pc) + (S - pc));
    r = (double)b[c] - Ap;
  }
  if (gated) return;  // uniform across the grid
  if (in) {
    r32[c] = (float)r;
    if (e) dst[c] = pc;
  }
  const double t = block_sum(r * r, sh);
  if (threadIdx.x == 0) partial[blk] = t;
}

// k_tb_resid with two cells per thread along y (ny even, so both cells share
// a row): the same arithmetic per cell, vector loads and stores; one partial
// sum per 512 cells.
template <typename T>
__global__ __launch_bounds__(256) void k_tb_resid2(const double* __restrict__ src,
                                                   double* __restrict__ dst,
                                                   const float* __restrict__ e,
                                                   const T* __restrict__ b, float* __restrict__ r32,
                                                   double* __restrict__ partial, int nx, int ny,
                                                   double rny, double ax, double ay,
                                                   const aj_mg_stats* gate) {
  __shared__ double sh[16];
  const long n = (long)nx * ny;
  {
    const long z = blockIdx.z;
    gate += z;
    src += z * n;
    if (dst) dst += z * n;
    if (e) e += z * n;
    b += z * n;
    r32 += z * n;
    partial += z * gridDim.x;
  }
  const int gated = gate->converged;
  const long blk = xcd_block(blockIdx.x, gridDim.x);
  const long c = (blk * 256 + threadIdx.x) * 2;
  double r0 = 0.0, r1 = 0.0, p0 = 0.0, p1 = 0.0;
  const bool in = c < n;
  int j = 0;
  const int i = in ? tb_row(c, ny, rny, &j) : 0;
  if (in && i < nx - 1) {
    auto P2 = [&](long q) {
      double2 v = *reinterpret_cast<const double2*>(src + q);
      if (e) {
        const float2 f = *reinterpret_cast<const float2*>(e + q);
        v.x += (double)f.x;
        v.y += (double)f.y;
      }
      return v;
    };
    auto P = [&](long q) { return e ? src[q] + (double)e[q] : src[q]; };
    const double2 C = P2(c);
    p0 = C.x;
    p1 = C.y;
    const double2 W = i > 0 ? P2(c - ny) : C;
    const double2 E = (i + 1 < nx - 1) ? P2(c + ny) : make_double2(0.0, 0.0);
    const double S = j > 0 ? P(c - 1) : p0;
    const double N = j + 2 < ny ? P(c + 2) : p1;
    const double A0 = ax * ((E.x - p0) + (W.x - p0)) + ay * ((p1 - p0) + (S - p0));
    const double A1 = ax * ((E.y - p1) + (W.y - p1)) + ay * ((N - p1) + (p0 - p1));
    r0 = (double)b[c] - A0;
    r1 = (double)b[c + 1] - A1;
  }
  if (gated) return;  // uniform across the grid
  if (in) {
    *reinterpret_cast<float2*>(r32 + c) = make_float2((float)r0, (float)r1);
    if (e) *reinterpret_cast<double2*>(dst + c) = make_double2(p0, p1);
  }
  const double t = block_sum(r0 * r0 + r1 * r1, sh);
  if (threadIdx.x == 0) partial[blk] = t;
}

// The start of a solve without an initial guess, ny even, in one launch:
// the state reset, the FP64 iterate P = 0, the copy of the rhs, and the
// initial residual r = b - A 0 = b (zero on the Dirichlet row) with its
// partial sums -- what k_tb_init followed by k_tb_resid2 on the zero
// iterate write, in the same cells, blocks and summation order.
template <typename T>
__global__ __launch_bounds__(256) void k_tb_start0(const T* __restrict__ rhs,
                                                   double* __restrict__ P,
                                                   T* __restrict__ bcopy, float* __restrict__ r32,
                                                   double* __restrict__ partial, int nx, int ny,
                                                   double rny, TbBegin bg) {
  __shared__ double sh[16];
  tb_begin(bg);
  const long n = (long)nx * ny;
  {
    const long z = blockIdx.z;
    rhs += z * n;
    P += z * n;
    bcopy += z * n;
    r32 += z * n;
    partial += z * gridDim.x;
  }
  const long blk = xcd_block(blockIdx.x, gridDim.x);
  const long c = (blk * 256 + threadIdx.x) * 2;
  double r0 = 0.0, r1 = 0.0;
  if (c < n) {
    int j = 0;
    const int i = tb_row(c, ny, rny, &j);
    const T b0 = rhs[c], b1 = rhs[c + 1];
    if (i < nx - 1) {
      r0 = (double)b0;
      r1 = (double)b1;
    }
    *reinterpret_cast<double2*>(P + c) = make_double2(0.0, 0.0);
    bcopy[c] = b0;
    bcopy[c + 1] = b1;
    *reinterpret_cast<float2*>(r32 + c) = make_float2((float)r0, (float)r1);
  }
  const double t = block_sum(r0 * r0 + r1 * r1, sh);
  if (threadIdx.x == 0) partial[blk] = t;
}

// residual launch: two cells per thread when ny is even; returns the number
// of partial sums per instance
template <typename T>
static int tb_resid_launch(const aj_mg_plan* p, const double* src, double* dst, const float* e,
                           hipStream_t st) {
  const int nx = p->nx, ny = p->ny;
  const unsigned B = (unsigned)p->batch;
  const double ax = 1.0 / (p->dx * p->dx), ay = 1.0 / (p->dy * p->dy);
  const long n = (long)nx * ny;
  if (ny % 2 == 0) {
    const int nb = (int)cdiv(n, 512);
    hipLaunchKernelGGL(k_tb_resid2<T>, dim3(nb, 1, B), dim3(256), 0, st, src, dst, e,
                       (const T*)p->bcopy, p->tB[0], p->partial, nx, ny, 1.0 / ny, ax, ay,
                       (const aj_mg_stats*)p->stats);
    return nb;
  }
  const int nb = p->npartial;
  hipLaunchKernelGGL(k_tb_resid<T>, dim3(nb, 1, B), dim3(256), 0, st, src, dst, e,
                     (const T*)p->bcopy, p->tB[0], p->partial, nx, ny, 1.0 / ny, ax, ay,
                     (const aj_mg_stats*)p->stats);
  return nb;
}

// The start of a solve with an initial guess in one launch: the state
// reset, the FP64 iterate P = guess (p0, or from the history when HIST;
// zero on the Dirichlet row), the copy of the rhs, r32 = r(guess) on the
// equation rows (rows < nx-1) and the partial sums of the guess test:
// r(guess)^2 and of the zero guess's residual b^2.  Neighbours' guess
// values are formed where they are read.
template <typename T, bool HIST>
__global__ __launch_bounds__(256) void k_tb_start_guess(
    const T* __restrict__ p0, const double* __restrict__ Hl, const double* __restrict__ Hp,
    int order, const T* __restrict__ rhs, double* __restrict__ P, T* __restrict__ bcopy,
    float* __restrict__ r32, double* __restrict__ part_g, double* __restrict__ part_z, int nx,
    int ny, double ax, double ay, TbBegin bg) {
  __shared__ double sh[16];
  tb_begin(bg);
  const long n = (long)nx * ny;
  {
    const long z = blockIdx.z;
    if (HIST) {
      Hl += z * n;
      Hp += z * n;
    } else {
      p0 += z * n;
    }
    rhs += z * n;
    P += z * n;
    bcopy += z * n;
    r32 += z * n;
    part_g += z * gridDim.x;
    part_z += z * gridDim.x;
  }
  // the guess at cell q of an equation row
  auto G = [&](long q) -> double {
    if (HIST) {
      const double hl = Hl[q];
      return order >= 2 ? 2.0 * hl - Hp[q] : hl;
    }
    return (double)p0[q];
  };
  double ag = 0.0, az = 0.0;
  // blocks numbered so that each XCD works through a contiguous band of
  // rows (the rows above and below a block are then in the same L2)
  const long blk = xcd_block(blockIdx.x, gridDim.x);
  const long c = blk * 256 + threadIdx.x;
  if (c < n) {
    int j;
    const int i = tb_row(c, ny, 1.0 / ny, &j);
    const T bt = rhs[c];
    bcopy[c] = bt;
    if (i == nx - 1) {
      P[c] = 0.0;
      r32[c] = 0.0f;
    } else {
      const double pc = G(c);
      P[c] = pc;
      const double W = i > 0 ? G(c - ny) : pc;
      const double E = (i + 1 < nx - 1) ? G(c + ny) : 0.0;
      const double S = j > 0 ? G(c - 1) : pc;
      const double N = j < ny - 1 ? G(c + 1) : pc;
      const double Ap = ax * ((E - pc) + (W - pc)) + ay * ((N - pc) + (S - pc));
      const double bv = (double)bt;
      const double r = bv - Ap;
      r32[c] = (float)r;
      ag = r * r;
      az = bv * bv;
    }
  }
  const double tg = block_sum(ag, sh);
  const double tz = block_sum(az, sh);
  if (threadIdx.x == 0) {
    part_g[blk] = tg;
    part_z[blk] = tz;
  }
}

// keep the guess only if its residual is not larger than the zero guess's
// (test = 0: keep any finite guess)
// the fields the adaptive driver reads (converged, nonfinite, outer), also
// written to its pinned host copy so that reading them needs no copy
__device__ __forceinline__ void tb_publish(const aj_mg_stats* st, aj_mg_stats* hs) {
  if (hs) {
    hs->converged = st->converged;
    hs->nonfinite = st->nonfinite;
    hs->outer = st->outer;
  }
}

__global__ void k_tb_guess_final(const double* __restrict__ part_g,
                                 const double* __restrict__ part_z, int np, double dxdy,
                                 int test, int rel, double* __restrict__ dtol, aj_mg_stats* st,
                                 int* state, aj_mg_stats* hs) {
  __shared__ double sh[16];
  {
    const int z = blockIdx.x;  // one workgroup per instance
    part_g += (long)z * np;
    part_z += (long)z * np;
    dtol += z;
    st += z;
    state += 2 * z;
    if (hs) hs += z;
  }
  double ag = 0.0, az = 0.0;
#pragma unroll 8
  for (int k = threadIdx.x; k < np; k += 256) {
    ag += part_g[k];
    az += part_z[k];
  }
  const double tg = block_sum(ag, sh);
  const double tz = block_sum(az, sh);
  if (threadIdx.x == 0) {
    const double ng = sqrt(tg * dxdy), nz = sqrt(tz * dxdy);
    const int use_guess = (test ? (ng <= nz) : isfinite(ng)) ? 1 : 0;
    st->guess_norm = ng;
    st->zero_norm = nz;
    st->used_guess = use_guess;
    state[1] = use_guess ? 0 : 1;
    const double nrm = use_guess ? ng : nz;
    st->norms[0] = nrm;
    st->norm = nrm;
    // a rejected (e.g. NaN) guess does not make the solve non-finite
    if (!isfinite(nrm)) st->nonfinite = 1;
    if (rel) dtol[0] *= nz;  // relative to ||b||
    st->converged = (nrm <= dtol[0]) ? 1 : 0;
    tb_publish(st, hs);
  }
}

// fallback to the zero guess: p64 = 0, r32 = b (rows < nx-1)
template <typename T>
__global__ __launch_bounds__(256) void k_tb_fallback(double* __restrict__ P,
                                                     const T* __restrict__ b,
                                                     float* __restrict__ r32,
                                                     const int* __restrict__ state, int nx,
                                                     int ny) {
  const long z = blockIdx.z;
  if (state[2 * z + 1] == 0) return;
  const long c = (long)blockIdx.x * 256 + threadIdx.x;
  if (c >= (long)nx * ny) return;
  int j;
  const int i = tb_row(c, ny, 1.0 / ny, &j);
  const long q = z * nx * ny + c;
  P[q] = 0.0;
  r32[q] = (i < nx - 1) ? (float)b[q] : 0.0f;
}

// flip = 0: initial residual (with `rel`, the tolerance becomes relative
// to it); flip = 1: after an outer iteration
__global__ void k_tb_norm_final(const double* __restrict__ partial, int np, double dxdy,
                                double* __restrict__ dtol, int flip, int cycles_added, int rel,
                                aj_mg_stats* st, int* state, aj_mg_stats* hs) {
  __shared__ double sh[16];
  {
    const int z = blockIdx.x;  // one workgroup per instance
    partial += (long)z * np;
    dtol += z;
    st += z;
    state += 2 * z;
    if (hs) hs += z;
  }
  const int skipped = st->converged;  // this outer iteration was skipped
  // the scalars thread 0 updates, read with the partial sums rather than one
  // after another behind the reduction (only thread 0 writes them, after it)
  const int outer0 = st->outer, cycles0 = st->cycles, parity0 = state[0];
  const double tol0 = dtol[0];
  double acc = 0.0;
#pragma unroll 8
  for (int k = threadIdx.x; k < np; k += 256) acc += partial[k];
  if (skipped) return;
  const double tot = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    const double nrm = sqrt(tot * dxdy);
    // norms[0] = initial residual, norms[k] = after outer iteration k
    const int k = outer0 + (flip ? 1 : 0);
    if (k < AJ_MG_MAX_NORMS) st->norms[k] = nrm;
    if (flip) {
      state[0] = parity0 ^ 1;
      st->outer = k;
      st->cycles = cycles0 + cycles_added;
    }
    st->norm = nrm;
    if (!isfinite(nrm)) st->nonfinite = 1;
    double tol = tol0;
    if (rel) {
      tol *= nrm;
      dtol[0] = tol;
    }
    st->converged = (nrm <= tol) ? 1 : 0;
    tb_publish(st, hs);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_tb_out(const double* __restrict__ P0,
                                                const double* __restrict__ P1,
                                                const int* __restrict__ state, T* __restrict__ out,
                                                double* __restrict__ H, long n) {
  const long c = (long)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const long z = blockIdx.z, q = z * n + c;
  const int s = state[2 * z];
  const double a0 = P0[q], a1 = P1[q];  // both loads in flight with the state
  const double r = s ? a1 : a0;
  out[q] = (T)r;
  if (H) H[q] = r;
}


// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static TbLevel make_level(const aj_mg_plan* p, int l) {
  TbLevel L;
  L.n = p->lnx[l];
  L.m = p->lny[l];
  L.level = l;
  const double s = std::ldexp(1.0, 2 * l);
  const double ax = 1.0 / (p->dx * p->dx) / s, ay = 1.0 / (p->dy * p->dy) / s;
  L.axl = (float)ax;
  L.ayl = (float)ay;
  const double g = l > 0 ? outlet_gamma(l) : 1.0;
  L.gam = (float)g;
  L.dE_last = (float)(1.0 - g) * L.axl;
  L.inv_int = 1.0f / (2.0f * L.axl + 2.0f * L.ayl);
  L.omega = (float)p->omega;
  return L;
}

struct Carve {
  char* base;
  size_t off = 0;
  void* take(size_t bytes) {
    void* r = base ? base + off : nullptr;
    off += (bytes + 255) & ~size_t(255);
    return r;
  }
};

static int tb_cycles(aj_mg_plan* p, hipStream_t st);

int tb_create(aj_mg_plan* p) {
  // hierarchy: levels = 0 coarsens while n*m > kTbCoarseCells; levels = L > 0
  // asks for exactly L levels
  if (p->nx < 3 || p->ny < 2) return fail(AJ_ERR_GRID, "textbook multigrid needs nx >= 3, ny >= 2");
  p->lnx.assign(1, p->nx);
  p->lny.assign(1, p->ny);
  {
    int n = p->nx, m = p->ny;
    auto more = [&]() {
      return p->levels > 0 ? (int)p->lnx.size() < p->levels
                           : ((long)n * m > kTbCoarseCells || n > kCMaxRows);
    };
    while (more() && n % 2 == 0 && m % 2 == 0 && n >= 4 && m >= 2) {
      n /= 2;
      m /= 2;
      p->lnx.push_back(n);
      p->lny.push_back(m);
    }
  }
  const int L = (int)p->lnx.size() - 1;
  const int nL = p->lnx[L], mL = p->lny[L];
  const int rows = (L == 0) ? nL - 1 : nL;
  if ((p->levels > 0 && L + 1 != p->levels) || L + 1 > kMaxLevels || rows < 1 ||
      rows > kCMaxRows) {
    char msg[240];
    std::snprintf(msg, sizeof(msg),
                  "Grid dimensions (%d, %d) not suitable for multigrid: %d levels (requested "
                  "%d), coarsest level %dx%d with %d unknown rows (at most %d)",
                  p->nx, p->ny, L + 1, p->levels, nL, mL, rows, kCMaxRows);
    return fail(AJ_ERR_GRID, msg);
  }
  p->coarse_rows = rows;
  int steps = 0;
  while ((1 << steps) < rows) ++steps;
  p->coarse_steps = steps;
  // exact coarsest solve: orthonormal DCT-II along y; per cosine mode the
  // tridiagonal T_k along x and its parallel-cyclic-reduction multipliers
  std::vector<double> Cm((size_t)mL * mL), Coef((size_t)std::max(steps, 1) * mL * 2 * rows),
      Dinv((size_t)mL * rows);
  std::vector<float> Cf((size_t)mL * mL);
  {
    const double pi = 3.14159265358979323846;
    const double sl = std::ldexp(1.0, 2 * L);
    const double axL = 1.0 / (p->dx * p->dx) / sl, ayL = 1.0 / (p->dy * p->dy) / sl;
    const double dE_last = (L == 0) ? axL : axL * (1.0 - outlet_gamma(L));
    for (int k = 0; k < mL; ++k)
      for (int j = 0; j < mL; ++j) {
        const double v = std::sqrt((k == 0 ? 1.0 : 2.0) / mL) * std::cos(pi * k * (j + 0.5) / mL);
        Cm[(size_t)k * mL + j] = v;
        Cf[(size_t)k * mL + j] = (float)v;
      }
    std::vector<double> a(rows), d(rows), c(rows), na(rows), nd(rows), nc(rows);
    for (int k = 0; k < mL; ++k) {
      const double sk = std::sin(pi * k / (2.0 * mL));
      const double mu = -4.0 * ayL * sk * sk;
      for (int i = 0; i < rows; ++i) {
        const double dW = i > 0 ? axL : 0.0;
        const double dE = i < rows - 1 ? axL : dE_last;
        a[i] = i > 0 ? axL : 0.0;
        c[i] = i < rows - 1 ? axL : 0.0;
        d[i] = mu - (dW + dE);
      }
      for (int t = 0; t < steps; ++t) {
        const int s = 1 << t;
        double* al = &Coef[(((size_t)t * mL + k) * 2) * rows];
        double* ga = al + rows;
        for (int i = 0; i < rows; ++i) {
          al[i] = i >= s ? -a[i] / d[i - s] : 0.0;
          ga[i] = i + s < rows ? -c[i] / d[i + s] : 0.0;
          na[i] = i >= s ? al[i] * a[i - s] : 0.0;
          nc[i] = i + s < rows ? ga[i] * c[i + s] : 0.0;
          nd[i] = d[i] + (i >= s ? al[i] * c[i - s] : 0.0) + (i + s < rows ? ga[i] * a[i + s] : 0.0);
        }
        a.swap(na);
        c.swap(nc);
        d.swap(nd);
      }
      for (int i = 0; i < rows; ++i) {
        if (d[i] == 0.0 || !std::isfinite(d[i]))
          return fail(AJ_ERR_GRID, "textbook multigrid: singular coarse operator");
        Dinv[(size_t)k * rows + i] = 1.0 / d[i];
      }
    }
  }
  p->ml = L;

  const long n0 = (long)p->nx * p->ny;
  const long nb = p->batch;  // instances: every per-instance buffer holds nb copies
  for (int pass = 0; pass < 2; ++pass) {
    Carve cv{pass == 0 ? nullptr : (char*)p->ws};
    p->tB.assign(L + 1, nullptr);
    p->tY.assign(L + 1, nullptr);
    p->tE.assign(L + 1, nullptr);
    for (int l = 0; l <= L; ++l) {
      const size_t n = (size_t)p->lnx[l] * p->lny[l] * nb;
      p->tB[l] = (float*)cv.take(n * 4);
      p->tY[l] = (float*)cv.take(n * 4);
      p->tE[l] = (float*)cv.take(n * 4);
    }
    p->p64[0] = (double*)cv.take(n0 * nb * 8);
    p->p64[1] = (double*)cv.take(n0 * nb * 8);
    p->coarse_C = (double*)cv.take(Cm.size() * 8);
    p->coarse_Cf = (float*)cv.take(Cf.size() * 4);
    p->coarse_coef = (double*)cv.take(Coef.size() * 8);
    p->coarse_dinv = (double*)cv.take(Dinv.size() * 8);
    p->coarse_W = (double*)cv.take((size_t)mL * rows * nb * 8);
    p->coarse_X = (double*)cv.take((size_t)mL * rows * nb * 8);
    p->npartial = (int)cdiv(n0, 256);  // per instance
    p->partial = (double*)cv.take((size_t)p->npartial * nb * 8);
    p->partial2 = (double*)cv.take((size_t)p->npartial * nb * 8);
    p->stats = (aj_mg_stats*)cv.take(sizeof(aj_mg_stats) * nb);
    p->state = (int*)cv.take(2 * sizeof(int) * nb);
    p->dtol = (double*)cv.take(sizeof(double) * nb);
    p->bcopy = cv.take(n0 * nb * (p->dtype == AJ_F64 ? 8 : 4));
    if (pass == 0) {
      p->ws_bytes = cv.off;
      AJ_HIP(hipMalloc(&p->ws, p->ws_bytes));
    }
  }
  AJ_HIP(hipMemcpy(p->coarse_C, Cm.data(), Cm.size() * 8, hipMemcpyHostToDevice));
  AJ_HIP(hipMemcpy(p->coarse_Cf, Cf.data(), Cf.size() * 4, hipMemcpyHostToDevice));
  AJ_HIP(hipMemcpy(p->coarse_coef, Coef.data(), Coef.size() * 8, hipMemcpyHostToDevice));
  AJ_HIP(hipMemcpy(p->coarse_dinv, Dinv.data(), Dinv.size() * 8, hipMemcpyHostToDevice));
  AJ_HIP(hipHostMalloc((void**)&p->host_stats, sizeof(aj_mg_stats) * nb, hipHostMallocDefault));
  // pick the smoother tiles of every level now (one cold and one warm
  // V-cycle on a zero right-hand side, gate open), so that the captured
  // outer iteration is built from the tuned launches
  if (p->ml > 0) {
    AJ_HIP(hipMemset(p->stats, 0, sizeof(aj_mg_stats) * nb));
    AJ_HIP(hipMemset(p->tB[0], 0, (size_t)n0 * nb * sizeof(float)));
    AJ_HIP(hipMemset(p->tE[0], 0, (size_t)n0 * nb * sizeof(float)));
    if (int rc = tb_cycles(p, nullptr)) return rc;
    AJ_HIP(hipStreamSynchronize(nullptr));
  }
  return AJ_OK;
}

// exact solve of the coarsest level: tB[L] -> tE[L]
static int tb_coarse(aj_mg_plan* p, hipStream_t st) {
  const int L = p->ml;
  TbCoarse c;
  c.n = p->lnx[L];
  c.m = p->lny[L];
  c.rows = p->coarse_rows;
  c.steps = p->coarse_steps;
  c.C = p->coarse_C;
  c.Cf = p->coarse_Cf;
  c.coef = p->coarse_coef;
  c.dinv = p->coarse_dinv;
  c.W = p->coarse_W;
  c.X = p->coarse_X;
  const aj_mg_stats* gate = p->stats;
  // b read, e written (the per-mode FP64 intermediates stay in L2)
  ProfScope prof(AJ_PROF_MG_COARSE, (double)c.n * c.m * 8 * p->batch, st);
  const unsigned nb = (unsigned)p->batch;
  hipLaunchKernelGGL(k_tb_cfwd, dim3(cdiv(c.rows, kCT), cdiv(c.m, kCT), nb), dim3(kCThreads), 0, st, c,
                     (const float*)p->tB[L], gate);
  AJ_LAUNCH_CHECK("k_tb_cfwd");
  const int nt = c.rows <= 128 ? 64 : (c.rows <= 256 ? 128 : (c.rows <= 512 ? 256 : 512));
  hipLaunchKernelGGL(k_tb_cpcr, dim3(c.m, 1, nb), dim3(nt), (size_t)2 * c.rows * sizeof(double), st,
                     c, gate);
  AJ_LAUNCH_CHECK("k_tb_cpcr");
  hipLaunchKernelGGL(k_tb_cinv, dim3(cdiv(c.n, kCT), cdiv(c.m, kCT), nb), dim3(kCThreads), 0, st, c,
                     p->tE[L], gate);
  AJ_LAUNCH_CHECK("k_tb_cinv");
  return AJ_OK;
}

// one FP32 V-cycle for A_lev e = tB[lev]; result in tE[lev]; warm = use
// tE[lev] as the start (the exact coarsest solve ignores it)
template <int TX, int TY, int NTH>
static int tb_launch_pre(const TbLevel& Lv, int nb, const float* pin, const float* b, float* Y,
                         float* Bc, const aj_mg_stats* gate, hipStream_t st) {
  hipLaunchKernelGGL((k_tb_pre<TX, TY, NTH>), dim3(cdiv(Lv.m, TY), cdiv(Lv.n, TX), nb), dim3(NTH),
                     0, st, Lv, pin, b, Y, Bc, gate);
  AJ_LAUNCH_CHECK("k_tb_pre");
  return AJ_OK;
}

template <int TX, int TY, int NTH>
static int tb_launch_post(const TbLevel& Lv, int nb, const float* Y, const float* Ec, int nc,
                          int mc, float gamc, const float* b, float* E, const aj_mg_stats* gate,
                          hipStream_t st) {
  hipLaunchKernelGGL((k_tb_post<TX, TY, NTH>), dim3(cdiv(Lv.m, TY), cdiv(Lv.n, TX), nb),
                     dim3(NTH), 0, st, Lv, Y, Ec, nc, mc, gamc, b, E, gate);
  AJ_LAUNCH_CHECK("k_tb_post");
  return AJ_OK;
}

template <int W, int R>
static int tb_launch_pre_rc(const TbLevel& Lv, int nb, const float* pin, const float* b, float* Y,
                            float* Bc, const aj_mg_stats* gate, hipStream_t st) {
  if ((Lv.n | Lv.m) & 1) return fail(AJ_ERR_INVALID, "register-column smoother: odd level shape");
  using G = rc::PreGeo<W, R>;
  const int XT = G::x_tiles(Lv.n), YT = (int)cdiv(Lv.m, G::UY);
  hipLaunchKernelGGL((k_tb_pre_rc<W, R>), dim3(8 * cdiv((long)XT * YT, 8), 1, nb), dim3(W * 64), 0,
                     st, Lv, XT, YT, pin, b, Y, Bc, gate);
  AJ_LAUNCH_CHECK("k_tb_pre_rc");
  return AJ_OK;
}

template <int W, int R>
static int tb_launch_post_rc(const TbLevel& Lv, int nb, const float* Y, const float* Ec, int nc,
                             int mc, float gamc, const float* b, float* E,
                             const aj_mg_stats* gate, hipStream_t st) {
  if ((Lv.n | Lv.m) & 1) return fail(AJ_ERR_INVALID, "register-column smoother: odd level shape");
  using G = rc::PostGeo<W, R>;
  const int XT = G::x_tiles(Lv.n), YT = (int)cdiv(Lv.m, G::UY);
  hipLaunchKernelGGL((k_tb_post_rc<W, R>), dim3(8 * cdiv((long)XT * YT, 8), 1, nb), dim3(W * 64), 0,
                     st, Lv, XT, YT, Y, Ec, nc, mc, gamc, b, E, gate);
  AJ_LAUNCH_CHECK("k_tb_post_rc");
  return AJ_OK;
}

// smoother candidates, tuned per level shape: LDS tiles (rows x columns,
// threads) 0-3, register columns (waves x rows per wave) 4-8
constexpr int kTbTileCands = 9;

// pre-smoothing on level lev (lev < coarsest): pin (null = zero start) ->
// Y, coarse rhs tB[lev + 1]
static int tb_pre(aj_mg_plan* p, int lev, const float* pin, float* Y, hipStream_t st) {
  const TbLevel Lv = make_level(p, lev);
  const int n = Lv.n, m = Lv.m, nb = p->batch;
  const aj_mg_stats* gate = p->stats;
  const float* b = p->tB[lev];
  float* Bc = p->tB[lev + 1];
  auto launch = [&](int c, hipStream_t s) -> int {
    switch (c) {
      case 1: return tb_launch_pre<16, 64, 512>(Lv, nb, pin, b, Y, Bc, gate, s);
      case 2: return tb_launch_pre<32, 32, 512>(Lv, nb, pin, b, Y, Bc, gate, s);
      case 3: return tb_launch_pre<16, 32, 256>(Lv, nb, pin, b, Y, Bc, gate, s);
      case 4: return tb_launch_pre_rc<8, 4>(Lv, nb, pin, b, Y, Bc, gate, s);
      case 5: return tb_launch_pre_rc<4, 8>(Lv, nb, pin, b, Y, Bc, gate, s);
      case 6: return tb_launch_pre_rc<8, 6>(Lv, nb, pin, b, Y, Bc, gate, s);
      case 7: return tb_launch_pre_rc<16, 4>(Lv, nb, pin, b, Y, Bc, gate, s);
      case 8: return tb_launch_pre_rc<16, 2>(Lv, nb, pin, b, Y, Bc, gate, s);
      default: return tb_launch_pre<32, 64, 1024>(Lv, nb, pin, b, Y, Bc, gate, s);
    }
  };
  char key[64];
  std::snprintf(key, sizeof(key), "tb_pre/%dx%dx%d", nb, n, m);
  // p_in (warm), b read; Y and the quarter-size coarse rhs written
  ProfScope prof(lev == 0 ? AJ_PROF_MG_PRE0 : AJ_PROF_MG_PRE,
                 (double)n * m * nb * (pin ? 13 : 9), st);
  int pick = 0;
  return tune_pick(key, kTbTileCands, launch, st, &pick);
}

// post-smoothing on level lev: Y + prolong(tE[lev + 1]) -> E
static int tb_post(aj_mg_plan* p, int lev, const float* Y, float* E, hipStream_t st) {
  const TbLevel Lv = make_level(p, lev);
  const int n = Lv.n, m = Lv.m, nb = p->batch;
  const aj_mg_stats* gate = p->stats;
  const float *Ec = p->tE[lev + 1], *b = p->tB[lev];
  const int nc = p->lnx[lev + 1], mc = p->lny[lev + 1];
  const float gamc = (float)outlet_gamma(lev + 1);
  auto launch = [&](int c, hipStream_t s) -> int {
    switch (c) {
      case 1: return tb_launch_post<16, 64, 512>(Lv, nb, Y, Ec, nc, mc, gamc, b, E, gate, s);
      case 2: return tb_launch_post<32, 32, 512>(Lv, nb, Y, Ec, nc, mc, gamc, b, E, gate, s);
      case 3: return tb_launch_post<16, 32, 256>(Lv, nb, Y, Ec, nc, mc, gamc, b, E, gate, s);
      case 4: return tb_launch_post_rc<8, 4>(Lv, nb, Y, Ec, nc, mc, gamc, b, E, gate, s);
      case 5: return tb_launch_post_rc<4, 8>(Lv, nb, Y, Ec, nc, mc, gamc, b, E, gate, s);
      case 6: return tb_launch_post_rc<8, 6>(Lv, nb, Y, Ec, nc, mc, gamc, b, E, gate, s);
      case 7: return tb_launch_post_rc<16, 4>(Lv, nb, Y, Ec, nc, mc, gamc, b, E, gate, s);
      case 8: return tb_launch_post_rc<16, 2>(Lv, nb, Y, Ec, nc, mc, gamc, b, E, gate, s);
      default: return tb_launch_post<32, 64, 1024>(Lv, nb, Y, Ec, nc, mc, gamc, b, E, gate, s);
    }
  };
  char key[64];
  std::snprintf(key, sizeof(key), "tb_post/%dx%dx%d", nb, n, m);
  // Y, quarter-size coarse correction, b read; E written
  ProfScope prof(lev == 0 ? AJ_PROF_MG_POST0 : AJ_PROF_MG_POST, (double)n * m * nb * 13, st);
  int pick = 0;
  return tune_pick(key, kTbTileCands, launch, st, &pick);
}

// one FP32 V-cycle for A_lev e = tB[lev]; result in tE[lev]; warm = start
// from tE[lev] (the exact coarsest solve ignores it)
static int tb_vcycle(aj_mg_plan* p, int lev, bool warm, hipStream_t st) {
  if (lev == p->ml) return tb_coarse(p, st);
  if (int rc = tb_pre(p, lev, warm ? p->tE[lev] : nullptr, p->tY[lev], st)) return rc;
  if (int rc = tb_vcycle(p, lev + 1, false, st)) return rc;
  return tb_post(p, lev, p->tY[lev], p->tE[lev], st);
}

// the V-cycles of an outer iteration for A e = tB[0] from zero, result in
// tE[0]; a single-level hierarchy is one exact solve
static int tb_cycles(aj_mg_plan* p, hipStream_t st) {
  if (p->ml == 0) return tb_coarse(p, st);
  for (int c = 0; c < p->cycles_per_outer; ++c)
    if (int rc = tb_vcycle(p, 0, c > 0, st)) return rc;
  return AJ_OK;
}

static int tb_cycles_per_outer(const aj_mg_plan* p) { return p->ml == 0 ? 1 : p->cycles_per_outer; }

// launches of one defect-correction outer iteration (2 V-cycles, FP64
// residual with the update p64 += e, norm and gate); only plan-owned
// buffers.  Iteration k of a solve reads p64[k & 1] and writes p64[~k & 1]:
// iterations run in order and only those after convergence are skipped, so
// the parity of the enqueue count names the live buffer.
template <typename T>
static int tb_outer_launches(aj_mg_plan* p, int parity, hipStream_t st) {
  const int nx = p->nx, ny = p->ny, B = p->batch;
  if (int rc = tb_cycles(p, st)) return rc;
  // p64, e, b read; p64', r32 written
  ProfScope prof(AJ_PROF_MG_RESID, (double)nx * ny * B * (8 + 4 + sizeof(T) + 8 + 4), st);
  const int np = tb_resid_launch<T>(p, (const double*)p->p64[parity], p->p64[parity ^ 1],
                                    (const float*)p->tE[0], st);
  AJ_LAUNCH_CHECK("k_tb_resid");
  hipLaunchKernelGGL(k_tb_norm_final, dim3(B), dim3(256), 0, st, (const double*)p->partial, np,
                     p->dx * p->dy, p->dtol, 1, tb_cycles_per_outer(p), 0, p->stats, p->state,
                     p->host_stats);
  AJ_LAUNCH_CHECK("k_tb_norm_final");
  return AJ_OK;
}

template <typename T>
static int tb_outer_iteration(aj_mg_plan* p, int parity, hipStream_t st) {
  // live event timing needs the individual launches, so it bypasses the graph
  if (!p->use_graph || prof_on()) return tb_outer_launches<T>(p, parity, st);
  if (!p->graph_exec[parity]) {
    // capture once per plan on a private stream; replays launch on `st`
    hipStream_t cs;
    AJ_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    int rc = AJ_OK;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      rc = tb_outer_launches<T>(p, parity, cs);
      hipGraph_t g = nullptr;
      e = hipStreamEndCapture(cs, &g);
      if (e == hipSuccess && rc == AJ_OK) {
        e = hipGraphInstantiate((hipGraphExec_t*)&p->graph_exec[parity], g, nullptr, nullptr, 0);
        p->graph[parity] = g;
      } else if (g) {
        (void)hipGraphDestroy(g);
      }
    }
    (void)hipStreamDestroy(cs);
    if (rc != AJ_OK) return rc;
    if (e != hipSuccess) return hip_fail(e, "capture of the multigrid outer iteration");
  }
  AJ_HIP(hipGraphLaunch((hipGraphExec_t)p->graph_exec[parity], st));
  return AJ_OK;
}

// state reset, FP64 copies of the guess and the rhs, initial residual (and
// the guess test)
template <typename T>
static int tb_start(aj_mg_plan* p, const T* rhs, const T* p0, double tol, hipStream_t st) {
  const int nx = p->nx, ny = p->ny, B = p->batch;
  const long n = (long)nx * ny;
  const double ax = 1.0 / (p->dx * p->dx), ay = 1.0 / (p->dy * p->dy);
  const int nb = p->npartial;
  // temporal guess: the previous solve's output wrote the next history slot
  if (p->temporal && p->hist_valid) {
    p->hist_last = (p->hist_last + 1) & 1;
    p->hist_count = std::min(p->hist_count + 1, 2);
  }
  p->hist_valid = p->temporal;
  const bool temporal_guess = p->temporal && !p0 && p->hist_count > 0;
  const TbBegin bg{p->stats, p->state, p->dtol, tol};
  if (!p0 && !temporal_guess && ny % 2 == 0) {
    // zero guess: reset, P = 0, rhs copy and the initial residual in one launch
    const int np = (int)cdiv(n, 512);
    hipLaunchKernelGGL(k_tb_start0<T>, dim3(np, 1, B), dim3(256), 0, st, rhs, p->p64[0],
                       (T*)p->bcopy, p->tB[0], p->partial, nx, ny, 1.0 / ny, bg);
    AJ_LAUNCH_CHECK("k_tb_start0");
    hipLaunchKernelGGL(k_tb_norm_final, dim3(B), dim3(256), 0, st, (const double*)p->partial, np,
                       p->dx * p->dy, p->dtol, 0, 0, p->relative, p->stats, p->state,
                       p->host_stats);
    AJ_LAUNCH_CHECK("k_tb_norm_final");
    return AJ_OK;
  }
  if (p0 || temporal_guess) {
    // reset, P = guess, rhs copy and the guess test's sums in one launch; a
    // guess is kept only if its residual is not above the zero guess's
    if (temporal_guess)
      hipLaunchKernelGGL((k_tb_start_guess<T, true>), dim3(nb, 1, B), dim3(256), 0, st,
                         (const T*)nullptr, (const double*)p->hist[p->hist_last],
                         (const double*)p->hist[(p->hist_last + 1) & 1], p->hist_count, rhs,
                         p->p64[0], (T*)p->bcopy, p->tB[0], p->partial, p->partial2, nx, ny, ax,
                         ay, bg);
    else
      hipLaunchKernelGGL((k_tb_start_guess<T, false>), dim3(nb, 1, B), dim3(256), 0, st, p0,
                         (const double*)nullptr, (const double*)nullptr, 0, rhs, p->p64[0],
                         (T*)p->bcopy, p->tB[0], p->partial, p->partial2, nx, ny, ax, ay, bg);
    AJ_LAUNCH_CHECK("k_tb_start_guess");
    hipLaunchKernelGGL(k_tb_guess_final, dim3(B), dim3(256), 0, st, (const double*)p->partial,
                       (const double*)p->partial2, nb, p->dx * p->dy, p->guess_test,
                       p->relative, p->dtol, p->stats, p->state, p->host_stats);
    AJ_LAUNCH_CHECK("k_tb_guess_final");
    hipLaunchKernelGGL(k_tb_fallback<T>, dim3(cdiv(n, 256), 1, B), dim3(256), 0, st, p->p64[0],
                       (const T*)p->bcopy, p->tB[0], (const int*)p->state, nx, ny);
    AJ_LAUNCH_CHECK("k_tb_fallback");
  } else {
    // zero guess with ny odd: reset, P = 0 and the rhs copy, then the
    // one-cell-per-thread residual
    hipLaunchKernelGGL(k_tb_init<T>, dim3(cdiv(n, 256), 1, B), dim3(256), 0, st, (const T*)nullptr,
                       rhs, p->p64[0], (T*)p->bcopy, nx, ny, bg);
    AJ_LAUNCH_CHECK("k_tb_init");
    const int np = tb_resid_launch<T>(p, (const double*)p->p64[0], p->p64[1], nullptr, st);
    AJ_LAUNCH_CHECK("k_tb_resid");
    hipLaunchKernelGGL(k_tb_norm_final, dim3(B), dim3(256), 0, st, (const double*)p->partial, np,
                       p->dx * p->dy, p->dtol, 0, 0, p->relative, p->stats, p->state,
                       p->host_stats);
    AJ_LAUNCH_CHECK("k_tb_norm_final");
  }
  return AJ_OK;
}

template <typename T>
static int tb_finish(aj_mg_plan* p, T* p_out, hipStream_t st) {
  const long n = (long)p->nx * p->ny;
  double* H = p->temporal ? p->hist[(p->hist_last + 1) & 1] : nullptr;
  hipLaunchKernelGGL(k_tb_out<T>, dim3(cdiv(n, 256), 1, p->batch), dim3(256), 0, st,
                     (const double*)p->p64[0], (const double*)p->p64[1], (const int*)p->state,
                     p_out, H, n);
  AJ_LAUNCH_CHECK("k_tb_out");
  return AJ_OK;
}

// host copy of the stats read: every instance converged or non-finite
static bool tb_settled(const aj_mg_plan* p) {
  for (int z = 0; z < p->batch; ++z) {
    const aj_mg_stats& s = p->host_stats[z];
    if (!s.converged && !s.nonfinite) return false;
  }
  return true;
}

static int tb_max_outer(const aj_mg_plan* p) {
  int o = 1;
  for (int z = 0; z < p->batch; ++z) o = std::max(o, p->host_stats[z].outer);
  return o;
}

template <typename T>
static int tb_solve_many_t(aj_mg_plan* const* plans, int count, const void* const* rhs,
                           const void* const* p0, void* const* p_out, int max_cycles, double tol,
                           hipStream_t st) {
  // max_cycles bounds the V-cycles (one exact solve counts as one)
  std::vector<int> done(count, 0), open(count, 0), max_outer(count, 0);
  for (int q = 0; q < count; ++q) {
    const int cpo = tb_cycles_per_outer(plans[q]);
    max_outer[q] = (max_cycles + cpo - 1) / cpo;
  }
  for (int q = 0; q < count; ++q)
    if (int rc = tb_start<T>(plans[q], (const T*)rhs[q], p0 ? (const T*)p0[q] : nullptr, tol, st))
      return rc;
  // non-adaptive plans: every iteration is enqueued (fully stream-ordered;
  // those after convergence exit at their first instruction).  Adaptive
  // plans: as many iterations as their previous solve needed, then the
  // device state of every open plan is read with one synchronisation and
  // each unsettled plan is topped up by one iteration.
  for (int q = 0; q < count; ++q) {
    aj_mg_plan* p = plans[q];
    done[q] = p->adaptive ? std::min(std::max(p->last_outer, 1), max_outer[q]) : max_outer[q];
    open[q] = p->adaptive;
    for (int k = 0; k < done[q]; ++k)
      if (int rc = tb_outer_iteration<T>(p, k & 1, st)) return rc;
  }
  for (;;) {
    bool any = false;
    for (int q = 0; q < count; ++q) any = any || open[q];
    if (!any) break;
    // the norm kernels wrote converged / nonfinite / outer of every
    // instance into the pinned host copy
    AJ_HIP(hipStreamSynchronize(st));
    for (int q = 0; q < count; ++q) {
      if (!open[q]) continue;
      aj_mg_plan* p = plans[q];
      if (tb_settled(p) || done[q] >= max_outer[q]) {
        open[q] = 0;
        p->last_outer = tb_max_outer(p);
        continue;
      }
      if (int rc = tb_outer_iteration<T>(p, done[q] & 1, st)) return rc;
      ++done[q];
    }
  }
  for (int q = 0; q < count; ++q)
    if (int rc = tb_finish<T>(plans[q], (T*)p_out[q], st)) return rc;
  return AJ_OK;
}

int tb_solve_many(aj_mg_plan* const* plans, int count, const void* const* rhs,
                  const void* const* p0, void* const* p_out, int max_cycles, double tol,
                  hipStream_t st) {
  if (count < 1) return AJ_OK;
  for (int q = 0; q < count; ++q)
    if (plans[q]->pend_active)
      return fail(AJ_ERR_INVALID, "textbook multigrid: a deferred solve is open");
  for (int q = 1; q < count; ++q)
    if (plans[q]->dtype != plans[0]->dtype)
      return fail(AJ_ERR_INVALID, "textbook multigrid: plans of one solve need one dtype");
  if (plans[0]->dtype == AJ_F32)
    return tb_solve_many_t<float>(plans, count, rhs, p0, p_out, max_cycles, tol, st);
  return tb_solve_many_t<double>(plans, count, rhs, p0, p_out, max_cycles, tol, st);
}

template <typename T>
static int tb_enqueue_t(aj_mg_plan* p, const T* rhs, const T* p0, T* p_out, int max_cycles,
                        double tol, hipStream_t st) {
  const int cpo = tb_cycles_per_outer(p);
  const int max_outer = (max_cycles + cpo - 1) / cpo;
  if (int rc = tb_start<T>(p, rhs, p0, tol, st)) return rc;
  const int n = std::min(std::max(p->last_outer, 1), max_outer);
  for (int k = 0; k < n; ++k)
    if (int rc = tb_outer_iteration<T>(p, k & 1, st)) return rc;
  if (int rc = tb_finish<T>(p, p_out, st)) return rc;
  // the completion waits for this marker only, not for work the caller
  // queues on the stream after the solve
  if (!p->pend_ev) AJ_HIP(hipEventCreateWithFlags(&p->pend_ev, hipEventDisableTiming));
  AJ_HIP(hipEventRecord(p->pend_ev, st));
  p->pend_active = 1;
  p->pend_done = n;
  p->pend_max = max_outer;
  p->pend_out = p_out;
  p->pend_st = st;
  return AJ_OK;
}

template <typename T>
static int tb_complete_t(aj_mg_plan* p, int* more) {
  hipStream_t st = p->pend_st;
  for (;;) {
    AJ_HIP(hipEventSynchronize(p->pend_ev));  // the norm kernels published the state
    if (tb_settled(p) || p->pend_done >= p->pend_max) break;
    if (int rc = tb_outer_iteration<T>(p, p->pend_done & 1, st)) return rc;
    ++p->pend_done;
    if (int rc = tb_finish<T>(p, (T*)p->pend_out, st)) return rc;
    AJ_HIP(hipEventRecord(p->pend_ev, st));
    *more = 1;
  }
  p->last_outer = tb_max_outer(p);
  p->pend_active = 0;
  return AJ_OK;
}

int tb_solve_enqueue(aj_mg_plan* p, const void* rhs, const void* p0, void* p_out, int max_cycles,
                     double tol, hipStream_t st) {
  if (p->pend_active) return fail(AJ_ERR_INVALID, "textbook multigrid: a deferred solve is open");
  if (p->dtype == AJ_F32)
    return tb_enqueue_t<float>(p, (const float*)rhs, (const float*)p0, (float*)p_out, max_cycles,
                               tol, st);
  return tb_enqueue_t<double>(p, (const double*)rhs, (const double*)p0, (double*)p_out,
                              max_cycles, tol, st);
}

int tb_solve_complete(aj_mg_plan* p, int* more) {
  *more = 0;
  if (!p->pend_active) return AJ_OK;
  return p->dtype == AJ_F32 ? tb_complete_t<float>(p, more) : tb_complete_t<double>(p, more);
}

int tb_solve(aj_mg_plan* p, const void* rhs, const void* p0, void* p_out, int max_cycles,
             double tol, hipStream_t st) {
  if (p->pend_active) return fail(AJ_ERR_INVALID, "textbook multigrid: a deferred solve is open");
  const void* r[1] = {rhs};
  const void* g[1] = {p0};
  void* o[1] = {p_out};
  return tb_solve_many(&p, 1, r, p0 ? g : nullptr, o, max_cycles, tol, st);
}

}  // namespace aj
