// Device side of the reference-exact multigrid (pressure_solvers/
// multigrid_solver.py:55-241), shared by the single-grid plan and the
// batched driver.
#pragma once

#include "aj_common.h"

namespace aj {

template <typename T>
struct RefJacobiArgs {
  const T* p_in;  // null -> zero initial guess
  const T* e_c;   // null -> no prolongation; else coarse correction
  int ncx, ncy;
  const T* b;
  int b_zero_last;  // level 0: b[-1,:] = 0
  T* p_out;
  int nx, ny;
  T ax, ay, inv_diag;
  int clamp_last;
  const aj_mg_stats* gate;
  // non-null (no coarse correction only): sums of the squared residual
  // b - A p of the input iterate over the tile cells of each of the 4 waves,
  // in FP64, written to res_out[0..3]
  double* res_out = nullptr;
};

__device__ __forceinline__ bool gated(const aj_mg_stats* g) {
  return g != nullptr && *(volatile const int*)&g->converged != 0;
}

// prolong (multigrid_solver.py:134-156) evaluated at fine cell (i, j)
template <typename T>
__device__ __forceinline__ T prolong_ref(const T* __restrict__ c, int ncx, int ncy, int i, int j) {
#pragma clang fp contract(off)
  const int nxf = 2 * ncx, nyf = 2 * ncy;
  if (i == nxf - 1 || j == nyf - 1) return T(0);
  const int I = i >> 1, J = j >> 1;
  const bool io = i & 1, jo = j & 1;
  const T c00 = c[(long)I * ncy + J];
  if (!io && !jo) return c00;
  if (io && !jo) return T(0.5) * (c00 + c[(long)(I + 1) * ncy + J]);
  if (!io && jo) return T(0.5) * (c00 + c[(long)I * ncy + J + 1]);
  return T(0.25) * (((c00 + c[(long)(I + 1) * ncy + J]) + c[(long)I * ncy + J + 1]) +
                    c[(long)(I + 1) * ncy + J + 1]);
}

// the same prolongation from a coarse patch in LDS whose (0, 0) is coarse
// cell (cI0, cJ0), row length CW.  All four corners are read and the case
// is selected, so the lanes of a wave (alternating parities) do not diverge;
// each case adds the same values in the same order as prolong_ref.  The
// patch holds the cells (I + 1, J + 1) of every cell of the region it serves.
template <typename T>
__device__ __forceinline__ T prolong_ref_lds(const T* __restrict__ c, int cI0, int cJ0, int CW,
                                             int ncx, int ncy, int i, int j) {
#pragma clang fp contract(off)
  const int nxf = 2 * ncx, nyf = 2 * ncy;
  const int I = i >> 1, J = j >> 1;
  const bool io = i & 1, jo = j & 1;
  const int l = (I - cI0) * CW + (J - cJ0);
  const T c00 = c[l], c10 = c[l + CW], c01 = c[l + 1], c11 = c[l + CW + 1];
  const T s1 = c00 + (io ? c10 : c01);
  const T r = (io && jo) ? T(0.25) * ((s1 + c01) + c11) : (io || jo) ? T(0.5) * s1 : c00;
  return (i == nxf - 1 || j == nyf - 1) ? T(0) : r;
}

// prolong_ref from a whole coarse grid (in LDS) with the case selected
// rather than branched on; coarse indices are clamped into the grid as the
// tile kernels' coarse patch is, and each case adds the same values in the
// same order as prolong_ref
template <typename T>
__device__ __forceinline__ T prolong_ref_sel(const T* __restrict__ c, int ncx, int ncy, int i,
                                             int j) {
#pragma clang fp contract(off)
  const int nxf = 2 * ncx, nyf = 2 * ncy;
  const int I = min(i >> 1, ncx - 1), J = min(j >> 1, ncy - 1);
  const int I1 = min(I + 1, ncx - 1), J1 = min(J + 1, ncy - 1);
  const bool io = i & 1, jo = j & 1;
  const T c00 = c[I * ncy + J], c10 = c[I1 * ncy + J], c01 = c[I * ncy + J1];
  const T c11 = c[I1 * ncy + J1];
  const T s1 = c00 + (io ? c10 : c01);
  const T r = (io && jo) ? T(0.25) * ((s1 + c01) + c11) : (io || jo) ? T(0.5) * s1 : c00;
  return (i == nxf - 1 || j == nyf - 1) ? T(0) : r;
}

// The per-cell arithmetic of the reference multigrid, written out operation
// by operation (explicit fused multiply-adds), so that every kernel that
// sweeps or forms residuals (LDS tiles, register columns, the in-LDS
// coarse cycles) gives the same bits.
//
// One Jacobi update (multigrid_solver.py:158-171) from the edge-padded
// neighbours: (ax (E + W) + ay (N + S) - b) / (2 (ax + ay)).
template <typename T>
__device__ __forceinline__ T ref_jacobi_val(T pW, T pE, T pS, T pN, T bv, T ax, T ay, T inv_diag) {
#pragma clang fp contract(off)
  return fma(ax, pE + pW, fma(ay, pN + pS, -bv)) * inv_diag;
}

// b - apply_laplacian(p) from the centre value, its edge-padded
// neighbours and b (multigrid_solver.py:173-183, 198)
template <typename T>
__device__ __forceinline__ T ref_residual_of(T pc, T pW, T pE, T pS, T pN, T bv, T ax, T ay) {
#pragma clang fp contract(off)
  const T lap = fma(ax, fma(T(-2), pc, pE + pW), ay * fma(T(-2), pc, pN + pS));
  return bv - lap;
}

// b - apply_laplacian(p) at (i, j)
template <typename T>
__device__ __forceinline__ T ref_residual(const T* __restrict__ p, const T* __restrict__ b,
                                          int bzl, int nx, int ny, T ax, T ay, int i, int j) {
  const long g = (long)i * ny + j;
  const T pc = p[g];
  const T pW = i > 0 ? p[g - ny] : pc, pE = i < nx - 1 ? p[g + ny] : pc;
  const T pS = j > 0 ? p[g - 1] : pc, pN = j < ny - 1 ? p[g + 1] : pc;
  const T bv = (bzl && i == nx - 1) ? T(0) : b[g];
  return ref_residual_of(pc, pW, pE, pS, pN, bv, ax, ay);
}

// NS Jacobi sweeps (smooth, multigrid_solver.py:158-171) on one tile; `c`
// holds the coarse patch the prolongation of the tile region reads
template <typename T, int TX, int TY, int NS>
struct JacobiLds {
  static constexpr int RN = (TX + 2 * NS) * (TY + 2 * NS);
  static constexpr int CH = (TX + 2 * NS) / 2 + 2, CW = (TY + 2 * NS) / 2 + 2;
  T p[2][RN];
  T b[RN];
  T c[CH * CW];
};

// NS Jacobi sweeps on the tile with origin (x0, y0), 256 threads.  Every
// pass over a region keeps each thread on one column and steps it over the
// rows (W columns per pass: 256 / W rows at a time), so a thread's column
// index, its wall tests and its y-neighbour offsets are fixed for the pass
// and only the row advances; the update of a cell is the same arithmetic in
// any thread layout.
template <typename T, int TX, int TY, int NS>
__device__ __forceinline__ void ref_jacobi_tile(const RefJacobiArgs<T>& a, int x0, int y0,
                                                JacobiLds<T, TX, TY, NS>& L) {
  constexpr int NT = 256;
  constexpr int RH = TX + 2 * NS, RW = TY + 2 * NS;
  static_assert(RW <= NT && TY <= NT, "tile wider than the workgroup");
  auto& s_p = L.p;
  auto& s_b = L.b;
  const int nx = a.nx, ny = a.ny;
  const int gx0 = x0 - NS, gy0 = y0 - NS;
  const int tid = threadIdx.x;
  {
    // all loads of a thread in flight together (clamped, always in bounds);
    // with a coarse correction, its patch is staged in LDS first and the
    // prolongation of every region cell reads it there
    using J = JacobiLds<T, TX, TY, NS>;
    const int cI0 = clampi(gx0, 0, nx - 1) >> 1, cJ0 = clampi(gy0, 0, ny - 1) >> 1;
    constexpr int NC = (J::CH * J::CW + NT - 1) / NT;
    T rc[NC];
    if (a.e_c) {
#pragma unroll
      for (int it = 0; it < NC; ++it) {
        const int idx = min(tid + it * NT, J::CH * J::CW - 1);
        const int cx = idx / J::CW, cy = idx - cx * J::CW;
        const int I = min(cI0 + cx, a.ncx - 1), Jc = min(cJ0 + cy, a.ncy - 1);
        rc[it] = a.e_c[(long)I * a.ncy + Jc];
      }
    }
    constexpr int RG = NT / RW;                 // region rows per pass
    constexpr int NST = (RH + RG - 1) / RG;     // passes
    const int col = tid % RW, rg = tid / RW;
    const int gyc = clampi(gy0 + col, 0, ny - 1);
    T rp[NST], rb[NST];
#pragma unroll
    for (int st = 0; st < NST; ++st) {
      const int gx = clampi(gx0 + rg + st * RG, 0, nx - 1);
      const long g = (long)gx * ny + gyc;
      rp[st] = a.p_in ? a.p_in[g] : T(0);
      rb[st] = (a.b_zero_last && gx == nx - 1) ? T(0) : a.b[g];
    }
    if (gated(a.gate)) return;
    if (a.e_c) {
#pragma unroll
      for (int it = 0; it < NC; ++it) {
        const int idx = tid + it * NT;
        if (idx < J::CH * J::CW) L.c[idx] = rc[it];
      }
      __syncthreads();
    }
#pragma unroll
    for (int st = 0; st < NST; ++st) {
      const int lx = rg + st * RG;
      if (rg < RG && lx < RH) {
        T pv = rp[st];
        if (a.e_c) {
          const int gx = clampi(gx0 + lx, 0, nx - 1);
          pv = pv + prolong_ref_lds(L.c, cI0, cJ0, J::CW, a.ncx, a.ncy, gx, gyc);
        }
        s_p[0][lx * RW + col] = pv;
        s_b[lx * RW + col] = rb[st];
      }
    }
  }
  __syncthreads();
  if (a.res_out) {
    // residual of the input iterate on the tile cells, from the loaded
    // region (b already zero on the last row when b_zero_last)
    constexpr int RGT = NT / TY;
    const int ty = tid % TY, rt = tid / TY;
    const int ly = NS + ty, gy = gy0 + ly;
    double acc = 0.0;
    if (rt < RGT && gy < ny) {
      const int dS = gy > 0 ? 1 : 0, dN = gy < ny - 1 ? 1 : 0;
      for (int tx = rt; tx < TX; tx += RGT) {
        const int lx = NS + tx, gx = gx0 + lx;
        if (gx >= nx) break;
        const int l = lx * RW + ly;
        const T* src = s_p[0];
        const T pc = src[l];
        const T pW = gx > 0 ? src[l - RW] : pc, pE = gx < nx - 1 ? src[l + RW] : pc;
        const T r = ref_residual_of(pc, pW, pE, src[l - dS], src[l + dN], s_b[l], a.ax, a.ay);
        acc += (double)r * (double)r;
      }
    }
    const double t = wave_sum(acc);
    if ((tid & 63) == 0) a.res_out[tid >> 6] = t;
  }
  int cur = 0;
#pragma unroll
  for (int s = 1; s <= NS; ++s) {
    const int m = NS - s;
    const int W = TY + 2 * m, Hh = TX + 2 * m;  // compile-time per unrolled sweep
    const int RG = NT / W;
    const T* src = s_p[cur];
    T* dst = s_p[cur ^ 1];
    const int c2 = tid % W, r2 = tid / W;
    const int ly = NS - m + c2, gy = gy0 + ly;
    const bool yok = r2 < RG && gy >= 0 && gy < ny;
    const int dS = gy > 0 ? 1 : 0, dN = gy < ny - 1 ? 1 : 0;
    if (yok) {
      for (int r = r2; r < Hh; r += RG) {
        const int lx = NS - m + r, gx = gx0 + lx;
        if (gx < 0 || gx >= nx) continue;
        const int l = lx * RW + ly;
        const T pW = src[gx > 0 ? l - RW : l], pE = src[gx < nx - 1 ? l + RW : l];
        const T pS = src[l - dS], pN = src[l + dN];
        dst[l] = ref_jacobi_val(pW, pE, pS, pN, s_b[l], a.ax, a.ay, a.inv_diag);
      }
    }
    __syncthreads();
    cur ^= 1;
  }
  {
    constexpr int RGT = NT / TY;
    const int ty = tid % TY, rt = tid / TY;
    const int ly = NS + ty, gy = gy0 + ly;
    if (rt < RGT && gy < ny) {
      for (int tx = rt; tx < TX; tx += RGT) {
        const int lx = NS + tx, gx = gx0 + lx;
        if (gx >= nx) break;
        const T val = (a.clamp_last && gx == nx - 1) ? T(0) : s_p[cur][lx * RW + ly];
        a.p_out[(long)gx * ny + gy] = val;
      }
    }
  }
}

// residual + restrict (multigrid_solver.py:108-132, 198-199) for coarse
// cell (I, J)
template <typename T>
__device__ __forceinline__ void ref_resrestrict_cell(const T* __restrict__ p,
                                                     const T* __restrict__ b, int bzl, int nx,
                                                     int ny, T ax, T ay, T* __restrict__ bc,
                                                     int ncy, int degenerate, int I, int J) {
  if (degenerate) {  // coarse side < 2: fine[:1, :1]
    if (I == 0 && J == 0) bc[0] = ref_residual(p, b, bzl, nx, ny, ax, ay, 0, 0);
    return;
  }
  const int i = 2 * I, j = 2 * J;
  const T r00 = ref_residual(p, b, bzl, nx, ny, ax, ay, i, j);
  const T r10 = ref_residual(p, b, bzl, nx, ny, ax, ay, i + 1, j);
  const T r01 = ref_residual(p, b, bzl, nx, ny, ax, ay, i, j + 1);
  const T r11 = ref_residual(p, b, bzl, nx, ny, ax, ay, i + 1, j + 1);
  bc[(long)I * ncy + J] = T(0.25) * (((r00 + r10) + r01) + r11);
}

// coarsest level: NSW Jacobi sweeps from zero in one workgroup
constexpr int kCoarseThreads = 1024;
constexpr int kCoarsePerThread = 8;
constexpr int kMaxRefLevels = 8;  // levels of one workgroup's in-LDS V-cycle

// nsw sweeps from zero on the whole coarsest grid (n <= kCoarseThreads *
// kCoarsePerThread cells) by one workgroup; s0 holds 2 n elements of LDS
template <typename T>
__device__ __forceinline__ void ref_coarse_wg(const T* __restrict__ b, T* __restrict__ out,
                                              int nx, int ny, T ax, T ay, T inv_diag, int nsw,
                                              T* s0) {
  const int n = nx * ny;
  T* s1 = s0 + n;
  T breg[kCoarsePerThread];
#pragma unroll
  for (int k = 0; k < kCoarsePerThread; ++k) {
    const int c = threadIdx.x + k * kCoarseThreads;
    breg[k] = c < n ? b[c] : T(0);
    if (c < n) s0[c] = T(0);
  }
  __syncthreads();
  T* src = s0;
  T* dst = s1;
  for (int s = 0; s < nsw; ++s) {
#pragma unroll
    for (int k = 0; k < kCoarsePerThread; ++k) {
      const int c = threadIdx.x + k * kCoarseThreads;
      if (c < n) {
        const int i = c / ny, j = c - i * ny;
        const T pc = src[c];
        const T pW = i > 0 ? src[c - ny] : pc, pE = i < nx - 1 ? src[c + ny] : pc;
        const T pS = j > 0 ? src[c - 1] : pc, pN = j < ny - 1 ? src[c + 1] : pc;
        dst[c] = ref_jacobi_val(pW, pE, pS, pN, breg[k], ax, ay, inv_diag);
      }
    }
    __syncthreads();
    T* t = src; src = dst; dst = t;
  }
#pragma unroll
  for (int k = 0; k < kCoarsePerThread; ++k) {
    const int c = threadIdx.x + k * kCoarseThreads;
    if (c < n) out[c] = src[c];
  }
}

// One Jacobi sweep (multigrid_solver.py:158-171) of a whole nx x ny grid in
// LDS by the kCoarseThreads threads of a workgroup: dst = J(src) with rhs bq
// (ZERO: src is zero everywhere and is not read; CLAMP: the last row of dst
// is stored as zero), with the per-cell arithmetic of ref_jacobi_tile and
// ref_coarse_wg.  Each thread steps its (row, column) by kCoarseThreads
// cells; the barrier at the end makes dst visible to the workgroup.
template <typename T, bool ZERO, bool CLAMP>
__device__ __forceinline__ void ref_sweep_lds(const T* __restrict__ src, T* __restrict__ dst,
                                              const T* __restrict__ bq, int nx, int ny, T ax,
                                              T ay, T inv_diag) {
  const int n = nx * ny;
  const int di = kCoarseThreads / ny, dj = kCoarseThreads - di * ny;
  int i = (int)threadIdx.x / ny, j = (int)threadIdx.x - ((int)threadIdx.x / ny) * ny;
#pragma unroll 4
  for (int c = threadIdx.x; c < n; c += kCoarseThreads) {
    T pc = T(0), pW = T(0), pE = T(0), pS = T(0), pN = T(0);
    if (!ZERO) {
      pc = src[c];
      pW = i > 0 ? src[c - ny] : pc;
      pE = i < nx - 1 ? src[c + ny] : pc;
      pS = j > 0 ? src[c - 1] : pc;
      pN = j < ny - 1 ? src[c + 1] : pc;
    }
    const T val = ref_jacobi_val(pW, pE, pS, pN, bq[c], ax, ay, inv_diag);
    dst[c] = (CLAMP && i == nx - 1) ? T(0) : val;
    i += di;
    j += dj;
    if (j >= ny) {
      j -= ny;
      ++i;
    }
  }
  __syncthreads();
}

// The reference V-cycle (multigrid_solver.py:186-203) from a level >= 1 down
// to the coarsest, for one grid by one workgroup of kCoarseThreads threads,
// every level in LDS: on each level but the coarsest 2 sweeps from zero
// (last row stored as zero after them), residual + restriction, the next
// level, then the prolonged correction added and 2 more sweeps (last row
// zero); on the coarsest 10 sweeps from zero.  The per-cell arithmetic is
// that of the per-level launches (ref_jacobi_tile, ref_resrestrict_cell,
// prolong_ref_lds, ref_coarse_wg), so the result is the same.  nl levels with
// dims nxl[q] x nyl[q] (q = 0 the top), rhs b (global, top level),
// correction out (global, top level).  LDS: w0, w1 (n0 each), then b_q and
// y_q per level.
template <typename T>
__device__ __forceinline__ void ref_subcycle_wg(const T* __restrict__ b, T* __restrict__ out,
                                                const int* nxl, const int* nyl, int nl, T ax,
                                                T ay, T inv_diag, T* sm) {
  const int n0 = nxl[0] * nyl[0];
  T* w0 = sm;
  T* w1 = w0 + n0;
  T* bq[kMaxRefLevels];
  T* yq[kMaxRefLevels];
  {
    T* at = w1 + n0;
    for (int q = 0; q < nl; ++q) {
      const int n = nxl[q] * nyl[q];
      bq[q] = at;
      yq[q] = at + n;
      at += 2 * n;
    }
  }
  for (int c = threadIdx.x; c < n0; c += kCoarseThreads) bq[0][c] = b[c];
  __syncthreads();
  // down: pre-smoothing from zero, residual + restriction
  for (int q = 0; q + 1 < nl; ++q) {
    const int nx = nxl[q], ny = nyl[q];
    ref_sweep_lds<T, true, false>(w1, w0, bq[q], nx, ny, ax, ay, inv_diag);
    ref_sweep_lds<T, false, true>(w0, yq[q], bq[q], nx, ny, ax, ay, inv_diag);
    const int ncx = nxl[q + 1], ncy = nyl[q + 1];
    const int degenerate = (nx / 2 < 2 || ny / 2 < 2) ? 1 : 0;
    for (int c = threadIdx.x; c < ncx * ncy; c += kCoarseThreads) {
      const int I = c / ncy, J = c - I * ncy;
      ref_resrestrict_cell(yq[q], bq[q], 0, nx, ny, ax, ay, bq[q + 1], ncy, degenerate, I, J);
    }
    __syncthreads();
  }
  // coarsest: 10 sweeps from zero, the last one into y
  {
    const int q = nl - 1, nx = nxl[q], ny = nyl[q];
    ref_sweep_lds<T, true, false>(w1, w0, bq[q], nx, ny, ax, ay, inv_diag);
    for (int s = 0; s < 4; ++s) {
      ref_sweep_lds<T, false, false>(w0, w1, bq[q], nx, ny, ax, ay, inv_diag);
      ref_sweep_lds<T, false, false>(w1, w0, bq[q], nx, ny, ax, ay, inv_diag);
    }
    ref_sweep_lds<T, false, false>(w0, yq[q], bq[q], nx, ny, ax, ay, inv_diag);
  }
  // up: y + prolong(correction), 2 sweeps, last row zero
  for (int q = nl - 2; q >= 0; --q) {
    const int nx = nxl[q], ny = nyl[q], n = nx * ny;
    const int ncx = nxl[q + 1], ncy = nyl[q + 1];
    {
      // prolong_ref from LDS with the case selected rather than branched on
      // (the lanes of a wave alternate parities)
      const int di = kCoarseThreads / ny, dj = kCoarseThreads - di * ny;
      int i = (int)threadIdx.x / ny, j = (int)threadIdx.x - ((int)threadIdx.x / ny) * ny;
#pragma unroll 4
      for (int c = threadIdx.x; c < n; c += kCoarseThreads) {
        w0[c] = yq[q][c] + prolong_ref_sel(yq[q + 1], ncx, ncy, i, j);
        i += di;
        j += dj;
        if (j >= ny) {
          j -= ny;
          ++i;
        }
      }
    }
    __syncthreads();
    ref_sweep_lds<T, false, false>(w0, w1, bq[q], nx, ny, ax, ay, inv_diag);
    ref_sweep_lds<T, false, true>(w1, q == 0 ? out : yq[q], bq[q], nx, ny, ax, ay, inv_diag);
  }
}

}  // namespace aj
