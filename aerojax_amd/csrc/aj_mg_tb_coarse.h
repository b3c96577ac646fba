// Exact solve of the coarsest level of the textbook hierarchy.
// The walls are Neumann, so the orthonormal DCT-II along y diagonalises the
// y operator (eigenvalues mu_k = -4 ay sin^2(pi k / 2m)); what is left is
// one tridiagonal system along x per cosine mode k:
//   T_k = tridiag(off, -(dW_i + dE_i) + mu_k, off) on the unknown rows, with
//   dW_0 = 0 (Neumann inlet) and dE of the last row = the level-0 Dirichlet
//   coupling or the coarse outlet ghost (1 - gamma) ax.
// Three kernels, each with one round trip to memory (every global load is
// issued before the first wait; the convergence gate is read alongside):
//   k_tb_cfwd : W[k][i] = sum_j C[k][j] b[i][j], FP64, tiles of 16 rows x
//               16 modes with the whole contraction (<= 128) staged in LDS;
//               four waves per SIMD: each of 4 groups of 256 threads takes a
//               quarter of the contraction, summed in a fixed order
//   k_tb_cpcr : T_k X_k = W_k by parallel cyclic reduction, one workgroup
//               per mode: with stride s = 1, 2, 4, ... row i eliminates its
//               couplings to rows i -+ s,
//                 r_i <- r_i + alpha_i r_{i-s} + gamma_i r_{i+s},
//               where alpha, gamma and the final 1/diagonal depend only on
//               T_k (precomputed in FP64 at plan creation)
//   k_tb_cinv : e[i][j] = sum_k C[k][j] X[k][i] in FP32 (an orthogonal
//               transform: FP32 rounding stays at FP32 level), split as
//               k_tb_cfwd
// Batched plans: instance blockIdx.z (W, X, b and e offset per instance).
// Included by aj_mg_tb.hip.
#pragma once

constexpr int kCT = 16;          // rows x modes (forward) / rows x columns (inverse) per tile
constexpr int kCJC = 128;        // contraction chunk staged in LDS
constexpr int kCMaxSteps = 10;   // PCR steps: rows <= 2^10
constexpr int kCMaxRows = 1024;  // two rows per thread, 512 threads
constexpr int kCSplit = 4;       // contraction groups of 256 threads per tile
constexpr int kCThreads = 256 * kCSplit;
constexpr int kCLoads = kCT * kCJC / kCThreads;  // per-thread staged elements per operand
constexpr int kCPart = kCJC / kCSplit;           // contraction length per group and chunk

struct TbCoarse {
  int n, m, rows, steps;  // level shape, unknown rows, PCR steps
  const double* C;        // [m][m] orthonormal DCT-II
  const float* Cf;        // same, FP32
  const double* coef;     // [steps][m][2][rows] alpha, gamma
  const double* dinv;     // [m][rows] 1 / final diagonal
  double* W;              // [m][rows] transformed right-hand side
  double* X;              // [m][rows] per-mode solution
};

// W[k][i] = sum_j C[k][j] b[i][j]
__global__ __launch_bounds__(kCThreads) void k_tb_cfwd(TbCoarse c, const float* __restrict__ b,
                                                       const aj_mg_stats* gate) {
  __shared__ float sb[kCT][kCJC + 1];
  __shared__ double sc[kCT][kCJC + 1];
  __shared__ double sred[kCSplit - 1][256];
  const int i0 = blockIdx.x * kCT, k0 = blockIdx.y * kCT;
  const int t = threadIdx.x, g = t >> 8, tt = t & 255, r = tt % kCT, kk = tt / kCT;
  {
    const long z = blockIdx.z;
    gate += z;
    b += z * c.n * c.m;
    c.W += z * c.m * c.rows;
  }
  const int gated = gate->converged;
  double acc = 0.0;
  for (int j0 = 0; j0 < c.m; j0 += kCJC) {
    float vb[kCLoads];
    double vc[kCLoads];
#pragma unroll
    for (int u = 0; u < kCLoads; ++u) {
      const int e = t + u * kCThreads, q = e / kCJC, jj = e - q * kCJC;
      const int i = i0 + q, k = k0 + q, j = j0 + jj;
      vb[u] = (i < c.rows && j < c.m) ? b[(long)i * c.m + j] : 0.0f;
      vc[u] = (k < c.m && j < c.m) ? c.C[(long)k * c.m + j] : 0.0;
    }
    if (gated) return;
    if (j0 > 0) __syncthreads();
#pragma unroll
    for (int u = 0; u < kCLoads; ++u) {
      const int e = t + u * kCThreads, q = e / kCJC, jj = e - q * kCJC;
      sb[q][jj] = vb[u];
      sc[q][jj] = vc[u];
    }
    __syncthreads();
    const int jb = g * kCPart;
#pragma unroll 8
    for (int jj = jb; jj < jb + kCPart; ++jj) acc += sc[kk][jj] * (double)sb[r][jj];
  }
  if (g > 0) sred[g - 1][tt] = acc;
  __syncthreads();
  if (g > 0) return;
#pragma unroll
  for (int q = 0; q < kCSplit - 1; ++q) acc += sred[q][tt];
  const int i = i0 + r, k = k0 + kk;
  if (i < c.rows && k < c.m) c.W[(long)k * c.rows + i] = acc;
}

// one workgroup per mode k; the thread's rows are threadIdx.x and
// threadIdx.x + blockDim.x
__global__ __launch_bounds__(512) void k_tb_cpcr(TbCoarse c, const aj_mg_stats* gate) {
  extern __shared__ double sr[];  // ping-pong 2 x rows
  const int k = blockIdx.x, nt = blockDim.x, rows = c.rows, m = c.m;
  const int ia = threadIdx.x, ib = threadIdx.x + nt;
  const bool ona = ia < rows, onb = ib < rows;
  const int ja = ona ? ia : 0, jb = onb ? ib : 0;
  {
    const long z = blockIdx.z;
    gate += z;
    c.W += z * m * rows;
    c.X += z * m * rows;
  }
  const int gated = gate->converged;
  double al[2][kCMaxSteps], ga[2][kCMaxSteps];
#pragma unroll
  for (int s = 0; s < kCMaxSteps; ++s) {
    const bool use = s < c.steps;
    const double* ct = c.coef + ((long)(use ? s : 0) * m + k) * 2 * rows;
    al[0][s] = use ? ct[ja] : 0.0;
    ga[0][s] = use ? ct[rows + ja] : 0.0;
    al[1][s] = use ? ct[jb] : 0.0;
    ga[1][s] = use ? ct[rows + jb] : 0.0;
  }
  const double* Wk = c.W + (long)k * rows;
  const double wa = Wk[ja], wb = Wk[jb];
  const double da = c.dinv[(long)k * rows + ja], db = c.dinv[(long)k * rows + jb];
  if (gated) return;
  if (ona) sr[ia] = wa;
  if (onb) sr[ib] = wb;
  __syncthreads();
  double* cur = sr;
  double* nxt = sr + rows;
#pragma unroll
  for (int s = 0; s < kCMaxSteps; ++s) {
    if (s < c.steps) {  // uniform across the workgroup
      const int d = 1 << s;
      if (ona)
        nxt[ia] = cur[ia] + al[0][s] * cur[ia >= d ? ia - d : ia] +
                  ga[0][s] * cur[ia + d < rows ? ia + d : ia];
      if (onb)
        nxt[ib] = cur[ib] + al[1][s] * cur[ib >= d ? ib - d : ib] +
                  ga[1][s] * cur[ib + d < rows ? ib + d : ib];
      __syncthreads();
      double* tmp = cur;
      cur = nxt;
      nxt = tmp;
    }
  }
  if (ona) c.X[(long)k * rows + ia] = cur[ia] * da;
  if (onb) c.X[(long)k * rows + ib] = cur[ib] * db;
}

// e[i][j] = sum_k C[k][j] X[k][i] on the unknown rows, 0 on the level-0
// Dirichlet row
__global__ __launch_bounds__(kCThreads) void k_tb_cinv(TbCoarse c, float* __restrict__ e,
                                                       const aj_mg_stats* gate) {
  __shared__ float sx[kCJC][kCT + 1];
  __shared__ float sq[kCJC][kCT + 1];
  __shared__ float sred[kCSplit - 1][256];
  const int i0 = blockIdx.x * kCT, j0 = blockIdx.y * kCT;
  const int t = threadIdx.x, g = t >> 8, tt = t & 255, jj = tt % kCT, r = tt / kCT;
  {
    const long z = blockIdx.z;
    gate += z;
    c.X += z * c.m * c.rows;
    e += z * c.n * c.m;
  }
  const int gated = gate->converged;
  float acc = 0.0f;
  for (int k0 = 0; k0 < c.m; k0 += kCJC) {
    float vx[kCLoads], vq[kCLoads];
#pragma unroll
    for (int u = 0; u < kCLoads; ++u) {
      const int e2 = t + u * kCThreads, q = e2 / kCT, rr = e2 - q * kCT;
      const int k = k0 + q, i = i0 + rr, j = j0 + rr;
      vx[u] = (k < c.m && i < c.rows) ? (float)c.X[(long)k * c.rows + i] : 0.0f;
      vq[u] = (k < c.m && j < c.m) ? c.Cf[(long)k * c.m + j] : 0.0f;
    }
    if (gated) return;
    if (k0 > 0) __syncthreads();
#pragma unroll
    for (int u = 0; u < kCLoads; ++u) {
      const int e2 = t + u * kCThreads, q = e2 / kCT, rr = e2 - q * kCT;
      sx[q][rr] = vx[u];
      sq[q][rr] = vq[u];
    }
    __syncthreads();
    const int qb = g * kCPart;
#pragma unroll 16
    for (int q = qb; q < qb + kCPart; ++q) acc += sq[q][jj] * sx[q][r];
  }
  if (g > 0) sred[g - 1][tt] = acc;
  __syncthreads();
  if (g > 0) return;
#pragma unroll
  for (int q = 0; q < kCSplit - 1; ++q) acc += sred[q][tt];
  const int i = i0 + r, j = j0 + jj;
  if (i < c.n && j < c.m) e[(long)i * c.m + j] = i < c.rows ? acc : 0.0f;
}
