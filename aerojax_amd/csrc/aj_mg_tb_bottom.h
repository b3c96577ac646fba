// Bottom of the textbook V-cycle: every level with <= kBottomCells cells
// runs in one workgroup with its iterates and right-hand sides in LDS; the
// coarsest level (<= 1024 unknowns; <= 64 cells wherever the sides stay
// even) is solved exactly in FP64 by a cosine
// transform along y (the walls are Neumann, so DCT-II diagonalises the y
// operator: eigenvalues -4 ay sin^2(pi k / 2m)) and one tridiagonal solve
// along x per cosine mode (LU factors precomputed on the host at plan
// creation).  Included by aj_mg_tb.hip after the level helpers.
#pragma once

constexpr int kMaxLevels = 24;

struct TbBottom {
  int nlev;                 // number of levels handled (bot..L)
  TbLevel lev[kMaxLevels];
  int off[kMaxLevels];      // LDS offsets of each level's arrays
  float gam_next[kMaxLevels];
  float inv_m[kMaxLevels];  // 1/m for the index split
  int coarse_rows;          // unknown rows of the coarsest level
  int total;                // total cells
};

// c / m for 0 <= c < 2^20 via a float reciprocal and one correction step
__device__ __forceinline__ int fdiv_idx(int c, int m, float inv_m) {
  int q = (int)((float)c * inv_m);
  if (q * m > c) --q;
  else if ((q + 1) * m <= c) ++q;
  return q;
}

__device__ __forceinline__ void tb_bottom_rbgs(const TbLevel L, float* p, const float* b,
                                               int sweeps) {
  const int n = L.n, m = L.m;
  const int active_rows = (L.level == 0) ? n - 1 : n;
  const int m2 = (m + 1) / 2;       // colour slots per row (m2 <= blockDim.x)
  const int rpp = blockDim.x / m2;  // rows per pass
  const int r0 = threadIdx.x / m2, q = threadIdx.x - r0 * m2;
  for (int s = 0; s < 2 * sweeps; ++s) {
    const int color = s & 1;
    if (r0 < rpp) {
      for (int i = r0; i < active_rows; i += rpp) {
        const int j = ((color - i) & 1) + 2 * q;
        if (j >= m) continue;
        const int c = i * m + j;
        p[c] = tb_relax(L, i, j, b[c], [&](int di, int dj) { return p[c + di * m + dj]; });
      }
    }
    __syncthreads();
  }
}

// Exact solve of the coarsest level: b (rows x m, float, LDS) -> p (n x m,
// float, LDS; rows beyond `rows` = 0).  sC: [m][m] orthonormal DCT-II,
// sl/sd: [m][rows] LU factors of the per-mode tridiagonal (all staged in
// LDS at kernel start), sup = ax.  Transforms accumulate in FP64.
__device__ __forceinline__ void tb_coarse_exact(const float* b, float* p, int n, int rows, int m,
                                                double* w, const float* sC, const float* sl,
                                                const float* sd, double sup) {
  const int N = rows * m;
  // forward transform along y: w[i][k] = sum_j C[k][j] b[i][j]
  for (int c = threadIdx.x; c < N; c += blockDim.x) {
    const int i = c / m, k = c - i * m;
    const float* Ck = sC + k * m;
    const float* bi = b + i * m;
    double acc = 0.0;
    for (int j = 0; j < m; ++j) acc += (double)Ck[j] * (double)bi[j];
    w[c] = acc;
  }
  __syncthreads();
  // one tridiagonal solve along x per mode (forward / backward substitution)
  for (int k = threadIdx.x; k < m; k += blockDim.x) {
    const float* l = sl + k * rows;
    const float* d = sd + k * rows;
    double y = w[k];
#pragma unroll 8
    for (int i = 1; i < rows; ++i) {
      y = w[i * m + k] - (double)l[i] * y;
      w[i * m + k] = y;
    }
    double x = y * (double)d[rows - 1];
    w[(rows - 1) * m + k] = x;
#pragma unroll 8
    for (int i = rows - 2; i >= 0; --i) {
      x = (w[i * m + k] - sup * x) * (double)d[i];
      w[i * m + k] = x;
    }
  }
  __syncthreads();
  // inverse transform: p[i][j] = sum_k C[k][j] w[i][k]
  for (int c = threadIdx.x; c < n * m; c += blockDim.x) {
    const int i = c / m, j = c - i * m;
    float v = 0.0f;
    if (i < rows) {
      double acc = 0.0;
      for (int k = 0; k < m; ++k) acc += (double)sC[k * m + j] * w[i * m + k];
      v = (float)acc;
    }
    p[c] = v;
  }
  __syncthreads();
}

// p_in (may be null = zero guess, may alias e_out) is the top level's start.
// kStage: copy the coarsest-solve tables into LDS (when they fit) so the
// serial per-mode substitutions read LDS instead of L2.
template <bool kStage>
__global__ __launch_bounds__(kBottomThreads) void k_tb_bottom(
    TbBottom B, const float* b_in, const float* p_in, float* e_out,
    const float* __restrict__ C, const float* __restrict__ lf, const float* __restrict__ invd,
    const aj_mg_stats* gate) {
  if (gate->converged) return;
  extern __shared__ double smd[];
  const int nl = B.nlev;
  const TbLevel Lc = B.lev[nl - 1];
  const int mc = Lc.m, rc = B.coarse_rows;
  double* w = smd;                                           // coarsest transform workspace
  float* tab = reinterpret_cast<float*>(smd + Lc.n * Lc.m);  // staged tables (kStage)
  const int ntab = kStage ? mc * (mc + 2 * rc) : 0;
  float* sp = tab + ntab;                                    // iterates
  float* sb = sp + B.total;                                  // right-hand sides
  const float* sC = kStage ? tab : C;                        // DCT table [m][m]
  const float* sl = kStage ? tab + mc * mc : lf;             // LU multipliers [m][rows]
  const float* sd = kStage ? tab + mc * (mc + rc) : invd;    // 1/diag [m][rows]
  {
    const int n0 = B.lev[0].n * B.lev[0].m;
    for (int c = threadIdx.x; c < n0; c += blockDim.x) {
      sb[c] = b_in[c];
      sp[c] = p_in ? p_in[c] : 0.0f;
    }
    if constexpr (kStage) {
      for (int c = threadIdx.x; c < mc * mc; c += blockDim.x) tab[c] = C[c];
      for (int c = threadIdx.x; c < mc * rc; c += blockDim.x) {
        tab[mc * mc + c] = lf[c];
        tab[mc * (mc + rc) + c] = invd[c];
      }
    }
  }
  __syncthreads();
  // down: smooth and restrict every level above the coarsest
  for (int k = 0; k < nl - 1; ++k) {
    const TbLevel L = B.lev[k];
    float* p = sp + B.off[k];
    float* bb = sb + B.off[k];
    tb_bottom_rbgs(L, p, bb, 2);
    const TbLevel Cl = B.lev[k + 1];
    float* bc = sb + B.off[k + 1];
    const int m = L.m;
    const float inv_mc = B.inv_m[k + 1];
    for (int c = threadIdx.x; c < Cl.n * Cl.m; c += blockDim.x) {
      const int I = fdiv_idx(c, Cl.m, inv_mc), J = c - I * Cl.m;
      const int i = 2 * I, j = 2 * J;
      auto R = [&](int ii, int jj) {
        const int q = ii * m + jj;
        return tb_resid(L, ii, jj, p[q], bb[q], [&](int di, int dj) { return p[q + di * m + dj]; });
      };
      const float r00 = R(i, j), r10 = R(i + 1, j), r01 = R(i, j + 1), r11 = R(i + 1, j + 1);
      bc[c] = 0.25f * ((r00 + r10) + (r01 + r11));
      sp[B.off[k + 1] + c] = 0.0f;  // coarse iterate starts from zero
    }
    __syncthreads();
  }
  tb_coarse_exact(sb + B.off[nl - 1], sp + B.off[nl - 1], Lc.n, rc, mc, w, sC, sl, sd,
                  (double)Lc.axl);
  // up: prolong, correct, smooth
  for (int k = nl - 2; k >= 0; --k) {
    const TbLevel L = B.lev[k];
    const TbLevel Cl = B.lev[k + 1];
    float* p = sp + B.off[k];
    const float* e = sp + B.off[k + 1];
    const float inv_m = B.inv_m[k];
    const float gam = B.gam_next[k];
    for (int c = threadIdx.x; c < L.n * L.m; c += blockDim.x) {
      const int i = fdiv_idx(c, L.m, inv_m), j = c - i * L.m;
      if (L.level == 0 && i == L.n - 1) continue;
      p[c] = p[c] + tb_prolong(e, Cl.n, Cl.m, gam, i, j);
    }
    __syncthreads();
    tb_bottom_rbgs(L, p, sb + B.off[k], 2);
  }
  {
    const int n0 = B.lev[0].n * B.lev[0].m;
    for (int c = threadIdx.x; c < n0; c += blockDim.x) e_out[c] = sp[c];
  }
}
